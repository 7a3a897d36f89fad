"""Writes tests/golden/config2_128_adaptive.npz: the FP64 oracle's run of the
drop-in configuration, for the GPU parity test of its full length.

128^3 grid on [-pi, pi]^3, eps = 2h, the union of two seeded ellipsoids
(vb_shape kind 2, seed 20260002), M1 = M2 = 1e4 with alpha/beta taken at t = 0,
ETDRK4 with adaptive dt from 1e-7 doubling every 10 steps and halving when E
rises (belastic.cpp:639-669), 1000 driver steps.

Stored: the accepted dt, E, V, A after every step (the oracle's trace), the
rejection count, and phi after 100 and 1000 steps sampled on every 4th grid
point in each direction (a 32^3 subgrid, float64).  Run from the repository
root: python tests/golden/gen_config2.py (about 10 minutes on 8 cores)."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import cudabubble_amd as vb          # host-only: the seeded shape generator
from oracle import pyoracle

N, SEED, STEPS, SUB = 128, 20260002, 1000, 4


def main(out=os.path.join(ROOT, "tests", "golden", "config2_128_adaptive.npz")):
    eps = 2.0 * 2 * np.pi / N
    phi, _ = vb.shape(vb.SHAPE_TWO_ELLIPSOIDS, SEED, N, N, 1.0, eps)
    o = pyoracle.Oracle(N, eps=eps, M1=1e4, M2=1e4, alpha=0.0, beta=0.0)
    d = phi.copy()
    st = o.fEIF_begin(d, 1e-7, dt_double_every=10)
    t0 = time.time()
    done1, tr1 = o.fEIF_steps(d, st, 100, trace=True)
    phi100 = d[::SUB, ::SUB, ::SUB].copy()
    rej100 = st.rejections
    print("100 steps, %.0f s" % (time.time() - t0), flush=True)
    done2, tr2 = o.fEIF_steps(d, st, STEPS - 100, trace=True)
    print("%d steps, %.0f s" % (STEPS, time.time() - t0), flush=True)
    q = o.params()
    E, V, A, Eb = o.energy(d)
    np.savez_compressed(out, trace=np.concatenate([tr1, tr2]), phi100=phi100, phi1000=d[::SUB, ::SUB, ::SUB],
                        rejections100=rej100, rejections=st.rejections, evaluations=st.evaluations,
                        steps=done1 + done2, status=st.status, time_current=st.time_current,
                        alpha=q.alpha, beta=q.beta, final=np.array([E, V, A, Eb]), sub=SUB)


if __name__ == "__main__":
    main()
