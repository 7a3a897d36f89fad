"""Build recipe of libvesicle.so (hand-written HIP for gfx950) and of the
C++ host programs.  Run as `python -m cudabubble_amd.build` or through
__graft_entry__.build().  Objects go to cudabubble_amd/_build/, the shared
library to cudabubble_amd/libvesicle.so."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvesicle.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -packed-fp32-ops: keep the compiler from forming v_pk_*_f32 instructions
DEV_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-I" + INCLUDE]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + INCLUDE,
              "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def sources():
    hips = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpps = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hips, cpps


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    return r.returncode, " ".join(cmd), out


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hips, cpps = sources()
    cmds = []
    objs = []
    for f in hips:
        o = os.path.join(OBJ, f[:-4] + ".o")
        objs.append(o)
        cmds.append([HIPCC] + DEV_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
    for f in cpps:
        o = os.path.join(OBJ, f[:-4] + ".o")
        objs.append(o)
        cmds.append([HIPCC] + HOST_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
    jobs = jobs or min(16, os.cpu_count() or 4)
    failed = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for rc, cmd, out in ex.map(_run, cmds):
            if rc != 0:
                failed.append((cmd, out))
            elif verbose and out.strip():
                print(out)
    if failed:
        for cmd, out in failed:
            sys.stderr.write("FAILED: %s\n%s\n" % (cmd, out))
        raise RuntimeError("libvesicle build failed (%d units)" % len(failed))
    tmp = LIB + ".tmp"
    rc, cmd, out = _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    if rc != 0:
        raise RuntimeError("link failed: %s\n%s" % (cmd, out))
    os.replace(tmp, LIB)
    build_host_programs()
    return LIB


def build_host_programs():
    """C++ host side above the C ABI (the belastic-shaped facade and its CLI)."""
    if not os.path.isdir(HOST):
        return
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return
    exe = os.path.join(HERE, "cudabubble")
    cmd = ["g++", "-O2", "-std=c++17", "-I" + INCLUDE, "-I" + HOST, "-o", exe] + srcs + \
          ["-L" + HERE, "-lvesicle", "-L/opt/rocm/lib", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    rc, c, out = _run(cmd)
    if rc != 0:
        raise RuntimeError("host program build failed: %s\n%s" % (c, out))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
