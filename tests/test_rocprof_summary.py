"""CPU checks of the rocprofv3 summary tool (cudabubble_amd/rocprof_summary.py):
kernel names map to the solver's own names, schedules and the probe stage stay
apart, and the traffic summary scales each counter by the copy calibration and
writes the JSON of the instantiation the run launched most at the given size."""
import csv
import json

from cudabubble_amd import rocprof_summary as rs


def test_solver_names():
    cases = {
        "void vb::k_xrow<float, 256, false, 1, false>(vb::XaArgs<float>, vb::XaArgs<float> const*)": ("xa", "fp32"),
        "void vb::k_xrow<double, 256, false, 0, true>(vb::XaArgs<double>, vb::XaArgs<double> const*)": ("xrow_sums", "fp64"),
        "void vb::k_xa<double, 256, true, false>(vb::XaArgs<double>, vb::XaArgs<double> const*)": ("xa_f", "fp64"),
        "void vb::k_xa<double, 256, false, false>(vb::XaArgs<double>, vb::XaArgs<double> const*)": ("xa_plain", "fp64"),
        "void vb::k_xsums<float, 256, false>(vb::XaArgs<float>, vb::XaArgs<float> const*)": ("xa_sums", "fp32"),
        "void vb::k_ypass<float, 256, -1, 16, false, 1, true>(vb::YArgs<float>, vb::YArgs<float> const*)": ("y_fwd", "fp32"),
        "void vb::k_ypass<float, 256, 1, 16, false, 1, false>(vb::YArgs<float>, vb::YArgs<float> const*)": ("y_inv", "fp32"),
        "void vb::k_zstage<float, 256, 4, 4, 8, false, 0, false>(vb::ZsArgs<float>, vb::ZsArgs<float> const*)": ("zstage4", "fp32"),
        "void vb::k_zstage<float, 256, 4, 4, 8, false, 0, true>(vb::ZsArgs<float>, vb::ZsArgs<float> const*)": ("zstage4p", "fp32"),
        "void vb::k_zstage<double, 64, 2, 1, 8, true, 0, false>(vb::ZsArgs<double>, vb::ZsArgs<double> const*)": ("zstage1_etdrk2", "fp64"),
        "void vb::k_finalize<false>(vb::FinArgs, vb::FinArgs const*)": ("finalize", None),
    }
    for name, want in cases.items():
        assert rs.solver_name(name) == want, name
    assert rs.solver_name("__amd_rocclr_copyBuffer") is None
    assert rs.kernel_size("void vb::k_xrow<float, 192, true, 0>(vb::XaArgs<float>)") == "192"
    assert rs.kernel_size("void vb::k_xr2c<double, float, 128>(double const*)") == "128"


def _counters(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Value"])
        w.writeheader()
        for k, v in rows:
            w.writerow({"Kernel_Name": k, "Counter_Value": v})


def test_traffic_scaling_and_json_choice(tmp_path):
    copy = "__amd_rocclr_copyBuffer"
    mib = 256 * 1024          # KiB in the calibration copy
    _counters(tmp_path / "cf.csv", [(copy, mib / 2)] * 3)   # FETCH_SIZE reads half the bytes
    _counters(tmp_path / "cw.csv", [(copy, mib)] * 3)
    main = "void vb::k_xrow<float, 256, false, 1>(vb::XaArgs<float>, vb::XaArgs<float> const*)"
    other = "void vb::k_xrow<float, 256, false, 0>(vb::XaArgs<float>, vb::XaArgs<float> const*)"
    small = "void vb::k_xrow<float, 64, true, 0>(vb::XaArgs<float>, vb::XaArgs<float> const*)"
    _counters(tmp_path / "f.csv", [(main, 1000.0)] * 5 + [(other, 1200.0)] * 2 + [(small, 10.0)] * 9)
    _counters(tmp_path / "w.csv", [(main, 500.0)] * 5 + [(other, 600.0)] * 2 + [(small, 5.0)] * 9)
    md = rs.traffic(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), str(tmp_path / "cf.csv"),
                    str(tmp_path / "cw.csv"), 256, ["cmd"], json_dir=str(tmp_path), json_prec="fp32")
    assert "FETCH_SIZE scaled by 2.000 and WRITE_SIZE by 1.000" in md
    assert md.count("| xa | fp32 |") == 3
    j = json.load(open(tmp_path / "traffic_xa_256_fp32.json"))
    # the most-launched instantiation at length 256, not the 64^3 one with more launches
    assert j["rocprof_name"] == main and j["launches"] == 5
    assert j["read_bytes_per_launch"] == 2000.0 * 1024 and j["write_bytes_per_launch"] == 500.0 * 1024


def test_counters_valu_issue(tmp_path):
    k32 = "void vb::k_xrow<float, 256, false, 1, false>(vb::XaArgs<float>, vb::XaArgs<float> const*)"
    k64 = "void vb::k_xrow<double, 256, false, 1, false>(vb::XaArgs<double>, vb::XaArgs<double> const*)"
    with open(tmp_path / "a.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for k, waves, valu in ((k32, 16384, 2.6e7), (k64, 16384, 2.4e7)):
            for _ in range(3):
                w.writerow({"Kernel_Name": k, "Counter_Name": "SQ_WAVES", "Counter_Value": waves})
                w.writerow({"Kernel_Name": k, "Counter_Name": "SQ_INSTS_VALU", "Counter_Value": valu})
    with open(tmp_path / "b.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        w.writerow({"Kernel_Name": k32, "Counter_Name": "GRBM_GUI_ACTIVE", "Counter_Value": 8 * 1.25e5})
    md = rs.counters([str(tmp_path / "a.csv"), str(tmp_path / "b.csv")], 256, ["cmd"])
    # 4 x 2.6e7 / (1024 SIMDs x 1.25e5 cycles) = 0.81
    row32 = [l for l in md.splitlines() if l.startswith("| xa | fp32 |")][0]
    assert "| 0.81 |" in row32 and "| 1587 |" in row32
    row64 = [l for l in md.splitlines() if l.startswith("| xa | fp64 |")][0]
    assert "| - |" in row64   # no cycles counted for it
