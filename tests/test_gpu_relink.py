"""The reference's basic driver (reconstruct/basic/XT_Main.c), compiled
against include/mbir4d.h and linked to phasecon_amd/libmbir4d.so instead of
the reference's libmbir4d.a (oracle/Makefile target _ref/xt_main_basic),
reconstructs from its input files: it reads projections.bin, weights.bin,
proj_angles.bin, proj_times.bin and recon_times.bin from its working
directory (basic/XT_Main.c:90-113, names XT_Main.h:4-8), calls reconstruct()
and leaves the reference's output files.  Its result must agree with the
library called through ctypes on the same scan (final cost within 1e-5,
images within 5e-3 relative RMS: the parallel sweep's order of
floating-point updates, and so the equit where the convergence test stops,
is not reproducible run to run)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

DRIVER = os.path.join(ROOT, "oracle", "_ref", "xt_main_basic")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DRIVER),
                                 reason="oracle/_ref/xt_main_basic not built (make -C oracle ref)")]


def test_relinked_reference_driver_reconstructs(tiny_scan, tmp_path):
    from phasecon_amd.mbir import reconstruct
    s = tiny_scan
    run = tmp_path / "driver"
    run.mkdir()
    for name, key in (("projections", "projections"), ("weights", "weights"),
                      ("proj_angles", "angles"), ("proj_times", "proj_times"),
                      ("recon_times", "recon_times")):
        np.ascontiguousarray(s[key], np.float32).tofile(run / f"{name}.bin")
    args = [DRIVER, "--proj_rows", str(s["proj_rows"]), "--proj_cols", str(s["proj_cols"]),
            "--proj_num", str(s["proj_num"]), "--recon_num", str(s["recon_num"]),
            "--vox_wid", repr(float(s["vox_wid"])), "--rot_center", repr(float(s["rot_center"])),
            "--sig_s", repr(float(s["sig_s"])), "--sig_t", repr(float(s["sig_t"])),
            "--c_s", repr(float(s["c_s"])), "--c_t", repr(float(s["c_t"])),
            "--convg_thresh", "1", "--remove_rings", "0", "--quad_convex"]
    print("running the relinked reference driver on tiny3d", flush=True)
    res = subprocess.run(args, cwd=run, capture_output=True, text=True, timeout=120)
    log = (run / "debug.log").read_text()
    assert res.returncode == 0, res.stderr + log
    assert "Cost value increased" not in log
    n = s["proj_rows"] * s["proj_cols"] * s["proj_cols"]
    drv = np.fromfile(run / "object_time_0.bin", np.float32)
    assert drv.size == n
    drv_cost = np.fromfile(run / "cost.bin", np.float64)[-1]
    lib = tmp_path / "lib"
    lib.mkdir()
    cwd = os.getcwd()
    os.chdir(lib)
    try:
        obj, _, _ = reconstruct(s, convg_thresh=1.0, log_path=str(lib / "debug.log"))
    finally:
        os.chdir(cwd)
    lib_cost = np.fromfile(lib / "cost.bin", np.float64)[-1]
    ref = obj[0, 1:-1].ravel()
    print(f"driver cost {drv_cost:.10g}, library cost {lib_cost:.10g}", flush=True)
    assert abs(drv_cost - lib_cost) <= 1e-5 * abs(lib_cost)
    assert np.sqrt(np.mean((drv - ref) ** 2) / np.mean(ref ** 2)) < 5e-3
