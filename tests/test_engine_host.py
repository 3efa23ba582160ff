"""CPU checks of the native library: exports and the host-side stage set-up.

The stage set-up in phasecon_amd/csrc/mbir_engine.c (geometry, prior
weights, detector response, view map) must reproduce the reference's
initStructures bit for bit; it is compared with the golden vectors written
from the reference build.  No GPU call is made here.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")
LIB = os.path.join(ROOT, "phasecon_amd", "libmbir4d.so")

needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="libmbir4d.so not built")


def _header_functions(headers=("mbir4d.h", "mbir_engine.h", "mbir_hip.h", "unet_hip.h")):
    names = set()
    for h in headers:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name in ("if", "sizeof", "defined", "extern") or name.isupper():
                continue
            names.add(name)
    return sorted(names)


@needs_lib
def test_library_exports_every_declared_function():
    from phasecon_amd import lib
    L = C.CDLL(LIB)
    declared = _header_functions()
    assert "reconstruct" in declared and "mbir_hip_icd_equits" in declared
    for name in declared:
        assert hasattr(L, name), name
    assert set(lib.exported_names()) <= set(declared)


PCLIB = os.path.join(ROOT, "phasecon_amd", "libpcmbir4d.so")


@pytest.mark.skipif(not os.path.exists(PCLIB), reason="libpcmbir4d.so not built")
def test_pcmbir_library_exports_every_declared_function():
    from phasecon_amd import pcmbir
    L = C.CDLL(PCLIB)
    declared = _header_functions(("pcmbir_hip.h", "pcmbir4d.h"))
    assert "pcmbir_hip_sd_run" in declared and "phcontomo_reconstruct" in declared
    for name in declared:
        assert hasattr(L, name), name
    assert set(pcmbir.exported_names()) | set(pcmbir.dropin_names()) == set(declared)


PC_ARGS = dict(vox_wid=1.0, rot_center=8.0, mag_sig_s=2e-2, mag_sig_t=1.0, mag_c_s=1e-2,
               mag_c_t=1e-2, phase_sig_s=2e-2, phase_sig_t=1.0, phase_c_s=1e-2, phase_c_t=1e-2,
               convg_thresh=1.0, obj2det_dist=1e5, light_energy=17.0, pag_regparam=50.0)


@pytest.mark.skipif(not os.path.exists(PCLIB), reason="libpcmbir4d.so not built")
@pytest.mark.parametrize("rows,cols,why", [(2, 16, "minimum number of projection rows"),
                                           (4, 15, "columns should be even"),
                                           (6, 16, "powers of two")])
def test_pcmbir_dropin_rejects_bad_sizes_before_any_gpu_call(rows, cols, why, tmp_path,
                                                             monkeypatch):
    """The reference's size checks (PCMBIR_3D/XT_Engine.c:100-103) and this
    build's power-of-two rule return -1 with the message in the log, on the
    host, before a device is touched."""
    from phasecon_amd import pcmbir
    monkeypatch.chdir(tmp_path)
    log = tmp_path / "debug.log"
    rc = pcmbir.phcontomo_reconstruct(np.ones((3, rows, cols)), np.ones((rows, cols)),
                                      np.zeros(3), np.arange(3.0), [0.0, 3.0], **PC_ARGS,
                                      recon_type=1, log_path=str(log))
    assert rc == -1
    assert why in log.read_text()
    assert not list(tmp_path.glob("*.bin"))


@pytest.mark.skipif(not os.path.exists(PCLIB), reason="libpcmbir4d.so not built")
def test_pcmbir_forward_project_needs_its_phantom_files(tmp_path, monkeypatch):
    """Without the phantom files (the reference reads fixed paths; here the
    environment names them) the simulator returns -1 and says what to set,
    before any GPU work."""
    monkeypatch.delenv("PHCONTOMO_MAG_PHANTOM", raising=False)
    monkeypatch.delenv("PHCONTOMO_PHASE_PHANTOM", raising=False)
    L = C.CDLL(PCLIB)
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen(str(tmp_path / "log").encode(), b"w")
    m, b = C.c_void_p(), C.c_void_p()
    L.phcontomo_forward_project.argtypes = [C.c_void_p] * 3 + [C.c_int32] * 3 + [C.c_float] * 5 + [C.c_void_p]
    rc = L.phcontomo_forward_project(C.byref(m), C.byref(b), None, 4, 16, 3, 1.0, 8.0, 1e5,
                                     17.0, 50.0, C.c_void_p(fp))
    libc.fclose(C.c_void_p(fp))
    assert rc == -1
    assert "PHCONTOMO_MAG_PHANTOM" in (tmp_path / "log").read_text()


@needs_lib
@pytest.mark.parametrize("name,gold", [("tiny3d", "mbir_tiny3d.npz"), ("ch4d_small", "mbir_ch4d_small.npz")])
def test_stage_setup_matches_reference(name, gold, request):
    from phasecon_amd import lib as plib
    from phasecon_amd.mbir import view_map
    scan = request.getfixturevalue("tiny_scan" if name == "tiny3d" else "small4d_scan")
    d = np.load(os.path.join(GOLD, gold))
    L = plib.load()
    ptr, views = view_map(scan["proj_times"], scan["recon_times"])
    assert np.array_equal(ptr, d["frame_ptr"])
    assert np.array_equal(views, d["frame_views"][: ptr[-1]])
    g = plib.HipGeom()
    f = C.c_float
    rc = L.mbir_stage_geometry(C.byref(g), C.c_int32(scan["proj_rows"]), C.c_int32(scan["proj_cols"]),
                               C.c_int32(scan["proj_num"]), C.c_int32(scan["recon_num"]),
                               f(scan["vox_wid"]), f(scan["rot_center"]), f(scan["sig_s"]),
                               f(scan["sig_t"]), f(scan["c_s"]), f(scan["c_t"]), C.c_int32(1),
                               C.c_int32(1), C.c_int32(1), f(0.5), f(4.0),
                               ptr.ctypes.data_as(C.c_void_p), C.c_double(1.2), C.c_double(1.2))
    assert rc == 0
    ref = d["dbls"]
    # reference order: delta_r delta_t R0 RMax T0 TMax OffsetR OffsetT x0 y0 z0
    # delta_xy delta_z radius delta_recon sSQ sSQP sTQ sTQP tfilt alpha T delta var, filt[27]
    assert g.delta_r == ref[0] and g.delta_t == ref[1] and g.R0 == ref[2] and g.RMax == ref[3]
    assert g.OffsetR == ref[6] and g.x0 == ref[8] and g.y0 == ref[9] and g.delta_xy == ref[11]
    assert [g.sig_s_q, g.sig_s_qp, g.sig_t_q, g.sig_t_qp] == list(ref[15:19])
    assert g.filt_t == ref[19] and g.alpha == ref[20]
    assert g.huber_T == ref[21] and g.huber_delta == ref[22] and g.var_est == ref[23]
    assert list(g.filt) == list(ref[24:51])
    angles = np.ascontiguousarray(scan["angles"], np.float32)
    H = np.zeros((scan["proj_num"], 65), np.float32)
    L.mbir_detector_response(C.byref(g), angles.ctypes.data_as(C.c_void_p), H.ctypes.data_as(C.c_void_p))
    assert np.array_equal(H, d["H"])


@needs_lib
def test_multires_schedule():
    from phasecon_amd.mbir import multires_schedule
    # XT_Engine.c:111-141 for one rank
    assert multires_schedule(8, 256) == [(4, 4), (2, 2), (1, 1)]
    assert multires_schedule(4, 128) == [(2, 2), (1, 1)]
    assert multires_schedule(32, 512) == [(8, 8), (4, 4), (2, 2), (1, 1)]
    assert multires_schedule(4, 64) == [(2, 2), (1, 1)]   # at least two stages
    assert multires_schedule(6, 256) == [(4, 2), (2, 2), (1, 1)]  # 6 rows halve once


@needs_lib
def test_view_map_rejects_decreasing_times():
    from phasecon_amd.mbir import view_map
    with pytest.raises(ValueError):
        view_map(np.array([0, 2, 1], np.float32), np.array([0, 3], np.float32))


@needs_lib
def test_view_map_examples():
    from phasecon_amd.mbir import view_map
    ex = os.path.join(GOLD, "examples")
    t = np.fromfile(os.path.join(ex, "ch_proj_times.bin"), np.float32)
    r = np.fromfile(os.path.join(ex, "ch_recon_times.bin"), np.float32)
    ptr, views = view_map(t, r)
    assert np.diff(ptr).tolist() == [32] * 32
    assert np.array_equal(views, np.arange(1024))


def test_interlaced_views_rule():
    """XT_Views.c:7-46 with K=4, N_theta=256: frames of 16 views cover pi/4
    and every angle index appears once per 4 half-turns."""
    from phasecon_amd.synth import interlaced_views
    a, t = interlaced_views(4, 256, 256)
    idx = np.rint(a / np.float32(np.pi / 256)).astype(int)
    assert idx[0] == 0 and idx[64] == 4 * 64 + 2 and idx[128] == 4 * 128 + 1 and idx[192] == 4 * 192 + 3
    assert np.all(np.diff(t) > 0)
    assert sorted(set(idx % 256 // 1)) == list(range(256))


@needs_lib
def test_cost_guard_actions():
    """The equit loops' cost-rise guard (mbir_guard_check): accepts
    decreases (asking for a save at each new lowest cost) and rises within
    the relative tolerance, asks for a recomputed error sinogram before it
    calls a larger rise real, recovers at most max_recoveries times, then
    fails as the reference does at any rise (XT_ICD_update.c:1171)."""
    from phasecon_amd import lib
    L = lib.load()
    g = lib.GuardState()
    L.mbir_guard_init(C.byref(g), 100.0, 1e-6, 2)
    chk = lambda c, exact=0: L.mbir_guard_check(C.byref(g), c, exact)
    assert chk(90.0) == lib.GUARD_SAVE and g.best == 90.0
    assert chk(90.0) == lib.GUARD_OK
    assert chk(90.0 * (1 + 0.5e-6)) == lib.GUARD_OK            # rounding noise
    assert chk(80.0, 1) == lib.GUARD_SAVE and g.best == 80.0
    assert chk(80.1) == lib.GUARD_CONFIRM                       # maintained-e rise
    assert chk(80.0 * (1 + 1e-7), 1) == lib.GUARD_OK            # gone after a refresh
    assert chk(85.0) == lib.GUARD_CONFIRM
    assert chk(85.0, 1) == lib.GUARD_RECOVER and g.recoveries == 1 and g.best == 80.0
    assert chk(float("nan")) == lib.GUARD_RECOVER               # NaN needs no confirmation
    assert chk(float("inf"), 1) == lib.GUARD_FAIL
    assert abs(L.mbir_guard_alpha(1.97) - 1.485) < 1e-12
    assert abs(L.mbir_guard_alpha(0.8) - 0.4) < 1e-12


DRIVER = os.path.join(ROOT, "oracle", "_ref", "xt_main_basic")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="oracle/_ref/xt_main_basic not built")
def test_reference_driver_links_to_this_library():
    """The reference's basic driver, compiled against include/mbir4d.h,
    takes reconstruct() from phasecon_amd/libmbir4d.so (the relink of the
    integration notes); the GPU test runs it."""
    import subprocess
    und = subprocess.run(["nm", "-D", "--undefined-only", DRIVER], capture_output=True,
                         text=True, check=True).stdout
    assert re.search(r"\breconstruct\b", und)
    ldd = subprocess.run(["ldd", DRIVER], capture_output=True, text=True, check=True).stdout
    assert "not found" not in ldd
    line = next(ln for ln in ldd.splitlines() if "libmbir4d.so" in ln)
    assert os.path.realpath(line.split("=>")[1].split("(")[0].strip()) == os.path.realpath(LIB)
