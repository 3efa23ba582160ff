"""The 4D benchmark workload at full size (256 x 256 x 8 voxels x 16 frames,
256 views interlaced K = 4, temporal p 1.1) through the benchmark's own
schedule and settings: 64x64x2 for 60 equits (tile 4, alpha 1.8),
128x128x4 for 80 equits (tile 6, alpha 1.97), then full resolution (tile
8 or 12, the two its autotuner picks between, alpha 1.97) until the FP64
MAP cost is within 1e-5 of C*, the FP64 oracle's converged cost
(tests/golden/cstar_ch4d.json, recorded with the scan's SHA-256), with the
benchmark's paired-parity launches.  Six runs over five seeds of the
visiting order, including the seed
offset at which alpha 1.97 on the coarsest stage diverged: every run must
reach the target with no recovery of the cost-rise guard (the guard checks
every 20 coarse equits and at every scheduled full-resolution check, so no
checked cost rose by more than 1e-6 relative within a stage)."""
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ch4d_stages():
    sys.path.insert(0, ROOT)
    import bench
    from phasecon_amd.mbir import Stage, multires_schedule
    from phasecon_amd.synth import make_scan
    scan = make_scan("ch4d", seed=0)
    rec = bench.load_cstar("ch4d")
    assert rec["digest"] == bench.scan_digest(scan) and rec["p_t"] == scan["p_t"]
    sched = multires_schedule(scan["proj_rows"], scan["proj_cols"])
    assert sched[:-1] == [(4, 4), (2, 2)]
    coarse = [(Stage(scan, mult_xy=4, mult_z=4, tile=4, alpha=1.8), 60),
              (Stage(scan, mult_xy=2, mult_z=2, tile=6, alpha=1.97), 80)]
    st = Stage(scan, tile=8, alpha=1.97)
    for s_ in [cs for cs, _ in coarse] + [st]:
        s_.option("pair_parities", 1)   # the benchmark's launches
    yield bench, coarse, st, rec["cstar"]
    for s_, _ in coarse:
        s_.close()
    st.close()


@pytest.mark.parametrize("full_tile,offset", [(8, 0), (8, 4), (8, 9), (12, 13), (12, 17),
                                              (12, 4)])
def test_benchmark_schedule_reaches_fp64_target(ch4d_stages, full_tile, offset):
    bench, coarse, st, cstar = ch4d_stages
    st.option("tile", full_tile)
    target = cstar * (1.0 + bench.REL_TOL)
    print(f"ch4d benchmark schedule, full-resolution tile {full_tile}, seed offset {offset}",
          flush=True)
    dt, n, cost = bench.run_once(st, target, 1500, 761521 + offset, 40, coarse)
    events = list(bench.GUARD_EVENTS)
    print(f"offset {offset}: {dt:.3f}s, {n} full-resolution equits, cost {cost:.10g} "
          f"(target {target:.10g}), guard recoveries {events}", flush=True)
    assert events == []
    assert cost <= target
    assert (cost - cstar) / cstar > -1e-6   # not below the FP64 optimum beyond rounding
    assert n < 200
