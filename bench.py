"""MBIR benchmark on one MI355X: seconds to the FP64-oracle MAP cost.

    python bench.py [--steps K] [--warmup W] [--workload ch4d] [--tile T]

A step is one full reconstruction of the workload from the plain (zero)
initialiser, until the MAP cost is within 1e-5 (relative) of C* -- the FP64
oracle's converged cost recorded in tests/golden/cstar_<workload>.json
(generated from the same seeded scan; its SHA-256 is checked) -- as confirmed
by a fresh FP64 cost evaluation.  With --schedule multires (the default) it
follows the reference's multi-resolution schedule (XT_Engine.c:111-141): the
coarse stages run a fixed number of equits each (--coarse-equits) from zero,
each stage's object is up-sampled on the device into the next with the
voxels outside the next stage's update mask zeroed (so the full-resolution
stage solves the same problem as a start from zero), and the full-resolution
stage runs to the target; --schedule single runs the full-resolution stage
alone from zero.  The stages over-relax their voxel updates by --alpha
(full resolution, default 1.97) and --coarse-alpha (per coarse stage,
default 1.8 for 64x64x2 and 1.97 for 128x128x4; the reference uses 1.5,
OVER_RELAXATION_FACTOR in XT_Constants.h): any factor in (0, 2) keeps the
same fixed point for exact sequential updates, but the parallel sweep's
concurrent updates diverge on the coarsest stage at 1.97 (measured on the
GPU and reproduced by the oracle's model of the sweep), where 1.8 converges
and was also the faster setting.  Every cost check goes through the
cost-rise guard (a confirmed rise restores the object of the lowest cost
and lowers alpha; recoveries are reported in the line).  Every stage's A
build and error sinogram are inside the timed region, as are the parallel
ICD equits and the cost checks (maintained error sinogram, FP32 prior
potentials summed in FP64).  Each equit's sweep launches -- 4 with the
paired-parity launches of every stage (--pair-parities, default 1), 8
colour-major -- are one replay of a captured HIP graph (mbir_hip_icd_graph;
--no-graph launches them directly); the coarse stages run 60 and 80 equits
(--coarse-equits) and each stage's tile comes from the autotuner
(--tile-candidates).  Inputs are resident in HBM before the timed region.
Prints one JSON line.

Learned warm start (skipped with --no-learned): training scans of the same
workload with seeds 1..--learned-scans are generated on the host before the
GPU is used; the GPU MBIR converges each (the schedule above plus 200
full-resolution equits) and a 3-level 3D U-Net (BF16 MFMA kernels) is trained
on (filtered back-projection of the frame and its mean over the frames,
converged image) patches.  Then the benchmark scan is reconstructed from the
network's estimate (A build, filtered back-projection, inference and, for
each --learned-coarse-equits count above 0, a two-grid correction of the
estimate's low frequencies on the 2x coarser grid, all timed) to the same
target, and once from zero at full resolution only (the plain initialiser).
Reported under "learned"; training and pair generation are timed apart.

Also reported: equits/s and the sweep kernel's algorithmic HBM bytes per
second with its roofline fraction (HIP events recorded by the library on its
own stream around every launch), and a CPU baseline: the reference's own ICD
(oracle/_ref, built from the reference sources) timed on this host's cores
over a bounded sample of equits of the same workload.

    python bench.py --longrun N --dump-x PATH   runs N equits and saves the
                                                object (C* warm start)
    python bench.py --profile-equits N          A build, error sinogram and N
                                                equits only, for counter runs
                                                under rocprofv3 (also runs the
                                                object upload, whose transpose
                                                kernel k_to_colmajor reads and
                                                writes 4 B per voxel: the
                                                traffic calibration copy)
    python bench.py --pcmbir                    PCMBIR phase retrieval, 360
                                                views x 512^2: s per SD
                                                iteration, FFT HBM rate,
                                                Nelder-Mead pixels/s
    python bench.py --pctomo [--steps K]        PCMBIR tomographic ICD, 256^3
                                                (delta, beta) x 360 views
                                                (tile 8): s per equit, sweep
                                                HBM rate
    python bench.py --pcmbir-admm [--steps H]   PCMBIR with its tomographic
                                                step, H head iterations end
                                                to end at that size
    python bench.py --aux                       ring offsets, variance,
                                                Paganin, phantom simulator
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sec to MAP cost ≤1e-5 rel of FP64 oracle; equits/s; HBM GB/s; learned vs plain"
PEAK_HBM_GBPS = 8000.0
REL_TOL = 1e-5

# HBM bytes per k_sweep launch (ch4d, full resolution, tile 8) measured with
# rocprofv3 on this code: FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (KiB per
# dispatch x 1024), the factors read off a cold 1 GiB device copy (both passes
# of profiles/k_sweep_ch4d_counters.md: bench.py --profile-equits 3
# --calibrate).  A counter run cannot happen inside the timed bench, so the
# line carries this figure with its source.
# Keyed by (workload, tile, FP16 A, paired parities): a paired launch holds
# two (colour, frame parity) sets.
MEASURED_TRAFFIC = {("ch4d", 8, False, False): 145609.5 * 1024 * 2.0 + 120815.1 * 1024 * 1.0,
                    ("ch4d", 8, False, True): 301410.4 * 1024 * 2.0 + 241630.7 * 1024 * 1.0,
                    ("ch4d", 12, False, True): 310021.1 * 1024 * 2.0 + 241499.1 * 1024 * 1.0}
TRAFFIC_SOURCE = "profiles/k_sweep_ch4d_counters.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"


def scan_digest(scan):
    h = hashlib.sha256()
    for k in ("projections", "weights", "angles", "proj_times", "recon_times"):
        h.update(np.ascontiguousarray(scan[k], np.float32).tobytes())
    return h.hexdigest()


def coarse_alphas(spec, n):
    """Per coarse stage over-relaxation from '--coarse-alpha': one value for
    all n stages, or one per stage (coarsest first); with n = 1 and a list,
    the finest coarse stage's."""
    vals = [float(v) for v in str(spec).split(",")]
    if len(vals) == 1:
        return vals * n
    if n == 1:
        return vals[-1:]
    if len(vals) != n:
        sys.exit(f"bench.py: --coarse-alpha needs 1 or {n} values")
    return vals


def load_cstar(workload, p_t=None):
    """The FP64 oracle's C* record of tests/golden/make_cstar.py for the
    workload (with a temporal p other than the workload's: its _pt file),
    or None."""
    name = f"cstar_{workload}.json" if p_t is None else f"cstar_{workload}_pt{p_t:g}.json"
    path = os.path.join(ROOT, "tests", "golden", name)
    return json.load(open(path)) if os.path.exists(path) else None


def algorithmic_bytes_per_equit(st):
    """Bytes one equit must move for the updated voxels: per (column, frame)
    visit the A window (first index + W values per view, the values in FP16
    plus one FP32 column scale when A is stored in FP16); per voxel the
    error sinogram read+write (8 B) and weight read (4 B) of each nonzero
    (view, bin, row) entry, plus the voxel's read and write (8 B)."""
    first, val = st.get_amatrix()
    g = st.g
    mask = st.mask.ravel().astype(bool)
    nnz = (val != 0).sum(axis=2)                       # [col][slot]
    total = 0.0
    for f in range(g.N_time):
        s0, s1 = int(st.frame_ptr[f]), int(st.frame_ptr[f + 1])
        per_col = nnz[:, s0:s1].sum(axis=1)[mask]       # nonzero entries per column
        nv = s1 - s0
        if st.afp16:
            total += mask.sum() * (nv * (2 * st.W + 4) + 4)
        else:
            total += mask.sum() * nv * (4 * st.W + 4)   # A window reads
        total += g.N_z * np.sum(12.0 * per_col * g.mult_z + 8.0)
    return total


GRAPH = True  # equits as replays of a captured HIP graph (--no-graph: direct launches)


NH = None  # (passes, fraction): non-homogeneous passes after each equit from the second
NH_STAGES = "full"  # stages that run them: "full", "coarse" or "all"


def _equit(st, it, seed, nh=False):
    nh = NH and (NH_STAGES == "all" or (NH_STAGES == "full") == nh)
    if nh and not getattr(st, "_nh_on", False):
        st.option("nhicd", 1)
        st._nh_on = True
    if GRAPH:
        st.icd_graph(1, it, seed)
    else:
        st.icd(1, it, seed)
    if nh and it > 1:
        st.nhicd(NH[0], NH[1], it, seed + 104729 * it)


PREDICT_CHECKS = True  # --check-every-equit turns the scheduled cost checks off
GUARD_RECOVER = 3      # mbir_guard_check's recovery code (phasecon_amd.lib.GUARD_RECOVER)


GUARD_EVENTS = []  # recoveries of the cost-rise guard in the current step


def to_target(st, target, max_equits, seed, refresh):
    """Full-resolution equits from the stage's current object until the
    FP64 MAP cost is <= target; returns (equits, cost).  Each cost check
    (a prior-cost pass and a host synchronisation) schedules the next one:
    from the decrease of C - C* since the last check, taken as geometric,
    the equits still needed are predicted and the next check comes after
    half of them (at least one equit later); with PREDICT_CHECKS off, after
    every equit.  Every check goes through the cost-rise guard (Guard): a
    confirmed rise restores the object of the lowest checked cost and halves
    alpha - 1 (recorded in GUARD_EVENTS; the time stays in the step)."""
    from phasecon_amd.mbir import Guard
    st.build_amatrix()
    st.init_error()
    cstar = target / (1.0 + REL_TOL)
    it = 0
    cost = st.cost(estimate=True)[0]
    guard = Guard(st, cost, max_recoveries=GUARD_RECOVERIES)
    prev = (0, cost - cstar)
    check_at = 1
    while it < max_equits:
        it += 1
        _equit(st, it, seed, nh=True)
        if refresh and it % refresh == 0:
            st.init_error()
        if it < check_at:
            continue
        cost = st.cost(estimate=True)[0]
        act, cost = guard.check(cost, where=f"{st.g.N_x}x{st.g.N_y}x{st.g.N_z} equit {it}")
        if guard.failed:
            break
        if act == GUARD_RECOVER:
            GUARD_EVENTS.extend(guard.events[-1:])
            prev = (it, cost - cstar)
            check_at = it + 1
            continue
        if cost <= target:
            fresh = st.cost(fresh=True)[0]   # FP64 check from a recomputed e
            if fresh <= target:
                cost = fresh
                break
        check_at = it + 1
        gap = cost - cstar
        if target <= 0:
            # no target (a reference run to max_equits): the guard's checks only
            check_at = it + COARSE_CHECK
        elif PREDICT_CHECKS and 0 < gap < prev[1]:
            rate = (gap / prev[1]) ** (1.0 / (it - prev[0]))
            need = math.log((target - cstar) / gap) / math.log(rate)
            check_at = it + max(1, int(0.5 * need))
        prev = (it, gap)
    return it, cost


def run_learned(st, net, target, max_equits, seed, refresh, source="fbp2", coarse=None,
                coarse_equits=0):
    """One timed reconstruction from the learned warm start: A, the
    network's input and estimate (with the two-grid correction when
    coarse_equits > 0), then full resolution to the target; returns
    (seconds, equits, cost, seconds to the start)."""
    from phasecon_amd.warmstart import learned_start
    GUARD_EVENTS.clear()
    for s_ in (st, coarse):
        if s_ is not None and s_.alpha != s_.alpha0:
            s_.option("alpha", s_.alpha0)
    st.synchronize()
    t0 = time.perf_counter()
    st.build_amatrix()
    learned_start(st, net, source, coarse, coarse_equits, seed)
    t_start = time.perf_counter() - t0
    it, cost = to_target(st, target, max_equits, seed, refresh)
    st.synchronize()
    return time.perf_counter() - t0, it, cost, t_start


GUARD_RECOVERIES = 3
COARSE_CHECK = 20  # equits between the cost-rise guard's checks in a coarse stage


def coarse_stage(cs, n, seed, refresh):
    """n equits of a coarse stage from its current object, the cost-rise
    guard checking every COARSE_CHECK equits and at the end (a recovery
    restores the object of the lowest checked cost with alpha - 1 halved
    and runs the equits since that check again)."""
    from phasecon_amd.mbir import Guard
    cs.build_amatrix()
    cs.init_error()
    guard = Guard(cs, cs.cost(estimate=True)[0], max_recoveries=GUARD_RECOVERIES)
    it, good = 0, 0
    while it < n:
        it += 1
        _equit(cs, it, seed)
        if refresh and it % refresh == 0:
            cs.init_error()
        if it % COARSE_CHECK == 0 or it == n:
            act, _ = guard.check(cs.cost(estimate=True)[0],
                                 where=f"{cs.g.N_x}x{cs.g.N_y}x{cs.g.N_z} equit {it}")
            if guard.failed:
                raise RuntimeError(f"coarse stage diverged: {guard.events}")
            if act == GUARD_RECOVER:
                GUARD_EVENTS.extend(guard.events[-1:])
                it = good
            else:
                good = it


def run_once(st, target, max_equits, seed, refresh, coarse=()):
    """One timed reconstruction from zero; returns (seconds, equits, cost).
    coarse: (stage, equits) pairs of the multi-resolution stages that run
    before st, coarsest first.  Every `refresh` equits the error sinogram is
    recomputed exactly (FP64 accumulation) so FP32 drift cannot bias the
    fixed point or the test.  Each stage starts at its configured alpha."""
    GUARD_EVENTS.clear()
    for s_ in [cs for cs, _ in coarse] + [st]:
        if s_.alpha != s_.alpha0:
            s_.option("alpha", s_.alpha0)
    first = coarse[0][0] if coarse else st
    first.set_object(np.zeros(first.shape_x, np.float32))
    first.synchronize()
    t0 = time.perf_counter()
    prev = None
    for cs, n in coarse:
        if prev is not None:
            cs.upsample_from(prev, masked=True)
        prev = cs
        if n == 0:
            continue
        coarse_stage(cs, n, seed, refresh)
    if prev is not None:
        st.upsample_from(prev, masked=True)
    it, cost = to_target(st, target, max_equits, seed, refresh)
    st.synchronize()
    return time.perf_counter() - t0, it, cost


def _make_train_scan(args):
    from phasecon_amd.synth import make_scan
    workload, seed = args
    return make_scan(workload, seed=seed)


def learned_leg(args, st, train_scans, target):
    """Config of the learned warm start: training pairs from the GPU MBIR on
    the training scans (other seeds of the workload), the U-Net trained on
    them, then timed reconstructions of the benchmark scan from its
    estimate (training and pair generation are reported, not timed)."""
    from phasecon_amd.warmstart import IN_CHANNELS, UNet, converged_pair, train
    log = lambda m: print("[bench] learned: " + m, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    src = args.learned_input
    pairs = [converged_pair(sc, alpha=args.alpha, source=src) for sc in train_scans]
    t_pairs = time.perf_counter() - t0
    log(f"{len(pairs)} training pairs in {t_pairs:.1f}s")
    net = UNet(channels=args.learned_channels, seed=0, in_channels=IN_CHANNELS[src])
    t0 = time.perf_counter()
    hist = train(net, pairs, steps=args.learned_steps, batch=8, size=64, lr=1e-3, seed=0, log=log)
    t_train = time.perf_counter() - t0
    from phasecon_amd.mbir import Stage, multires_schedule
    mxy, mz = multires_schedule(st.scan["proj_rows"], st.scan["proj_cols"])[-2]
    coarse = Stage(st.scan, mult_xy=mxy, mult_z=mz, tile=args.coarse_tile,
                   alpha=coarse_alphas(args.coarse_alpha, 1)[-1])
    variants = []
    for nc in [int(v) for v in str(args.learned_coarse_equits).split(",")]:
        run_learned(st, net, target, args.max_equits, args.seed, args.refresh, src, coarse, nc)
        times, equits, costs, starts = [], [], [], []
        for k in range(args.steps):
            dt, n, c, ts = run_learned(st, net, target, args.max_equits, args.seed + k,
                                       args.refresh, src, coarse, nc)
            times.append(dt)
            equits.append(n)
            costs.append(c)
            starts.append(ts)
            log(f"coarse {nc}: step {k}: {dt:.3f}s ({ts * 1e3:.1f} ms to the start), {n} "
                f"equits, cost {c:.10g}")
        variants.append({
            "coarse_equits": nc,
            "value": float(np.mean(times)) if all(c <= target for c in costs) else None,
            "equits_to_target": equits, "ms_to_start": float(np.mean(starts)) * 1e3})
    coarse.close()
    net.close()
    # the plain initialiser the network replaces: zero, full resolution only
    st.set_object(np.zeros(st.shape_x, np.float32))
    st.synchronize()
    t0 = time.perf_counter()
    n0, c0 = to_target(st, target, args.max_equits, args.seed, args.refresh)
    st.synchronize()
    zero = {"value": time.perf_counter() - t0 if c0 <= target else None, "equits_to_target": n0}
    log(f"zero start: {zero['value']}s, {n0} equits")
    best = min((v for v in variants if v["value"]), key=lambda v: v["value"], default=variants[0])
    return {
        "value": best["value"],
        "unit": "s",
        "equits_to_target": best["equits_to_target"],
        "ms_to_start": best["ms_to_start"],
        "coarse_equits": best["coarse_equits"],
        "variants": variants,
        "zero_start": zero,
        "network": f"3-level 3D U-Net, channels {args.learned_channels}/"
                   f"{2 * args.learned_channels}/{4 * args.learned_channels}, BF16 MFMA",
        "training": {"scans": len(train_scans), "seeds": list(range(1, len(train_scans) + 1)),
                     "pairs_s": t_pairs, "steps": args.learned_steps, "batch": 8,
                     "patch": "full depth x 64 x 64", "train_s": t_train,
                     "loss_first": float(hist[:20].mean()), "loss_last": float(hist[-20:].mean())},
        "input": {"fbp3": "filtered back-projection of the frame, its mean over the frame and "
                          "its neighbours, and its mean over the frames",
                  "fbp2": "filtered back-projection of the frame and its mean over the frames",
                  "fbp": "filtered back-projection of the frame",
                  "bp": "normalised back-projection A^T(w y) / A^T(w A m)"}[src],
        "start": ("U-Net estimate; with coarse_equits > 0 its low frequencies corrected by "
                  "that many equits on the 2x coarser grid from its block mean (two-grid), "
                  "then full resolution to the target"),
    }


def profile_unet(n, c=16, in_ch=2, N=16, D=8, H=256, W=256):
    """n forward passes of the network on a random volume of the benchmark's
    shape (frames as the batch); prints the mean time and TFLOP/s."""
    from phasecon_amd.warmstart import UNet
    net = UNet(channels=c, seed=0, in_channels=in_ch)
    x = np.random.default_rng(0).random((N, D, H, W, in_ch)).astype(np.float32)
    net.forward(x)  # allocates
    t0 = time.perf_counter()
    for _ in range(n):
        net.forward(x)
    dt = (time.perf_counter() - t0) / n
    net.close()
    shapes = [(in_ch, c), (c, c), (c, 2 * c), (2 * c, 2 * c), (2 * c, 4 * c), (4 * c, 4 * c),
              (6 * c, 2 * c), (2 * c, 2 * c), (3 * c, c), (c, c), (c, 1)]
    lev = [0, 0, 1, 1, 2, 2, 1, 1, 0, 0, 0]
    V = N * D * H * W
    macs = sum(27 * ci * co * V // 8 ** l for (ci, co), l in zip(shapes, lev))
    print(json.dumps({"profile_unet": n, "shape": [N, D, H, W, in_ch], "channels": c,
                      "ms_per_forward": dt * 1e3, "gmac": macs / 1e9,
                      "tflops_incl_transfers": 2 * macs / dt / 1e12}))


def config4(args):
    """The mixed-precision 4D workload (512 x 512 x 32 voxels x 24 frames, 720
    views interlaced K = 8, 28-32 views per frame): seconds to (1 + 1e-5) C*
    with FP32 A and with FP16 A + per-column FP32 scale, every stage's equits
    as graph replays and tiles from the autotuner; C* and the reference image
    from a long FP32-A run (FP64 projector and cost), and the FP16 run's
    converged cost and image error against it.  Prints one JSON line."""
    from phasecon_amd.autotune import tune_tile
    from phasecon_amd.lib import K_SWEEP
    from phasecon_amd.mbir import Stage, multires_schedule
    from phasecon_amd.synth import make_cfg4_scan
    log = lambda m: print("[bench] config4: " + m, file=sys.stderr, flush=True)
    t0 = time.time()
    scan = make_cfg4_scan(0)
    log(f"scan generated in {time.time() - t0:.1f}s, views per frame "
        f"{int(scan['views_per_frame'].min())}-{int(scan['views_per_frame'].max())}")
    sched = multires_schedule(scan["proj_rows"], scan["proj_cols"])
    counts = [int(v) for v in args.config4_coarse.split(",")]
    out = {"metric": METRIC, "unit": "s", "n_gpus": 1, "higher_is_better": False,
           "dtype": "fp32", "data": "synthetic", "steps": args.steps,
           "config": {"workload": "cfg4: 512x512x32 voxels x 24 frames, 720 views interlaced "
                                  "K=8, 28-32 views per frame (seeded), I0=5000, drifting "
                                  "ellipsoids projected by the GPU projector",
                      "schedule": " -> ".join(f"{mx}:{mz}" for mx, mz in sched),
                      "coarse_equits": counts, "alpha": args.alpha}}
    results = {}
    ref_x = None
    for afp16 in (False, True):
        coarse = [(Stage(scan, mult_xy=mx, mult_z=mz, tile=8, alpha=a, afp16=afp16), n)
                  for (mx, mz), n, a in zip(sched[:-1], counts,
                                            coarse_alphas(args.config4_coarse_alpha,
                                                          len(sched) - 1))]
        st = Stage(scan, tile=8, alpha=args.alpha, afp16=afp16)
        tiles = {}
        for s_ in [cs for cs, _ in coarse] + [st]:
            s_.set_object(np.zeros(s_.shape_x, np.float32))
            s_.build_amatrix()
            # tile 4 is left out: with 30 views per frame and 2-8 detector rows
            # per voxel, that many concurrent columns make the coarse stages
            # diverge at alpha >= 1.5 (measured; tiles 8 and up converge)
            tiles[f"{s_.g.N_x}x{s_.g.N_y}x{s_.g.N_z}"] = tune_tile(s_, candidates=(8, 16))[0]
        if ref_x is None:
            # C*: the multi-resolution schedule, then a long full-resolution run
            run_once(st, -1.0, args.config4_ref_equits, args.seed, args.refresh, coarse)
            cstar = st.cost(fresh=True)[0]
            ref_x = st.get_object().astype(np.float64)
            log(f"C* = {cstar:.12g} after {args.config4_ref_equits} full-resolution equits")
            if not np.isfinite(cstar) or cstar > st.cost(fresh=True)[0] * 1.000001 or \
                    not np.all(np.isfinite(ref_x)) or np.abs(ref_x).max() > 1.0:
                raise SystemExit("bench.py: the config-4 reference run diverged")
            target = cstar * (1.0 + REL_TOL)
        st.reset_kernel_time()
        times, equits = [], []
        for k in range(args.steps):
            dt, n, c = run_once(st, target, args.max_equits, args.seed + k, args.refresh, coarse)
            times.append(dt)
            equits.append(n)
            log(f"{'fp16' if afp16 else 'fp32'} A step {k}: {dt:.3f}s, {n} equits, cost {c:.12g}")
        sweep_ms, sweep_n = st.kernel_time(K_SWEEP)
        bytes_launch = algorithmic_bytes_per_equit(st) * sum(equits) / max(sweep_n, 1)
        rec = {"value": float(np.mean(times)), "equits_to_target": equits, "tiles": tiles,
               "sweep_ms_per_launch": sweep_ms / max(sweep_n, 1),
               "sweep_gbps": bytes_launch / (sweep_ms / max(sweep_n, 1) * 1e-3) / 1e9}
        if afp16:
            # converged FP16 run against the FP32-A reference
            run_once(st, -1.0, args.config4_ref_equits, args.seed, args.refresh, coarse)
            c16 = st.cost(fresh=True)[0]
            x16 = st.get_object().astype(np.float64)
            rec["error_vs_fp32_a"] = {
                "cost_rel": (c16 - cstar) / cstar,
                "image_rmse_rel": float(np.sqrt(np.mean((x16 - ref_x) ** 2) /
                                                np.mean(ref_x ** 2)))}
        results["fp16_a" if afp16 else "fp32_a"] = rec
        for cs, _ in coarse:
            cs.close()
        st.close()
    out.update(results)
    out["cstar"] = cstar
    out["cstar_source"] = (f"FP32-A GPU run, {args.config4_ref_equits} full-resolution equits "
                           "after the coarse stages (FP64 projector and cost)")
    out["value"] = results["fp16_a"]["value"]
    out["vs_fp32_a"] = results["fp32_a"]["value"] / results["fp16_a"]["value"]
    out["roofline"] = {"bound": "hbm", "achieved": results["fp16_a"]["sweep_gbps"],
                       "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                       "frac": results["fp16_a"]["sweep_gbps"] / PEAK_HBM_GBPS,
                       "traffic": None, "kernel": "k_sweep<3, true, true>"}
    print(json.dumps(out))


def _up2(a, axis):
    """2x along axis by the reference's rule for the next stage's initial
    object (XT_ICD_update.c:358-390): ends copied, interior samples 3/4 and
    1/4 of the two nearest coarse samples."""
    a = np.moveaxis(a.astype(np.float32), axis, -1)
    n = a.shape[-1]
    out = np.empty(a.shape[:-1] + (2 * n,), np.float32)
    out[..., 0], out[..., -1] = a[..., 0], a[..., -1]
    out[..., 1] = (3 * a[..., 0] + a[..., 1]) / 4
    out[..., -2] = (a[..., -2] + 3 * a[..., -1]) / 4
    out[..., 2:-2:2] = (a[..., :-2] + 3 * a[..., 1:-1]) / 4
    out[..., 3:-2:2] = (3 * a[..., 1:-1] + a[..., 2:]) / 4
    return np.moveaxis(out, -1, axis)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(scan, budget_s=120.0, coarse_cap=100, convg=1.0):
    """The reference's own ICD (oracle/_ref, its sources compiled in place)
    on this host's cores, timed to its own target: its multi-resolution
    schedule (XT_Engine.c:111-141: each coarse stage from the previous one's
    up-sampled object until its convergence test at convg_thresh = convg %,
    at most coarse_cap equits), then full resolution until the FP64 MAP cost
    is <= (1 + 1e-5) C*_ref, or budget_s of reference time.  The reference
    has one qGGMRF p (MRF_P = 1.2, Q6), so it solves the workload with p = 1.2
    for both priors and C*_ref is the FP64 oracle's converged cost of that
    problem (tests/golden/cstar_<workload>_pt1.2.json).  Voxels outside each
    stage's update mask are zeroed after up-sampling, as on the GPU (the
    reference leaves them at their interpolated values and so converges to a
    slightly different cost).  Time = stage set-up (initStructures, error
    sinogram, up-sampling) + the equits; the reference's per-equit cost
    evaluations, which its shipped build does not run (NO_COST_CALCULATE,
    MBIR_4D/Makefile:3), are excluded."""
    try:
        from oracle.oracle import Oracle, Reference, reference_available
    except Exception as exc:  # pragma: no cover
        return {"value": None, "unit": "s", "cores": 0, "kind": "reference",
                "sample": f"unavailable: {exc}"}
    if not reference_available():
        return {"value": None, "unit": "s", "cores": 0, "kind": "reference",
                "sample": "oracle/_ref not built"}
    from phasecon_amd.mbir import multires_schedule
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    s = dict(scan)
    s["p_t"] = s["p_s"]
    rec = load_cstar(s["name"], s["p_t"])
    if rec is None or rec.get("digest") != scan_digest(scan):
        return {"value": None, "unit": "s", "cores": threads, "kind": "reference",
                "sample": "no FP64 C* for the reference's p = 1.2 problem"}
    cstar_ref = rec["cstar"]
    target = cstar_ref * (1.0 + REL_TOL)
    stages = multires_schedule(s["proj_rows"], s["proj_cols"])
    spent, x, per_stage, reached, trace = 0.0, None, [], False, []
    own_stop = None
    for k, (mxy, mz) in enumerate(stages):
        last = k == len(stages) - 1
        t0 = time.perf_counter()
        ref = Reference(s, mult_xy=mxy, mult_z=mz, convg=convg, threads=threads)
        if x is None:
            x = np.zeros(ref.shape_x, np.float32)
        else:
            x = _up2(_up2(x, 3), 2)
            if x.shape[1] != ref.shape_x[1]:
                x = _up2(x, 1)
            x[..., ~Oracle(s, mult_xy=mxy, mult_z=mz).mask.astype(bool)] = 0.0
        ref.set_object(x)
        ref.init_error()
        spent += time.perf_counter() - t0
        n = 0
        while spent < budget_s:
            costs, conv, secs = ref.icd_run(1, first_iter=n + 1)
            n += 1
            spent += float(secs[0])
            print(f"[bench] cpu reference {ref.N_x}x{ref.N_y}x{ref.N_z} equit {n}: "
                  f"{secs[0]:.2f}s, cost {costs[0]:.10g}, {spent:.1f}s so far",
                  file=sys.stderr, flush=True)
            if last:
                trace.append((round(spent, 3), float(costs[0])))
                if own_stop is None and conv[0] and n > 1:
                    # where the reference's own stopping rule ends its run
                    own_stop = {"s": round(spent, 3), "full_res_equits": n,
                                "cost_rel_cstar": float(costs[0]) / cstar_ref - 1.0}
                if costs[0] <= target:
                    reached = True
                    break
            elif (conv[0] and n > 1) or n >= coarse_cap:
                break
        per_stage.append({"grid": f"{ref.N_x}x{ref.N_y}x{ref.N_z}", "equits": n})
        x = ref.get_object()
        ref.close()
        if spent >= budget_s:
            break
    out = {"value": spent if reached else None, "unit": "s", "cores": threads,
           "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(), "kind": "reference",
           "reached_target": reached, "stages": per_stage, "budget_s": budget_s,
           "cstar_ref": rec["cstar"],
           "sample": ("the reference's ICD (OpenMP, its HICD + NHICD schedule, alpha 1.5) "
                      "from zero through its multi-resolution stages to (1 + 1e-5) C* of the "
                      "same scan with p = 1.2 for both priors (the reference has one p), "
                      f"within a budget of {budget_s:g} s")}
    if own_stop is not None:
        # the same run where the reference's own convergence test (average
        # update below convg % at full resolution) would have ended it: the
        # stopping rule of the drop-in reconstruct()
        out["own_stop"] = own_stop
    if not reached and len(trace) >= 3:
        # geometric extrapolation of the last full-resolution equits (labelled)
        (t1, c1), (t2, c2) = trace[-3], trace[-1]
        g1, g2 = c1 - rec["cstar"], c2 - rec["cstar"]
        if 0 < g2 < g1:
            rate = (g2 / g1) ** (1.0 / 2)
            need = math.log((target - rec["cstar"]) / g2) / math.log(rate)
            out["s_to_target_est"] = spent + need * (t2 - t1) / 2
            out["note"] = ("budget spent before the target: s_to_target_est extrapolates the "
                           "last equits' geometric cost decrease")
            if scan.get("name") == "ch4d":
                out["full_run"] = {"value": 178.96, "unit": "s", "cores": 16,
                                   "source": "profiles/cpu_reference_ch4d.md (bench.py "
                                             "--cpu-only --cpu-budget 600, the same path "
                                             "run to its target)"}
    return out


def calibration_copy(nbytes=1 << 30):
    """A plain device-to-device copy of nbytes (the HIP runtime's copy
    kernel) whose source was evicted from L2 and the Infinity Cache first (a
    second buffer of the same size is filled after it): the reference reading
    for FETCH_SIZE / WRITE_SIZE counter scaling.  Returns the bytes read
    (= bytes written)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    bufs = [C.c_void_p() for _ in range(3)]
    for b in bufs:
        if hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    src, dst, other = bufs
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipMemset(other, 2, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    for b in bufs:
        hip.hipFree(b)
    if rc != 0:
        raise RuntimeError("hipMemcpy failed")
    return nbytes


def pcmbir_bench(views, n, sd_steps, nms, fft_reg=1):
    """Phase-retrieval throughput on one GPU at the PCMBIR workload's shape
    (views x n x n transforms, n = 512 the padded 256^2 detector): seconds
    per steepest-descent iteration over all views (6 Fresnel transforms of
    3 passes each per view), the FFT passes' HBM rate against 8 TB/s, and one
    Nelder-Mead pass (pixels per second).  Synthetic data: a weak random phase
    object propagated by the same operator."""
    from phasecon_amd.pcmbir import PhaseRetrieval, wavelength, fresnel_window
    lam, dist = wavelength(17.0), 1e5
    win = fresnel_window(n, n, 1.0, 1.0, 2.0 / np.sqrt(lam * dist))
    rng = np.random.default_rng(0)
    shape = (views, n, n)
    t0 = time.time()
    ctx = PhaseRetrieval(views, n, n, 1.0, 1.0, lam, dist, win)
    ctx.option("fft_reg", fft_reg)
    obj = np.exp(-0.01 * rng.random(shape) - 0.2j * rng.random(shape))
    y = np.abs(ctx.fresnel(obj)).astype(np.complex64)
    D = np.ones(shape, np.complex64)
    lamw = np.ones(shape, np.float32)
    w = np.ones(shape, np.complex64)
    om = np.ones(shape, np.complex64)
    ctx.set_problem(y, D, lamw, w, w, om, 0.5)
    print(f"[bench] pcmbir setup {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    ctx.sd_run(1, 0.0)            # warm-up
    ctx.fft_time()
    t0 = time.perf_counter()
    ctx.sd_run(sd_steps, 0.0)
    dt = (time.perf_counter() - t0) / sd_steps
    fft_ms, launches = ctx.fft_time()
    per_transform_bytes = 3 * 2 * n * n * 8          # 3 passes, read + write, complex64
    transforms = 6 * views * sd_steps                # 5 batched calls, one over 2 x views
    gbps = per_transform_bytes * transforms / (fft_ms * 1e-3) / 1e9
    out = {"pcmbir": {"views": views, "n": n, "fft": "register" if fft_reg else "lds-radix2",
                      "s_per_sd_iter": dt,
                      "fft_ms_per_sd_iter": fft_ms / sd_steps,
                      "fft_hbm_gbps": gbps, "fft_frac": gbps / PEAK_HBM_GBPS}}
    ctx.close()
    if nms:
        nv = min(views, 8)  # the per-pixel FP64 search is timed on a few views
        sub = (nv, n, n)
        cn = PhaseRetrieval(nv, n, n, 1.0, 1.0, lam, dist, win)
        cn.set_nms_prior(2e-2, 1e-2, 2e-2, 1e-2, 50.0, True, 1.2, 0.5)
        zr = rng.random(sub) * 0.05 + 0.02
        zi = rng.random(sub) * 0.5 + 0.2
        cn.nms_pass(zr, zi, obj[:nv])  # warm-up (first launches load the code object)
        passes = []
        for _ in range(3):  # the same pass three times, the median reported
            t0 = time.perf_counter()
            _, _, its = cn.nms_pass(zr, zi, obj[:nv])
            passes.append(time.perf_counter() - t0)
        dt_nms = float(np.median(passes))
        out["pcmbir"].update({"nms_views": nv, "nms_pass_s": dt_nms,
                              "nms_pixels_per_s": nv * n * n / dt_nms,
                              "nms_avg_iters": its / (nv * n * n)})
        cn.close()
    print(json.dumps(out))


def pctomo_bench(n, views, equits, tile, seed=761521):
    """The PCMBIR tomographic ICD (k_pc_sweep) on one GPU at the PCMBIR
    workload's object size: an n^3 (delta, beta) object from seeded
    ellipsoids (delta up to 1e-6, beta = delta / 1000), views parallel-beam
    views over 180 degrees on an n x n detector, wavenumber of 17 keV per
    1 um voxel.  The targets are k A x of that object (on the GPU); the
    equits run from zero.  Reports seconds per equit, the sweep's launches
    (library events) and its algorithmic HBM bytes against 8 TB/s."""
    from phasecon_amd.synth import geometry_scan, drifting_ellipsoids
    from phasecon_amd.mbir import TomoStage, tomo_prior
    from phasecon_amd.lib import K_SWEEP
    t0 = time.time()
    scan = geometry_scan(n, n, views, seed=0)
    dob = 1000.0
    kw = 2.0 * np.pi * 1e-6 / (1.23984e-9 / 17.0)   # 2 pi / lambda, per 1 um voxel
    prior = tomo_prior((2e-7, 1e-3, 0.01, 0.01), (2e-10, 1e-3, 0.01, 0.01), 1,
                       float(views), kw, delta_over_beta=dob, alpha=1.5)
    st = TomoStage(scan, prior, tile=tile)
    xd = drifting_ellipsoids(n, n, 1, seed=3, lo=2e-7, hi=1e-6, count=20)
    xb = xd / dob
    zero = np.zeros(st.shape_e, np.float32)
    st.set_objects(xd, xb)
    st.init_errors(zero, zero)
    ed, eb = st.get_errors()
    td, tb = -ed, -eb                                 # k A x of the object
    st.set_objects(np.zeros(st.shape_x, np.float32), np.zeros(st.shape_x, np.float32))
    st.init_errors(td, tb)
    c0 = st.tomo_cost()[0]
    print(f"[bench] pctomo setup {time.time() - t0:.1f}s (tile {tile})", file=sys.stderr,
          flush=True)
    st.equits(1, 1, seed)                              # warm-up
    st.synchronize()
    st.reset_kernel_time()
    t0 = time.perf_counter()
    st.equits(equits, 2, seed)
    st.synchronize()
    dt = (time.perf_counter() - t0) / equits
    ms, launches = st.kernel_time(K_SWEEP)
    c1 = st.tomo_cost()[0]
    # algorithmic bytes of one equit: per column visit the A window (first
    # index + W values per view); per voxel both error sinograms' read +
    # write (16 B) on each nonzero (view, bin, row) entry and both objects'
    # read + write (16 B)
    first, val = st.get_amatrix()
    nnz = int((val != 0).sum())
    ncol = int(st.mask.sum())
    g = st.g
    bytes_equit = ncol * int(st.frame_ptr[-1]) * (4 * st.W + 4) + \
        g.N_z * (16.0 * nnz * g.mult_z + 16.0 * ncol)
    bytes_launch = bytes_equit * equits / launches
    gbps = bytes_launch / (ms / launches * 1e-3) / 1e9
    print(json.dumps({"pctomo": {
        "object": f"{n}^3 (delta, beta)", "views": views, "tile": tile, "equits": equits,
        "s_per_equit": dt, "sweep_launches": launches, "sweep_avg_ms": ms / launches,
        "bytes_per_launch": bytes_launch, "hbm_gbps": gbps, "frac": gbps / PEAK_HBM_GBPS,
        "cost_start": c0, "cost_end": c1, "data": "synthetic (seeded ellipsoids, targets k A x)"}}))
    st.close()


def pcmbir_admm_bench(n, views, heads, tile, n_icd=2, pret_iters=2, sd_iters=3,
                      seed=761521):
    """The PCMBIR reconstruction with its tomographic step, end to end on one
    GPU at the PCMBIR workload's size: an n^3 (delta, beta) object from
    seeded ellipsoids (delta up to 1e-6, beta = delta / 1000, 17 keV, 1 um
    voxels, 0.1 m propagation), views parallel-beam views of n x n, the
    intensity measurements |H exp(-z)| of its projections z = k A x
    (simulated on the GPU, D = 1, unit weights).  From zero objects and the
    Paganin estimate of z it runs `heads` head iterations (n_icd ICD equits, the coupled
    phase retrieval of every view with pret_iters Nelder-Mead passes and
    sd_iters steepest-descent iterations each, the dual update) and reports
    seconds per head iteration and the part of each stage."""
    from phasecon_amd.synth import geometry_scan, drifting_ellipsoids
    from phasecon_amd.mbir import TomoStage, tomo_prior
    from phasecon_amd.pcmbir import PhaseRetrieval, wavelength, fresnel_window
    from phasecon_amd.pctomo import TomoADMM
    t0 = time.time()
    scan = geometry_scan(n, n, views, seed=0)
    dob, lam, dist = 1000.0, wavelength(17.0), 1e5
    kw = 2.0 * np.pi / lam
    prior = tomo_prior((2e-7, 1e-3, 0.01, 0.01), (2e-10, 1e-3, 0.01, 0.01), 1,
                       float(views), kw, delta_over_beta=dob, mu=0.5, alpha=1.5)
    # projections z = k A x of the object, by the tomographic stage
    st = TomoStage(scan, prior, tile=tile)
    xd = drifting_ellipsoids(n, n, 1, seed=3, lo=2e-7, hi=1e-6, count=20)
    zero = np.zeros(st.shape_e, np.float32)
    st.set_objects(xd, xd / dob)
    st.init_errors(zero, zero)
    ed, eb = st.get_errors()
    st.close()
    n_p, n_r, n_t = ed.shape
    win = fresnel_window(n_r, n_t, 1.0, 1.0, 2.0 / np.sqrt(lam * dist))
    sim = PhaseRetrieval(n_p, n_r, n_t, 1.0, 1.0, lam, dist, win)
    zb, zd = -eb.astype(np.float64), -ed.astype(np.float64)     # magnitude, phase
    meas = np.abs(sim.fresnel(np.exp(-zb) * (np.cos(-zd) + 1j * np.sin(-zd)))).astype(np.complex64)
    sim.close()
    del ed, eb, zb, zd
    nms = dict(phase_sig=2e-1, phase_c=1e-2, mag_sig=2e-4, mag_c=1e-2, delta_over_beta=dob,
               decorrelate=True)
    ones = np.ones(meas.shape, np.complex64)
    run = TomoADMM(scan, prior, meas, ones, 1.0, (1.0, 1.0, lam, dist, win), 0.5, nms,
                   tile=tile)
    # the reference's start: the Paganin estimate of z, zero objects (the
    # simplex search cannot leave z = 0: its start simplex (z, 2 z) collapses)
    z_mag, z_phase = run.pr.paganin(meas, ones, dob, kw)
    zeros = np.zeros(run.st.shape_x, np.float32)
    run.start(zeros, zeros, z_phase, z_mag)
    print(f"[bench] pcmbir admm setup {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    rows = []
    for h in range(heads):
        t_h = time.perf_counter()
        c0 = run.st.tomo_cost()[0]
        t_i = time.perf_counter()
        for i in range(n_icd):
            run.st.equits(1, i + 1, seed + run.equits)
            run.equits += 1
        run.st.synchronize()
        t_icd = time.perf_counter() - t_i
        c1 = run.st.tomo_cost()[0]
        t_r = time.perf_counter()
        ad, ab = run.st.admm_target()
        run.pr.set_tomo_target(prior.mu, ab + 1j * ad)
        out = run.pr.estimate(run.zb, run.zd, pret_iters, 0.0, sd_iters, 0.0)
        run.zb, run.zd = out["zr"], out["zi"]
        t_ret = time.perf_counter() - t_r
        t_u = time.perf_counter()
        res = run.st.admm_update(run.zd, run.zb)
        t_upd = time.perf_counter() - t_u
        dt = time.perf_counter() - t_h
        rows.append(dict(s=dt, icd_s=t_icd, retrieval_s=t_ret, dual_s=t_upd,
                         icd_cost=[c0, c1], residuals=[float(v) for v in np.ravel(res)]))
        print(f"[bench] head {h}: {dt:.3f}s (ICD {t_icd:.3f}, retrieval {t_ret:.3f}, "
              f"dual {t_upd:.3f})", file=sys.stderr, flush=True)
    run.close()
    print(json.dumps({"pcmbir_admm": {
        "object": f"{n}^3 (delta, beta)", "views": views, "view": f"{n_r}x{n_t}", "tile": tile,
        "icd_equits": n_icd, "pret_iters": pret_iters, "sd_iters": sd_iters,
        "s_per_head": float(np.mean([r["s"] for r in rows[1:] or rows])), "heads": rows,
        "data": "synthetic (seeded ellipsoids; measurements |H exp(-k A x)| on the GPU)"}}))


def aux_bench():
    """Timings of the stages around the sweeps, each after one untimed call:
    the ring-offset update (modes 1-3) and the noise-variance re-estimate on
    the ch4d workload (once per equit when enabled; device only), the
    Paganin start over 360 views of 512^2 and the phase-contrast phantom
    simulator over a 256^3 object and 360 views of 256^2, and the APS
    flat-field normalisation of 720 raw frames of 64 x 1024 (once per scan;
    host arrays in and out, so the PCIe copies are included); the drop-in
    reconstruct() on ch4d with the reference's schedule and stopping rule."""
    from phasecon_amd.synth import make_scan, drifting_ellipsoids
    from phasecon_amd.mbir import Stage, phantom_project
    from phasecon_amd.pcmbir import PhaseRetrieval, wavelength, fresnel_window
    out = {}
    st = Stage(make_scan("ch4d", seed=0), tile=12)
    st.build_amatrix()
    st.init_error()
    for mode in (1, 2, 3):
        st.offset_update(mode)
        st.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            st.offset_update(mode)
        st.synchronize()
        out[f"ring_offset_mode{mode}_ms"] = (time.perf_counter() - t0) / 10 * 1e3
    st.variance_update()
    t0 = time.perf_counter()
    for _ in range(10):
        st.variance_update()
    st.synchronize()
    out["variance_ms"] = (time.perf_counter() - t0) / 10 * 1e3
    st.close()
    lam, dist, n, nv = wavelength(17.0), 1e5, 512, 360
    win = fresnel_window(n, n, 1.0, 1.0, 2.0 / np.sqrt(lam * dist))
    pr = PhaseRetrieval(nv, n, n, 1.0, 1.0, lam, dist, win)
    y = (1.0 + 0.01 * np.random.default_rng(0).random((nv, n, n))).astype(np.complex64)
    D = np.ones((nv, n, n), np.complex64)
    pr.paganin(y, D, 1000.0, 2 * np.pi / lam)
    t0 = time.perf_counter()
    pr.paganin(y, D, 1000.0, 2 * np.pi / lam)
    out["paganin_360x512sq_s"] = time.perf_counter() - t0
    pr.close()
    obj = drifting_ellipsoids(256, 256, 1, seed=3, lo=2e-7, hi=1e-6, count=20)[0]
    ang = (np.arange(360) * np.pi / 360).astype(np.float32)
    args = (obj / 1000, obj, ang, 256, 256, 1.0, 128.0, 2 * np.pi / lam)
    phantom_project(*args)
    t0 = time.perf_counter()
    phantom_project(*args)
    out["phantom_project_256cubed_x360_s"] = time.perf_counter() - t0
    # APS input stage at the mixed-precision workload's detector: 720 raw
    # uint16 frames of 64 x 1024 (down-sampled by 2 to 32 x 512), 20 white
    # and 20 dark frames
    from phasecon_amd.mbir import aps_normalize
    rng = np.random.default_rng(1)
    raw = rng.integers(1000, 9000, (720, 64, 1024), dtype=np.uint16)
    white = rng.integers(9000, 10000, (20, 64, 1024), dtype=np.uint16)
    dark = rng.integers(0, 100, (20, 64, 1024), dtype=np.uint16)
    aps_normalize(raw, white, dark, 2)
    t0 = time.perf_counter()
    aps_normalize(raw, white, dark, 2)
    out["aps_normalize_720x64x1024_s"] = time.perf_counter() - t0
    # the drop-in reconstruct() on ch4d with the reference's own schedule and
    # stopping rule (average update below 1 % per stage), host arrays in and
    # out, against the FP64 oracle's C* of the problem it solves (the
    # reference has one p: 1.2 for both priors)
    from phasecon_amd.mbir import reconstruct
    scan = make_scan("ch4d", seed=0)
    t0 = time.perf_counter()
    _, eq, cost = reconstruct(scan, convg_thresh=1.0)
    out["reconstruct_ch4d_s"] = time.perf_counter() - t0
    out["reconstruct_ch4d_last_stage_equits"] = eq
    rec = load_cstar("ch4d", 1.2)
    if rec is not None:
        out["reconstruct_ch4d_cost_rel_cstar_p1.2"] = cost / rec["cstar"] - 1.0
    print(json.dumps({"aux": out}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ch4d")
    ap.add_argument("--tile", type=int, default=8)
    ap.add_argument("--schedule", choices=("multires", "single"), default="multires")
    ap.add_argument("--coarse-equits", default="60,80",
                    help="equits of the coarse stages of the multi-resolution schedule, "
                         "coarsest first: one count for every stage or a comma list")
    ap.add_argument("--coarse-tile", type=int, default=4)
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every equit's sweeps directly instead of replaying a "
                         "captured HIP graph")
    ap.add_argument("--nh", default="",
                    help="non-homogeneous ICD after every full-resolution equit from the second: "
                         "'passes,fraction' (e.g. '1,0.25'); default off")
    ap.add_argument("--nh-stages", choices=("full", "coarse", "all"), default="full")
    ap.add_argument("--check-every-equit", action="store_true",
                    help="evaluate the stopping cost after every full-resolution equit")
    ap.add_argument("--tile-candidates", default="4,6,8,12,16",
                    help="tiles the autotuner times per stage")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep --tile / --coarse-tile instead of timing candidates per stage")
    ap.add_argument("--stages", default="",
                    help="explicit coarse stages, coarsest first: mult_xy:mult_z:equits,...")
    ap.add_argument("--max-equits", type=int, default=3000)
    ap.add_argument("--seed", type=int, default=761521)
    ap.add_argument("--refresh", type=int, default=40,
                    help="equits between exact error-sinogram refreshes (10, 20, 40 measured "
                         "on ch4d: 0.500, 0.492, 0.488 s, the same equits to target)")
    ap.add_argument("--alpha", type=float, default=1.97,
                    help="over-relaxation factor of the full-resolution sweeps "
                         "(the reference's is 1.5)")
    ap.add_argument("--coarse-alpha", default="1.8,1.97",
                    help="over-relaxation of the coarse stages, coarsest first: one value "
                         "for every stage or a comma list")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned warm start leg")
    ap.add_argument("--learned-scans", type=int, default=8)
    ap.add_argument("--learned-steps", type=int, default=3000)
    ap.add_argument("--learned-channels", type=int, default=16)
    ap.add_argument("--config4", action="store_true",
                    help="the mixed-precision 512^2 x 32 x 24 workload, FP32 vs FP16 A")
    ap.add_argument("--config4-coarse", default="40,60,100")
    ap.add_argument("--config4-coarse-alpha", default="1.97",
                    help="over-relaxation of config 4's coarse stages (one value or one per "
                         "stage; tiles 8 and 16 there)")
    ap.add_argument("--config4-ref-equits", type=int, default=300)
    ap.add_argument("--profile-unet", type=int, default=0,
                    help="only time N forward passes of the warm-start network")
    ap.add_argument("--learned-coarse-equits", default="0,60",
                    help="comma list: equits of the two-grid correction of the learned start")
    ap.add_argument("--learned-input", choices=("fbp3", "fbp2", "fbp", "bp"), default="fbp2")
    ap.add_argument("--afp16", action="store_true",
                    help="store A in FP16 with a per-column FP32 scale")
    ap.add_argument("--placement", type=int, choices=(0, 1), default=0,
                    help="sweep work items on the XCDs: 0 tile-major, 1 each frame on one XCD")
    ap.add_argument("--pair-parities", default="1",
                    help="paired-parity launches (4 per equit instead of 8): one value for "
                         "every stage or a comma list, coarsest first, full resolution last")
    ap.add_argument("--xcd-check", action="store_true",
                    help="count sweep work items that ran off their intended XCD")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-only", action="store_true",
                    help="time only the reference's CPU path to its target (--cpu-budget)")
    ap.add_argument("--cpu-budget", type=float, default=120.0,
                    help="seconds of reference CPU time for the CPU baseline's run to its target")
    ap.add_argument("--p-t", type=float, default=None,
                    help="temporal qGGMRF p instead of the workload's (C* from "
                         "tests/golden/cstar_<workload>_pt<p>.json)")
    ap.add_argument("--longrun", type=int, default=0)
    ap.add_argument("--dump-x", default=None)
    ap.add_argument("--profile-equits", type=int, default=0)
    ap.add_argument("--calibrate", action="store_true",
                    help="with --profile-equits: also run the 1 GiB calibration copy")
    ap.add_argument("--pcmbir", action="store_true",
                    help="phase-retrieval throughput line instead of the MBIR metric")
    ap.add_argument("--pcmbir-views", type=int, default=360)
    ap.add_argument("--pcmbir-n", type=int, default=512)
    ap.add_argument("--pctomo", action="store_true",
                    help="the PCMBIR tomographic ICD at the PCMBIR object size instead")
    ap.add_argument("--pctomo-n", type=int, default=256)
    ap.add_argument("--pctomo-views", type=int, default=360)
    ap.add_argument("--pctomo-tile", type=int, default=8,
                    help="tile of the tomographic sweep (4 diverges at 256^3 x 360 views: too many concurrent columns per slice stack)")
    ap.add_argument("--aux", action="store_true",
                    help="timings of ring offsets, variance, Paganin and the simulator")
    ap.add_argument("--pcmbir-admm", action="store_true",
                    help="the PCMBIR reconstruction with its tomographic step, end to end")
    ap.add_argument("--pcmbir-fft", choices=("register", "lds"), default="register",
                    help="FFT passes: register-resident radix-8 (default) or in-LDS radix-2")
    args = ap.parse_args()
    global GRAPH, NH, NH_STAGES, PREDICT_CHECKS
    GRAPH = not args.no_graph
    PREDICT_CHECKS = not args.check_every_equit
    NH_STAGES = args.nh_stages
    if args.nh:
        p_, f_ = args.nh.split(",")
        NH = (int(p_), float(f_))
    if args.profile_unet and args.gpus == 1:
        profile_unet(args.profile_unet)
        return
    if args.config4 and args.gpus == 1:
        config4(args)
        return
    if args.aux and args.gpus == 1:
        aux_bench()
        return
    if args.pcmbir_admm and args.gpus == 1:
        pcmbir_admm_bench(args.pctomo_n, args.pctomo_views, max(args.steps, 1), args.pctomo_tile)
        return
    if args.pctomo and args.gpus == 1:
        pctomo_bench(args.pctomo_n, args.pctomo_views, max(args.steps, 1), args.pctomo_tile)
        return
    if args.pcmbir and args.gpus == 1:
        pcmbir_bench(args.pcmbir_views, args.pcmbir_n, max(args.steps, 1), nms=True,
                     fft_reg=int(args.pcmbir_fft == "register"))
        return
    if args.gpus != 1:
        print("bench.py: the MBIR sweep runs as one process on one GPU; --gpus must be 1 "
              "(multi-GPU runs are not supported).", file=sys.stderr)
        sys.exit(2)

    from phasecon_amd.synth import make_scan, WORKLOADS
    from phasecon_amd.mbir import Stage
    from phasecon_amd.lib import K_SWEEP, K_COST, K_AMATRIX, K_FWDPROJ

    if args.cpu_only:
        # the reference's CPU time to its own target, nothing on the GPU
        scan = make_scan(args.workload, seed=0)
        print(json.dumps({"cpu_baseline": cpu_baseline(scan, args.cpu_budget)}))
        return

    learned = not (args.no_learned or args.profile_equits or args.longrun)
    train_scans = []
    if learned:
        # training scans (other seeds) on the host, before the GPU is used
        import multiprocessing as mp
        t0 = time.time()
        with mp.get_context("fork").Pool(min(args.learned_scans, 8)) as pool:
            train_scans = pool.map(_make_train_scan,
                                   [(args.workload, s) for s in range(1, args.learned_scans + 1)])
        print(f"[bench] {len(train_scans)} training scans generated in {time.time() - t0:.1f}s",
              file=sys.stderr, flush=True)
    t0 = time.time()
    scan = make_scan(args.workload, seed=0)
    if args.p_t is not None:
        scan["p_t"] = args.p_t
    print(f"[bench] scan {args.workload} generated in {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)
    st = Stage(scan, tile=args.tile, afp16=args.afp16, alpha=args.alpha)
    coarse = []
    if args.schedule == "multires" and not (args.profile_equits or args.longrun):
        from phasecon_amd.mbir import multires_schedule
        stages = multires_schedule(scan["proj_rows"], scan["proj_cols"])[:-1]
        counts = [int(v) for v in str(args.coarse_equits).split(",")]
        if args.stages:  # explicit coarse stages "mxy:mz:equits,..."
            spec = [tuple(int(v) for v in item.split(":")) for item in args.stages.split(",")]
            stages = [(a, b) for a, b, _ in spec]
            counts = [n for _, _, n in spec]
        counts = (counts * len(stages))[:len(stages)] if len(counts) == 1 else counts
        if len(counts) != len(stages):
            sys.exit(f"bench.py: --coarse-equits needs 1 or {len(stages)} counts")
        alphas = coarse_alphas(args.coarse_alpha, len(stages))
        for (mxy, mz), n, a in zip(stages, counts, alphas):
            # a count of 0 leaves a leading stage out; a later one only
            # passes the object through (up-sampling doubles the grid)
            if n > 0 or coarse:
                coarse.append((Stage(scan, mult_xy=mxy, mult_z=mz, tile=args.coarse_tile,
                                     afp16=args.afp16, alpha=a), n))

    stages_all = [cs for cs, _ in coarse] + [st]
    pairs = [int(v) for v in str(args.pair_parities).split(",")]
    if len(pairs) == 1:
        pairs = pairs * len(stages_all)
    if len(pairs) != len(stages_all):
        sys.exit(f"bench.py: --pair-parities needs 1 or {len(stages_all)} values")
    for s_, pp in zip(stages_all, pairs):
        s_.option("pair_parities", pp)
    for s_ in stages_all:
        s_.option("placement", args.placement)
        if args.xcd_check:
            s_.option("xcd_check", 1)

    tuned = {}

    def tile_of(s_):
        key = f"{s_.g.N_x}x{s_.g.N_y}x{s_.g.N_z}"
        return tuned[key]["tile"] if key in tuned else (args.tile if s_ is st else args.coarse_tile)

    if not (args.no_autotune or args.profile_equits or args.longrun):
        # tile of every stage by timing a few equits at each candidate (untimed)
        from phasecon_amd.autotune import tune_tile
        for s_ in [cs for cs, n in coarse if n > 0] + [st]:
            s_.set_object(np.zeros(s_.shape_x, np.float32))
            s_.build_amatrix()
            best, tms = tune_tile(s_, candidates=tuple(int(v) for v in
                                                       args.tile_candidates.split(",")))
            tuned[f"{s_.g.N_x}x{s_.g.N_y}x{s_.g.N_z}"] = {
                "tile": best, "ms_per_launch": {str(k): round(v, 4) for k, v in tms.items()}}
        print(f"[bench] autotuned tiles: {tuned}", file=sys.stderr, flush=True)

    if args.profile_equits:
        st.set_object(np.zeros(st.shape_x, np.float32))
        st.build_amatrix()
        st.init_error()
        st.reset_kernel_time()
        for it in range(1, args.profile_equits + 1):
            st.icd(1, it, args.seed)
        ms, n = st.kernel_time(K_SWEEP)
        bytes_equit = algorithmic_bytes_per_equit(st)
        rec = {"profile_equits": args.profile_equits, "sweep_launches": n,
               "sweep_avg_ms": ms / n, "bytes_per_launch": bytes_equit * args.profile_equits / n}
        if args.calibrate:
            rec["calibration_copy_bytes"] = calibration_copy()
        print(json.dumps(rec))
        return

    if args.longrun:
        st.build_amatrix()
        st.init_error()
        for it in range(1, args.longrun + 1):
            st.icd(1, it, args.seed)
            if args.refresh and it % args.refresh == 0:
                st.init_error()
            if it % 500 == 0:
                print(f"[bench] longrun {it}: cost {st.cost()[0]:.12g}", file=sys.stderr, flush=True)
        c = st.cost(fresh=True)[0]
        x = st.get_object()
        if args.dump_x:
            np.save(args.dump_x, x)
        print(json.dumps({"longrun_equits": args.longrun, "fresh_cost": c, "digest": scan_digest(scan)}))
        return

    cstar, csrc = None, None
    rec = load_cstar(args.workload, args.p_t)
    if rec is not None:
        if rec.get("digest") == scan_digest(scan) and rec.get("p_t") == scan["p_t"]:
            cstar, csrc = rec["cstar"], "fp64-oracle"
    if cstar is None:
        # no matching oracle value: fall back to a long GPU run (reported as such)
        st.build_amatrix()
        st.init_error()
        for it in range(1, 3001):
            st.icd(1, it, args.seed)
        cstar, csrc = st.cost(fresh=True)[0], "gpu-3000-equits"
    target = cstar * (1.0 + REL_TOL)

    for _ in range(args.warmup):
        run_once(st, target, args.max_equits, args.seed, args.refresh, coarse)
    st.synchronize()
    st.reset_kernel_time()
    times, equits, costs, recoveries = [], [], [], []
    for k in range(args.steps):
        dt, n, c = run_once(st, target, args.max_equits, args.seed + k, args.refresh, coarse)
        times.append(dt)
        equits.append(n)
        costs.append(c)
        recoveries.extend({"step": k, **ev} for ev in GUARD_EVENTS)
        print(f"[bench] step {k}: {dt:.3f}s, {n} equits, cost {c:.10g} (target {target:.10g})"
              + (f", guard recoveries {GUARD_EVENTS}" if GUARD_EVENTS else ""),
              file=sys.stderr, flush=True)
    st.synchronize()
    sweep_ms, sweep_n = st.kernel_time(K_SWEEP)
    cost_ms, _ = st.kernel_time(K_COST)
    amat_ms, _ = st.kernel_time(K_AMATRIX)
    fwd_ms, _ = st.kernel_time(K_FWDPROJ)
    total_equits = int(np.sum(equits))
    bytes_equit = algorithmic_bytes_per_equit(st)
    launches_per_equit = sweep_n / max(total_equits, 1)
    bytes_launch = bytes_equit / launches_per_equit
    avg_launch_ms = sweep_ms / max(sweep_n, 1)
    achieved = bytes_launch / (avg_launch_ms * 1e-3) / 1e9
    value = float(np.mean(times))
    reached = all(c <= target for c in costs)
    w = WORKLOADS[args.workload]
    out = {
        "metric": METRIC,
        "value": value if reached else None,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": value * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.workload}: {w['cols']}x{w['cols']}x{w['rows']} voxels x {w['frames']} "
                         f"frames, {w['views']} views" + (f" interlaced K={w['K']}" if w['K'] else "")
                         + f", I0={w['I0']:g}, qGGMRF p_s={w['p_s']} p_t={scan['p_t']} q=2, from zero"),
            "tile": tile_of(st),
            "autotuned_tiles": tuned or None,
            "alpha": args.alpha,
            "coarse_alpha": [cs.alpha0 for cs, _ in coarse],
            "a_storage": "fp16 + fp32 column scale" if args.afp16 else "fp32",
            "schedule": (("multires: " + " -> ".join(
                f"{cs.g.N_x}x{cs.g.N_y}x{cs.g.N_z} ({n} equits, tile {tile_of(cs)})"
                for cs, n in coarse) + f" -> {st.g.N_x}x{st.g.N_y}x{st.g.N_z} to target, "
                "up-sampled on the device with out-of-mask voxels zeroed")
                if coarse else "single full-resolution stage from zero"),
            "refresh_every": args.refresh,
            "equit_launch": "captured HIP graph" if GRAPH else "direct launches",
            "cost_checks": ("scheduled from the observed decrease" if PREDICT_CHECKS
                            else "after every equit"),
            "nhicd": ({"passes": NH[0], "fraction": NH[1], "stages": NH_STAGES} if NH else None),
            "cstar": cstar,
            "cstar_source": csrc,
            "target_rel": REL_TOL,
        },
        "reached_target": reached,
        "equits_to_target": equits,
        "pair_parities": {f"{s_.g.N_x}x{s_.g.N_y}x{s_.g.N_z}": pp
                          for s_, pp in zip(stages_all, pairs)},
        "placement": {"frames_on_xcds": args.placement == 1,
                      "xcd_check": ({f"{s_.g.N_x}x{s_.g.N_y}x{s_.g.N_z}": s_.xcd_check()
                                     for s_ in [cs for cs, _ in coarse] + [st]}
                                    if args.xcd_check else None)},
        "guard_recoveries": recoveries,
        "coarse_equits": [n for _, n in coarse],
        "equits_per_s": total_equits / (sweep_ms * 1e-3),
        "hbm_gbps": achieved,
        "kernel_ms": {"sweep": sweep_ms, "cost": cost_ms, "amatrix": amat_ms, "init_error": fwd_ms},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS,
            "traffic": MEASURED_TRAFFIC.get((args.workload, tile_of(st), args.afp16,
                                             bool(pairs[-1]))),
            "traffic_source": TRAFFIC_SOURCE,
            "kernel": "k_sweep",
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": avg_launch_ms,
        },
    }
    if learned:
        try:
            out["learned"] = learned_leg(args, st, train_scans, target)
            lv = out["learned"]["value"]
            # ratios only between legs whose every step reached the target
            if lv and reached:
                out["learned"]["vs_multires"] = value / lv
            z = out["learned"]["zero_start"]["value"]
            if lv and z:
                out["learned"]["vs_zero_start"] = z / lv
        except Exception as exc:  # the headline line must still be printed
            out["learned"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scan, args.cpu_budget)
        cb = out["cpu_baseline"]
        if cb.get("value") and reached:
            cb["gpu_speedup"] = cb["value"] / value
    print(json.dumps(out))


if __name__ == "__main__":
    main()
