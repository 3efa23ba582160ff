"""Writes the golden vectors under tests/golden/ from the reference build.

Run in a container where the reference sources exist:
    make -C oracle ref && python tests/golden/make_golden.py

Every expected value comes from the reference's own C functions (compiled
from its sources into oracle/_ref/libmbir4d_ref.so, driven through
oracle/ref_harness.c); inputs come from the seeded generator
phasecon_amd/synth.py.  The fixtures are plain arrays (inputs and outputs)
in .npz files:

  mbir_<scan>.npz  geometry dump, view map, detector response table,
                   A columns of a column subset, e = y - A x for a seeded
                   object, cost of that state, functional-substitution cases,
                   single-voxel updates, and a reference ICD cost trajectory
                   with its final object (tiny3d only).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Reference  # noqa: E402
from phasecon_amd.synth import make_scan  # noqa: E402


def neighbourhoods(rng, n):
    V = rng.random(n) * 0.02
    nb = rng.random((n, 27)) * 0.02
    nb[::4] = V[::4, None]
    nbf = (rng.random((n, 27)) > 0.15).astype(np.uint8)
    tn = rng.random((n, 2)) * 0.02
    tf = (rng.random((n, 2)) > 0.3).astype(np.uint8)
    th1 = rng.standard_normal(n) * 20
    th2 = rng.random(n) * 2e4
    return V, nb, nbf, tn, tf, th1, th2


def golden(name, seed, icd_iters):
    scan = make_scan(name, seed=seed)
    rng = np.random.default_rng(1234)
    out = {}
    ref = Reference(scan, threads=1)
    out["ints"], out["dbls"] = ref.ints, ref.dbls
    out["frame_ptr"], out["frame_views"] = ref.view_map()
    out["H"] = ref.detector_response()
    cols = rng.choice(ref.N_x * ref.N_y, size=min(32, ref.N_x * ref.N_y), replace=False)
    idx = np.full((len(cols), ref.N_p, 4), -1, np.int32)
    val = np.zeros((len(cols), ref.N_p, 4), np.float64)
    for i, c in enumerate(cols):
        for p in range(ref.N_p):
            ii, vv = ref.amatrix_column(int(c // ref.N_x), int(c % ref.N_x), p)
            idx[i, p, :len(ii)] = ii
            val[i, p, :len(vv)] = vv
    out["a_cols"], out["a_idx"], out["a_val"] = cols.astype(np.int32), idx, val
    # object on a 1e-4 grid so the fixture compresses
    x = (np.round(rng.random(ref.shape_x) * 200) * 1e-4).astype(np.float32)
    ref.set_object(x)
    out["x"] = x
    out["e"] = ref.init_error()
    out["w"] = ref.get_weight()
    c, parts = ref.cost()
    out["cost"] = np.array([c, parts[0], parts[1]])
    V, nb, nbf, tn, tf, th1, th2 = neighbourhoods(rng, 256)
    out["sub_in"] = np.concatenate([V[:, None], th1[:, None], th2[:, None], nb, tn], axis=1)
    out["sub_flags"] = np.concatenate([nbf, tf], axis=1)
    out["sub_out"] = np.array([ref.substitution(V[i], th1[i], th2[i], nb[i], nbf[i], tn[i], tf[i])
                               for i in range(len(V))])
    # single-voxel updates through the reference's voxel-line driver
    vox = np.array([(rng.integers(ref.N_time), rng.integers(ref.N_z), rng.integers(ref.N_y),
                     rng.integers(ref.N_x)) for _ in range(8)], np.int32)
    after_x, after_e = [], []
    for t, z, yy, xx in vox:
        ref.update_voxel(int(t), int(z), int(yy), int(xx), 1)
        after_x.append(ref.get_object())
        after_e.append(ref.get_error())
    out["vox"] = vox
    out["vox_x"] = np.array([a[tuple(v)] for a, v in zip(after_x, vox)], np.float32)
    out["vox_e_final"] = after_e[-1]
    ref.close()
    if icd_iters:
        r2 = Reference(scan, threads=1)
        r2.set_object(np.zeros(r2.shape_x, np.float32))
        r2.init_error()
        costs, conv, _ = r2.icd_run(icd_iters)
        out["icd_costs"] = costs
        out["icd_x"] = r2.get_object()
        r2.close()
    path = os.path.join(HERE, f"mbir_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


if __name__ == "__main__":
    golden("tiny3d", 0, 150)
    golden("ch4d_small", 3, 0)
