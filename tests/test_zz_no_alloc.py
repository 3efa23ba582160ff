"""mbir_hip_substitution_batch_noalloc against mbir_hip_substitution_batch.

The no-allocation entry point stages its inputs and output through the
library's fixed device array instead of allocating; it runs the same kernel
as the allocating entry point, so the outputs must be identical bit for bit.
Two calls in a row, the second smaller, show that a call does not depend on
what the previous one left in the array; a call above the documented maximum
is refused with a message.  Called through the C-ABI (ctypes).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    V = (rng.random(n) * 0.02).astype(np.float32)
    nb = (rng.random((n, 27)) * 0.02).astype(np.float32)
    nb[::5] = V[::5, None]
    nbf = (rng.random((n, 27)) > 0.15).astype(np.uint8)
    tn = (rng.random((n, 2)) * 0.02).astype(np.float32)
    tf = (rng.random((n, 2)) > 0.3).astype(np.uint8)
    th2 = (rng.random(n) * 2e4).astype(np.float32)
    th1 = (rng.standard_normal(n) * 20).astype(np.float32)
    return V, th1, th2, nb, nbf, tn, tf


def _call(fn, g, args):
    V, th1, th2, nb, nbf, tn, tf = args
    out = np.full(len(V), np.nan, np.float32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = fn(C.byref(g), C.c_int32(len(V)), p(V), p(th1), p(th2), p(nb), p(nbf), p(tn), p(tf),
            p(out))
    return rc, out


def test_noalloc_substitution_matches_allocating_entry(small4d_scan):
    from phasecon_amd import lib
    from phasecon_amd.mbir import Stage
    L = lib.load()
    st = Stage(small4d_scan, alpha=1.7)
    g = st.g
    for n, seed in ((lib_max(), 1), (1000, 2)):   # the second call smaller
        args = _inputs(n, seed)
        print(f"noalloc substitution: n = {n}", flush=True)
        rc0, want = _call(L.mbir_hip_substitution_batch, g, args)
        assert rc0 == 0, lib.last_error()
        rc1, got = _call(L.mbir_hip_substitution_batch_noalloc, g, args)
        assert rc1 == 0, lib.last_error()
        assert np.isfinite(want).all()
        assert np.array_equal(got, want)
    rc, _ = _call(L.mbir_hip_substitution_batch_noalloc, g, _inputs(lib_max() + 1, 3))
    assert rc == -1 and "MBIR_SUBST_NOALLOC_MAX" in lib.last_error()
    st.close()


def lib_max():
    """MBIR_SUBST_NOALLOC_MAX from include/mbir_hip.h."""
    import os
    import re
    from conftest import ROOT
    text = open(os.path.join(ROOT, "include", "mbir_hip.h")).read()
    return int(re.search(r"#define MBIR_SUBST_NOALLOC_MAX (\d+)", text).group(1))
