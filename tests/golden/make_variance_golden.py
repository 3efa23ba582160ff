"""Writes tests/golden/variance_shepp3d.json: the noise-variance estimate the
reference's own reconstruct() reaches on the seeded Shepp-Logan workload
with the generalized-Huber data term (quad_convex = 0, T = 4, delta = 0.5,
convergence threshold 0.1 %), read from its variance_estimate.bin, with the
scan's SHA-256.  The reference run takes minutes on the CPU, too long for a
GPU-box test, hence the fixture.

Run in a container where the reference sources exist:
    make -C oracle ref && python tests/golden/make_variance_golden.py
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle.oracle import reference_reconstruct  # noqa: E402
from phasecon_amd.synth import make_scan  # noqa: E402


def digest(scan):
    h = hashlib.sha256()
    for k in ("projections", "weights", "angles", "proj_times", "recon_times"):
        h.update(np.ascontiguousarray(scan[k], np.float32).tobytes())
    return h.hexdigest()


if __name__ == "__main__":
    scan = make_scan("shepp3d", seed=0)
    d = tempfile.mkdtemp(prefix="varref_")
    reference_reconstruct(scan, convg=0.1, quad_convex=0, huber_T=4.0, huber_delta=0.5,
                          scratch=d)
    var = float(np.fromfile(os.path.join(d, "variance_estimate.bin"), np.float64)[0])
    out = {"workload": "shepp3d", "seed": 0, "quad_convex": 0, "huber_T": 4.0,
           "huber_delta": 0.5, "convg_thresh": 0.1, "variance_estimate": var,
           "digest": digest(scan), "source": "reference reconstruct() (oracle/_ref)"}
    with open(os.path.join(HERE, "variance_shepp3d.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)
