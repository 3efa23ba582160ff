"""The phase-contrast phantom simulator on the GPU: the phantom projection
(mbir_hip_phantom_project) and the noise-free measurements
(phcontomo_forward_project_ex) against the oracle's restatement of
ForwardProject (PCMBIR_3D/XT_genSinogram.c:146-264), the noise against its
distribution, and the drop-in phcontomo_forward_project at the reference's
phantom size (128 x 512 x 512)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _phantoms(n_xy, n_z, seed=0):
    """Absorption index ~1e-9 and refractive-index decrement ~1e-7 (hard
    X-rays: at 17 keV, k = 8.6e5 per um, a few 1e-2 of absorption and a few
    radians of phase across the object)."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[:n_xy, :n_xy] - (n_xy - 1) / 2
    disk = (xx ** 2 + (yy / 1.3) ** 2 < (n_xy / 3) ** 2).astype(np.float32)
    mag = disk[None] * rng.random((n_z, n_xy, n_xy)).astype(np.float32) * 2e-9
    phase = disk[None] * (1 + rng.random((n_z, 1, 1))).astype(np.float32) * 1e-7
    return mag, phase


@pytest.mark.parametrize("n_xy,n_z,rows,cols", [(64, 16, 8, 32), (48, 40, 4, 16)])
def test_phantom_projection_matches_oracle(n_xy, n_z, rows, cols):
    from oracle.oracle import phantom_project as ref
    from phasecon_amd.mbir import phantom_project
    ang = np.linspace(0, np.pi, 12, endpoint=False).astype(np.float32)
    mag, phase = _phantoms(n_xy, n_z)
    print(f"phantom {n_z}x{n_xy}x{n_xy}, detector {rows}x{cols}", flush=True)
    pr, pi = phantom_project(mag, phase, ang, rows, cols, 1.5, cols / 2 - 0.3, 2.0e4)
    rr, ri = ref(mag, phase, ang, rows, cols, 1.5, cols / 2 - 0.3, 2.0e4)
    for got, want in ((pr, rr), (pi, ri)):
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"max error / max {err:.3g}", flush=True)
        assert err <= 2e-6
        assert np.allclose(got, want, rtol=1e-4, atol=1e-6 * np.abs(want).max())


def test_simulated_measurements_match_oracle(tmp_path, monkeypatch):
    from oracle import pcmbir_oracle as P
    from phasecon_amd.pcmbir import forward_project
    monkeypatch.chdir(tmp_path)
    n_xy, n_z, rows, cols = 64, 16, 8, 32
    ang = np.linspace(0, np.pi, 6, endpoint=False).astype(np.float32)
    mag, phase = _phantoms(n_xy, n_z, 1)
    args = (ang, rows, cols, 1.0, cols / 2, 1.0e5, 17.0)
    meas, br = forward_project(mag, phase, *args)
    want = P.simulate(mag, phase, *args)
    err = np.abs(meas - want).max() / want.max()
    print(f"measurements: max error / max {err:.3g}, range {want.min():.1f}..{want.max():.1f}",
          flush=True)
    assert np.array_equal(br, np.full((rows, cols), P.EXPECTED_COUNT, np.float32))
    assert err <= 2e-5
    assert os.path.getsize("SimMagProjs.bin") == 4 * ang.size * n_z * n_xy
    # the noise: |m + sqrt(m) n| with n standard normal
    noisy, _ = forward_project(mag, phase, *args, noise_seed=12345)
    z = (noisy.astype(np.float64) - meas) / np.sqrt(meas)
    print(f"noise: mean {z.mean():.3f}, std {z.std():.3f} over {z.size}", flush=True)
    assert want.min() > 100           # |.| of the noise model is inactive
    assert abs(z.mean()) < 0.1 and abs(z.std() - 1) < 0.1
    # the same seed, the same noise (the projection's LDS sums may differ in
    # their last bits from run to run)
    again, _ = forward_project(mag, phase, *args, noise_seed=12345)
    assert np.allclose(again, noisy, rtol=1e-5)


def test_dropin_forward_project_at_the_reference_phantom_size(tmp_path, monkeypatch):
    from phasecon_amd.pcmbir import phcontomo_forward_project
    mag, phase = _phantoms(512, 128, 2)
    mag.tofile(tmp_path / "mag.bin")
    phase.tofile(tmp_path / "phase.bin")
    del mag, phase
    monkeypatch.setenv("PHCONTOMO_MAG_PHANTOM", str(tmp_path / "mag.bin"))
    monkeypatch.setenv("PHCONTOMO_PHASE_PHANTOM", str(tmp_path / "phase.bin"))
    monkeypatch.chdir(tmp_path)
    ang = np.linspace(0, np.pi, 4, endpoint=False).astype(np.float32)
    rc, meas, br = phcontomo_forward_project(ang, 32, 128, 1.0, 64.0, 1.0e5, 17.0)
    print(f"drop-in: rc {rc}, measurements {meas.min():.1f}..{meas.max():.1f}", flush=True)
    assert rc == 0
    assert np.all(np.isfinite(meas)) and meas.min() >= 0 and meas.max() > 0
    assert np.all(br == 2500.0)
    assert os.path.getsize("SimPhaseProjs.bin") == 4 * 4 * 128 * 512
