"""The phase-contrast phantom simulator in the oracle (phantom_project,
pcmbir_oracle.simulate; PCMBIR_3D/XT_genSinogram.c:59-264).  The
reference's simulator needs FFTW (absent here) and fixed phantom files, so
it is pinned by what fixes it whatever the code: the z-line response is the
identity when the phantom's slices are the detector rows, each view's
projection conserves the phantom's mass (the detector response has unit
area), projections are linear in the phantoms and the wavenumber, an empty
phantom measures the expected count everywhere, and at zero propagation
distance the intensity is I0 exp(-2 p_mag) whatever the phase."""
import numpy as np

from oracle import oracle as O
from oracle import pcmbir_oracle as P

N_XY, N_Z, ROWS, COLS = 32, 8, 4, 16
VOX, RC = 2.0, 8.0


def _disk(seed=0):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[:N_XY, :N_XY] - (N_XY - 1) / 2
    disk = (xx ** 2 + yy ** 2 < (N_XY / 3) ** 2).astype(np.float32)
    mag = disk[None] * rng.random((N_Z, 1, 1)).astype(np.float32) * 1e-4
    phase = disk[None] * rng.random((N_Z, N_XY, N_XY)).astype(np.float32) * 1e-3
    return mag, phase


def test_z_response_is_the_identity_on_matched_slices():
    ang = np.linspace(0, np.pi, 5, endpoint=False).astype(np.float32)
    mag, _ = _disk()
    for z in (0, 3, N_Z - 1):
        one = np.zeros_like(mag)
        one[z] = mag[z]
        pr, _ = O.phantom_project(one, np.zeros_like(one), ang, ROWS, COLS, VOX, RC, 1.0)
        rows = np.nonzero(np.abs(pr).sum(axis=(0, 2)))[0]
        assert list(rows) == [z]


def test_projection_conserves_mass_per_view_and_row():
    ang = np.linspace(0, np.pi, 7, endpoint=False).astype(np.float32)
    mag, phase = _disk()
    k = 3.5
    pr, pi = O.phantom_project(mag, phase, ang, ROWS, COLS, VOX, RC, k)
    d = VOX * COLS / N_XY                        # detector pitch = voxel width
    for src, out in ((mag, pr), (phase, pi)):
        mass = k * d * src.astype(np.float64).sum(axis=(1, 2))   # k * area * value / pitch
        got = out.sum(axis=2)                                    # [view][row]
        # a voxel's response sums to its area / pitch up to the tabulated
        # profile's sampling (64 bins, linear interpolation): within 1 % per
        # row, a few 1e-3 over a view
        assert np.allclose(got, mass[None], rtol=1e-2)
        assert np.allclose(got.sum(axis=1), mass.sum(), rtol=5e-3)
        assert np.allclose(got[0], mass, rtol=1e-6)        # view 0: exact


def test_projection_is_linear():
    ang = np.array([0.3, 1.1, 2.0], np.float32)
    m1, p1 = _disk(1)
    m2, p2 = _disk(2)
    a = O.phantom_project(m1, p1, ang, ROWS, COLS, VOX, RC, 2.0)
    b = O.phantom_project(m2, p2, ang, ROWS, COLS, VOX, RC, 2.0)
    c = O.phantom_project(m1 + 3 * m2, p1 + 3 * p2, ang, ROWS, COLS, VOX, RC, 4.0)
    for i in range(2):
        assert np.allclose(c[i], 2 * (a[i] + 3 * b[i]), rtol=1e-6, atol=1e-12)


def test_simulate_empty_phantom_and_zero_distance():
    ang = np.linspace(0, np.pi, 4, endpoint=False).astype(np.float32)
    zero = np.zeros((N_Z, N_XY, N_XY), np.float32)
    m = P.simulate(zero, zero, ang, ROWS, COLS, VOX, RC, 1e5, 17.0)
    assert np.allclose(m, P.EXPECTED_COUNT, rtol=1e-12)
    mag, phase = _disk()
    m = P.simulate(mag, phase, ang, ROWS, COLS, VOX, RC, 1e-12, 17.0)
    k = 2 * np.pi / P.wavelength(17.0)
    pr, _ = O.phantom_project(mag, phase, ang, ROWS, COLS, VOX, RC, k)
    inten = P.EXPECTED_COUNT * np.exp(-2 * pr)            # [view][t][r]
    rs, ts = N_XY // COLS, N_Z // ROWS
    want = inten.reshape(len(ang), ROWS, ts, COLS, rs).mean(axis=(2, 4))
    assert np.allclose(m, want, rtol=1e-9)
