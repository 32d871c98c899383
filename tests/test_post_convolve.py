"""MCS post-convolution of EdepC rows (gpu_transport_wrapper.cu:551-616), in the library
(sm2d_post_convolve, host code) and the oracle, against the discrete Gaussian it must apply:
a unit impulse in every row comes back as the normalised kernel exp(-k^2 / (2 s^2)),
s = sqrt(sigma_total^2 - sigma_x0^2) / dx, truncated at ceil(3 s) cells (clamped to [1, Nx/2]),
zero-padded at the row ends; rows with sigma_total^2 - sigma_x0^2 <= 1e-6 are left alone.
The kernel's second moment is that of a Gaussian of width s truncated at 3 s (a wrong width, for
instance sigma_total in place of the MCS part, fails it), and away from the edges a row keeps its
sum."""
import numpy as np
import pytest

from oracle import pyoracle
from sm_2d_amd import runtime


@pytest.fixture(scope="module", params=["hip", "oracle"])
def lib(request):
    return runtime.load_library() if request.param == "hip" else pyoracle.load()


def _expected_row(Nx, pos, s):
    half = int(np.ceil(np.float32(3.0) * np.float32(s)))
    half = max(1, min(half, Nx // 2))
    k = np.arange(-half, half + 1, dtype=np.float64)
    ker = np.exp(-0.5 * k * k / (float(np.float32(s)) ** 2))
    ker /= ker.sum()
    row = np.zeros(Nx)
    for i, v in zip(pos + k.astype(int), ker):
        if 0 <= i < Nx:
            row[i] += v
    return row, half, ker


def test_impulse_response_is_the_discrete_gaussian(lib):
    Nx, Nz, dx, sx0 = 101, 6, 0.5, 3.0
    st = np.array([3.0, 3.0000001, 3.2, 4.0, 6.5, 40.0], np.float32)  # last row: half-width clamps
    ed = np.zeros((Nz, Nx))
    ed[:, 50] = 1.0
    out = lib.post_convolve(ed, dx, st, sx0)
    for iz in range(Nz):
        mcs2 = np.float32(st[iz]) * np.float32(st[iz]) - np.float32(sx0) * np.float32(sx0)
        if mcs2 <= np.float32(1e-6):
            assert np.array_equal(out[iz], ed[iz]), iz
            continue
        s = np.float32(np.sqrt(np.float32(mcs2))) / np.float32(dx)
        row, half, ker = _expected_row(Nx, 50, s)
        assert np.allclose(out[iz], row, rtol=1e-12, atol=1e-15), iz
        if half < Nx // 2:
            k = np.arange(-half, half + 1)
            m2 = float((ker * k * k).sum())
            # truncation at +-3 s lowers the variance of a Gaussian to 0.973 s^2
            assert 0.95 * float(s) ** 2 < m2 < float(s) ** 2 + 0.1, (iz, m2, float(s) ** 2)
        else:
            assert half == Nx // 2


def test_row_sums_kept_away_from_edges(lib):
    rng = np.random.default_rng(3)
    Nx, Nz = 200, 40
    ed = np.zeros((Nz, Nx))
    ed[:, 60:140] = rng.random((Nz, 80))
    st = np.linspace(3.0, 8.0, Nz).astype(np.float32)
    out = lib.post_convolve(ed, 0.5, st, 3.0)
    assert np.allclose(out.sum(axis=1), ed.sum(axis=1), rtol=1e-12)
    assert np.all(out >= 0.0)
