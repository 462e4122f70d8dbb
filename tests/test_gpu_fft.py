"""GPU parity of the half-shifted 3D transform (FFT3D::execute, fft3d.cu:72-163)
against the FP64 oracle (rism_oracle.c or_fft3d, itself pinned to a brute-force
DFT in test_oracle_fft.py).

Tolerance: FP64, relative to the largest spectrum magnitude, 1e-12 (the two
sides use different radix orders; the mismatch is rounding only).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [
    ((16, 16, 16), (8.0, 8.0, 8.0)),
    ((32, 24, 24), (10.0, 9.0, 7.5)),
    ((64, 64, 64), (32.0, 32.0, 32.0)),
    ((48, 40, 40), (20.0, 17.0, 15.0)),   # radix 3 and 5 along x, y and z/2
    ((96, 60, 48), (40.0, 30.0, 24.0)),
    ((128, 128, 64), (64.0, 64.0, 32.0)),
]


@pytest.mark.parametrize("shape,box", SHAPES)
def test_r2c_matches_oracle(rism, oracle, gpu, shape, box):
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 7 + ny * 3 + nz)
    f = rng.standard_normal((nz, ny, nx))
    F = oracle.fft3d(f, box, -1)
    G = rism.fft_r2c(f, box)
    ref = F[nz // 2:]
    err = np.abs(G - ref).max() / np.abs(ref).max()
    print(shape, "r2c rel err", err)
    assert err < 1e-12


@pytest.mark.parametrize("shape,box", SHAPES)
def test_c2r_matches_oracle(rism, oracle, gpu, shape, box):
    nx, ny, nz = shape
    rng = np.random.default_rng(nx + ny + nz)
    f = rng.standard_normal((nz, ny, nx))
    F = oracle.fft3d(f, box, -1)
    back = rism.fft_c2r(F[nz // 2:], box)
    ref = oracle.fft3d(F, box, +1).real
    assert np.abs(back - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(back - f).max() < 1e-12 * np.abs(f).max()


def test_roundtrip_512_identity(rism, gpu):
    """Size-independent property at the 512^3 production shape: inverse(forward(f)) = f
    and linearity F(a f + b g) = a F(f) + b F(g)."""
    nx = ny = nz = 512
    box = (256.0, 256.0, 256.0)
    rng = np.random.default_rng(5)
    f = rng.standard_normal((nz, ny, nx))
    F = rism.fft_r2c(f, box)
    back = rism.fft_c2r(F, box)
    assert np.abs(back - f).max() < 1e-12 * np.abs(f).max()
    g = np.cos(np.arange(nz))[:, None, None] * np.ones((nz, ny, nx))
    Fg = rism.fft_r2c(g, box)
    Fs = rism.fft_r2c(2.0 * f - 3.0 * g, box)
    assert np.abs(Fs - (2.0 * F - 3.0 * Fg)).max() < 1e-12 * np.abs(Fs).max()


def test_gaussian_closed_form(rism, gpu):
    """A Gaussian exp(-a r^2) centred on the grid origin has the closed-form
    transform (pi/a)^{3/2} exp(-k^2/(4a)); with a = 1/4 on a 0.5 A grid in a
    32 A box both the k-space aliases (e^{-(2 pi/dx - k)^2}) and the r-space
    truncation (e^{-a (L/2)^2}) are below 1e-16."""
    nx, ny, nz = 64, 64, 64
    L = 32.0
    box = (L, L, L)
    a = 0.25
    r = (np.arange(nx) - nx // 2) * (L / nx)
    z, y, x = np.meshgrid(r, r, r, indexing="ij")
    f = np.exp(-a * (x * x + y * y + z * z))
    G = rism.fft_r2c(f, box)
    k = 2 * np.pi * (np.arange(nx) - nx / 2 + 0.5) / L
    kz, ky, kx = np.meshgrid(k[nz // 2:], k, k, indexing="ij")
    exact = (np.pi / a) ** 1.5 * np.exp(-(kx * kx + ky * ky + kz * kz) / (4 * a))
    assert np.abs(G - exact).max() < 1e-12 * exact.max()
