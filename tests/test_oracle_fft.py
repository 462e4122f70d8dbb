"""Pins the oracle's half-shifted transform (or_fft3d) on the CPU:
  * against the definition F(k_j) = dv sum_n f(r_n) e^{-i k_j . r_n} evaluated
    by brute force (O(N^2)) on small grids, including N % 4 == 2 axes;
  * against a literal numpy restatement of the reference's own recipe
    (fft3d_cuda.cu:56-128 twiddles around an unnormalised cuFFT-convention
    DFT, fft3d.cu:81-163), which equals the definition times
    S = prod_d (-1)^{N_d/2}: identical whenever every N_d % 4 == 0;
  * round trip, Parseval, linearity."""
import numpy as np
import pytest


def axis(n, b):
    d = b / n
    r = (np.arange(n) - n // 2) * d
    k = 2 * np.pi * (np.arange(n) - n / 2 + 0.5) / b
    return r, k


def brute(f, box):
    nz, ny, nx = f.shape
    rx, kx = axis(nx, box[0])
    ry, ky = axis(ny, box[1])
    rz, kz = axis(nz, box[2])
    Ex, Ey, Ez = (np.exp(-1j * np.outer(k, r)) for k, r in ((kx, rx), (ky, ry), (kz, rz)))
    dv = np.prod(box) / f.size
    return dv * np.einsum("az,by,cx,zyx->abc", Ez, Ey, Ex, f)


def reference_recipe(f, box):
    """fft3d.cu forward at P = 1: timekf (x dir, x dkf), 3 x cuFFT forward,
    timeirvolf (x dik, x dv).  dkf = exp(-i pi sum_d (n_d - N_d/2.0)/N_d)
    truncated to int as in setkf (fft3d_cuda.cu:94-106); dir/dik = (-1)^{sum}."""
    nz, ny, nx = f.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    xs = np.trunc(x - nx / 2.0)
    ys = np.trunc(y - ny / 2.0)
    zs = np.trunc(z - nz / 2.0)
    dkr = np.pi / nx * xs + np.pi / ny * ys + np.pi / nz * zs
    dkf = np.cos(dkr) - 1j * np.sin(dkr)
    sgn = np.where((x + y + z) % 2 == 0, 1.0, -1.0)
    a = f * sgn * dkf
    a = np.fft.fftn(a, axes=(0, 1, 2))
    dv = np.prod(box) / f.size
    return a * sgn * dv


SMALL = [((4, 4, 4), (2.0, 2.0, 2.0)), ((8, 8, 8), (3.1, 4.7, 5.3)),
         ((6, 10, 12), (3.0, 5.0, 6.5)), ((8, 12, 6), (4.0, 4.0, 4.0)),
         ((12, 8, 16), (6.0, 2.5, 7.0))]


@pytest.mark.parametrize("shape,box", SMALL)
def test_matches_definition(oracle, shape, box):
    nx, ny, nz = shape
    rng = np.random.default_rng(sum(shape))
    f = rng.standard_normal((nz, ny, nx)) + 1j * rng.standard_normal((nz, ny, nx))
    F = oracle.fft3d(f, box, -1)
    B = brute(f, box)
    assert np.abs(F - B).max() < 1e-13 * np.abs(B).max()


@pytest.mark.parametrize("shape,box", SMALL)
def test_reference_recipe_phase(oracle, shape, box):
    nx, ny, nz = shape
    rng = np.random.default_rng(7 + sum(shape))
    f = rng.standard_normal((nz, ny, nx))
    F = oracle.fft3d(f, box, -1)
    R = reference_recipe(f, box)
    S = np.prod([(-1) ** (n // 2) for n in shape])
    assert np.abs(R - S * F).max() < 1e-12 * np.abs(F).max()
    if all(n % 4 == 0 for n in shape):
        assert S == 1


@pytest.mark.parametrize("shape,box", SMALL + [((32, 32, 32), (16.0, 16.0, 16.0))])
def test_roundtrip_parseval_linearity(oracle, shape, box):
    nx, ny, nz = shape
    rng = np.random.default_rng(3)
    f = rng.standard_normal((nz, ny, nx))
    g = rng.standard_normal((nz, ny, nx))
    F = oracle.fft3d(f, box, -1)
    back = oracle.fft3d(F, box, +1)
    assert np.abs(back - f).max() < 1e-13 * np.abs(f).max()
    V = np.prod(box)
    dv = V / f.size
    assert np.isclose((np.abs(F) ** 2).sum() / V, dv * (f * f).sum(), rtol=1e-12)
    Fs = oracle.fft3d(2 * f - 0.5 * g, box, -1)
    Fg = oracle.fft3d(g, box, -1)
    assert np.abs(Fs - (2 * F - 0.5 * Fg)).max() < 1e-13 * np.abs(Fs).max()


def test_real_input_hermitian(oracle):
    """k_{N-1-j} = -k_j on every axis, so a real field has F(-k) = conj F(k);
    the GPU path stores only k_z > 0 because of this."""
    f = np.random.default_rng(1).standard_normal((8, 6, 10))
    F = oracle.fft3d(f, (4.0, 3.0, 5.0), -1)
    assert np.abs(F[::-1, ::-1, ::-1] - np.conj(F)).max() < 1e-13 * np.abs(F).max()
