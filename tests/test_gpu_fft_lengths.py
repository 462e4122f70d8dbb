"""GPU parity of the half-shifted 3D transform (FFT3D::execute, fft3d.cu:72-163)
against the FP64 oracle (rism_oracle.c or_fft3d, pinned to a brute-force DFT
in test_oracle_fft.py) for every FFT length the kernels are instantiated for
(RISM_LENGTHS in csrc/rism_common.h), along each axis: as nx (x pass), as ny
(y passes) and as nz/2 (the z passes' half-complex columns, including the
radix-4 forward plan used for nz/2 >= 128 in FP64).  Thin grids keep the
oracle cheap; the two other axes are the smallest supported sizes.

Tolerance: FP64, relative to the largest magnitude, 1e-12 (different radix
orders on the two sides; the mismatch is rounding only).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LENGTHS = [8, 12, 16, 20, 24, 30, 32, 36, 40, 48, 60, 64, 72, 80, 90, 96, 128, 144, 160,
           180, 192, 256, 512, 1024]

SHAPES = []
for L in LENGTHS:
    SHAPES.append(("x", (L, 8, 16)))
    SHAPES.append(("y", (8, L, 16)))
    SHAPES.append(("z", (8, 8, 2 * L)))
# the batched medium grids' lengths on all three axes at once
SHAPES += [("xyz", (144, 160, 360)), ("xyz", (90, 72, 160)), ("xyz", (60, 96, 384))]


def _check(rism, oracle, shape, seed):
    nx, ny, nz = shape
    box = (0.5 * nx, 0.5 * ny, 0.5 * nz)
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((nz, ny, nx))
    F = oracle.fft3d(f, box, -1)
    G = rism.fft_r2c(f, box)
    ref = F[nz // 2:]
    err_f = np.abs(G - ref).max() / np.abs(ref).max()
    back = rism.fft_c2r(ref, box)
    inv = oracle.fft3d(F, box, +1).real
    err_b = np.abs(back - inv).max() / np.abs(inv).max()
    err_rt = np.abs(back - f).max() / np.abs(f).max()
    print(shape, "r2c", err_f, "c2r", err_b, "round trip", err_rt, flush=True)
    assert err_f < 1e-12
    assert err_b < 1e-12
    assert err_rt < 1e-12


@pytest.mark.parametrize("axis,shape", SHAPES, ids=[f"{a}-{s[0]}x{s[1]}x{s[2]}" for a, s in SHAPES])
def test_length_matches_oracle(rism, oracle, gpu, axis, shape):
    _check(rism, oracle, shape, seed=shape[0] * 1009 + shape[1] * 31 + shape[2])

