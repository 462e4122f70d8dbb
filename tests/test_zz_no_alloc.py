"""rism_k_fft_noalloc (no counterpart in the reference): the half-shifted
transform of rism_k_fft_r2c / rism_k_fft_c2r on caller-owned device buffers,
with its tables and partial sums in a static device array instead of
allocations.  Checked against those allocating entry points on two calls in
a row, the second with a smaller grid (the static array keeps the first
call's contents), in both directions, at the FP64 rounding gate of the
transform tests (1e-12 of the largest value; the tables are computed on the
device here and on the host there, so they may differ in the last bit).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(rism, grid, box, seed):
    import torch

    nx, ny, nz = grid
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((nz, ny, nx))
    ref_spec = rism.fft_r2c(f, box)
    ref_back = rism.fft_c2r(ref_spec, box)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    real = torch.from_numpy(f.reshape(-1).copy()).to(dev)
    spec = torch.zeros(nz // 2 * ny * nx * 2, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    rism.fft_noalloc(-1, grid, box, real.data_ptr(), spec.data_ptr(), stream)
    torch.cuda.synchronize()
    got = spec.cpu().numpy().view(np.complex128).reshape(nz // 2, ny, nx)
    err_f = np.abs(got - ref_spec).max() / np.abs(ref_spec).max()
    # the inverse consumes its spectrum: give it a copy of the reference one
    spec2 = torch.from_numpy(ref_spec.reshape(-1).view(np.float64).copy()).to(dev)
    out = torch.zeros(nz * ny * nx, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    rism.fft_noalloc(+1, grid, box, out.data_ptr(), spec2.data_ptr(), stream)
    torch.cuda.synchronize()
    back = out.cpu().numpy().reshape(nz, ny, nx)
    err_b = np.abs(back - ref_back).max() / np.abs(ref_back).max()
    err_rt = np.abs(back - f).max() / np.abs(f).max()
    print(grid, "r2c", err_f, "c2r", err_b, "round trip", err_rt, flush=True)
    assert err_f < 1e-12
    assert err_b < 1e-12
    assert err_rt < 1e-12


def test_noalloc_matches_allocating_entry_points(rism, gpu):
    _call(rism, (64, 48, 80), (32.0, 24.0, 40.0), seed=1)
    _call(rism, (32, 24, 40), (16.0, 12.0, 20.0), seed=2)


def test_noalloc_refuses_grids_beyond_its_scratch(rism, gpu):
    import torch

    t = torch.zeros(16, dtype=torch.float64, device=torch.device("cuda:0"))
    with pytest.raises(rism.RismError) as e:
        rism.fft_noalloc(-1, (1024, 1024, 2048), (512.0, 512.0, 1024.0), t.data_ptr(),
                         t.data_ptr(), 0)
    assert e.value.code == rism.RISM_ERR_ARG
    assert "too large" in str(e.value)
