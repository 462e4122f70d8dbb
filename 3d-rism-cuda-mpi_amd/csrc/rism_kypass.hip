// rism_kypass.hip -- complex DFT along y of the half spectrum (passes 2, 4).
#include "rism_common.h"

namespace rism {

// ---------------------------------------------------------------------------
// 2/4. complex DFT along y of the half spectrum (strided columns)
// ---------------------------------------------------------------------------
// Forward: pre-twiddle a_n = (-1)^n e^{-i pi n/ny}, DFT.  Inverse: DFT^-1
// (unnormalised), post-twiddle conj(a_n).  The output phases are "raw".
template <int L, int B, int S>
__global__ void __launch_bounds__(B *FftShape<L>::T)
    k_ypass(cplx *__restrict__ spec, int nx, const cplx *__restrict__ twL,
            const cplx *__restrict__ pre) {
  constexpr int CS = FftShape<L>::CS;
  __shared__ cplx lds[B * CS];
  const int tid = threadIdx.x;
  const int c = tid % B, j = tid / B;
  const int x = blockIdx.x * B + c;
  const bool ok = x < nx;
  cplx *col = spec + (size_t)blockIdx.y * L * nx + (ok ? x : 0);
  auto load = [&](int i) -> cplx {
    if (!ok) return mk(0.0, 0.0);
    const cplx vv = col[(size_t)i * nx];
    return (S < 0) ? cmul(vv, pre[i]) : vv;
  };
  auto store = [&](int i, cplx vv) {
    if (ok) col[(size_t)i * nx] = (S < 0) ? vv : cmulc(vv, pre[i]);
  };
  fft_column<L, S, true, true>(lds + c * CS, j, twL, load, store);
}

template <int L, int S>
static void launch_y(const Params &P, cplx *spec, int nplanes, const cplx *twL,
                     const cplx *pre, hipStream_t st) {
  constexpr int B = tile_cols<L>();
  dim3 grid((P.nx + B - 1) / B, nplanes);
  hipLaunchKernelGGL((k_ypass<L, B, S>), grid, dim3(B * FftShape<L>::T), 0, st, spec, P.nx,
                     twL, pre);
}
hipError_t run_y(const Params &P, cplx *spec, int dir, const cplx *twL, const cplx *pre,
                 hipStream_t st) {
  const int nplanes = P.natv * (P.nz / 2);
  switch (P.ny) {
#define X(n)                                           \
  case n:                                              \
    if (dir < 0)                                       \
      launch_y<n, -1>(P, spec, nplanes, twL, pre, st); \
    else                                               \
      launch_y<n, 1>(P, spec, nplanes, twL, pre, st);  \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace rism
