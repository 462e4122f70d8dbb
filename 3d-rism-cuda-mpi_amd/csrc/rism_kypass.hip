// rism_kypass.hip -- complex DFT along y of the half spectrum (passes 2, 4).
#include "rism_common.h"

namespace rism {

// ---------------------------------------------------------------------------
// 2/4. complex DFT along y of the half spectrum (strided columns)
// ---------------------------------------------------------------------------
// Forward: pre-twiddle a_n = (-1)^n e^{-i pi n/ny}, DFT.  Inverse: DFT^-1
// (unnormalised), post-twiddle conj(a_n).  The output phases are "raw".
template <int L, int B, int S, class Re>
__global__ void __launch_bounds__(B *FftShape<L>::T)
    k_ypass(cx<Re> *__restrict__ spec, int nx, const cx<Re> *__restrict__ twL,
            const cx<Re> *__restrict__ pre, const int *done) {
  using C = cx<Re>;
  constexpr int CS = FftShape<L>::CS;
  __shared__ C lds[B * CS];
  if (done && *done) return;
  const int tid = threadIdx.x;
  const int c = tid % B, j = tid / B;
  const int x = blockIdx.x * B + c;
  const bool ok = x < nx;
  C *col = spec + (size_t)blockIdx.y * L * nx + (ok ? x : 0);
  // the spectrum is read once and written once per pass: non-temporal
  // accesses leave the caches to the twiddle tables
  auto load = [&](int i) -> C {
    if (!ok) return C{0, 0};
    const C *q = col + (size_t)i * nx;
    const C vv{__builtin_nontemporal_load(&q->x), __builtin_nontemporal_load(&q->y)};
    return (S < 0) ? cmul(vv, pre[i]) : vv;
  };
  auto store = [&](int i, C vv) {
    if (!ok) return;
    const C o = (S < 0) ? vv : cmulc(vv, pre[i]);
    C *q = col + (size_t)i * nx;
    __builtin_nontemporal_store(o.x, &q->x);
    __builtin_nontemporal_store(o.y, &q->y);
  };
  fft_column<L, S, true, true>(lds + c * CS, j, twL, load, store);
}

template <int L, int S, class Re>
static void launch_y(const Params &P, cx<Re> *spec, int nplanes, const cx<Re> *twL,
                     const cx<Re> *pre, hipStream_t st) {
  constexpr int B = tile_cols<L, Re>();
  dim3 grid((P.nx + B - 1) / B, nplanes);
  hipLaunchKernelGGL((k_ypass<L, B, S, Re>), grid, dim3(B * FftShape<L>::T), 0, st, spec, P.nx,
                     twL, pre, P.done);
}
template <class Re>
hipError_t run_y(const Params &P, cx<Re> *spec, int dir, const cx<Re> *twL, const cx<Re> *pre,
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
template hipError_t run_y<double>(const Params &, cplx *, int, const cplx *, const cplx *,
                                  hipStream_t);
template hipError_t run_y<float>(const Params &, cplxf *, int, const cplxf *, const cplxf *,
                                 hipStream_t);

}  // namespace rism
