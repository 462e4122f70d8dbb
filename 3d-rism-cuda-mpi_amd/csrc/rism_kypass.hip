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
            const cx<Re> *__restrict__ pre, const int *done, const GroupPtrs *grp) {
  using C = cx<Re>;
  constexpr int CS = FftShape<L>::CS;
  __shared__ C lds[B * CS];
  if (grp) {  // grouped launch: this context's spectrum
    const GroupPtrs &g = grp[blockIdx.z];
    done = g.done;
    spec = (C *)g.spec;
  }
  if (done && *done) return;
  const int tid = threadIdx.x;
  const int c = tid % B, j = tid / B;
  const int x = blockIdx.x * B + c;
  const bool ok = x < nx;
  C *col = spec + (size_t)blockIdx.y * L * nx + (ok ? x : 0);
  // the spectrum is read once and written once per pass: non-temporal
  // accesses leave the caches to the twiddle tables
  // pre[j + c] = pre[j] pre[c] (an exponential; c is the same for every
  // thread in the first and last stages): one per-thread load
  const C prej = pre[j];
  // col points at column 0 when x is past nx, so every load stays in bounds
  // and is issued unguarded: a load guarded by `ok` waits for its data inside
  // the guard, one element after another
  auto load = [&](int i) -> C {
    const C *q = col + (size_t)i * nx;
    const C vv{__builtin_nontemporal_load(&q->x), __builtin_nontemporal_load(&q->y)};
    const C w = (S < 0) ? cmul(vv, cmul(prej, pre[i - j])) : vv;
    return ok ? w : C{0, 0};
  };
  auto store = [&](int i, C vv) {
    if (!ok) return;
    const C o = (S < 0) ? vv : cmulc(vv, cmul(prej, pre[i - j]));
    C *q = col + (size_t)i * nx;
    __builtin_nontemporal_store(o.x, &q->x);
    __builtin_nontemporal_store(o.y, &q->y);
  };
  fft_column<L, S, true, true>(lds + c * CS, j, twL, load, store);
}

// Launch plans of the y passes (chosen per grid by the autotuner): 0 the
// default tile_cols columns, 1 half as many, 2 twice as many (when the
// workgroup and its LDS tile still fit: 1024 threads, 160 KB).  A plan
// narrower than 64 bytes of contiguous x per row is not offered.
template <int L, class Re, int V>
struct YPlan {
  static constexpr int B0 = tile_cols<L, Re>();
  static constexpr int B = V == 0 ? B0 : (V == 1 ? (B0 > 1 ? B0 / 2 : 1) : 2 * B0);
};
template <int L, class Re, int V>
constexpr bool y_offered() {
  if constexpr (V == 0) {
    return true;
  } else {
    constexpr int B = YPlan<L, Re, V>::B;
    return L >= 48 && B != YPlan<L, Re, 0>::B && B * (int)sizeof(cx<Re>) >= 64 &&
           B * FftShape<L>::T <= 1024 &&
           (long)B * FftShape<L>::CS * (long)sizeof(cx<Re>) <= 160 * 1024;
  }
}

template <int L, int S, class Re, int V = 0>
static void launch_y(const Params &P, cx<Re> *spec, int nplanes, const cx<Re> *twL,
                     const cx<Re> *pre, hipStream_t st) {
  constexpr int B = YPlan<L, Re, V>::B;
  dim3 grid((P.nx + B - 1) / B, nplanes, P.grp ? P.ngrp : 1);
  hipLaunchKernelGGL((k_ypass<L, B, S, Re>), grid, dim3(B * FftShape<L>::T), 0, st, spec, P.nx,
                     twL, pre, P.done, P.grp);
}
template <int L, int S, class Re>
static void launch_y_plan(const Params &P, cx<Re> *spec, int nplanes, const cx<Re> *twL,
                          const cx<Re> *pre, hipStream_t st) {
  const int v = P.var_y[sizeof(Re) == 4];
  if constexpr (y_offered<L, Re, 1>()) {
    if (v == 1) return launch_y<L, S, Re, 1>(P, spec, nplanes, twL, pre, st);
  }
  if constexpr (y_offered<L, Re, 2>()) {
    if (v == 2) return launch_y<L, S, Re, 2>(P, spec, nplanes, twL, pre, st);
  }
  launch_y<L, S, Re>(P, spec, nplanes, twL, pre, st);
}
template <class Re>
hipError_t run_y(const Params &P, cx<Re> *spec, int dir, const cx<Re> *twL, const cx<Re> *pre,
                 hipStream_t st) {
  const int nplanes = P.natv * (P.nz / 2);
  switch (P.ny) {
#define X(n)                                                \
  case n:                                                   \
    if (dir < 0)                                            \
      launch_y_plan<n, -1>(P, spec, nplanes, twL, pre, st); \
    else                                                    \
      launch_y_plan<n, 1>(P, spec, nplanes, twL, pre, st);  \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <class Re>
static bool y_offered_at(int L, int v) {
  switch (L) {
#define X(n)                                                                      \
  case n:                                                                         \
    return v == 0 ? true                                                          \
                  : (v == 1 ? y_offered<n, Re, 1>() : (v == 2 ? y_offered<n, Re, 2>() : false));
    RISM_LENGTHS(X)
#undef X
    default:
      return false;
  }
}
bool y_plan_offered(int L, int fp32, int v) {
  return fp32 ? y_offered_at<float>(L, v) : y_offered_at<double>(L, v);
}
template hipError_t run_y<double>(const Params &, cplx *, int, const cplx *, const cplx *,
                                  hipStream_t);
template hipError_t run_y<float>(const Params &, cplxf *, int, const cplxf *, const cplxf *,
                                 hipStream_t);

}  // namespace rism
