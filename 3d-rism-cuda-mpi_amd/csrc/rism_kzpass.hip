// rism_kzpass.hip -- the z passes of one step: closure + c_sr + real to
// half-complex DFT along z (pass 1), and half-complex to real inverse DFT
// along z + residual + FP64 Gram partial sums (pass 5).  Conventions and the
// five-pass structure are described in rism_kxpass.hip.
#include "rism_common.h"

namespace rism {

// ---------------------------------------------------------------------------
// 1. closure + c_sr + real->half-complex DFT along z
// ---------------------------------------------------------------------------
// Column = the z-line at (x, y) of one site: nz reals packed as M = nz/2
// complex z_p = c(2p) + i c(2p+1), pre-twiddled by e^{-i pi p/M}; after the
// M-point DFT, G(m) = E(m) + w_m O(m) with E, O from Z(m) and Z(M-1-m) and
// w_m = e^{-2 pi i (m+1/2)/nz}: the z-DFT at the positive half-shifted
// frequencies.
template <int M, int B, int TZ, int MAXR, class Re>
__global__ void __launch_bounds__(B *TZ)
    __attribute__((amdgpu_waves_per_eu(MAXR == 4 ? 8 : 1)))
    k_zfwd_closure(const Re *__restrict__ t, const Re *__restrict__ u,
                   const Re *__restrict__ fr, Params P, cx<Re> *__restrict__ spec,
                   const cx<Re> *__restrict__ twM, const cx<Re> *__restrict__ zpre,
                   const cx<Re> *__restrict__ zpost) {
  using C = cx<Re>;
  constexpr int T = TZ, CS = FftShape<M>::CS;
  __shared__ C lds[B * CS];
  if (P.done && *P.done) return;
  const int tid = threadIdx.x;
  const int c = tid % B, j = tid / B;
  // logical workgroup order: site fastest, then x tile, then y (sites of
  // one tile share fr in L2)
  const int ntx = (P.nx + B - 1) / B;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int v = lb % P.natv;
  const int tx = (lb / P.natv) % ntx;
  const int y = lb / (P.natv * ntx);
  const int x = tx * B + c;
  const bool ok = x < P.nx;
  const size_t plane = (size_t)P.nx * P.ny;
  const size_t N = plane * (size_t)P.nz;
  const Re *tv = t + (size_t)v * N;
  const Re *uv = u + (size_t)v * N;
  const Re q = (Re)P.qv[v];
  const size_t base = (size_t)y * P.nx + (ok ? x : 0);
  const int clos = P.clos;
  const Re dv = (Re)P.dv;
  C *col = lds + c * CS;

  // t and u are read once per step: non-temporal loads keep them out of the
  // caches that fr (shared by the sites) and the twiddles use
  auto cval = [&](int z) -> Re {
    const size_t i = base + (size_t)z * plane;
    const Re tt = __builtin_nontemporal_load(&tv[i]);
    // difference form of a correction sweep: t holds the correction d, u
    // the exponent a = -u + t0 at the base point, and c(t0 + d) - c(t0) =
    // g(a + d) - g(a) - d (the q fr term cancels)
    const Re uu = __builtin_nontemporal_load(&uv[i]);
    if (P.mode == 3) return (closure_dg(clos, uu, tt) - tt) * dv;
    if (P.mode != 0) return tt * dv;  // plain transform of t
    const Re g = closure_g(clos, tt, uu);
    return (g - (Re)1 - tt + q * fr[i]) * dv;  // trm1mt, times the forward dv
  };
  auto load = [&](int p) -> C {
    if (!ok) return C{0, 0};
    return cmul(C{cval(2 * p), cval(2 * p + 1)}, zpre[p]);
  };
  auto store = [&](int, C) {};
  fft_column<M, -1, true, false, TZ, MAXR>(col, j, twM, load, store);

  C *out = spec + ((size_t)v * M * P.ny + y) * P.nx + x;
  const size_t mstride = (size_t)P.ny * P.nx;
  for (int m = j; m < M; m += T) {
    const C a = col[lpad(m)];
    const C b = conj(col[lpad(M - 1 - m)]);
    const C E = (Re)0.5 * (a + b);
    const C O = mul_si<-1>((Re)0.5 * (a - b));
    if (ok) out[(size_t)m * mstride] = E + cmul(zpost[m], O);
  }
}

// ---------------------------------------------------------------------------
// 5. half-complex->real inverse DFT along z, residual, FP64 partial sums
// ---------------------------------------------------------------------------
// Partial sums per workgroup, HM = 0 (AN2, 6 doubles): R.R, R.(R-R0),
// R.(R-R1), (R-R0)^2, (R-R0)(R-R1), (R-R1)^2 (anderson_cuda.cu:34-40,
// cal_rms.cu:20); HM = 1 (Anderson(m), 1 + RISM_MAX_HIST doubles): R.R and
// R.R_p for the nh history residuals.
template <int HM>
struct ZinvSums {
  static constexpr int W = HM == 0 ? 6 : 1 + RISM_MAX_HIST;
};

// Waves per SIMD the residual pass is compiled for.  An FP32 tile (<= 74 KB
// of LDS, 1024 threads) leaves room for a second workgroup per CU once the
// kernel fits in 64 VGPRs, and the FP64 Anderson(m) tile (512 threads, 74 KB)
// once it fits in 128; with one workgroup per CU the load, transform and
// residual phases of a tile do not overlap.  The compiler spills 36-116 B
// per lane to meet these limits, far cheaper than the lost overlap (512^3,
// 4 sites: FP32 4.7 -> 3.2 ms, FP64 Anderson(10) 24.4 -> 14.9 ms).  The FP64
// AN2 tile needs 148 KB of LDS and stays at one workgroup per CU.
template <int HM, class Re>
constexpr int zinv_waves_per_eu() {
  return sizeof(Re) == 4 ? 8 : (HM ? 4 : 1);
}

template <int M, int B, int TZ, int HM, class Re>
__global__ void __launch_bounds__(B *TZ)
    __attribute__((amdgpu_waves_per_eu(zinv_waves_per_eu<HM, Re>())))
    k_zinv_resid(const cx<Re> *__restrict__ spec, const Re *__restrict__ t,
                 const Re *__restrict__ u, Re *__restrict__ R, const Re *__restrict__ R0,
                 const Re *__restrict__ R1, const Re *__restrict__ r0, HistT<Re> H, Params P,
                 double *__restrict__ partial,
                 const cx<Re> *__restrict__ twM, const cx<Re> *__restrict__ zpre,
                 const cx<Re> *__restrict__ zpost, int use_hist) {
  using C = cx<Re>;
  constexpr int T = TZ, CS = FftShape<M>::CS;
  constexpr int NT = B * T;
  constexpr int W = ZinvSums<HM>::W;
  // the LDS tile doubles as the reduction buffer: W doubles per wave, or per
  // thread when the workgroup is not made of whole waves
  constexpr int NRED = (NT % 64 == 0) ? NT / 64 : NT;
  constexpr int NC = (int)((NRED * W * sizeof(double) + sizeof(C) - 1) / sizeof(C));
  constexpr int LDSN = (B * CS > NC) ? B * CS : NC;
  __shared__ C lds[LDSN];
  if (P.done && *P.done) return;
  const int tid = threadIdx.x;
  const int c = tid % B, j = tid / B;
  const int ntx = (P.nx + B - 1) / B;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int v = lb % P.natv;
  const int tx = (lb / P.natv) % ntx;
  const int y = lb / (P.natv * ntx);
  const int x = tx * B + c;
  const bool ok = x < P.nx;
  const size_t plane = (size_t)P.nx * P.ny;
  const size_t N = plane * (size_t)P.nz;
  C *col = lds + c * CS;
  const C *in = spec + ((size_t)v * M * P.ny + y) * P.nx + (ok ? x : 0);
  for (int m = j; m < M; m += T) col[lpad(m)] = ok ? in[(size_t)m * plane] : C{0, 0};
  __syncthreads();
  // G(m) -> Z(m) = E(m) + i O(m), pairs (m, M-1-m) handled by one thread
  for (int m = j; 2 * m <= M - 1; m += T) {
    const int mm = M - 1 - m;
    const C a = col[lpad(m)], b = col[lpad(mm)];
    const C Em = (Re)0.5 * (a + conj(b));
    const C Om = cmulc((Re)0.5 * (a - conj(b)), zpost[m]);
    const C Emm = (Re)0.5 * (b + conj(a));
    const C Omm = cmulc((Re)0.5 * (b - conj(a)), zpost[mm]);
    col[lpad(m)] = Em + mul_si<1>(Om);
    if (mm != m) col[lpad(mm)] = Emm + mul_si<1>(Omm);
  }
  __syncthreads();

  const Re *tv = t + (size_t)v * N;
  const Re *uv = u + (size_t)v * N;
  Re *Rv = R + (size_t)v * N;
  const Re *R0v = R0 + (size_t)v * N;
  const Re *R1v = R1 + (size_t)v * N;
  const size_t base = (size_t)y * P.nx + (ok ? x : 0);
  const int clos = P.clos;
  const Re scale = (Re)P.inv_scale;
  double s[W];
#pragma unroll
  for (int k = 0; k < W; ++k) s[k] = 0.0;
  const int plain = P.mode == 1 || P.mode == 2;
  const int lin = P.mode == 3;
  const size_t voff = (size_t)v * N;
  auto resid = [&](int z, Re h) {
    const size_t i = base + (size_t)z * plane;
    if (plain) {  // plain inverse transform: write h itself
      Rv[i] = h;
      return;
    }
    // difference form r0 + dh - dg, else tr (calculate_cuda.cu:58)
    const Re ti = __builtin_nontemporal_load(&tv[i]), ui = __builtin_nontemporal_load(&uv[i]);
    const Re rr = lin ? r0[voff + i] + h - closure_dg(clos, ui, ti)
                      : h + (Re)1 - closure_g(clos, ti, ui);
    __builtin_nontemporal_store(rr, &Rv[i]);
    // the sums are FP64 whatever the bulk type
    const double r = (double)rr;
    s[0] += r * r;
    if constexpr (HM == 0) {
      if (use_hist) {
        const double d0 = r - (double)__builtin_nontemporal_load(&R0v[i]),
                     d1 = r - (double)__builtin_nontemporal_load(&R1v[i]);
        s[1] += r * d0;
        s[2] += r * d1;
        s[3] += d0 * d0;
        s[4] += d0 * d1;
        s[5] += d1 * d1;
      }
    } else {
#pragma unroll
      for (int q = 0; q < RISM_MAX_HIST; ++q)
        if (q < H.nh) {
          // non-temporal in FP64 (-3 %); in the spilling FP32 tile it measured slower
          const Re hq = (sizeof(Re) == 8) ? __builtin_nontemporal_load(&H.r[q][voff + i])
                                          : H.r[q][voff + i];
          s[1 + q] += r * (double)hq;
        }
    }
  };
  auto load = [&](int) -> C { return C{0, 0}; };
  auto store = [&](int p, C zv) {
    if (!ok) return;
    const C zz = scale * cmulc(zv, zpre[p]);
    resid(2 * p, zz.x);
    resid(2 * p + 1, zz.y);
  };
  fft_column<M, +1, false, true, TZ>(col, j, twM, load, store);

  // workgroup reduction of the W sums: wave shuffles then LDS when the
  // workgroup is made of whole waves, LDS only otherwise (tiny grids)
  __syncthreads();
  double *red = reinterpret_cast<double *>(lds);
  int nparts;
  if constexpr (NT % 64 == 0) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
      double a = s[k];
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
      s[k] = a;
    }
    const int lane = tid & 63, wv = tid >> 6;
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < W; ++k) red[wv * W + k] = s[k];
    }
    nparts = NT / 64;
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) red[tid * W + k] = s[k];
    nparts = NT;
  }
  __syncthreads();
  if (tid < W) {
    double a = 0.0;
    for (int w = 0; w < nparts; ++w) a += red[w * W + tid];
    const size_t blk = (size_t)v * ntx * P.ny + (size_t)y * ntx + tx;  // [site][tile]
    partial[blk * W + tid] = a;
  }
}

static inline unsigned zgrid(const Params &P, int B) {
  return (unsigned)(((P.nx + B - 1) / B) * P.ny * P.natv);
}
template <int M, class Re>
static void launch_zfwd(const Params &P, const Re *t, const Re *u, const Re *fr, cx<Re> *spec,
                        const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost,
                        hipStream_t st) {
  using Z = ZTile<M, Re>;
  constexpr int B = Z::BF, TZ = Z::TF;
  hipLaunchKernelGGL((k_zfwd_closure<M, B, TZ, Z::MAXRF, Re>), dim3(zgrid(P, B)), dim3(B * TZ), 0,
                     st, t, u, fr, P, spec, twM, zpre, zpost);
}
template <int M, int HM, class Re>
static void launch_zinv(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                        const Re *R0, const Re *R1, const Re *r0, const HistT<Re> &H,
                        double *partial, const cx<Re> *twM, const cx<Re> *zpre,
                        const cx<Re> *zpost, int use_hist, hipStream_t st) {
  constexpr int B = HM ? ZTile<M, Re>::BH : ZTile<M, Re>::B, TZ = ZTile<M, Re>::T;
  hipLaunchKernelGGL((k_zinv_resid<M, B, TZ, HM, Re>), dim3(zgrid(P, B)), dim3(B * TZ), 0, st, spec,
                     t, u, R, R0, R1, r0, H, P, partial, twM, zpre, zpost, use_hist);
}
bool fft_length_supported(int L) {
  switch (L) {
#define X(n) case n:
    RISM_LENGTHS(X)
#undef X
    return true;
    default:
      return false;
  }
}
// Workgroups of one z pass; the same for both bulk types (the FP32 tiles are
// twice as wide, so pass the type the solver runs in).
template <class Re>
size_t zinv_num_blocks(const Params &P, int hm) {
  const int M = P.nz / 2;
  int B = 16;
  switch (M) {
#define X(n)                                         \
  case n:                                            \
    B = hm ? ZTile<n, Re>::BH : ZTile<n, Re>::B;     \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      break;
  }
  return (size_t)((P.nx + B - 1) / B) * P.ny * P.natv;
}
template <class Re>
hipError_t run_zfwd(const Params &P, const Re *t, const Re *u, const Re *fr, cx<Re> *spec,
                    const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost, hipStream_t st) {
  switch (P.nz / 2) {
#define X(n)                                                 \
  case n:                                                    \
    launch_zfwd<n>(P, t, u, fr, spec, twM, zpre, zpost, st); \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <class Re>
hipError_t run_zinv(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                    const Re *R0, const Re *R1, const Re *r0, double *partial,
                    const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost, int use_hist,
                    hipStream_t st) {
  if (P.mode == 3 && !r0) return hipErrorInvalidValue;
  switch (P.nz / 2) {
#define X(n)                                                                                   \
  case n:                                                                                      \
    launch_zinv<n, 0>(P, spec, t, u, R, R0, R1, r0, HistT<Re>{}, partial, twM, zpre, zpost,   \
                      use_hist, st);                                                           \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class Re>
hipError_t run_zinv_hist(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                         const Re *r0, const HistT<Re> &H, double *partial, const cx<Re> *twM,
                         const cx<Re> *zpre, const cx<Re> *zpost, hipStream_t st) {
  if (P.mode == 3 && !r0) return hipErrorInvalidValue;
  switch (P.nz / 2) {
#define X(n)                                                                                 \
  case n:                                                                                    \
    launch_zinv<n, 1>(P, spec, t, u, R, (const Re *)nullptr, (const Re *)nullptr, r0, H,     \
                      partial, twM, zpre, zpost, 0, st);                                     \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define INST(Re)                                                                              \
  template size_t zinv_num_blocks<Re>(const Params &, int);                                       \
  template hipError_t run_zfwd<Re>(const Params &, const Re *, const Re *, const Re *,       \
                                   cx<Re> *, const cx<Re> *, const cx<Re> *, const cx<Re> *, \
                                   hipStream_t);                                             \
  template hipError_t run_zinv<Re>(const Params &, const cx<Re> *, const Re *, const Re *,   \
                                   Re *, const Re *, const Re *, const Re *, double *,       \
                                   const cx<Re> *, const cx<Re> *, const cx<Re> *, int,      \
                                   hipStream_t);                                             \
  template hipError_t run_zinv_hist<Re>(const Params &, const cx<Re> *, const Re *,          \
                                        const Re *, Re *, const Re *, const HistT<Re> &,     \
                                        double *,                                            \
                                        const cx<Re> *, const cx<Re> *, const cx<Re> *,      \
                                        hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace rism
