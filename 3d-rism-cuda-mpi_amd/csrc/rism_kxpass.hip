// rism_kxpass.hip -- pass 3 of one step (x DFT + OZ + inverse x DFT) and
// the layout conventions shared by all hand-written CDNA4 (gfx950) kernels.
//
// One fixed-point step (reference: RISM3D::calculate, calculate.cu:4-65,
// and RISM3D::cal_rms, cal_rms.cu:14-26) runs as five passes over HBM:
//
//   1. zfwd_closure  closure (KH/HNC) + c_sr + real-to-half-complex DFT along
//                    z, per site   (kh/hnc/trm1mt + 1/3 of FFT3D forward)
//   2. ypass<-1>     complex DFT along y                 (1/3 of forward)
//   3. xpass_oz      DFT along x, -q_v fk, site-site OZ contraction with
//                    chi_vv(|k|), inverse DFT along x    (mqvfk + oz)
//   4. ypass<+1>     inverse DFT along y
//   5. zinv_resid    half-complex-to-real inverse DFT along z, closure again,
//                    residual R = h + 1 - g, FP64 partial sums of R.R and of
//                    the AN2 Gram entries                (tr + theta30)
//
// followed by an elementwise AN2 update (anderson_cuda.cu:1-22).
//
// Transform convention (fft3d.cu:81-163, fft3d_cuda.cu:56-128, with the
// S = +1 phase): F(k) = dv sum_r f(r) e^{-ik.r}, f(r) = V^-1 sum_k F(k) e^{ik.r},
// r_n = (n - N/2) L/N, k_j = 2 pi (j - N/2 + 1/2)/L.  Because f is real,
// only the k_z > 0 half is stored.  Spectra in HBM are kept "raw": the
// unit-modulus per-axis output phases phi(k) of the exact transform are left
// out (they cancel between the forward and the inverse, and the OZ operator is
// real), and the stored fk is multiplied by conj(phi) once at set-up.
//
// Data layout: real fields [site][z][y][x] (x fastest) in FP64; spectra
// [site][m][y][x] complex FP64 with m in [0, nz/2) <-> k_z index m + nz/2.

#include "rism_common.h"

namespace rism {

// ---------------------------------------------------------------------------
// 3. DFT along x, OZ with chi_vv(|k|), inverse DFT along x
// ---------------------------------------------------------------------------
// chi_vv(|k|) sources for the OZ step.  ChiTable: the cubic-box shell table
// (one row of natv^2 values per k^2 shell, built once by k_chi_table);
// ChiSpline: the natural cubic spline of the .xvk table evaluated at |k|
// (solvent_spline.cu:44-55 evaluates the same spline at the same |k| shell
// values; splint.cc:4-20 picks the interval).
template <class Tb>
struct ChiTable {
  const Tb *row;
  __device__ __forceinline__ Tb operator()(int pw) const { return row[pw]; }
};
struct ChiSpline {
  const double *cf;
  double tt;
  int nv2;
  __device__ __forceinline__ double operator()(int pw) const {
    return ((cf[3 * nv2 + pw] * tt + cf[2 * nv2 + pw]) * tt + cf[nv2 + pw]) * tt + cf[pw];
  }
};

// OZ at NP points that share |k| and therefore chi (here (kx, ky) and
// (-kx, ky) of one row): H_v = sum_w X[v][w] (C_w - q_w fk)
// (mqvfk + oz, calculate_cuda.cu:32-52).  off[k] is the LDS offset of point
// k in site 0's column; site w sits w*CS further.  NVM is a compile-time
// bound on the site count so the per-site values stay in registers.
template <int NVM, int NP, class Re, class Chi>
__device__ __forceinline__ void oz_points(cx<Re> *lds, int CS, const int (&off)[NP],
                                          const cx<Re> (&fk)[NP], const Params &P, int nv,
                                          const Chi &X) {
  cx<Re> c[NP][NVM];
#pragma unroll
  for (int w = 0; w < NVM; ++w) {
    if (w < nv) {
#pragma unroll
      for (int k = 0; k < NP; ++k) c[k][w] = lds[off[k] + w * CS] - (Re)P.qv[w] * fk[k];
    }
  }
  // one output site at a time keeps the live chi values to one row of NVM
#pragma unroll 1
  for (int v = 0; v < NVM; ++v) {
    if (v < nv) {
      cx<Re> a[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) a[k] = cx<Re>{0, 0};
#pragma unroll
      for (int w = 0; w < NVM; ++w) {
        if (w < nv) {
          const Re x = (Re)X(v * nv + w);
#pragma unroll
          for (int k = 0; k < NP; ++k) a[k] = a[k] + x * c[k][w];
        }
      }
#pragma unroll
      for (int k = 0; k < NP; ++k) lds[off[k] + v * CS] = a[k];
    }
  }
}

// NVMAX (4 or 8) bounds the site counts compiled into one kernel: the
// register allocation follows the largest case, so natv <= 4 gets a kernel
// without the 5..8-site paths.
template <int NVMAX, int NP, class Re, class Chi>
__device__ __forceinline__ void oz_dispatch(cx<Re> *lds, int CS, const int (&off)[NP],
                                            const cx<Re> (&fk)[NP], const Params &P,
                                            const Chi &X) {
  switch (P.natv) {
    case 1: oz_points<1>(lds, CS, off, fk, P, 1, X); break;
    case 2: oz_points<2>(lds, CS, off, fk, P, 2, X); break;
    case 3: oz_points<3>(lds, CS, off, fk, P, 3, X); break;
    case 4: oz_points<4>(lds, CS, off, fk, P, 4, X); break;
    default:
      if constexpr (NVMAX > 4) {
        switch (P.natv) {
          case 5: oz_points<5>(lds, CS, off, fk, P, 5, X); break;
          case 6: oz_points<6>(lds, CS, off, fk, P, 6, X); break;
          case 7: oz_points<7>(lds, CS, off, fk, P, 7, X); break;
          default: oz_points<8>(lds, CS, off, fk, P, 8, X); break;
        }
      }
      break;
  }
}

template <int NVMAX, class Re, class Chi>
__device__ __forceinline__ void oz_row(cx<Re> *lds, int CS, int jx, int L, const cx<Re> *fkr,
                                       const Params &P, const Chi &X) {
  const int jm = L - 1 - jx;
  const int off[2] = {lpad(jx), lpad(jm)};
  const cx<Re> fk[2] = {fkr ? fkr[jx] : cx<Re>{0, 0}, fkr ? fkr[jm] : cx<Re>{0, 0}};
  oz_dispatch<NVMAX, 2>(lds, CS, off, fk, P, X);
}

// The same contraction with the chi row already in registers (chr[v][w],
// v, w < NV): used when every thread owns at most one (kx, -kx) pair, so the
// row is fetched before the forward transform and its latency is hidden.
template <int NV, class Re>
__device__ __forceinline__ void oz_points_reg(cx<Re> *lds, int CS, const int (&off)[2],
                                              const cx<Re> (&fk)[2], const Params &P,
                                              const Re (&chr)[4][4]) {
  cx<Re> c[2][NV];
#pragma unroll
  for (int w = 0; w < NV; ++w) {
#pragma unroll
    for (int k = 0; k < 2; ++k) c[k][w] = lds[off[k] + w * CS] - (Re)P.qv[w] * fk[k];
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      cx<Re> a{0, 0};
#pragma unroll
      for (int w = 0; w < NV; ++w) a = a + chr[v][w] * c[k][w];
      lds[off[k] + v * CS] = a;
    }
  }
}

// One workgroup per m and pair of mirror rows (y, ny-1-y) -- several pairs
// side by side for short rows, see the row slots below -- all sites: natv
// columns of nx complex in LDS per pair, natv * T threads per pair, the two
// rows one after the other.  The rows of a pair have the same |ky| and so the same chi rows
// (the shell index and |k| depend on ky^2 only): the register path loads
// them once for both rows.  SHELL selects the chi source at compile time (the
// cubic-box shell table or the spline at |k|) and NVMAX bounds the site
// count, so each instantiation carries only the registers it needs.
// dynamic LDS of the x passes, shared by all instantiations; 16-byte
// aligned, so FP64 complex values move as ds_read_b128 / ds_write_b128
// (with the default 8-byte alignment they split into ds_read2_b64 pairs,
// with about eight bank-conflict cycles per LDS instruction)
extern __shared__ __attribute__((aligned(16))) double xsm[];

template <int L, int NVMAX, bool SHELL, class Re>
__global__ void __launch_bounds__(512)
    k_xpass_oz(cx<Re> *__restrict__ spec, const cx<Re> *__restrict__ fkraw, Params P,
               const cx<Re> *__restrict__ twL, const cx<Re> *__restrict__ prex,
               const double *__restrict__ knots, const double *__restrict__ coef,
               const Re *__restrict__ chitab) {
  using C = cx<Re>;
  constexpr int T = FftShape<L>::T, CS = FftShape<L>::CS;
  C *lds = reinterpret_cast<C *>(xsm);
  if (P.grp) {  // grouped launch: this context's spectrum and fk
    const GroupPtrs &g = P.grp[blockIdx.z];
    if (g.done && *g.done) return;
    spec = (C *)g.spec;
    fkraw = (const C *)g.fk;
  } else if (P.done && *P.done) return;
  const int tid = threadIdx.x;
  const int nv = P.natv;
  // row slots: a workgroup of short rows (T < 64) holds several row pairs
  // side by side (one wave at least), each slot with its own natv LDS
  // columns; with T >= 64 a workgroup is one slot, and the row indices stay
  // workgroup-uniform (scalar branches, the code of one row per workgroup)
  constexpr bool kSlots = T < 64;
  const int nthr = nv * T;  // threads of one row slot
  const int slot = kSlots ? tid / nthr : 0, tl = kSlots ? tid % nthr : tid;
  const int v = tl / T, j = tl % T;
  const int y0 = kSlots ? blockIdx.x * (int)(blockDim.x / nthr) + slot : (int)blockIdx.x;
  const int m = blockIdx.y;
  const bool live = !kSlots || y0 < (P.ny + 1) / 2;  // the last workgroup's slots may be spare
  const int y1 = P.ny - 1 - y0;                       // the mirror row
  const bool pair = live && y1 > y0;                  // not for the middle row of odd ny
  const int M = P.nz / 2;
  const size_t rowstride = (size_t)M * P.ny * L;  // site stride in spec
  C *slds = lds + slot * nv * CS;
  C *mylds = slds + v * CS;
  // the spectrum row is read once and written once: non-temporal accesses
  // leave the caches to the chi table, fk and the twiddles.  The element
  // indices of the first and last FFT stages are j + c with c the same for
  // every thread, and the pre-twiddle is an exponential, so prex[j + c] =
  // prex[j] prex[c]: one per-thread load, the rest uniform (scalar) loads.
  const C prej = prex[j];
  auto nostore = [&](int, C) {};
  const int nv2 = nv * nv;
  const long oy = 2 * y0 - P.ny + 1, oz = 2 * m + 1;  // odd lattice coordinates (|oy| of both rows)
  // With several slots both rows are run by every slot (a spare row
  // transforms zeros and stores nothing), so that all threads meet the same
  // barriers.
  // at most one (kx, -kx) pair per thread and at most 4 sites: fetch this
  // thread's chi row (shared by the two rows) and the fk pair of each row
  // before its forward transform
  if constexpr (SHELL && NVMAX == 4) {
    if (L / 2 <= nthr) {
      Re chr[4][4];
      const bool mine = live && tl < L / 2;
      if (mine) {
        const long ox = 2 * tl - L + 1;
        const long sidx = (ox * ox + oy * oy + oz * oz - 3) >> 3;
        const Re *rowp = chitab + sidx * nv2;
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int w = 0; w < 4; ++w) chr[v][w] = (v < nv && w < nv) ? rowp[v * nv + w] : (Re)0;
      }
#pragma unroll 1
      for (int r = 0; r < 2; ++r) {
        // an opaque zero: the table loads of the row's transforms stay in the
        // loop (hoisted out of it, they held ~100 more VGPRs)
        int zero = 0;
        asm volatile("" : "+s"(zero));
        const C *tw = twL + zero, *px = prex + zero;
        // a spare row: the workgroup stops before it (one slot), or the slot
        // runs it on zeros (several slots: per-element guards)
        const bool rowok = r ? pair : live;
        if (!kSlots && !rowok) break;
        const size_t rowoff = rowok ? ((size_t)m * P.ny + (r ? y1 : y0)) * L : 0;
        C *row = spec + (size_t)v * rowstride + rowoff;
        auto load = [&](int i) -> C {
          // rowoff is 0 for a spare row: the load stays in bounds, unguarded
          const C a{__builtin_nontemporal_load(&row[i].x), __builtin_nontemporal_load(&row[i].y)};
          const C w = cmul(a, cmul(prej, px[i - j]));
          return (kSlots && !rowok) ? C{0, 0} : w;
        };
        auto store = [&](int i, C vv) {
          if (kSlots && !rowok) return;
          const C o = cmulc(vv, cmul(prej, px[i - j]));
          __builtin_nontemporal_store(o.x, &row[i].x);
          __builtin_nontemporal_store(o.y, &row[i].y);
        };
        // the difference-form step (mode 3): the long-range term cancels, fk = 0
        const C *fkr = (P.mode == 3) ? nullptr : fkraw + rowoff;
        C fkp[2] = {C{0, 0}, C{0, 0}};
        if (mine && rowok && fkr) {
          fkp[0] = fkr[tl];
          fkp[1] = fkr[L - 1 - tl];
        }
        if (r) __syncthreads();  // the first row's inverse stage has read its LDS
        fft_column<L, -1, true, false>(mylds, j, tw, load, nostore);
        if (mine && rowok) {
          const int off[2] = {lpad(tl), lpad(L - 1 - tl)};
          switch (nv) {
            case 1: oz_points_reg<1>(slds, CS, off, fkp, P, chr); break;
            case 2: oz_points_reg<2>(slds, CS, off, fkp, P, chr); break;
            case 3: oz_points_reg<3>(slds, CS, off, fkp, P, chr); break;
            default: oz_points_reg<4>(slds, CS, off, fkp, P, chr); break;
          }
        }
        __syncthreads();
        fft_column<L, +1, false, true>(mylds, j, tw, load, store);
      }
      return;
    }
  }
#pragma unroll 1
  for (int r = 0; r < 2; ++r) {
    int zero = 0;  // an opaque zero, as above
    asm volatile("" : "+s"(zero));
    const C *tw = twL + zero, *px = prex + zero;
    const bool rowok = r ? pair : live;
    if (!kSlots && !rowok) break;  // as above
    const int y = r ? y1 : y0;
    const size_t rowoff = rowok ? ((size_t)m * P.ny + y) * L : 0;
    C *row = spec + (size_t)v * rowstride + rowoff;
    auto load = [&](int i) -> C {
      // rowoff is 0 for a spare row: the load stays in bounds, unguarded
      const C a{__builtin_nontemporal_load(&row[i].x), __builtin_nontemporal_load(&row[i].y)};
      const C w = cmul(a, cmul(prej, px[i - j]));
      return (kSlots && !rowok) ? C{0, 0} : w;
    };
    auto store = [&](int i, C vv) {
      if (kSlots && !rowok) return;
      const C o = cmulc(vv, cmul(prej, px[i - j]));
      __builtin_nontemporal_store(o.x, &row[i].x);
      __builtin_nontemporal_store(o.y, &row[i].y);
    };
    const C *fkr = (P.mode == 3) ? nullptr : fkraw + rowoff;
    if (r) __syncthreads();  // the first row's inverse stage has read its LDS
    fft_column<L, -1, true, false>(mylds, j, tw, load, nostore);

    // OZ at every k of the row, the mirror points (kx, -kx) together
    for (int jx = tl; rowok && jx < L / 2; jx += nthr) {
      if constexpr (SHELL) {
        const long ox = 2 * jx - L + 1;
        const long sidx = (ox * ox + oy * oy + oz * oz - 3) >> 3;  // |k|^2 = (pi/L)^2 (8 sidx + 3)
        oz_row<NVMAX>(slds, CS, jx, L, fkr, P, ChiTable<Re>{chitab + sidx * nv2});
      } else {
        const double ky = P.kfac[1] * ((double)y - (double)P.ny / 2.0 + 0.5);
        const double kz = P.kfac[2] * ((double)(m + P.nz / 2) - (double)P.nz / 2.0 + 0.5);
        const double kx = P.kfac[0] * ((double)jx - (double)L / 2.0 + 0.5);
        const double kk = sqrt(kx * kx + ky * ky + kz * kz);
        const int ii = spline_interval(knots, P.np, P.x0, P.idk, kk);
        oz_row<NVMAX>(slds, CS, jx, L, fkr, P,
                      ChiSpline{coef + (size_t)ii * 4 * nv2, kk - knots[ii], nv2});
      }
    }
    __syncthreads();
    fft_column<L, +1, false, true>(mylds, j, tw, load, store);
  }
}

// Plain transform along x of the half spectrum (kernel-level entry points):
// S = -1 forward (pre-twiddle, DFT), S = +1 inverse (DFT^-1, post-twiddle).
template <int L, int S>
__global__ void __launch_bounds__(512)
    k_xpass_plain(cplx *__restrict__ spec, Params P, const cplx *__restrict__ twL,
                  const cplx *__restrict__ prex) {
  constexpr int T = FftShape<L>::T, CS = FftShape<L>::CS;
  cplx *lds = reinterpret_cast<cplx *>(xsm);
  const int v = threadIdx.x / T, j = threadIdx.x % T;
  const size_t rowoff = ((size_t)blockIdx.y * P.ny + blockIdx.x) * L;
  cplx *row = spec + (size_t)v * (P.nz / 2) * P.ny * L + rowoff;
  auto load = [&](int i) -> cplx { return S < 0 ? cmul(row[i], prex[i]) : row[i]; };
  auto store = [&](int i, cplx vv) { row[i] = S < 0 ? vv : cmulc(vv, prex[i]); };
  fft_column<L, S, true, true>(lds + v * CS, j, twL, load, store);
}

// chi_vv on every k^2 shell of a cubic box: |k| = (pi/L) sqrt(8 i + 3), i.e.
// the shells of the reference's sorted k^2 list (initialize_g.cu:26-45) with
// the spline of solvent_spline.cu:44-55 evaluated at each.
__global__ void k_chi_table(const double *__restrict__ knots, const double *__restrict__ coef,
                            Params P, double kscale, size_t nshell, double *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshell) return;
  const double kk = kscale * sqrt((double)(8 * i + 3));
  const int ii = spline_interval(knots, P.np, P.x0, P.idk, kk);
  const int nv2 = P.natv * P.natv;
  const ChiSpline X{coef + (size_t)ii * 4 * nv2, kk - knots[ii], nv2};
  for (int pw = 0; pw < nv2; ++pw) out[i * nv2 + pw] = X(pw);
}

template <int L, class Re>
static void launch_xoz(const Params &P, cx<Re> *spec, const cx<Re> *fkraw, const cx<Re> *twL,
                       const cx<Re> *prex, const double *knots, const double *coef,
                       const Re *chitab, hipStream_t st) {
  dim3 grid(P.ny, P.nz / 2, P.grp ? P.ngrp : 1);
  const int threads = P.natv * FftShape<L>::T;
  const size_t lds = (size_t)P.natv * FftShape<L>::CS * sizeof(cx<Re>);
  // the OZ pass takes the rows in mirror pairs (y, ny-1-y), and short rows
  // several pairs to a workgroup (row slots), so that it fills a wave
  const int slots = (FftShape<L>::T >= 64 || threads >= 64) ? 1 : 64 / threads;
  const dim3 pgrid(((P.ny + 1) / 2 + slots - 1) / slots, P.nz / 2, P.grp ? P.ngrp : 1);
  const int pthreads = slots * threads;
  const size_t plds = (size_t)slots * lds;
  if constexpr (sizeof(Re) == sizeof(double)) {
    if (P.mode == 1 || P.mode == 2) {
      if (P.mode == 1)
        hipLaunchKernelGGL((k_xpass_plain<L, -1>), grid, dim3(threads), lds, st, spec, P, twL,
                           prex);
      else
        hipLaunchKernelGGL((k_xpass_plain<L, 1>), grid, dim3(threads), lds, st, spec, P, twL,
                           prex);
      return;
    }
  }
  auto go = [&](auto kern) {
    if (plds > 65536)
      (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)plds);
    hipLaunchKernelGGL(kern, pgrid, dim3(pthreads), plds, st, spec, fkraw, P, twL, prex, knots,
                       coef, chitab);
  };
  if (P.natv <= 4)
    P.shell ? go(k_xpass_oz<L, 4, true, Re>) : go(k_xpass_oz<L, 4, false, Re>);
  else
    P.shell ? go(k_xpass_oz<L, 8, true, Re>) : go(k_xpass_oz<L, 8, false, Re>);
}
int xoz_threads(int L, int natv) {
  switch (L) {
#define X(n) \
  case n:    \
    return natv * FftShape<n>::T;
    RISM_LENGTHS(X)
#undef X
    default:
      return 0;
  }
}
template <class Re>
hipError_t run_xoz(const Params &P, cx<Re> *spec, const cx<Re> *fkraw, const cx<Re> *twL,
                   const cx<Re> *prex, const double *knots, const double *coef,
                   const Re *chitab, hipStream_t st) {
  switch (P.nx) {
#define X(n)                                                           \
  case n:                                                              \
    launch_xoz<n>(P, spec, fkraw, twL, prex, knots, coef, chitab, st); \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template hipError_t run_xoz<double>(const Params &, cplx *, const cplx *, const cplx *,
                                    const cplx *, const double *, const double *, const double *,
                                    hipStream_t);
template hipError_t run_xoz<float>(const Params &, cplxf *, const cplxf *, const cplxf *,
                                   const cplxf *, const double *, const double *,
                                   const float *, hipStream_t);
hipError_t run_chi_table(const double *knots, const double *coef, const Params &P,
                          double kscale, size_t nshell, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_chi_table, dim3((unsigned)((nshell + 255) / 256)), dim3(256), 0, st, knots,
                     coef, P, kscale, nshell, out);
  return hipGetLastError();
}

}  // namespace rism
