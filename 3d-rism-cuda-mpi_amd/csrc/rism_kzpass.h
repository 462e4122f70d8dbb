// rism_kzpass.h -- the z passes of one step: closure + c_sr + real to
// half-complex DFT along z (pass 1), and half-complex to real inverse DFT
// along z + residual + FP64 Gram partial sums (pass 5).  Conventions and the
// five-pass structure are described in rism_kxpass.hip.  Kernel templates
// and launchers; rism_kzpass64.hip and rism_kzpass32.hip instantiate them
// for the FP64 and FP32 bulk types (two translation units that compile in
// parallel).
#pragma once
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
// What a z pass computes, fixed at compile time so that each instantiation
// carries only its own arithmetic and registers (a runtime switch between
// them cost 12 % in the residual pass).
enum ZKind {
  ZK_STEP = 0,    // solver step: closure + c_sr (forward); residual + Gram sums (inverse)
  ZK_PICARD = 1,  // inverse only: residual with R.R alone (no history yet)
  ZK_DIFF = 2,    // difference form of a correction sweep (mixed solve, FP32 level)
  ZK_PLAIN = 3,   // plain transform of t (forward) or to h (inverse)
};

// x when A, else y (two objects of different types)
template <bool A, class X, class Y>
__device__ __forceinline__ auto &pick_ref(X &x, Y &y) {
  if constexpr (A)
    return x;
  else
    return y;
}

// WPE: waves per SIMD the kernel is compiled for (register budget)
template <int M, int B, int TZ, int MAXR, int KIND, class Re, int WPE = (MAXR == 4 ? 8 : 1)>
__global__ void __launch_bounds__(B *TZ)
    __attribute__((amdgpu_waves_per_eu(WPE)))
    k_zfwd_closure(const Re *__restrict__ t, const Re *__restrict__ u,
                   const Re *__restrict__ fr, Params P, cx<Re> *__restrict__ spec,
                   const cx<Re> *__restrict__ twM, const cx<Re> *__restrict__ zpre,
                   const cx<Re> *__restrict__ zpost) {
  using C = cx<Re>;
  constexpr int T = TZ, CS = FftShape<M>::CS;
  __shared__ C lds[B * CS];
  if (P.grp) {  // grouped launch: this context's fields
    const GroupPtrs &g = P.grp[blockIdx.z];
    if (g.done && *g.done) return;
    t = (const Re *)g.t;
    u = (const Re *)g.u;
    fr = (const Re *)g.fr;
    spec = (C *)g.spec;
  } else if (P.done && *P.done) return;
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
  // caches that fr (shared by the sites) and the twiddles use.  (Issuing the
  // loads of several elements before their closures -- the staged protocol
  // of rism_fft.h -- spills in the 64-VGPR tiles and measured 15-40 % slower
  // there; the tiles without a register limit use it, below.)
  auto cval = [&](int z) -> Re {
    const size_t i = base + (size_t)z * plane;
    const Re tt = __builtin_nontemporal_load(&tv[i]);
    if constexpr (KIND == ZK_PLAIN) {
      return tt * dv;  // plain transform of t
    } else {
      const Re uu = __builtin_nontemporal_load(&uv[i]);
      // difference form of a correction sweep: t holds the correction d, u
      // the exponent a = -u + t0 at the base point, and c(t0 + d) - c(t0) =
      // g(a + d) - g(a) - d (the q fr term cancels)
      if constexpr (KIND == ZK_DIFF) return (closure_dg(clos, uu, tt) - tt) * dv;
      const Re g = closure_g(clos, tt, uu);
      return (g - (Re)1 - tt + q * fr[i]) * dv;  // trm1mt, times the forward dv
    }
  };
  // zpre[p] = e^{-i pi p/M} is an exponential and the first stage reads
  // elements p = j + c with c the same for every thread: zpre[p] = zpre[j]
  // zpre[c], one per-thread load and uniform (scalar) loads
  const C zprej = zpre[j];
  // base points at column 0 when x is past nx: the loads stay in bounds and
  // are issued unguarded (a guarded load waits for its data inside the guard,
  // one element after another), except in FP32 and the register-capped
  // tiles: there the unguarded loads spill or push the FP32 tile past the 64
  // VGPRs that let two workgroups share a CU
  auto plain_load = [&](int p) -> C {
    if constexpr (WPE > 1 || sizeof(Re) == 4) {
      if (!ok) return C{0, 0};
      return cmul(C{cval(2 * p), cval(2 * p + 1)}, cmul(zprej, zpre[p - j]));
    } else {
      const C w = cmul(C{cval(2 * p), cval(2 * p + 1)}, cmul(zprej, zpre[p - j]));
      return ok ? w : C{0, 0};
    }
  };
  // Solver steps in tiles compiled without a register limit: the t, u and
  // fr values of four elements are fetched together before their closures,
  // so that their loads are in flight at once (512^3, 4 sites: FP64 wide
  // radix-8 tile 3.60 -> 3.33 ms, FP32 1.72 -> 1.44 ms)
  constexpr bool STG = KIND == ZK_STEP && WPE == 1;
  struct StagedLoad {
    using staged = void;
    enum { group = 4 };
    struct raw_t {
      Re t0, t1, u0, u1, f0, f1;
    };
    const Re *tv, *uv, *fr;
    const C *zpre;
    C zprej;
    size_t base, plane;
    Re q, dv;
    int j, clos;
    bool ok;
    __device__ __forceinline__ raw_t fetch(int p) const {
      raw_t r;
      const size_t i0 = base + (size_t)(2 * p) * plane, i1 = i0 + plane;
      r.t0 = __builtin_nontemporal_load(&tv[i0]);
      r.t1 = __builtin_nontemporal_load(&tv[i1]);
      r.u0 = __builtin_nontemporal_load(&uv[i0]);
      r.u1 = __builtin_nontemporal_load(&uv[i1]);
      r.f0 = fr[i0];
      r.f1 = fr[i1];
      return r;
    }
    __device__ __forceinline__ Re cv(Re tt, Re uu, Re ff) const {
      const Re g = closure_g(clos, tt, uu);
      return (g - (Re)1 - tt + q * ff) * dv;  // trm1mt, times the forward dv
    }
    __device__ __forceinline__ C make(const raw_t &r, int p) const {
      if (!ok) return C{0, 0};
      return cmul(C{cv(r.t0, r.u0, r.f0), cv(r.t1, r.u1, r.f1)}, cmul(zprej, zpre[p - j]));
    }
  } staged_load{tv, uv, fr, zpre, zprej, base, plane, q, dv, j, clos, ok};
  auto &load = pick_ref<STG>(staged_load, plain_load);
  C *out = spec + ((size_t)v * M * P.ny + y) * P.nx + x;
  const size_t mstride = (size_t)P.ny * P.nx;
  if constexpr (mirror_last_stage<M, TZ, MAXR>()) {
    // the last stage leaves Z(m) and Z(M-1-m) in the same thread (butterflies
    // j and NS-1-j): the untangle G(m) = E(m) + w_m O(m) is formed from the
    // registers and written out (no LDS write, read and barrier)
    constexpr RadixList rl = make_radices(M, MAXR);
    constexpr int RL = rl.r[rl.n - 1], NS = M / RL;
    struct Untangle {
      using mirror_block = void;
      C *out;
      size_t mstride;
      const C *zpre;
      C zp0, zp1;  // zpost at the two butterflies' first elements
      bool ok;
      __device__ __forceinline__ void half(int m, C a, C b, C zp, C zr) {
        const C E = (Re)0.5 * (a + conj(b));
        const C O = mul_si<-1>((Re)0.5 * (a - conj(b)));
        // zpost[m0 + c] = zpost[m0] zpre[c]
        out[(size_t)m * mstride] = E + cmul(cmul(zp, zr), O);
      }
      __device__ __forceinline__ void block(int j0, int j1, C (&w)[2][RL]) {
        if (!ok) return;
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const C zr = zpre[r * NS];
          // element j0 + r NS pairs with M-1-(j0 + r NS) = j1 + (RL-1-r) NS
          half(j0 + r * NS, w[0][r], w[1][RL - 1 - r], zp0, zr);
          half(j1 + r * NS, w[1][r], w[0][RL - 1 - r], zp1, zr);
        }
      }
    } store{out, mstride, zpre, zpost[j], zpost[NS - 1 - j], ok};
    fft_column<M, -1, true, true, TZ, MAXR>(col, j, twM, load, store);
  } else {
    auto store = [&](int, C) {};
    fft_column<M, -1, true, false, TZ, MAXR>(col, j, twM, load, store);
    // zpost[m] = e^{-i pi (2m+1)/nz} = zpost[j] zpre[m - j]
    const C zpostj = zpost[j];
    for (int m = j; m < M; m += T) {
      const C a = col[lpad(m)];
      const C b = conj(col[lpad(M - 1 - m)]);
      const C E = (Re)0.5 * (a + b);
      const C O = mul_si<-1>((Re)0.5 * (a - b));
      if (ok) out[(size_t)m * mstride] = E + cmul(cmul(zpostj, zpre[m - j]), O);
    }
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
// residual phases of a tile do not overlap (512^3, 4 sites: FP32 4.7 ->
// 3.2 ms, FP64 Anderson(10) 24.4 -> 14.9 ms, even when the limit made the
// compiler spill; the FP32 AN2 and Picard tiles and the FP64 Anderson(m)
// tile now fit without spills; the FP32 difference-form and Anderson(m)
// tiles spill 8-112 B per lane).  The FP64 AN2 tile needs 148 KB of LDS and stays at
// one workgroup per CU.
template <int HM, class Re>
constexpr int zinv_waves_per_eu() {
  return sizeof(Re) == 4 ? 8 : (HM ? 4 : 1);
}

template <int M, int B, int TZ, int HM, int KIND, class Re, int WPE = zinv_waves_per_eu<HM, Re>(),
          int MAXRI = 8>
__global__ void __launch_bounds__(B *TZ)
    __attribute__((amdgpu_waves_per_eu(WPE)))
    k_zinv_resid(const cx<Re> *__restrict__ spec, const Re *__restrict__ t,
                 const Re *__restrict__ u, Re *__restrict__ R, const Re *__restrict__ R0,
                 const Re *__restrict__ R1, const Re *__restrict__ r0, HistT<Re> H, Params P,
                 double *__restrict__ partial,
                 const cx<Re> *__restrict__ twM, const cx<Re> *__restrict__ zpre,
                 const cx<Re> *__restrict__ zpost) {
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
  if (P.grp) {  // grouped launch (AN2 and Picard steps): this context's fields
    const GroupPtrs &g = P.grp[blockIdx.z];
    if (g.done && *g.done) return;
    spec = (const C *)g.spec;
    t = (const Re *)g.t;
    u = (const Re *)g.u;
    R = (Re *)g.R;
    R0 = (const Re *)g.R0;
    R1 = (const Re *)g.R1;
    partial = g.partial;
  } else if (P.done && *P.done) return;
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
  // G(m) -> Z(m) = E(m) + i O(m) as the column is read: one thread loads the
  // pair (m, M-1-m) and writes both Z values to LDS (one LDS pass and one
  // barrier fewer than staging G first: 6.6 -> 5.7 ms at 512^3, 4 sites)
  // zpost[m] = zpost[j] zpre[m - j] and zpost[M-1-m] = -conj(zpost[m])
  // (e^{-i pi (2m+1)/nz}, nz = 2M)
  const C zpostj = zpost[j];
  auto untangle = [&](int m, C a, C b) {
    const int mm = M - 1 - m;
    const C wm = cmul(zpostj, zpre[m - j]);
    const C wmm{-wm.x, wm.y};
    const C Em = (Re)0.5 * (a + conj(b));
    const C Om = cmulc((Re)0.5 * (a - conj(b)), wm);
    const C Emm = (Re)0.5 * (b + conj(a));
    const C Omm = cmulc((Re)0.5 * (b - conj(a)), wmm);
    col[lpad(m)] = Em + mul_si<1>(Om);
    if (mm != m) col[lpad(mm)] = Emm + mul_si<1>(Omm);
  };
  if constexpr (sizeof(Re) == 8) {
    // all pairs of this thread are loaded before any is untangled, so their
    // loads are in flight together (in points at column 0 when x is past
    // nx: the loads stay in bounds and need no guard); the register-capped
    // FP32 tiles would spill for it and load pair by pair
    constexpr int NPAIR = (M + 1) / 2;  // m with 2m <= M-1
    constexpr int PPT = (NPAIR + T - 1) / T;
    C av[PPT], bv[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int m = j + k * T;
      if (NPAIR % T == 0 || m < NPAIR) {
        av[k] = in[(size_t)m * plane];
        bv[k] = in[(size_t)(M - 1 - m) * plane];
      }
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int m = j + k * T;
      if (NPAIR % T == 0 || m < NPAIR)
        untangle(m, ok ? av[k] : C{0, 0}, ok ? bv[k] : C{0, 0});
    }
  } else {
    for (int m = j; 2 * m <= M - 1; m += T) {
      const int mm = M - 1 - m;
      untangle(m, ok ? in[(size_t)m * plane] : C{0, 0}, ok ? in[(size_t)mm * plane] : C{0, 0});
    }
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
  const size_t voff = (size_t)v * N;
  auto resid = [&](int z, Re h) {
    const size_t i = base + (size_t)z * plane;
    if constexpr (KIND == ZK_PLAIN) {  // plain inverse transform: write h itself
      Rv[i] = h;
      return;
    }
    // difference form r0 + dh - dg, else tr (calculate_cuda.cu:58)
    const Re ti = __builtin_nontemporal_load(&tv[i]), ui = __builtin_nontemporal_load(&uv[i]);
    Re rr;
    if constexpr (KIND == ZK_DIFF)
      rr = r0[voff + i] + h - closure_dg(clos, ui, ti);
    else
      rr = h + (Re)1 - closure_g(clos, ti, ui);
    __builtin_nontemporal_store(rr, &Rv[i]);
    // the sums are FP64 whatever the bulk type
    const double r = (double)rr;
    s[0] += r * r;
    if constexpr (HM == 0) {
      if constexpr (KIND != ZK_PICARD) {
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
  if constexpr (HM == 0 && (sizeof(Re) == 8 || KIND != ZK_DIFF)) {
    // AN2 and Picard steps and the plain inverse: the loads of two outputs at
    // a time are issued together (prep), then the residuals formed (put):
    // FP64 5.7 -> 5.5 ms at 512^3, 4 sites.  In the 64-VGPR FP32 tile this
    // form also fits without spilling, where the plain chain spills one
    // VGPR per output (2 GB of scratch traffic a launch at 512^3); the FP32
    // difference form keeps the plain chain (the staged one spills more)
    struct Store {
      using staged = void;
      enum { group = 2 };
      struct aux_t {
        Re t0, t1, u0, u1, a0, a1, b0, b1, c0, c1;
      };
      const Re *tv, *uv, *R0v, *R1v, *r0v;
      Re *Rv;
      const C *zpre;
      C zprej;
      int j;
      size_t base, plane;
      Re scale;
      int clos;
      bool ok;
      double s[W];
      __device__ __forceinline__ aux_t prep(int p) const {
        aux_t a;
        if constexpr (KIND != ZK_PLAIN) {
          const size_t i0 = base + (size_t)(2 * p) * plane, i1 = i0 + plane;
          a.t0 = __builtin_nontemporal_load(&tv[i0]);
          a.t1 = __builtin_nontemporal_load(&tv[i1]);
          a.u0 = __builtin_nontemporal_load(&uv[i0]);
          a.u1 = __builtin_nontemporal_load(&uv[i1]);
          if constexpr (KIND != ZK_PICARD) {
            a.a0 = __builtin_nontemporal_load(&R0v[i0]);
            a.a1 = __builtin_nontemporal_load(&R0v[i1]);
            a.b0 = __builtin_nontemporal_load(&R1v[i0]);
            a.b1 = __builtin_nontemporal_load(&R1v[i1]);
          }
          if constexpr (KIND == ZK_DIFF) {
            a.c0 = r0v[i0];
            a.c1 = r0v[i1];
          }
        }
        return a;
      }
      __device__ __forceinline__ void one(size_t i, Re h, Re ti, Re ui, Re ra, Re rb, Re rc) {
        if constexpr (KIND == ZK_PLAIN) {  // plain inverse transform: write h itself
          Rv[i] = h;
          return;
        }
        // difference form r0 + dh - dg, else tr (calculate_cuda.cu:58)
        Re rr;
        if constexpr (KIND == ZK_DIFF)
          rr = rc + h - closure_dg(clos, ui, ti);
        else
          rr = h + (Re)1 - closure_g(clos, ti, ui);
        __builtin_nontemporal_store(rr, &Rv[i]);
        // the sums are FP64 whatever the bulk type
        const double r = (double)rr;
        s[0] += r * r;
        if constexpr (KIND != ZK_PICARD) {
          const double d0 = r - (double)ra, d1 = r - (double)rb;
          s[1] += r * d0;
          s[2] += r * d1;
          s[3] += d0 * d0;
          s[4] += d0 * d1;
          s[5] += d1 * d1;
        }
      }
      __device__ __forceinline__ void put(int p, C zv, const aux_t &a) {
        if (!ok) return;
        // zpre[p] = zpre[j] zpre[p - j], p - j the same for every thread
        const C zz = scale * cmulc(zv, cmul(zprej, zpre[p - j]));
        const size_t i0 = base + (size_t)(2 * p) * plane;
        one(i0, zz.x, a.t0, a.u0, a.a0, a.b0, a.c0);
        one(i0 + plane, zz.y, a.t1, a.u1, a.a1, a.b1, a.c1);
      }
    } store{tv,   uv,   R0v,  R1v,   r0 ? r0 + voff : nullptr, Rv, zpre, zpre[j], j, base,
            plane, scale, clos, ok, {}};
    fft_column<M, +1, false, true, TZ, MAXRI>(col, j, twM, load, store);
#pragma unroll
    for (int k = 0; k < W; ++k) s[k] = store.s[k];
  } else {
    const C zprej = zpre[j];
    auto store = [&](int p, C zv) {
      if (!ok) return;
      const C zz = scale * cmulc(zv, cmul(zprej, zpre[p - j]));
      resid(2 * p, zz.x);
      resid(2 * p + 1, zz.y);
    };
    fft_column<M, +1, false, true, TZ, MAXRI>(col, j, twM, load, store);
  }

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
  // every column of the W sums, also when the workgroup has fewer than W
  // threads (FP64 Anderson(m) tiles of tiny grids)
  for (int k = tid; k < W; k += NT) {
    double a = 0.0;
    for (int w = 0; w < nparts; ++w) a += red[w * W + k];
    const size_t blk = (size_t)v * ntx * P.ny + (size_t)y * ntx + tx;  // [site][tile]
    partial[blk * W + k] = a;
  }
}

static inline unsigned zgrid(const Params &P, int B) {
  return (unsigned)(((P.nx + B - 1) / B) * P.ny * P.natv);
}

// Launch plans of the z passes, chosen per grid by the autotuner (plan 0 is
// the ZTile default).  Forward pass: 0 the ZTile tile (FP64 columns of
// M >= 128 in radix-4 stages, 16 columns, two workgroups per CU); 1 the wide
// radix-8 tile (ZTile::B columns, one workgroup per CU when it needs the
// LDS); 2 half as many columns in radix 8, compiled for two workgroups per
// CU.  Residual pass: 0 the ZTile tile; 1 half as many columns, compiled for
// twice the waves.  A plan narrower than 64 bytes of contiguous x per row
// is not offered.
template <int M, class Re, int V>
struct ZfPlan {
  using Z = ZTile<M, Re>;
  static constexpr int B = V == 0 ? Z::BF : (V == 1 ? Z::B : (Z::B > 1 ? Z::B / 2 : 1));
  static constexpr int T = V == 0 ? Z::TF : Z::T;
  static constexpr int MAXR = V == 0 ? Z::MAXRF : 8;
  static constexpr int WPE = V == 0 ? (Z::MAXRF == 4 ? 8 : 1) : (V == 1 ? 1 : 2);
};
template <int M, class Re, int V>
constexpr bool zf_plan_offered() {
  if constexpr (V == 0) {
    return true;
  } else {
    using A = ZfPlan<M, Re, 0>;
    using P1 = ZfPlan<M, Re, V>;
    if (P1::B * (int)sizeof(Re) < 64) return false;
    const bool same0 = A::B == P1::B && A::T == P1::T && A::MAXR == P1::MAXR;
    if constexpr (V == 1) {
      return !same0;
    } else {
      using Q = ZfPlan<M, Re, 1>;
      const bool same1 = Q::B == P1::B && Q::T == P1::T && Q::MAXR == P1::MAXR;
      return !same0 && !same1;
    }
  }
}
template <int M, class Re, int HM, int V>
struct ZiPlan {
  static constexpr int B0 = HM ? ZTile<M, Re>::BH : ZTile<M, Re>::B;
  static constexpr int B = V == 0 ? B0 : (B0 > 1 ? B0 / 2 : 1);
  static constexpr int T = ZTile<M, Re>::T;
  static constexpr int W0 = zinv_waves_per_eu<HM, Re>();
  static constexpr int WPE = V == 0 ? W0 : (2 * W0 > 8 ? 8 : 2 * W0);
};
template <int M, class Re, int V>
constexpr bool zi_plan_offered() {
  return V == 0 || (ZiPlan<M, Re, 0, V>::B * (int)sizeof(Re) >= 64 &&
                    ZiPlan<M, Re, 0, V>::B != ZiPlan<M, Re, 0, 0>::B);
}

template <int M, int KIND, class Re, int V = 0>
static void launch_zfwd(const Params &P, const Re *t, const Re *u, const Re *fr, cx<Re> *spec,
                        const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost,
                        hipStream_t st) {
  using Z = ZfPlan<M, Re, V>;
  hipLaunchKernelGGL((k_zfwd_closure<M, Z::B, Z::T, Z::MAXR, KIND, Re, Z::WPE>),
                     dim3(zgrid(P, Z::B), 1, P.grp ? P.ngrp : 1), dim3(Z::B * Z::T), 0, st, t, u,
                     fr, P, spec, twM, zpre, zpost);
}
template <int M, int HM, int KIND, class Re, int V = 0>
static void launch_zinv(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                        const Re *R0, const Re *R1, const Re *r0, const HistT<Re> &H,
                        double *partial, const cx<Re> *twM, const cx<Re> *zpre,
                        const cx<Re> *zpost, hipStream_t st) {
  using Z = ZiPlan<M, Re, HM, V>;
  hipLaunchKernelGGL((k_zinv_resid<M, Z::B, Z::T, HM, KIND, Re, Z::WPE>),
                     dim3(zgrid(P, Z::B), 1, P.grp ? P.ngrp : 1), dim3(Z::B * Z::T), 0, st, spec,
                     t, u, R, R0, R1, r0, H, P, partial, twM, zpre, zpost);
}

// Lengths with more than one plan: the autotuner's search space (smaller
// columns keep their single plan).
#define RISM_TUNED(n) ((n) >= 48)

// The kinds each bulk type is built for: FP64 runs solver steps, Picard
// steps and the plain transforms of the kernel-level entry points; FP32 (the
// bulk of the mixed solve) runs solver steps, Picard steps and the
// difference form.  Other combinations report hipErrorInvalidValue.
template <class Re>
constexpr bool zkind_built(int kind) {
  return kind == ZK_STEP || kind == ZK_PICARD ||
         (sizeof(Re) == 8 ? kind == ZK_PLAIN : kind == ZK_DIFF);
}

// solver steps of the forward pass in the tuned plan
template <int M, class Re>
static void zfwd_step_plan(const Params &P, const Re *t, const Re *u, const Re *fr, cx<Re> *spec,
                           const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost,
                           hipStream_t st) {
  const int v = P.var_zf[sizeof(Re) == 4];
  if constexpr (RISM_TUNED(M) && zf_plan_offered<M, Re, 1>()) {
    if (v == 1) return launch_zfwd<M, ZK_STEP, Re, 1>(P, t, u, fr, spec, twM, zpre, zpost, st);
  }
  if constexpr (RISM_TUNED(M) && zf_plan_offered<M, Re, 2>()) {
    if (v == 2) return launch_zfwd<M, ZK_STEP, Re, 2>(P, t, u, fr, spec, twM, zpre, zpost, st);
  }
  launch_zfwd<M, ZK_STEP, Re>(P, t, u, fr, spec, twM, zpre, zpost, st);
}

template <class Re>
hipError_t zfwd_dispatch(const Params &P, const Re *t, const Re *u, const Re *fr, cx<Re> *spec,
                         const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost,
                         hipStream_t st) {
  const int kind = P.mode == 3 ? ZK_DIFF : (P.mode != 0 ? ZK_PLAIN : ZK_STEP);
  if (!zkind_built<Re>(kind)) return hipErrorInvalidValue;
  switch (P.nz / 2) {
#define X(n)                                                                      \
  case n:                                                                         \
    if (kind == ZK_STEP)                                                          \
      zfwd_step_plan<n, Re>(P, t, u, fr, spec, twM, zpre, zpost, st);             \
    else if constexpr (sizeof(Re) == 8)                                           \
      launch_zfwd<n, ZK_PLAIN>(P, t, u, fr, spec, twM, zpre, zpost, st);          \
    else                                                                          \
      launch_zfwd<n, ZK_DIFF>(P, t, u, fr, spec, twM, zpre, zpost, st);           \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// AN2 and Picard steps of the residual pass in the tuned plan
template <int M, int KIND, class Re>
static void zinv_step_plan(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                           const Re *R0, const Re *R1, const Re *r0, const HistT<Re> &H,
                           double *partial, const cx<Re> *twM, const cx<Re> *zpre,
                           const cx<Re> *zpost, hipStream_t st) {
  if constexpr (RISM_TUNED(M) && zi_plan_offered<M, Re, 1>()) {
    if (P.var_zi[sizeof(Re) == 4] == 1)
      return launch_zinv<M, 0, KIND, Re, 1>(P, spec, t, u, R, R0, R1, r0, H, partial, twM,
                                            zpre, zpost, st);
  }
  launch_zinv<M, 0, KIND, Re>(P, spec, t, u, R, R0, R1, r0, H, partial, twM, zpre, zpost, st);
}

template <class Re>
hipError_t zinv_dispatch(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                         const Re *R0, const Re *R1, const Re *r0, double *partial,
                         const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost,
                         int use_hist, hipStream_t st) {
  if (P.mode == 3 && !r0) return hipErrorInvalidValue;
  const int kind = P.mode == 3 ? ZK_DIFF
                               : (P.mode != 0 ? ZK_PLAIN : (use_hist ? ZK_STEP : ZK_PICARD));
  if (!zkind_built<Re>(kind)) return hipErrorInvalidValue;
  const HistT<Re> H{};
  switch (P.nz / 2) {
#define X(n)                                                                              \
  case n:                                                                                 \
    if (kind == ZK_STEP)                                                                  \
      zinv_step_plan<n, ZK_STEP, Re>(P, spec, t, u, R, R0, R1, r0, H, partial, twM, zpre, \
                                     zpost, st);                                          \
    else if (kind == ZK_PICARD)                                                           \
      zinv_step_plan<n, ZK_PICARD, Re>(P, spec, t, u, R, R0, R1, r0, H, partial, twM,     \
                                       zpre, zpost, st);                                  \
    else if constexpr (sizeof(Re) == 8)                                                   \
      launch_zinv<n, 0, ZK_PLAIN>(P, spec, t, u, R, R0, R1, r0, H, partial, twM, zpre,    \
                                  zpost, st);                                             \
    else                                                                                  \
      launch_zinv<n, 0, ZK_DIFF>(P, spec, t, u, R, R0, R1, r0, H, partial, twM, zpre,     \
                                 zpost, st);                                              \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Anderson(m): solver steps (and the difference form in FP32); R.R_p for
// the nh history residuals, nh = 0 during the Picard steps
template <class Re>
hipError_t zinv_hist_dispatch(const Params &P, const cx<Re> *spec, const Re *t, const Re *u,
                              Re *R, const Re *r0, const HistT<Re> &H, double *partial,
                              const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost,
                              hipStream_t st) {
  if (P.mode == 3 && !r0) return hipErrorInvalidValue;
  if (P.mode != 0 && P.mode != 3) return hipErrorInvalidValue;
  if (P.mode == 3 && sizeof(Re) == 8) return hipErrorInvalidValue;
  const Re *nul = nullptr;
  switch (P.nz / 2) {
#define X(n)                                                                                \
  case n:                                                                                   \
    if constexpr (sizeof(Re) == 4) {                                                        \
      if (P.mode == 3) {                                                                    \
        launch_zinv<n, 1, ZK_DIFF>(P, spec, t, u, R, nul, nul, r0, H, partial, twM, zpre,   \
                                   zpost, st);                                              \
        break;                                                                              \
      }                                                                                     \
    }                                                                                       \
    launch_zinv<n, 1, ZK_STEP>(P, spec, t, u, R, nul, nul, r0, H, partial, twM, zpre, zpost, \
                               st);                                                         \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Workgroups of the residual pass in the plan it runs with (the tuned plan
// for AN2 and Picard steps, the default for the Anderson(m) sums, the
// difference form and the plain inverse); plan < 0: the most workgroups any
// plan of this type launches (the size of the partial-sum buffer).
template <class Re>
size_t zinv_blocks_of(const Params &P, int hm, int plan) {
  const int M = P.nz / 2;
  int B = 16;
  const int v = plan >= 0 ? plan : 1;
  switch (M) {
#define X(n)                                                                         \
  case n:                                                                            \
    B = hm ? ZTile<n, Re>::BH : ZTile<n, Re>::B;                                     \
    if constexpr (RISM_TUNED(n) && zi_plan_offered<n, Re, 1>())                      \
      if (!hm && v == 1) B = ZiPlan<n, Re, 0, 1>::B;                                 \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      break;
  }
  return (size_t)((P.nx + B - 1) / B) * P.ny * P.natv;
}

// The launch interface (rism_kernels.h); each bulk type is instantiated in
// its own translation unit.
template <class Re>
size_t zinv_num_blocks(const Params &P, int hm) {
  if (hm < 0) return zinv_blocks_of<Re>(P, 0, -1);  // largest over the plans
  const bool tuned = !hm && (P.mode == 0);
  return zinv_blocks_of<Re>(P, hm, tuned ? P.var_zi[sizeof(Re) == 4] : 0);
}
template <class Re>
hipError_t run_zfwd(const Params &P, const Re *t, const Re *u, const Re *fr, cx<Re> *spec,
                    const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost, hipStream_t st) {
  return zfwd_dispatch<Re>(P, t, u, fr, spec, twM, zpre, zpost, st);
}
template <class Re>
hipError_t run_zinv(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                    const Re *R0, const Re *R1, const Re *r0, double *partial,
                    const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost, int use_hist,
                    hipStream_t st) {
  return zinv_dispatch<Re>(P, spec, t, u, R, R0, R1, r0, partial, twM, zpre, zpost, use_hist,
                           st);
}
template <class Re>
hipError_t run_zinv_hist(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                         const Re *r0, const HistT<Re> &H, double *partial, const cx<Re> *twM,
                         const cx<Re> *zpre, const cx<Re> *zpost, hipStream_t st) {
  return zinv_hist_dispatch<Re>(P, spec, t, u, R, r0, H, partial, twM, zpre, zpost, st);
}

#define RISM_KZPASS_INSTANTIATE(Re)                                                           \
  template size_t zinv_num_blocks<Re>(const Params &, int);                                   \
  template hipError_t run_zfwd<Re>(const Params &, const Re *, const Re *, const Re *,        \
                                   cx<Re> *, const cx<Re> *, const cx<Re> *, const cx<Re> *,  \
                                   hipStream_t);                                              \
  template hipError_t run_zinv<Re>(const Params &, const cx<Re> *, const Re *, const Re *,    \
                                   Re *, const Re *, const Re *, const Re *, double *,        \
                                   const cx<Re> *, const cx<Re> *, const cx<Re> *, int,       \
                                   hipStream_t);                                              \
  template hipError_t run_zinv_hist<Re>(const Params &, const cx<Re> *, const Re *,           \
                                        const Re *, Re *, const Re *, const HistT<Re> &,      \
                                        double *, const cx<Re> *, const cx<Re> *,             \
                                        const cx<Re> *, hipStream_t);

}  // namespace rism
