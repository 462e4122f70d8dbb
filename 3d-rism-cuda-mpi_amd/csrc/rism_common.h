// rism_common.h -- device helpers and compile-time tables shared by the
// kernel translation units (rism_kzpass64.hip, rism_kzpass32.hip, rism_kypass.hip,
// rism_kxpass.hip, rism_kmisc.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rism_fft.h"
#include "rism_kernels.h"

namespace rism {

__device__ __forceinline__ double rexp(double a) { return exp(a); }
__device__ __forceinline__ float rexp(float a) { return expf(a); }

// g from t and u in the bulk type: clos 0 KH, 1 HNC (kh / hnc,
// calculate_cuda.cu:1-16), 10 + n PSE-n (Kast-Kloss partial series: exp(e)
// for e = -u + t < 0, else sum_{i <= n} e^i / i!; PSE-1 is KH)
template <class Re>
__device__ __forceinline__ Re closure_g(int clos, Re t, Re u) {
  const Re e = -u + t;  // calculate_cuda.cu:4
  if (clos == 1 || e < (Re)0) return rexp(e);
  if (clos == 0) return (Re)1 + e;
  Re s = (Re)1;
  for (int k = clos - 10; k >= 1; --k) s = (Re)1 + e * s / (Re)k;
  return s;
}

__device__ __forceinline__ double rexpm1(double a) { return expm1(a); }
__device__ __forceinline__ float rexpm1(float a) { return expm1f(a); }

// g(a + d) - g(a) for e = a -> a + d, evaluated without cancellation when
// |d| << |a| (the difference form of a correction sweep): exp(a) expm1(d)
// on the exponential branch, d on KH's linear one, d times the divided
// difference of the PSE-n polynomial on its series branch; across a branch
// point |a|, |a + d| <= |d| and the plain difference is exact enough.
template <class Re>
__device__ __forceinline__ Re closure_dg(int clos, Re a, Re d) {
  const Re b = a + d;
  if (clos == 1 || (a < (Re)0 && b < (Re)0)) return rexp(a) * rexpm1(d);
  if (a >= (Re)0 && b >= (Re)0) {
    if (clos == 0) return d;
    // PSE-n: P[a, b] = sum_{i>=1} h_{i-1}(a, b) / i!, h_k = sum_j a^j b^(k-j)
    Re pa = (Re)1, h = (Re)1, fact = (Re)1, s = (Re)1;
    for (int i = 2; i <= clos - 10; ++i) {
      pa *= a;
      h = b * h + pa;
      fact *= (Re)i;
      s += h / fact;
    }
    return d * s;
  }
  return closure_g(clos, b, (Re)0) - closure_g(clos, a, (Re)0);
}

// the PSE-n correction to the HNC integrand of the excess chemical
// potential: e^(n+1) / (n+1)! where e = -u + t > 0
__device__ __forceinline__ double pse_excess(int n, double e) {
  if (e <= 0.0) return 0.0;
  double p = 1.0;
  for (int k = 1; k <= n + 1; ++k) p *= e / (double)k;
  return p;
}

__device__ __forceinline__ int spline_interval(const double *__restrict__ knots, int np,
                                               double x0, double idk, double kk) {
  int i = (int)floor((kk - x0) * idk);
  i = i < 0 ? 0 : (i > np - 2 ? np - 2 : i);
  if (i > 0 && knots[i] > kk) --i;
  if (i < np - 2 && knots[i + 1] <= kk) ++i;
  return i;
}


// FFT lengths (nx, ny and nz/2) the kernels are instantiated for: powers of
// two and the 2^a 3^b 5^c sizes of the uneven batched grids.
#define RISM_LENGTHS(X) \
  X(8) X(12) X(16) X(20) X(24) X(30) X(32) X(36) X(40) X(48) X(60) X(64) X(72) X(80) X(90) \
  X(96) X(128) X(144) X(160) X(180) X(192) X(256) X(512) X(1024)

// Columns per workgroup of the strided (y, z) passes: one workgroup's LDS
// tile stays at <= ~74 KB, and a tile row is 128 B of contiguous x, in both
// bulk types (FP32 tiles are twice as wide), within 1024 threads.
template <int L, class Re = double>
constexpr int tile_cols() {
  int b = (L >= 1024 ? 4 : (L >= 512 ? 8 : 16)) * (int)(sizeof(double) / sizeof(Re));
  while (b * FftShape<L>::T > 1024) b /= 2;
  return b;
}

// Tiles of the z passes, which read and write real fields with the z-plane
// stride: for M >= 128 and FP64 a tile row is 256 B of contiguous x (32
// columns) within ~148 KB of LDS (one workgroup per CU, 1024 threads); FP32
// keeps 32 columns (128 B), which measured faster than 64 columns with two
// butterflies per thread; smaller columns keep the tiles of the y pass.
template <int M, class Re>
struct ZTile {
  static constexpr int CS = FftShape<M>::CS;
  static constexpr int wide() {
    int b = 32;
    while (b > 1 && (long)b * CS * (long)sizeof(cx<Re>) > 150 * 1024) b /= 2;
    return b;
  }
  static constexpr int B = (M >= 128) ? wide() : tile_cols<M, Re>();
  static constexpr int T = (B * FftShape<M>::T > 1024) ? 1024 / B : FftShape<M>::T;
  // the residual pass with the Anderson(m) sums needs more registers: half
  // as many columns (512 threads)
  static constexpr int BH = (M >= 128 && B > 1) ? B / 2 : B;
  // FP64 forward pass of long columns in radix-4 stages: M/4 threads per
  // column and half as many columns (74 KB of LDS), so that two 1024-thread
  // workgroups share a CU (53 VGPRs, no spills; 4.8 -> 4.3 ms at 512^3, 4
  // sites).  The residual pass gains nothing from the same plan (it would
  // spill to fit 64 VGPRs) and keeps radix 8.
  static constexpr bool R4F = sizeof(Re) == 8 && M >= 128;
  static constexpr int MAXRF = R4F ? 4 : 8;
  static constexpr int TF = R4F ? FftShape<M, 4>::T : T;
  static constexpr int BF = R4F ? 1024 / TF : B;
};

static inline unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Workgroups of a 1D grid are handed to the 8 XCDs of the MI355X round-robin
// by linear id.  This bijection renumbers them so that each XCD executes one
// contiguous range of logical ids, in order: neighbouring tiles (which share
// cache lines and read-only fields) then run on the same XCD and meet in
// its L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  constexpr int NX = 8;
  const int per = (nblocks + NX - 1) / NX;
  int tall = nblocks % NX;
  if (tall == 0) tall = NX;
  const int xcd = bid % NX, local = bid / NX;
  return xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
}

}  // namespace rism
