// rism_common.h -- device helpers and compile-time tables shared by the
// kernel translation units (rism_kzpass.hip, rism_kypass.hip,
// rism_kxpass.hip, rism_kmisc.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rism_fft.h"
#include "rism_kernels.h"

namespace rism {

__device__ __forceinline__ double closure_g(int clos, double t, double u) {
  const double e = -u + t;  // calculate_cuda.cu:4
  if (clos == 0 && e >= 0.0) return 1.0 + e;
  return exp(e);
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

template <int L>
constexpr int tile_cols() {
  // keep one workgroup's LDS tile at <= ~74 KB of FP64 complex
  return L >= 1024 ? 4 : (L >= 512 ? 8 : 16);
}

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
