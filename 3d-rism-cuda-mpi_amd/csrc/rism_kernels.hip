// rism_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the 3D-RISM solve.
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

// ---------------------------------------------------------------------------
// 1. closure + c_sr + real->half-complex DFT along z
// ---------------------------------------------------------------------------
// Column = the z-line at (x, y) of one site: nz reals packed as M = nz/2
// complex z_p = c(2p) + i c(2p+1), pre-twiddled by e^{-i pi p/M}; after the
// M-point DFT, G(m) = E(m) + w_m O(m) with E, O from Z(m) and Z(M-1-m) and
// w_m = e^{-2 pi i (m+1/2)/nz}: the z-DFT at the positive half-shifted
// frequencies.
template <int M, int B>
__global__ void __launch_bounds__(B *FftShape<M>::T)
    k_zfwd_closure(const double *__restrict__ t, const double *__restrict__ u,
                   const double *__restrict__ fr, Params P, cplx *__restrict__ spec,
                   const cplx *__restrict__ twM, const cplx *__restrict__ zpre,
                   const cplx *__restrict__ zpost) {
  constexpr int T = FftShape<M>::T, CS = FftShape<M>::CS;
  __shared__ cplx lds[B * CS];
  const int tid = threadIdx.x;
  const int c = tid % B, j = tid / B;
  const int x = blockIdx.x * B + c;
  const int y = blockIdx.y;
  const int v = blockIdx.z;
  const bool ok = x < P.nx;
  const size_t plane = (size_t)P.nx * P.ny;
  const size_t N = plane * (size_t)P.nz;
  const double *tv = t + (size_t)v * N;
  const double *uv = u + (size_t)v * N;
  const double q = P.qv[v];
  const size_t base = (size_t)y * P.nx + (ok ? x : 0);
  const int clos = P.clos;
  const double dv = P.dv;
  cplx *col = lds + c * CS;

  auto cval = [&](int z) -> double {
    const size_t i = base + (size_t)z * plane;
    const double tt = tv[i];
    if (P.mode != 0) return tt * dv;  // plain transform of t
    const double g = closure_g(clos, tt, uv[i]);
    return (g - 1.0 - tt + q * fr[i]) * dv;  // trm1mt, times the forward dv
  };
  auto load = [&](int p) -> cplx {
    if (!ok) return mk(0.0, 0.0);
    return cmul(mk(cval(2 * p), cval(2 * p + 1)), zpre[p]);
  };
  auto store = [&](int, cplx) {};
  fft_column<M, -1, true, false>(col, j, twM, load, store);

  cplx *out = spec + ((size_t)v * M * P.ny + y) * P.nx + x;
  const size_t mstride = (size_t)P.ny * P.nx;
  for (int m = j; m < M; m += T) {
    const cplx a = col[lpad(m)];
    const cplx b = conj(col[lpad(M - 1 - m)]);
    const cplx E = 0.5 * (a + b);
    const cplx O = mul_si<-1>(0.5 * (a - b));
    if (ok) out[(size_t)m * mstride] = E + cmul(zpost[m], O);
  }
}

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

// ---------------------------------------------------------------------------
// 3. DFT along x, OZ with chi_vv(|k|), inverse DFT along x
// ---------------------------------------------------------------------------
// One workgroup per (m, y) row, all sites: natv rows of nx complex in LDS.
// chi_vv is evaluated from the natural cubic spline of the .xvk table at
// |k| = sqrt(kx^2 + ky^2 + kz^2) (solvent_spline.cu:44-55 evaluates the same
// spline at the same |k| shell values, splint.cc:4-20 picks the interval).
__device__ __forceinline__ int spline_interval(const double *__restrict__ knots, int np,
                                               double x0, double idk, double kk) {
  int i = (int)floor((kk - x0) * idk);
  i = i < 0 ? 0 : (i > np - 2 ? np - 2 : i);
  if (i > 0 && knots[i] > kk) --i;
  if (i < np - 2 && knots[i + 1] <= kk) ++i;
  return i;
}

template <int L>
__global__ void __launch_bounds__(1024)
    k_xpass_oz(cplx *__restrict__ spec, const cplx *__restrict__ fkraw, Params P,
               const cplx *__restrict__ twL, const cplx *__restrict__ prex,
               const double *__restrict__ knots, const double *__restrict__ coef) {
  constexpr int T = FftShape<L>::T, CS = FftShape<L>::CS;
  extern __shared__ cplx lds[];
  const int tid = threadIdx.x;
  const int nv = P.natv;
  const int v = tid / T, j = tid % T;
  const int y = blockIdx.x, m = blockIdx.y;
  const int M = P.nz / 2;
  const size_t rowstride = (size_t)M * P.ny * L;  // site stride in spec
  const size_t rowoff = ((size_t)m * P.ny + y) * L;
  cplx *row = spec + (size_t)v * rowstride + rowoff;
  cplx *mylds = lds + v * CS;
  auto load = [&](int i) -> cplx { return cmul(row[i], prex[i]); };
  auto store = [&](int i, cplx vv) { row[i] = cmulc(vv, prex[i]); };
  auto nostore = [&](int, cplx) {};
  if (P.mode == 1) {  // plain forward transform along x only
    auto st = [&](int i, cplx vv) { row[i] = vv; };
    fft_column<L, -1, true, true>(mylds, j, twL, load, st);
    return;
  }
  if (P.mode == 2) {  // plain inverse transform along x only
    auto ld = [&](int i) -> cplx { return row[i]; };
    fft_column<L, +1, true, true>(mylds, j, twL, ld, store);
    return;
  }
  fft_column<L, -1, true, false>(mylds, j, twL, load, nostore);

  // OZ at every k of the row
  const double ky = P.kfac[1] * ((double)y - (double)P.ny / 2.0 + 0.5);
  const double kz = P.kfac[2] * ((double)(m + P.nz / 2) - (double)P.nz / 2.0 + 0.5);
  const int nthr = nv * T;
  const int nv2 = nv * nv;
  for (int jx = tid; jx < L; jx += nthr) {
    const double kx = P.kfac[0] * ((double)jx - (double)L / 2.0 + 0.5);
    const double kk = sqrt(kx * kx + ky * ky + kz * kz);
    const int ii = spline_interval(knots, P.np, P.x0, P.idk, kk);
    const double tt = kk - knots[ii];
    const double *cf = coef + (size_t)ii * 4 * nv2;
    const cplx fk = fkraw[rowoff + jx];
    cplx cc[RISM_MAX_SITES];
    for (int w = 0; w < nv; ++w) cc[w] = lds[w * CS + lpad(jx)] - P.qv[w] * fk;
    for (int vv = 0; vv < nv; ++vv) {
      cplx acc = mk(0.0, 0.0);
      for (int w = 0; w < nv; ++w) {
        const int pw = vv * nv + w;
        const double X =
            ((cf[3 * nv2 + pw] * tt + cf[2 * nv2 + pw]) * tt + cf[nv2 + pw]) * tt + cf[pw];
        acc = acc + X * cc[w];
      }
      lds[vv * CS + lpad(jx)] = acc;
    }
  }
  __syncthreads();
  fft_column<L, +1, false, true>(mylds, j, twL, load, store);
}

// ---------------------------------------------------------------------------
// 5. half-complex->real inverse DFT along z, residual, FP64 partial sums
// ---------------------------------------------------------------------------
// Partial sums per workgroup (6 doubles): R.R, R.(R-R0), R.(R-R1),
// (R-R0)^2, (R-R0)(R-R1), (R-R1)^2 (anderson_cuda.cu:34-40, cal_rms.cu:20).
template <int M, int B>
__global__ void __launch_bounds__(B *FftShape<M>::T)
    k_zinv_resid(const cplx *__restrict__ spec, const double *__restrict__ t,
                 const double *__restrict__ u, double *__restrict__ R,
                 const double *__restrict__ R0, const double *__restrict__ R1, Params P,
                 double *__restrict__ partial, const cplx *__restrict__ twM,
                 const cplx *__restrict__ zpre, const cplx *__restrict__ zpost, int use_hist) {
  constexpr int T = FftShape<M>::T, CS = FftShape<M>::CS;
  constexpr int NT = B * T;
  __shared__ cplx lds[B * CS];
  const int tid = threadIdx.x;
  const int c = tid % B, j = tid / B;
  const int x = blockIdx.x * B + c;
  const int y = blockIdx.y;
  const int v = blockIdx.z;
  const bool ok = x < P.nx;
  const size_t plane = (size_t)P.nx * P.ny;
  const size_t N = plane * (size_t)P.nz;
  cplx *col = lds + c * CS;
  const cplx *in = spec + ((size_t)v * M * P.ny + y) * P.nx + (ok ? x : 0);
  for (int m = j; m < M; m += T) col[lpad(m)] = ok ? in[(size_t)m * plane] : mk(0.0, 0.0);
  __syncthreads();
  // G(m) -> Z(m) = E(m) + i O(m), pairs (m, M-1-m) handled by one thread
  for (int m = j; 2 * m <= M - 1; m += T) {
    const int mm = M - 1 - m;
    const cplx a = col[lpad(m)], b = col[lpad(mm)];
    const cplx Em = 0.5 * (a + conj(b));
    const cplx Om = cmulc(0.5 * (a - conj(b)), zpost[m]);
    const cplx Emm = 0.5 * (b + conj(a));
    const cplx Omm = cmulc(0.5 * (b - conj(a)), zpost[mm]);
    col[lpad(m)] = Em + mul_si<1>(Om);
    if (mm != m) col[lpad(mm)] = Emm + mul_si<1>(Omm);
  }
  __syncthreads();

  const double *tv = t + (size_t)v * N;
  const double *uv = u + (size_t)v * N;
  double *Rv = R + (size_t)v * N;
  const double *R0v = R0 + (size_t)v * N;
  const double *R1v = R1 + (size_t)v * N;
  const size_t base = (size_t)y * P.nx + (ok ? x : 0);
  const int clos = P.clos;
  const double scale = P.inv_scale;
  double s[6] = {0, 0, 0, 0, 0, 0};
  const int plain = P.mode != 0;
  auto resid = [&](int z, double h) {
    const size_t i = base + (size_t)z * plane;
    if (plain) {  // plain inverse transform: write h itself
      Rv[i] = h;
      return;
    }
    const double g = closure_g(clos, tv[i], uv[i]);
    const double r = h + 1.0 - g;  // tr (calculate_cuda.cu:58)
    Rv[i] = r;
    s[0] += r * r;
    if (use_hist) {
      const double d0 = r - R0v[i], d1 = r - R1v[i];
      s[1] += r * d0;
      s[2] += r * d1;
      s[3] += d0 * d0;
      s[4] += d0 * d1;
      s[5] += d1 * d1;
    }
  };
  auto load = [&](int) -> cplx { return mk(0.0, 0.0); };
  auto store = [&](int p, cplx zv) {
    if (!ok) return;
    const cplx zz = scale * cmulc(zv, zpre[p]);
    resid(2 * p, zz.x);
    resid(2 * p + 1, zz.y);
  };
  fft_column<M, +1, false, true>(col, j, twM, load, store);

  // workgroup reduction of the six sums: wave shuffles then LDS when the
  // workgroup is made of whole waves, LDS only otherwise (tiny grids)
  __syncthreads();
  double *red = reinterpret_cast<double *>(lds);
  int nparts;
  if constexpr (NT % 64 == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      double a = s[k];
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
      s[k] = a;
    }
    const int lane = tid & 63, wv = tid >> 6;
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) red[wv * 6 + k] = s[k];
    }
    nparts = NT / 64;
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k) red[tid * 6 + k] = s[k];
    nparts = NT;
  }
  __syncthreads();
  if (tid < 6) {
    double a = 0.0;
    for (int w = 0; w < nparts; ++w) a += red[w * 6 + tid];
    const size_t blk = ((size_t)v * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    partial[blk * 6 + tid] = a;
  }
}

// ---------------------------------------------------------------------------
// FP64 reductions
// ---------------------------------------------------------------------------
// Sums nblk rows of `width` doubles into out[width] (one workgroup per column
// group; pairwise tree within the workgroup, sequential across strides).
__global__ void __launch_bounds__(256)
    k_reduce_cols(const double *__restrict__ in, size_t nrow, int width,
                  double *__restrict__ out) {
  __shared__ double sh[256];
  const int col = blockIdx.x;
  double a = 0.0, comp = 0.0;  // Kahan-compensated per-thread accumulation
  for (size_t r = threadIdx.x; r < nrow; r += blockDim.x) {
    const double yv = in[r * width + col] - comp;
    const double tv = a + yv;
    comp = (tv - a) - yv;
    a = tv;
  }
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[col] = sh[0];
}

// ---------------------------------------------------------------------------
// AN2 (anderson_cuda.cu:1-22)
// ---------------------------------------------------------------------------
// Picard warm-up: slot <- (t, R) is a pointer rotation on the host; the new
// t = t + R lands in the buffer being retired.
__global__ void k_an2_picard(double *__restrict__ tnew, const double *__restrict__ t,
                             const double *__restrict__ R, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tnew[i] = t[i] + R[i];
}

// tnew aliases T0 or T1 (the slot being retired); each element is read before
// it is written by the same thread.
__global__ void k_an2_update(double *tnew, const double *__restrict__ t,
                             const double *__restrict__ R, const double *T0, const double *T1,
                             const double *__restrict__ R0, const double *__restrict__ R1,
                             double s1, double s2, double mcoef, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ti = t[i], ri = R[i];
  const double t0 = T0[i], t1 = T1[i];
  const double uu = ti + s1 * (t0 - ti) + s2 * (t1 - ti);
  const double vv = ti + ri + s1 * (t0 + R0[i] - ti - ri) + s2 * (t1 + R1[i] - ti - ri);
  tnew[i] = uu + mcoef * (vv - uu);
}

// ---------------------------------------------------------------------------
// set-up: potentials (cal_LJ.cu:56-81, cal_Coulomb.cu:40-62, cal_potential.cu:39-43)
// ---------------------------------------------------------------------------
// Each thread owns PX consecutive x points of one (y, z) line and all natv
// sites; solute atoms stream through LDS in tiles of ATILE.  Outputs
// u[v] = beta_LJ sum 4 eps (s^-12 - s^-6) + q_v de and fr.
constexpr int PX = 4;
constexpr int ATILE = 256;

__global__ void __launch_bounds__(256)
    k_potential_r(const double *__restrict__ atoms, int natu, Params P, double ikbt,
                  double ubeta, int do_coul, double *__restrict__ u, double *__restrict__ fr) {
  // atoms layout per atom: x, y, z, q, then natv sigma_uv^2, then natv 4 eps_uv
  __shared__ double sa[ATILE * (4 + 2 * RISM_MAX_SITES)];
  const int nv = P.natv;
  const int stride = 4 + 2 * nv;
  const int nxq = (P.nx + PX - 1) / PX;
  const size_t lines = (size_t)P.ny * P.nz;
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = gid < lines * (size_t)nxq;
  const int xq = active ? (int)(gid % nxq) : 0;
  const size_t line = active ? gid / nxq : 0;
  const int y = (int)(line % P.ny), z = (int)(line / P.ny);
  const double ry = (double)(y - P.ny / 2) * P.dr[1];
  const double rz = (double)(z - P.nz / 2) * P.dr[2];
  double rx[PX];
#pragma unroll
  for (int p = 0; p < PX; ++p) rx[p] = (double)(xq * PX + p - P.nx / 2) * P.dr[0];
  double lj[PX][RISM_MAX_SITES];
  double de[PX], ff[PX];
#pragma unroll
  for (int p = 0; p < PX; ++p) {
    de[p] = 0.0;
    ff[p] = 0.0;
    for (int w = 0; w < RISM_MAX_SITES; ++w) lj[p][w] = 0.0;
  }
  constexpr double icut2 = 1.0 / (2.0e-3 * 2.0e-3);
  for (int a0 = 0; a0 < natu; a0 += ATILE) {
    const int na = min(ATILE, natu - a0);
    __syncthreads();
    for (int i = threadIdx.x; i < na * stride; i += blockDim.x)
      sa[i] = atoms[(size_t)a0 * stride + i];
    __syncthreads();
    for (int a = 0; a < na; ++a) {
      const double *at = sa + a * stride;
      const double ax = at[0], ay = at[1], az = at[2], aq = at[3];
      const double dy = ry - ay, dz = rz - az;
      const double dyz = dy * dy + dz * dz;
#pragma unroll
      for (int p = 0; p < PX; ++p) {
        const double dx = rx[p] - ax;
        const double r2 = dx * dx + dyz;
        const double ir2 = 1.0 / r2;
        for (int w = 0; w < nv; ++w) {
          double s2 = at[4 + w] * ir2;
          s2 = s2 > icut2 ? icut2 : s2;
          const double s6 = s2 * s2 * s2;
          lj[p][w] += at[4 + nv + w] * s6 * (s6 - 1.0);
        }
        if (do_coul) {
          const double ra = sqrt(r2);
          if (ra >= 1.0e-5) {
            const double qr = aq / ra;
            de[p] += qr;
            // 1 - e^{-r} rounds to exactly 1 in FP64 once r > 37.5
            ff[p] += (ra < 38.0) ? qr * (1.0 - exp(-ra)) : qr;
          } else {
            ff[p] += aq;
          }
        }
      }
    }
  }
  if (!active) return;
  const size_t N = (size_t)P.nx * P.ny * P.nz;
  const size_t rowbase = line * P.nx;
#pragma unroll
  for (int p = 0; p < PX; ++p) {
    const int x = xq * PX + p;
    if (x >= P.nx) break;
    const double e = de[p] * ubeta;
    for (int w = 0; w < nv; ++w) u[(size_t)w * N + rowbase + x] = lj[p][w] * ikbt + P.qv[w] * e;
    fr[rowbase + x] = ff[p] * ubeta;
  }
}

// fk on the half spectrum (cal_Coulomb.cu:65-79, 90-95), stored raw:
// fkraw = conj(phi(k)) fk, phi(k) = -i (-1)^(nx/2 + ny/2 + jx + jy + m).
// Each thread owns FKP consecutive jx of one (jy, m) row; e^{-i kx x_u} is
// advanced along the row by the recurrence e^{-i (kx + dkx) x_u} =
// e^{-i kx x_u} e^{-i dkx x_u}.
constexpr int FKP = 16;

__global__ void __launch_bounds__(256)
    k_fk_half(const double *__restrict__ atoms, int natu, int astride, Params P, double ubeta,
              cplx *__restrict__ fkraw) {
  __shared__ double sa[ATILE * 4];
  const int M = P.nz / 2;
  const int nxq = (P.nx + FKP - 1) / FKP;
  const size_t rows = (size_t)P.ny * M;
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = gid < rows * (size_t)nxq;
  const int xq = active ? (int)(gid % nxq) : 0;
  const size_t row = active ? gid / nxq : 0;
  const int jy = (int)(row % P.ny), m = (int)(row / P.ny);
  const int jx0 = xq * FKP;
  const double kx0 = P.kfac[0] * ((double)jx0 - (double)P.nx / 2.0 + 0.5);
  const double ky = P.kfac[1] * ((double)jy - (double)P.ny / 2.0 + 0.5);
  const double kz = P.kfac[2] * ((double)(m + P.nz / 2) - (double)P.nz / 2.0 + 0.5);
  const double dkx = P.kfac[0];
  double re[FKP], im[FKP];
#pragma unroll
  for (int p = 0; p < FKP; ++p) re[p] = im[p] = 0.0;
  for (int a0 = 0; a0 < natu; a0 += ATILE) {
    const int na = min(ATILE, natu - a0);
    __syncthreads();
    for (int i = threadIdx.x; i < na; i += blockDim.x) {
      const double *src = atoms + (size_t)(a0 + i) * astride;
      sa[4 * i] = src[0];
      sa[4 * i + 1] = src[1];
      sa[4 * i + 2] = src[2];
      sa[4 * i + 3] = src[3];
    }
    __syncthreads();
    for (int a = 0; a < na; ++a) {
      const double ax = sa[4 * a], ay = sa[4 * a + 1], az = sa[4 * a + 2], aq = sa[4 * a + 3];
      double sn, cs, dsn, dcs;
      sincos(kx0 * ax + ky * ay + kz * az, &sn, &cs);
      sincos(dkx * ax, &dsn, &dcs);
      // e = cos - i sin; step = dcs - i dsn
      double er = cs, ei = -sn;
#pragma unroll
      for (int p = 0; p < FKP; ++p) {
        re[p] += aq * er;
        im[p] += aq * ei;
        const double nr = er * dcs + ei * dsn;
        const double ni = ei * dcs - er * dsn;
        er = nr;
        ei = ni;
      }
    }
  }
  if (!active) return;
  const int sgn0 = ((P.nx / 2 + P.ny / 2 + jy + m) & 1) ? -1 : 1;
#pragma unroll
  for (int p = 0; p < FKP; ++p) {
    const int jx = jx0 + p;
    if (jx >= P.nx) break;
    const double kx = P.kfac[0] * ((double)jx - (double)P.nx / 2.0 + 0.5);
    const double rk2 = kx * kx + ky * ky + kz * kz;
    const double rk4i = 1.0 / (rk2 * (rk2 + 1.0));
    const double f = 4.0 * M_PI * rk4i * ubeta;
    const double fre = f * re[p], fim = f * im[p];
    // conj(phi) = i * s
    const double s = ((jx & 1) ? -sgn0 : sgn0);
    fkraw[row * P.nx + jx] = mk(-s * fim, s * fre);
  }
}

// t = q_v fr (initialize_tuv.cu:19-23)
__global__ void k_init_t(double *__restrict__ t, const double *__restrict__ fr, Params P,
                         size_t N) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * P.natv) return;
  const int v = (int)(i / N);
  t[i] = P.qv[v] * fr[i - (size_t)v * N];
}

// g = closure(t, u), h = R + g - 1 (the h of the step that produced R)
__global__ void k_fields(const double *__restrict__ t, const double *__restrict__ u,
                         const double *__restrict__ R, Params P, size_t n, int which,
                         double *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double g = closure_g(P.clos, t[i], u[i]);
  if (which == 1)
    out[i] = g;
  else if (which == 2)
    out[i] = R[i] + g - 1.0;
  else
    out[i] = R[i] + g - 1.0 - t[i];
}

// ---------------------------------------------------------------------------
// thermodynamics (cal_exchem.cu:55-155, cal_pmv.cu:27-57): per-block FP64
// sums of SC (closure-matched), HNC, GF integrands and h - t, per site.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_thermo_partial(const double *__restrict__ t, const double *__restrict__ th,
                     const double *__restrict__ u, const double *__restrict__ R, Params P,
                     size_t N, double *__restrict__ partial) {
  // t: the current t (the integrand's t); th: the t of the step that produced
  // R, from which h = R + g(th, u) - 1 is rebuilt (equal unless the loop
  // stopped at maxstep, after an AN2 update)
  __shared__ double sh[4][256 / 64];
  const int v = blockIdx.y;
  const double *tv = t + (size_t)v * N, *uv = u + (size_t)v * N, *Rv = R + (size_t)v * N;
  const double *thv = th + (size_t)v * N;
  double s[4] = {0, 0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (size_t)gridDim.x * blockDim.x) {
    const double tt = tv[i];
    const double g = closure_g(P.clos, thv[i], uv[i]);
    const double h = Rv[i] + g - 1.0;
    const double ehnc = h * 0.5 * tt - (h - tt);
    const double egf = -(h * 0.5 + 1.0) * (h - tt);
    const double ekh = (h > 0.0) ? egf : ehnc;
    s[0] += (P.clos == 0) ? ekh : ehnc;
    s[1] += ehnc;
    s[2] += egf;
    s[3] += h - tt;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double a = s[k];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    s[k] = a;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 4; ++k) sh[k][wv] = s[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    double a = 0.0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) a += sh[threadIdx.x][w];
    partial[((size_t)v * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = a;
  }
}

// reference-convention spectrum from a raw one (tests): F = phi * D,
// phi = -i (-1)^(nx/2 + ny/2 + jx + jy + m); scaled by `scale`.
__global__ void k_apply_phase(cplx *__restrict__ s, Params P, int inverse) {
  const size_t M = P.nz / 2;
  const size_t n = M * P.ny * P.nx;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int jx = (int)(i % P.nx), jy = (int)((i / P.nx) % P.ny), m = (int)(i / ((size_t)P.nx * P.ny));
  const int sg = ((P.nx / 2 + P.ny / 2 + jx + jy + m) & 1) ? -1 : 1;
  const cplx a = s[i];
  // phi = -i sg : (x, y) -> sg * (y, -x); conj(phi) = i sg : sg * (-y, x)
  s[i] = inverse ? mk(-sg * a.y, sg * a.x) : mk(sg * a.y, -sg * a.x);
}

// ---------------------------------------------------------------------------
// host-side launchers (dispatch on the FFT lengths)
// ---------------------------------------------------------------------------
template <int L>
constexpr int tile_cols() {
  // keep one workgroup's LDS tile at <= ~74 KB of FP64 complex
  return L >= 1024 ? 4 : (L >= 512 ? 8 : 16);
}

#define RISM_LENGTHS(X) \
  X(8) X(12) X(16) X(20) X(24) X(30) X(32) X(36) X(40) X(48) X(60) X(64) X(72) X(80) X(90) \
  X(96) X(120) X(128) X(144) X(160) X(180) X(192) X(240) X(256) X(288) X(320) X(360)     \
  X(384) X(480) X(512) X(576) X(640) X(720) X(768) X(960) X(1024)

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

template <int M>
static void launch_zfwd(const Params &P, const double *t, const double *u, const double *fr,
                        cplx *spec, const cplx *twM, const cplx *zpre, const cplx *zpost,
                        hipStream_t st) {
  constexpr int B = tile_cols<M>();
  dim3 grid((P.nx + B - 1) / B, P.ny, P.natv);
  hipLaunchKernelGGL((k_zfwd_closure<M, B>), grid, dim3(B * FftShape<M>::T), 0, st, t, u, fr,
                     P, spec, twM, zpre, zpost);
}

template <int L, int S>
static void launch_y(const Params &P, cplx *spec, int nplanes, const cplx *twL,
                     const cplx *pre, hipStream_t st) {
  constexpr int B = tile_cols<L>();
  dim3 grid((P.nx + B - 1) / B, nplanes);
  hipLaunchKernelGGL((k_ypass<L, B, S>), grid, dim3(B * FftShape<L>::T), 0, st, spec, P.nx,
                     twL, pre);
}

template <int L>
static void launch_xoz(const Params &P, cplx *spec, const cplx *fkraw, const cplx *twL,
                       const cplx *prex, const double *knots, const double *coef,
                       hipStream_t st) {
  dim3 grid(P.ny, P.nz / 2);
  const int threads = P.natv * FftShape<L>::T;
  const size_t lds = (size_t)P.natv * FftShape<L>::CS * sizeof(cplx);
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_xpass_oz<L>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_xpass_oz<L>), grid, dim3(threads), lds, st, spec, fkraw, P, twL, prex,
                     knots, coef);
}

template <int M>
static void launch_zinv(const Params &P, const cplx *spec, const double *t, const double *u,
                        double *R, const double *R0, const double *R1, double *partial,
                        const cplx *twM, const cplx *zpre, const cplx *zpost, int use_hist,
                        hipStream_t st) {
  constexpr int B = tile_cols<M>();
  dim3 grid((P.nx + B - 1) / B, P.ny, P.natv);
  hipLaunchKernelGGL((k_zinv_resid<M, B>), grid, dim3(B * FftShape<M>::T), 0, st, spec, t, u,
                     R, R0, R1, P, partial, twM, zpre, zpost, use_hist);
}

size_t zinv_num_blocks(const Params &P) {
  const int M = P.nz / 2;
  int B = 16;
  switch (M) {
#define X(n) \
  case n:    \
    B = tile_cols<n>(); \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      break;
  }
  return (size_t)((P.nx + B - 1) / B) * P.ny * P.natv;
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

hipError_t run_zfwd(const Params &P, const double *t, const double *u, const double *fr,
                    cplx *spec, const cplx *twM, const cplx *zpre, const cplx *zpost,
                    hipStream_t st) {
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

hipError_t run_xoz(const Params &P, cplx *spec, const cplx *fkraw, const cplx *twL,
                   const cplx *prex, const double *knots, const double *coef, hipStream_t st) {
  switch (P.nx) {
#define X(n)                                                       \
  case n:                                                          \
    launch_xoz<n>(P, spec, fkraw, twL, prex, knots, coef, st);     \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t run_zinv(const Params &P, const cplx *spec, const double *t, const double *u,
                    double *R, const double *R0, const double *R1, double *partial,
                    const cplx *twM, const cplx *zpre, const cplx *zpost, int use_hist,
                    hipStream_t st) {
  switch (P.nz / 2) {
#define X(n)                                                                             \
  case n:                                                                                \
    launch_zinv<n>(P, spec, t, u, R, R0, R1, partial, twM, zpre, zpost, use_hist, st); \
    break;
    RISM_LENGTHS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t run_reduce_cols(const double *in, size_t nrow, int width, double *out,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_cols, dim3(width), dim3(256), 0, st, in, nrow, width, out);
  return hipGetLastError();
}

static inline unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t run_an2_picard(double *tnew, const double *t, const double *R, size_t n,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_an2_picard, dim3(nblk(n, 256)), dim3(256), 0, st, tnew, t, R, n);
  return hipGetLastError();
}

hipError_t run_an2_update(double *tnew, const double *t, const double *R, const double *T0,
                          const double *T1, const double *R0, const double *R1, double s1,
                          double s2, double m, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_an2_update, dim3(nblk(n, 256)), dim3(256), 0, st, tnew, t, R, T0, T1, R0,
                     R1, s1, s2, m, n);
  return hipGetLastError();
}

hipError_t run_potential_r(const double *atoms, int natu, const Params &P, double ikbt,
                           double ubeta, int do_coul, double *u, double *fr, hipStream_t st) {
  const size_t nth = (size_t)P.ny * P.nz * ((P.nx + PX - 1) / PX);
  hipLaunchKernelGGL(k_potential_r, dim3(nblk(nth, 256)), dim3(256), 0, st, atoms, natu, P,
                     ikbt, ubeta, do_coul, u, fr);
  return hipGetLastError();
}

hipError_t run_fk_half(const double *atoms, int natu, int astride, const Params &P,
                       double ubeta, cplx *fkraw, hipStream_t st) {
  const size_t nth = (size_t)P.ny * (P.nz / 2) * ((P.nx + FKP - 1) / FKP);
  hipLaunchKernelGGL(k_fk_half, dim3(nblk(nth, 256)), dim3(256), 0, st, atoms, natu, astride, P,
                     ubeta, fkraw);
  return hipGetLastError();
}

hipError_t run_init_t(double *t, const double *fr, const Params &P, size_t N, hipStream_t st) {
  hipLaunchKernelGGL(k_init_t, dim3(nblk(N * P.natv, 256)), dim3(256), 0, st, t, fr, P, N);
  return hipGetLastError();
}

hipError_t run_fields(const double *t, const double *u, const double *R, const Params &P,
                      size_t n, int which, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_fields, dim3(nblk(n, 256)), dim3(256), 0, st, t, u, R, P, n, which, out);
  return hipGetLastError();
}

int thermo_blocks(size_t N) {
  size_t b = (N + 255) / 256;
  return (int)(b < 1024 ? b : 1024);
}

hipError_t run_thermo_partial(const double *t, const double *th, const double *u,
                              const double *R, const Params &P, size_t N, double *partial,
                              hipStream_t st) {
  dim3 grid(thermo_blocks(N), P.natv);
  hipLaunchKernelGGL(k_thermo_partial, grid, dim3(256), 0, st, t, th, u, R, P, N, partial);
  return hipGetLastError();
}

hipError_t run_apply_phase(cplx *s, const Params &P, int inverse, hipStream_t st) {
  const size_t n = (size_t)(P.nz / 2) * P.ny * P.nx;
  hipLaunchKernelGGL(k_apply_phase, dim3(nblk(n, 256)), dim3(256), 0, st, s, P, inverse);
  return hipGetLastError();
}

}  // namespace rism
