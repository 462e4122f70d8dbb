// rism_kmisc.hip -- FP64 reductions, the AN2 update, set-up potentials,
// k-space Coulomb, initial t, derived fields and thermodynamic partial sums.
#include <algorithm>

#include "rism_common.h"

namespace rism {

// ---------------------------------------------------------------------------
// FP64 reductions
// ---------------------------------------------------------------------------
// Sums nrow rows of `width` doubles into out[width], for each of
// gridDim.y consecutive blocks of rows (one workgroup per column and block;
// Kahan-compensated per thread, pairwise tree within the workgroup).
__global__ void __launch_bounds__(256)
    k_reduce_cols(const double *__restrict__ in, size_t nrow, int width,
                  double *__restrict__ out, const int *done, const GroupPtrs *grp) {
  __shared__ double sh[256];
  if (grp) {  // grouped launch: this context's partial sums
    const GroupPtrs &g = grp[blockIdx.z];
    done = g.done;
    in = g.partial;
    out = g.sums;
  }
  if (done && *done) return;
  const int col = blockIdx.x;
  in += (size_t)blockIdx.y * nrow * width;
  out += (size_t)blockIdx.y * width;
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
// The reference keeps the slot pairs (T_p, R_p) and forms
//   u = t + s1 (T0 - t) + s2 (T1 - t)
//   v = t + R + s1 (T0 + R0 - t - R) + s2 (T1 + R1 - t - R)
//   t' = u + m (v - u).
// Writing Q = t + m R and P_p = T_p + m R_p this is
//   t' = Q + s1 (P0 - Q) + s2 (P1 - Q),
// so each slot stores P_p (for the update) and R_p (for the Gram matrix,
// kept by pointer rotation): the update reads t, R, P0, P1 and writes t'
// and the retired slot's P = Q, 6 words per point instead of 7.
//
// Picard warm-up (newdt0): t' = t + R, P_b = t + m R.
// Both kernels move two consecutive elements per thread as 16-byte accesses
// (n = natv * N is even: every grid size is even).
// Elementwise kernels move 16 B per lane per field (two FP64 or four FP32
// values).
template <class Re>
struct alignas(16) V16 {
  static constexpr int K = 16 / sizeof(Re);
  Re a[K];
};

// 16-byte non-temporal accesses of the single-use streams of the updates
template <class Re>
struct NTVec;
template <>
struct NTVec<double> {
  typedef double T __attribute__((ext_vector_type(2)));
};
template <>
struct NTVec<float> {
  typedef float T __attribute__((ext_vector_type(4)));
};
template <class Re>
__device__ __forceinline__ V16<Re> ld_nt(const V16<Re> *p) {
  using T = typename NTVec<Re>::T;
  const T v = __builtin_nontemporal_load(reinterpret_cast<const T *>(p));
  V16<Re> r;
#pragma unroll
  for (int k = 0; k < V16<Re>::K; ++k) r.a[k] = v[k];
  return r;
}
template <class Re>
__device__ __forceinline__ void st_nt(V16<Re> *p, const V16<Re> &x) {
  using T = typename NTVec<Re>::T;
  T v;
#pragma unroll
  for (int k = 0; k < V16<Re>::K; ++k) v[k] = x.a[k];
  __builtin_nontemporal_store(v, reinterpret_cast<T *>(p));
}

template <class Re>
__global__ void k_an2_picard(V16<Re> *__restrict__ tnew, const V16<Re> *__restrict__ t,
                             const V16<Re> *__restrict__ R, V16<Re> *__restrict__ Pb, Re mcoef,
                             size_t nv, const int *done, const GroupPtrs *grp) {
  if (grp) {
    const GroupPtrs &g = grp[blockIdx.z];
    done = g.done;
    tnew = (V16<Re> *)g.tnew;
    t = (const V16<Re> *)g.t;
    R = (const V16<Re> *)g.R;
    Pb = (V16<Re> *)g.Pb;
  }
  if (done && *done) return;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const V16<Re> ti = t[i], ri = R[i];
  V16<Re> a, b;
#pragma unroll
  for (int k = 0; k < V16<Re>::K; ++k) {
    a.a[k] = ti.a[k] + ri.a[k];
    b.a[k] = ti.a[k] + mcoef * ri.a[k];
  }
  tnew[i] = a;
  Pb[i] = b;
}

// Pb aliases P0 or P1 (the slot being retired); each element is read before
// it is written by the same thread.
template <class Re>
__global__ void k_an2_update(V16<Re> *__restrict__ tnew, const V16<Re> *__restrict__ t,
                             const V16<Re> *__restrict__ R, const V16<Re> *P0,
                             const V16<Re> *P1, V16<Re> *Pb, const double *__restrict__ scal,
                             Re mcoef, size_t nv, const int *done, const GroupPtrs *grp) {
  if (grp) {
    const GroupPtrs &g = grp[blockIdx.z];
    done = g.done;
    tnew = (V16<Re> *)g.tnew;
    t = (const V16<Re> *)g.t;
    R = (const V16<Re> *)g.R;
    P0 = (const V16<Re> *)g.P0;
    P1 = (const V16<Re> *)g.P1;
    Pb = (V16<Re> *)g.Pb;
    scal = g.scal;
  }
  if (done && *done) return;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const Re s1 = (Re)scal[3], s2 = (Re)scal[4];
  const V16<Re> ti = ld_nt(t + i), ri = ld_nt(R + i), p0 = ld_nt(P0 + i), p1 = ld_nt(P1 + i);
  V16<Re> a, q;
#pragma unroll
  for (int k = 0; k < V16<Re>::K; ++k) {
    q.a[k] = ti.a[k] + mcoef * ri.a[k];
    a.a[k] = q.a[k] + s1 * (p0.a[k] - q.a[k]) + s2 * (p1.a[k] - q.a[k]);
  }
  st_nt(tnew + i, a);
  st_nt(Pb + i, q);
}

// End of a step: rms (cal_rms.cu:20-24), convergence / divergence flag
// (iterate.cu:50-57) and the AN2 mixing coefficients for the next update
// (cal_theta anderson.cu:81-102 with the rho^-1 weights, leg()'s Gaussian
// elimination without pivoting anderson.cu:105-130, the mp shift
// anderson.cu:66-67).  One thread; the sums are 6 doubles per site.
__global__ void k_step_finish(const double *__restrict__ sums, const double *__restrict__ rho,
                              int natv, size_t N, double tol, int mixing, double mp, int binary,
                              double *__restrict__ scal, double *__restrict__ rms_hist,
                              int *__restrict__ done, const GroupPtrs *grp) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (grp) {
    const GroupPtrs &g = grp[blockIdx.z];
    sums = g.sums;
    scal = g.scal;
    rms_hist = g.rms_hist;
    done = g.done;
  }
  if (done && done[0]) return;
  // step number: done[2] counts the finished steps of the run
  int step = 0;
  if (done) {
    step = done[2] + 1;
    done[2] = step;
  }
  double ss = 0.0, cc0 = 0.0, cc1 = 0.0, a0 = 0.0, a1 = 0.0, a3 = 0.0;
  for (int v = 0; v < natv; ++v) {
    const double *s = sums + 6 * v;
    const double ir = 1.0 / rho[v];
    ss += s[0];
    cc0 += s[1] * ir;
    cc1 += s[2] * ir;
    a0 += s[3] * ir;
    a1 += s[4] * ir;
    a3 += s[5] * ir;
  }
  const double rms = sqrt(ss / ((double)N * natv));
  scal[0] = rms;
  if (rms_hist) rms_hist[step] = rms;
  if (done) {
    if (!isfinite(rms)) {
      done[0] = 2;
      done[1] = step;
    } else if (rms <= tol) {
      done[0] = 1;
      done[1] = step;
    }
  }
  if (mixing) {
    const double a2 = a1;
    double b1 = a1, b3 = a3, c0 = cc0, c1 = cc1;
    const double p = a0;
    b1 /= p;
    const double r = c0 /= p;
    const double q = a2;
    b3 -= q * b1;
    c1 -= q * r;
    const double x1 = c1 / b3;
    const double x0 = c0 - b1 * x1;
    scal[1] = x0;
    scal[2] = x1;
    scal[3] = x0 + mp * ((binary == 0) ? 0 : 1);
    scal[4] = x1 + mp * ((binary == 0) ? 1 : 0);
  }
}

// ---------------------------------------------------------------------------
// End of one Anderson(m) step on the device (no counterpart in the
// reference): rms and the convergence / divergence flag as k_step_finish;
// the weighted dots of the new residual R (buffer cur) with itself and the
// nh history residuals (buffers rb[0..nh), oldest first) enter the dot matrix
// D[m1][m1]; and, when the next update mixes, the coefficients th of
//   min |R - sum_p th_p (R - R_p)|  over the newest k history entries:
// G th = b with G_pq = <R - R_p, R - R_q>, b_p = <R, R - R_p>, G scaled to a
// unit diagonal, Gaussian elimination with partial pivoting; when a pivot
// falls below 1e-12 (or a diagonal is not positive) the oldest entry is left
// out and the solve repeated.  th (zeros for entries left out) goes to
// xcoef[0..nh).  One thread, FP64; W = 1 + RISM_MAX_HIST sums per site.
// ---------------------------------------------------------------------------
struct AndmHist {
  int rb[RISM_MAX_HIST];
  int nh;
};

__global__ void k_andm_finish(const double *__restrict__ sums, const double *__restrict__ rho,
                              int natv, size_t N, double tol, int W, double *__restrict__ D,
                              int m1, int cur, AndmHist H, int mixing,
                              double *__restrict__ xcoef, double *__restrict__ scal,
                              double *__restrict__ rms_hist, int *__restrict__ done) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (done && done[0]) return;
  int step = 0;
  if (done) {
    step = done[2] + 1;
    done[2] = step;
  }
  double ss = 0.0;
  for (int v = 0; v < natv; ++v) ss += sums[W * v];
  const double rms = sqrt(ss / ((double)N * natv));  // cal_rms.cu:24
  scal[0] = rms;
  if (rms_hist) rms_hist[step] = rms;
  if (done) {
    if (!isfinite(rms)) {
      done[0] = 2;
      done[1] = step;
    } else if (rms <= tol) {
      done[0] = 1;
      done[1] = step;
    }
  }
  // weighted dots <R, R> and <R, R_p>, weights 1 / rho_v (anderson.cu:37-39)
  auto wdot = [&](int k) {
    double a = 0.0;
    for (int v = 0; v < natv; ++v) a += sums[W * v + k] / rho[v];
    return a;
  };
  D[(size_t)cur * m1 + cur] = wdot(0);
  for (int p = 0; p < H.nh; ++p) {
    const int bb = H.rb[p];
    D[(size_t)cur * m1 + bb] = D[(size_t)bb * m1 + cur] = wdot(1 + p);
  }
  if (!mixing) return;
  const int nh = H.nh;
  for (int p = 0; p < nh; ++p) xcoef[p] = 0.0;
  const double rr = D[(size_t)cur * m1 + cur];
  double G[RISM_MAX_HIST][RISM_MAX_HIST], b[RISM_MAX_HIST], sc[RISM_MAX_HIST], x[RISM_MAX_HIST];
  for (int first = 0; first < nh; ++first) {
    const int k = nh - first;
    bool ok = true;
    for (int p = 0; p < k && ok; ++p) {
      const int bp = H.rb[first + p];
      const double dp = D[(size_t)cur * m1 + bp];
      b[p] = rr - dp;
      for (int q = 0; q < k; ++q) {
        const int bq = H.rb[first + q];
        G[p][q] = rr - dp - D[(size_t)cur * m1 + bq] + D[(size_t)bp * m1 + bq];
      }
      ok = G[p][p] > 0.0;
      if (ok) sc[p] = 1.0 / sqrt(G[p][p]);
    }
    if (!ok) continue;
    for (int p = 0; p < k; ++p) {
      b[p] *= sc[p];
      for (int q = 0; q < k; ++q) G[p][q] *= sc[p] * sc[q];
    }
    for (int col = 0; col < k && ok; ++col) {
      int piv = col;
      for (int r = col + 1; r < k; ++r)
        if (fabs(G[r][col]) > fabs(G[piv][col])) piv = r;
      if (fabs(G[piv][col]) < 1e-12) {
        ok = false;
        break;
      }
      if (piv != col) {
        for (int q = 0; q < k; ++q) {
          const double tq = G[piv][q];
          G[piv][q] = G[col][q];
          G[col][q] = tq;
        }
        const double tb = b[piv];
        b[piv] = b[col];
        b[col] = tb;
      }
      for (int r = col + 1; r < k; ++r) {
        const double f = G[r][col] / G[col][col];
        for (int q = col; q < k; ++q) G[r][q] -= f * G[col][q];
        b[r] -= f * b[col];
      }
    }
    if (!ok) continue;
    for (int r = k - 1; r >= 0; --r) {
      double a = b[r];
      for (int q = r + 1; q < k; ++q) a -= G[r][q] * x[q];
      x[r] = a / G[r][r];
    }
    for (int p = 0; p < k; ++p) xcoef[first + p] = x[p] * sc[p];
    return;
  }
}

// ---------------------------------------------------------------------------
// Anderson(m): the AN2 update generalised to nh history slots,
//   t' = Q + sum_p x_p (P_p - Q),  Q = t + beta R,
// with the retired slot's P overwritten by Q.  Pb may alias one of the P_p;
// each element is read before the same thread writes it.
// ---------------------------------------------------------------------------
template <class Re>
__global__ void k_andm_update(V16<Re> *__restrict__ tnew, const V16<Re> *__restrict__ t,
                              const V16<Re> *__restrict__ R, HistUpdateT<Re> H, V16<Re> *Pb,
                              const double *__restrict__ xcoef, Re beta, size_t nv,
                              const int *done) {
  if (done && *done) return;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const V16<Re> ti = ld_nt(t + i), ri = ld_nt(R + i);
  V16<Re> q, a;
#pragma unroll
  for (int k = 0; k < V16<Re>::K; ++k) a.a[k] = q.a[k] = ti.a[k] + beta * ri.a[k];
#pragma unroll
  for (int h = 0; h < RISM_MAX_HIST; ++h) {
    if (h < H.nh) {
      // Pb may be the oldest P_h: this thread reads element i before it writes it
      const V16<Re> p = ld_nt(reinterpret_cast<const V16<Re> *>(H.p[h]) + i);
      const Re x = (Re)xcoef[h];
#pragma unroll
      for (int k = 0; k < V16<Re>::K; ++k) a.a[k] = fma(x, p.a[k] - q.a[k], a.a[k]);
    }
  }
  st_nt(tnew + i, a);
  st_nt(Pb + i, q);
}

// the closure exponent -u + t, formed in FP64 and rounded once
__global__ void k_exponent(const double *__restrict__ t, const double *__restrict__ u, size_t n,
                           float *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)(-u[i] + t[i]);
}

// Moves a field between the FP64 frame and the FP32 frame of a correction
// block, whose values are differences from the base point t0:
// up: out = t0 + d (out may be t0 itself); down: out = x - t0 in FP32.
__global__ void k_rebase_up(double *out, const double *t0, const float *__restrict__ d,
                            size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = t0[i] + (double)d[i];
}
__global__ void k_rebase_down(const double *__restrict__ x, const double *__restrict__ t0,
                              float *__restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)(x[i] - t0[i]);
}

// conversion between the FP64 and FP32 copies of a field
template <class From, class To>
__global__ void k_convert(const From *__restrict__ in, To *__restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (To)in[i];
}

// ---------------------------------------------------------------------------
// set-up: potentials (cal_LJ.cu:56-81, cal_Coulomb.cu:40-62, cal_potential.cu:39-43)
// ---------------------------------------------------------------------------
// Each thread owns PX consecutive x points of one (y, z) line and a group of
// G <= 8 solvent sites starting at v0; solute atoms stream through LDS in
// tiles of ATILE records (atom_stride(natv) doubles: x, y, z, q, rc2,
// sigma_uv^2[natv], A[natv] = 4 eps sigma^12, B[natv] = 4 eps sigma^6).
// 4 eps (s^-12 - s^-6) with s^2 = r^2 / sigma^2 is A r^-12 - B r^-6, so the
// powers of 1/r are formed once per pair and each site costs two FMAs.  The
// reference's floor s^2 >= 4e-6 (cal_LJ.cu:11-12) can only bind when
// r^2 < rc2 = 4e-6 max_v sigma_uv^2; those (rare) pairs take the clamped form.
// Outputs u[v] = beta sum LJ + q_v de, and fr (first group).
// 1/r comes from the v_rsq_f64 seed refined by two Newton steps, so the pair
// loop has no FP64 division or square root.
constexpr int PX = 4;
constexpr int ATILE = 256;

// 1/sqrt(x) in FP64: the hardware estimate (relative error up to 5.2e-8 on
// this part) and one third-order (Halley-type) correction, y (1 + e/2 +
// 3e^2/8) with e = 1 - x y^2: measured within 1.4e-16 relative over
// x in [1e-4, 1e6], against 2.4e-16 for two Newton steps, in 5 operations
// instead of 7
__device__ __forceinline__ double rsq_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// 1/x in FP64 for the rare clamped pairs: the hardware estimate and two
// Newton steps, within an ulp of the IEEE quotient (with the IEEE division
// sequence the 4-site pair loop spills at the four-wave register limit)
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(fma(-x, y, 1.0), y, y);
  return fma(fma(-x, y, 1.0), y, y);
}

// Up to 4 sites per launch the pair loop fits 128 VGPRs without spilling and
// runs four waves per SIMD instead of the three its unconstrained 138 VGPRs
// allow (512^3, 12877 atoms, 4 sites: 1.88 -> 1.65 s); 5-8 sites would
// spill at that limit and keep the default.
template <int G>
constexpr int potential_waves_per_eu() {
  return G <= 4 ? 4 : 1;
}

template <int G, bool COUL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(potential_waves_per_eu<G>())))
    k_potential_r(const double *__restrict__ atoms, int natu, int v0, Params P, double ikbt,
                  double ubeta, int write_fr, double *__restrict__ u,
                  double *__restrict__ fr) {
  extern __shared__ double sa[];
  const int nv = P.natv;
  const int stride = atom_stride(nv);
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
  double lj[PX][G];
  double de[PX], ff[PX];
#pragma unroll
  for (int p = 0; p < PX; ++p) {
    de[p] = 0.0;
    ff[p] = 0.0;
#pragma unroll
    for (int w = 0; w < G; ++w) lj[p][w] = 0.0;
  }
  constexpr double icut2 = 1.0 / (2.0e-3 * 2.0e-3);  // s^2 >= 4e-6 (cal_LJ.cu:11-12)
  for (int a0 = 0; a0 < natu; a0 += ATILE) {
    const int na = min(ATILE, natu - a0);
    __syncthreads();
    for (int i = threadIdx.x; i < na * stride; i += blockDim.x)
      sa[i] = atoms[(size_t)a0 * stride + i];
    __syncthreads();
    for (int a = 0; a < na; ++a) {
      const double *at = sa + a * stride;
      const double ax = at[0], ay = at[1], az = at[2], aq = at[3], rc2 = at[4];
      double A[G], B[G];
#pragma unroll
      for (int w = 0; w < G; ++w) {
        A[w] = at[5 + nv + v0 + w];
        B[w] = at[5 + 2 * nv + v0 + w];
      }
      const double dy = ry - ay, dz = rz - az;
      const double dyz = dy * dy + dz * dz;
#pragma unroll
      for (int p = 0; p < PX; ++p) {
        const double dx = rx[p] - ax;
        const double r2 = dx * dx + dyz;
        const double ir = rsq_nr(r2);  // NaN at r2 = 0: both uses below handle it
        const double ir2 = ir * ir;
        if (r2 >= rc2) {
          // A r^-12 - B r^-6 = r^-6 (A r^-6 - B)
          const double ir6 = ir2 * ir2 * ir2;
#pragma unroll
          for (int w = 0; w < G; ++w) lj[p][w] = fma(fma(A[w], ir6, -B[w]), ir6, lj[p][w]);
        } else {
#pragma unroll
          for (int w = 0; w < G; ++w) {
            const double sg2 = at[5 + v0 + w];
            const double e4 = B[w] * rcp_nr(sg2 * sg2 * sg2);
            const double s2 = fmin(sg2 * ir2, icut2);  // fmin drops the NaN of r = 0
            const double s6 = s2 * s2 * s2;
            lj[p][w] += e4 * s6 * (s6 - 1.0);
          }
        }
        if constexpr (COUL) {
          if (r2 >= 1.0e-10) {  // r >= 1e-5
            const double qr = aq * ir;
            de[p] += qr;
            // 1 - e^{-r} rounds to exactly 1 in FP64 once r > 37.5
            ff[p] += (r2 < 38.0 * 38.0) ? qr * (1.0 - exp(-(r2 * ir))) : qr;
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
#pragma unroll
    for (int w = 0; w < G; ++w)
      u[(size_t)(v0 + w) * N + rowbase + x] = lj[p][w] * ikbt + P.qv[v0 + w] * e;
    if (write_fr) fr[rowbase + x] = ff[p] * ubeta;
  }
}

// ---------------------------------------------------------------------------
// per-atom integrals of the converged g (cal_grad.cu:44-90, cal_euv.cu:30-70,
// cal_ad1.cu:40-80, cal_ad2.cu:40-75): for each solute atom u, sums over the
// grid points r with |r - r_u| >= sigma_uv / 2 (sigma_uv lambda / 2 for ad1)
// of g_v(r) times
//   kind 0 (grad):  [24 eps s6 (2 s6 - 1) / r^2 + cc q_u q_v / r^3] (r - r_u)
//   kind 1 (euv):   4 eps s6 (s6 - 1) + cc q_u q_v / r          (per site v)
//   kind 2 (ad1):   24 eps s6 (2 s6 - 1) / lambda
//   kind 3 (ad2):   cc q_u q_v / r
// with s6 = (sigma_uv^2 / r^2)^3, eps = sqrt(eps_u eps_v) in J/mol, cc =
// E_h a_0 N_A; kinds 0, 2, 3 weight each site by rho_v.  One thread per atom
// (atoms in registers), grid points with their g values staged in LDS; each
// workgroup covers one slab of points and writes its sums per atom.
// ---------------------------------------------------------------------------
constexpr int AS_TILE = 256;

template <int KIND>
__global__ void __launch_bounds__(256)
    k_atom_sums(const double *__restrict__ t, const double *__restrict__ u, Params P,
                const double *__restrict__ adata, int natu, double lambda, double cc,
                const double *__restrict__ rho, size_t slab, int width,
                double *__restrict__ partial) {
  __shared__ double sp[AS_TILE * (3 + RISM_MAX_SITES / 2)];
  const int nv = P.natv;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = a < natu;
  // atom record: x y z q, then sigma_uv[nv], eps_uv[nv], q_v folded below
  const double *ad = adata + (size_t)(active ? a : 0) * (4 + 2 * nv);
  const double ax = ad[0], ay = ad[1], az = ad[2], aq = ad[3];
  double acc[KIND == 0 ? 3 : (KIND == 1 ? 8 : 1)];
#pragma unroll
  for (int k = 0; k < (int)(sizeof(acc) / sizeof(double)); ++k) acc[k] = 0.0;
  const size_t N = (size_t)P.nx * P.ny * P.nz;
  const size_t p0 = (size_t)blockIdx.y * slab;
  const size_t p1 = p0 + slab < N ? p0 + slab : N;
  double *px = sp, *py = sp + AS_TILE, *pz = sp + 2 * AS_TILE, *pg = sp + 3 * AS_TILE;
  for (size_t q0 = p0; q0 < p1; q0 += AS_TILE) {
    const int nq = (int)((p1 - q0) < (size_t)AS_TILE ? (p1 - q0) : (size_t)AS_TILE);
    __syncthreads();
    for (int k = threadIdx.x; k < nq; k += blockDim.x) {
      const size_t i = q0 + k;
      const int x = (int)(i % P.nx), y = (int)((i / P.nx) % P.ny), z = (int)(i / ((size_t)P.nx * P.ny));
      px[k] = (double)(x - P.nx / 2) * P.dr[0];
      py[k] = (double)(y - P.ny / 2) * P.dr[1];
      pz[k] = (double)(z - P.nz / 2) * P.dr[2];
    }
    // g of every site for the tile, AS_TILE doubles per site (in two halves
    // of the LDS when nv > 4)
    for (int vh = 0; vh < nv; vh += 4) {
      const int nvh = nv - vh < 4 ? nv - vh : 4;
      for (int k = threadIdx.x; k < nq * nvh; k += blockDim.x) {
        const int w = k / nq, kk = k % nq;
        const size_t i = (size_t)(vh + w) * N + q0 + kk;
        pg[w * AS_TILE + kk] = closure_g(P.clos, t[i], u[i]);
      }
      __syncthreads();
      if (active) {
        for (int w = 0; w < nvh; ++w) {
          const int v = vh + w;
          const double sg = ad[4 + v], eps = ad[4 + nv + v];
          const double cut = (KIND == 2 ? sg * lambda : sg) * 0.5;
          const double cut2 = cut * cut;
          const double sg2 = sg * sg;
          const double qq = cc * aq * P.qv[v];
          const double wv = (KIND == 1) ? 1.0 : rho[v];
          double s0 = 0.0, s1 = 0.0, s2 = 0.0;
          for (int k = 0; k < nq; ++k) {
            const double dx = px[k] - ax, dy = py[k] - ay, dz = pz[k] - az;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < cut2) continue;  // r < sigma / 2 (r = 0 included)
            const double g = pg[w * AS_TILE + k];
            const double ir = rsq_nr(r2);
            const double ir2 = ir * ir;
            const double rs2 = sg2 * ir2, rs6 = rs2 * rs2 * rs2;
            if (KIND == 0) {
              const double f = (eps * 24.0 * rs6 * ir2 * (2.0 * rs6 - 1.0) + qq * ir2 * ir) * g;
              s0 += f * dx;
              s1 += f * dy;
              s2 += f * dz;
            } else if (KIND == 1) {
              s0 += (eps * 4.0 * rs6 * (rs6 - 1.0) + qq * ir) * g;
            } else if (KIND == 2) {
              s0 += eps * 24.0 * rs6 * (2.0 * rs6 - 1.0) / lambda * g;
            } else {
              s0 += qq * ir * g;
            }
          }
          if (KIND == 0) {
            acc[0] += wv * s0;
            acc[1] += wv * s1;
            acc[2] += wv * s2;
          } else if (KIND == 1) {
#pragma unroll
            for (int k = 0; k < 8; ++k)
              if (k == v) acc[k] += s0;
          } else {
            acc[0] += wv * s0;
          }
        }
      }
      __syncthreads();
    }
  }
  if (!active) return;
  double *out = partial + ((size_t)blockIdx.y * natu + a) * width;
  if (KIND == 1) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (k < nv) out[k] = acc[k];
  } else {
#pragma unroll
    for (int k = 0; k < (int)(sizeof(acc) / sizeof(double)); ++k) out[k] = acc[k];
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
  // per atom: x, y, z, q and the x-phase step e^{-i dkx x} (cos, sin), formed
  // once per atom by the thread that stages it
  __shared__ double sa[ATILE * 6];
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
      double dsn, dcs;
      sincos(dkx * src[0], &dsn, &dcs);
      sa[6 * i] = src[0];
      sa[6 * i + 1] = src[1];
      sa[6 * i + 2] = src[2];
      sa[6 * i + 3] = src[3];
      sa[6 * i + 4] = dcs;
      sa[6 * i + 5] = dsn;
    }
    __syncthreads();
    for (int a = 0; a < na; ++a) {
      const double ax = sa[6 * a], ay = sa[6 * a + 1], az = sa[6 * a + 2], aq = sa[6 * a + 3];
      const double dcs = sa[6 * a + 4], dsn = sa[6 * a + 5];
      double sn, cs;
      sincos(kx0 * ax + ky * ay + kz * az, &sn, &cs);
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
template <class Re>
__global__ void k_fields(const Re *__restrict__ t, const Re *__restrict__ u,
                         const Re *__restrict__ R, Params P, size_t n, int which,
                         double *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ti = (double)t[i];
  const double g = closure_g(P.clos, ti, (double)u[i]);
  if (which == 1)
    out[i] = g;
  else if (which == 2)
    out[i] = (double)R[i] + g - 1.0;
  else
    out[i] = (double)R[i] + g - 1.0 - ti;
}

// ---------------------------------------------------------------------------
// thermodynamics (cal_exchem.cu:55-155, cal_pmv.cu:27-57): per-block FP64
// sums of SC (closure-matched), HNC, GF integrands and h - t, per site.
// ---------------------------------------------------------------------------
template <class Re>
__global__ void __launch_bounds__(256)
    k_thermo_partial(const Re *__restrict__ t, const Re *__restrict__ th,
                     const Re *__restrict__ u, const Re *__restrict__ R, Params P,
                     size_t N, double *__restrict__ partial) {
  // t: the current t (the integrand's t); th: the t of the step that produced
  // R, from which h = R + g(th, u) - 1 is rebuilt (equal unless the loop
  // stopped at maxstep, after an AN2 update)
  __shared__ double sh[4][256 / 64];
  const int v = blockIdx.y;
  const Re *tv = t + (size_t)v * N, *uv = u + (size_t)v * N, *Rv = R + (size_t)v * N;
  const Re *thv = th + (size_t)v * N;
  double s[4] = {0, 0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (size_t)gridDim.x * blockDim.x) {
    const double tt = (double)tv[i];
    const double g = closure_g(P.clos, (double)thv[i], (double)uv[i]);
    const double h = (double)Rv[i] + g - 1.0;
    const double ehnc = h * 0.5 * tt - (h - tt);
    const double egf = -(h * 0.5 + 1.0) * (h - tt);
    const double ekh = (h > 0.0) ? egf : ehnc;
    // PSE-n (clos = 10 + n): HNC integrand minus (t*)^(n+1)/(n+1)!, t* = -u + t
    const double epse =
        (P.clos >= 10) ? ehnc - pse_excess(P.clos - 10, (double)thv[i] - (double)uv[i]) : ehnc;
    s[0] += (P.clos == 0) ? ekh : epse;
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

// Plain streaming copy of n 16-byte vectors: the reference point for the HBM
// traffic counters and for the achievable copy bandwidth (reads n, writes n).
__global__ void __launch_bounds__(256)
    k_copy_stream(const V16<double> *__restrict__ in, V16<double> *__restrict__ out, size_t nv) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride)
    st_nt(out + i, ld_nt(in + i));
}

hipError_t run_copy(const double *in, double *out, size_t n, hipStream_t st) {
  if (n % 2) return hipErrorInvalidValue;
  const size_t nv = n / 2;
  // 256 CUs x 8 workgroups, grid-stride beyond that
  const unsigned blocks = (unsigned)std::min<size_t>(nblk(nv, 256), 2048);
  hipLaunchKernelGGL(k_copy_stream, dim3(blocks), dim3(256), 0, st,
                     (const V16<double> *)in, (V16<double> *)out, nv);
  return hipGetLastError();
}

hipError_t run_reduce_cols(const double *in, size_t nrow, int width, int nsets, double *out,
                           const int *done, hipStream_t st, const GroupPtrs *grp, int ngrp) {
  hipLaunchKernelGGL(k_reduce_cols, dim3(width, nsets, grp ? ngrp : 1), dim3(256), 0, st, in,
                     nrow, width, out, done, grp);
  return hipGetLastError();
}

hipError_t run_step_finish(const double *sums, const double *rho, int natv, size_t N,
                           double tol, int mixing, double mp, int binary, double *scal,
                           double *rms_hist, int *done, hipStream_t st, const GroupPtrs *grp,
                           int ngrp) {
  hipLaunchKernelGGL(k_step_finish, dim3(1, 1, grp ? ngrp : 1), dim3(64), 0, st, sums, rho, natv,
                     N, tol, mixing, mp, binary, scal, rms_hist, done, grp);
  return hipGetLastError();
}


template <class Re>
static size_t nvec(size_t n) {
  return n / V16<Re>::K;
}
template <class Re>
hipError_t run_an2_picard(Re *tnew, const Re *t, const Re *R, Re *Pb, double m, size_t n,
                          const int *done, hipStream_t st, const GroupPtrs *grp, int ngrp) {
  if (n % V16<Re>::K) return hipErrorInvalidValue;
  const size_t nv = nvec<Re>(n);
  hipLaunchKernelGGL(k_an2_picard<Re>, dim3(nblk(nv, 256), 1, grp ? ngrp : 1), dim3(256), 0, st,
                     (V16<Re> *)tnew, (const V16<Re> *)t, (const V16<Re> *)R, (V16<Re> *)Pb,
                     (Re)m, nv, done, grp);
  return hipGetLastError();
}

template <class Re>
hipError_t run_an2_update(Re *tnew, const Re *t, const Re *R, const Re *P0, const Re *P1, Re *Pb,
                          const double *scal, double m, size_t n, const int *done,
                          hipStream_t st, const GroupPtrs *grp, int ngrp) {
  if (n % V16<Re>::K) return hipErrorInvalidValue;
  const size_t nv = nvec<Re>(n);
  hipLaunchKernelGGL(k_an2_update<Re>, dim3(nblk(nv, 256), 1, grp ? ngrp : 1), dim3(256), 0, st,
                     (V16<Re> *)tnew, (const V16<Re> *)t, (const V16<Re> *)R,
                     (const V16<Re> *)P0, (const V16<Re> *)P1, (V16<Re> *)Pb, scal, (Re)m, nv,
                     done, grp);
  return hipGetLastError();
}

template <class Re>
hipError_t run_andm_update(Re *tnew, const Re *t, const Re *R, const HistUpdateT<Re> &H, Re *Pb,
                           const double *xcoef, double beta, size_t n, const int *done,
                           hipStream_t st) {
  if ((n % V16<Re>::K) || H.nh < 0 || H.nh > RISM_MAX_HIST) return hipErrorInvalidValue;
  const size_t nv = nvec<Re>(n);
  hipLaunchKernelGGL(k_andm_update<Re>, dim3(nblk(nv, 256)), dim3(256), 0, st, (V16<Re> *)tnew,
                     (const V16<Re> *)t, (const V16<Re> *)R, H, (V16<Re> *)Pb, xcoef, (Re)beta,
                     nv, done);
  return hipGetLastError();
}

hipError_t run_andm_finish(const double *sums, const double *rho, int natv, size_t N, double tol,
                           int width, double *D, int m1, int cur, const int *rb, int nh,
                           int mixing, double *xcoef, double *scal, double *rms_hist, int *done,
                           hipStream_t st) {
  if (nh < 0 || nh > RISM_MAX_HIST || m1 > RISM_MAX_HIST + 1 || width < 1 + nh)
    return hipErrorInvalidValue;
  AndmHist H{};
  H.nh = nh;
  for (int p = 0; p < nh; ++p) H.rb[p] = rb[p];
  hipLaunchKernelGGL(k_andm_finish, dim3(1), dim3(64), 0, st, sums, rho, natv, N, tol, width, D,
                     m1, cur, H, mixing, xcoef, scal, rms_hist, done);
  return hipGetLastError();
}

template <class From, class To>
hipError_t run_convert(const From *in, To *out, size_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_convert<From, To>), dim3(nblk(n, 256)), dim3(256), 0, st, in, out, n);
  return hipGetLastError();
}

#define INST(Re)                                                                              \
  template hipError_t run_an2_picard<Re>(Re *, const Re *, const Re *, Re *, double, size_t,  \
                                         const int *, hipStream_t, const GroupPtrs *, int);   \
  template hipError_t run_an2_update<Re>(Re *, const Re *, const Re *, const Re *, const Re *, \
                                         Re *, const double *, double, size_t, const int *,   \
                                         hipStream_t, const GroupPtrs *, int);                \
  template hipError_t run_andm_update<Re>(Re *, const Re *, const Re *,                       \
                                          const HistUpdateT<Re> &, Re *, const double *,      \
                                          double, size_t, const int *, hipStream_t);          \
  template hipError_t run_fields<Re>(const Re *, const Re *, const Re *, const Params &,      \
                                     size_t, int, double *, hipStream_t);                     \
  template hipError_t run_thermo_partial<Re>(const Re *, const Re *, const Re *, const Re *,  \
                                             const Params &, size_t, double *, hipStream_t);
template hipError_t run_convert<double, float>(const double *, float *, size_t, hipStream_t);
hipError_t run_exponent(const double *t, const double *u, size_t n, float *out, hipStream_t st) {
  hipLaunchKernelGGL(k_exponent, dim3(nblk(n, 256)), dim3(256), 0, st, t, u, n, out);
  return hipGetLastError();
}
hipError_t run_rebase_up(double *out, const double *t0, const float *d, size_t n,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_rebase_up, dim3(nblk(n, 256)), dim3(256), 0, st, out, t0, d, n);
  return hipGetLastError();
}
hipError_t run_rebase_down(const double *x, const double *t0, float *out, size_t n,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_rebase_down, dim3(nblk(n, 256)), dim3(256), 0, st, x, t0, out, n);
  return hipGetLastError();
}
template hipError_t run_convert<float, double>(const float *, double *, size_t, hipStream_t);

hipError_t run_potential_r(const double *atoms, int natu, const Params &P, double ikbt,
                           double ubeta, int do_coul, double *u, double *fr, hipStream_t st) {
  const size_t nth = (size_t)P.ny * P.nz * ((P.nx + PX - 1) / PX);
  const size_t lds = (size_t)ATILE * atom_stride(P.natv) * sizeof(double);
  const dim3 grid(nblk(nth, 256)), block(256);
  for (int v0 = 0; v0 < P.natv; v0 += 8) {
    const int g = min(8, P.natv - v0);
    const int wf = v0 == 0;
    switch (g) {
#define G_CASE(n)                                                                              \
  case n:                                                                                      \
    if (do_coul)                                                                               \
      hipLaunchKernelGGL((k_potential_r<n, true>), grid, block, lds, st, atoms, natu, v0, P,   \
                         ikbt, ubeta, wf, u, fr);                                              \
    else                                                                                       \
      hipLaunchKernelGGL((k_potential_r<n, false>), grid, block, lds, st, atoms, natu, v0, P,  \
                         ikbt, ubeta, wf, u, fr);                                              \
    break;
      G_CASE(1) G_CASE(2) G_CASE(3) G_CASE(4) G_CASE(5) G_CASE(6) G_CASE(7) G_CASE(8)
#undef G_CASE
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
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

template <class Re>
hipError_t run_fields(const Re *t, const Re *u, const Re *R, const Params &P,
                      size_t n, int which, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_fields<Re>, dim3(nblk(n, 256)), dim3(256), 0, st, t, u, R, P, n, which,
                     out);
  return hipGetLastError();
}

int thermo_blocks(size_t N) {
  size_t b = (N + 255) / 256;
  return (int)(b < 1024 ? b : 1024);
}

template <class Re>
hipError_t run_thermo_partial(const Re *t, const Re *th, const Re *u,
                              const Re *R, const Params &P, size_t N, double *partial,
                              hipStream_t st) {
  dim3 grid(thermo_blocks(N), P.natv);
  hipLaunchKernelGGL(k_thermo_partial<Re>, grid, dim3(256), 0, st, t, th, u, R, P, N, partial);
  return hipGetLastError();
}

size_t atom_sum_slab(size_t N) {
  size_t s = N / 1024;
  if (s < 4096) s = 4096;
  return (s + AS_TILE - 1) / AS_TILE * AS_TILE;
}

hipError_t run_atom_sums(int kind, const double *t, const double *u, const Params &P,
                         const double *adata, int natu, double lambda, double cc,
                         const double *rho, double *partial, hipStream_t st) {
  const size_t N = (size_t)P.nx * P.ny * P.nz;
  const size_t slab = atom_sum_slab(N);
  const int width = kind == 0 ? 3 : (kind == 1 ? P.natv : 1);
  dim3 grid((natu + 255) / 256, (unsigned)((N + slab - 1) / slab));
  switch (kind) {
    case 0:
      hipLaunchKernelGGL(k_atom_sums<0>, grid, dim3(256), 0, st, t, u, P, adata, natu, lambda,
                         cc, rho, slab, width, partial);
      break;
    case 1:
      hipLaunchKernelGGL(k_atom_sums<1>, grid, dim3(256), 0, st, t, u, P, adata, natu, lambda,
                         cc, rho, slab, width, partial);
      break;
    case 2:
      hipLaunchKernelGGL(k_atom_sums<2>, grid, dim3(256), 0, st, t, u, P, adata, natu, lambda,
                         cc, rho, slab, width, partial);
      break;
    default:
      hipLaunchKernelGGL(k_atom_sums<3>, grid, dim3(256), 0, st, t, u, P, adata, natu, lambda,
                         cc, rho, slab, width, partial);
      break;
  }
  return hipGetLastError();
}

// The twiddle and phase tables of the three axis passes computed on the
// device (the host computes the same values for a context): consecutive
// arrays roots(nx), roots(ny), roots(M), pre(nx), pre(ny), zpre(M),
// zpost(M), M = nz/2, with roots(L)_q = e^{-2 pi i q/L}, pre(n)_j =
// (-1)^j e^{-i pi j/n}, zpre_p = e^{-i pi p/M}, zpost_p = e^{-i pi (2p+1)/nz}.
__global__ void k_fft_tables(int nx, int ny, int nz, cplx *__restrict__ tw) {
  const int M = nz / 2;
  const int total = 2 * nx + 2 * ny + 3 * M;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int k = i;
  double a;  // angle in units of pi
  double sg = 1.0;
  if (k < nx) {
    a = -2.0 * k / nx;
  } else if ((k -= nx) < ny) {
    a = -2.0 * k / ny;
  } else if ((k -= ny) < M) {
    a = -2.0 * k / M;
  } else if ((k -= M) < nx) {
    a = -(double)k / nx;
    sg = (k & 1) ? -1.0 : 1.0;
  } else if ((k -= nx) < ny) {
    a = -(double)k / ny;
    sg = (k & 1) ? -1.0 : 1.0;
  } else if ((k -= ny) < M) {
    a = -(double)k / M;
  } else {
    k -= M;
    a = -(2.0 * k + 1.0) / nz;
  }
  double sn, cs;
  sincospi(a, &sn, &cs);
  tw[i] = cplx{sg * cs, sg * sn};
}

hipError_t run_fft_tables(int nx, int ny, int nz, cplx *tw, hipStream_t st) {
  const int total = 2 * nx + 2 * ny + 3 * (nz / 2);
  hipLaunchKernelGGL(k_fft_tables, dim3(nblk((size_t)total, 256)), dim3(256), 0, st, nx, ny, nz,
                     tw);
  return hipGetLastError();
}

hipError_t run_apply_phase(cplx *s, const Params &P, int inverse, hipStream_t st) {
  const size_t n = (size_t)(P.nz / 2) * P.ny * P.nx;
  hipLaunchKernelGGL(k_apply_phase, dim3(nblk(n, 256)), dim3(256), 0, st, s, P, inverse);
  return hipGetLastError();
}


INST(double)
INST(float)
#undef INST

}  // namespace rism
