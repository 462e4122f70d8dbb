// rism_kernels.h -- launch interface of the HIP kernels (rism_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#define RISM_MAX_SITES 16

namespace rism {

template <class Re>
struct cx;
using cplx = cx<double>;
using cplxf = cx<float>;


// Everything a kernel needs to know about the grid, the solvent sites and
// the chi spline, passed by value.
// Grouped launches: contexts of one grid shape, solvent and options stepped
// together by one launch per kernel, one context per blockIdx.z.  Each
// kernel then takes the pointers of its context from grp[blockIdx.z]; its
// own pointer arguments are the first context's (the twiddle and chi tables
// are equal across the group).
struct GroupPtrs {
  const void *t, *u, *fr;  // forward z pass (bulk type)
  void *spec;
  const void *fk;  // x pass
  void *R;         // residual pass: output and the two AN2 slot residuals
  const void *R0, *R1;
  double *partial, *sums, *scal, *rms_hist;
  int *done;
  void *tnew;  // AN2 update
  const void *P0, *P1;
  void *Pb;
};

struct Params {
  int nx, ny, nz;
  int natv;
  int clos;          // 0 KH, 1 HNC
  int mode;          // 0 solver step; 1 plain forward transform; 2 plain inverse;
                     // 3 difference-form step (correction sweep of the mixed solve)
  int shell;         // 1: chi from the cubic-box shell table, 0: spline at |k|
  double dv;         // volume element (forward scale)
  double inv_scale;  // 1 / (dv nx ny (nz/2)): inverse scale of the three passes
  double dr[3];      // grid spacing
  double kfac[3];    // 2 pi / box
  double qv[RISM_MAX_SITES];
  // chi spline: knots x_i = x0 + i/idk approximately (exact values in `knots`)
  int np;
  double x0, idk;
  // device flag of a solve driven without host synchronisation: once it is
  // non-zero (converged or diverged) every step kernel returns at entry
  const int *done;
  // launch plans of the forward z, residual z and y passes picked by the
  // autotuner, per bulk type ([0] FP64, [1] FP32); 0 is the default plan
  int var_zf[2], var_zi[2], var_y[2];
  // grouped launch (GroupPtrs): ngrp contexts, or grp = null and ngrp = 0
  const GroupPtrs *grp;
  int ngrp;
};

// Launch plans the autotuner chooses between (rism_ctx.cpp): whether plan
// v (0, 1, 2) of a pass is built for FFT length L (nz/2 for the z passes,
// ny for the y passes) and the bulk type; plan 0 always is.
enum TunedPass { TP_ZFWD = 0, TP_ZINV = 1, TP_Y = 2, TP_COUNT = 3 };
constexpr int kPlans = 3;
bool plan_offered(int pass, int L, int fp32, int v);
bool y_plan_offered(int L, int fp32, int v);

// History of the Anderson(m) accelerator passed to kernels by value.
#define RISM_MAX_HIST 16
template <class Re>
struct HistT {
  const Re *r[RISM_MAX_HIST];  // residual history R_p, p < nh
  int nh;
};
template <class Re>
struct HistUpdateT {
  const Re *p[RISM_MAX_HIST];  // P_p = T_p + beta R_p, p < nh
  int nh;
};

// doubles per solute atom record of the potential kernels: x, y, z, q,
// rc2, sigma_uv^2[natv], 4 eps sigma^12 [natv], 4 eps sigma^6 [natv]
inline __host__ __device__ int atom_stride(int natv) { return 5 + 3 * natv; }

bool fft_length_supported(int L);
// workgroups of the residual pass (hm: with the Anderson(m) sums)
template <class Re>
size_t zinv_num_blocks(const Params &P, int hm);
int xoz_threads(int L, int natv);
int thermo_blocks(size_t N);

// The hot-loop launchers are instantiated for the FP64 path (Re = double)
// and the FP32 bulk of the mixed-precision path (Re = float).  Sums of
// squares, Gram entries and thermodynamic integrals are FP64 in both.
template <class Re>
hipError_t run_zfwd(const Params &P, const Re *t, const Re *u, const Re *fr, cx<Re> *spec,
                    const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost, hipStream_t st);
template <class Re>
hipError_t run_y(const Params &P, cx<Re> *spec, int dir, const cx<Re> *twL, const cx<Re> *pre,
                 hipStream_t st);
// chitab: the cubic-box shell table in the bulk type (P.shell = 1)
template <class Re>
hipError_t run_xoz(const Params &P, cx<Re> *spec, const cx<Re> *fkraw, const cx<Re> *twL,
                   const cx<Re> *prex, const double *knots, const double *coef,
                   const Re *chitab, hipStream_t st);
hipError_t run_chi_table(const double *knots, const double *coef, const Params &P,
                         double kscale, size_t nshell, double *out, hipStream_t st);
// AN2 partial sums (6 per workgroup) of the residual pass
// P.mode 3 (difference form): t = correction d, u = exponent a = -u + t0 at
// the base point, r0 = the full residual at t0; R = r0 + h - (g(a+d) - g(a)).
template <class Re>
hipError_t run_zinv(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                    const Re *R0, const Re *R1, const Re *r0, double *partial,
                    const cx<Re> *twM, const cx<Re> *zpre, const cx<Re> *zpost, int use_hist,
                    hipStream_t st);
// Anderson(m) variant: partial sums R.R and R.R_p for the nh history vectors
// (width 1 + RISM_MAX_HIST per workgroup).
template <class Re>
hipError_t run_zinv_hist(const Params &P, const cx<Re> *spec, const Re *t, const Re *u, Re *R,
                         const Re *r0, const HistT<Re> &H, double *partial, const cx<Re> *twM,
                         const cx<Re> *zpre, const cx<Re> *zpost, hipStream_t st);
// xcoef: the nh mixing coefficients in device memory (run_andm_finish)
template <class Re>
hipError_t run_andm_update(Re *tnew, const Re *t, const Re *R, const HistUpdateT<Re> &H, Re *Pb,
                           const double *xcoef, double beta, size_t n, const int *done,
                           hipStream_t st);
// End of an Anderson(m) step on the device: rms, flags and rms_hist as
// run_step_finish (done may be null); the dots of the residual in buffer
// `cur` with itself and the history buffers rb[0..nh) (oldest first) into
// D[m1][m1]; when mixing, the nh coefficients into xcoef.  sums: [site][width]
// FP64, R.R then R.R_p.
hipError_t run_andm_finish(const double *sums, const double *rho, int natv, size_t N, double tol,
                           int width, double *D, int m1, int cur, const int *rb, int nh,
                           int mixing, double *xcoef, double *scal, double *rms_hist, int *done,
                           hipStream_t st);
// nsets consecutive blocks of nrow x width partial sums -> nsets x width
hipError_t run_reduce_cols(const double *in, size_t nrow, int width, int nsets, double *out,
                           const int *done, hipStream_t st, const GroupPtrs *grp = nullptr,
                           int ngrp = 0);
// End of one AN2 step on the device (cal_rms.cu:20-24 and, when the next
// update mixes, cal_theta + leg of anderson.cu:81-130): from the per-site
// sums (6 per site) it writes scal = {rms, x0, x1, s1, s2}.  With a flag
// array done = {state, step, count} it also counts the step (count + 1),
// writes rms_hist[step] and, when rms <= tol or rms is not finite, sets
// state to 1 or 2 and step; no argument depends on the step number, so a
// run of steps can be captured once and replayed.  binary: the slot the
// next update retires.
hipError_t run_step_finish(const double *sums, const double *rho, int natv, size_t N,
                           double tol, int mixing, double mp, int binary, double *scal,
                           double *rms_hist, int *done, hipStream_t st,
                           const GroupPtrs *grp = nullptr, int ngrp = 0);
template <class Re>
hipError_t run_an2_picard(Re *tnew, const Re *t, const Re *R, Re *Pb, double m, size_t n,
                          const int *done, hipStream_t st, const GroupPtrs *grp = nullptr,
                          int ngrp = 0);
// s1, s2 from scal[3], scal[4] (run_step_finish)
template <class Re>
hipError_t run_an2_update(Re *tnew, const Re *t, const Re *R, const Re *P0, const Re *P1, Re *Pb,
                          const double *scal, double m, size_t n, const int *done,
                          hipStream_t st,
                          const GroupPtrs *grp = nullptr, int ngrp = 0);
template <class From, class To>
hipError_t run_convert(const From *in, To *out, size_t n, hipStream_t st);
// closure exponent a = -u + t at the base point t0 of a correction sweep,
// rounded to FP32
hipError_t run_exponent(const double *t, const double *u, size_t n, float *out, hipStream_t st);
// fields of a correction block, whose FP32 values are differences from the
// FP64 base point t0: up writes out = t0 + d (out may be t0), down writes
// out = x - t0 rounded to FP32
hipError_t run_rebase_up(double *out, const double *t0, const float *d, size_t n,
                         hipStream_t st);
hipError_t run_rebase_down(const double *x, const double *t0, float *out, size_t n,
                           hipStream_t st);
// streaming copy of n doubles (n even): reads n, writes n
hipError_t run_copy(const double *in, double *out, size_t n, hipStream_t st);
hipError_t run_potential_r(const double *atoms, int natu, const Params &P, double ikbt,
                           double ubeta, int do_coul, double *u, double *fr, hipStream_t st);
hipError_t run_fk_half(const double *atoms, int natu, int astride, const Params &P,
                       double ubeta, cplx *fkraw, hipStream_t st);
hipError_t run_init_t(double *t, const double *fr, const Params &P, size_t N, hipStream_t st);
template <class Re>
hipError_t run_fields(const Re *t, const Re *u, const Re *R, const Params &P, size_t n,
                      int which, double *out, hipStream_t st);
template <class Re>
hipError_t run_thermo_partial(const Re *t, const Re *th, const Re *u, const Re *R,
                              const Params &P, size_t N, double *partial, hipStream_t st);
hipError_t run_apply_phase(cplx *s, const Params &P, int inverse, hipStream_t st);
// twiddle and phase tables of the axis passes, computed on the device into
// tw[2 nx + 2 ny + 3 nz/2]: roots(nx), roots(ny), roots(nz/2), pre(nx),
// pre(ny), zpre, zpost (the order of a context's host-computed tables)
hipError_t run_fft_tables(int nx, int ny, int nz, cplx *tw, hipStream_t st);
// per-atom integrals of g (kind 0 grad, 1 euv, 2 ad1, 3 ad2); adata: per atom
// x y z q sigma_uv[natv] eps_uv[natv]; partial: [slab][natu][width] with
// width 3, natv, 1, 1 and atom_sum_slab(N) points per slab
size_t atom_sum_slab(size_t N);
hipError_t run_atom_sums(int kind, const double *t, const double *u, const Params &P,
                         const double *adata, int natu, double lambda, double cc,
                         const double *rho, double *partial, hipStream_t st);

}  // namespace rism
