/*
 * rism3d_c.h -- C ABI of the MI355X-native 3D-RISM solver (librism3d.so).
 *
 * The reference (rismical-dev/3D-RISM-CUDA-MPI) is one executable with no
 * plugin or FFI layer; its in-process hot-path boundary is the set of
 * RISM3D / AN2 / FFT3D member calls made by RISM3D::iterate
 * (src-cuda-mpi/iterate.cu:48-72).  Each entry point below names the
 * reference interface it replaces.  Plain pointers and sizes only; host
 * buffers in, host buffers out; the context owns every device allocation.
 *
 * Errors are returned as codes (the reference calls exit(1) on input errors,
 * read_input.cu:35,46,67, solvent_read.cc:18, solvent_spline.cu:19).
 * Divergence is not an error: it is reported through a non-finite rms and
 * RISM_STATUS_DIVERGED, as the reference does (iterate.cu:50-54,73-76).
 *
 * Threading: one context is used by one host thread at a time; it owns one
 * HIP stream on the device that was current at rism_create().
 */
#ifndef RISM3D_C_H
#define RISM3D_C_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------ */
#define RISM_OK 0
#define RISM_ERR_ARG 1        /* bad argument / unsupported grid           */
#define RISM_ERR_IO 2         /* file could not be opened / parsed         */
#define RISM_ERR_VERSION 3    /* .inp version string mismatch              */
#define RISM_ERR_CLOSURE 4    /* unknown closure keyword                   */
#define RISM_ERR_NATOM 5      /* solute atom count is not a natural number */
#define RISM_ERR_KRANGE 6     /* "insufficient maximal T tabulated"        */
#define RISM_ERR_HIP 7        /* HIP runtime error                          */
#define RISM_ERR_NOMEM 8      /* device allocation failed                   */
#define RISM_ERR_STATE 9      /* call out of order                          */

/* ---- solve status ------------------------------------------------------ */
#define RISM_STATUS_CONVERGED 0
#define RISM_STATUS_DIVERGED 1
#define RISM_STATUS_MAXSTEP 2

/* ---- closures (read_input.cu:38-47; PSE-n is an extension) -------------- */
#define RISM_CLOSURE_KH 0
#define RISM_CLOSURE_HNC 1
/* Kast-Kloss partial series expansion of order n, 1 <= n <= 20 (.inp keyword
 * PSEn or PSE-n); PSE-1 equals KH and PSE-n tends to HNC as n grows. */
#define RISM_CLOSURE_PSE(n) (10 + (n))

/* ---- fields for rism_get_field / rism_set_field ------------------------ */
#define RISM_FIELD_T 0   /* t_uv (r-space), [natv][nz][ny][nx]                */
#define RISM_FIELD_G 1   /* g_uv = closure(t, u) of the current t             */
#define RISM_FIELD_H 2   /* h_uv of the last step                             */
#define RISM_FIELD_R 3   /* residual of the last step (h + 1 - g)             */
#define RISM_FIELD_U 4   /* beta u_uv (LJ + q_v de)                           */
#define RISM_FIELD_FR 5  /* smooth Coulomb fr, [nz][ny][nx]                   */
#define RISM_FIELD_C 6   /* c_uv = h - t of the last step                     */

/* Grid: Cell (cell.h:5-26) minus the MPI decomposition. */
typedef struct {
  int nx, ny, nz;        /* grid points per axis (ce->grid)                 */
  double box[3];         /* box lengths, Angstrom (ce->box)                 */
} rism_grid;

/* Solvent: Solvent (solvent.h:9-30) after site merging (solvent_read.cc:67-79). */
typedef struct {
  int natv;              /* merged sites                                     */
  int ntab;              /* chi table records                                */
  double dr;             /* r spacing; k_i = (i+1) pi / (dr ntab)            */
  double temper;         /* K                                                */
  double xt;             /* isothermal compressibility, 1/Pa                 */
  const double *sig;     /* [natv] Angstrom                                  */
  const double *eps;     /* [natv] J/mol                                     */
  const double *q;       /* [natv] e                                         */
  const double *rho;     /* [natv] 1/Angstrom^3                              */
  const double *xvv;     /* [ntab][natv][natv], X[v2][v1] in file order      */
} rism_solvent;

/* Solute: Solute (solute.h:6-22), coordinates already centred. */
typedef struct {
  int natu;
  const double *sig;     /* [natu] Angstrom */
  const double *eps;     /* [natu] J/mol    */
  const double *q;       /* [natu] e        */
  const double *xyz;     /* [natu][3] Angstrom */
} rism_solute;

/* Options: closure, AN2 parameters (.inp line 5, anderson.h:18-20) and the
 * -l / -e lambda switches (main.cu:27-32, set_ad.cu). */
typedef struct {
  int closure;           /* RISM_CLOSURE_KH / _HNC / _PSE(n)                 */
  int an2_count;         /* Picard warm-up steps                             */
  double an2_m;          /* mixing coefficient                               */
  double an2_mp;         /* extra weight on the newest history entry         */
  int ad_switch;         /* 0 none, 1 = -l (LJ sigma scaling, no Coulomb),   */
                         /* 2 = -e (Coulomb scaling)                          */
  double lambda;         /* the -l / -e value                                 */
  int verbose;           /* print the reference's progress lines            */
  int chi_spline;        /* 1: evaluate chi_vv(|k|) from the spline at every  */
                         /* k; 0: per-shell table when the box is cubic       */
  int anderson_m;        /* 0: the reference's two-slot AN2 (anderson.cu);    */
                         /* 1..16: Anderson mixing over that many history    */
                         /* pairs, beta = an2_m, after an2_count Picard      */
                         /* steps (no counterpart in the reference)          */
  int precision;         /* 0: FP64 throughout (the reference); 1: mixed --   */
                         /* FFTs, closure, OZ and the accelerator's fields in */
                         /* FP32 with FP64 sums, then FP64 steps to tol       */
  double mixed_switch;   /* rms at which the mixed solve leaves plain FP32    */
                         /* steps for blocks of difference-form FP32 steps    */
                         /* between FP64 steps (0: 1e-3); it also leaves when */
                         /* the FP32 rms stops improving                      */
  int mixed_block;       /* FP32 steps per block (0: 48)                      */
} rism_opts;

/* Thermodynamics of the converged state (output.cu:23-46, output_xmu.cu). */
typedef struct {
  int natv;
  double sfe_sc;            /* SFE_SC, J/mol                                  */
  double sfe_sc_hnc;        /* SFE_SC_HNC (KH only), J/mol                    */
  double sfe_gf;            /* SFE_GF, J/mol                                  */
  double pmv;               /* PMV, L/mol                                     */
  double pressure;          /* Pressure, J/m^3                                */
  double correction;        /* Correction_Term, J/mol                         */
  double sfec_sc[16];       /* per site SFEC_SC(v), J/mol                     */
  double sfec_sc_hnc[16];   /* per site SFEC_SC_HNC(v)                        */
  double sfec_gf[16];       /* per site SFEC_GF(v)                            */
} rism_thermo_out;

/* Timing of the last rism_solve (hipEvents on the context's stream), ms. */
typedef struct {
  double setup_ms;          /* potentials + initial t                        */
  double iterate_ms;        /* fixed-point loop                              */
  double thermo_ms;         /* reductions for the .xmu quantities            */
  int steps;
  int status;
  double rms;
  int steps_fp32;           /* steps of a mixed solve taken at FP32         */
} rism_timing;

typedef struct rism_ctx rism_ctx;

/* ---- library ----------------------------------------------------------- */
/* Version string of the .inp format this library reads ("20220301",
 * version.h:5). */
const char *rism_version(void);
/* Text of the last error on this thread. */
const char *rism_last_error(void);

/* ---- context lifetime -------------------------------------------------- */
/* Replaces RISM3D::initialize (initialize.cu:4-44) up to, not including,
 * cal_potential: uploads the solute, builds the k-grid shells and the chi
 * spline table (initialize_g.cu:7-50, solvent_spline.cu:7-96) and allocates
 * every device buffer.  Returns NULL and sets *err on failure. */
rism_ctx *rism_create(const rism_grid *grid, const rism_solvent *solvent,
                      const rism_solute *solute, const rism_opts *opts, int *err);
/* Frees the context and all its device memory. */
void rism_destroy(rism_ctx *ctx);

/* ---- setup ------------------------------------------------------------- */
/* Replaces RISM3D::cal_potential (cal_potential.cu:3-36: cal_LJ, cal_Coulomb,
 * set_du) followed by initialize_tuv (initialize_tuv.cu:7-17). */
int rism_setup_potentials(rism_ctx *ctx);

/* ---- the hot loop ------------------------------------------------------ */
/* Replaces RISM3D::calculate (calculate.cu:4-65) + RISM3D::cal_rms
 * (cal_rms.cu:14-26): one closure / forward transform / OZ / inverse
 * transform / residual step; returns the rms of the residual. */
int rism_calculate(rism_ctx *ctx, double *rms);
/* Replaces AN2::calculate (anderson.cu:43-78): Picard warm-up or the
 * two-slot modified Anderson update of t from the last residual. */
int rism_anderson(rism_ctx *ctx);
/* ---- batches of independent solves (no counterpart in the reference,
 * which runs one solute per process) ------------------------------------ */
/* Runs rism_solve on n distinct contexts at once from `threads` host
 * threads.  FP64 contexts (precision 0, verbose 0) have their steps
 * enqueued in chunks without a host round trip per step (a device flag
 * stops a context's kernels at convergence); AN2 contexts of equal grid,
 * box, solvent, options and accelerator state are stepped in groups of up
 * to 16 that share one launch per kernel, on the first member's stream.
 * Mixed-precision and verbose contexts run rism_solve's step-by-step loop
 * on their worker thread.  The results are those of n separate rism_solve
 * calls.  status/steps/rms: n entries each (may be NULL). */
int rism_solve_batch(rism_ctx **ctx, int n, int maxstep, double tol, int threads,
                     int *status, int *steps, double *rms);
/* Potentials, batched solve and thermodynamics for n contexts (rism_run for
 * a batch); out, status, steps: n entries each (may be NULL). */
int rism_run_batch(rism_ctx **ctx, int n, int maxstep, double tol, int threads,
                   rism_thermo_out *out, int *status, int *steps);

/* Mixed-precision contexts (opts.precision = 1; no counterpart in the
 * reference): moves the loop state between the FP64 level (prec 0) and the
 * FP32 level (prec 1), keeping the accelerator's history.  rism_solve moves
 * to FP64 by itself; rism_get_precision returns the level the next
 * rism_calculate runs at. */
int rism_set_precision(rism_ctx *ctx, int prec);
int rism_get_precision(rism_ctx *ctx);
/* Replaces the loop of RISM3D::iterate (iterate.cu:48-82): calculate, rms,
 * divergence and convergence tests, AN2 update; at most maxstep steps.
 * *status is a RISM_STATUS_* value, *steps the steps taken, *rms the last
 * rms.  Prints the reference's progress lines when opts.verbose != 0. */
int rism_solve(rism_ctx *ctx, int maxstep, double tol, int *status, int *steps,
               double *rms);

/* ---- post-processing --------------------------------------------------- */
/* Replaces cal_pmv + cal_pressure + cal_exchem (output.cu:23-46,
 * cal_exchem.cu:4-52, cal_pmv.cu:5-24, cal_pressure.cu:6-20), with FP64
 * reductions on the device. */
int rism_thermo(rism_ctx *ctx, rism_thermo_out *out);
/* Copies a field (RISM_FIELD_*) of one site (ignored for FR) to host memory
 * of nx*ny*nz doubles.  Replaces the D2H copies of iterate.cu:86-91. */
int rism_get_field(rism_ctx *ctx, int which, int site, double *host);
/* Uploads t (RISM_FIELD_T) or u (RISM_FIELD_U) of one site; used for
 * restart (read_tuv.cu, with the missing H2D copy done) and tests.  A mixed
 * context at its FP32 level keeps the FP64 value of t as well: a solve whose
 * FP32 phase overflows restarts in FP64 from it. */
int rism_set_field(rism_ctx *ctx, int which, int site, const double *host);
/* Timing of the last rism_solve / rism_run. */
int rism_get_timing(rism_ctx *ctx, rism_timing *out);
/* One full timed solve on an existing context: potentials, initial t,
 * loop, thermodynamics (the work of initialize.cu:43 + iterate + output
 * 'm'), all timed with hipEvents on the context stream. */
int rism_run(rism_ctx *ctx, int maxstep, double tol, rism_thermo_out *out,
             rism_timing *timing);

/* ---- per-kernel timing --------------------------------------------------
 * The reference's -DTEST build prints MPI_Wtime splits per stage
 * (initialize.cu:38-40, cal_potential.cu:30-32, iterate.cu:97-104).  Here
 * each hot-path kernel launch can be bracketed by hipEvents on the context
 * stream; durations accumulate per kernel until the next reset. */
#define RISM_K_ZFWD 0     /* closure + z forward (real -> half complex)      */
#define RISM_K_YFWD 1     /* y forward                                        */
#define RISM_K_XOZ 2      /* x forward + OZ + x inverse                       */
#define RISM_K_YINV 3     /* y inverse                                        */
#define RISM_K_ZINV 4     /* z inverse + residual + Gram partial sums         */
#define RISM_K_REDUCE 5   /* FP64 column reductions of the partial sums       */
#define RISM_K_AN2 6      /* AN2 / Picard update                              */
#define RISM_K_POT 7      /* LJ + Coulomb r-space potentials                  */
#define RISM_K_FK 8       /* k-space Coulomb fk                               */
#define RISM_K_THERMO 9   /* thermodynamic partial sums                       */
#define RISM_K_COUNT 10

typedef struct {
  long launches;
  double total_ms;
} rism_kstat;

/* on != 0 enables event timing and clears the counters; 0 disables it. */
int rism_set_kernel_timing(rism_ctx *ctx, int on);
/* Copies RISM_K_COUNT entries (indexed by RISM_K_*) into out. */
int rism_get_kernel_stats(rism_ctx *ctx, rism_kstat *out);

/* ---- input files (host only, no device use) ---------------------------- */
/* A parsed .inp/.xvk pair (read_input.cu:21-97, solvent_read.cc:10-94,
 * solute.cu:18-54).  Arrays are owned by the library. */
typedef struct {
  char outlist[64];
  int ksave;
  int closure;
  char fsolvent[1024];
  double convergence;
  int maxstep;
  int an2_count;
  double an2_m, an2_mp;
  double box[3];
  int grid[3];
  int natu;
  double *usig, *ueps, *uq, *uxyz; /* solute, centred when requested */
  double shift[3];                 /* centring shift (solute.cu:36-38) */
  /* solvent */
  int natv, ntab;
  double dr, temper, xt;
  double *vsig, *veps, *vq, *vrho;
  double *xvv;
} rism_input;

/* Reads the control file (and the solute file when structure != NULL or
 * ""), applies centring (unless center == 0) and -z charge zeroing, then
 * reads the .xvk named in it (relative to the working directory, like the
 * reference, or to solvent_dir when given).  Returns a RISM_ERR_* code. */
int rism_read_input(const char *control, const char *structure, int center, int q0,
                    const char *solvent_dir, rism_input **out);
void rism_free_input(rism_input *in);
/* Parses one .xvk file into the solvent part of a rism_input. */
int rism_read_xvk(const char *path, rism_input **out);
/* Writes <base>.xmu (or .xmu0) in the format of output_xmu.cu:28-79. */
int rism_write_xmu(const char *path, int closure, const rism_thermo_out *th);
/* Writes the binary .guv file of output_guv.cu:16-47 for the context's
 * final g. */
int rism_write_guv(rism_ctx *ctx, const char *path, const double shift[3]);
/* Writes the text .tuv checkpoint of write_tuv.cu:19-28. */
int rism_write_tuv(rism_ctx *ctx, const char *path);
/* Replace RISM3D::output_huv / output_cuv (output_huv.cu:7-50,
 * output_cuv.cu:7-50; the CLI names the files <base>.huv0 / <base>.cuv0 as
 * the reference's rank 0 does): text dumps of h and c = h - t of the last
 * step with the solute header; .cuv values in the reference's "(re,im)" form. */
int rism_write_huv(rism_ctx *ctx, const char *path);

/* ---- analysis outputs (output.cu:52-105) -------------------------------- */
#define RISM_ANALYSIS_GRAD 0 /* cal_grad + output_grad: natu x 3, kcal/mol/A  */
#define RISM_ANALYSIS_EUV 1  /* cal_euv + output_euv: natu x (1 + natv), J/mol:
                                the atom's total, then per solvent site      */
#define RISM_ANALYSIS_AD 2   /* cal_ad1 (-l) or cal_ad2 + output_ad: natu,
                                kcal/mol                                    */
/* Per-atom integrals of the g of the last FP64 step (cal_grad.cu,
 * cal_euv.cu, cal_ad1.cu, cal_ad2.cu), scaled as the reference writes them
 * to .gra / .euv / <base>-<switch>-<lambda>. */
int rism_analysis(rism_ctx *ctx, int kind, double *out);
int rism_write_cuv(rism_ctx *ctx, const char *path);
/* Reads a .tuv checkpoint and uploads it as t (read_tuv.cu:6-29, with the
 * H2D copy the reference omits). */
int rism_read_tuv(rism_ctx *ctx, const char *path);

/* ---- kernel-level entry points (parity checks of single stages) -------- */
/* Forward half-shifted 3D transform of a real field (FFT3D::execute(-1),
 * fft3d.cu:81-120) on the device: out holds the k_z > 0 half of the
 * spectrum, [nz/2][ny][nx] complex (interleaved), reference convention
 * (dv * sum f e^{-ik.r}); entry (m, jy, jx) is reference k-index
 * (jx, jy, m + nz/2). */
int rism_k_fft_r2c(int nx, int ny, int nz, const double box[3], const double *in,
                   double *out);
/* Inverse (FFT3D::execute(+1), fft3d.cu:123-163) of such a half spectrum
 * back to a real field, 1/V normalisation. */
int rism_k_fft_c2r(int nx, int ny, int nz, const double box[3], const double *in,
                   double *out);
/* The k-space Coulomb term fk on the same half spectrum (cal_Coulomb.cu:65-79,
 * 90-95), reference convention, for a context after rism_setup_potentials. */
int rism_k_get_fk(rism_ctx *ctx, double *out);
/* AN2 scalars of the last Anderson step: x[2] (leg solution), s[2] (s1, s2),
 * count (remaining warm-up steps), binary (next slot). */
int rism_k_an2_state(rism_ctx *ctx, double *x, double *s, int *count, int *binary);

/* ---- diagnostics (no counterpart in the reference) ---------------------- */
/* Copies `bytes` (rounded down to whole 16-byte vectors) between two device
 * buffers of its own `reps` times with a plain streaming kernel and returns
 * the average time per copy: the reference point for HBM traffic counters
 * (each copy reads and writes `bytes`) and for the achievable copy
 * bandwidth.  Allocates and frees its buffers. */
int rism_k_copy_bench(size_t bytes, int reps, double *avg_ms);

/* Autotuner report (no counterpart in the reference, whose transforms are
 * cuFFT plans, fft3d.cu:61-63).  The forward z, residual z and y passes are
 * built in up to three launch plans; each context creation picks, per grid
 * shape and bulk type, the fastest plan by timing them on its buffers (once
 * per process and shape; RISM_TUNE=0 keeps the default plans,
 * RISM_TUNE_FILE=<path> persists the choices, RISM_TUNE_PLANS=zf,zi,y forces
 * plans without timing where they are built).  Writes up to cap entries
 * (plan 0 = the built-in default, ms[v] < 0 = plan v not built for that
 * length) and returns how many there are. */
typedef struct {
  int pass; /* 0 forward z, 1 residual z, 2 y */
  int nx, ny, nz, natv, fp32;
  int choice;
  float ms[3]; /* mean launch time of each plan */
} rism_tune_entry;
int rism_tune_report(rism_tune_entry *out, int cap);

/* The transform of rism_k_fft_r2c (dir = -1) or rism_k_fft_c2r (dir = +1)
 * on caller-owned DEVICE buffers, enqueued on `stream` (a hipStream_t, NULL
 * for the default stream), allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight loops, harnesses that forbid hipMalloc).  real: nx*ny*nz
 * doubles [z][y][x]; spec: the k_z > 0 half spectrum, nx*ny*nz/2 complex
 * (as rism_k_fft_r2c writes it).  dir = -1 reads real and writes spec;
 * dir = +1 reads spec, uses it as workspace (it is overwritten) and writes
 * real.  The tables and partial sums live in a static device array of
 * 2 MiB per GPU (kept until the process ends): grids whose partial sums
 * exceed 1.875 MiB (beyond 512^3) are refused with RISM_ERR_ARG.  Calls
 * that overlap on one GPU share that array and race.  Returns once the
 * work is enqueued. */
int rism_k_fft_noalloc(int dir, int nx, int ny, int nz, const double box[3], double *real,
                       double *spec, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* RISM3D_C_H */
