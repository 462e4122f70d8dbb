// rism_ctx.cpp -- the C ABI (include/rism3d_c.h): context, set-up, the
// fixed-point loop, AN2 and the thermodynamics, driving rism_kernels.hip.
//
// Host-side counterpart of the reference's RISM3D orchestration
// (initialize.cu, iterate.cu, calculate.cu, anderson.cu, output.cu) for one
// solve on one GPU.  The context owns every device buffer and one stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/rism3d_c.h"
#include "rism_fft.h"
#include "rism_internal.h"
#include "rism_kernels.h"

namespace {

thread_local std::string g_err;

const double kAvogadro = 6.02214076e23;          // physical.h:5
const double kBohr = 0.5291771500001;            // physical.h:8
const double kBoltzmann = 1.380649e-23;          // physical.h:11
const double kHartree = 4.359744722207185e-18;   // physical.h:14
// width of the per-workgroup partial sums of the residual pass
const int kSumWidth = 1 + RISM_MAX_HIST;

#define HIPCHECK(expr)                                                                \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return rism::set_error(RISM_ERR_HIP, std::string(#expr) + ": " +               \
                                               hipGetErrorString(e_));                \
  } while (0)

using rism::cplx;
using rism::cplxf;
using rism::cx;

// Natural cubic spline through (xp, yp): the same tridiagonal elimination as
// spline.cc:4-43 (second derivatives s, then per-interval coefficients
// coe0 (linear), coe1 (quadratic), coe2 (cubic)).
void natural_spline(const double *xp, const double *yp, int np, double *c0, double *c1,
                    double *c2) {
  std::vector<double> a(np), b(np), c(np), d(np), s(np);
  const int n = np - 2;
  for (int k = 1; k <= n; ++k) {
    const double hl = xp[k] - xp[k - 1], hr = xp[k + 1] - xp[k];
    a[k] = hl / (hl + hr);
    b[k] = 1.0 - a[k];
    const double g1 = (yp[k + 1] - yp[k]) / hr, g2 = (yp[k] - yp[k - 1]) / hl;
    c[k] = 6.0 * ((g1 - g2) / (hl + hr));
    d[k] = 2.0;
  }
  for (int k = 1; k < n; ++k) {
    b[k] = b[k] / d[k];
    c[k] = c[k] / d[k];
    d[k + 1] = d[k + 1] - a[k + 1] * b[k];
    c[k + 1] = c[k + 1] - a[k + 1] * c[k];
  }
  if (n >= 1) s[n] = c[n] / d[n];
  for (int k = n - 1; k >= 1; --k) s[k] = c[k] - b[k] * s[k + 1];
  s[0] = 0.0;
  s[n + 1] = 0.0;
  for (int k = 0; k <= n; ++k) {
    const double h = xp[k + 1] - xp[k];
    c1[k] = s[k] / 2.0;
    c2[k] = (s[k + 1] - s[k]) / (6.0 * h);
    c0[k] = (yp[k + 1] - yp[k]) / h - c1[k] * h - c2[k] * h * h;
  }
}

// k component of grid index j (initialize_g.cu:58-60)
inline double kcomp(int j, int n, double b) {
  return 2.0 * M_PI * ((double)j - (double)n / 2.0 + 0.5) / b;
}

// Twiddle and phase tables of the three axis passes in one bulk type.
template <class Re>
struct TablesT {
  cx<Re> *twx = nullptr, *twy = nullptr, *twz = nullptr;
  cx<Re> *prex = nullptr, *prey = nullptr, *zpre = nullptr, *zpost = nullptr;
};

// The device fields of the fixed-point loop in one bulk type: FP64 (the
// reference's precision) or FP32 (the bulk of the mixed-precision solve,
// before it hands over to FP64).
template <class Re>
struct Level {
  Re *u = nullptr, *fr = nullptr, *T[2] = {}, *Pslot[2] = {}, *R[3] = {};
  cx<Re> *fk = nullptr, *spec = nullptr;
  std::vector<Re *> Rb, Pb;  // Anderson(m) residuals and mixed iterates
  Re *Rlast = nullptr;       // residual written by the last step at this level
  Re *aux = nullptr;         // FP32 level: the FP64 residual r0 of a correction sweep
  Re *chitab = nullptr;      // chi_vv shell table in this level's type
  TablesT<Re> tb;
};

std::vector<cplx> roots(int L) {  // e^{-2 pi i q / L}
  std::vector<cplx> w(L);
  for (int q = 0; q < L; ++q) {
    const double a = -2.0 * M_PI * (double)q / (double)L;
    w[q] = cplx{std::cos(a), std::sin(a)};
  }
  return w;
}

std::vector<cplx> axis_pre(int n) {  // (-1)^j e^{-i pi j / n}
  std::vector<cplx> w(n);
  for (int j = 0; j < n; ++j) {
    const double sg = (j & 1) ? -1.0 : 1.0;
    const double a = -M_PI * (double)j / (double)n;
    w[j] = cplx{sg * std::cos(a), sg * std::sin(a)};
  }
  return w;
}

bool grid_ok(int nx, int ny, int nz, std::string &why) {
  if (nx < 8 || ny < 8 || nz < 16 || (nx & 1) || (ny & 1) || (nz & 1)) {
    why = "grid sizes must be even, nx, ny >= 8 and nz >= 16";
    return false;
  }
  if (!rism::fft_length_supported(nx) || !rism::fft_length_supported(ny) ||
      !rism::fft_length_supported(nz / 2)) {
    why = "unsupported FFT length (nx, ny and nz/2 must be 2^a 3^b 5^c in the built set)";
    return false;
  }
  return true;
}

}  // namespace

int rism::set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct rism_ctx {
  int dev = 0;
  hipStream_t st = nullptr;
  rism::Params P{};
  int nx = 0, ny = 0, nz = 0, natv = 0;
  size_t N = 0;
  double box[3] = {0, 0, 0};
  std::vector<double> rho, qv, sigv, epsv;
  double temper = 0, xt = 0;
  int natu = 0;
  std::vector<double> atoms_host;  // per atom record, rism::atom_stride(natv) doubles
  std::vector<double> usig, ueps, uq, uxyz;  // the solute as given (for .huv/.cuv headers)
  rism_opts opts{};
  // device: the FP64 level, and the FP32 level of a mixed-precision context
  Level<double> d;
  Level<float> f;
  int prec = 0;  // level the loop runs at: 0 FP64, 1 FP32
  double *partial = nullptr, *tpartial = nullptr, *sums = nullptr, *scratch = nullptr;
  double *atoms = nullptr, *knots = nullptr, *coef = nullptr, *chitab = nullptr;
  size_t nchitab = 0;  // doubles in chitab
  size_t zinv_blocks = 0;
  std::vector<void *> allocs;
  // AN2 state (anderson.h:18-40)
  int count = 0, binary = 0;
  double lastx[2] = {0, 0}, lasts[2] = {0, 0};
  // t ping-pongs between T[0] and T[1]; slot p holds P_p = T_p + m R_p in
  // Pslot[p] and R_p in R[iR_slot[p]] (rotated by pointer)
  int iT_cur = 0, iR_cur = 0, iR_slot[2] = {1, 2};
  int iT_last = 0;
  int last_prec = 0;  // level of the last rism_calculate
  bool fresh = false;  // no rism_anderson since that rism_calculate
  // Anderson(m) state (opts.anderson_m > 0): residuals in Rb[m + 1], mixed
  // iterates P_p = t_p + beta R_p in Pb[m]; hist lists the (R, P) buffer
  // pairs oldest first.  On the device: dD, the weighted dots <R_a, R_b> by
  // R buffer ((m + 1)^2), and xcoef, the mixing coefficients of the next
  // update (both written by the step's finish kernel).
  struct HEnt {
    int rb, pb;
  };
  std::vector<HEnt> hist;
  int cur_rb = 0, next_pb = 0;
  double *dD = nullptr, *xcoef = nullptr;
  // device-side step state: scal = {rms, x0, x1, s1, s2} of the last step,
  // dflag = {done (1 converged, 2 diverged), step}, rms_hist[step]; hscal
  // and hflag are their pinned host mirrors
  double *scal = nullptr, *rms_hist = nullptr, *rho_dev = nullptr, *hscal = nullptr;
  int *dflag = nullptr, *hflag = nullptr;
  // FNV-1a hash of the solvent's data (sites, chi table, T, compressibility):
  // contexts with equal grids, boxes, options and solvents share every table
  // and step together in batches
  uint64_t solv_hash = 0;
  size_t rms_cap = 0;
  bool async = false;
  // one period of mixing steps captured as a graph (see async_enqueue), with
  // the state it was captured from
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  struct GraphKey {
    int iT, iR, s0, s1, b, prec;
    double tol;
    const double *hist;
    bool operator==(const GraphKey &o) const {
      return iT == o.iT && iR == o.iR && s0 == o.s0 && s1 == o.s1 && b == o.b &&
             prec == o.prec && tol == o.tol && hist == o.hist;
    }
  } gkey{};
  bool have_step = false;
  rism_timing timing{};
  // per-kernel hipEvent timing (rism_set_kernel_timing)
  bool ktime = false;
  std::vector<hipEvent_t> evpool;
  size_t evnext = 0;
  struct Pending {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  rism_kstat kstat[RISM_K_COUNT] = {};

  template <class Tp>
  int alloc(Tp **p, size_t count_elems) {
    void *q = nullptr;
    const size_t bytes = sizeof(Tp) * (count_elems ? count_elems : 1);
    if (hipMalloc(&q, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return rism::set_error(RISM_ERR_NOMEM,
                             "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    allocs.push_back(q);
    *p = (Tp *)q;
    return RISM_OK;
  }
  ~rism_ctx() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (hscal) (void)hipHostFree(hscal);
    if (hflag) (void)hipHostFree(hflag);
    for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    for (void *q : allocs) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {
template <class Re>
Level<Re> &lv(rism_ctx *c);
template <>
Level<double> &lv<double>(rism_ctx *c) {
  return c->d;
}
template <>
Level<float> &lv<float>(rism_ctx *c) {
  return c->f;
}
}  // namespace

namespace {

template <class Tp>
int upload(rism_ctx *c, Tp **dst, const std::vector<Tp> &src) {
  int rc = c->alloc(dst, src.size());
  if (rc) return rc;
  HIPCHECK(hipMemcpy(*dst, src.data(), sizeof(Tp) * src.size(), hipMemcpyHostToDevice));
  return RISM_OK;
}

// tables computed in FP64 and rounded once to the bulk type
template <class Re>
std::vector<cx<Re>> round_to(const std::vector<cplx> &w) {
  std::vector<cx<Re>> o(w.size());
  for (size_t i = 0; i < w.size(); ++i) o[i] = cx<Re>{(Re)w[i].x, (Re)w[i].y};
  return o;
}

template <class Re>
int make_tables(rism_ctx *c, TablesT<Re> &tb) {
  const int M = c->nz / 2;
  int rc;
  if ((rc = upload(c, &tb.twx, round_to<Re>(roots(c->nx))))) return rc;
  if ((rc = upload(c, &tb.twy, round_to<Re>(roots(c->ny))))) return rc;
  if ((rc = upload(c, &tb.twz, round_to<Re>(roots(M))))) return rc;
  if ((rc = upload(c, &tb.prex, round_to<Re>(axis_pre(c->nx))))) return rc;
  if ((rc = upload(c, &tb.prey, round_to<Re>(axis_pre(c->ny))))) return rc;
  std::vector<cplx> zp(M), zq(M);
  for (int p = 0; p < M; ++p) {
    const double a = -M_PI * (double)p / (double)M;
    zp[p] = cplx{std::cos(a), std::sin(a)};
    const double b = -M_PI * (double)(2 * p + 1) / (double)c->nz;
    zq[p] = cplx{std::cos(b), std::sin(b)};
  }
  if ((rc = upload(c, &tb.zpre, round_to<Re>(zp)))) return rc;
  if ((rc = upload(c, &tb.zpost, round_to<Re>(zq)))) return rc;
  return RISM_OK;
}

void fill_params(rism_ctx *c) {
  rism::Params &P = c->P;
  P.nx = c->nx;
  P.ny = c->ny;
  P.nz = c->nz;
  P.natv = c->natv;
  P.mode = 0;
  const double volume = c->box[0] * c->box[1] * c->box[2];
  P.dv = volume / (double)c->N;  // cell.cc:13-15
  P.inv_scale = 1.0 / (P.dv * (double)c->nx * (double)c->ny * (double)(c->nz / 2));
  P.dr[0] = c->box[0] / c->nx;
  P.dr[1] = c->box[1] / c->ny;
  P.dr[2] = c->box[2] / c->nz;
  for (int d = 0; d < 3; ++d) P.kfac[d] = 2.0 * M_PI / c->box[d];
  for (int v = 0; v < RISM_MAX_SITES; ++v) P.qv[v] = v < c->natv ? c->qv[v] : 0.0;
}

// chi spline on the knots that bracket the grid's |k| range
// (solvent_spline.cu:17-55), uploaded as [interval][coef][v][w].
int make_chi(rism_ctx *c, const rism_solvent *s) {
  double k2min = 0.0, k2max = 0.0;
  {
    // |k_j| is smallest at j = n/2 and largest at j = 0 on every axis
    const double kx0 = kcomp(c->nx / 2, c->nx, c->box[0]), ky0 = kcomp(c->ny / 2, c->ny, c->box[1]),
                 kz0 = kcomp(c->nz / 2, c->nz, c->box[2]);
    const double kx1 = kcomp(0, c->nx, c->box[0]), ky1 = kcomp(0, c->ny, c->box[1]),
                 kz1 = kcomp(0, c->nz, c->box[2]);
    k2min = kx0 * kx0 + ky0 * ky0 + kz0 * kz0;
    k2max = kx1 * kx1 + ky1 * ky1 + kz1 * kz1;
  }
  const double gmin = std::sqrt(k2min), gmax = std::sqrt(k2max);
  const int ntab = s->ntab;
  const double dk = M_PI / (s->dr * ntab);
  std::vector<double> ttab(ntab);
  for (int i = 0; i < ntab; ++i) ttab[i] = dk * (i + 1);
  if (gmax > ttab[ntab - 1]) return rism::set_error(RISM_ERR_KRANGE, "insufficient maximal T tabulated");
  int ntab1 = 0;
  for (int n = 0; n < ntab; ++n)
    if (ttab[n] < gmin) ntab1 = n;
  int ntab2 = ntab - 1;
  for (int n = ntab - 1; n > 0; --n)
    if (ttab[n] > gmax) ntab2 = n;
  const int np = ntab2 - ntab1 + 1;
  if (np < 3) return rism::set_error(RISM_ERR_KRANGE, "chi table has fewer than 3 knots in range");
  const int nv = c->natv, nv2 = nv * nv;
  std::vector<double> coef((size_t)(np - 1) * 4 * nv2), xk(np), yk(np), a0(np), a1(np), a2(np);
  for (int n = 0; n < np; ++n) xk[n] = ttab[ntab1 + n];
  for (int v = 0; v < nv; ++v)
    for (int w = 0; w < nv; ++w) {
      for (int n = 0; n < np; ++n) yk[n] = s->xvv[((size_t)(ntab1 + n) * nv + v) * nv + w];
      natural_spline(xk.data(), yk.data(), np, a0.data(), a1.data(), a2.data());
      for (int i = 0; i < np - 1; ++i) {
        double *cf = coef.data() + (size_t)i * 4 * nv2 + v * nv + w;
        cf[0] = yk[i];
        cf[nv2] = a0[i];
        cf[2 * nv2] = a1[i];
        cf[3 * nv2] = a2[i];
      }
    }
  int rc;
  if ((rc = upload(c, &c->knots, xk))) return rc;
  if ((rc = upload(c, &c->coef, coef))) return rc;
  c->P.np = np;
  c->P.x0 = xk[0];
  c->P.idk = 1.0 / dk;
  // cubic box: every |k|^2 is (pi/L)^2 (ox^2 + oy^2 + oz^2) with odd o, i.e.
  // (pi/L)^2 (8 i + 3); tabulate chi once per shell i
  c->P.shell = 0;
  if (!c->opts.chi_spline && c->box[0] == c->box[1] && c->box[1] == c->box[2]) {
    const long smax = (long)(c->nx - 1) * (c->nx - 1) + (long)(c->ny - 1) * (c->ny - 1) +
                      (long)(c->nz - 1) * (c->nz - 1);
    const size_t nshell = (size_t)((smax - 3) / 8 + 1);
    if ((rc = c->alloc(&c->chitab, nshell * nv2))) return rc;
    c->nchitab = nshell * nv2;
    HIPCHECK(rism::run_chi_table(c->knots, c->coef, c->P, M_PI / c->box[0], nshell, c->chitab, c->st));
    HIPCHECK(hipStreamSynchronize(c->st));
    c->P.shell = 1;
  }
  return RISM_OK;
}

int check_sizes(int nx, int ny, int nz, int natv) {
  std::string why;
  if (!grid_ok(nx, ny, nz, why)) return rism::set_error(RISM_ERR_ARG, why);
  if (natv < 1 || natv > RISM_MAX_SITES)
    return rism::set_error(RISM_ERR_ARG, "natv must be in [1, 16]");
  // the x pass keeps one row of every site in one workgroup, with the
  // per-site OZ values in registers
  if (natv > 8 || rism::xoz_threads(nx, natv) > 512)
    return rism::set_error(RISM_ERR_ARG, "natv must be <= 8 (and <= 4 for nx = 1024)");
  return RISM_OK;
}

void reset_an2(rism_ctx *c) {
  c->count = c->opts.an2_count;
  c->binary = 0;
  c->iT_cur = 0;
  c->iR_cur = 0;
  c->iR_slot[0] = 1;
  c->iR_slot[1] = 2;
  c->iT_last = 0;
  c->d.Rlast = nullptr;
  c->f.Rlast = nullptr;
  c->last_prec = c->prec;
  c->fresh = false;
  c->have_step = false;
  c->hist.clear();
  c->cur_rb = 0;
  c->next_pb = 0;
  c->lastx[0] = c->lastx[1] = c->lasts[0] = c->lasts[1] = 0.0;
}

// Accumulates the durations of every bracketed launch; the stream must have
// been synchronised (or the events completed) before this is called.
void ktime_resolve(rism_ctx *c) {
  for (const auto &p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->kstat[p.id].launches += 1;
      c->kstat[p.id].total_ms += ms;
    } else {
      (void)hipGetLastError();
    }
  }
  c->pending.clear();
  c->evnext = 0;
}

hipEvent_t ktime_event(rism_ctx *c) {
  if (c->evnext == c->evpool.size()) {
    if (c->evpool.size() >= 256) {  // pool exhausted: drain and reuse
      (void)hipStreamSynchronize(c->st);
      ktime_resolve(c);
    } else {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      c->evpool.push_back(e);
    }
  }
  return c->evpool[c->evnext++];
}

// RAII bracket: events around one launch when timing is on
struct KTime {
  rism_ctx *c;
  int id;
  hipEvent_t a = nullptr;
  KTime(rism_ctx *c_, int id_) : c(c_), id(id_) {
    if (c->ktime && (a = ktime_event(c))) (void)hipEventRecord(a, c->st);
  }
  ~KTime() {
    if (!a) return;
    hipEvent_t b = ktime_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->st);
    c->pending.push_back({id, a, b});
  }
};

// One Anderson(m) step: the first an2_count steps are the reference's
// Picard steps t' = t + R (anderson.cu:54-60); after that
//   t' = Q + sum_p th_p (P_p - Q),  Q = t + beta R,  beta = an2_m.
// Either way the new pair (R, Q) enters the history, retiring the oldest
// pair once m are held.
template <class Re>
int anderson_m_step(rism_ctx *c) {
  Level<Re> &L = lv<Re>(c);
  const size_t nn = c->N * c->natv;
  const int m = c->opts.anderson_m, m1 = m + 1;
  Re *t = L.T[c->iT_cur], *tnew = L.T[1 - c->iT_cur];
  (void)m1;
  Re *R = L.Rb[c->cur_rb];
  // slot for the new pair: the oldest when full, else a fresh one
  const bool full = (int)c->hist.size() == m;
  const rism_ctx::HEnt old = full ? c->hist.front() : rism_ctx::HEnt{-1, c->next_pb};
  Re *Pdst = L.Pb[old.pb];
  const int *done = c->async ? c->dflag : nullptr;
  if (c->count > 0) {
    --c->count;
    KTime kt(c, RISM_K_AN2);
    HIPCHECK(rism::run_an2_picard(tnew, t, R, Pdst, c->opts.an2_m, nn, done, c->st));
  } else {
    // coefficients from the step's finish kernel (run_andm_finish)
    rism::HistUpdateT<Re> H{};
    H.nh = (int)c->hist.size();
    for (int p = 0; p < H.nh; ++p) H.p[p] = L.Pb[c->hist[p].pb];
    KTime kt(c, RISM_K_AN2);
    HIPCHECK(rism::run_andm_update(tnew, t, R, H, Pdst, c->xcoef, c->opts.an2_m, nn, done,
                                   c->st));
  }
  if (full) c->hist.erase(c->hist.begin());
  else ++c->next_pb;
  c->hist.push_back({c->cur_rb, old.pb});
  // next residual buffer: the retired one, else the lowest free index
  if (full) {
    c->cur_rb = old.rb;
  } else {
    std::vector<char> used(m1, 0);
    for (const auto &h : c->hist) used[h.rb] = 1;
    int fi = 0;
    while (used[fi]) ++fi;
    c->cur_rb = fi;
  }
  c->iT_cur = 1 - c->iT_cur;
  return RISM_OK;
}

// One AN2::calculate (anderson.cu:43-78) at level Re: Picard warm-up, or
// the two-slot update with the coefficients that the step's finish kernel
// left in scal (cal_theta + leg on the device).
template <class Re>
int an2_step(rism_ctx *c) {
  Level<Re> &L = lv<Re>(c);
  const size_t nn = c->N * c->natv;
  const int b = c->binary;
  const int oldR = c->iR_slot[b];
  const int *done = c->async ? c->dflag : nullptr;
  Re *t = L.T[c->iT_cur], *R = L.R[c->iR_cur], *tnew = L.T[1 - c->iT_cur];
  if (c->count > 0) {  // anderson.cu:54-60
    --c->count;
    KTime kt(c, RISM_K_AN2);
    HIPCHECK(rism::run_an2_picard(tnew, t, R, L.Pslot[b], c->opts.an2_m, nn, done, c->st));
  } else {  // anderson.cu:61-72
    KTime kt(c, RISM_K_AN2);
    HIPCHECK(rism::run_an2_update(tnew, t, R, L.Pslot[0], L.Pslot[1], L.Pslot[b], c->scal,
                                  c->opts.an2_m, nn, done, c->st));
  }
  // slot[b] <- (t, R): P_b was rewritten by the kernel, R_b by pointer
  c->iR_slot[b] = c->iR_cur;
  c->iR_cur = oldR;
  c->iT_cur = 1 - c->iT_cur;
  c->binary = (b == 0) ? 1 : 0;
  return RISM_OK;
}

// One calculate() + cal_rms() at level Re (calculate.cu:4-65, cal_rms.cu:14-26)
// lin: the difference-form step of a correction sweep (t = the correction
// d, u = the closure exponent at t0, aux = the residual at t0).
// In async mode (AN2 only) nothing is read back: the finish kernel records
// rms_hist[step] and raises the done flag, which stops every later kernel.
template <class Re>
int calc_step(rism_ctx *c, double *rms, bool lin = false, double tol = -1.0) {
  Level<Re> &L = lv<Re>(c);
  rism::Params P = c->P;
  P.mode = lin ? 3 : 0;
  P.done = c->async ? c->dflag : nullptr;
  Re *t = L.T[c->iT_cur];
  const Re *r0 = lin ? L.aux : nullptr;
  {
    KTime kt(c, RISM_K_ZFWD);
    HIPCHECK(rism::run_zfwd(P, t, L.u, L.fr, L.spec, L.tb.twz, L.tb.zpre, L.tb.zpost, c->st));
  }
  {
    KTime kt(c, RISM_K_YFWD);
    HIPCHECK(rism::run_y(P, L.spec, -1, L.tb.twy, L.tb.prey, c->st));
  }
  {
    KTime kt(c, RISM_K_XOZ);
    HIPCHECK(rism::run_xoz(P, L.spec, L.fk, L.tb.twx, L.tb.prex, c->knots, c->coef, L.chitab,
                           c->st));
  }
  {
    KTime kt(c, RISM_K_YINV);
    HIPCHECK(rism::run_y(P, L.spec, +1, L.tb.twy, L.tb.prey, c->st));
  }
  const bool andm = c->opts.anderson_m > 0;
  const int W = andm ? kSumWidth : 6;
  Re *Rout = andm ? L.Rb[c->cur_rb] : L.R[c->iR_cur];
  {
    KTime kt(c, RISM_K_ZINV);
    if (andm) {
      rism::HistT<Re> H{};
      H.nh = (int)c->hist.size();
      for (int p = 0; p < H.nh; ++p) H.r[p] = L.Rb[c->hist[p].rb];
      HIPCHECK(rism::run_zinv_hist(P, L.spec, t, L.u, Rout, r0, H, c->partial, L.tb.twz,
                                   L.tb.zpre, L.tb.zpost, c->st));
    } else {
      const int use_hist = (c->count == 0);
      HIPCHECK(rism::run_zinv(P, L.spec, t, L.u, Rout, L.R[c->iR_slot[0]], L.R[c->iR_slot[1]],
                              r0, c->partial, L.tb.twz, L.tb.zpre, L.tb.zpost, use_hist, c->st));
    }
  }
  const size_t nb = rism::zinv_num_blocks<Re>(P, andm ? 1 : 0) / c->natv;
  {
    KTime kt(c, RISM_K_REDUCE);
    HIPCHECK(rism::run_reduce_cols(c->partial, nb, W, c->natv, c->sums, P.done, c->st));
    if (andm) {
      int rb[RISM_MAX_HIST];
      const int nh = (int)c->hist.size();
      for (int p = 0; p < nh; ++p) rb[p] = c->hist[p].rb;
      HIPCHECK(rism::run_andm_finish(c->sums, c->rho_dev, c->natv, c->N, tol, W, c->dD,
                                     c->opts.anderson_m + 1, c->cur_rb, rb, nh, c->count == 0,
                                     c->xcoef, c->scal, c->async ? c->rms_hist : nullptr,
                                     c->async ? c->dflag : nullptr, c->st));
    } else {
      HIPCHECK(rism::run_step_finish(c->sums, c->rho_dev, c->natv, c->N, tol,
                                     c->count == 0, c->opts.an2_mp, c->binary, c->scal,
                                     c->async ? c->rms_hist : nullptr,
                                     c->async ? c->dflag : nullptr, c->st));
    }
  }
  c->iT_last = c->iT_cur;
  L.Rlast = Rout;
  c->last_prec = c->prec;
  c->fresh = true;
  c->have_step = true;
  if (c->async) return RISM_OK;
  double ss = 0.0;
  HIPCHECK(hipMemcpyAsync(c->hscal, c->scal, sizeof(double) * 5, hipMemcpyDeviceToHost, c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  ss = c->hscal[0];
  if (!andm && c->count == 0) {
    c->lastx[0] = c->hscal[1];
    c->lastx[1] = c->hscal[2];
    c->lasts[0] = c->hscal[3];
    c->lasts[1] = c->hscal[4];
  }
  if (c->ktime) ktime_resolve(c);
  if (rms) *rms = ss;
  return RISM_OK;
}

// Moves the loop state (t, the AN2 or Anderson(m) slots, the last residual)
// from one level to the other, keeping every slot index, so the iteration
// continues where it stands.
template <class From, class To>
int convert_state(rism_ctx *c) {
  Level<From> &A = lv<From>(c);
  Level<To> &B = lv<To>(c);
  const size_t nn = c->N * c->natv;
  auto cv = [&](From *src, To *dst) -> int {
    if (src && dst) HIPCHECK(rism::run_convert(src, dst, nn, c->st));
    return RISM_OK;
  };
  int rc;
  for (int i = 0; i < 2; ++i)
    if ((rc = cv(A.T[i], B.T[i]))) return rc;
  for (int i = 0; i < 2; ++i)
    if ((rc = cv(A.Pslot[i], B.Pslot[i]))) return rc;
  for (int i = 0; i < 3; ++i)
    if ((rc = cv(A.R[i], B.R[i]))) return rc;
  for (size_t i = 0; i < A.Rb.size(); ++i)
    if ((rc = cv(A.Rb[i], B.Rb[i]))) return rc;
  for (size_t i = 0; i < A.Pb.size(); ++i)
    if ((rc = cv(A.Pb[i], B.Pb[i]))) return rc;
  B.Rlast = nullptr;
  for (int i = 0; i < 3; ++i)
    if (A.Rlast && A.Rlast == A.R[i]) B.Rlast = B.R[i];
  for (size_t i = 0; i < A.Rb.size(); ++i)
    if (A.Rlast && A.Rlast == A.Rb[i]) B.Rlast = B.Rb[i];
  HIPCHECK(hipStreamSynchronize(c->st));
  c->last_prec = c->prec = (sizeof(To) == sizeof(double)) ? 0 : 1;
  return RISM_OK;
}

template <class Re>
int accel_step(rism_ctx *c) {
  c->fresh = false;
  return c->opts.anderson_m > 0 ? anderson_m_step<Re>(c) : an2_step<Re>(c);
}

// The accelerator's fields at level Re, slot for slot alike at both levels:
// the mixed iterates (P = t + m R), which move with the base point of a
// correction block, the history residuals, which do not, and the residual
// of the current point.
template <class Re>
struct AccelFields {
  std::vector<Re *> iter, resid;
  Re *cur = nullptr;
};
template <class Re>
AccelFields<Re> accel_fields(rism_ctx *c) {
  Level<Re> &L = lv<Re>(c);
  AccelFields<Re> a;
  if (c->opts.anderson_m > 0) {
    for (const auto &h : c->hist) {
      a.iter.push_back(L.Pb[h.pb]);
      a.resid.push_back(L.Rb[h.rb]);
    }
    a.cur = L.Rb[c->cur_rb];
  } else {
    for (int i = 0; i < 2; ++i) {
      a.iter.push_back(L.Pslot[i]);
      a.resid.push_back(L.R[c->iR_slot[i]]);
    }
    a.cur = L.R[c->iR_cur];
  }
  return a;
}

// FP64 state -> FP32 frame of a correction block based at t0 = d.T[k0], the
// point the last FP64 step evaluated: the closure exponent and residual r0
// at t0, d = 0, iterates relative to t0, residuals as they are.
int rebase_down(rism_ctx *c, int k0) {
  const size_t nn = c->N * c->natv;
  const double *t0 = c->d.T[k0];
  const AccelFields<double> D = accel_fields<double>(c);
  const AccelFields<float> F = accel_fields<float>(c);
  HIPCHECK(rism::run_exponent(t0, c->d.u, nn, c->f.u, c->st));
  HIPCHECK(rism::run_convert((const double *)D.cur, F.cur, nn, c->st));
  HIPCHECK(hipMemcpyAsync(c->f.aux, F.cur, sizeof(float) * nn, hipMemcpyDeviceToDevice, c->st));
  for (size_t i = 0; i < D.iter.size(); ++i)
    HIPCHECK(rism::run_rebase_down(D.iter[i], t0, F.iter[i], nn, c->st));
  for (size_t i = 0; i < D.resid.size(); ++i)
    HIPCHECK(rism::run_convert((const double *)D.resid[i], F.resid[i], nn, c->st));
  HIPCHECK(hipMemsetAsync(c->f.T[k0], 0, sizeof(float) * nn, c->st));
  c->prec = 1;
  return RISM_OK;
}

// FP32 block frame -> FP64 state at t0 + d; t0 is overwritten last when the
// current t shares its buffer.
int rebase_up(rism_ctx *c, int k0) {
  const size_t nn = c->N * c->natv;
  const double *t0 = c->d.T[k0];
  const AccelFields<double> D = accel_fields<double>(c);
  const AccelFields<float> F = accel_fields<float>(c);
  for (size_t i = 0; i < D.iter.size(); ++i)
    HIPCHECK(rism::run_rebase_up(D.iter[i], t0, F.iter[i], nn, c->st));
  for (size_t i = 0; i < D.resid.size(); ++i)
    HIPCHECK(rism::run_convert((const float *)F.resid[i], D.resid[i], nn, c->st));
  HIPCHECK(rism::run_rebase_up(c->d.T[c->iT_cur], t0, c->f.T[c->iT_cur], nn, c->st));
  c->prec = c->last_prec = 0;
  return RISM_OK;
}

// The mixed-precision solve (no counterpart in the reference).
//  1. FP32 steps of the fixed-point map with the accelerator, while the FP32
//     residual tracks the FP64 one: until rms <= hand_rms, or no progress
//     for kStall steps.
//  2. One accelerated iteration in which every (block + 1)-th residual is
//     evaluated in FP64 and the others in FP32 in difference form about the
//     last FP64 point t0: R(t0 + d) = r0 + [h(t0 + d) - h(t0)] -
//     [g(t0 + d) - g(t0)], with r0 the FP64 residual at t0.  The brackets are
//     formed from d and the base exponent without cancellation, so the FP32
//     rounding scales with |d| (a block's worth of movement), not with |t|;
//     no linearisation is involved.  The accelerator state moves between the
//     two frames at every FP64 step (rebase_down / rebase_up), so its history
//     is never discarded, and the convergence test is the FP64 residual.
// When the FP64 residual stops falling for kNoGain blocks it returns with
// *fallback set and the state at the FP64 level, and the caller continues
// with plain FP64 steps from there.  A non-finite residual in phase 1 (the
// FP32 exponent overflows above 88, FP64 only above 709) sets *restart: the
// caller solves in FP64 from the t phase 1 started from (its FP64 copy,
// which phase 1 leaves untouched).
int solve_mixed(rism_ctx *c, int maxstep, double tol, int &istep, double &rms, bool &conver,
                int &low_steps, bool &fallback, bool &restart) {
  const bool verbose = c->opts.verbose != 0;
  const double hand_rms = std::max(tol, c->opts.mixed_switch > 0 ? c->opts.mixed_switch : 1e-3);
  // 48 FP32 steps per FP64 step: over cfg2, cfg3 and test2 (AN2 and
  // Anderson(10)) 48 was fastest in 4 of 5 cases against 24 and 96
  const int block = c->opts.mixed_block > 0 ? c->opts.mixed_block : 48;
  constexpr int kStall = 25, kNoGain = 3;
  int rc;
  // 1. FP32 iteration
  double best = HUGE_VAL;
  int since_best = 0;
  while (istep < maxstep) {
    ++istep;
    if ((rc = calc_step<float>(c, &rms))) return rc;
    ++low_steps;
    if (!std::isfinite(rms)) {
      if (verbose)
        std::cout << " Step = " << istep << " FP32 overflow: FP64 from the starting Tuv"
                  << std::endl;
      restart = true;
      return RISM_OK;
    }
    if (verbose) std::cout << " Step = " << istep << " Reside = " << rms << " (FP32)" << std::endl;
    if (rms < 0.99 * best) {
      best = rms;
      since_best = 0;
    } else {
      ++since_best;
    }
    if ((rc = accel_step<float>(c))) return rc;
    if (rms <= hand_rms || since_best >= kStall) break;
  }
  // the accelerator state of the last FP32 step, at the FP64 level
  if ((rc = convert_state<float, double>(c))) return rc;
  // 2. blocks of FP32 difference-form steps between FP64 steps
  best = HUGE_VAL;
  int no_gain = 0;
  while (istep < maxstep) {
    ++istep;
    const int k0 = c->iT_cur;
    if ((rc = calc_step<double>(c, &rms))) return rc;
    if (!std::isfinite(rms)) return RISM_OK;  // diverged: the caller's test sees rms
    if (verbose) std::cout << " Step = " << istep << " Reside = " << rms << std::endl;
    if (rms <= tol) {
      conver = true;
      return RISM_OK;
    }
    if (rms < 0.95 * best) {
      best = rms;
      no_gain = 0;
    } else if (++no_gain >= kNoGain) {
      fallback = true;
      return RISM_OK;
    }
    if ((rc = rebase_down(c, k0))) return rc;
    if ((rc = accel_step<float>(c))) return rc;
    for (int inner = 0; inner < block && istep < maxstep; ++inner) {
      ++istep;
      double r = 0.0;
      if ((rc = calc_step<float>(c, &r, true))) return rc;
      ++low_steps;
      if (verbose)
        std::cout << " Step = " << istep << " Reside = " << r << " (FP32 correction)" << std::endl;
      // a non-finite FP32 residual is not fed to the accelerator: the FP64
      // step re-evaluates the same point
      if (!std::isfinite(r)) break;
      if ((rc = accel_step<float>(c))) return rc;
      if (r <= 0.5 * tol) break;
    }
    if ((rc = rebase_up(c, k0))) return rc;
  }
  return RISM_OK;
}

// ---- solves without a host round trip per step (FP64, AN2 and Anderson(m)) ---
// The host enqueues steps in chunks; each step's finish kernel writes
// rms_hist[step] and raises dflag at convergence or divergence, after which
// every enqueued kernel of the context returns at entry.  The host reads the
// flag once per chunk.  The arithmetic is that of rism_solve's step-by-step
// loop, so the results are identical; only the iteration count is found
// afterwards, from the flag.
constexpr int kAsyncChunk = 24;  // steps enqueued between reads of the flag

struct AsyncRun {
  rism_ctx *c = nullptr;
  int maxstep = 0, istep = 0;
  double tol = 0.0;
  bool finished = false;
  int status = RISM_STATUS_MAXSTEP, steps = 0;
  double rms = 0.0;
  struct Log {  // the accelerator's indices when the step was enqueued
    int iT, iR, s0, s1, b, count;
    int cur_rb, next_pb;                  // Anderson(m)
    std::vector<rism_ctx::HEnt> hist;     // Anderson(m)
  };
  std::vector<Log> log;  // per step
};

bool async_capable(const rism_ctx *c) {
  return c->opts.precision == 0 && !c->opts.verbose;
}

int async_begin(AsyncRun &a, rism_ctx *c, int maxstep, double tol) {
  a = AsyncRun{};
  a.c = c;
  a.maxstep = maxstep;
  a.tol = tol;
  a.log.assign(maxstep + 1, AsyncRun::Log{0, 0, 0, 0, 0, 0, 0, 0, {}});
  if (c->rms_cap < (size_t)maxstep + 1) {
    size_t cap = std::max<size_t>(1024, (size_t)maxstep + 1);
    int rc = c->alloc(&c->rms_hist, cap);
    if (rc) return rc;
    c->rms_cap = cap;
  }
  HIPCHECK(hipMemsetAsync(c->dflag, 0, 3 * sizeof(int), c->st));
  c->async = true;
  return RISM_OK;
}

// The buffer indices of the AN2 slots repeat every kGraphSteps mixing steps
// (t ping-pong x the rotation of three residual buffers), and no kernel
// argument depends on the step number, so one period is captured once per
// context (and tolerance) as a hipGraph and replayed.
constexpr int kGraphSteps = 6;

// the host-side index bookkeeping of one AN2 update (an2_step) without
// launching it
void advance_indices(rism_ctx *c) {
  const int b = c->binary;
  const int oldR = c->iR_slot[b];
  if (c->count > 0) --c->count;
  c->iR_slot[b] = c->iR_cur;
  c->iR_cur = oldR;
  c->iT_cur = 1 - c->iT_cur;
  c->binary = (b == 0) ? 1 : 0;
}

template <class Re>
int ensure_graph(rism_ctx *c, double tol) {
  const rism_ctx::GraphKey key{c->iT_cur, c->iR_cur, c->iR_slot[0], c->iR_slot[1], c->binary,
                               c->prec, tol, c->rms_hist};
  if (c->gexec && key == c->gkey) return RISM_OK;
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  // capture from the current state, then restore it (capture runs nothing)
  const int sT = c->iT_cur, sR = c->iR_cur, s0 = c->iR_slot[0], s1 = c->iR_slot[1],
            sb = c->binary, scount = c->count;
  HIPCHECK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  int rc = RISM_OK;
  for (int k = 0; k < kGraphSteps && !rc; ++k) {
    rc = calc_step<Re>(c, nullptr, false, tol);
    if (!rc) rc = an2_step<Re>(c);
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->st, &g);
  c->iT_cur = sT;
  c->iR_cur = sR;
  c->iR_slot[0] = s0;
  c->iR_slot[1] = s1;
  c->binary = sb;
  c->count = scount;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return rism::set_error(RISM_ERR_HIP, "hipStreamEndCapture failed");
  c->graph = g;
  HIPCHECK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  c->gkey = key;
  return RISM_OK;
}

AsyncRun::Log index_log(const rism_ctx *c) {
  AsyncRun::Log L{c->iT_cur, c->iR_cur, c->iR_slot[0], c->iR_slot[1], c->binary, c->count,
                  c->cur_rb, c->next_pb, {}};
  if (c->opts.anderson_m > 0) L.hist = c->hist;
  return L;
}

// enqueues up to `chunk` steps at level Re and the read-back of the flag:
// Picard steps and remainders directly, whole periods of AN2 mixing steps
// as graph launches.  Anderson(m) steps are enqueued directly (their buffer
// rotation repeats only every lcm(m + 1, m, 2) steps; the coefficients come
// from the device-side finish kernel, so nothing is read back).
template <class Re>
int async_enqueue(AsyncRun &a, int chunk) {
  rism_ctx *c = a.c;
  int rc;
  const bool andm = c->opts.anderson_m > 0;
  for (int k = 0; k < chunk && a.istep < a.maxstep;) {
    if (andm) {
      a.log[++a.istep] = index_log(c);
      if ((rc = calc_step<Re>(c, nullptr, false, a.tol))) return rc;
      c->fresh = false;
      if ((rc = anderson_m_step<Re>(c))) return rc;
      ++k;
      continue;
    }
    if (c->count == 0 && k + kGraphSteps <= chunk && a.istep + kGraphSteps <= a.maxstep) {
      if ((rc = ensure_graph<Re>(c, a.tol))) return rc;
      HIPCHECK(hipGraphLaunch(c->gexec, c->st));
      for (int j = 0; j < kGraphSteps; ++j) {
        a.log[++a.istep] = index_log(c);
        advance_indices(c);
      }
      k += kGraphSteps;
      continue;
    }
    a.log[++a.istep] = index_log(c);
    if ((rc = calc_step<Re>(c, nullptr, false, a.tol))) return rc;
    c->fresh = false;
    if ((rc = an2_step<Re>(c))) return rc;
    ++k;
  }
  HIPCHECK(hipMemcpyAsync(c->hflag, c->dflag, 2 * sizeof(int), hipMemcpyDeviceToHost, c->st));
  return RISM_OK;
}

// after the stream has drained: finishes the run if the flag is up or the
// step limit is reached, restoring the state of the deciding step
template <class Re>
int async_check(AsyncRun &a) {
  rism_ctx *c = a.c;
  if (c->ktime) ktime_resolve(c);
  const int flag = c->hflag[0];
  if (!flag && a.istep < a.maxstep) return RISM_OK;
  const int k = flag ? c->hflag[1] : a.maxstep;
  a.finished = true;
  a.steps = k;
  a.status = flag == 1 ? RISM_STATUS_CONVERGED
                       : (flag == 2 ? RISM_STATUS_DIVERGED : RISM_STATUS_MAXSTEP);
  HIPCHECK(hipMemcpyAsync(c->hscal, c->rms_hist + k, sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  a.rms = c->hscal[0];
  c->async = false;
  const AsyncRun::Log &L = a.log[k];
  const bool andm = c->opts.anderson_m > 0;
  c->iT_last = L.iT;
  lv<Re>(c).Rlast = andm ? lv<Re>(c).Rb[L.cur_rb] : lv<Re>(c).R[L.iR];
  c->last_prec = c->prec;
  c->have_step = true;
  c->fresh = false;
  // converged or diverged at k: no update followed (iterate.cu:50-72), so the
  // accelerator stands where step k found it; at the step limit the last
  // update did, as in the step-by-step loop
  if (flag) {
    c->iT_cur = L.iT;
    c->iR_cur = L.iR;
    c->iR_slot[0] = L.s0;
    c->iR_slot[1] = L.s1;
    c->binary = L.b;
    c->count = L.count;
    if (andm) {
      c->cur_rb = L.cur_rb;
      c->next_pb = L.next_pb;
      c->hist = L.hist;
    }
  }
  return RISM_OK;
}

// ---- grouped batch solves -------------------------------------------------
// FP64 AN2 contexts of one grid shape, box, solvent and options are stepped
// together: every kernel of a step is launched once for the group, one
// context per blockIdx.z, with each context's pointers taken from a device
// table (rism::GroupPtrs).  All members start from the same accelerator
// state and follow the same slot rotation, so one table per state serves a
// step of every member, and a 6-step period of grouped launches is captured
// once per group as a hipGraph.  A member that converges or diverges raises
// its own flag and its workgroups return at entry; the group runs until
// every member has stopped.  Per context the arithmetic is that of
// calc_step + an2_step, so the results equal separate solves bit for bit.
constexpr int kMaxGroup = 16;

// the accelerator's slot indices of a group (as in rism_ctx)
struct GroupIdx {
  int iT, iR, s0, s1, b, count;
  bool operator<(const GroupIdx &o) const {
    return std::tie(iT, iR, s0, s1, b, count) < std::tie(o.iT, o.iR, o.s0, o.s1, o.b, o.count);
  }
  bool operator==(const GroupIdx &o) const { return !(*this < o) && !(o < *this); }
};

GroupIdx group_idx(const rism_ctx *c) {
  return GroupIdx{c->iT_cur, c->iR_cur, c->iR_slot[0], c->iR_slot[1], c->binary, c->count};
}

// the state after one AN2 update (an2_step's bookkeeping)
GroupIdx group_next(GroupIdx n) {
  if (n.count > 0) --n.count;
  int *slot = n.b == 0 ? &n.s0 : &n.s1;
  const int oldR = *slot;
  *slot = n.iR;
  n.iR = oldR;
  n.iT = 1 - n.iT;
  n.b = n.b == 0 ? 1 : 0;
  return n;
}

// contexts whose steps can share launches: equal kernel parameters (grid,
// box, solvent, closure, chi source, launch plans), AN2 options and state
bool groupable(const rism_ctx *a, const rism_ctx *b) {
  const rism::Params &p = a->P, &q = b->P;
  return a->nx == b->nx && a->ny == b->ny && a->nz == b->nz && a->natv == b->natv &&
         a->box[0] == b->box[0] && a->box[1] == b->box[1] && a->box[2] == b->box[2] &&
         a->solv_hash == b->solv_hash && p.clos == q.clos && p.shell == q.shell &&
         a->opts.an2_count == b->opts.an2_count && a->opts.an2_count <= 32 &&
         a->opts.an2_m == b->opts.an2_m &&
         a->opts.an2_mp == b->opts.an2_mp && p.var_zf[0] == q.var_zf[0] &&
         p.var_zi[0] == q.var_zi[0] && p.var_y[0] == q.var_y[0] && a->prec == 0 &&
         b->prec == 0 && group_idx(a) == group_idx(b);
}

struct GroupRun {
  std::vector<rism_ctx *> m;
  std::vector<AsyncRun *> runs;
  GroupIdx ix{};
  int istep = 0, maxstep = 0;
  double tol = 0.0;
  bool finished = false;
  // device tables by state (one allocation, a table per state)
  static constexpr int kMaxStates = 64;
  rism::GroupPtrs *dtab = nullptr;
  std::map<GroupIdx, int> tabs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  GroupIdx gix{};
  GroupRun() = default;
  GroupRun(const GroupRun &) = delete;
  GroupRun &operator=(const GroupRun &) = delete;
  ~GroupRun() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (dtab) (void)hipFree(dtab);
  }
  rism_ctx *c0() const { return m[0]; }
};

// the device table of state ix: each member's buffers for one step
int group_table(GroupRun &g, const GroupIdx &ix, rism::GroupPtrs **out) {
  auto it = g.tabs.find(ix);
  if (it == g.tabs.end())
    return rism::set_error(RISM_ERR_STATE, "group step in an unexpected accelerator state");
  *out = g.dtab + (size_t)it->second * g.m.size();
  return RISM_OK;
}

// every state the group's steps pass through (the Picard steps, then the
// period of the slot rotation), their tables built and uploaded at once
int group_tables(GroupRun &g) {
  const size_t k = g.m.size();
  std::vector<rism::GroupPtrs> h;
  GroupIdx ix = g.ix;
  while (g.tabs.find(ix) == g.tabs.end()) {
    const int slot = (int)g.tabs.size();
    if (slot >= GroupRun::kMaxStates)
      return rism::set_error(RISM_ERR_STATE, "too many accelerator states in a group");
    g.tabs[ix] = slot;
    for (size_t i = 0; i < k; ++i) {
      rism_ctx *c = g.m[i];
      Level<double> &L = c->d;
      rism::GroupPtrs e{};
      e.t = L.T[ix.iT];
      e.u = L.u;
      e.fr = L.fr;
      e.spec = L.spec;
      e.fk = L.fk;
      e.R = L.R[ix.iR];
      e.R0 = L.R[ix.s0];
      e.R1 = L.R[ix.s1];
      e.partial = c->partial;
      e.sums = c->sums;
      e.scal = c->scal;
      e.rms_hist = c->rms_hist;
      e.done = c->dflag;
      e.tnew = L.T[1 - ix.iT];
      e.P0 = L.Pslot[0];
      e.P1 = L.Pslot[1];
      e.Pb = L.Pslot[ix.b];
      h.push_back(e);
    }
    ix = group_next(ix);
  }
  if (hipMalloc(&g.dtab, sizeof(rism::GroupPtrs) * h.size()) != hipSuccess)
    return rism::set_error(RISM_ERR_NOMEM, "hipMalloc of the group tables failed");
  hipStream_t st = g.c0()->st;
  HIPCHECK(hipMemcpyAsync(g.dtab, h.data(), sizeof(rism::GroupPtrs) * h.size(),
                          hipMemcpyHostToDevice, st));
  HIPCHECK(hipStreamSynchronize(st));
  return RISM_OK;
}

// one step (calc_step) and one AN2 update (an2_step) of every member at the
// group's state g.ix, which then advances as an2_step advances a context's
int group_step(GroupRun &g) {
  rism_ctx *c = g.c0();
  rism::GroupPtrs *tab = nullptr;
  int rc = group_table(g, g.ix, &tab);
  if (rc) return rc;
  Level<double> &L = c->d;
  const GroupIdx ix = g.ix;
  rism::Params P = c->P;
  P.mode = 0;
  P.done = nullptr;
  P.grp = tab;
  P.ngrp = (int)g.m.size();
  hipStream_t st = c->st;
  // the pointer arguments are the first member's (its tables serve all)
  HIPCHECK(rism::run_zfwd(P, L.T[ix.iT], L.u, L.fr, L.spec, L.tb.twz, L.tb.zpre, L.tb.zpost, st));
  HIPCHECK(rism::run_y(P, L.spec, -1, L.tb.twy, L.tb.prey, st));
  HIPCHECK(rism::run_xoz(P, L.spec, L.fk, L.tb.twx, L.tb.prex, c->knots, c->coef, L.chitab, st));
  HIPCHECK(rism::run_y(P, L.spec, +1, L.tb.twy, L.tb.prey, st));
  HIPCHECK(rism::run_zinv(P, L.spec, L.T[ix.iT], L.u, L.R[ix.iR], L.R[ix.s0], L.R[ix.s1],
                          (const double *)nullptr, c->partial, L.tb.twz, L.tb.zpre, L.tb.zpost,
                          ix.count == 0, st));
  const size_t nb = rism::zinv_num_blocks<double>(P, 0) / c->natv;
  HIPCHECK(rism::run_reduce_cols(c->partial, nb, 6, c->natv, c->sums, nullptr, st, tab, P.ngrp));
  HIPCHECK(rism::run_step_finish(c->sums, c->rho_dev, c->natv, c->N, g.tol, ix.count == 0,
                                 c->opts.an2_mp, ix.b, c->scal, c->rms_hist, c->dflag, st, tab,
                                 P.ngrp));
  const size_t nn = c->N * c->natv;
  if (ix.count > 0) {
    HIPCHECK(rism::run_an2_picard(L.T[1 - ix.iT], L.T[ix.iT], L.R[ix.iR], L.Pslot[ix.b],
                                  c->opts.an2_m, nn, (const int *)nullptr, st, tab, P.ngrp));
  } else {
    HIPCHECK(rism::run_an2_update(L.T[1 - ix.iT], L.T[ix.iT], L.R[ix.iR], L.Pslot[0],
                                  L.Pslot[1], L.Pslot[ix.b], c->scal, c->opts.an2_m, nn,
                                  (const int *)nullptr, st, tab, P.ngrp));
  }
  g.ix = group_next(g.ix);
  return RISM_OK;
}

// the 6-step period from the current state as one graph (capture runs
// nothing; the state is restored afterwards)
int group_graph(GroupRun &g) {
  if (g.gexec && g.gix == g.ix) return RISM_OK;
  if (g.gexec) (void)hipGraphExecDestroy(g.gexec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g.gexec = nullptr;
  g.graph = nullptr;
  const GroupIdx probe = g.ix;
  hipStream_t st = g.c0()->st;
  HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = RISM_OK;
  for (int k = 0; k < kGraphSteps && !rc; ++k) rc = group_step(g);
  hipGraph_t gr = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &gr);
  g.ix = probe;
  if (rc) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return rism::set_error(RISM_ERR_HIP, "hipStreamEndCapture failed");
  g.graph = gr;
  HIPCHECK(hipGraphInstantiate(&g.gexec, g.graph, nullptr, nullptr, 0));
  g.gix = g.ix;
  return RISM_OK;
}

int group_begin(GroupRun &g, int maxstep, double tol) {
  g.maxstep = maxstep;
  g.tol = tol;
  g.istep = 0;
  g.ix = group_idx(g.c0());
  for (size_t i = 0; i < g.m.size(); ++i) {
    int rc = async_begin(*g.runs[i], g.m[i], maxstep, tol);
    if (rc) return rc;
  }
  // the members' set-up and flag resets ran on their own streams
  for (rism_ctx *c : g.m)
    if (hipStreamSynchronize(c->st) != hipSuccess)
      return rism::set_error(RISM_ERR_HIP, "hipStreamSynchronize failed");
  return group_tables(g);  // after async_begin: the tables hold rms_hist
}

// enqueues up to `chunk` steps of the group and the read-back of every
// running member's flag (async_enqueue for a group)
int group_enqueue(GroupRun &g, int chunk) {
  rism_ctx *c = g.c0();
  auto log_step = [&]() {
    ++g.istep;
    for (size_t i = 0; i < g.m.size(); ++i)
      if (!g.runs[i]->finished)
        g.runs[i]->log[g.istep] =
            AsyncRun::Log{g.ix.iT, g.ix.iR, g.ix.s0, g.ix.s1, g.ix.b, g.ix.count, 0, 0, {}};
  };
  int rc;
  for (int k = 0; k < chunk && g.istep < g.maxstep;) {
    if (g.ix.count == 0 && k + kGraphSteps <= chunk && g.istep + kGraphSteps <= g.maxstep) {
      if ((rc = group_graph(g))) return rc;
      HIPCHECK(hipGraphLaunch(g.gexec, c->st));
      for (int j = 0; j < kGraphSteps; ++j) {
        log_step();
        g.ix = group_next(g.ix);
      }
      k += kGraphSteps;
      continue;
    }
    log_step();
    if ((rc = group_step(g))) return rc;
    ++k;
  }
  for (size_t i = 0; i < g.m.size(); ++i)
    if (!g.runs[i]->finished)
      HIPCHECK(hipMemcpyAsync(g.m[i]->hflag, g.m[i]->dflag, 2 * sizeof(int),
                              hipMemcpyDeviceToHost, c->st));
  return RISM_OK;
}

// after the group's stream has drained: each running member takes the
// group's state and step count, then async_check decides it
int group_check(GroupRun &g) {
  bool left = false;
  for (size_t i = 0; i < g.m.size(); ++i) {
    AsyncRun &a = *g.runs[i];
    if (a.finished) continue;
    rism_ctx *c = g.m[i];
    c->iT_cur = g.ix.iT;
    c->iR_cur = g.ix.iR;
    c->iR_slot[0] = g.ix.s0;
    c->iR_slot[1] = g.ix.s1;
    c->binary = g.ix.b;
    c->count = g.ix.count;
    a.istep = g.istep;
    int rc = async_check<double>(a);
    if (rc) return rc;
    left = left || !a.finished;
  }
  g.finished = !left;
  return RISM_OK;
}

int setup_enqueue(rism_ctx *c) {
  const size_t nn = c->N * c->natv;
  c->prec = 0;
  reset_an2(c);
  // history slots start at zero (the reference leaves them uninitialised)
  if (c->opts.anderson_m == 0) {
    for (int i = 0; i < 2; ++i) {
      HIPCHECK(hipMemsetAsync(c->d.Pslot[i], 0, sizeof(double) * nn, c->st));
      HIPCHECK(hipMemsetAsync(c->d.R[c->iR_slot[i]], 0, sizeof(double) * nn, c->st));
      if (c->opts.precision == 1) {
        HIPCHECK(hipMemsetAsync(c->f.Pslot[i], 0, sizeof(float) * nn, c->st));
        HIPCHECK(hipMemsetAsync(c->f.R[c->iR_slot[i]], 0, sizeof(float) * nn, c->st));
      }
    }
  }
  if (c->opts.verbose) {
    std::cout << "tabulating solute Lennard-Jones potential ..." << std::endl;
    if (c->opts.ad_switch != 1) std::cout << "synthesizing solute Coulomb potential ..." << std::endl;
  }
  const double ikbt = 1.0 / (kAvogadro * kBoltzmann * c->temper);  // cal_LJ.cu:47
  const double lam2 = (c->opts.ad_switch == 2) ? c->opts.lambda : 1.0;
  const double ubeta = kHartree * kBohr / (kBoltzmann * c->temper) * lam2;  // cal_Coulomb.cu:33
  const int do_coul = c->opts.ad_switch != 1;  // cal_Coulomb.cu:23
  {
    KTime kt(c, RISM_K_POT);
    HIPCHECK(rism::run_potential_r(c->atoms, c->natu, c->P, ikbt, ubeta, do_coul, c->d.u,
                                   c->d.fr, c->st));
  }
  if (do_coul && c->natu > 0) {
    KTime kt(c, RISM_K_FK);
    HIPCHECK(rism::run_fk_half(c->atoms, c->natu, rism::atom_stride(c->natv), c->P, ubeta,
                               c->d.fk, c->st));
  } else {
    HIPCHECK(hipMemsetAsync(c->d.fk, 0, sizeof(cplx) * (c->N / 2), c->st));
  }
  if (c->opts.verbose) std::cout << "synthesizing initial estimate for Tuv ..." << std::endl;
  HIPCHECK(rism::run_init_t(c->d.T[c->iT_cur], c->d.fr, c->P, c->N, c->st));
  if (c->opts.precision == 1) {  // the FP32 level starts from the FP64 set-up
    HIPCHECK(rism::run_convert(c->d.u, c->f.u, nn, c->st));
    HIPCHECK(rism::run_convert(c->d.fr, c->f.fr, c->N, c->st));
    HIPCHECK(rism::run_convert((const double *)c->d.fk, (float *)c->f.fk, c->N, c->st));
    HIPCHECK(rism::run_convert(c->d.T[c->iT_cur], c->f.T[c->iT_cur], nn, c->st));
    c->prec = c->last_prec = 1;
  }
  return RISM_OK;
}

// ---- autotuner ----------------------------------------------------------------
// The forward z, residual z and y passes are built in up to three launch
// plans each (tile widths, radix-4 or radix-8 stages, waves per SIMD; see
// rism_kzpass.h and rism_kypass.hip).  Which is fastest depends on the grid
// (the FFT length, the row width nx, the site count) and the bulk type, so
// the plan is picked per grid shape by timing every plan the build offers on
// the context's own buffers (one warm-up launch, then the mean of three),
// once per process and shape (the choice is cached for the contexts that
// follow).  RISM_TUNE=0 keeps the default plans; RISM_TUNE_FILE=<path>
// reads earlier choices from that file and appends new ones, so a later
// process reuses them without timing.  Plans differ in their radix orders,
// so results can differ in the last bits between plans, never beyond FP64
// (FP32) rounding.
namespace tune {
using Key = std::tuple<int, int, int, int, int, int>;  // pass, nx, ny, nz, natv, fp32
struct Rec {
  int choice;
  float ms[rism::kPlans];  // < 0: plan not offered
};
std::mutex mu;
std::map<Key, Rec> table;
bool loaded = false;

bool enabled() {
  const char *e = std::getenv("RISM_TUNE");
  return !(e && e[0] == '0');
}

// RISM_TUNE_PLANS=zf,zi,y forces these plans (where the build offers them,
// else the default) instead of timing: reproducible plans across processes
// and a way to check each plan.  Returns -1 when unset.
int forced(int pass) {
  const char *e = std::getenv("RISM_TUNE_PLANS");
  if (!e) return -1;
  int v[3] = {0, 0, 0};
  if (std::sscanf(e, "%d,%d,%d", &v[0], &v[1], &v[2]) != 3) return -1;
  return v[pass];
}

// RISM_TUNE_FILE lines: pass nx ny nz natv fp32 choice ms0 ms1 ms2
void load_locked() {
  if (loaded) return;
  loaded = true;
  const char *path = std::getenv("RISM_TUNE_FILE");
  if (!path) return;
  std::ifstream in(path);
  int a[7];
  float m[rism::kPlans];
  while (in >> a[0] >> a[1] >> a[2] >> a[3] >> a[4] >> a[5] >> a[6] >> m[0] >> m[1] >> m[2]) {
    if (a[6] < 0 || a[6] >= rism::kPlans) continue;
    Rec r{a[6], {m[0], m[1], m[2]}};
    table[Key{a[0], a[1], a[2], a[3], a[4], a[5]}] = r;
  }
}

void save_locked(const Key &k, const Rec &r) {
  const char *path = std::getenv("RISM_TUNE_FILE");
  if (!path) return;
  std::ofstream out(path, std::ios::app);
  out << std::get<0>(k) << " " << std::get<1>(k) << " " << std::get<2>(k) << " "
      << std::get<3>(k) << " " << std::get<4>(k) << " " << std::get<5>(k) << " " << r.choice;
  for (float m : r.ms) out << " " << m;
  out << "\n";
}
}  // namespace tune

// one launch of `pass` at level Re in the plan set in c->P
template <class Re>
int tune_launch(rism_ctx *c, int pass) {
  Level<Re> &L = lv<Re>(c);
  rism::Params P = c->P;
  P.mode = 0;
  P.done = nullptr;
  switch (pass) {
    case rism::TP_ZFWD:
      HIPCHECK(rism::run_zfwd(P, L.T[0], L.u, L.fr, L.spec, L.tb.twz, L.tb.zpre, L.tb.zpost,
                              c->st));
      break;
    case rism::TP_ZINV:
      HIPCHECK(rism::run_zinv(P, L.spec, L.T[0], L.u, L.R[0], L.R[1], L.R[2], (const Re *)nullptr,
                              c->partial, L.tb.twz, L.tb.zpre, L.tb.zpost, 1, c->st));
      break;
    default:
      HIPCHECK(rism::run_y(P, L.spec, -1, L.tb.twy, L.tb.prey, c->st));
      break;
  }
  return RISM_OK;
}

// Picks the plan of each tuned pass for level Re (the cache first, else by
// timing) and records it in c->P.  The level's fields hold no data yet:
// they are set first so every plan runs on the same values (zero, u = 0.5).
template <class Re>
int tune_level(rism_ctx *c) {
  const int f = sizeof(Re) == 4 ? 1 : 0;
  Level<Re> &L = lv<Re>(c);
  const size_t nn = c->N * c->natv;
  bool cleared = false;
  for (int pass = 0; pass < rism::TP_COUNT; ++pass) {
    int *slot = pass == rism::TP_ZFWD ? &c->P.var_zf[f]
                                      : (pass == rism::TP_ZINV ? &c->P.var_zi[f] : &c->P.var_y[f]);
    *slot = 0;
    // the residual pass of Anderson(m) contexts keeps its single plan
    if (pass == rism::TP_ZINV && c->opts.anderson_m > 0) continue;
    const int len = pass == rism::TP_Y ? c->ny : c->nz / 2;
    const int fv = tune::forced(pass);
    if (fv >= 0) {
      *slot = (fv < rism::kPlans && rism::plan_offered(pass, len, f, fv)) ? fv : 0;
      continue;
    }
    int offered = 0;
    for (int v = 0; v < rism::kPlans; ++v) offered += rism::plan_offered(pass, len, f, v) ? 1 : 0;
    if (offered < 2) continue;
    const tune::Key key{pass, c->nx, c->ny, c->nz, c->natv, f};
    {
      std::lock_guard<std::mutex> g(tune::mu);
      tune::load_locked();
      auto it = tune::table.find(key);
      if (it != tune::table.end()) {
        *slot = it->second.choice;
        continue;
      }
    }
    if (!cleared) {
      HIPCHECK(hipMemsetAsync(L.T[0], 0, sizeof(Re) * nn, c->st));
      // u = 0.5 (the 32-bit word of 0.5f repeated: 0.5 + 2^-30 in FP64), so
      // that t - u < 0 and the closure takes its exponential branch, as it
      // does over much of a real grid: with u = 0 the KH and PSE closures
      // take the linear branch, the z passes skip their exponentials, and a
      // plan that cannot hide them (the 512^3 forward pass: 3.37 ms on zeros,
      // 3.65 ms on a solve's fields) is timed too fast
      HIPCHECK(hipMemsetD32Async((hipDeviceptr_t)L.u, 0x3F000000u, sizeof(Re) * nn / 4, c->st));
      HIPCHECK(hipMemsetAsync(L.fr, 0, sizeof(Re) * c->N, c->st));
      HIPCHECK(hipMemsetAsync(L.spec, 0, sizeof(cx<Re>) * nn / 2, c->st));
      for (int i = 0; i < 3; ++i)
        if (L.R[i]) HIPCHECK(hipMemsetAsync(L.R[i], 0, sizeof(Re) * nn, c->st));
      cleared = true;
    }
    tune::Rec rec{0, {-1.f, -1.f, -1.f}};
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    int rc = RISM_OK;
    for (int v = 0; v < rism::kPlans && !rc; ++v) {
      if (!rism::plan_offered(pass, len, f, v)) continue;
      *slot = v;
      rc = tune_launch<Re>(c, pass);  // warm-up (first launch of the plan)
      if (!rc && hipEventRecord(e0, c->st) != hipSuccess) rc = RISM_ERR_HIP;
      for (int r = 0; r < 3 && !rc; ++r) rc = tune_launch<Re>(c, pass);
      if (!rc && (hipEventRecord(e1, c->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        rc = RISM_ERR_HIP;
      float ms = 0.f;
      if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
      rec.ms[v] = ms / 3.f;
      if (rec.ms[v] < rec.ms[rec.choice] || rec.ms[rec.choice] < 0.f) rec.choice = v;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc == RISM_ERR_HIP ? rism::set_error(rc, "autotuner launch failed") : rc;
    *slot = rec.choice;
    std::lock_guard<std::mutex> g(tune::mu);
    if (tune::table.emplace(key, rec).second) tune::save_locked(key, rec);
  }
  // u back to zero after the timing (the other fields stay cleared)
  if (cleared) HIPCHECK(hipMemsetAsync(L.u, 0, sizeof(Re) * nn, c->st));
  return RISM_OK;
}

// Device buffers of one level: u, fr, fk, t (ping-pong) and the AN2 or
// Anderson(m) slots; the FP32 level's spectrum lives in the FP64 one's memory
// (the two levels never run at the same time).
template <class Re>
int alloc_level(rism_ctx *c, Level<Re> &L, cplx *spec_alias) {
  const size_t nn = c->N * c->natv;
  int rc;
  if ((rc = c->alloc(&L.u, nn))) return rc;
  if ((rc = c->alloc(&L.fr, c->N))) return rc;
  if ((rc = c->alloc(&L.fk, c->N / 2))) return rc;
  if (spec_alias) {  // the FP32 level of a mixed context
    L.spec = reinterpret_cast<cx<Re> *>(spec_alias);
    if ((rc = c->alloc(&L.aux, nn))) return rc;
  } else if ((rc = c->alloc(&L.spec, nn / 2))) {
    return rc;
  }
  for (int i = 0; i < 2; ++i)
    if ((rc = c->alloc(&L.T[i], nn))) return rc;
  const int am = c->opts.anderson_m;
  if (am == 0) {
    for (int i = 0; i < 2; ++i)
      if ((rc = c->alloc(&L.Pslot[i], nn))) return rc;
    for (int i = 0; i < 3; ++i)
      if ((rc = c->alloc(&L.R[i], nn))) return rc;
  } else {
    L.Rb.assign(am + 1, nullptr);
    L.Pb.assign(am, nullptr);
    for (auto &b : L.Rb)
      if ((rc = c->alloc(&b, nn))) return rc;
    for (auto &b : L.Pb)
      if ((rc = c->alloc(&b, nn))) return rc;
  }
  return RISM_OK;
}

}  // namespace

extern "C" {

const char *rism_last_error(void) { return g_err.c_str(); }

int rism_set_kernel_timing(rism_ctx *c, int on) {
  if (!c) return rism::set_error(RISM_ERR_ARG, "null context");
  HIPCHECK(hipStreamSynchronize(c->st));
  c->pending.clear();
  c->evnext = 0;
  c->ktime = on != 0;
  for (auto &k : c->kstat) k = rism_kstat{0, 0.0};
  return RISM_OK;
}

int rism_get_kernel_stats(rism_ctx *c, rism_kstat *out) {
  if (!c || !out) return rism::set_error(RISM_ERR_ARG, "null argument");
  HIPCHECK(hipStreamSynchronize(c->st));
  ktime_resolve(c);
  for (int i = 0; i < RISM_K_COUNT; ++i) out[i] = c->kstat[i];
  return RISM_OK;
}

rism_ctx *rism_create(const rism_grid *grid, const rism_solvent *solvent,
                      const rism_solute *solute, const rism_opts *opts, int *err) {
  int dummy;
  if (!err) err = &dummy;
  *err = RISM_OK;
  if (!grid || !solvent || !solute || !opts) {
    *err = rism::set_error(RISM_ERR_ARG, "null argument");
    return nullptr;
  }
  int rc = check_sizes(grid->nx, grid->ny, grid->nz, solvent->natv);
  if (rc) {
    *err = rc;
    return nullptr;
  }
  if (opts->closure != RISM_CLOSURE_KH && opts->closure != RISM_CLOSURE_HNC &&
      !(opts->closure >= RISM_CLOSURE_PSE(1) && opts->closure <= RISM_CLOSURE_PSE(20))) {
    *err = rism::set_error(RISM_ERR_CLOSURE, "Error: Unexpected closure switch ");
    return nullptr;
  }
  if (opts->anderson_m < 0 || opts->anderson_m > RISM_MAX_HIST) {
    *err = rism::set_error(RISM_ERR_ARG, "anderson_m must be in [0, 16]");
    return nullptr;
  }
  if (opts->mixed_block < 0) {
    *err = rism::set_error(RISM_ERR_ARG, "mixed_block must be >= 0");
    return nullptr;
  }
  rism_ctx *c = new rism_ctx();
  auto fail = [&](int code) {
    *err = code;
    delete c;
    return (rism_ctx *)nullptr;
  };
  if (hipGetDevice(&c->dev) != hipSuccess) {
    (void)hipGetLastError();
    return fail(rism::set_error(RISM_ERR_HIP, "no HIP device"));
  }
  if (hipStreamCreate(&c->st) != hipSuccess) return fail(rism::set_error(RISM_ERR_HIP, "hipStreamCreate failed"));
  c->nx = grid->nx;
  c->ny = grid->ny;
  c->nz = grid->nz;
  c->natv = solvent->natv;
  c->N = (size_t)c->nx * c->ny * c->nz;
  for (int d = 0; d < 3; ++d) c->box[d] = grid->box[d];
  c->opts = *opts;
  c->temper = solvent->temper;
  c->xt = solvent->xt;
  c->rho.assign(solvent->rho, solvent->rho + c->natv);
  c->qv.assign(solvent->q, solvent->q + c->natv);
  c->sigv.assign(solvent->sig, solvent->sig + c->natv);
  c->epsv.assign(solvent->eps, solvent->eps + c->natv);
  {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t bytes) {
      const unsigned char *q = static_cast<const unsigned char *>(p);
      for (size_t i = 0; i < bytes; ++i) h = (h ^ q[i]) * 1099511628211ull;
    };
    const size_t nv = (size_t)solvent->natv;
    mix(&solvent->natv, sizeof(int));
    mix(&solvent->ntab, sizeof(int));
    mix(&solvent->dr, sizeof(double));
    mix(&solvent->temper, sizeof(double));
    mix(&solvent->xt, sizeof(double));
    for (const double *a : {solvent->sig, solvent->eps, solvent->q, solvent->rho})
      mix(a, sizeof(double) * nv);
    mix(solvent->xvv, sizeof(double) * nv * nv * (size_t)solvent->ntab);
    c->solv_hash = h;
  }
  c->natu = solute->natu;
  c->usig.assign(solute->sig, solute->sig + c->natu);
  c->ueps.assign(solute->eps, solute->eps + c->natu);
  c->uq.assign(solute->q, solute->q + c->natu);
  c->uxyz.assign(solute->xyz, solute->xyz + 3 * (size_t)c->natu);
  c->P.clos = opts->closure;
  fill_params(c);

  // solute table: cal_LJ.cu:26-40 mixing rules, -l scales sigma
  const double lam1 = (opts->ad_switch == 1) ? opts->lambda : 1.0;
  const int stride = rism::atom_stride(c->natv);
  c->atoms_host.assign((size_t)(c->natu > 0 ? c->natu : 1) * stride, 0.0);
  for (int a = 0; a < c->natu; ++a) {
    double *p = c->atoms_host.data() + (size_t)a * stride;
    p[0] = solute->xyz[3 * a];
    p[1] = solute->xyz[3 * a + 1];
    p[2] = solute->xyz[3 * a + 2];
    p[3] = solute->q[a];
    double rc2 = 0.0;
    for (int v = 0; v < c->natv; ++v) {  // cal_LJ.cu:33-44: sigma_uv, 4 eps_uv
      const double su = (solute->sig[a] + c->sigv[v]) * 0.5 * lam1;
      const double s2 = su * su, s6 = s2 * s2 * s2;
      const double e4 = 4.0 * std::sqrt(solute->eps[a] * c->epsv[v]);
      p[5 + v] = s2;
      p[5 + c->natv + v] = e4 * s6 * s6;
      p[5 + 2 * c->natv + v] = e4 * s6;
      rc2 = std::max(rc2, 4.0e-6 * s2);
    }
    p[4] = rc2;
  }
  if ((rc = upload(c, &c->atoms, c->atoms_host))) return fail(rc);
  if ((rc = make_tables(c, c->d.tb))) return fail(rc);
  if ((rc = make_chi(c, solvent))) return fail(rc);

  const size_t nn = c->N * c->natv;
  if ((rc = alloc_level(c, c->d, nullptr))) return fail(rc);
  c->d.chitab = c->chitab;
  if (c->opts.precision == 1) {
    if ((rc = make_tables(c, c->f.tb))) return fail(rc);
    if ((rc = alloc_level(c, c->f, c->d.spec))) return fail(rc);
    if (c->chitab) {
      if ((rc = c->alloc(&c->f.chitab, c->nchitab))) return fail(rc);
      if (rism::run_convert((const double *)c->chitab, c->f.chitab, c->nchitab, c->st) !=
              hipSuccess ||
          hipStreamSynchronize(c->st) != hipSuccess)
        return fail(rism::set_error(RISM_ERR_HIP, "chi table conversion failed"));
    }
  }
  if (c->opts.anderson_m > 0) {
    const size_t m1 = (size_t)c->opts.anderson_m + 1;
    if ((rc = c->alloc(&c->dD, m1 * m1))) return fail(rc);
    if ((rc = c->alloc(&c->xcoef, RISM_MAX_HIST))) return fail(rc);
    if (hipMemsetAsync(c->dD, 0, sizeof(double) * m1 * m1, c->st) != hipSuccess ||
        hipMemsetAsync(c->xcoef, 0, sizeof(double) * RISM_MAX_HIST, c->st) != hipSuccess)
      return fail(rism::set_error(RISM_ERR_HIP, "clear of the Anderson(m) state failed"));
  }
  // the partial-sum buffer holds the most workgroups any plan launches
  c->zinv_blocks = std::max(std::max(rism::zinv_num_blocks<double>(c->P, -1),
                                     rism::zinv_num_blocks<double>(c->P, 1)),
                            std::max(rism::zinv_num_blocks<float>(c->P, -1),
                                     rism::zinv_num_blocks<float>(c->P, 1)));
  if ((rc = c->alloc(&c->partial, c->zinv_blocks * kSumWidth))) return fail(rc);
  const int tb = rism::thermo_blocks(c->N);
  if ((rc = c->alloc(&c->tpartial, (size_t)tb * 4 * c->natv))) return fail(rc);
  if ((rc = c->alloc(&c->sums, kSumWidth * RISM_MAX_SITES))) return fail(rc);
  if ((rc = c->alloc(&c->scratch, c->N))) return fail(rc);
  if ((rc = c->alloc(&c->scal, 8))) return fail(rc);
  if ((rc = c->alloc(&c->dflag, 3))) return fail(rc);
  if ((rc = upload(c, &c->rho_dev, c->rho))) return fail(rc);
  if (hipHostMalloc((void **)&c->hscal, 8 * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void **)&c->hflag, 3 * sizeof(int)) != hipSuccess)
    return fail(rism::set_error(RISM_ERR_NOMEM, "hipHostMalloc failed"));
  if (tune::enabled() || tune::forced(0) >= 0) {
    if ((rc = tune_level<double>(c))) return fail(rc);
    if (c->opts.precision == 1 && (rc = tune_level<float>(c))) return fail(rc);
  }
  reset_an2(c);
  if (hipMemsetAsync(c->d.T[0], 0, sizeof(double) * nn, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(rism::set_error(RISM_ERR_HIP, "initial clear failed"));
  return c;
}

void rism_destroy(rism_ctx *ctx) { delete ctx; }

int rism_setup_potentials(rism_ctx *c) {
  if (!c) return rism::set_error(RISM_ERR_ARG, "null context");
  int rc = setup_enqueue(c);
  if (rc) return rc;
  HIPCHECK(hipStreamSynchronize(c->st));
  return RISM_OK;
}

int rism_calculate(rism_ctx *c, double *rms) {
  if (!c) return rism::set_error(RISM_ERR_ARG, "null context");
  return c->prec ? calc_step<float>(c, rms) : calc_step<double>(c, rms);
}

int rism_anderson(rism_ctx *c) {
  if (!c) return rism::set_error(RISM_ERR_ARG, "null context");
  if (!c->have_step || !c->fresh || c->last_prec != c->prec)
    return rism::set_error(RISM_ERR_STATE, "rism_anderson needs a preceding rism_calculate");
  c->fresh = false;
  if (c->opts.anderson_m > 0)
    return c->prec ? anderson_m_step<float>(c) : anderson_m_step<double>(c);
  return c->prec ? an2_step<float>(c) : an2_step<double>(c);
}

int rism_set_precision(rism_ctx *c, int prec) {
  if (!c) return rism::set_error(RISM_ERR_ARG, "null context");
  if (prec != 0 && prec != 1) return rism::set_error(RISM_ERR_ARG, "precision must be 0 or 1");
  if (prec == 1 && c->opts.precision != 1)
    return rism::set_error(RISM_ERR_STATE, "the context was created without the FP32 level");
  if (prec == c->prec) return RISM_OK;
  return prec ? convert_state<double, float>(c) : convert_state<float, double>(c);
}

int rism_get_precision(rism_ctx *c) { return c ? c->prec : -1; }

// RISM_SYNC_FREE=0 keeps rism_solve in the step-by-step loop (the rms read
// back after every step, as iterate.cu:48-72 does): the loop structure of
// the reference, for comparisons; any other value or none allows the
// sync-free path.
static bool sync_free_allowed() {
  const char *e = std::getenv("RISM_SYNC_FREE");
  return !(e && e[0] == '0');
}

int rism_solve(rism_ctx *c, int maxstep, double tol, int *status, int *steps, double *rms_out) {
  if (!c) return rism::set_error(RISM_ERR_ARG, "null context");
  if (async_capable(c) && !c->ktime && c->prec == 0 && maxstep > 0 && sync_free_allowed()) {
    // same steps without a host round trip per step
    AsyncRun a;
    int rc = async_begin(a, c, maxstep, tol);
    while (!rc && !a.finished) {
      rc = async_enqueue<double>(a, kAsyncChunk);
      if (!rc && hipStreamSynchronize(c->st) != hipSuccess)
        rc = rism::set_error(RISM_ERR_HIP, "hipStreamSynchronize failed");
      if (!rc) rc = async_check<double>(a);
    }
    c->async = false;
    if (rc) return rc;
    if (status) *status = a.status;
    if (steps) *steps = a.steps;
    if (rms_out) *rms_out = a.rms;
    c->timing.steps = a.steps;
    c->timing.status = a.status;
    c->timing.rms = a.rms;
    c->timing.steps_fp32 = 0;
    return RISM_OK;
  }
  if (c->opts.verbose) std::cout << "relaxing 3D UV RISM:" << std::endl;
  bool conver = false, diverge = false;
  int istep = 0, last = 0, low_steps = 0;
  double rms = 0.0;
  if (c->opts.precision == 1 && c->prec == 1) {
    bool fallback = false, restart = false;
    const int iT0 = c->iT_cur;  // the FP64 copy of the t the FP32 phase starts from
    int rc = solve_mixed(c, maxstep, tol, istep, rms, conver, low_steps, fallback, restart);
    if (rc) return rc;
    last = istep;
    diverge = !std::isfinite(rms) && !restart;
    if (restart) {
      // FP64 from the t the FP32 phase started from (the FP32 phase leaves
      // the FP64 level untouched) with a fresh accelerator; the FP32 steps
      // spent count against maxstep
      c->prec = 0;
      reset_an2(c);
      if (iT0 != c->iT_cur)
        HIPCHECK(hipMemcpyAsync(c->d.T[c->iT_cur], c->d.T[iT0], sizeof(double) * c->N * c->natv,
                                hipMemcpyDeviceToDevice, c->st));
    } else if (fallback) {  // plain FP64 steps from the last FP64 evaluation on
      if ((rc = accel_step<double>(c))) return rc;
    }
  }
  for (istep = istep + 1; !conver && !diverge && istep <= maxstep; ++istep) {  // iterate.cu:48-72
    int rc = rism_calculate(c, &rms);
    if (rc) return rc;
    last = istep;
    diverge = !std::isfinite(rms);
    if (diverge) break;
    if (rms <= tol) {
      conver = true;
    } else {
      rc = rism_anderson(c);
      if (rc) return rc;
    }
    if (c->opts.verbose) std::cout << " Step = " << istep << " Reside = " << rms << std::endl;
    if (conver) break;
  }
  if (c->opts.verbose) {
    if (diverge)
      std::cout << "Calculation diverged." << std::endl;
    else if (!conver)
      std::cout << "3D UV RISM: reached limit # of relaxation steps: " << maxstep << std::endl;
  }
  const int st = conver ? RISM_STATUS_CONVERGED : (diverge ? RISM_STATUS_DIVERGED : RISM_STATUS_MAXSTEP);
  if (status) *status = st;
  if (steps) *steps = last;
  if (rms_out) *rms_out = rms;
  c->timing.steps = last;
  c->timing.status = st;
  c->timing.rms = rms;
  c->timing.steps_fp32 = low_steps;
  return RISM_OK;
}

int rism_thermo(rism_ctx *c, rism_thermo_out *out) {
  if (!c || !out) return rism::set_error(RISM_ERR_ARG, "null argument");
  if (!c->have_step) return rism::set_error(RISM_ERR_STATE, "rism_thermo needs a solved state");
  const int tb = rism::thermo_blocks(c->N);
  {
    KTime kt(c, RISM_K_THERMO);
    if (c->last_prec == 1)
      HIPCHECK(rism::run_thermo_partial(c->f.T[c->iT_cur], c->f.T[c->iT_last], c->f.u,
                                        (const float *)c->f.Rlast, c->P, c->N, c->tpartial,
                                        c->st));
    else
      HIPCHECK(rism::run_thermo_partial(c->d.T[c->iT_cur], c->d.T[c->iT_last], c->d.u, c->d.Rlast,
                                      c->P, c->N, c->tpartial, c->st));
  }
  HIPCHECK(rism::run_reduce_cols(c->tpartial, tb, 4, c->natv, c->sums, nullptr, c->st));
  std::vector<double> s(4 * c->natv);
  HIPCHECK(hipMemcpyAsync(s.data(), c->sums, sizeof(double) * s.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  const int n = c->natv;
  const double dv = c->P.dv;
  const double ibeta = kAvogadro * kBoltzmann * c->temper;
  std::memset(out, 0, sizeof(*out));
  out->natv = n;
  double xsc = 0, xhnc = 0, xgf = 0, cuv = 0;
  for (int v = 0; v < n; ++v) {
    const double sc = s[4 * v] * c->rho[v] * dv;  // cal_exchem.cu:15,46
    const double sh = s[4 * v + 1] * c->rho[v] * dv;
    const double gf = s[4 * v + 2] * c->rho[v] * dv;
    cuv += s[4 * v + 3] * c->rho[v];  // cal_pmv.cu:14
    out->sfec_sc[v] = ibeta * sc;
    out->sfec_sc_hnc[v] = (c->P.clos == 0) ? ibeta * sh : 0.0;
    out->sfec_gf[v] = ibeta * gf;
    xsc += sc;
    xhnc += sh;
    xgf += gf;
  }
  cuv *= dv;
  const double pmv = c->xt * (1.0 - cuv) * ibeta;  // cal_pmv.cu:21
  double pres = 0.0;                                // cal_pressure.cu:8-19
  for (int v = 0; v < n; ++v) pres += c->rho[v];
  pres *= 1.0e30 / kAvogadro;
  pres = 0.5 * (pres + 1.0 / (c->xt * ibeta));
  out->sfe_sc = ibeta * xsc;  // output_xmu.cu:21-31
  out->sfe_sc_hnc = (c->P.clos == 0) ? ibeta * xhnc : 0.0;
  out->sfe_gf = ibeta * xgf;
  out->pmv = pmv * 1000.0;
  out->pressure = ibeta * pres;
  out->correction = -pres * pmv * ibeta;
  return RISM_OK;
}

int rism_get_field(rism_ctx *c, int which, int site, double *host) {
  if (!c || !host) return rism::set_error(RISM_ERR_ARG, "null argument");
  if (which != RISM_FIELD_FR && (site < 0 || site >= c->natv))
    return rism::set_error(RISM_ERR_ARG, "site out of range");
  const size_t off = (size_t)site * c->N;
  const bool low = (which == RISM_FIELD_T) ? c->prec == 1 : c->last_prec == 1;
  const double *src = nullptr;
  switch (which) {
    case RISM_FIELD_T:
      if (low) {
        HIPCHECK(rism::run_convert(c->f.T[c->iT_cur] + off, c->scratch, c->N, c->st));
        src = c->scratch;
      } else {
        src = c->d.T[c->iT_cur] + off;
      }
      break;
    case RISM_FIELD_U:
      src = c->d.u + off;
      break;
    case RISM_FIELD_FR:
      src = c->d.fr;
      break;
    case RISM_FIELD_R:
      if (!c->have_step) return rism::set_error(RISM_ERR_STATE, "no step taken yet");
      if (low) {
        HIPCHECK(rism::run_convert(c->f.Rlast + off, c->scratch, c->N, c->st));
        src = c->scratch;
      } else {
        src = c->d.Rlast + off;
      }
      break;
    case RISM_FIELD_G:
    case RISM_FIELD_H:
    case RISM_FIELD_C: {
      const int w = which == RISM_FIELD_G ? 1 : (which == RISM_FIELD_H ? 2 : 3);
      if (w != 1 && !c->have_step) return rism::set_error(RISM_ERR_STATE, "no step taken yet");
      const int it = c->have_step ? c->iT_last : c->iT_cur;
      const bool lw = c->have_step ? c->last_prec == 1 : c->prec == 1;
      // g alone before any step: R is not read
      if (lw) {
        const float *R = c->f.Rlast ? c->f.Rlast + off : c->f.T[it] + off;
        HIPCHECK(rism::run_fields(c->f.T[it] + off, c->f.u + off, R, c->P, c->N, w, c->scratch,
                                  c->st));
      } else {
        const double *R = c->d.Rlast ? c->d.Rlast + off : c->d.T[it] + off;
        HIPCHECK(rism::run_fields(c->d.T[it] + off, c->d.u + off, R, c->P, c->N, w, c->scratch,
                                  c->st));
      }
      src = c->scratch;
      break;
    }
    default:
      return rism::set_error(RISM_ERR_ARG, "unknown field");
  }
  HIPCHECK(hipMemcpyAsync(host, src, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  return RISM_OK;
}

int rism_set_field(rism_ctx *c, int which, int site, const double *host) {
  if (!c || !host) return rism::set_error(RISM_ERR_ARG, "null argument");
  if (site < 0 || site >= c->natv) return rism::set_error(RISM_ERR_ARG, "site out of range");
  const size_t off = (size_t)site * c->N;
  if (which == RISM_FIELD_T) {
    if (c->prec == 1) {
      // the FP64 copy is kept too: a mixed solve whose FP32 phase overflows
      // restarts in FP64 from it
      HIPCHECK(hipMemcpyAsync(c->d.T[c->iT_cur] + off, host, sizeof(double) * c->N,
                              hipMemcpyHostToDevice, c->st));
      HIPCHECK(rism::run_convert((const double *)c->d.T[c->iT_cur] + off,
                                 c->f.T[c->iT_cur] + off, c->N, c->st));
    } else {
      HIPCHECK(hipMemcpyAsync(c->d.T[c->iT_cur] + off, host, sizeof(double) * c->N,
                              hipMemcpyHostToDevice, c->st));
    }
  } else if (which == RISM_FIELD_U) {
    HIPCHECK(hipMemcpyAsync(c->d.u + off, host, sizeof(double) * c->N, hipMemcpyHostToDevice, c->st));
    if (c->f.u) HIPCHECK(rism::run_convert((const double *)c->d.u + off, c->f.u + off, c->N, c->st));
  } else {
    return rism::set_error(RISM_ERR_ARG, "only T and U can be set");
  }
  HIPCHECK(hipStreamSynchronize(c->st));
  return RISM_OK;
}

int rism_get_timing(rism_ctx *c, rism_timing *out) {
  if (!c || !out) return rism::set_error(RISM_ERR_ARG, "null argument");
  *out = c->timing;
  return RISM_OK;
}

int rism_solve_batch(rism_ctx **ctx, int n, int maxstep, double tol, int threads, int *status,
                     int *steps, double *rms) {
  if (!ctx || n < 0) return rism::set_error(RISM_ERR_ARG, "null argument");
  for (int i = 0; i < n; ++i) {
    if (!ctx[i]) return rism::set_error(RISM_ERR_ARG, "null context in batch");
    for (int j = 0; j < i; ++j)
      if (ctx[j] == ctx[i]) return rism::set_error(RISM_ERR_ARG, "context repeated in batch");
  }
  if (maxstep <= 0) return rism::set_error(RISM_ERR_ARG, "maxstep must be positive");
  if (n == 0) return RISM_OK;
  std::vector<AsyncRun> runs(n);
  const int nt = std::max(1, std::min(threads > 0 ? threads : 1, n));
  std::vector<int> rcs(nt, RISM_OK);
  std::vector<std::string> msgs(nt);  // error messages are per thread
  // FP64 AN2 contexts run without a host round trip per step, in groups of
  // equal grid, box, solvent and options (one launch per kernel and group,
  // up to kMaxGroup contexts); other FP64 contexts that can run sync-free
  // (Anderson(m), kernel timing on) run so one by one; contexts whose steps
  // need the host (mixed precision, verbose) run rism_solve's loop.  The
  // groups and contexts are dealt to the worker threads, each of which
  // enqueues a chunk on every one of its groups and contexts, then drains
  // them in turn, until every run has finished.
  std::vector<std::vector<int>> groups;
  std::vector<int> single, slow;
  for (int i = 0; i < n; ++i) {
    rism_ctx *c = ctx[i];
    if (!async_capable(c)) {
      slow.push_back(i);
    } else if (c->opts.anderson_m > 0 || c->ktime || c->prec != 0) {
      single.push_back(i);
    } else {
      bool put = false;
      for (auto &gr : groups)
        if ((int)gr.size() < kMaxGroup && groupable(ctx[gr[0]], c)) {
          gr.push_back(i);
          put = true;
          break;
        }
      if (!put) groups.push_back({i});
    }
  }
  // a group of one runs the per-context sync-free path (its kernels read
  // their pointers from their arguments rather than from a table)
  for (size_t q = 0; q < groups.size();) {
    if (groups[q].size() == 1) {
      single.push_back(groups[q][0]);
      groups.erase(groups.begin() + q);
    } else {
      ++q;
    }
  }
  std::vector<std::unique_ptr<GroupRun>> gruns(groups.size());
  auto work = [&](int w) {
    int rc = RISM_OK;
    std::vector<GroupRun *> mine;
    for (size_t q = w; q < groups.size(); q += nt) {
      gruns[q].reset(new GroupRun());
      for (int i : groups[q]) {
        gruns[q]->m.push_back(ctx[i]);
        gruns[q]->runs.push_back(&runs[i]);
      }
      mine.push_back(gruns[q].get());
    }
    std::vector<int> fast, slw;
    for (size_t q = w; q < single.size(); q += nt) fast.push_back(single[q]);
    for (size_t q = w; q < slow.size(); q += nt) slw.push_back(slow[q]);
    for (GroupRun *g : mine) {
      if (rc) break;
      rc = group_begin(*g, maxstep, tol);
    }
    for (int i : fast) {
      if (rc) break;
      rc = async_begin(runs[i], ctx[i], maxstep, tol);
    }
    bool left = true;
    while (!rc && left) {
      left = false;
      for (GroupRun *g : mine) {
        if (rc) break;
        if (!g->finished) rc = group_enqueue(*g, kAsyncChunk);
      }
      for (int i : fast) {
        if (rc) break;
        if (!runs[i].finished) rc = async_enqueue<double>(runs[i], kAsyncChunk);
      }
      for (GroupRun *g : mine) {
        if (rc) break;
        if (g->finished) continue;
        if (hipStreamSynchronize(g->c0()->st) != hipSuccess)
          rc = rism::set_error(RISM_ERR_HIP, "hipStreamSynchronize failed");
        if (!rc) rc = group_check(*g);
        left = left || !g->finished;
      }
      for (int i : fast) {
        if (rc) break;
        if (runs[i].finished) continue;
        if (hipStreamSynchronize(ctx[i]->st) != hipSuccess)
          rc = rism::set_error(RISM_ERR_HIP, "hipStreamSynchronize failed");
        if (!rc) rc = async_check<double>(runs[i]);
        left = left || !runs[i].finished;
      }
    }
    for (GroupRun *g : mine) {
      // a failed group must not leave work in flight on its stream
      (void)hipStreamSynchronize(g->c0()->st);
      for (rism_ctx *c : g->m) c->async = false;
    }
    for (int i : fast) ctx[i]->async = false;
    for (int i : slw) {
      if (rc) break;
      rc = rism_solve(ctx[i], maxstep, tol, &runs[i].status, &runs[i].steps, &runs[i].rms);
    }
    rcs[w] = rc;
    if (rc) msgs[w] = rism_last_error();
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < nt; ++w) pool.emplace_back(work, w);
  work(0);
  for (auto &th : pool) th.join();
  for (int w = 0; w < nt; ++w)
    if (rcs[w]) return rism::set_error(rcs[w], msgs[w]);
  for (int i = 0; i < n; ++i) {
    if (status) status[i] = runs[i].status;
    if (steps) steps[i] = runs[i].steps;
    if (rms) rms[i] = runs[i].rms;
    ctx[i]->timing.steps = runs[i].steps;
    ctx[i]->timing.status = runs[i].status;
    ctx[i]->timing.rms = runs[i].rms;
  }
  return RISM_OK;
}

int rism_run_batch(rism_ctx **ctx, int n, int maxstep, double tol, int threads,
                   rism_thermo_out *out, int *status, int *steps) {
  if (!ctx || n < 0) return rism::set_error(RISM_ERR_ARG, "null argument");
  for (int i = 0; i < n; ++i) {
    if (!ctx[i]) return rism::set_error(RISM_ERR_ARG, "null context in batch");
    int rc = setup_enqueue(ctx[i]);
    if (rc) return rc;
  }
  int rc = rism_solve_batch(ctx, n, maxstep, tol, threads, status, steps, nullptr);
  if (rc) return rc;
  for (int i = 0; i < n; ++i) {
    rism_thermo_out th;
    if ((rc = rism_thermo(ctx[i], &th))) return rc;
    if (out) out[i] = th;
  }
  return RISM_OK;
}

int rism_run(rism_ctx *c, int maxstep, double tol, rism_thermo_out *out, rism_timing *timing) {
  if (!c) return rism::set_error(RISM_ERR_ARG, "null context");
  hipEvent_t e[4];
  for (auto &ev : e) HIPCHECK(hipEventCreate(&ev));
  int rc = RISM_OK;
  int status = 0, steps = 0;
  double rms = 0.0;
  rism_thermo_out th;
  if (hipEventRecord(e[0], c->st) != hipSuccess) rc = RISM_ERR_HIP;
  if (!rc) rc = rism_setup_potentials(c);
  if (!rc && hipEventRecord(e[1], c->st) != hipSuccess) rc = RISM_ERR_HIP;
  if (!rc) rc = rism_solve(c, maxstep, tol, &status, &steps, &rms);
  if (!rc && hipEventRecord(e[2], c->st) != hipSuccess) rc = RISM_ERR_HIP;
  if (!rc) rc = rism_thermo(c, &th);
  if (!rc && hipEventRecord(e[3], c->st) != hipSuccess) rc = RISM_ERR_HIP;
  if (!rc && hipEventSynchronize(e[3]) != hipSuccess) rc = RISM_ERR_HIP;
  if (!rc) {
    float ms[3];
    for (int i = 0; i < 3; ++i) (void)hipEventElapsedTime(&ms[i], e[i], e[i + 1]);
    c->timing.setup_ms = ms[0];
    c->timing.iterate_ms = ms[1];
    c->timing.thermo_ms = ms[2];
    c->timing.steps = steps;
    c->timing.status = status;
    c->timing.rms = rms;
    if (out) *out = th;
    if (timing) *timing = c->timing;
  } else if (rc == RISM_ERR_HIP) {
    rism::set_error(rc, "hipEvent failure in rism_run");
  }
  for (auto &ev : e) (void)hipEventDestroy(ev);
  return rc;
}

// ---- .guv / .tuv -----------------------------------------------------------

int rism_write_guv(rism_ctx *c, const char *path, const double shift[3]) {
  if (!c || !path) return rism::set_error(RISM_ERR_ARG, "null argument");
  std::ofstream o(path, std::ios::binary);
  if (!o) return rism::set_error(RISM_ERR_IO, std::string("cannot write ") + path);
  const int32_t hdr[4] = {c->natv, c->nx, c->ny, c->nz};  // output_guv.cu:24-27
  const double s0[3] = {0, 0, 0};
  o.write((const char *)hdr, sizeof(hdr));
  o.write((const char *)c->box, sizeof(double) * 3);
  o.write((const char *)(shift ? shift : s0), sizeof(double) * 3);
  std::vector<double> g(c->N);
  for (int v = 0; v < c->natv; ++v) {
    int rc = rism_get_field(c, RISM_FIELD_G, v, g.data());
    if (rc) return rc;
    o.write((const char *)g.data(), sizeof(double) * c->N);
  }
  return o ? RISM_OK : rism::set_error(RISM_ERR_IO, std::string("write failed: ") + path);
}

// .huv / .cuv text dumps (output_huv.cu:7-50, output_cuv.cu:7-50, one rank):
// box, grid, natu, per atom "q sig eps x y z", natv, then one line per grid
// point with the natv values, default stream formatting.  The .cuv values
// keep the reference's "(re,im)" form (it writes the complex h minus t).
// The reference's host h array is never filled from the device, so its
// files hold zeros; these hold the fields of the last step.
static int write_text_field(rism_ctx *c, const char *path, int which) {
  if (!c || !path) return rism::set_error(RISM_ERR_ARG, "null argument");
  if (!c->have_step) return rism::set_error(RISM_ERR_STATE, "no step taken yet");
  std::vector<double> f((size_t)c->natv * c->N);
  for (int v = 0; v < c->natv; ++v) {
    int rc = rism_get_field(c, which, v, f.data() + (size_t)v * c->N);
    if (rc) return rc;
  }
  std::ofstream o(path);
  if (!o) return rism::set_error(RISM_ERR_IO, std::string("cannot write ") + path);
  o << c->box[0] << " " << c->box[1] << " " << c->box[2] << " " << std::endl
    << c->nx << " " << c->ny << " " << c->nz << " " << std::endl
    << c->natu << std::endl;
  for (int a = 0; a < c->natu; ++a)
    o << c->uq[a] << " " << c->usig[a] << " " << c->ueps[a] << " " << c->uxyz[3 * a] << " "
      << c->uxyz[3 * a + 1] << " " << c->uxyz[3 * a + 2] << std::endl;
  o << c->natv << std::endl;
  for (size_t i = 0; i < c->N; ++i) {
    for (int v = 0; v < c->natv; ++v) {
      const double x = f[(size_t)v * c->N + i];
      if (which == RISM_FIELD_C)
        o << "(" << x << ",0) ";
      else
        o << x << " ";
    }
    o << "\n";
  }
  o.flush();
  return o ? RISM_OK : rism::set_error(RISM_ERR_IO, std::string("write failed: ") + path);
}

int rism_write_huv(rism_ctx *c, const char *path) { return write_text_field(c, path, RISM_FIELD_H); }

int rism_analysis(rism_ctx *c, int kind, double *out) {
  if (!c || !out) return rism::set_error(RISM_ERR_ARG, "null argument");
  if (kind < RISM_ANALYSIS_GRAD || kind > RISM_ANALYSIS_AD)
    return rism::set_error(RISM_ERR_ARG, "unknown analysis kind");
  if (!c->have_step || c->last_prec != 0)
    return rism::set_error(RISM_ERR_STATE, "the analysis needs a solved FP64 state");
  const int nv = c->natv, nu = c->natu;
  if (nu == 0) return RISM_OK;
  // cal_LJ.cu:33-44: sigma_uv (lambda-scaled for -l), eps_uv = sqrt(eps_u eps_v)
  const double lam1 = (c->opts.ad_switch == 1) ? c->opts.lambda : 1.0;
  std::vector<double> adata((size_t)nu * (4 + 2 * nv));
  for (int a = 0; a < nu; ++a) {
    double *p = adata.data() + (size_t)a * (4 + 2 * nv);
    p[0] = c->uxyz[3 * a];
    p[1] = c->uxyz[3 * a + 1];
    p[2] = c->uxyz[3 * a + 2];
    p[3] = c->uq[a];
    for (int v = 0; v < nv; ++v) {
      p[4 + v] = (c->usig[a] + c->sigv[v]) * 0.5 * lam1;
      p[4 + nv + v] = std::sqrt(c->ueps[a] * c->epsv[v]);
    }
  }
  const int k = kind == RISM_ANALYSIS_GRAD ? 0
              : kind == RISM_ANALYSIS_EUV  ? 1
              : (c->opts.ad_switch == 1 ? 2 : 3);
  const int width = k == 0 ? 3 : (k == 1 ? nv : 1);
  const size_t nslab = (c->N + rism::atom_sum_slab(c->N) - 1) / rism::atom_sum_slab(c->N);
  double *dad = nullptr, *dpart = nullptr, *dout = nullptr;
  auto cleanup = [&]() {
    if (dad) (void)hipFree(dad);
    if (dpart) (void)hipFree(dpart);
    if (dout) (void)hipFree(dout);
  };
  if (hipMalloc(&dad, sizeof(double) * adata.size()) != hipSuccess ||
      hipMalloc(&dpart, sizeof(double) * nslab * nu * width) != hipSuccess ||
      hipMalloc(&dout, sizeof(double) * nu * width) != hipSuccess) {
    (void)hipGetLastError();
    cleanup();
    return rism::set_error(RISM_ERR_NOMEM, "hipMalloc failed in rism_analysis");
  }
  std::vector<double> sums((size_t)nu * width);
  const double cc = kHartree * kBohr * kAvogadro;  // cal_grad.cu:47
  hipError_t e = hipMemcpyAsync(dad, adata.data(), sizeof(double) * adata.size(),
                                hipMemcpyHostToDevice, c->st);
  if (e == hipSuccess)
    e = rism::run_atom_sums(k, c->d.T[c->iT_last], c->d.u, c->P, dad, nu,
                            c->opts.lambda, cc, c->rho_dev, dpart, c->st);
  if (e == hipSuccess)
    e = rism::run_reduce_cols(dpart, nslab, nu * width, 1, dout, nullptr, c->st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(sums.data(), dout, sizeof(double) * sums.size(), hipMemcpyDeviceToHost,
                       c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  cleanup();
  if (e != hipSuccess) return rism::set_error(RISM_ERR_HIP, hipGetErrorString(e));
  const double dv = c->P.dv;
  constexpr double kcal2J = 4184.0;  // physical.h:17
  if (k == 1) {  // output_euv.cu:13-27: dv * (total, per site), rho_v-weighted
    for (int a = 0; a < nu; ++a) {
      double tot = 0.0;
      for (int v = 0; v < nv; ++v) {
        const double ev = sums[(size_t)a * nv + v] * c->rho[v];
        out[(size_t)a * (nv + 1) + 1 + v] = dv * ev;
        tot += ev;
      }
      out[(size_t)a * (nv + 1)] = dv * tot;
    }
  } else {  // output_grad.cu:13-20, output_ad.cu:15-19: * dv / kcal2J
    for (size_t i = 0; i < sums.size(); ++i) out[i] = sums[i] * dv / kcal2J;
  }
  return RISM_OK;
}

int rism_write_cuv(rism_ctx *c, const char *path) { return write_text_field(c, path, RISM_FIELD_C); }

int rism_write_tuv(rism_ctx *c, const char *path) {
  if (!c || !path) return rism::set_error(RISM_ERR_ARG, "null argument");
  std::ofstream o(path);
  if (!o) return rism::set_error(RISM_ERR_IO, std::string("cannot write ") + path);
  o << c->nx << " " << c->ny << " " << c->nz << " " << c->natv << std::endl;  // write_tuv.cu:19-22
  std::vector<double> t(c->N);
  for (int v = 0; v < c->natv; ++v) {
    int rc = rism_get_field(c, RISM_FIELD_T, v, t.data());
    if (rc) return rc;
    for (size_t i = 0; i < c->N; ++i) o << t[i] << "\n";
  }
  o.flush();
  return o ? RISM_OK : rism::set_error(RISM_ERR_IO, std::string("write failed: ") + path);
}

int rism_read_tuv(rism_ctx *c, const char *path) {
  if (!c || !path) return rism::set_error(RISM_ERR_ARG, "null argument");
  std::ifstream f(path);
  if (!f) return rism::set_error(RISM_ERR_IO, std::string("cannot open ") + path);
  int g0, g1, g2, nv;
  f >> g0 >> g1 >> g2 >> nv;
  if (!f) return rism::set_error(RISM_ERR_IO, "bad .tuv header");
  if (g0 != c->nx || g1 != c->ny || g2 != c->nz || nv != c->natv) {  // read_tuv.cu:14-20
    std::cout << " RXRISM:  actual   NGr = " << c->N << "   NatV = " << c->natv << std::endl
              << "           saved   NGr = " << (long long)g0 * g1 * g2 << "   NatV = " << nv
              << std::endl;
    return rism::set_error(RISM_ERR_ARG, ".tuv grid does not match the input");
  }
  std::vector<double> t(c->N);
  for (int v = 0; v < c->natv; ++v) {
    for (size_t i = 0; i < c->N; ++i) f >> t[i];
    if (!f) return rism::set_error(RISM_ERR_IO, "truncated .tuv");
    int rc = rism_set_field(c, RISM_FIELD_T, v, t.data());
    if (rc) return rc;
  }
  return RISM_OK;
}

// ---- kernel-level entry points ----------------------------------------------

}  // extern "C"

namespace {
int transform_ctx(int nx, int ny, int nz, const double box[3], rism_ctx **out) {
  int rc = check_sizes(nx, ny, nz, 1);
  if (rc) return rc;
  rism_ctx *c = new rism_ctx();
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->natv = 1;
  c->N = (size_t)nx * ny * nz;
  for (int d = 0; d < 3; ++d) c->box[d] = box[d];
  c->qv.assign(1, 0.0);
  fill_params(c);
  if (hipStreamCreate(&c->st) != hipSuccess) {
    delete c;
    return rism::set_error(RISM_ERR_HIP, "hipStreamCreate failed");
  }
  if ((rc = make_tables(c, c->d.tb))) {
    delete c;
    return rc;
  }
  *out = c;
  return RISM_OK;
}
}  // namespace

extern "C" {

int rism_k_fft_r2c(int nx, int ny, int nz, const double box[3], const double *in, double *out) {
  if (!box || !in || !out) return rism::set_error(RISM_ERR_ARG, "null argument");
  rism_ctx *c = nullptr;
  int rc = transform_ctx(nx, ny, nz, box, &c);
  if (rc) return rc;
  struct Del {
    rism_ctx *c;
    ~Del() { delete c; }
  } del{c};
  double *t;
  cplx *spec;
  if ((rc = c->alloc(&t, c->N)) || (rc = c->alloc(&spec, c->N / 2))) return rc;
  HIPCHECK(hipMemcpy(t, in, sizeof(double) * c->N, hipMemcpyHostToDevice));
  rism::Params P = c->P;
  P.mode = 1;
  HIPCHECK(rism::run_zfwd(P, t, t, t, spec, c->d.tb.twz, c->d.tb.zpre, c->d.tb.zpost, c->st));
  HIPCHECK(rism::run_y(P, spec, -1, c->d.tb.twy, c->d.tb.prey, c->st));
  HIPCHECK(rism::run_xoz(P, spec, spec, c->d.tb.twx, c->d.tb.prex, nullptr, nullptr,
                         (const double *)nullptr, c->st));
  HIPCHECK(rism::run_apply_phase(spec, P, 0, c->st));
  HIPCHECK(hipMemcpyAsync(out, spec, sizeof(cplx) * (c->N / 2), hipMemcpyDeviceToHost, c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  return RISM_OK;
}

int rism_k_fft_c2r(int nx, int ny, int nz, const double box[3], const double *in, double *out) {
  if (!box || !in || !out) return rism::set_error(RISM_ERR_ARG, "null argument");
  rism_ctx *c = nullptr;
  int rc = transform_ctx(nx, ny, nz, box, &c);
  if (rc) return rc;
  struct Del {
    rism_ctx *c;
    ~Del() { delete c; }
  } del{c};
  double *h;
  cplx *spec;
  double *partial;
  if ((rc = c->alloc(&h, c->N)) || (rc = c->alloc(&spec, c->N / 2)) ||
      (rc = c->alloc(&partial, rism::zinv_num_blocks<double>(c->P, 0) * 6)))
    return rc;
  HIPCHECK(hipMemcpy(spec, in, sizeof(cplx) * (c->N / 2), hipMemcpyHostToDevice));
  rism::Params P = c->P;
  P.mode = 2;
  HIPCHECK(rism::run_apply_phase(spec, P, 1, c->st));
  HIPCHECK(rism::run_xoz(P, spec, spec, c->d.tb.twx, c->d.tb.prex, nullptr, nullptr,
                         (const double *)nullptr, c->st));
  HIPCHECK(rism::run_y(P, spec, +1, c->d.tb.twy, c->d.tb.prey, c->st));
  HIPCHECK(rism::run_zinv(P, spec, h, h, h, h, h, (const double *)nullptr, partial, c->d.tb.twz,
                          c->d.tb.zpre, c->d.tb.zpost, 0, c->st));
  HIPCHECK(hipMemcpyAsync(out, h, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  return RISM_OK;
}

int rism_k_get_fk(rism_ctx *c, double *out) {
  if (!c || !out) return rism::set_error(RISM_ERR_ARG, "null argument");
  cplx *tmp = c->d.spec;  // borrow the spectrum buffer
  HIPCHECK(hipMemcpyAsync(tmp, c->d.fk, sizeof(cplx) * (c->N / 2), hipMemcpyDeviceToDevice, c->st));
  HIPCHECK(rism::run_apply_phase(tmp, c->P, 0, c->st));
  HIPCHECK(hipMemcpyAsync(out, tmp, sizeof(cplx) * (c->N / 2), hipMemcpyDeviceToHost, c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  return RISM_OK;
}

int rism_k_copy_bench(size_t bytes, int reps, double *avg_ms) {
  if (bytes < 16 || reps < 1) return rism::set_error(RISM_ERR_ARG, "bytes >= 16 and reps >= 1");
  const size_t n = (bytes / 16) * 2;  // doubles, whole 16-byte vectors
  double *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = RISM_OK;
  if (hipMalloc((void **)&a, n * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&b, n * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    rc = rism::set_error(RISM_ERR_NOMEM, "hipMalloc failed in rism_k_copy_bench");
  }
  if (!rc && (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
              hipEventCreate(&e1) != hipSuccess))
    rc = rism::set_error(RISM_ERR_HIP, "stream/event creation failed");
  if (!rc && (hipMemsetAsync(a, 0, n * sizeof(double), st) != hipSuccess ||
              rism::run_copy(a, b, n, st) != hipSuccess))  // first touch, untimed
    rc = rism::set_error(RISM_ERR_HIP, "copy warm-up failed");
  if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = RISM_ERR_HIP;
  for (int r = 0; r < reps && !rc; ++r)
    if (rism::run_copy(a, b, n, st) != hipSuccess) rc = RISM_ERR_HIP;
  if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
    rc = RISM_ERR_HIP;
  if (!rc) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (avg_ms) *avg_ms = ms / reps;
  } else if (rc == RISM_ERR_HIP) {
    rism::set_error(rc, "copy launch failed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

int rism_tune_report(rism_tune_entry *out, int cap) {
  std::lock_guard<std::mutex> g(tune::mu);
  tune::load_locked();
  int n = 0;
  for (const auto &kv : tune::table) {
    if (out && n < cap) {
      rism_tune_entry &e = out[n];
      e.pass = std::get<0>(kv.first);
      e.nx = std::get<1>(kv.first);
      e.ny = std::get<2>(kv.first);
      e.nz = std::get<3>(kv.first);
      e.natv = std::get<4>(kv.first);
      e.fp32 = std::get<5>(kv.first);
      e.choice = kv.second.choice;
      for (int v = 0; v < 3; ++v) e.ms[v] = kv.second.ms[v];
    }
    ++n;
  }
  return n;
}

int rism_k_an2_state(rism_ctx *c, double *x, double *s, int *count, int *binary) {
  if (!c) return rism::set_error(RISM_ERR_ARG, "null context");
  if (x) {
    x[0] = c->lastx[0];
    x[1] = c->lastx[1];
  }
  if (s) {
    s[0] = c->lasts[0];
    s[1] = c->lasts[1];
  }
  if (count) *count = c->count;
  if (binary) *binary = c->binary;
  return RISM_OK;
}

}  // extern "C"
