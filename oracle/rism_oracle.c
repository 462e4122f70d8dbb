/*
 * rism_oracle.c -- FP64 CPU restatement of the reference 3D-RISM algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product
 * path is checked against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or the
 * timed CPU baseline), never as the thing that ships.
 *
 * Every function cites the reference file:line it restates.  Layouts follow
 * the reference: real fields are [site][z][y][x] (x fastest), k-space fields
 * are full complex arrays of N points per site, interleaved (re, im).
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   or_fft3d          brute-force O(N^2) shifted DFT, numpy.fft with explicit
 *                     twiddles, round trip, Parseval, FFT(fr) ~ fk.
 *   or_kshells        brute-force sort of k^2 (numpy.unique).
 *   or_spline_*       scipy CubicSpline(bc_type="natural").
 *   or_lj/or_coulomb/ brute-force numpy sums on 1-3 atoms, single-atom
 *   or_fk             closed forms, r<1e-5 / s^2 floor guards.
 *   or_closure        closed forms; KH == HNC where -u+t < 0.
 *   or_an2_*          closed-form 2x2 solve (numpy), linear fixed point.
 *   or_thermo         closed-form integrands; Pressure known answer from
 *                     example_milc/mobley_3053621.xmu.
 *   or_solve          end-to-end: example-pegasus/test1.xmu (128^3) and
 *                     example_milc/mobley_3053621.xmu/.xmu0 (256^3).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* physical.h:5-17 */
#define OR_AVOGADRO 6.02214076e23
#define OR_BOHR 0.5291771500001
#define OR_BOLTZMANN 1.380649e-23
#define OR_HARTREE 4.359744722207185e-18

/*
 * Compensated (Neumaier) accumulation for the long sums over the grid (rms,
 * Gram entries, thermodynamic integrals).  The reference reduces these with
 * thrust::reduce, a tree whose rounding error does not grow with the length;
 * a plain running sum over 10^8 terms per thread drifts by ~1e-11 relative,
 * which is what the check of a 512^3 step would then see.  Each thread keeps
 * OR_NSUM compensated sums over its static chunk; the per-thread results are
 * added in thread order, so a run is reproducible for a fixed thread count.
 */
#define OR_NSUM 5
#define OR_MAXTHR 1024
typedef struct { double s[OR_NSUM], c[OR_NSUM]; } or_ksum;
static inline void ks_add(or_ksum *k, int j, double x) {
  const double t = k->s[j] + x;
  if (fabs(k->s[j]) >= fabs(x)) k->c[j] += (k->s[j] - t) + x;
  else k->c[j] += (x - t) + k->s[j];
  k->s[j] = t;
}
static inline int or_thread_id(void) {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}
static inline int or_thread_count(void) {
#ifdef _OPENMP
  int n = omp_get_max_threads();
  return n < OR_MAXTHR ? n : OR_MAXTHR;
#else
  return 1;
#endif
}
/* sums the per-thread partials part[0..nt) of component j, in order */
static double ks_total(const or_ksum *part, int nt, int j) {
  or_ksum k;
  memset(&k, 0, sizeof(k));
  for (int i = 0; i < nt; ++i) {
    ks_add(&k, j, part[i].s[j]);
    ks_add(&k, j, part[i].c[j]);
  }
  return k.s[j] + k.c[j];
}

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) {
  cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}

/* ------------------------------------------------------------------------- */
/* 1D complex DFT of any length, VB lines at a time: an iterative mixed-radix */
/* self-sorting (Stockham) transform on planar blocks [i][re|im][VB lane].    */
/* Radix 4 and 2 have unrolled butterflies; radix 3, 5 and any other prime  */
/* are evaluated from the DFT definition.  Used by or_fft3d.                 */
/* ------------------------------------------------------------------------- */
#define OR_VB 8

typedef struct {
  int n;
  int nf;
  int fac[32];
  cplx *tw[32]; /* stage s: tw[s][k * (r - 1) + q - 1] = e^{-2 pi i q k / (Ns r)}, k < Ns */
  cplx *wr[32]; /* stage s: wr[s][t] = e^{-2 pi i t / r}, t < r                          */
} fftplan;

static void plan_init(fftplan *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  /* prefer radix 4, then 2, 3, 5, then any prime */
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
  }
  int ns = 1;
  for (int st = 0; st < p->nf; ++st) {
    int r = p->fac[st];
    p->tw[st] = (cplx *)malloc(sizeof(cplx) * (size_t)ns * (r - 1) + sizeof(cplx));
    for (int k = 0; k < ns; ++k)
      for (int q = 1; q < r; ++q) {
        double a = -2.0 * M_PI * (double)q * (double)k / ((double)ns * r);
        p->tw[st][k * (r - 1) + q - 1].re = cos(a);
        p->tw[st][k * (r - 1) + q - 1].im = sin(a);
      }
    p->wr[st] = (cplx *)malloc(sizeof(cplx) * (size_t)r);
    for (int t = 0; t < r; ++t) {
      double a = -2.0 * M_PI * (double)t / (double)r;
      p->wr[st][t].re = cos(a);
      p->wr[st][t].im = sin(a);
    }
    ns *= r;
  }
}

static void plan_free(fftplan *p) {
  for (int st = 0; st < p->nf; ++st) {
    free(p->tw[st]);
    free(p->wr[st]);
  }
}

/* element i of a planar block: re at blk + i*2*OR_VB, im OR_VB further */
#define RE(blk, i) ((blk) + (size_t)(i) * 2 * OR_VB)
#define IM(blk, i) ((blk) + (size_t)(i) * 2 * OR_VB + OR_VB)

/* One Stockham stage of radix r over the block: butterfly j (j < n/r, k = j mod
 * Ns) reads elements j + q n/r, multiplies element q by e^{sign 2 pi i q k /
 * (Ns r)}, takes the length-r DFT and writes output t to (j/Ns) Ns r + k + t Ns.
 * sign = -1 forward, +1 inverse (unnormalised). */
static void stage(const fftplan *p, int st, int ns, const double *in, double *out, int sign) {
  const int n = p->n, r = p->fac[st], m = n / r;
  const cplx *tw = p->tw[st], *wr = p->wr[st];
  double vr[8][OR_VB], vi[8][OR_VB];
  double *gr = NULL, *gi = NULL;
  if (r > 8) {
    gr = (double *)malloc(sizeof(double) * (size_t)r * OR_VB);
    gi = (double *)malloc(sizeof(double) * (size_t)r * OR_VB);
  }
  for (int j = 0; j < m; ++j) {
    const int k = j % ns;
    const int o0 = (j / ns) * ns * r + k;
    double *xr = (r > 8) ? gr : &vr[0][0];
    double *xi = (r > 8) ? gi : &vi[0][0];
    for (int q = 0; q < r; ++q) {
      const double *sr = RE(in, j + q * m), *si = IM(in, j + q * m);
      double wre = 1.0, wim = 0.0;
      if (q) {
        wre = tw[k * (r - 1) + q - 1].re;
        wim = -sign * tw[k * (r - 1) + q - 1].im; /* conj for the inverse */
      }
      for (int b = 0; b < OR_VB; ++b) {
        xr[q * OR_VB + b] = sr[b] * wre - si[b] * wim;
        xi[q * OR_VB + b] = sr[b] * wim + si[b] * wre;
      }
    }
    if (r == 2) {
      double *a0r = RE(out, o0), *a0i = IM(out, o0), *a1r = RE(out, o0 + ns), *a1i = IM(out, o0 + ns);
      for (int b = 0; b < OR_VB; ++b) {
        a0r[b] = xr[b] + xr[OR_VB + b];
        a0i[b] = xi[b] + xi[OR_VB + b];
        a1r[b] = xr[b] - xr[OR_VB + b];
        a1i[b] = xi[b] - xi[OR_VB + b];
      }
    } else if (r == 4) {
      /* y0 = (x0 + x2) + (x1 + x3), y2 = (x0 + x2) - (x1 + x3),
         y1 = (x0 - x2) + sign i (x1 - x3), y3 = (x0 - x2) - sign i (x1 - x3) */
      double *y0r = RE(out, o0), *y0i = IM(out, o0);
      double *y1r = RE(out, o0 + ns), *y1i = IM(out, o0 + ns);
      double *y2r = RE(out, o0 + 2 * ns), *y2i = IM(out, o0 + 2 * ns);
      double *y3r = RE(out, o0 + 3 * ns), *y3i = IM(out, o0 + 3 * ns);
      for (int b = 0; b < OR_VB; ++b) {
        const double ar = xr[b] + xr[2 * OR_VB + b], ai = xi[b] + xi[2 * OR_VB + b];
        const double cr = xr[b] - xr[2 * OR_VB + b], ci = xi[b] - xi[2 * OR_VB + b];
        const double br = xr[OR_VB + b] + xr[3 * OR_VB + b], bi = xi[OR_VB + b] + xi[3 * OR_VB + b];
        const double dr = xr[OR_VB + b] - xr[3 * OR_VB + b], di = xi[OR_VB + b] - xi[3 * OR_VB + b];
        /* sign i (dr + i di) = (-sign di) + i (sign dr) */
        const double er = -sign * di, ei = sign * dr;
        y0r[b] = ar + br;
        y0i[b] = ai + bi;
        y2r[b] = ar - br;
        y2i[b] = ai - bi;
        y1r[b] = cr + er;
        y1i[b] = ci + ei;
        y3r[b] = cr - er;
        y3i[b] = ci - ei;
      }
    } else {
      /* the definition: y_t = sum_q x_q e^{sign 2 pi i q t / r} */
      for (int t = 0; t < r; ++t) {
        double *yr = RE(out, o0 + t * ns), *yi = IM(out, o0 + t * ns);
        for (int b = 0; b < OR_VB; ++b) {
          yr[b] = 0.0;
          yi[b] = 0.0;
        }
        for (int q = 0; q < r; ++q) {
          const cplx w = wr[(q * t) % r];
          const double wre = w.re, wim = -sign * w.im;
          for (int b = 0; b < OR_VB; ++b) {
            yr[b] += xr[q * OR_VB + b] * wre - xi[q * OR_VB + b] * wim;
            yi[b] += xr[q * OR_VB + b] * wim + xi[q * OR_VB + b] * wre;
          }
        }
      }
    }
  }
  free(gr);
  free(gi);
}

/* unnormalised DFT of the OR_VB lines of a block; returns the buffer that
 * holds the result (blk or tmp) */
static double *fft_block(const fftplan *p, double *blk, double *tmp, int sign) {
  double *a = blk, *b = tmp;
  int ns = 1;
  for (int st = 0; st < p->nf; ++st) {
    stage(p, st, ns, a, b, sign);
    ns *= p->fac[st];
    double *t = a;
    a = b;
    b = t;
  }
  return a;
}

/*
 * or_fft3d: the reference's half-shifted 3D transform (fft3d.cu:81-163,
 * fft3d_cuda.cu:56-128), restated by its definition with the S = +1 phase
 * (exact when every N_d % 4 == 0, which is where the reference is
 * self-consistent):
 *   key = -1 (forward):  F(k_j) = dv * sum_n f(r_n) exp(-i k_j . r_n)
 *   key = +1 (backward): f(r_n) = (1/V) * sum_j F(k_j) exp(+i k_j . r_n)
 * with r_n = (n - N/2) * L/N and k_j = 2 pi (j - N/2 + 1/2) / L per axis
 * (cal_LJ.cu:61-63, initialize_g.cu:58-60).
 * Per axis: F(j) = C (-1)^j sum_n [f_n (-1)^n e^{-i pi n/N}] e^{-2 pi i j n/N},
 * C = i (-1)^{N/2}; the backward transform is its exact inverse.
 */
void or_fft3d(double *data, int nx, int ny, int nz, double bx, double by, double bz,
              int key) {
  int dims[3] = {nx, ny, nz};
  size_t strides[3] = {1, (size_t)nx, (size_t)nx * ny};
  size_t ntot = (size_t)nx * ny * nz;
  cplx *a = (cplx *)data;
  int sign = (key < 0) ? -1 : 1;
  for (int ax = 0; ax < 3; ++ax) {
    int n = dims[ax];
    size_t st = strides[ax];
    fftplan plan;
    plan_init(&plan, n);
    /* pre and post twiddles */
    cplx *pre = (cplx *)malloc(sizeof(cplx) * n);
    cplx *post = (cplx *)malloc(sizeof(cplx) * n);
    double cs = ((n / 2) % 2 == 0) ? 1.0 : -1.0; /* (-1)^{N/2} */
    for (int j = 0; j < n; ++j) {
      double sg = (j % 2 == 0) ? 1.0 : -1.0;
      double ang = -M_PI * (double)j / (double)n;
      /* forward pre: (-1)^n e^{-i pi n / N}; post: i (-1)^{N/2} (-1)^j */
      pre[j].re = sg * cos(ang);
      pre[j].im = sg * sin(ang);
      post[j].re = 0.0;
      post[j].im = cs * sg;
      if (sign > 0) { /* inverse: conj(post) before, conj(pre) after */
        pre[j].im = -pre[j].im;
        post[j].im = -post[j].im;
      }
    }
    size_t nlines = ntot / (size_t)n;
    size_t nblocks = (nlines + OR_VB - 1) / OR_VB;
#pragma omp parallel
    {
      double *blk = (double *)malloc(sizeof(double) * (size_t)n * 2 * OR_VB);
      double *tmp = (double *)malloc(sizeof(double) * (size_t)n * 2 * OR_VB);
      size_t base[OR_VB];
#pragma omp for schedule(static)
      for (size_t bl = 0; bl < nblocks; ++bl) {
        /* lines bl*OR_VB .. +OR_VB-1; for y and z these are neighbouring x */
        int nb = 0;
        for (int b = 0; b < OR_VB; ++b) {
          size_t l = bl * OR_VB + b;
          if (l >= nlines) break;
          if (ax == 0) {
            base[b] = l * (size_t)n;
          } else if (ax == 1) {
            size_t x = l % nx, z = l / nx;
            base[b] = x + z * (size_t)nx * ny;
          } else {
            base[b] = l; /* l = x + y*nx */
          }
          ++nb;
        }
        /* gather with the pre-twiddle (forward) or post-twiddle (inverse) */
        const cplx *tin = (sign < 0) ? pre : post;
        const cplx *tout = (sign < 0) ? post : pre;
        for (int i = 0; i < n; ++i) {
          double *br = RE(blk, i), *bi = IM(blk, i);
          for (int b = 0; b < OR_VB; ++b) {
            if (b < nb) {
              cplx v = cmul(a[base[b] + (size_t)i * st], tin[i]);
              br[b] = v.re;
              bi[b] = v.im;
            } else {
              br[b] = 0.0;
              bi[b] = 0.0;
            }
          }
        }
        double *res = fft_block(&plan, blk, tmp, sign);
        for (int i = 0; i < n; ++i) {
          const double *rr = RE(res, i), *ri = IM(res, i);
          for (int b = 0; b < nb; ++b) {
            cplx v = {rr[b], ri[b]};
            a[base[b] + (size_t)i * st] = cmul(v, tout[i]);
          }
        }
      }
      free(blk);
      free(tmp);
    }
    free(pre);
    free(post);
    plan_free(&plan);
  }
  /* normalisation: forward * dv; backward * 1/V = 1/(dv * N) */
  double dv = (bx / nx) * (by / ny) * (bz / nz);
  double scale = (sign < 0) ? dv : 1.0 / (dv * (double)ntot);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < ntot; ++i) {
    a[i].re *= scale;
    a[i].im *= scale;
  }
}

/* ------------------------------------------------------------------------- */
/* k-grid and shells: initialize_g.cu:7-63                                   */
/* ------------------------------------------------------------------------- */

/* k components per axis index: setg (initialize_g.cu:58-60) */
static inline double kcomp(int j, int n, double b) {
  return 2.0 * M_PI * ((double)j - (double)n / 2.0 + 0.5) / b;
}

typedef struct { double g2; int idx; } g2key;
static int cmp_g2(const void *a, const void *b) {
  double x = ((const g2key *)a)->g2, y = ((const g2key *)b)->g2;
  if (x < y) return -1;
  if (x > y) return 1;
  int i = ((const g2key *)a)->idx, j = ((const g2key *)b)->idx;
  return (i > j) - (i < j);
}

/*
 * or_kshells: k^2 per point, sorted; unique values -> ga (|k| per shell),
 * indga (shell id per point).  initialize_g.cu:21-45.
 * Returns the number of shells; ga must hold ntot entries.
 */
int or_kshells(int nx, int ny, int nz, double bx, double by, double bz, double *ga,
               int *indga) {
  size_t ntot = (size_t)nx * ny * nz;
  g2key *k = (g2key *)malloc(sizeof(g2key) * ntot);
#pragma omp parallel for schedule(static)
  for (size_t ip = 0; ip < ntot; ++ip) {
    int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((size_t)nx * ny));
    double kx = kcomp(x, nx, bx), ky = kcomp(y, ny, by), kz = kcomp(z, nz, bz);
    k[ip].g2 = kx * kx + ky * ky + kz * kz;
    k[ip].idx = (int)ip;
  }
  qsort(k, ntot, sizeof(g2key), cmp_g2);
  double prev = -1.0;
  int nga = 0;
  for (size_t i = 0; i < ntot; ++i) {
    if (k[i].g2 > prev) {
      ga[nga++] = sqrt(k[i].g2);
      prev = k[i].g2;
    }
    indga[k[i].idx] = nga - 1;
  }
  free(k);
  return nga;
}

/* ------------------------------------------------------------------------- */
/* natural cubic spline: spline.cc:4-43, splint.cc:4-20                      */
/* ------------------------------------------------------------------------- */

/* coe[0..2][np] (row-major, 3 rows of np) as in spline.cc */
void or_spline_coef(const double *xp, const double *yp, int np, double *coe) {
  double *a = (double *)calloc(np, sizeof(double));
  double *b = (double *)calloc(np, sizeof(double));
  double *c = (double *)calloc(np, sizeof(double));
  double *d = (double *)calloc(np, sizeof(double));
  double *s = (double *)calloc(np, sizeof(double));
  int n = np - 2;
  for (int k = 1; k <= n; ++k) {
    double hl = xp[k] - xp[k - 1];
    double hr = xp[k + 1] - xp[k];
    a[k] = hl / (hl + hr);
    b[k] = 1.0 - a[k];
    double c1 = (yp[k + 1] - yp[k]) / hr;
    double c2 = (yp[k] - yp[k - 1]) / hl;
    c[k] = 6.0 * ((c1 - c2) / (hl + hr));
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
    double h = xp[k + 1] - xp[k];
    coe[np + k] = s[k] / 2.0;
    coe[2 * np + k] = (s[k + 1] - s[k]) / (6.0 * h);
    coe[k] = (yp[k + 1] - yp[k]) / h - coe[np + k] * h - coe[2 * np + k] * h * h;
  }
  free(a); free(b); free(c); free(d); free(s);
}

double or_splint(const double *xp, const double *yp, const double *coe, int np, double x) {
  int klo = 0, khi = np - 1;
  while (khi - klo > 1) {
    int k = (khi + klo) >> 1;
    if (xp[k] > x) khi = k; else klo = k;
  }
  int k = klo;
  double t = x - xp[k];
  return ((coe[2 * np + k] * t + coe[np + k]) * t + coe[k]) * t + yp[k];
}

/*
 * or_chi_shells: solvent_spline.cu:7-55.  xvv is the .xvk table in file
 * order, xvv[(i*natv + v2)*natv + v1] = X[v2][v1](k_i), k_i = (i+1) pi/(dr ntab).
 * Output chi[(s*natv + v2)*natv + v1] at each shell |k| = ga[s].
 * Returns 0, or 1 for "insufficient maximal T tabulated".
 */
int or_chi_shells(const double *xvv, int natv, int ntab, double dr, const double *ga,
                  int nga, double *chi) {
  if (ntab < 1 || nga < 1) return 1;
  double dk = M_PI / (dr * ntab);
  double *ttab = (double *)malloc(sizeof(double) * ntab);
  for (int i = 0; i < ntab; ++i) ttab[i] = dk * (i + 1);
  if (ga[nga - 1] > ttab[ntab - 1]) { free(ttab); return 1; }
  int ntab1 = 0;
  for (int n = 0; n < ntab; ++n) if (ttab[n] < ga[0]) ntab1 = n;
  int ntab2 = ntab - 1;
  for (int n = ntab - 1; n > 0; --n) if (ttab[n] > ga[nga - 1]) ntab2 = n;
  int np = ntab2 - ntab1 + 1;
  double *x = (double *)malloc(sizeof(double) * np);
  double *y = (double *)malloc(sizeof(double) * np);
  double *coe = (double *)malloc(sizeof(double) * 3 * np);
  for (int v2 = 0; v2 < natv; ++v2)
    for (int v1 = 0; v1 < natv; ++v1) {
      for (int n = 0; n < np; ++n) {
        x[n] = ttab[ntab1 + n];
        y[n] = xvv[((size_t)(ntab1 + n) * natv + v2) * natv + v1];
      }
      or_spline_coef(x, y, np, coe);
      for (int s = 0; s < nga; ++s)
        chi[((size_t)s * natv + v2) * natv + v1] = or_splint(x, y, coe, np, ga[s]);
    }
  free(x); free(y); free(coe); free(ttab);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* potentials: cal_LJ.cu:4-81, cal_Coulomb.cu:4-95, cal_potential.cu:3-43    */
/* ------------------------------------------------------------------------- */

/* r-space coordinate of grid index j: ((int)j - n/2) * d (cal_LJ.cu:61-63) */
static inline double rcomp(int j, int n, double d) { return (double)(j - n / 2) * d; }

/*
 * or_lj: u[v][p] = beta * sum_u 4 eps_uv (s^-12 - s^-6),
 * s^2 = max(|r - r_u|^2 / sig_uv^2, 4e-6); sig_uv = lam1 (sig_u + sig_v)/2,
 * eps_uv = sqrt(eps_u eps_v), beta = 1/(N_A k_B T).  cal_LJ.cu:26-80.
 */
/* or_lj_rows: the same on the sub-block of rows z in [z0, z0 + nzs),
 * y in [y0, y0 + nys) of the full grid (every x), u[v][(z - z0)][(y - y0)][x];
 * or_lj is the whole grid. */
void or_lj_rows(int nx, int ny, int nz, double bx, double by, double bz, int z0, int nzs,
                int y0, int nys, int natu, const double *xyz, const double *sigu,
                const double *epsu, int natv, const double *sigv, const double *epsv,
                double temper, double lam1, double *u) {
  const double cut2 = 2.0e-3 * 2.0e-3;
  double dx = bx / nx, dy = by / ny, dz = bz / nz;
  double ikbt = 1.0 / (OR_AVOGADRO * OR_BOLTZMANN * temper);
  size_t ntot = (size_t)nx * nys * nzs;
  double *su = (double *)malloc(sizeof(double) * natu * natv);
  double *eu = (double *)malloc(sizeof(double) * natu * natv);
  for (int v = 0; v < natv; ++v)
    for (int a = 0; a < natu; ++a) {
      su[a + natu * v] = (sigu[a] + sigv[v]) * 0.5 * lam1;
      eu[a + natu * v] = sqrt(epsu[a] * epsv[v]);
    }
  for (int v = 0; v < natv; ++v) {
#pragma omp parallel for schedule(static)
    for (size_t ip = 0; ip < ntot; ++ip) {
      int x = (int)(ip % nx), y = y0 + (int)((ip / nx) % nys),
          z = z0 + (int)(ip / ((size_t)nx * nys));
      double rx = rcomp(x, nx, dx), ry = rcomp(y, ny, dy), rz = rcomp(z, nz, dz);
      double acc = 0.0;
      for (int a = 0; a < natu; ++a) {
        double ddx = rx - xyz[3 * a], ddy = ry - xyz[3 * a + 1], ddz = rz - xyz[3 * a + 2];
        double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
        double sg = su[a + natu * v];
        double rs2 = r2 / (sg * sg);
        if (rs2 < cut2) rs2 = cut2;
        double irs6 = 1.0 / (rs2 * rs2 * rs2);
        acc += eu[a + natu * v] * 4.0 * irs6 * (irs6 - 1.0);
      }
      u[(size_t)v * ntot + ip] = acc * ikbt;
    }
  }
  free(su);
  free(eu);
}

void or_lj(int nx, int ny, int nz, double bx, double by, double bz, int natu,
           const double *xyz, const double *sigu, const double *epsu, int natv,
           const double *sigv, const double *epsv, double temper, double lam1, double *u) {
  or_lj_rows(nx, ny, nz, bx, by, bz, 0, nz, 0, ny, natu, xyz, sigu, epsu, natv, sigv, epsv,
             temper, lam1, u);
}

/* ubeta = E_h a0 / (k_B T) * lam2, in Angstrom (cal_Coulomb.cu:31-33) */
double or_ubeta(double temper, double lam2) {
  return OR_HARTREE * OR_BOHR / (OR_BOLTZMANN * temper) * lam2;
}

/*
 * or_coulomb: de = ubeta sum_{r>=1e-5} q/r;  fr = ubeta sum [r>=1e-5 ?
 * q (1 - e^{-r})/r : q].  cal_Coulomb.cu:40-62, 82-87.
 */
void or_coulomb_rows(int nx, int ny, int nz, double bx, double by, double bz, int z0, int nzs,
                     int y0, int nys, int natu, const double *xyz, const double *q,
                     double ubeta, double *de, double *fr) {
  double dx = bx / nx, dy = by / ny, dz = bz / nz;
  size_t ntot = (size_t)nx * nys * nzs;
#pragma omp parallel for schedule(static)
  for (size_t ip = 0; ip < ntot; ++ip) {
    int x = (int)(ip % nx), y = y0 + (int)((ip / nx) % nys),
        z = z0 + (int)(ip / ((size_t)nx * nys));
    double rx = rcomp(x, nx, dx), ry = rcomp(y, ny, dy), rz = rcomp(z, nz, dz);
    double e = 0.0, f = 0.0;
    for (int a = 0; a < natu; ++a) {
      double ddx = rx - xyz[3 * a], ddy = ry - xyz[3 * a + 1], ddz = rz - xyz[3 * a + 2];
      double ra = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
      if (ra >= 1.0e-5) {
        double qr = q[a] / ra;
        e += qr;
        f += qr * (1.0 - exp(-ra));
      } else {
        f += q[a];
      }
    }
    de[ip] = e * ubeta;
    fr[ip] = f * ubeta;
  }
}

void or_coulomb(int nx, int ny, int nz, double bx, double by, double bz, int natu,
                const double *xyz, const double *q, double ubeta, double *de, double *fr) {
  or_coulomb_rows(nx, ny, nz, bx, by, bz, 0, nz, 0, ny, natu, xyz, q, ubeta, de, fr);
}

/*
 * or_fk: fk(k) = ubeta sum_u 4 pi q_u e^{-i k.r_u} / (k^2 (k^2 + 1)) on the
 * full k-grid (complex, interleaved).  cal_Coulomb.cu:65-79, 90-95.
 */
void or_fk_rows(int nx, int ny, int nz, double bx, double by, double bz, int z0, int nzs,
                int y0, int nys, int natu, const double *xyz, const double *q, double ubeta,
                double *fk) {
  size_t ntot = (size_t)nx * nys * nzs;
#pragma omp parallel for schedule(static)
  for (size_t ip = 0; ip < ntot; ++ip) {
    int x = (int)(ip % nx), y = y0 + (int)((ip / nx) % nys),
        z = z0 + (int)(ip / ((size_t)nx * nys));
    double kx = kcomp(x, nx, bx), ky = kcomp(y, ny, by), kz = kcomp(z, nz, bz);
    double rk2 = kx * kx + ky * ky + kz * kz;
    double rk4i = 1.0 / (rk2 * (rk2 + 1.0));
    double re = 0.0, im = 0.0;
    for (int a = 0; a < natu; ++a) {
      double ruk = kx * xyz[3 * a] + ky * xyz[3 * a + 1] + kz * xyz[3 * a + 2];
      double tmp = 4.0 * M_PI * q[a] * rk4i;
      re += tmp * cos(ruk);
      im -= tmp * sin(ruk);
    }
    fk[2 * ip] = re * ubeta;
    fk[2 * ip + 1] = im * ubeta;
  }
}

void or_fk(int nx, int ny, int nz, double bx, double by, double bz, int natu,
           const double *xyz, const double *q, double ubeta, double *fk) {
  or_fk_rows(nx, ny, nz, bx, by, bz, 0, nz, 0, ny, natu, xyz, q, ubeta, fk);
}

/* ------------------------------------------------------------------------- */
/* one step: calculate.cu:4-65, calculate_cuda.cu:1-60, cal_rms.cu:14-26     */
/* ------------------------------------------------------------------------- */

/* closure: clos 0 = KH, 1 = HNC (calculate_cuda.cu:1-16) */
/* clos 0 KH, 1 HNC (calculate_cuda.cu:1-16); 10 + n: PSE-n (Kast & Kloss,
 * J. Chem. Phys. 129, 214110 (2008)), the partial sum of exp's series
 * sum_{i=0}^{n} e^i / i! where e > 0 (an extension; not in the reference) */
static inline double closure_g(int clos, double t, double u) {
  double earg = -u + t;
  if (clos == 0 && earg >= 0.0) return 1.0 + earg;
  if (clos >= 10 && earg > 0.0) {
    double term = 1.0, sum = 1.0;
    for (int i = 1; i <= clos - 10; ++i) {
      term *= earg / (double)i;
      sum += term;
    }
    return sum;
  }
  return exp(earg);
}

void or_closure(int clos, size_t n, const double *t, const double *u, double *g) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) g[i] = closure_g(clos, t[i], u[i]);
}

/*
 * or_calculate: one reference step.  chi_shell[(s*natv+v)*natv+v'] = X[v][v'](ga[s]).
 * Outputs: g (closure), R = h + 1 - g (residual), h (real part of inverse
 * transform), all [natv][N].  Returns the rms (cal_rms.cu:20-24).
 * work must hold 2*natv*N doubles (complex c per site) + 2*natv*N (complex h).
 */
double or_calculate(int nx, int ny, int nz, double bx, double by, double bz, int natv,
                    int clos, const double *qv, const double *t, const double *u,
                    const double *fr, const double *fk, const double *chi_shell,
                    const int *indga, double *g, double *R, double *h, double *work) {
  size_t N = (size_t)nx * ny * nz;
  double *cw = work;                     /* natv complex fields */
  double *hw = work + 2 * (size_t)natv * N; /* natv complex fields */
  for (int v = 0; v < natv; ++v) {
    const double *tv = t + (size_t)v * N, *uv = u + (size_t)v * N;
    double *gv = g + (size_t)v * N;
    double *cv = cw + 2 * (size_t)v * N;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) {
      double gg = closure_g(clos, tv[i], uv[i]);
      gv[i] = gg;
      cv[2 * i] = gg - 1.0 - tv[i] + qv[v] * fr[i]; /* trm1mt */
      cv[2 * i + 1] = 0.0;
    }
    or_fft3d(cv, nx, ny, nz, bx, by, bz, -1);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) { /* mqvfk */
      cv[2 * i] -= qv[v] * fk[2 * i];
      cv[2 * i + 1] -= qv[v] * fk[2 * i + 1];
    }
  }
  for (int v = 0; v < natv; ++v) { /* oz */
    double *hv = hw + 2 * (size_t)v * N;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) {
      const double *X = chi_shell + ((size_t)indga[i] * natv + v) * natv;
      double hr = 0.0, hi = 0.0;
      for (int w = 0; w < natv; ++w) {
        hr += cw[2 * ((size_t)w * N + i)] * X[w];
        hi += cw[2 * ((size_t)w * N + i) + 1] * X[w];
      }
      hv[2 * i] = hr;
      hv[2 * i + 1] = hi;
    }
  }
  double ss = 0.0;
  for (int v = 0; v < natv; ++v) {
    double *hv = hw + 2 * (size_t)v * N;
    or_fft3d(hv, nx, ny, nz, bx, by, bz, +1);
    const double *gv = g + (size_t)v * N;
    double *Rv = R + (size_t)v * N, *hr = h + (size_t)v * N;
    or_ksum part[OR_MAXTHR];
    const int nt = or_thread_count();
    memset(part, 0, sizeof(or_ksum) * nt);
#pragma omp parallel num_threads(nt)
    {
      or_ksum k;
      memset(&k, 0, sizeof(k));
#pragma omp for schedule(static)
      for (size_t i = 0; i < N; ++i) { /* tr */
        hr[i] = hv[2 * i];
        Rv[i] = hv[2 * i] + 1.0 - gv[i];
        ks_add(&k, 0, Rv[i] * Rv[i]);
      }
      part[or_thread_id()] = k;
    }
    ss += ks_total(part, nt, 0);
  }
  return sqrt(ss / ((double)N * natv));
}

/* ------------------------------------------------------------------------- */
/* AN2: anderson.cu:20-130, anderson_cuda.cu:1-67                            */
/* ------------------------------------------------------------------------- */

typedef struct {
  int count;
  double m, mp;
  int binary;
  size_t nn;   /* natv * N */
  int natv;
  size_t N;
  double *tp;  /* 2 slots of nn */
  double *rp;
  double irho[64];
  double last_x[2];
  double last_s[2];
} or_an2;

or_an2 *or_an2_create(int natv, size_t N, const double *rhov, int count, double m,
                      double mp) {
  or_an2 *a = (or_an2 *)calloc(1, sizeof(or_an2));
  a->count = count;
  a->m = m;
  a->mp = mp;
  a->binary = 0;
  a->natv = natv;
  a->N = N;
  a->nn = (size_t)natv * N;
  /* the reference leaves the slots uninitialised (cudaMalloc); zero them */
  a->tp = (double *)calloc(2 * a->nn, sizeof(double));
  a->rp = (double *)calloc(2 * a->nn, sizeof(double));
  for (int v = 0; v < natv; ++v) a->irho[v] = 1.0 / rhov[v];
  return a;
}

void or_an2_destroy(or_an2 *a) {
  free(a->tp);
  free(a->rp);
  free(a);
}

/* leg(): Gaussian elimination without pivoting (anderson.cu:105-130), n = 2 */
void or_leg2(const double *a_in, const double *c_in, double *x) {
  double a[4] = {a_in[0], a_in[1], a_in[2], a_in[3]};
  double b[2] = {c_in[0], c_in[1]};
  double p = a[0];
  a[1] /= p;
  double r = b[0] /= p;
  double q = a[2];
  a[3] -= q * a[1];
  b[1] -= q * r;
  x[1] = b[1] / a[3];
  x[0] = b[0] - a[1] * x[1];
}

/* one AN2::calculate (anderson.cu:43-78); t updated in place */
void or_an2_step(or_an2 *A, double *t, const double *R) {
  size_t nn = A->nn;
  if (A->count > 0) {
    --A->count;
    double *T = A->tp + (size_t)A->binary * nn, *Rp = A->rp + (size_t)A->binary * nn;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nn; ++i) { /* newdt0 */
      T[i] = t[i];
      Rp[i] = R[i];
      t[i] += R[i];
    }
  } else {
    /* cal_theta (anderson.cu:81-102) */
    double c[2] = {0, 0}, a[4] = {0, 0, 0, 0};
    for (int v = 0; v < A->natv; ++v) {
      size_t off = (size_t)v * A->N;
      or_ksum part[OR_MAXTHR];
      const int nt = or_thread_count();
      memset(part, 0, sizeof(or_ksum) * nt);
#pragma omp parallel num_threads(nt)
      {
        or_ksum k;
        memset(&k, 0, sizeof(k));
#pragma omp for schedule(static)
        for (size_t i = 0; i < A->N; ++i) {
          double r = R[off + i];
          double t1 = r - A->rp[off + i];
          double t2 = r - A->rp[nn + off + i];
          ks_add(&k, 0, r * t1);
          ks_add(&k, 1, r * t2);
          ks_add(&k, 2, t1 * t1);
          ks_add(&k, 3, t1 * t2);
          ks_add(&k, 4, t2 * t2);
        }
        part[or_thread_id()] = k;
      }
      const double s0 = ks_total(part, nt, 0), s1 = ks_total(part, nt, 1),
                   s2 = ks_total(part, nt, 2), s3 = ks_total(part, nt, 3),
                   s4 = ks_total(part, nt, 4);
      c[0] += s0 * A->irho[v];
      c[1] += s1 * A->irho[v];
      a[0] += s2 * A->irho[v];
      a[1] += s3 * A->irho[v];
      a[3] += s4 * A->irho[v];
    }
    a[2] = a[1];
    double x[2];
    or_leg2(a, c, x);
    double s1 = x[0] + A->mp * ((A->binary == 0) ? 0 : 1);
    double s2 = x[1] + A->mp * ((A->binary == 0) ? 1 : 0);
    A->last_x[0] = x[0];
    A->last_x[1] = x[1];
    A->last_s[0] = s1;
    A->last_s[1] = s2;
    double m = A->m;
    double *T0 = A->tp, *T1 = A->tp + nn, *R0 = A->rp, *R1 = A->rp + nn;
    double *Tb = A->tp + (size_t)A->binary * nn, *Rb = A->rp + (size_t)A->binary * nn;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nn; ++i) { /* newdt */
      double ti = t[i], ri = R[i];
      double uu = ti + s1 * (T0[i] - ti) + s2 * (T1[i] - ti);
      double vv = ti + ri + s1 * (T0[i] + R0[i] - ti - ri) + s2 * (T1[i] + R1[i] - ti - ri);
      Tb[i] = ti;
      Rb[i] = ri;
      t[i] = uu + m * (vv - uu);
    }
  }
  A->binary = (A->binary == 0) ? 1 : 0;
}

void or_an2_get(const or_an2 *A, int *count, int *binary, double *x, double *s) {
  *count = A->count;
  *binary = A->binary;
  x[0] = A->last_x[0]; x[1] = A->last_x[1];
  s[0] = A->last_s[0]; s[1] = A->last_s[1];
}

/* ------------------------------------------------------------------------- */
/* thermodynamics: cal_exchem.cu:4-155, cal_pmv.cu:5-57, cal_pressure.cu     */
/* ------------------------------------------------------------------------- */

/*
 * Per-site integrals (before the RT factor):
 *   xmu[v]        SC (closure-matched: KH for clos 0, HNC for clos 1)
 *   xmu[natv+v]   GF
 *   xmu2[v]       SC_HNC (KH only; zero otherwise)
 * each = rho_v * dv * sum_r integrand.  Also returns cuv = dv sum_v rho_v sum (h-t).
 */
void or_thermo(int natv, size_t N, double dv, int clos, const double *rhov,
               const double *h, const double *t, double *xmu, double *xmu2, double *cuv) {
  double cu = 0.0;
  for (int v = 0; v < natv; ++v) {
    const double *hv = h + (size_t)v * N, *tv = t + (size_t)v * N;
    or_ksum part[OR_MAXTHR];
    const int nt = or_thread_count();
    memset(part, 0, sizeof(or_ksum) * nt);
#pragma omp parallel num_threads(nt)
    {
      or_ksum k;
      memset(&k, 0, sizeof(k));
#pragma omp for schedule(static)
      for (size_t i = 0; i < N; ++i) {
        double hh = hv[i], tt = tv[i];
        double ehnc = hh * 0.5 * tt - (hh - tt);
        double egf = -(hh * 0.5 + 1.0) * (hh - tt);
        double ekh = (hh > 0.0) ? egf : ehnc;
        ks_add(&k, 0, (clos == 0) ? ekh : ehnc);
        ks_add(&k, 1, ehnc);
        ks_add(&k, 2, egf);
        ks_add(&k, 3, hh - tt);
      }
      part[or_thread_id()] = k;
    }
    const double sc = ks_total(part, nt, 0), sh = ks_total(part, nt, 1),
                 gf = ks_total(part, nt, 2), pm = ks_total(part, nt, 3);
    xmu[v] = sc * rhov[v] * dv;
    xmu[natv + v] = gf * rhov[v] * dv;
    xmu2[v] = (clos == 0) ? sh * rhov[v] * dv : 0.0;
    cu += pm * rhov[v];
  }
  *cuv = cu * dv;
}

/* cal_pmv.cu:20-21 (m^3/mol) */
double or_pmv(double xt, double temper, double cuv) {
  double ibeta = OR_AVOGADRO * OR_BOLTZMANN * temper;
  return xt * (1.0 - cuv) * ibeta;
}

/* cal_pressure.cu:8-19: P/kT (mol/m^3) */
double or_pressure(int natv, const double *rhov, double xt, double temper) {
  double p = 0.0;
  for (int v = 0; v < natv; ++v) p += rhov[v];
  p *= 1.0e30 / OR_AVOGADRO;
  double ibeta = OR_AVOGADRO * OR_BOLTZMANN * temper;
  return 0.5 * (p + 1.0 / (xt * ibeta));
}

/* ------------------------------------------------------------------------- */
/* the whole solve: iterate.cu:6-107                                        */
/* ------------------------------------------------------------------------- */

/*
 * or_solve: runs the reference loop from t (initial guess, updated in place).
 * Outputs g, h, R of the last step; rms history (rms_hist[maxstep]).
 * Returns the number of steps taken; *status = 0 converged, 1 diverged,
 * 2 step limit.  verbose != 0 prints the reference's progress lines.
 */
int or_solve(int nx, int ny, int nz, double bx, double by, double bz, int natv, int clos,
             const double *qv, const double *rhov, const double *u, const double *fr,
             const double *fk, const double *chi_shell, const int *indga, int count,
             double mcoef, double mp, double tol, int maxstep, double *t, double *g,
             double *h, double *R, double *rms_hist, int *status, int verbose) {
  size_t N = (size_t)nx * ny * nz;
  double *work = (double *)malloc(sizeof(double) * 4 * (size_t)natv * N);
  or_an2 *A = or_an2_create(natv, N, rhov, count, mcoef, mp);
  int conver = 0, diverge = 0, istep, last = 0;
  if (verbose) printf("relaxing 3D UV RISM:\n");
  for (istep = 1; istep <= maxstep; ++istep) {
    double rms = or_calculate(nx, ny, nz, bx, by, bz, natv, clos, qv, t, u, fr, fk,
                              chi_shell, indga, g, R, h, work);
    last = istep;
    if (rms_hist) rms_hist[istep - 1] = rms;
    diverge = !isfinite(rms);
    if (diverge) break;
    if (rms <= tol) conver = 1;
    else or_an2_step(A, t, R);
    if (verbose) { printf(" Step = %d Reside = %g\n", istep, rms); fflush(stdout); }
    if (conver) break;
  }
  *status = conver ? 0 : (diverge ? 1 : 2);
  or_an2_destroy(A);
  free(work);
  return last;
}

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
