// rism_noalloc.hip -- the half-shifted 3D transform on caller-owned device
// buffers without any allocation (rism_k_fft_noalloc in rism3d_c.h).
//
// The kernels of the kernel-level transform (rism_k_fft_r2c / c2r) need
// device memory beyond their input and output: the twiddle and phase
// tables of the three axis passes and the residual pass's per-workgroup
// partial sums.  Here those live in one file-scope device array, `scratch`,
// of kScratchBytes per GPU (held until the process ends); the tables are
// computed into it on the device by the first kernel of each call, and the
// partial sums are written, never accumulated, so nothing needs clearing.
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (other streams, other host threads) race and must be serialised by
// the caller.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/rism3d_c.h"
#include "rism_fft.h"
#include "rism_internal.h"
#include "rism_kernels.h"

namespace {

constexpr size_t kScratchBytes = 2u << 20;
constexpr size_t kTableBytes = 128u << 10;  // 2 nx + 2 ny + 3 nz/2 complex <= 7 * 1024 * 16 B
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};

// the address of this GPU's copy of `scratch`, looked up once per device
// with that device current (host code must not use the symbol itself)
int scratch_ptr(unsigned char **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return rism::set_error(RISM_ERR_HIP, "rism_k_fft_noalloc: no current device");
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return rism::set_error(RISM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                               hipGetErrorString(e));
    g_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_ptr[dev];
  return RISM_OK;
}

#define NA_CHECK(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return rism::set_error(RISM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" int rism_k_fft_noalloc(int dir, int nx, int ny, int nz, const double box[3],
                                  double *real, double *spec, void *stream) {
  if (!box || !real || !spec) return rism::set_error(RISM_ERR_ARG, "null argument");
  if (dir != -1 && dir != 1) return rism::set_error(RISM_ERR_ARG, "dir must be -1 or +1");
  if (nx < 8 || ny < 8 || nz < 16 || (nx & 1) || (ny & 1) || (nz & 1) ||
      !rism::fft_length_supported(nx) || !rism::fft_length_supported(ny) ||
      !rism::fft_length_supported(nz / 2))
    return rism::set_error(RISM_ERR_ARG, "unsupported grid (nx, ny, nz/2 must be built FFT "
                                         "lengths, nx, ny >= 8, nz >= 16, all even)");
  if (!(box[0] > 0 && box[1] > 0 && box[2] > 0))
    return rism::set_error(RISM_ERR_ARG, "box lengths must be positive");
  rism::Params P{};
  P.nx = nx;
  P.ny = ny;
  P.nz = nz;
  P.natv = 1;
  const size_t N = (size_t)nx * ny * nz;
  P.dv = box[0] * box[1] * box[2] / (double)N;
  P.inv_scale = 1.0 / (P.dv * (double)nx * (double)ny * (double)(nz / 2));
  for (int d = 0; d < 3; ++d) {
    P.dr[d] = box[d] / (d == 0 ? nx : (d == 1 ? ny : nz));
    P.kfac[d] = 2.0 * M_PI / box[d];
  }
  // scratch: the tables, then the partial sums of the residual pass (6
  // doubles per workgroup of its default plan)
  const size_t ntab = 2 * (size_t)nx + 2 * (size_t)ny + 3 * (size_t)(nz / 2);
  const size_t npart = rism::zinv_num_blocks<double>(P, 0) * 6;
  if (ntab * sizeof(rism::cplx) > kTableBytes ||
      kTableBytes + npart * sizeof(double) > kScratchBytes)
    return rism::set_error(RISM_ERR_ARG,
                           "grid too large for the no-allocation transform: its partial sums "
                           "need " + std::to_string(npart * sizeof(double)) + " of " +
                               std::to_string(kScratchBytes - kTableBytes) + " bytes");
  unsigned char *base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc) return rc;
  rism::cplx *tw = reinterpret_cast<rism::cplx *>(base);
  double *partial = reinterpret_cast<double *>(base + kTableBytes);
  const rism::cplx *twx = tw, *twy = twx + nx, *twz = twy + ny, *prex = twz + nz / 2,
                   *prey = prex + nx, *zpre = prey + ny, *zpost = zpre + nz / 2;
  hipStream_t st = static_cast<hipStream_t>(stream);
  rism::cplx *sp = reinterpret_cast<rism::cplx *>(spec);
  NA_CHECK(rism::run_fft_tables(nx, ny, nz, tw, st));
  if (dir < 0) {  // the steps of rism_k_fft_r2c
    P.mode = 1;
    NA_CHECK(rism::run_zfwd(P, (const double *)real, (const double *)real,
                            (const double *)real, sp, twz, zpre, zpost, st));
    NA_CHECK(rism::run_y(P, sp, -1, twy, prey, st));
    NA_CHECK(rism::run_xoz(P, sp, sp, twx, prex, nullptr, nullptr, (const double *)nullptr, st));
    NA_CHECK(rism::run_apply_phase(sp, P, 0, st));
  } else {  // the steps of rism_k_fft_c2r; the spectrum is the workspace
    P.mode = 2;
    NA_CHECK(rism::run_apply_phase(sp, P, 1, st));
    NA_CHECK(rism::run_xoz(P, sp, sp, twx, prex, nullptr, nullptr, (const double *)nullptr, st));
    NA_CHECK(rism::run_y(P, sp, +1, twy, prey, st));
    NA_CHECK(rism::run_zinv(P, sp, real, real, real, real, real, (const double *)nullptr,
                            partial, twz, zpre, zpost, 0, st));
  }
  return RISM_OK;
}
