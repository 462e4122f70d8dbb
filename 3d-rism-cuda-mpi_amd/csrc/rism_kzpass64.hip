// rism_kzpass64.hip -- the z passes (rism_kzpass.h) for the FP64 bulk type,
// and the list of FFT lengths the build supports.
#include "rism_kzpass.h"

namespace rism {

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

RISM_KZPASS_INSTANTIATE(double)

template <class Re>
static bool z_offered(int pass, int L, int v) {
  switch (L) {
#define X(n)                                                              \
  case n:                                                                 \
    if (v == 0) return true;                                              \
    if (!RISM_TUNED(n)) return false;                                     \
    if (pass == TP_ZFWD)                                                  \
      return v == 1 ? zf_plan_offered<n, Re, 1>()                         \
                    : (v == 2 ? zf_plan_offered<n, Re, 2>() : false);     \
    return v == 1 ? zi_plan_offered<n, Re, 1>() : false;
    RISM_LENGTHS(X)
#undef X
    default:
      return false;
  }
}

bool plan_offered(int pass, int L, int fp32, int v) {
  if (pass == TP_Y) return y_plan_offered(L, fp32, v);
  if (pass != TP_ZFWD && pass != TP_ZINV) return false;
  return fp32 ? z_offered<float>(pass, L, v) : z_offered<double>(pass, L, v);
}

}  // namespace rism
