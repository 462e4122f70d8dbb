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

}  // namespace rism
