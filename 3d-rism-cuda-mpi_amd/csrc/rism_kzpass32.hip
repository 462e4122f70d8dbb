// rism_kzpass32.hip -- the z passes (rism_kzpass.h) for the FP32 bulk of the
// mixed-precision solve.
#include "rism_kzpass.h"

namespace rism {

RISM_KZPASS_INSTANTIATE(float)

}  // namespace rism
