// F32 sweep kernels, variant 1 (autotuner candidate).
#define TDSE_REAL float
#define TDSE_SUFFIX f32
#define TDSE_VARIANT_UNIT 1
#include "kernels_impl.hpp"
