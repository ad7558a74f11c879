// F32 sweep kernels, variant 3 (autotuner candidate).
#define TDSE_REAL float
#define TDSE_SUFFIX f32
#define TDSE_VARIANT_UNIT 3
#include "kernels_impl.hpp"
