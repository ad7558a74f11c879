// F32 sweep kernels, variant 2 (autotuner candidate).
#define TDSE_REAL float
#define TDSE_SUFFIX f32
#define TDSE_VARIANT_UNIT 2
#include "kernels_impl.hpp"
