// F64 sweep kernels, variant 1 (autotuner candidate).
#define TDSE_REAL double
#define TDSE_SUFFIX f64
#define TDSE_VARIANT_UNIT 1
#include "kernels_impl.hpp"
