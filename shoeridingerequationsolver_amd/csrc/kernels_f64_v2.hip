// F64 sweep kernels, variant 2 (autotuner candidate).
#define TDSE_REAL double
#define TDSE_SUFFIX f64
#define TDSE_VARIANT_UNIT 2
#include "kernels_impl.hpp"
