// FP64 (complex128 state) batched sweep kernels.
#define TDSE_REAL double
#define TDSE_SUFFIX f64
#define TDSE_BATCH_UNIT 1
#include "kernels_impl.hpp"
