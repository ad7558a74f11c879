// FP32 (complex64 state) batched sweep kernels.
#define TDSE_REAL float
#define TDSE_SUFFIX f32
#define TDSE_BATCH_UNIT 1
#include "kernels_impl.hpp"
