// FP32 (complex64 state) sweep kernels.
#define TDSE_REAL float
#define TDSE_SUFFIX f32
#include "kernels_impl.hpp"
