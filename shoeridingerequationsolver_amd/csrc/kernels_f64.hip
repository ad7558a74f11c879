// FP64 (complex128 state) sweep kernels.
#define TDSE_REAL double
#define TDSE_SUFFIX f64
#include "kernels_impl.hpp"
