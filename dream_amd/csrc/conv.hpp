// conv.hpp -- batched linear convolution (fftconv_p, sum_fftconv).
#pragma once

#include <cstdint>

#include "common.hpp"

namespace dream {

// ConvMode of the reference (include/dream.h:91-95): '=', '+=', '-='.
enum { CONV_EQU = 0, CONV_SUM = 1, CONV_NEG = 2 };

// Y[:, n] (mode) sum_{l < L} A_l[:, n] (*) B_l[:, n] for n < N, all device
// pointers, column-major.  A_l[:, n] = A + l*a_layer + n*a_col (MA samples),
// B_l[:, n] = B + l*b_layer + n*b_col (MB samples; b_col = 0 shares one
// filter across the columns), Y[:, n] = Y + n*(MA+MB-1).
template <typename T>
void launch_conv_sum(T *Y, const T *A, int64_t MA, int64_t N, int64_t L, int64_t a_col,
                     int64_t a_layer, const T *B, int64_t MB, int64_t b_col, int64_t b_layer,
                     int mode, hipStream_t s);

} // namespace dream
