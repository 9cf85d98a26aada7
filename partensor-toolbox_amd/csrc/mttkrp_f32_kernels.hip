// FP32-stream contractions on matrix cores (placeholder until the bf16x3 / FP32
// MFMA kernels land).
#include "ptn_internal.hpp"
