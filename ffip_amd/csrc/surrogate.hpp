// Host entry points of the surrogate kernels (surrogate.hip).  Device pointers
// and the caller's HIP stream (as void*); no allocation.
#pragma once

#include <cstdint>

namespace ffip_amd {

// C (+)= op(A) op(B), A and B bf16, C FP32 (see k_gemm); ksplit > 1 splits K
// over workgroups, their partial tiles added atomically or, with a workspace of
// at least ksplit * M * N floats, in a fixed order.
void sg_gemm(int ta, int tb, int64_t M, int64_t N, int64_t K, const void* A, int64_t lda, const void* B, int64_t ldb,
             float* C, int64_t ldc, int accumulate, int ksplit, void* stream, float* work = nullptr,
             int64_t work_elems = 0);
// geom = {batch, C, in_z, in_y, in_x, s_z, s_y, s_x, d_z, d_y, d_x}
void sg_im2col(const int64_t geom[11], const void* X, void* col, void* stream);
void sg_col2im(const int64_t geom[11], const float* dcol, float* dX, void* stream);
void sg_bias_act(int64_t rows, int64_t n, const float* Y, const float* bias, int relu, void* out, float* zkeep,
                 void* stream);
void sg_act_back(int64_t total, const float* dA, const float* Z, int relu, void* dZb, float* dZf, void* stream);
void sg_colsum(int64_t rows, int64_t n, const float* dZ, float* db, int accumulate, void* stream);
void sg_to_bf16(int64_t n, const float* x, void* y, void* stream);
// h = {batch, nz, ny_in, nx_in, C, y0, x0, ny, nx}
void sg_head(const int64_t h[9], const void* A, float* out, void* stream);
void sg_head_back(const int64_t h[9], const float* dout, float* dA, void* stream);
// hp = {lr, beta1, beta2, eps}; t = step number from 1
void sg_adam(int64_t n, float* w, const float* g, float* m, float* v, const float hp[4], int64_t t, void* wb,
             void* stream);

}  // namespace ffip_amd
