// mlp.hpp -- launchers of the warm-start network kernels (k_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace atus {

// floats of workspace mlp_gemm needs for an M x N x K product
size_t mlp_gemm_work(int M, int N, int K);
// C (M x N, row-major) = act(A B + bias); A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn]; operands rounded to bf16, fp32 accumulation;
// bias may be null; act 0: none, 1: tanh
hipError_t mlp_gemm(const float* A, long long sam, long long sak, const float* B, long long sbk, long long sbn,
                    float* C, int M, int N, int K, const float* bias, int act, float* work, size_t work_floats,
                    hipStream_t s);
hipError_t mlp_tanh_back(const float* dX, const float* X, float* dZ, long long n, hipStream_t s);
hipError_t mlp_colsum(const float* dZ, int M, int N, float* db, hipStream_t s);
// dY = 2 (Y - T) / n; part[0..nblocks) partial sums of (Y - T)^2
hipError_t mlp_mse(const float* Y, const float* T, long long n, float* dY, float* part, int nblocks, hipStream_t s);
hipError_t mlp_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                    float eps, float c1, float c2, hipStream_t s);
hipError_t mlp_gather(const float* src, const int* idx, int rows, long long w, float* dst, hipStream_t s);
hipError_t mlp_upsample32(const float* y, double2* psi, const Grid& g, const double half_width[3], hipStream_t s);

}  // namespace atus
