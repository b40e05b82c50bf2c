// mlp.hpp -- launchers of the warm-start network kernels (k_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace atus {

// floats of workspace mlp_gemm needs for an M x N x K product
size_t mlp_gemm_work(int M, int N, int K);
// C (M x N, row-major) = act(A B + bias); A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn]; operands rounded to bf16 (split: hi + lo
// bf16 pairs, three products), fp32 accumulation; bias may be null; act 0:
// none, 1: tanh
hipError_t mlp_gemm(const float* A, long long sam, long long sak, const float* B, long long sbk, long long sbn,
                    float* C, int M, int N, int K, const float* bias, int act, float* work, size_t work_floats,
                    hipStream_t s, bool split);
hipError_t mlp_tanh_back(const float* dX, const float* X, float* dZ, long long n, hipStream_t s);
hipError_t mlp_colsum(const float* dZ, int M, int N, float* db, hipStream_t s);
// dY = 2 (Y - T) / n; part[0..nblocks) partial sums of (Y - T)^2
hipError_t mlp_mse(const float* Y, const float* T, long long n, float* dY, float* part, int nblocks, hipStream_t s);
hipError_t mlp_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                    float eps, float c1, float c2, hipStream_t s);
hipError_t mlp_gather(const float* src, const int* idx, int rows, long long w, float* dst, hipStream_t s);
// psi[n0][n1][n2] = sum E0[i0][j0] E1[i1][j1] E2[i2][j2] y[j0][j1][j2] over the
// 32^3 values y, through t1 (n0 x 32 x 32) and t2 (n0 x n1 x 32); E_a: n_a x 32
hipError_t mlp_interp32(const float* y, const double* E0, const double* E1, const double* E2, double* t1, double* t2,
                        double2* psi, const int n[3], hipStream_t s);

}  // namespace atus
