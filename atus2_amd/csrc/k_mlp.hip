// k_mlp.hip -- kernels of the warm-start network: BF16 MFMA matrix
// products with FP32 accumulation (v_mfma_f32_16x16x32_bf16), their
// epilogues, the loss, the optimiser, and the separable interpolation of
// the network's 32^3 output onto a solver grid.
#include "mlp.hpp"

namespace atus {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// P[z][m][n] = sum over k in [z*kchunk, (z+1)*kchunk) of A(m,k) B(k,n), with
// A(m,k) = A[m*sam + k*sak] and B(k,n) = B[k*sbk + n*sbn] rounded to bf16
// (SPLIT false), or split as a = hi + lo, hi = bf16(a), lo = bf16(a - hi),
// with a b ~ hi_a hi_b + hi_a lo_b + lo_a hi_b (SPLIT true: three products
// per slice, about 16 significant bits per operand).
// One wavefront per 16x16 tile; lane l holds A rows l%16 and B columns
// l%16 at k = 8*(l/16) + 0..7 of each 32-wide slice, and receives rows
// 4*(l/16) + 0..3 of column l%16 of the tile.
template <bool SPLIT>
__global__ __launch_bounds__(64) void k_gemm_bf16(const float* __restrict__ A, long long sam, long long sak,
                                                  const float* __restrict__ B, long long sbk, long long sbn,
                                                  float* __restrict__ P, int M, int N, int K, int kchunk) {
  const int lane = threadIdx.x;
  const int n0 = blockIdx.x * 16, m0 = blockIdx.y * 16;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int i = lane & 15, grp = lane >> 4;
  const int m = m0 + i, n = n0 + i;
  const bool mok = m < M, nok = n < N;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 32) {
    bf16x8 a, b, al, bl;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + 8 * grp + j;
      const bool kin = k < ke;
      const float fa = (kin && mok) ? A[(long long)m * sam + (long long)k * sak] : 0.f;
      const float fb = (kin && nok) ? B[(long long)k * sbk + (long long)n * sbn] : 0.f;
      a[j] = (__bf16)fa;
      b[j] = (__bf16)fb;
      if constexpr (SPLIT) {
        al[j] = (__bf16)(fa - (float)a[j]);
        bl[j] = (__bf16)(fb - (float)b[j]);
      }
    }
    if constexpr (SPLIT) {
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl, acc, 0, 0, 0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  float* out = P + (long long)blockIdx.z * M * N;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = m0 + 4 * grp + r, col = n0 + i;
    if (row < M && col < N) out[(long long)row * N + col] = acc[r];
  }
}

// C[m][n] = act(sum_z P[z][m][n] + bias[n]); act 1: tanh.  Split order fixed.
__global__ __launch_bounds__(256) void k_gemm_sum(const float* __restrict__ P, int splits, int M, int N,
                                                  const float* __restrict__ bias, int act, float* __restrict__ C) {
  const long long MN = (long long)M * N;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < MN; e += (long long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += P[z * MN + e];
    if (bias) s += bias[e % N];
    C[e] = act == 1 ? tanhf(s) : s;
  }
}

// dZ = dX (1 - X^2), X = tanh(Z)
__global__ __launch_bounds__(256) void k_tanh_back(const float* __restrict__ dX, const float* __restrict__ X,
                                                   float* __restrict__ dZ, long long n) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    dZ[e] = dX[e] * (1.f - X[e] * X[e]);
}

// db[n] = sum_m dZ[m][n], rows in order
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dZ, int M, int N, float* __restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dZ[(long long)m * N + n];
  db[n] = s;
}

// mean squared error over n values: dY = 2 (Y - T) / n; per-block partial
// sums of (Y - T)^2 (fixed order within and across blocks)
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T, long long n,
                                             float* __restrict__ dY, float* __restrict__ part) {
  __shared__ float sh[256];
  float acc = 0.f;
  const float sc = 2.f / (float)n;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const float d = Y[e] - T[e];
    dY[e] = sc * d;
    acc += d * d;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam (Kingma & Ba 2015) with bias corrections c1 = 1 - beta1^t, c2 = 1 - beta2^t
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, long long n, float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const float ge = g[e];
    const float me = b1 * m[e] + (1.f - b1) * ge;
    const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    p[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
  }
}

// rows idx[r] of src (n x w) into dst (rows x w)
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ src, const int* __restrict__ idx, int rows,
                                                long long w, float* __restrict__ dst) {
  const long long tot = (long long)rows * w;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x)
    dst[e] = src[(long long)idx[e / w] * w + e % w];
}

// one axis of a separable interpolation: out[p][i][q] = sum_j E[i][j] in[p][j][q]
// (p < outer, i < nout, j < nin, q < inner); `in` is fp32 (the network's
// output) or fp64, out is fp64 or, for the last axis, the complex field
template <typename Tin, bool TO_FIELD>
__global__ __launch_bounds__(256) void k_interp_axis(const Tin* __restrict__ in, const double* __restrict__ E,
                                                     void* __restrict__ out, long long outer, int nin, int nout,
                                                     long long inner) {
  const long long tot = outer * nout * inner;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const long long q = e % inner;
    const long long r = e / inner;
    const int i = (int)(r % nout);
    const long long p = r / nout;
    const double* Ei = E + (long long)i * nin;
    const Tin* src = in + p * nin * inner + q;
    double acc = 0.0;
    for (int j = 0; j < nin; ++j) acc += Ei[j] * (double)src[(long long)j * inner];
    if constexpr (TO_FIELD)
      static_cast<double2*>(out)[e] = make_double2(acc, 0.0);
    else
      static_cast<double*>(out)[e] = acc;
  }
}

}  // namespace atus

namespace atus {

static unsigned ew_grid(long long n) {
  long long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// split of K into 32-wide slices so that tiles x splits is about 2048
static void split_plan(int M, int N, int K, int* splits, int* kchunk) {
  const long long tiles = (long long)((M + 15) / 16) * ((N + 15) / 16);
  const int slices = (K + 31) / 32;
  long long want = (2048 + tiles - 1) / tiles;
  if (want > slices) want = slices;
  if (want < 1) want = 1;
  const int per = (int)((slices + want - 1) / want);
  *kchunk = per * 32;
  *splits = (slices + per - 1) / per;
}

size_t mlp_gemm_work(int M, int N, int K) {
  int splits, kchunk;
  split_plan(M, N, K, &splits, &kchunk);
  return (size_t)splits * M * N;
}

hipError_t mlp_gemm(const float* A, long long sam, long long sak, const float* B, long long sbk, long long sbn,
                    float* C, int M, int N, int K, const float* bias, int act, float* work, size_t work_floats,
                    hipStream_t s, bool split) {
  if (M < 1 || N < 1 || K < 1) return hipErrorInvalidValue;
  int splits, kchunk;
  split_plan(M, N, K, &splits, &kchunk);
  if ((size_t)splits * M * N > work_floats) return hipErrorInvalidValue;
  dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16), (unsigned)splits);
  if (split)
    hipLaunchKernelGGL(k_gemm_bf16<true>, grid, dim3(64), 0, s, A, sam, sak, B, sbk, sbn, work, M, N, K, kchunk);
  else
    hipLaunchKernelGGL(k_gemm_bf16<false>, grid, dim3(64), 0, s, A, sam, sak, B, sbk, sbn, work, M, N, K, kchunk);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gemm_sum, dim3(ew_grid((long long)M * N)), dim3(256), 0, s, work, splits, M, N, bias, act, C);
  return hipGetLastError();
}

hipError_t mlp_tanh_back(const float* dX, const float* X, float* dZ, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_tanh_back, dim3(ew_grid(n)), dim3(256), 0, s, dX, X, dZ, n);
  return hipGetLastError();
}

hipError_t mlp_colsum(const float* dZ, int M, int N, float* db, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, dZ, M, N, db);
  return hipGetLastError();
}

hipError_t mlp_mse(const float* Y, const float* T, long long n, float* dY, float* part, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3((unsigned)nblocks), dim3(256), 0, s, Y, T, n, dY, part);
  return hipGetLastError();
}

hipError_t mlp_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                    float eps, float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(ew_grid(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t mlp_gather(const float* src, const int* idx, int rows, long long w, float* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3(ew_grid((long long)rows * w)), dim3(256), 0, s, src, idx, rows, w, dst);
  return hipGetLastError();
}

hipError_t mlp_interp32(const float* y, const double* E0, const double* E1, const double* E2, double* t1, double* t2,
                        double2* psi, const int n[3], hipStream_t s) {
  // [32][32][32] -> [n0][32][32] -> [n0][n1][32] -> [n0][n1][n2]
  hipLaunchKernelGGL((k_interp_axis<float, false>), dim3(ew_grid((long long)n[0] * 32 * 32)), dim3(256), 0, s, y, E0,
                     (void*)t1, 1LL, 32, n[0], 32LL * 32);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_interp_axis<double, false>), dim3(ew_grid((long long)n[0] * n[1] * 32)), dim3(256), 0, s, t1,
                     E1, (void*)t2, (long long)n[0], 32, n[1], 32LL);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_interp_axis<double, true>), dim3(ew_grid((long long)n[0] * n[1] * n[2])), dim3(256), 0, s, t2,
                     E2, (void*)psi, (long long)n[0] * n[1], 32, n[2], 1LL);
  return hipGetLastError();
}

}  // namespace atus
