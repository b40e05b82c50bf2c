// k_mlp.hip -- kernels of the warm-start network: BF16 MFMA matrix
// products with FP32 accumulation (v_mfma_f32_16x16x32_bf16), their
// epilogues, the loss, the optimiser, and the map of the network's 32^3
// output onto a solver grid.
#include "mlp.hpp"

namespace atus {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// P[z][m][n] = sum over k in [z*kchunk, (z+1)*kchunk) of A(m,k) B(k,n), with
// A(m,k) = A[m*sam + k*sak] and B(k,n) = B[k*sbk + n*sbn] rounded to bf16.
// One wavefront per 16x16 tile; lane l holds A rows l%16 and B columns
// l%16 at k = 8*(l/16) + 0..7 of each 32-wide slice, and receives rows
// 4*(l/16) + 0..3 of column l%16 of the tile.
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
    bf16x8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + 8 * grp + j;
      const bool kin = k < ke;
      a[j] = (__bf16)((kin && mok) ? A[(long long)m * sam + (long long)k * sak] : 0.f);
      b[j] = (__bf16)((kin && nok) ? B[(long long)k * sbk + (long long)n * sbn] : 0.f);
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

// psi(x) = trilinear interpolation of the 32^3 values y at grid
// coordinates u_a = x_a * 16 / s_a + 16 (the 32^3 grid spans [-s_a, s_a)
// with x_j = (j - 16) s_a / 16), 0 outside it; imaginary part 0
__global__ __launch_bounds__(256) void k_upsample32(const float* __restrict__ y, double2* __restrict__ psi, Grid g,
                                                    double s0, double s1, double s2) {
  const double s[3] = {s0, s1, s2};
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < g.npts;
       l += (long long)gridDim.x * blockDim.x) {
    long long rem = l;
    int idx[3] = {0, 0, 0};
    for (int a = g.dim - 1; a >= 0; --a) {
      idx[a] = (int)(rem % g.n[a]);
      rem /= g.n[a];
    }
    int j0[3];
    double f[3];
    bool inside = true;
    for (int a = 0; a < 3; ++a) {
      if (a >= g.dim) {
        j0[a] = 16;
        f[a] = 0.0;
        continue;
      }
      const double x = (double)(idx[a] - g.n[a] / 2) * g.dx[a];
      const double u = x * 16.0 / s[a] + 16.0;
      const double fl = floor(u);
      j0[a] = (int)fl;
      f[a] = u - fl;
      if (j0[a] < 0 || j0[a] > 31) inside = false;
    }
    double val = 0.0;
    if (inside) {
      for (int c = 0; c < 8; ++c) {
        double w = 1.0;
        int jj[3];
        for (int a = 0; a < 3; ++a) {
          const int bit = (c >> a) & 1;
          jj[a] = j0[a] + bit;
          w *= bit ? f[a] : 1.0 - f[a];
        }
        if (w == 0.0) continue;
        if (jj[0] > 31 || jj[1] > 31 || jj[2] > 31) continue;
        val += w * (double)y[(jj[0] * 32 + jj[1]) * 32 + jj[2]];
      }
    }
    psi[l] = make_double2(val, 0.0);
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
                    hipStream_t s) {
  if (M < 1 || N < 1 || K < 1) return hipErrorInvalidValue;
  int splits, kchunk;
  split_plan(M, N, K, &splits, &kchunk);
  if ((size_t)splits * M * N > work_floats) return hipErrorInvalidValue;
  dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16), (unsigned)splits);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(64), 0, s, A, sam, sak, B, sbk, sbn, work, M, N, K, kchunk);
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

hipError_t mlp_upsample32(const float* y, double2* psi, const Grid& g, const double sc[3], hipStream_t s) {
  hipLaunchKernelGGL(k_upsample32, dim3(ew_grid(g.npts)), dim3(256), 0, s, y, psi, g, sc[0], sc[1], sc[2]);
  return hipGetLastError();
}

}  // namespace atus
