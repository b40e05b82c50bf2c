// fft_core.hpp -- register/LDS Stockham FFT building blocks for gfx950.
//
// One pencil of N complex128 points is owned by P = N/R threads; thread t
// holds R values in registers (R = min(RM, N), RM = 8 or 16 chosen per
// kernel: 16 halves the LDS exchanges, 8 halves the registers).  Between stages the values of a pencil are
// exchanged through LDS.  The layout invariant at the start of every stage is
//     register i of thread t  <->  element t + P*i
// which is also the layout the last stage leaves behind (natural order), so
//   * the first stage loads straight from HBM (consecutive t -> consecutive
//     addresses along a contiguous axis),
//   * the last stage stores straight to HBM in the same pattern, and
//   * a forward transform can be followed by a pointwise operation and an
//     inverse transform without any extra exchange.
// Stage radices are R repeated, then one smaller power of two.
#pragma once
#include <hip/hip_runtime.h>

namespace atus {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// multiply by DIR*i  (DIR = -1: forward, e^{-i}; DIR = +1: backward, e^{+i})
template <int DIR>
__device__ __forceinline__ double2 mul_dir_i(double2 a) {
  if constexpr (DIR < 0) return make_double2(a.y, -a.x);
  else return make_double2(-a.y, a.x);
}

// cos / sin of 2*pi*k/16, k = 0..15
__host__ __device__ constexpr double cos16(int k) {
  constexpr double c[16] = {1.0,
                            0.92387953251128675613,
                            0.70710678118654752440,
                            0.38268343236508977173,
                            0.0,
                            -0.38268343236508977173,
                            -0.70710678118654752440,
                            -0.92387953251128675613,
                            -1.0,
                            -0.92387953251128675613,
                            -0.70710678118654752440,
                            -0.38268343236508977173,
                            0.0,
                            0.38268343236508977173,
                            0.70710678118654752440,
                            0.92387953251128675613};
  return c[k & 15];
}
__host__ __device__ constexpr double sin16(int k) { return cos16(k + 12); }

// o * W_R^k with W_R = exp(DIR * 2 pi i / R), R <= 16, k < R/2
template <int R, int DIR>
__device__ __forceinline__ double2 tw_const(int k, double2 o) {
  const int e = k * (16 / R);  // angle index in sixteenths of a turn
  if (e == 0) return o;
  if (e == 4) return mul_dir_i<DIR>(o);
  double2 w = make_double2(cos16(e), DIR * sin16(e));
  return cmul(o, w);
}

// in-place R-point DFT of v[0..R-1]: v[k] <- sum_j v[j] W_R^{jk}, natural order
template <int R, int DIR>
__device__ __forceinline__ void dft(double2* v) {
  if constexpr (R == 2) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R > 2) {
    constexpr int H = R / 2;
    double2 e[H], o[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    dft<H, DIR>(e);
    dft<H, DIR>(o);
#pragma unroll
    for (int k = 0; k < H; ++k) {
      double2 t = tw_const<R, DIR>(k, o[k]);
      v[k] = cadd(e[k], t);
      v[k + H] = csub(e[k], t);
    }
  }
}

// ---------------------------------------------------------------- plan
__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
__host__ __device__ constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

template <int N, int RM = 16>
struct Plan {
  static constexpr int R = N < RM ? N : RM;           // values per thread
  static constexpr int P = N / R;                      // threads per pencil
  static constexpr int A = ilog2c(N) / ilog2c(R);      // stages of radix R
  static constexpr int Q = N / ipow(R, A);             // radix of the tail stage (1: none)
  static constexpr int S = A + (Q > 1 ? 1 : 0);        // number of stages
  template <int s> static constexpr int radix() { return s < A ? R : Q; }
  template <int s> static constexpr int span() { return ipow(R, s); }  // Ls of stage s
  // start of stage s's twiddles in the stage-major table (stage 0 has none)
  template <int s> static constexpr int tw_off() {
    if constexpr (s <= 1) return 0;
    else return tw_off<s - 1>() + (radix<s - 1>() - 1) * span<s - 1>();
  }
};

// Twiddle tables hold, behind the natural table W^j (j < N), one
// stage-major table per plan (values per thread 8, then 16): for every
// stage s >= 1 of the plan, entry tw_off<s>() + (m - 1) Ls + k is
// W^(k m N / (Ls r)), m = 1..r-1, k = 0..Ls-1 -- the values the stage reads,
// in the order the threads of a wavefront read them (consecutive k), which
// spreads them over the LDS banks and over fewer cache lines.  The entries
// of a plan sum to N - 1.
template <int RM>
__host__ __device__ constexpr int tw_table_slot() {
  static_assert(RM == 8 || RM == 16, "plans of 8 or 16 values per thread");
  return RM == 8 ? 1 : 2;
}
constexpr int TW_TABLES = 3;  // natural, then the plans of 8 and 16
template <int N, int RM>
__device__ __forceinline__ const double2* fft_tw(const double2* tw) {
  return tw + (size_t)N * tw_table_slot<RM>();
}

// LDS position of pencil element j: XOR swizzle inside rows of 16 elements
// (16 B each) so the stride-16 scatter of the first stage hits distinct banks
__device__ __forceinline__ int lds_pos(int j) { return (j & ~15) | ((j ^ (j >> 4)) & 15); }

template <bool WAVE_ONLY>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE_ONLY) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// element i of a twiddle table in LDS or (address-space qualified) global
// memory
typedef double __attribute__((ext_vector_type(2))) tw_vec2;
template <class TW>
__device__ __forceinline__ double2 load_tw(TW p, int i) {
  if constexpr (__is_same(TW, const double2*)) {
    return p[i];
  } else {
    typedef const __attribute__((address_space(1))) tw_vec2* gvec;
    const tw_vec2 w = ((gvec)p)[i];
    return make_double2(w.x, w.y);
  }
}

// compute of stage s (twiddle + r-point DFTs); tw: the plan's stage-major
// twiddle table (fft_tw)
template <int N, int RM, int s, int DIR, class TW>
__device__ __forceinline__ void stage_compute(double2 (&v)[(Plan<N, RM>::R)], int t, TW tw) {
  using PL = Plan<N, RM>;
  constexpr int R = PL::R, P = PL::P;
  constexpr int r = PL::template radix<s>();
  constexpr int Ls = PL::template span<s>();
  constexpr int Qn = R / r;  // butterflies per thread
#pragma unroll
  for (int q = 0; q < Qn; ++q) {
    double2 x[r];
#pragma unroll
    for (int m = 0; m < r; ++m) x[m] = v[q + m * Qn];
    if constexpr (Ls > 1) {
      const int b = t + q * P;
      const int k = b & (Ls - 1);
      const TW tws = tw + PL::template tw_off<s>() + k;
#pragma unroll
      for (int m = 1; m < r; ++m) {
        double2 w = load_tw(tws, (m - 1) * Ls);
        x[m] = (DIR < 0) ? cmul(x[m], w) : cmulc(x[m], w);
      }
    }
    dft<r, DIR>(x);
#pragma unroll
    for (int m = 0; m < r; ++m) v[q + m * Qn] = x[m];
  }
}

// scatter the outputs of stage s to their Stockham positions and gather the
// next stage's inputs (element t + P*i into register i)
template <int N, int RM, int s, bool WAVE_ONLY>
__device__ __forceinline__ void stage_exchange(double2 (&v)[(Plan<N, RM>::R)], int t, double2* lds) {
  using PL = Plan<N, RM>;
  constexpr int R = PL::R, P = PL::P;
  constexpr int r = PL::template radix<s>();
  constexpr int Ls = PL::template span<s>();
  constexpr int Qn = R / r;
  lds_sync<WAVE_ONLY>();  // previous readers of this region are done
#pragma unroll
  for (int q = 0; q < Qn; ++q) {
    const int b = t + q * P;
    const int k = b & (Ls - 1);
    const int base = (b - k) * r + k;
#pragma unroll
    for (int m = 0; m < r; ++m) lds[lds_pos(base + m * Ls)] = v[q + m * Qn];
  }
  lds_sync<WAVE_ONLY>();
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = lds[lds_pos(t + P * i)];
}

template <int N, int RM, int s, int DIR, bool WAVE_ONLY, class TW>
__device__ __forceinline__ void run_stages(double2 (&v)[(Plan<N, RM>::R)], int t, double2* lds, TW tw) {
  stage_compute<N, RM, s, DIR>(v, t, tw);
  if constexpr (s + 1 < Plan<N, RM>::S) {
    stage_exchange<N, RM, s, WAVE_ONLY>(v, t, lds);
    run_stages<N, RM, s + 1, DIR, WAVE_ONLY, TW>(v, t, lds, tw);
  }
}

// Twiddles from global memory (N > 1024) at up to 8 values per thread: a
// stage's twiddles (at most R - 1 per thread) are loaded before the
// previous stage's exchange, so their latency hides behind its barriers
// instead of following them.  Same values, same arithmetic.
template <int N, int RM, int s, class TW>
__device__ __forceinline__ void load_stage_tw(double2 (&w)[8], int t, TW tw) {
  using PL = Plan<N, RM>;
  constexpr int R = PL::R, P = PL::P;
  constexpr int r = PL::template radix<s>();
  constexpr int Ls = PL::template span<s>();
  constexpr int Qn = R / r;
  static_assert(Qn * (r - 1) <= 8, "prefetched twiddles of up to 8 values per thread");
  if constexpr (Ls > 1) {
#pragma unroll
    for (int q = 0; q < Qn; ++q) {
      const int k = (t + q * P) & (Ls - 1);
      const TW tws = tw + PL::template tw_off<s>() + k;
#pragma unroll
      for (int m = 1; m < r; ++m) w[q * (r - 1) + m - 1] = load_tw(tws, (m - 1) * Ls);
    }
  }
}

template <int N, int RM, int s, int DIR>
__device__ __forceinline__ void stage_compute_w(double2 (&v)[(Plan<N, RM>::R)], const double2 (&w)[8]) {
  using PL = Plan<N, RM>;
  constexpr int R = PL::R;
  constexpr int r = PL::template radix<s>();
  constexpr int Ls = PL::template span<s>();
  constexpr int Qn = R / r;
#pragma unroll
  for (int q = 0; q < Qn; ++q) {
    double2 x[r];
#pragma unroll
    for (int m = 0; m < r; ++m) x[m] = v[q + m * Qn];
    if constexpr (Ls > 1) {
#pragma unroll
      for (int m = 1; m < r; ++m) {
        const double2 wm = w[q * (r - 1) + m - 1];
        x[m] = (DIR < 0) ? cmul(x[m], wm) : cmulc(x[m], wm);
      }
    }
    dft<r, DIR>(x);
#pragma unroll
    for (int m = 0; m < r; ++m) v[q + m * Qn] = x[m];
  }
}

template <int N, int RM, int s, int DIR, bool WAVE_ONLY, class TW>
__device__ __forceinline__ void run_stages_pf(double2 (&v)[(Plan<N, RM>::R)], int t, double2* lds, TW tw,
                                              const double2 (&w)[8]) {
  stage_compute_w<N, RM, s, DIR>(v, w);
  if constexpr (s + 1 < Plan<N, RM>::S) {
    double2 wn[8];
    load_stage_tw<N, RM, s + 1>(wn, t, tw);
    stage_exchange<N, RM, s, WAVE_ONLY>(v, t, lds);
    run_stages_pf<N, RM, s + 1, DIR, WAVE_ONLY, TW>(v, t, lds, tw, wn);
  }
}

// full unnormalised N-point DFT of the pencil held in registers (layout
// t + P*i in and out).  lds: this pencil's region of at least N elements;
// tw: the plan's stage-major twiddle table (fft_tw), a pointer to LDS or
// global memory.
template <int N, int RM, int DIR, bool WAVE_ONLY, class TW>
__device__ __forceinline__ void fft_pencil(double2 (&v)[(Plan<N, RM>::R)], int t, double2* lds, TW tw) {
  if constexpr (N > 1) {
    if constexpr (!__is_same(TW, const double2*) && Plan<N, RM>::R <= 8) {
      double2 w0[8];
      load_stage_tw<N, RM, 0>(w0, t, tw);  // stage 0 reads no twiddles
      run_stages_pf<N, RM, 0, DIR, WAVE_ONLY, TW>(v, t, lds, tw, w0);
    } else {
      run_stages<N, RM, 0, DIR, WAVE_ONLY, TW>(v, t, lds, tw);
    }
  }
}

}  // namespace atus
