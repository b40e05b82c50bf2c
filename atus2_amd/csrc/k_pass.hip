// k_pass.hip -- single-direction FFT passes along one axis.
#include "launch.hpp"

namespace atus {

bool supported_len(int N) { return N >= 2 && N <= 4096 && (N & (N - 1)) == 0; }

PassCfg make_pass_cfg(int N, int rm, long long inner, long long outer, int ycount, size_t max_lds, int ntables,
                      int max_threads) {
  PassCfg c{};
  const int P = threads_per_pencil(N, rm);
  const size_t pencil_bytes = (size_t)(N + 1) * sizeof(double2);  // lds_stride<N>()
  const size_t table_bytes = N <= 1024 ? (size_t)ntables * N * sizeof(double2) : 0;
  long long T = max_threads / P;
  if (T < 1) T = 1;
  while (T > 1 && (size_t)T * pencil_bytes + table_bytes > max_lds) T >>= 1;
  if (inner == 1) {
    long long tiles = (outer + T - 1) / T;
    c.g.inner = 1;
    c.g.outer = outer;
    c.g.col_tiles = 1;
    c.g.ntiles = tiles;
    c.g.T = (int)T;
  } else {
    if (T > inner) T = inner;
    c.g.inner = inner;
    c.g.outer = outer;
    c.g.col_tiles = (inner + T - 1) / T;
    c.g.ntiles = outer * c.g.col_tiles;
    c.g.T = (int)T;
  }
  // more threads than the kernel's launch bounds cannot be launched; such
  // a configuration is returned with an empty grid and rejected
  if (T * P > max_threads) T = 0;
  c.block = dim3((unsigned)(T * P));
  c.grid = dim3((unsigned)c.g.ntiles, (unsigned)ycount);
  c.lds = (size_t)T * pencil_bytes + table_bytes;
  if (c.lds > max_lds) c.block = dim3(0);  // rejected by the launchers
  return c;
}

template <int N, int RM>
static hipError_t pass_nr(bool contig, int dir, const double2* src, double2* dst, long long cs, int ncomp,
                          long long inner, long long outer, const double2* tw, double scale, const Limits& max_lds,
                          hipStream_t s) {
  PassCfg c = make_pass_cfg(N, RM, inner, outer, ncomp, max_lds);
  if (c.block.x == 0 || c.block.x > 256 || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok[4] = {false, false, false, false};
  hipError_t e = hipSuccess;
  if (contig) {
    if (dir < 0) {
      e = allow_lds(k_pass<N, RM, true, -1>, c.lds, ok[0]);
      if (e == hipSuccess)
        hipLaunchKernelGGL((k_pass<N, RM, true, -1>), c.grid, c.block, c.lds, s, src, dst, cs, c.g, tw, scale);
    } else {
      e = allow_lds(k_pass<N, RM, true, 1>, c.lds, ok[1]);
      if (e == hipSuccess)
        hipLaunchKernelGGL((k_pass<N, RM, true, 1>), c.grid, c.block, c.lds, s, src, dst, cs, c.g, tw, scale);
    }
  } else {
    if (dir < 0) {
      e = allow_lds(k_pass<N, RM, false, -1>, c.lds, ok[2]);
      if (e == hipSuccess)
        hipLaunchKernelGGL((k_pass<N, RM, false, -1>), c.grid, c.block, c.lds, s, src, dst, cs, c.g, tw, scale);
    } else {
      e = allow_lds(k_pass<N, RM, false, 1>, c.lds, ok[3]);
      if (e == hipSuccess)
        hipLaunchKernelGGL((k_pass<N, RM, false, 1>), c.grid, c.block, c.lds, s, src, dst, cs, c.g, tw, scale);
    }
  }
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

template <int N>
static hipError_t pass_n(int rm, bool contig, int dir, const double2* src, double2* dst, long long cs, int ncomp,
                         long long inner, long long outer, const double2* tw, double scale, const Limits& max_lds,
                         hipStream_t s) {
  if (rm == 8 && N > 8 && N / 8 <= 256) return pass_nr<N, 8>(contig, dir, src, dst, cs, ncomp, inner, outer, tw, scale, max_lds, s);
  return pass_nr<N, 16>(contig, dir, src, dst, cs, ncomp, inner, outer, tw, scale, max_lds, s);
}

hipError_t launch_pass(int N, int rm, bool contig, int dir, const double2* src, double2* dst, long long comp_stride,
                       int ncomp, long long inner, long long outer, const double2* tw, double scale,
                       const Limits& max_lds, hipStream_t s) {
  switch (N) {
#define ATUS_CASE(n) \
  case n: return pass_n<n>(rm, contig, dir, src, dst, comp_stride, ncomp, inner, outer, tw, scale, max_lds, s);
    ATUS_CASE(2) ATUS_CASE(4) ATUS_CASE(8) ATUS_CASE(16) ATUS_CASE(32) ATUS_CASE(64) ATUS_CASE(128)
    ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024) ATUS_CASE(2048) ATUS_CASE(4096)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace atus
