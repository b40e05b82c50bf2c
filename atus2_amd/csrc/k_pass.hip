// k_pass.hip -- single-direction FFT passes along one axis.
#include "launch.hpp"

namespace atus {

bool supported_len(int N) { return N >= 2 && N <= 4096 && (N & (N - 1)) == 0; }

// LDS elements per pencil region.  The bank-conflict counters of the
// passes fit a model where LDS serves a 16-byte read of a wavefront 16
// lanes at a time from 64 banks of 4 bytes (16 slots of 16 B), and a
// 16-byte write 8 lanes at a time from 32 banks (8 slots).  Within a pencil
// the exchange positions are XOR-swizzled in rows of 16 elements.  When the
// lanes of a group span several pencils, the regions of those pencils are
// padded so that they start on different slots:
//   * strided passes (lane = t*T + p): an 8-lane write group holds
//     g = min(T, 8) pencils, padding 8/g elements (1 for T = 1 or T >= 8);
//   * contiguous passes with P < 16 threads per pencil: a 16-lane read
//     group holds 16/P pencils, padding P elements;
//   * otherwise one element.
static int pencil_lds_elems(int N, int P, bool contig, long long T) {
  int pad = 1;
  if (contig) {
    if (P < 16) pad = P;
  } else if (T > 1 && T < 8) {
    pad = 8 / (int)T;
  }
  return N + pad;
}

PassCfg make_pass_cfg(int N, int rm, long long inner, long long outer, int ycount, size_t max_lds, int ntables,
                      int max_threads, long long table_elems) {
  PassCfg c{};
  const int P = threads_per_pencil(N, rm);
  const bool contig = inner == 1;
  const size_t table_bytes = table_elems >= 0   ? (size_t)table_elems * sizeof(double2)
                             : N <= 1024 ? (size_t)ntables * N * sizeof(double2)
                                         : 0;
  long long T = max_threads / P;
  if (T < 1) T = 1;
  if (!contig && T > inner) T = inner;
  auto pencil_bytes = [&](long long t) { return (size_t)pencil_lds_elems(N, P, contig, t) * sizeof(double2); };
  while (T > 1 && (size_t)T * pencil_bytes(T) + table_bytes > max_lds) T >>= 1;
  if (inner == 1) {
    long long tiles = (outer + T - 1) / T;
    c.g.inner = 1;
    c.g.outer = outer;
    c.g.col_tiles = 1;
    c.g.ntiles = tiles;
    c.g.T = (int)T;
  } else {
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
  c.g.ls = pencil_lds_elems(N, P, contig, T);
  c.lds = (size_t)T * pencil_bytes(T) + table_bytes;
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

// ------------------------------------------------ Rabi integrals of a 3D box

template <int N, int RM>
static hipError_t rabi_nr(int axis, const double2* src, long long seg, int nseg, double2* zb, double2* yb,
                          const double2* tw, const RabiBox& rb, const Limits& lim, double* partials,
                          unsigned* nblocks, hipStream_t s) {
  PassCfg c;
  if (axis == 2) c = make_pass_cfg(N, RM, 1, (long long)rb.n[0] * rb.n[1], nseg, lim);
  else if (axis == 1) c = make_pass_cfg(N, RM, rb.len[2], rb.n[0], nseg, lim);
  else c = make_pass_cfg(N, RM, (long long)rb.len[1] * rb.len[2], 1, nseg, lim);
  if (c.block.x == 0 || c.block.x > 256 || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok[3] = {false, false, false};
  hipError_t e;
  if (axis == 2) {
    e = allow_lds(k_rabi_z<N, RM>, c.lds, ok[0]);
    if (e == hipSuccess) hipLaunchKernelGGL((k_rabi_z<N, RM>), c.grid, c.block, c.lds, s, src, seg, zb, c.g, tw, rb);
  } else if (axis == 1) {
    e = allow_lds(k_rabi_y<N, RM>, c.lds, ok[1]);
    if (e == hipSuccess) hipLaunchKernelGGL((k_rabi_y<N, RM>), c.grid, c.block, c.lds, s, zb, yb, c.g, tw, rb);
  } else {
    if (nblocks) *nblocks = c.grid.x;
    e = allow_lds(k_rabi_x<N, RM>, c.lds, ok[2]);
    if (e == hipSuccess && partials)
      hipLaunchKernelGGL((k_rabi_x<N, RM>), c.grid, c.block, c.lds, s, yb, c.g, tw, rb, partials);
  }
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

template <int N>
static hipError_t rabi_n(int rm, int axis, const double2* src, long long seg, int nseg, double2* zb, double2* yb,
                         const double2* tw, const RabiBox& rb, const Limits& lim, double* partials, unsigned* nblocks,
                         hipStream_t s) {
  if (rm == 8 && N > 8 && N / 8 <= 256)
    return rabi_nr<N, 8>(axis, src, seg, nseg, zb, yb, tw, rb, lim, partials, nblocks, s);
  return rabi_nr<N, 16>(axis, src, seg, nseg, zb, yb, tw, rb, lim, partials, nblocks, s);
}

hipError_t launch_rabi_box(int axis, int rm, const double2* src, long long seg, int nseg, double2* zb, double2* yb,
                           const double2* tw, const RabiBox& rb, const Limits& lim, double* partials,
                           unsigned* nblocks, hipStream_t s) {
  if (axis < 0 || axis > 2) return hipErrorInvalidValue;
  switch (rb.n[axis]) {
#define ATUS_CASE(n) \
  case n: return rabi_n<n>(rm, axis, src, seg, nseg, zb, yb, tw, rb, lim, partials, nblocks, s);
    ATUS_CASE(2) ATUS_CASE(4) ATUS_CASE(8) ATUS_CASE(16) ATUS_CASE(32) ATUS_CASE(64) ATUS_CASE(128)
    ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024) ATUS_CASE(2048) ATUS_CASE(4096)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace atus
