// k_sob.hip -- Sobolev filter passes along axis 0 and the pointwise V launcher.
#include "launch.hpp"

namespace atus {

template <int N, int RM>
static hipError_t sob_nr(bool contig, const double2* src, double2* o1, double2* o2, long long cs, int ncomp,
                         long long inner, const double2* tw, const SobArgs& sa, const Limits& max_lds, hipStream_t s) {
  PassCfg c = make_pass_cfg(N, RM, contig ? 1 : inner, 1, ncomp, max_lds);
  if (c.block.x == 0 || c.block.x > 256 || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok[2] = {false, false};
  hipError_t e;
  if (contig) {
    e = allow_lds(k_sobf<N, RM, true>, c.lds, ok[0]);
    if (e == hipSuccess)
      hipLaunchKernelGGL((k_sobf<N, RM, true>), c.grid, c.block, c.lds, s, src, o1, o2, cs, c.g, tw, sa);
  } else {
    e = allow_lds(k_sobf<N, RM, false>, c.lds, ok[1]);
    if (e == hipSuccess)
      hipLaunchKernelGGL((k_sobf<N, RM, false>), c.grid, c.block, c.lds, s, src, o1, o2, cs, c.g, tw, sa);
  }
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

template <int N>
static hipError_t sob_n(int rm, bool contig, const double2* src, double2* o1, double2* o2, long long cs, int ncomp,
                        long long inner, const double2* tw, const SobArgs& sa, const Limits& max_lds, hipStream_t s) {
  if (rm == 8 && N > 8 && N / 8 <= 256) return sob_nr<N, 8>(contig, src, o1, o2, cs, ncomp, inner, tw, sa, max_lds, s);
  return sob_nr<N, 16>(contig, src, o1, o2, cs, ncomp, inner, tw, sa, max_lds, s);
}

hipError_t launch_sobf(int N, int rm, bool contig, const double2* src, double2* out_sob, double2* out_lap,
                       long long comp_stride, int ncomp, long long inner, const double2* tw, const SobArgs& sa,
                       const Limits& max_lds, hipStream_t s) {
  switch (N) {
#define ATUS_CASE(n) \
  case n: return sob_n<n>(rm, contig, src, out_sob, out_lap, comp_stride, ncomp, inner, tw, sa, max_lds, s);
    ATUS_CASE(2) ATUS_CASE(4) ATUS_CASE(8) ATUS_CASE(16) ATUS_CASE(32) ATUS_CASE(64) ATUS_CASE(128)
    ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024) ATUS_CASE(2048) ATUS_CASE(4096)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

template <int N, int RM, int NW>
static hipError_t gf_go(const double2* lap, const double2* psi, const double* pot, double2* out, long long npts,
                        int comp, double g0, double g1, const double2* tw, const Limits& max_lds, hipStream_t s) {
  PassCfg c = make_pass_cfg(N, RM, 1, npts / N, 1, max_lds);
  if (c.block.x == 0 || c.block.x > 256 || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok = false;
  hipError_t e = allow_lds(k_sob_gradfuse<N, RM, NW>, c.lds, ok);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_sob_gradfuse<N, RM, NW>), c.grid, c.block, c.lds, s, lap, psi, pot, out, npts, c.g, tw, comp,
                     g0, g1);
  return hipGetLastError();
}

template <int N>
static hipError_t gf_n(int rm, int nw, const double2* lap, const double2* psi, const double* pot, double2* out,
                       long long npts, int comp, double g0, double g1, const double2* tw, const Limits& max_lds,
                       hipStream_t s) {
  const bool r8 = rm == 8 && N > 8 && N / 8 <= 256;
  if (nw == 1)
    return r8 ? gf_go<N, 8, 1>(lap, psi, pot, out, npts, comp, g0, g1, tw, max_lds, s)
              : gf_go<N, 16, 1>(lap, psi, pot, out, npts, comp, g0, g1, tw, max_lds, s);
  if (nw == 2)
    return r8 ? gf_go<N, 8, 2>(lap, psi, pot, out, npts, comp, g0, g1, tw, max_lds, s)
              : gf_go<N, 16, 2>(lap, psi, pot, out, npts, comp, g0, g1, tw, max_lds, s);
  return hipErrorInvalidValue;
}

hipError_t launch_sob_gradfuse(int N, int rm, int nw, const double2* lap, const double2* psi, const double* pot,
                               double2* out, long long npts, int comp, double g0, double g1, const double2* tw,
                               const Limits& max_lds, hipStream_t s) {
  switch (N) {
#define ATUS_CASE(n) \
  case n: return gf_n<n>(rm, nw, lap, psi, pot, out, npts, comp, g0, g1, tw, max_lds, s);
    ATUS_CASE(2) ATUS_CASE(4) ATUS_CASE(8) ATUS_CASE(16) ATUS_CASE(32) ATUS_CASE(64) ATUS_CASE(128)
    ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024) ATUS_CASE(2048) ATUS_CASE(4096)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

static unsigned point_blocks(long long n) {
  long long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

hipError_t launch_point(int vk, int nc, double2* psi, const Grid& gr, const VArgs& va, hipStream_t s) {
  dim3 grid(point_blocks(gr.npts)), block(256);
#define ATUS_PT(V, C) \
  if (vk == V && nc == C) { hipLaunchKernelGGL((k_point<V, C>), grid, block, 0, s, psi, gr, va); return hipGetLastError(); }
  ATUS_PT(V_NL, 1) ATUS_PT(V_NL, 2) ATUS_PT(V_NL, 3) ATUS_PT(V_NL, 4)
  ATUS_PT(V_NL_IF, 1) ATUS_PT(V_NL_IF, 2) ATUS_PT(V_NL_IF, 3) ATUS_PT(V_NL_IF, 4)
  ATUS_PT(V_BRAGG, 2)
  ATUS_PT(V_NL_IF2, 2) ATUS_PT(V_NL_IF2, 4) ATUS_PT(V_BRAGG_DS, 4)
  ATUS_PT(V_DBRAGG, 3) ATUS_PT(V_NUM_BRAGG, 2) ATUS_PT(V_NUM_BRAGG, 3) ATUS_PT(V_RAMAN, 3) ATUS_PT(V_RAMAN, 4)
#undef ATUS_PT
  return hipErrorInvalidValue;
}

}  // namespace atus
