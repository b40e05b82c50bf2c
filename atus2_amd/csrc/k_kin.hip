// k_kin.hip -- fused kinetic step along axis 0 (forward DFT, phase, inverse).
#include "launch.hpp"

namespace atus {

// LB: threads per workgroup at most (the kernel's launch bounds)
template <int N, int RM, bool CONTIG, int VK, int NC, int LB = 256>
static hipError_t kin_go(double2* psi, long long cs, int ycount, long long inner, long long outer,
                         const double2* tw, const KinTabs& kt, const VArgs& va, double dx0, const Limits& max_lds,
                         hipStream_t s) {
  PassCfg c = make_pass_cfg(N, RM, inner, outer, ycount, max_lds, 0, LB, kin_lds_table_elems<N, RM, CONTIG>());
  if (c.block.x == 0 || c.block.x > LB || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok = false;
  hipError_t e = allow_lds(k_kin<N, RM, CONTIG, VK, NC, false, LB>, c.lds, ok);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_kin<N, RM, CONTIG, VK, NC, false, LB>), c.grid, c.block, c.lds, s, psi, cs, c.g, tw, kt, va,
                     dx0);
  return hipGetLastError();
}

template <int N, int RM>
static hipError_t kin_nr(bool contig, int vk, int nc, double2* psi, long long cs, int ncomp, long long inner,
                        const double2* tw, const KinTabs& kt, const VArgs& va, double dx0, const Limits& max_lds,
                        hipStream_t s, long long rows) {
  if (!contig) {
    // 2D/3D: axis 0 is strided, one component per blockIdx.y, no V prologue
    if (vk != V_NONE) return hipErrorInvalidValue;
    // columns of 2048 and more: workgroups of up to 512 threads hold twice
    // the columns, so each row is read in segments twice as long (at 1024
    // points, 512 threads leave one workgroup per CU: slower)
    if constexpr (N >= 2048)
      return kin_go<N, RM, false, V_NONE, 1, 512>(psi, cs, ncomp, inner, rows, tw, kt, va, dx0, max_lds, s);
    return kin_go<N, RM, false, V_NONE, 1>(psi, cs, ncomp, inner, rows, tw, kt, va, dx0, max_lds, s);
  }
  // 1D: the pencil is the whole grid
  if (vk == V_NONE) return kin_go<N, RM, true, V_NONE, 1>(psi, cs, ncomp, 1, rows, tw, kt, va, dx0, max_lds, s);
  if (nc == 1) {
    if (vk == V_NL) return kin_go<N, RM, true, V_NL, 1>(psi, cs, 1, 1, 1, tw, kt, va, dx0, max_lds, s);
    if (vk == V_NL_IF) return kin_go<N, RM, true, V_NL_IF, 1>(psi, cs, 1, 1, 1, tw, kt, va, dx0, max_lds, s);
  } else if (nc == 2) {
    if (vk == V_NL) return kin_go<N, RM, true, V_NL, 2>(psi, cs, 1, 1, 1, tw, kt, va, dx0, max_lds, s);
    if (vk == V_NL_IF) return kin_go<N, RM, true, V_NL_IF, 2>(psi, cs, 1, 1, 1, tw, kt, va, dx0, max_lds, s);
    if (vk == V_BRAGG) return kin_go<N, RM, true, V_BRAGG, 2>(psi, cs, 1, 1, 1, tw, kt, va, dx0, max_lds, s);
  }
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t kin_n(int rm, bool contig, int vk, int nc, double2* psi, long long cs, int ncomp, long long inner,
                        const double2* tw, const KinTabs& kt, const VArgs& va, double dx0, const Limits& max_lds,
                        hipStream_t s, long long rows) {
  if (rm == 8 && N > 8 && N / 8 <= 256)
    return kin_nr<N, 8>(contig, vk, nc, psi, cs, ncomp, inner, tw, kt, va, dx0, max_lds, s, rows);
  return kin_nr<N, 16>(contig, vk, nc, psi, cs, ncomp, inner, tw, kt, va, dx0, max_lds, s, rows);
}

hipError_t launch_kin(int N, int rm, bool contig, int vk, int nc, double2* psi, long long comp_stride, int ncomp,
                      long long inner, const double2* tw, const KinTabs& kt, const VArgs& va, double dx0,
                      const Limits& max_lds, hipStream_t s, long long rows) {
  switch (N) {
#define ATUS_CASE(n) \
  case n: return kin_n<n>(rm, contig, vk, nc, psi, comp_stride, ncomp, inner, tw, kt, va, dx0, max_lds, s, rows);
    ATUS_CASE(2) ATUS_CASE(4) ATUS_CASE(8) ATUS_CASE(16) ATUS_CASE(32) ATUS_CASE(64) ATUS_CASE(128)
    ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024) ATUS_CASE(2048) ATUS_CASE(4096)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace atus
