// k_vfuse.hip -- inverse DFT / pointwise V / forward DFT along the last axis.
#include "launch.hpp"

namespace atus {

template <int N, int RM, int VK, int NC, int EDGE>
static hipError_t vf_go(double2* psi, long long cs, long long outer, const double2* tw, const VArgs& va,
                        long long ppx, int n0, double dx0, const double2* kz, const Limits& max_lds, hipStream_t s) {
  PassCfg c = make_pass_cfg(N, RM, 1, outer, 1, max_lds, EDGE ? 2 : 1);
  if (c.block.x == 0 || c.block.x > 256 || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok = false;
  hipError_t e = allow_lds(k_vfuse<N, RM, VK, NC, false, EDGE>, c.lds, ok);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_vfuse<N, RM, VK, NC, false, EDGE>), c.grid, c.block, c.lds, s, psi, cs, c.g, tw, va, ppx, n0,
                     dx0, kz);
  return hipGetLastError();
}

template <int N, int RM, int EDGE>
static hipError_t vf_nr(int vk, int nc, double2* psi, long long cs, long long outer, const double2* tw,
                        const VArgs& va, long long ppx, int n0, double dx0, const double2* kz, const Limits& max_lds,
                        hipStream_t s) {
#define ATUS_VF(V, C) \
  if (vk == V && nc == C) return vf_go<N, RM, V, C, EDGE>(psi, cs, outer, tw, va, ppx, n0, dx0, kz, max_lds, s);
  ATUS_VF(V_NL, 1) ATUS_VF(V_NL_IF, 1) ATUS_VF(V_NL, 2) ATUS_VF(V_NL_IF, 2) ATUS_VF(V_BRAGG, 2)
#undef ATUS_VF
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t vf_n(int rm, int vk, int nc, double2* psi, long long cs, long long outer, const double2* tw,
                       const VArgs& va, long long ppx, int n0, double dx0, const double2* kz, int edge,
                       const Limits& max_lds, hipStream_t s) {
  constexpr bool r8 = N > 8 && N / 8 <= 256;
  // the block edges (once per block) always take the plan of 8 values per
  // thread where it exists, so that the tuned choice does not double the
  // edge kernels
  constexpr int RE = r8 ? 8 : 16;
  if (edge == 1) return vf_nr<N, RE, 1>(vk, nc, psi, cs, outer, tw, va, ppx, n0, dx0, kz, max_lds, s);
  if (edge == 2) return vf_nr<N, RE, 2>(vk, nc, psi, cs, outer, tw, va, ppx, n0, dx0, kz, max_lds, s);
  if (edge == 3) return vf_nr<N, RE, 3>(vk, nc, psi, cs, outer, tw, va, ppx, n0, dx0, kz, max_lds, s);
  if (rm == 8 && r8) return vf_nr<N, 8, 0>(vk, nc, psi, cs, outer, tw, va, ppx, n0, dx0, nullptr, max_lds, s);
  return vf_nr<N, 16, 0>(vk, nc, psi, cs, outer, tw, va, ppx, n0, dx0, nullptr, max_lds, s);
}

hipError_t launch_vfuse(int N, int rm, int vk, int nc, double2* psi, long long comp_stride, long long outer,
                        const double2* tw, const VArgs& va, long long pencils_per_x0, int n0, double dx0,
                        const Limits& max_lds, hipStream_t s, const double2* kz, int edge) {
  if (edge < 0 || edge > 3 || (edge && !kz)) return hipErrorInvalidValue;
  switch (N) {
#define ATUS_CASE(n) \
  case n:            \
    return vf_n<n>(rm, vk, nc, psi, comp_stride, outer, tw, va, pencils_per_x0, n0, dx0, kz, edge, max_lds, s);
    ATUS_CASE(2) ATUS_CASE(4) ATUS_CASE(8) ATUS_CASE(16) ATUS_CASE(32) ATUS_CASE(64) ATUS_CASE(128)
    ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024) ATUS_CASE(2048) ATUS_CASE(4096)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace atus
