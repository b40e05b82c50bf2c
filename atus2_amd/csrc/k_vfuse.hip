// k_vfuse.hip -- inverse DFT / pointwise V / forward DFT along the last axis.
#include "launch.hpp"

namespace atus {

template <int N, int RM, int VK, int NC>
static hipError_t vf_go(double2* psi, long long cs, long long outer, const double2* tw, const VArgs& va,
                        long long ppx, int n0, double dx0, const Limits& max_lds, hipStream_t s) {
  PassCfg c = make_pass_cfg(N, RM, 1, outer, 1, max_lds);
  if (c.block.x == 0 || c.block.x > 256 || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok = false;
  hipError_t e = allow_lds(k_vfuse<N, RM, VK, NC>, c.lds, ok);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_vfuse<N, RM, VK, NC>), c.grid, c.block, c.lds, s, psi, cs, c.g, tw, va, ppx, n0, dx0);
  return hipGetLastError();
}

template <int N, int RM>
static hipError_t vf_nr(int vk, int nc, double2* psi, long long cs, long long outer, const double2* tw,
                       const VArgs& va, long long ppx, int n0, double dx0, const Limits& max_lds, hipStream_t s) {
  if (nc == 1) {
    if (vk == V_NL) return vf_go<N, RM, V_NL, 1>(psi, cs, outer, tw, va, ppx, n0, dx0, max_lds, s);
    if (vk == V_NL_IF) return vf_go<N, RM, V_NL_IF, 1>(psi, cs, outer, tw, va, ppx, n0, dx0, max_lds, s);
  } else if (nc == 2) {
    if (vk == V_NL) return vf_go<N, RM, V_NL, 2>(psi, cs, outer, tw, va, ppx, n0, dx0, max_lds, s);
    if (vk == V_NL_IF) return vf_go<N, RM, V_NL_IF, 2>(psi, cs, outer, tw, va, ppx, n0, dx0, max_lds, s);
    if (vk == V_BRAGG) return vf_go<N, RM, V_BRAGG, 2>(psi, cs, outer, tw, va, ppx, n0, dx0, max_lds, s);
  }
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t vf_n(int rm, int vk, int nc, double2* psi, long long cs, long long outer, const double2* tw,
                       const VArgs& va, long long ppx, int n0, double dx0, const Limits& max_lds, hipStream_t s) {
  if (rm == 8 && N > 8 && N / 8 <= 256) return vf_nr<N, 8>(vk, nc, psi, cs, outer, tw, va, ppx, n0, dx0, max_lds, s);
  return vf_nr<N, 16>(vk, nc, psi, cs, outer, tw, va, ppx, n0, dx0, max_lds, s);
}

hipError_t launch_vfuse(int N, int rm, int vk, int nc, double2* psi, long long comp_stride, long long outer,
                        const double2* tw, const VArgs& va, long long pencils_per_x0, int n0, double dx0,
                        const Limits& max_lds, hipStream_t s) {
  switch (N) {
#define ATUS_CASE(n) \
  case n: return vf_n<n>(rm, vk, nc, psi, comp_stride, outer, tw, va, pencils_per_x0, n0, dx0, max_lds, s);
    ATUS_CASE(2) ATUS_CASE(4) ATUS_CASE(8) ATUS_CASE(16) ATUS_CASE(32) ATUS_CASE(64) ATUS_CASE(128)
    ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024) ATUS_CASE(2048) ATUS_CASE(4096)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace atus
