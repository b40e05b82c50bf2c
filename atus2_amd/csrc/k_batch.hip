// k_batch.hip -- V-step kernels of batches of independent two-component
// runs: the 1D fused step (whole pencil = one run) and the contiguous-axis
// pass of 2D/3D runs, with the parameters of each run read per pencil.
#include "launch.hpp"

namespace atus {

template <int N, int RM, int VK>
static hipError_t kinb_go(double2* psi, long long cs, long long nbatch, const double2* tw, const KinTabs& kt,
                          const VArgs& va, double dx0, const Limits& max_lds, hipStream_t s) {
  PassCfg c = make_pass_cfg(N, RM, 1, nbatch, 1, max_lds, 0, 256, kin_lds_table_elems<N, RM, true>());
  if (c.block.x == 0 || c.block.x > 256 || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok = false;
  hipError_t e = allow_lds(k_kin<N, RM, true, VK, 2, true>, c.lds, ok);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_kin<N, RM, true, VK, 2, true>), c.grid, c.block, c.lds, s, psi, cs, c.g, tw, kt, va, dx0);
  return hipGetLastError();
}

template <int N, int RM, int VK, int EDGE>
static hipError_t vfb_go(double2* psi, long long cs, long long outer, const double2* tw, const VArgs& va,
                         long long ppx, int n0, double dx0, const double2* kz, const Limits& max_lds, hipStream_t s) {
  PassCfg c = make_pass_cfg(N, RM, 1, outer, 1, max_lds, EDGE ? 2 : 1);
  if (c.block.x == 0 || c.block.x > 256 || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok = false;
  hipError_t e = allow_lds(k_vfuse<N, RM, VK, 2, true, EDGE>, c.lds, ok);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_vfuse<N, RM, VK, 2, true, EDGE>), c.grid, c.block, c.lds, s, psi, cs, c.g, tw, va, ppx, n0,
                     dx0, kz);
  return hipGetLastError();
}

template <int N, int RM, int EDGE>
static hipError_t batch_nr(bool one_d, int vk, double2* psi, long long cs, long long count, const double2* tw,
                           const KinTabs& kt, const VArgs& va, long long ppx, int n0, double dx0, const double2* kz,
                           const Limits& max_lds, hipStream_t s) {
  if (one_d) {
    if constexpr (EDGE == 0) {
      if (vk == V_BRAGG) return kinb_go<N, RM, V_BRAGG>(psi, cs, count, tw, kt, va, dx0, max_lds, s);
      if (vk == V_NL_IF) return kinb_go<N, RM, V_NL_IF>(psi, cs, count, tw, kt, va, dx0, max_lds, s);
    }
  } else {
    if (vk == V_BRAGG) return vfb_go<N, RM, V_BRAGG, EDGE>(psi, cs, count, tw, va, ppx, n0, dx0, kz, max_lds, s);
    if (vk == V_NL_IF) return vfb_go<N, RM, V_NL_IF, EDGE>(psi, cs, count, tw, va, ppx, n0, dx0, kz, max_lds, s);
  }
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t batch_n(int rm, bool one_d, int vk, double2* psi, long long cs, long long count, const double2* tw,
                          const KinTabs& kt, const VArgs& va, long long ppx, int n0, double dx0, const double2* kz,
                          int edge, const Limits& max_lds, hipStream_t s) {
  constexpr bool r8 = N > 8 && N / 8 <= 256;
  // block edges: the plan of 8 values per thread where it exists (as for
  // single runs)
  constexpr int RE = r8 ? 8 : 16;
  if (edge == 1) return batch_nr<N, RE, 1>(one_d, vk, psi, cs, count, tw, kt, va, ppx, n0, dx0, kz, max_lds, s);
  if (edge == 2) return batch_nr<N, RE, 2>(one_d, vk, psi, cs, count, tw, kt, va, ppx, n0, dx0, kz, max_lds, s);
  if (edge == 3) return batch_nr<N, RE, 3>(one_d, vk, psi, cs, count, tw, kt, va, ppx, n0, dx0, kz, max_lds, s);
  if (rm == 8 && r8)
    return batch_nr<N, 8, 0>(one_d, vk, psi, cs, count, tw, kt, va, ppx, n0, dx0, nullptr, max_lds, s);
  return batch_nr<N, 16, 0>(one_d, vk, psi, cs, count, tw, kt, va, ppx, n0, dx0, nullptr, max_lds, s);
}

hipError_t launch_vstep_batch(int N, int rm, bool one_d, int vk, double2* psi, long long comp_stride, long long count,
                              const double2* tw, const KinTabs& kt, const VArgs& va, long long pencils_per_x0, int n0,
                              double dx0, const Limits& max_lds, hipStream_t s, const double2* kz, int edge) {
  if (edge < 0 || edge > 3 || (edge && (!kz || one_d))) return hipErrorInvalidValue;
  switch (N) {
#define ATUS_CASE(n)                                                                                                  \
  case n:                                                                                                             \
    return batch_n<n>(rm, one_d, vk, psi, comp_stride, count, tw, kt, va, pencils_per_x0, n0, dx0, kz, edge, max_lds, \
                      s);
    ATUS_CASE(2) ATUS_CASE(4) ATUS_CASE(8) ATUS_CASE(16) ATUS_CASE(32) ATUS_CASE(64) ATUS_CASE(128)
    ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024) ATUS_CASE(2048) ATUS_CASE(4096)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace atus
