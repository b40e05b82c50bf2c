// k_block.hip -- whole blocks of 1D runs in one launch (k_block1d).
#include "launch.hpp"

namespace atus {

template <int N, int RM, int VK, int NC, bool BATCH>
static hipError_t block_go(double2* psi, long long cs, long long runs, const double2* tw, const double2* eh,
                           const double2* ef, const VArgs& va, const double* t1tab, const double* Ftab, double dx0,
                           int Nk, const Limits& max_lds, hipStream_t s) {
  // pencils of more than 256 threads (4096 points at 8 values per thread)
  // take workgroups of 512
  constexpr int LB = Plan<N, RM>::P > 256 ? 512 : 256;
  // a single run: one pencil per workgroup; a batch: as many pencils as
  // the workgroup holds
  const int max_threads = BATCH ? LB : threads_per_pencil(N, RM);
  PassCfg c = make_pass_cfg(N, RM, 1, runs, 1, max_lds, 3, max_threads);
  // above 1024 points the twiddle table alone is staged in LDS
  if constexpr (!tables_in_lds<N>()) c.lds += (size_t)N * sizeof(double2);
  if (c.block.x == 0 || c.block.x > LB || c.grid.x == 0 || c.lds > max_lds.lds) return hipErrorInvalidConfiguration;
  static bool ok = false;
  hipError_t e = allow_lds(k_block1d<N, RM, VK, NC, BATCH, LB>, c.lds, ok);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_block1d<N, RM, VK, NC, BATCH, LB>), c.grid, c.block, c.lds, s, psi, cs, c.g, tw, eh, ef, va,
                     t1tab, Ftab, dx0, Nk);
  return hipGetLastError();
}

template <int N, int RM>
static hipError_t block_nr(int vk, int nc, bool batch, double2* psi, long long cs, long long runs,
                           const double2* tw, const double2* eh, const double2* ef, const VArgs& va,
                           const double* t1tab, const double* Ftab, double dx0, int Nk, const Limits& max_lds,
                           hipStream_t s) {
  if (batch) {
    if (nc != 2) return hipErrorInvalidValue;
    if (vk == V_BRAGG)
      return block_go<N, RM, V_BRAGG, 2, true>(psi, cs, runs, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
    if (vk == V_NL_IF)
      return block_go<N, RM, V_NL_IF, 2, true>(psi, cs, runs, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
    return hipErrorInvalidValue;
  }
  if (nc == 1) {
    if (vk == V_NL) return block_go<N, RM, V_NL, 1, false>(psi, cs, 1, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
    if (vk == V_NL_IF)
      return block_go<N, RM, V_NL_IF, 1, false>(psi, cs, 1, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
  } else if (nc == 2) {
    if (vk == V_NL) return block_go<N, RM, V_NL, 2, false>(psi, cs, 1, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
    if (vk == V_NL_IF)
      return block_go<N, RM, V_NL_IF, 2, false>(psi, cs, 1, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
    if (vk == V_BRAGG)
      return block_go<N, RM, V_BRAGG, 2, false>(psi, cs, 1, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
  }
  return hipErrorInvalidValue;
}

// values per thread as tuned for the kinetic family (8 or 16); 8 also at
// 4096 points, where the per-step kinetic pass falls back to 16
template <int N>
static hipError_t block_n(int rm, int vk, int nc, bool batch, double2* psi, long long cs, long long runs,
                          const double2* tw, const double2* eh, const double2* ef, const VArgs& va,
                          const double* t1tab, const double* Ftab, double dx0, int Nk, const Limits& max_lds,
                          hipStream_t s) {
  if constexpr (N > 8 && N / 8 <= 512) {
    if (rm == 8) return block_nr<N, 8>(vk, nc, batch, psi, cs, runs, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
  }
  return block_nr<N, 16>(vk, nc, batch, psi, cs, runs, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
}

hipError_t launch_block1d(int N, int rm, int vk, int nc, bool batch, double2* psi, long long comp_stride,
                          long long runs, const double2* tw, const double2* eh, const double2* ef, const VArgs& va,
                          const double* t1tab, const double* Ftab, double dx0, int Nk, const Limits& max_lds,
                          hipStream_t s) {
  switch (N) {
#define ATUS_CASE(n)                                                                                              \
  case n:                                                                                                         \
    return block_n<n>(rm, vk, nc, batch, psi, comp_stride, runs, tw, eh, ef, va, t1tab, Ftab, dx0, Nk, max_lds, s);
    ATUS_CASE(2) ATUS_CASE(4) ATUS_CASE(8) ATUS_CASE(16) ATUS_CASE(32) ATUS_CASE(64) ATUS_CASE(128)
    ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024) ATUS_CASE(2048) ATUS_CASE(4096)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace atus
