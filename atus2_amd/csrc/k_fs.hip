// k_fs.hip -- column passes of the four-step 1D transform (N = N1 * N2,
// 2^13 <= N <= 2^20, N1 = 64 .. 1024).
#include "launch.hpp"

namespace atus {

template <int N1, int RM, int MODE, int VK, int NC, bool BATCH>
static hipError_t fs_go(double2* psi, long long cs, int ycount, long long n2, const double2* tw1,
                        const double2* wN, const VArgs& va, double dx, double scale, const Limits& max_lds,
                        hipStream_t s) {
  PassCfg c = make_pass_cfg(N1, RM, n2, 1, ycount, max_lds);
  if (c.block.x == 0 || c.block.x > 256 || c.grid.x == 0) return hipErrorInvalidConfiguration;
  static bool ok = false;
  hipError_t e = allow_lds(k_fs_col<N1, RM, MODE, VK, NC, BATCH>, c.lds, ok);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_fs_col<N1, RM, MODE, VK, NC, BATCH>), c.grid, c.block, c.lds, s, psi, cs, c.g, tw1, wN, va,
                     dx, scale);
  return hipGetLastError();
}

template <int N1, int RM>
static hipError_t fs_nr(int mode, int vk, int nc, bool batch, double2* psi, long long cs, int ycount, long long n2,
                        const double2* tw1, const double2* wN, const VArgs& va, double dx, double scale,
                        const Limits& max_lds, hipStream_t s) {
#define ATUS_FS(M, V, C, B) fs_go<N1, RM, M, V, C, B>(psi, cs, ycount, n2, tw1, wN, va, dx, scale, max_lds, s)
  if (mode == FS_ENTER) return ATUS_FS(FS_ENTER, V_NONE, 1, false);
  if (mode == FS_EXIT) return ATUS_FS(FS_EXIT, V_NONE, 1, false);
  if (mode != FS_STEP) return hipErrorInvalidValue;
  if (batch) {
    if (nc != 2) return hipErrorInvalidValue;
    if (vk == V_BRAGG) return ATUS_FS(FS_STEP, V_BRAGG, 2, true);
    if (vk == V_NL_IF) return ATUS_FS(FS_STEP, V_NL_IF, 2, true);
    return hipErrorInvalidValue;
  }
  if (nc == 1) {
    if (vk == V_NL) return ATUS_FS(FS_STEP, V_NL, 1, false);
    if (vk == V_NL_IF) return ATUS_FS(FS_STEP, V_NL_IF, 1, false);
  } else if (nc == 2) {
    if (vk == V_NL) return ATUS_FS(FS_STEP, V_NL, 2, false);
    if (vk == V_NL_IF) return ATUS_FS(FS_STEP, V_NL_IF, 2, false);
    if (vk == V_BRAGG) return ATUS_FS(FS_STEP, V_BRAGG, 2, false);
  }
#undef ATUS_FS
  return hipErrorInvalidValue;
}

template <int N1>
static hipError_t fs_n(int rm, int mode, int vk, int nc, bool batch, double2* psi, long long cs, int ycount,
                       long long n2, const double2* tw1, const double2* wN, const VArgs& va, double dx, double scale,
                       const Limits& max_lds, hipStream_t s) {
  if (rm == 8) return fs_nr<N1, 8>(mode, vk, nc, batch, psi, cs, ycount, n2, tw1, wN, va, dx, scale, max_lds, s);
  return fs_nr<N1, 16>(mode, vk, nc, batch, psi, cs, ycount, n2, tw1, wN, va, dx, scale, max_lds, s);
}

hipError_t launch_fs_col(int n1, int rm, int mode, int vk, int nc, bool batch, double2* psi, long long comp_stride,
                         int ycount, long long n2, const double2* tw1, const double2* wN, const VArgs& va, double dx,
                         double scale, const Limits& max_lds, hipStream_t s) {
  switch (n1) {
#define ATUS_CASE(n) \
  case n: return fs_n<n>(rm, mode, vk, nc, batch, psi, comp_stride, ycount, n2, tw1, wN, va, dx, scale, max_lds, s);
    ATUS_CASE(64) ATUS_CASE(128) ATUS_CASE(256) ATUS_CASE(512) ATUS_CASE(1024)
#undef ATUS_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace atus
