// launch.hpp -- host-side launchers of the kernel families (one translation
// unit each, so the template instantiations compile in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace atus {

// workgroup shape of one axis pass: T pencils of P = N/min(16,N) threads
struct PassCfg {
  dim3 grid;
  dim3 block;
  size_t lds;
  AxisGeo g;
};

// values per thread R = min(rm, N); threads per pencil P = N / R
inline int threads_per_pencil(int N, int rm) { return N / (N < rm ? N : rm); }

// what a launch may use: LDS bytes per workgroup, and threads per
// workgroup at most (0: as many as the kernel's launch bounds allow; the
// pencils per workgroup follow as threads / P, at least one pencil)
struct Limits {
  size_t lds;
  int threads;
};

// contiguous pencils if inner == 1; the pass covers `outer` slabs of
// [N][inner]; ycount = gridDim.y (components)
// ntables: lookup tables of length N the kernel stages in LDS (N <= 1024);
// table_elems >= 0 gives the LDS table elements exactly instead
PassCfg make_pass_cfg(int N, int rm, long long inner, long long outer, int ycount, size_t max_lds,
                      int ntables = 1, int max_threads = 256, long long table_elems = -1);
// the same within Limits (max_threads: the kernel's launch bounds)
inline PassCfg make_pass_cfg(int N, int rm, long long inner, long long outer, int ycount, const Limits& lim,
                             int ntables = 1, int max_threads = 256, long long table_elems = -1) {
  int t = max_threads;
  if (lim.threads > 0) {
    const int P = threads_per_pencil(N, rm);
    t = lim.threads < max_threads ? lim.threads : max_threads;
    if (t < P) t = P < max_threads ? P : max_threads;
  }
  return make_pass_cfg(N, rm, inner, outer, ycount, lim.lds, ntables, t, table_elems);
}

bool supported_len(int N);  // power of two in [2, 4096]

// raises the dynamic-LDS limit of `kernel` once when a launch needs more
// than the 64 KiB default (gfx950 allows 160 KiB per workgroup)
template <class K>
inline hipError_t allow_lds(K kernel, size_t bytes, bool& done) {
  if (bytes <= 65536 || done) return hipSuccess;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)bytes);
  done = e == hipSuccess;
  return e;
}

// rm: values per thread (8 or 16) of the FFT kernels
hipError_t launch_pass(int N, int rm, bool contig, int dir, const double2* src, double2* dst, long long comp_stride,
                       int ncomp, long long inner, long long outer, const double2* tw, double scale,
                       const Limits& max_lds, hipStream_t s);

// kinetic step along axis 0; vk/nc select the fused V prologue (1D only).
// Strided pencils (contig false): `inner` points apart, `rows` blocks of
// inner * N points (1 for axis 0, the planes of a slab for the middle axis)
hipError_t launch_kin(int N, int rm, bool contig, int vk, int nc, double2* psi, long long comp_stride, int ncomp,
                      long long inner, const double2* tw, const KinTabs& kt, const VArgs& va, double dx0,
                      const Limits& max_lds, hipStream_t s, long long rows = 1);

// inverse / V / forward along the last (contiguous) axis of a 2D/3D grid;
// edge (bit 0: before, bit 1: after the V step) adds this axis's half
// kinetic step with the phase table kz (k_vfuse's EDGE form)
hipError_t launch_vfuse(int N, int rm, int vk, int nc, double2* psi, long long comp_stride, long long outer,
                        const double2* tw, const VArgs& va, long long pencils_per_x0, int n0, double dx0,
                        const Limits& max_lds, hipStream_t s, const double2* kz = nullptr, int edge = 0);

// V step of a batch of independent two-component runs (VArgs::runs, ::Fm,
// ::pencils_per_run, ::run_shift set).  one_d: the fused 1D step over
// `count` runs (the kinetic phase of slot kt follows the V step); otherwise
// the contiguous-axis pass (inverse DFT, V, forward DFT) over `count`
// pencils of all runs.
// edge: as launch_vfuse (2D/3D only)
hipError_t launch_vstep_batch(int N, int rm, bool one_d, int vk, double2* psi, long long comp_stride, long long count,
                              const double2* tw, const KinTabs& kt, const VArgs& va, long long pencils_per_x0, int n0,
                              double dx0, const Limits& max_lds, hipStream_t s, const double2* kz = nullptr,
                              int edge = 0);

// a whole block of a 1D grid (N <= 4096) in one launch: one run (nc
// components, runs = 1) or a batch of `runs` two-component runs; eh / ef
// the half and full kinetic phase tables, t1tab [Nk] and Ftab ([Nk] or
// [Nk][runs]) the per-step t1 and amplitudes in device memory
hipError_t launch_block1d(int N, int rm, int vk, int nc, bool batch, double2* psi, long long comp_stride,
                          long long runs, const double2* tw, const double2* eh, const double2* ef, const VArgs& va,
                          const double* t1tab, const double* Ftab, double dx0, int Nk, const Limits& max_lds,
                          hipStream_t s);

// column pass of the four-step 1D transform (k_fs_col): n1 in 64..1024,
// mode FS_ENTER / FS_EXIT / FS_STEP; ycount components (NC == 1, no
// batch), runs (batch) or 1
hipError_t launch_fs_col(int n1, int rm, int mode, int vk, int nc, bool batch, double2* psi, long long comp_stride,
                         int ycount, long long n2, const double2* tw1, const double2* wN, const VArgs& va, double dx,
                         double scale, const Limits& max_lds, hipStream_t s);

// one pass of the Rabi integrals of a 3D box (k_rabi_z / _y / _x for axis
// 2 / 1 / 0) over nseg segments: src (axis 2) `seg` elements apart, the
// compact buffers zb [nseg][n0][n1][len2] and yb [nseg][n0][len1][len2];
// axis 0 writes nseg * (*nblocks) * RED_MAX_NV partials (partials NULL:
// only *nblocks is set)
hipError_t launch_rabi_box(int axis, int rm, const double2* src, long long seg, int nseg, double2* zb, double2* yb,
                           const double2* tw, const RabiBox& rb, const Limits& lim, double* partials,
                           unsigned* nblocks, hipStream_t s);

// Sobolev filters along axis 0
hipError_t launch_sobf(int N, int rm, bool contig, const double2* src, double2* out_sob, double2* out_lap,
                       long long comp_stride, int ncomp, long long inner, const double2* tw, const SobArgs& sa,
                       const Limits& max_lds, hipStream_t s);

// the Sobolev iteration's L2 gradient of component `comp` fused into the
// last-axis inverse (Laplacian) and forward (gradient filter) passes of a
// 2D/3D grid of nw (1 or 2) wave functions
hipError_t launch_sob_gradfuse(int N, int rm, int nw, const double2* lap, const double2* psi, const double* pot,
                               double2* out, long long npts, int comp, double g0, double g1, const double2* tw,
                               const Limits& max_lds, hipStream_t s);

// pointwise V step on all points (unfused path)
hipError_t launch_point(int vk, int nc, double2* psi, const Grid& gr, const VArgs& va, hipStream_t s);

}  // namespace atus
