// atus_abi.hip -- the extern "C" boundary of libatus_hip.so (include/atus_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/atus_hip.h"
#include "ctx.hpp"
#include "kernels_misc.hpp"

using namespace atus;
using atus_host::fail;
using atus_host::g_err;

thread_local std::string atus_host::g_err;

namespace {

// the scale() factor of cft_Nd (cft_1d.cpp:193, cft_2d.cpp:451, cft_3d.cpp:629)
double ft_factor(int dim, const double* s) {
  switch (dim) {
    case 1: return s[0] / sqrt(2.0 * M_PI);
    case 2: return 0.5 * s[0] * s[1] / M_PI;
    default: return s[0] * s[1] * s[2] / pow(2 * M_PI, 1.5);
  }
}

unsigned red_blocks(long long n) {
  long long b = (n + RED_THREADS - 1) / RED_THREADS;
  if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
  if (b < 1) b = 1;
  return (unsigned)b;
}

unsigned ew_blocks(long long n) {
  long long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

// Device scratch of atus_batch_particle_numbers_noalloc, the entry point for
// callers that cannot allocate device memory: the partial sums of its
// segmented reduction and their per-segment results.  Each GPU holds its
// own copy for the life of the process.  Host code never names the array
// itself (in host code the name is a stand-in, not a device address); it
// uses the per-device address hipGetSymbolAddress returns, taken once per
// device when the first context is created there.  Calls on one GPU share
// the array: calls that overlap in time on one GPU race.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = 32768;  // 256 KiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int MAX_SCRATCH_DEVICES = 64;
double* g_scratch_ptr[MAX_SCRATCH_DEVICES] = {};

// the address of `scratch` on the current device `dev`, looked up once
int ensure_scratch_symbol(int dev) {
  if (dev < 0 || dev >= MAX_SCRATCH_DEVICES) return fail(ATUS_ERR_UNSUPPORTED, "device index beyond 63");
  if (g_scratch_ptr[dev]) return ATUS_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(ATUS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = (double*)p;
  return ATUS_OK;
}

}  // namespace

namespace {

// launch limits of kernel family fam: the device's LDS per workgroup and
// the family's tuned threads per workgroup (atus_tune_threads)
Limits lim(const atus_ctx* c, int fam) { return Limits{c->max_lds, c->wg[fam]}; }

int check_ctx(atus_ctx* c) {
  if (!c) return fail(ATUS_ERR_ARG, "null context");
  return hipSetDevice(c->device) == hipSuccess ? ATUS_OK : fail(ATUS_ERR_HIP, "hipSetDevice failed");
}

// entry points written for one run of up to four coupled components
int check_single(atus_ctx* c, const char* what) {
  if (c->nbatch > 1)
    return fail(ATUS_ERR_STATE, std::string(what) + " acts on a single run; this context holds a batch of " +
                                    std::to_string(c->nbatch) + " runs (use the atus_batch_* entry points)");
  return ATUS_OK;
}

int check_comp(atus_ctx* c, int comp) {
  if (comp < 0 || comp >= c->ncomp)
    return fail(ATUS_ERR_ARG, "component " + std::to_string(comp) + " out of range [0," +
                                  std::to_string(c->ncomp) + ")");
  return ATUS_OK;
}

// brackets one launch with an event pair when the context is profiling
struct ProfScope {
  atus_ctx* c;
  bool on;
  hipStream_t st;
  // events on the stream the launch goes to (the context's stream unless
  // given)
  ProfScope(atus_ctx* cc, int fam, double frac = 1.0, hipStream_t s = nullptr)
      : c(cc), on(cc->prof), st(s ? s : cc->stream) {
    if (!on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      on = false;
      return;
    }
    (void)hipEventRecord(e, st);
    c->prof_ev.push_back(e);
    c->prof_fam.push_back(fam);
    c->prof_frac.push_back(frac);
  }
  ~ProfScope() {
    if (!on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, st);
    c->prof_ev.push_back(e);
  }
};

// twiddle tables of length N: W^j = e^{-2 pi i j / N} (j < N), then the
// stage-major tables of the plans of 8 and 16 values per thread (fft_tw):
// the same values, in the order the stages read them (lengths beyond the
// largest pencil, 4096, only serve as four-step factors: natural only)
int ensure_twiddles(atus_ctx* c, int N) {
  if (c->tw.count(N)) return ATUS_OK;
  const int ntab = N <= 4096 ? TW_TABLES : 1;
  std::vector<double2> h((size_t)ntab * N);
  for (int j = 0; j < N; ++j) {
    long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)N;
    h[j] = make_double2((double)cosl(a), (double)sinl(a));
  }
  for (int slot = 1; slot < ntab; ++slot) {
    const int rm = slot == 1 ? 8 : 16;
    const int R = N < rm ? N : rm;
    int A = 0, RA = 1;
    while (RA * R <= N) { RA *= R; ++A; }
    const int Q = N / RA;
    const int S = A + (Q > 1 ? 1 : 0);
    double2* t = h.data() + (size_t)slot * N;
    int off = 0, Ls = 1;
    for (int s = 0; s < S; ++s) {
      const int r = s < A ? R : Q;
      if (s >= 1) {
        const int step = N / (Ls * r);
        for (int m = 1; m < r; ++m)
          for (int k = 0; k < Ls; ++k) t[off + (m - 1) * Ls + k] = h[(size_t)k * m * step];
        off += (r - 1) * Ls;
      }
      Ls *= r;
    }
    if (off > N - 1) return fail(ATUS_ERR_UNSUPPORTED, "stage-major twiddle table overflow");
  }
  double2* d = nullptr;
  if (hipMalloc(&d, sizeof(double2) * h.size()) != hipSuccess) return fail(ATUS_ERR_NOMEM, "twiddle table allocation");
  HIP_TRY(hipMemcpy(d, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
  c->tw[N] = d;
  return ATUS_OK;
}

// per-axis kinetic phase tables for frac 0.5 (slot 0) and 1 (slot 1):
//   E_a[j] = exp(i * s * phi_a(j)),  phi_a = -dt * (k_a alpha_a) k_a
// (CRT_Base::Init, CRT_Base.h:253-274, factorised over the axes); axis 0
// carries the ft(-1) * ft(+1) scale.  Slot f = 0: half step, 1: full step,
// 2: half step for a transform pair along one axis alone (every axis's
// table carries that axis's share dx dk / 2pi of the scale).
void kin_tables(const atus_ctx* c, const double* alpha, double dt, std::vector<double2>& h) {
  const int nm = c->kin_nmax;
  h.assign((size_t)atus_ctx::KIN_SLOTS_N * 3 * nm, make_double2(1.0, 0.0));
  const double scale = ft_factor(c->dim, c->dx) * ft_factor(c->dim, c->dk);
  const double mdt = -dt;
  for (int f = 0; f < atus_ctx::KIN_SLOTS_N; ++f) {
    for (int a = 0; a < c->dim; ++a) {
      const int N = c->n[a], sh = N / 2;
      const double sa = 0.5 * c->dx[a] * c->dk[a] / M_PI;
      for (int j = 0; j < N; ++j) {
        const double k = c->dk[a] * (double)(((j + sh) % N) - sh);
        const double phi = mdt * ((k * alpha[a]) * k);
        const double ph = (f == 1) ? phi : 0.5 * phi;
        double2 e = make_double2(cos(ph), sin(ph));
        if (f == 2) e = make_double2(e.x * sa, e.y * sa);
        else if (a == 0) e = make_double2(e.x * scale, e.y * scale);
        h[((size_t)f * 3 + a) * nm + j] = e;
      }
    }
  }
}

int ensure_kin(atus_ctx* c, double dt) {
  if (c->kin_valid && c->kin_dt == dt) return ATUS_OK;
  const int nm = c->kin_nmax;
  std::vector<double2> h, h2;
  kin_tables(c, c->alpha, dt, h);
  HIP_TRY(hipMemcpyAsync(c->kin, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, c->stream));
  if (c->alpha2_set) {
    // CRT_Base_2::Init (CRT_Base_2.h:195-221): the tables of Alpha_2
    if (!c->kin2 && hipMalloc(&c->kin2, sizeof(double2) * atus_ctx::KIN_SLOTS_N * 3 * nm) != hipSuccess)
      return fail(ATUS_ERR_NOMEM, "second-species tables");
    kin_tables(c, c->alpha2, dt, h2);
    HIP_TRY(hipMemcpyAsync(c->kin2, h2.data(), sizeof(double2) * h2.size(), hipMemcpyHostToDevice, c->stream));
  }
  std::vector<double2> hz;
  if (c->fs) {
    // four-step spectrum: row k1, column k2 holds frequency k1 + N1 k2
    const long long N = c->npts, n1 = c->fs_n1, n2 = c->fs_n2;
    hz.resize((size_t)2 * N);
    for (int f = 0; f < 2; ++f)
      for (long long k1 = 0; k1 < n1; ++k1)
        for (long long k2 = 0; k2 < n2; ++k2) hz[(size_t)f * N + k1 * n2 + k2] = h[((size_t)f * 3) * nm + k1 + n1 * k2];
    HIP_TRY(hipMemcpyAsync(c->kinz, hz.data(), sizeof(double2) * hz.size(), hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->kin_dt = dt;
  c->kin_valid = true;
  return ATUS_OK;
}

KinTabs kin_tabs(atus_ctx* c, int f) {
  KinTabs kt{};
  const int nm = c->kin_nmax;
  kt.e0 = c->kin + ((size_t)f * 3 + 0) * nm;
  kt.e1 = c->dim > 1 ? c->kin + ((size_t)f * 3 + 1) * nm : nullptr;
  kt.e2 = c->dim > 2 ? c->kin + ((size_t)f * 3 + 2) * nm : nullptr;
  kt.n2 = c->dim > 2 ? c->n[2] : 1;
  if (c->fs) {
    kt.ez = c->kinz + (size_t)f * c->npts;
    kt.ez_rows = c->fs_n1;
  }
  if (c->alpha2_set && c->kin2) {
    kt.f0 = c->kin2 + ((size_t)f * 3 + 0) * nm;
    kt.f1 = c->dim > 1 ? c->kin2 + ((size_t)f * 3 + 1) * nm : nullptr;
    kt.f2 = c->dim > 2 ? c->kin2 + ((size_t)f * 3 + 2) * nm : nullptr;
    kt.split = c->ncomp / 2;
  }
  return kt;
}

// the half step of axis a alone (slot 2): the pencil's own axis only
KinTabs axis_tabs(atus_ctx* c, int a) {
  KinTabs kt{};
  kt.e0 = c->kin + ((size_t)2 * 3 + a) * c->kin_nmax;
  kt.n2 = 1;
  return kt;
}

int ensure_scratch(atus_ctx* c) {
  if (c->scratch) return ATUS_OK;
  if (hipMalloc(&c->scratch, sizeof(double2) * c->npts) != hipSuccess)
    return fail(ATUS_ERR_NOMEM, "scratch allocation");
  return ATUS_OK;
}

// one pass along axis a of `ncomp` components starting at src/dst, `cs`
// elements apart (0: npts, i.e. consecutive components)
int pass(atus_ctx* c, int a, int dir, const double2* src, double2* dst, int ncomp, double scale, long long cs = 0) {
  const int N = c->n[a];
  ATUS_TRY(ensure_twiddles(c, N));
  const long long inner = c->inner_of(a), outer = c->outer_of(a);
  const int fam = inner == 1 ? ATUS_KFAM_PASS_CONTIG : ATUS_KFAM_PASS_STRIDED;
  ProfScope ps(c, fam);
  HIP_TRY(launch_pass(N, c->rm[fam], inner == 1, dir, src, dst, cs ? cs : c->npts, ncomp, inner, outer, c->tw[N], scale,
                      lim(c, fam), c->stream));
  return ATUS_OK;
}

// one pass along axis a (1 or 2) of a 3D grid over the x-planes
// [s0, s0 + S) of all components, in place
int pass_slab(atus_ctx* c, int a, int dir, int s0, int S, hipStream_t st) {
  const int N = c->n[a];
  ATUS_TRY(ensure_twiddles(c, N));
  const long long plane = (long long)c->n[1] * c->n[2];
  const long long inner = c->inner_of(a), outer = c->outer_of(a) / c->n[0] * S;
  const int fam = inner == 1 ? ATUS_KFAM_PASS_CONTIG : ATUS_KFAM_PASS_STRIDED;
  double2* p = c->psi + (long long)s0 * plane;
  ProfScope ps(c, fam, (double)S / c->n[0], st);
  HIP_TRY(launch_pass(N, c->rm[fam], inner == 1, dir, p, p, c->npts, c->ncomp, inner, outer, c->tw[N], 1.0, lim(c, fam),
                      st));
  return ATUS_OK;
}

// the fused inverse-z / V / forward-z pass over the x-planes [s0, s0 + S):
// the kernel's x index counts from the slab, so n0 is shifted to keep
// x0 = (i0 - n0/2) dx0 (n0 even), and the potential pointer moves along
// edge: 1 / 2 / 3 adds z's share of the half kinetic step before / after /
// on both sides of the V step (the first and last steps of a block)
int vfuse_slab(atus_ctx* c, int vk, const VArgs& va, int s0, int S, hipStream_t st, int edge = 0) {
  const int NL = c->n[2];
  const long long plane = (long long)c->n[1] * c->n[2];
  VArgs vs = va;
  if (vs.pot) vs.pot += (long long)s0 * plane;
  ProfScope ps(c, ATUS_KFAM_VFUSE, (double)S / c->n[0], st);
  HIP_TRY(launch_vfuse(NL, c->rm[ATUS_KFAM_VFUSE], vk, c->ncomp, c->psi + (long long)s0 * plane, c->npts,
                       (long long)S * c->n[1], c->tw[NL], vs, c->n[1], c->n[0] - 2 * s0, c->dx[0], lim(c, ATUS_KFAM_VFUSE),
                       st, edge ? axis_tabs(c, 2).e0 : nullptr, edge));
  return ATUS_OK;
}

// y's share of the half kinetic step (forward DFT, phase, inverse DFT along
// axis 1) over the x-planes [s0, s0 + S) of all components.  It takes the
// place of a strided y pass in the slab chain, beside the other stream's
// launches, so it runs (and is timed and tuned) as that family.
int kin_y_slab(atus_ctx* c, const VArgs& va, int s0, int S, hipStream_t st) {
  const int N = c->n[1];
  ATUS_TRY(ensure_twiddles(c, N));
  const long long plane = (long long)c->n[1] * c->n[2];
  ProfScope ps(c, ATUS_KFAM_PASS_STRIDED, (double)S / c->n[0], st);
  HIP_TRY(launch_kin(N, c->rm[ATUS_KFAM_PASS_STRIDED], false, V_NONE, 1, c->psi + (long long)s0 * plane, c->npts,
                     c->ncomp, c->n[2], c->tw[N], axis_tabs(c, 1), va, c->dx[0], lim(c, ATUS_KFAM_PASS_STRIDED), st,
                     S));
  return ATUS_OK;
}

// four-step column pass (FS_ENTER / FS_EXIT: ycount components cs apart;
// FS_STEP: V step of nc components, or of every run of a batch)
int fs_col(atus_ctx* c, int mode, double2* p, int ycount, long long cs, int vk, int nc, bool batch, const VArgs& va,
           double scale) {
  ProfScope ps(c, mode == FS_STEP ? ATUS_KFAM_VFUSE : ATUS_KFAM_PASS_STRIDED);
  HIP_TRY(launch_fs_col(c->fs_n1, c->rm[mode == FS_STEP ? ATUS_KFAM_VFUSE : ATUS_KFAM_PASS_STRIDED], mode, vk, nc, batch,
                        p, cs, ycount, c->fs_n2, c->tw[c->fs_n1], c->tw[(int)c->npts], va, c->dx[0], scale,
                        lim(c, mode == FS_STEP ? ATUS_KFAM_VFUSE : ATUS_KFAM_PASS_STRIDED), c->stream));
  return ATUS_OK;
}

// four-step row pass of ncomp components: kinetic phase of slot f
int fs_row_kin(atus_ctx* c, int f, const VArgs& va) {
  ProfScope ps(c, ATUS_KFAM_KIN);
  HIP_TRY(launch_kin(c->fs_n2, c->rm[ATUS_KFAM_KIN], true, V_NONE, 1, c->psi, c->npts, c->ncomp, 1, c->tw[c->fs_n2],
                     kin_tabs(c, f), va, c->dx[0], lim(c, ATUS_KFAM_KIN), c->stream, c->fs_n1));
  return ATUS_OK;
}

int transpose(atus_ctx* c, const double2* src, double2* dst, long long rows, long long cols, int ncomp, long long cs,
              double scale) {
  const long long tiles = ((rows + 31) / 32) * ((cols + 31) / 32) * ncomp;
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)tiles), dim3(256), 0, c->stream, src, dst, rows, cols, cs, scale);
  HIP_TRY(hipGetLastError());
  return ATUS_OK;
}

int ensure_tbuf(atus_ctx* c) {
  if (c->tbuf) return ATUS_OK;
  if (hipMalloc(&c->tbuf, sizeof(double2) * c->npts * c->ncomp) != hipSuccess)
    return fail(ATUS_ERR_NOMEM, "transpose buffer");
  return ATUS_OK;
}

// full forward (dir -1) or backward (+1) transform with the cft scale,
// src -> dst (dst may equal src); axes from fastest to slowest
int full_ft(atus_ctx* c, int dir, const double2* src, double2* dst, int ncomp, long long cs = 0) {
  const double scale = ft_factor(c->dim, dir < 0 ? c->dx : c->dk);
  if (c->fs) {
    // natural order in and out: [N1][N2] spectrum <-> frequency k1 + N1 k2
    ATUS_TRY(ensure_tbuf(c));
    const long long s_ = cs ? cs : c->npts;
    // tbuf mirrors psi: dst's offset when dst lies in psi, else its start
    const uintptr_t d = (uintptr_t)dst, p0 = (uintptr_t)c->psi;
    const uintptr_t p1 = p0 + sizeof(double2) * (uintptr_t)(c->npts * c->ncomp);
    double2* t = c->tbuf + ((d >= p0 && d < p1) ? (long long)((d - p0) / sizeof(double2)) : 0);
    VArgs none{};
    if (dir < 0) {
      if (dst != src) HIP_TRY(hipMemcpy2DAsync(dst, s_ * sizeof(double2), src, s_ * sizeof(double2), c->npts * sizeof(double2),
                                               ncomp, hipMemcpyDeviceToDevice, c->stream));
      ATUS_TRY(fs_col(c, FS_ENTER, dst, ncomp, s_, V_NONE, 1, false, none, 1.0));
      {
        ProfScope ps(c, ATUS_KFAM_PASS_CONTIG);
        HIP_TRY(launch_pass(c->fs_n2, c->rm[ATUS_KFAM_PASS_CONTIG], true, -1, dst, dst, s_, ncomp, 1, c->fs_n1,
                            c->tw[c->fs_n2], 1.0, lim(c, ATUS_KFAM_PASS_CONTIG), c->stream));
      }
      ATUS_TRY(transpose(c, dst, t, c->fs_n1, c->fs_n2, ncomp, s_, scale));
      HIP_TRY(hipMemcpy2DAsync(dst, s_ * sizeof(double2), t, s_ * sizeof(double2), c->npts * sizeof(double2), ncomp,
                               hipMemcpyDeviceToDevice, c->stream));
    } else {
      ATUS_TRY(transpose(c, src, t, c->fs_n2, c->fs_n1, ncomp, s_, scale));
      {
        ProfScope ps(c, ATUS_KFAM_PASS_CONTIG);
        HIP_TRY(launch_pass(c->fs_n2, c->rm[ATUS_KFAM_PASS_CONTIG], true, +1, t, t, s_, ncomp, 1, c->fs_n1,
                            c->tw[c->fs_n2], 1.0, lim(c, ATUS_KFAM_PASS_CONTIG), c->stream));
      }
      ATUS_TRY(fs_col(c, FS_EXIT, t, ncomp, s_, V_NONE, 1, false, none, 1.0));
      HIP_TRY(hipMemcpy2DAsync(dst, s_ * sizeof(double2), t, s_ * sizeof(double2), c->npts * sizeof(double2), ncomp,
                               hipMemcpyDeviceToDevice, c->stream));
    }
    return ATUS_OK;
  }
  const double2* s = src;
  for (int a = c->dim - 1; a >= 0; --a) {
    ATUS_TRY(pass(c, a, dir, s, dst, ncomp, a == 0 ? scale : 1.0, cs));
    s = dst;
  }
  return ATUS_OK;
}

// kinetic step on all components with table slot f (0 half, 1 full)
int kinetic_unfused(atus_ctx* c, int f) {
  if (c->fs) {
    VArgs none{};
    ATUS_TRY(fs_col(c, FS_ENTER, c->psi, c->ncomp, c->npts, V_NONE, 1, false, none, 1.0));
    ATUS_TRY(fs_row_kin(c, f, none));
    ATUS_TRY(fs_col(c, FS_EXIT, c->psi, c->ncomp, c->npts, V_NONE, 1, false, none, 1.0));
    return ATUS_OK;
  }
  // forward along the fast axes, fused phase along axis 0, inverse back
  for (int a = c->dim - 1; a >= 1; --a) ATUS_TRY(pass(c, a, -1, c->psi, c->psi, c->ncomp, 1.0));
  const int N0 = c->n[0];
  ATUS_TRY(ensure_twiddles(c, N0));
  VArgs va{};
  {
    ProfScope ps(c, ATUS_KFAM_KIN);
    HIP_TRY(launch_kin(N0, c->rm[ATUS_KFAM_KIN], c->dim == 1, V_NONE, 1, c->psi, c->npts, c->ncomp, c->inner_of(0), c->tw[N0],
                       kin_tabs(c, f), va, c->dx[0], lim(c, ATUS_KFAM_KIN), c->stream));
  }
  for (int a = 1; a < c->dim; ++a) ATUS_TRY(pass(c, a, +1, c->psi, c->psi, c->ncomp, 1.0));
  return ATUS_OK;
}

template <class Fn>
int reduce(atus_ctx* c, const Fn& fn, long long n, double* out) {
  constexpr int NV = Fn::NV;
  const unsigned nb = red_blocks(n);
  hipLaunchKernelGGL((k_reduce<Fn>), dim3(nb), dim3(RED_THREADS), 0, c->stream, fn, n, c->red_part);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL((k_reduce_final<NV>), dim3(1), dim3(RED_THREADS), 0, c->stream, c->red_part, (int)nb,
                     c->red_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(c->red_host, c->red_out, sizeof(double) * NV, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int v = 0; v < NV; ++v) out[v] = c->red_host[v];
  return ATUS_OK;
}

// workgroups per segment of a segmented reduction: about 2048 over all
// segments, at least one per segment
unsigned seg_blocks(long long n, int nseg) {
  unsigned nb = red_blocks(n);
  const unsigned want = (unsigned)std::max(1, 2048 / nseg);
  return nb > want ? want : nb;
}

// the two launches of a segmented reduction into caller-given device
// buffers: `part` holds nseg * nb * NV partials, `res` nseg * NV results
// (no host copy: capturable into a graph)
template <class Fn>
int reduce_seg_launch(atus_ctx* c, const Fn& fn, long long n, long long seg_stride, int nseg, unsigned nb, double* part,
                      double* res) {
  constexpr int NV = Fn::NV;
  hipLaunchKernelGGL((k_reduce_seg<Fn>), dim3(nb, (unsigned)nseg), dim3(RED_THREADS), 0, c->stream, fn, n, seg_stride,
                     part);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL((k_reduce_seg_final<NV>), dim3((unsigned)nseg), dim3(RED_THREADS), 0, c->stream, part, (int)nb,
                     res);
  HIP_TRY(hipGetLastError());
  return ATUS_OK;
}

// the launches of reduce_seg with caller-given device buffers: `part`
// holds nseg * nb * NV partials, `res` nseg * NV results; out (host)
// receives res
template <class Fn>
int reduce_seg_into(atus_ctx* c, const Fn& fn, long long n, long long seg_stride, int nseg, unsigned nb, double* part,
                    double* res, double* out) {
  constexpr int NV = Fn::NV;
  ATUS_TRY(reduce_seg_launch(c, fn, n, seg_stride, nseg, nb, part, res));
  HIP_TRY(hipMemcpyAsync(out, res, sizeof(double) * (size_t)nseg * NV, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

// room for `need` partials and nseg * RED_MAX_NV results in the
// segmented-reduction buffers (allocated outside any graph capture)
int ensure_sred(atus_ctx* c, size_t need, int nseg) {
  if (c->sred_cap >= need && c->sred_nseg >= nseg) return ATUS_OK;
  HIP_TRY(hipStreamSynchronize(c->stream));
  need = std::max(need, c->sred_cap);
  nseg = std::max(nseg, c->sred_nseg);
  (void)hipFree(c->sred_part);
  (void)hipFree(c->sred_out);
  c->sred_part = c->sred_out = nullptr;
  c->sred_cap = 0;
  c->sred_nseg = 0;
  if (hipMalloc(&c->sred_part, sizeof(double) * need) != hipSuccess ||
      hipMalloc(&c->sred_out, sizeof(double) * (size_t)nseg * RED_MAX_NV) != hipSuccess)
    return fail(ATUS_ERR_NOMEM, "segmented reduction buffers");
  c->sred_cap = need;
  c->sred_nseg = nseg;
  return ATUS_OK;
}

// nseg reductions of n elements each, segment y starting y * seg_stride
// elements after the functor's data: out[y * NV + v]
template <class Fn>
int reduce_seg(atus_ctx* c, const Fn& fn, long long n, long long seg_stride, int nseg, double* out) {
  constexpr int NV = Fn::NV;
  const unsigned nb = seg_blocks(n, nseg);
  ATUS_TRY(ensure_sred(c, (size_t)nseg * nb * NV, nseg));
  return reduce_seg_into(c, fn, n, seg_stride, nseg, nb, c->sred_part, c->sred_out, out);
}

// The Rabi integrals' reduction (RedRabi): over the index box of the balls
// |k - k0_q| < thr when that box is under half the grid (RedRabiBox, the
// same sums over far fewer points), else over every point.  Per axis the
// box spans the centred indices m with |dk m - k0_a| < thr for some q, one
// index of margin on each side.
struct RabiReduction {
  RedRabi full{};
  RedRabiBox box{};
  bool use_box = false;
  long long npts = 0;  // points the reduction visits
};

RabiReduction rabi_reduction(const atus_ctx* c, const double2* psik, const double* k0, int n, double thr) {
  RabiReduction r;
  r.full.psik = psik;
  r.full.gr = c->grid();
  r.full.nk0 = n;
  for (int q = 0; q < n; ++q)
    for (int a = 0; a < 3; ++a) r.full.k0[q][a] = k0[3 * q + a];
  r.full.thr = thr;
  r.npts = c->npts;
  long long vol = 1;
  bool ok = std::isfinite(thr) && thr > 0;
  for (int a = 0; a < 3 && ok; ++a) {
    if (a >= c->dim) {
      r.box.lo[a] = 0;
      r.box.len[a] = 1;
      continue;
    }
    const long long N = c->n[a], s = N / 2;
    double lo = INFINITY, hi = -INFINITY;
    for (int q = 0; q < n; ++q) {
      lo = std::min(lo, std::floor((k0[3 * q + a] - thr) / c->dk[a]) - 1.0);
      hi = std::max(hi, std::ceil((k0[3 * q + a] + thr) / c->dk[a]) + 1.0);
    }
    if (!(lo <= hi)) {
      ok = false;
      break;
    }
    const long long mlo = (long long)std::max(lo, (double)-s), mhi = (long long)std::min(hi, (double)(N - s - 1));
    if (mhi < mlo) {
      // no grid point near any ball along this axis: the plain reduction
      // (all sums zero)
      ok = false;
      break;
    }
    r.box.lo[a] = (int)mlo;
    r.box.len[a] = (int)(mhi - mlo + 1);
    vol *= r.box.len[a];
  }
  if (ok && vol > 0 && vol * 2 < c->npts && vol < (1LL << 31)) {
    r.use_box = true;
    r.box.psik = psik;
    r.box.gr = r.full.gr;
    r.box.nk0 = n;
    for (int q = 0; q < n; ++q)
      for (int a = 0; a < 3; ++a) r.box.k0[q][a] = r.full.k0[q][a];
    r.box.thr = thr;
    r.npts = vol;
  }
  return r;
}

// compact-spectrum buffers of the 3D box Rabi integrals over nseg segments
size_t rabi_box_elems(const atus_ctx* c, const RabiReduction& rr, int nseg) {
  const long long n0 = c->n[0], n1 = c->n[1];
  const long long l1 = rr.box.len[1], l2 = rr.box.len[2];
  return (size_t)nseg * (size_t)(n0 * n1 * l2 + n0 * l1 * l2);
}

// whether the Rabi integrals of this context take the 3D box passes
bool rabi_box3_ok(const atus_ctx* c, const RabiReduction& rr) { return c->dim == 3 && !c->fs && rr.use_box; }

// the buffers of rabi_box3 (allocated outside any graph capture)
int ensure_rabi_box(atus_ctx* c, const RabiReduction& rr, int nseg) {
  const size_t need = rabi_box_elems(c, rr, nseg);
  if (c->rbuf_cap < need) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(c->rbuf);
    c->rbuf = nullptr;
    c->rbuf_cap = 0;
    if (hipMalloc(&c->rbuf, sizeof(double2) * need) != hipSuccess) return fail(ATUS_ERR_NOMEM, "Rabi spectrum buffers");
    c->rbuf_cap = need;
  }
  for (int a = 0; a < 3; ++a) ATUS_TRY(ensure_twiddles(c, c->n[a]));
  RabiBox rb{};
  for (int a = 0; a < 3; ++a) rb.n[a] = c->n[a];
  unsigned nb = 0;
  rb.len[1] = rr.box.len[1];
  rb.len[2] = rr.box.len[2];
  HIP_TRY(launch_rabi_box(0, c->rm[ATUS_KFAM_PASS_STRIDED], nullptr, 0, nseg, nullptr, nullptr, c->tw[c->n[0]], rb,
                          lim(c, ATUS_KFAM_PASS_STRIDED), nullptr, &nb, c->stream));
  return ensure_sred(c, (size_t)nseg * nb * RED_MAX_NV, nseg);
}

// compute_rabi_integrals of nseg segments (`seg` elements apart from src)
// of a 3D grid from the spectrum inside the balls' index box: the three
// k_rabi passes and the segmented final sum into res (device, nseg *
// RED_MAX_NV, before the factor ar_k); src is not modified.  The buffers
// come from ensure_rabi_box.
int rabi_box3(atus_ctx* c, const double2* src, long long seg, int nseg, const RabiReduction& rr, double* res) {
  if (c->rbuf_cap < rabi_box_elems(c, rr, nseg)) return fail(ATUS_ERR_STATE, "Rabi spectrum buffers not sized");
  RabiBox rb{};
  for (int a = 0; a < 3; ++a) {
    rb.n[a] = c->n[a];
    rb.lo[a] = rr.box.lo[a];
    rb.len[a] = rr.box.len[a];
    rb.dk[a] = c->dk[a];
  }
  rb.nk0 = rr.box.nk0;
  for (int q = 0; q < RED_MAX_NV; ++q)
    for (int a = 0; a < 3; ++a) rb.k0[q][a] = rr.box.k0[q][a];
  rb.thr = rr.box.thr;
  rb.scale = ft_factor(c->dim, c->dx);
  double2* zb = c->rbuf;
  double2* yb = c->rbuf + (size_t)nseg * c->n[0] * c->n[1] * rb.len[2];
  unsigned nb = 0;
  HIP_TRY(launch_rabi_box(2, c->rm[ATUS_KFAM_PASS_CONTIG], src, seg, nseg, zb, yb, c->tw[c->n[2]], rb,
                          lim(c, ATUS_KFAM_PASS_CONTIG), nullptr, nullptr, c->stream));
  HIP_TRY(launch_rabi_box(1, c->rm[ATUS_KFAM_PASS_STRIDED], src, seg, nseg, zb, yb, c->tw[c->n[1]], rb,
                          lim(c, ATUS_KFAM_PASS_STRIDED), nullptr, nullptr, c->stream));
  HIP_TRY(launch_rabi_box(0, c->rm[ATUS_KFAM_PASS_STRIDED], src, seg, nseg, zb, yb, c->tw[c->n[0]], rb,
                          lim(c, ATUS_KFAM_PASS_STRIDED), nullptr, &nb, c->stream));
  if (c->sred_cap < (size_t)nseg * nb * RED_MAX_NV) return fail(ATUS_ERR_STATE, "reduction buffers not sized");
  HIP_TRY(launch_rabi_box(0, c->rm[ATUS_KFAM_PASS_STRIDED], src, seg, nseg, zb, yb, c->tw[c->n[0]], rb,
                          lim(c, ATUS_KFAM_PASS_STRIDED), c->sred_part, &nb, c->stream));
  hipLaunchKernelGGL((k_reduce_seg_final<RED_MAX_NV>), dim3((unsigned)nseg), dim3(RED_THREADS), 0, c->stream,
                     c->sred_part, (int)nb, res);
  HIP_TRY(hipGetLastError());
  return ATUS_OK;
}

int vkind_of(int kind) {
  switch (kind) {
    case ATUS_STEP_FREEPROP: return V_NL;
    case ATUS_STEP_FREEPROP_IF: return V_NL_IF;
    case ATUS_STEP_BRAGG_AD: return V_BRAGG;
    case ATUS_STEP_DOUBLE_BRAGG: return V_DBRAGG;
    case ATUS_STEP_NUM_BRAGG: return V_NUM_BRAGG;
    case ATUS_STEP_RAMAN: return V_RAMAN;
    default: return -1;
  }
}

// the components each coupled step acts on
int check_vk_ncomp(const atus_ctx* c, int vk) {
  const int n = c->ncomp;
  if (vk == V_BRAGG && n != 2) return fail(ATUS_ERR_STATE, "bragg_ad needs exactly two components");
  if (vk == V_DBRAGG && n != 3) return fail(ATUS_ERR_STATE, "the double Bragg step needs three components");
  if (vk == V_NUM_BRAGG && n != 2 && n != 3)
    return fail(ATUS_ERR_STATE, "the numerical Bragg step needs two or three components");
  if (vk == V_RAMAN && n != 3 && n != 4) return fail(ATUS_ERR_STATE, "the Raman step needs three or four components");
  return ATUS_OK;
}

void fill_vargs(atus_ctx* c, VArgs& va, double dt, const atus_laser* L) {
  va = VArgs{};
  va.dt = dt;
  for (int i = 0; i < 16; ++i) va.gs[i] = c->gs[i];
  va.pot = c->pot;
  va.pot_stride = c->npts;
  if (L) {
    va.beta0 = L->beta0;
    for (int i = 0; i < 4; ++i) va.deltaL[i] = L->deltaL[i];
    va.amp0 = L->amp0;
    va.laser_k = L->laser_k;
    va.laser_domh = L->laser_domh;
    va.laser_dk = L->laser_dk;
    va.chirp = L->chirp;
    va.chirp_rate0 = L->chirp_rate0;
    va.phase0 = L->phase0;
    va.amp1 = L->amp1;
    for (int a = 0; a < 3; ++a) va.beta[a] = a < c->dim ? L->beta[a] : 0.0;
  }
  va.F = 1.0;
}

}  // namespace

extern "C" {

const char* atus_last_error(void) { return g_err.c_str(); }

const char* atus_version(void) { return "atus_hip 0.1 (gfx950, complex128)"; }

int atus_device_info(char* name, int len, int* cu_count, double* hbm_gbps_peak) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, dev));
  if (name && len > 0) {
    snprintf(name, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (hbm_gbps_peak) *hbm_gbps_peak = 2.0 * (double)p.memoryClockRate * 1e3 * (p.memoryBusWidth / 8) / 1e9;
  return ATUS_OK;
}

}  // extern "C"

namespace {

int ctx_create(atus_ctx** out, int dim, const long long n[3], const double dx[3], const double dk[3], int nc_run,
               int nbatch, bool batch) {
  if (!out) return fail(ATUS_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (dim < 1 || dim > 3) return fail(ATUS_ERR_ARG, "dim must be 1, 2 or 3");
  if (nc_run < 1 || nc_run > 4) return fail(ATUS_ERR_ARG, "ncomp must be 1..4");
  // every component is one blockIdx.y of the per-component passes (<= 65535)
  if (nbatch < 1 || nbatch > 16384) return fail(ATUS_ERR_ARG, "nbatch must be 1..16384");
  const int ncomp = nc_run * nbatch;
  for (int a = 0; a < 3; ++a) {
    if (a < dim) {
      const bool four_step = dim == 1 && n[a] > 4096 && n[a] <= (1 << 20) && (n[a] & (n[a] - 1)) == 0;
      if (!four_step && (n[a] > 4096 || !supported_len((int)n[a])))
        return fail(ATUS_ERR_UNSUPPORTED, "axis " + std::to_string(a) + " length " + std::to_string(n[a]) +
                                              " is not a power of two in [2, 4096] (1D: up to 2^20)");
    } else if (n[a] != 1) {
      return fail(ATUS_ERR_ARG, "unused axes must have length 1");
    }
  }
  atus_ctx* c = new atus_ctx();
  c->dim = dim;
  c->ncomp = ncomp;
  c->nc_run = nc_run;
  c->nbatch = nbatch;
  c->npts = 1;
  for (int a = 0; a < 3; ++a) {
    c->n[a] = (int)n[a];
    c->dx[a] = dx[a];
    c->dk[a] = dk[a];
    c->npts *= n[a];
    c->kin_nmax = std::max(c->kin_nmax, (int)n[a]);
  }
  auto bail = [&](int code, const std::string& m) {
    atus_ctx_destroy(c);
    return fail(code, m);
  };
  if (dim == 1 && c->n[0] > 4096) {
    // N1 = 2^floor(log2(N)/2) <= N2 = N / N1, both in 64..1024
    int l = 0;
    while ((1 << (l + 1)) <= c->n[0]) ++l;
    c->fs = true;
    c->fs_n1 = 1 << (l / 2);
    c->fs_n2 = c->n[0] / c->fs_n1;
  }
  if (hipGetDevice(&c->device) != hipSuccess) return bail(ATUS_ERR_HIP, "no HIP device");
  if (ensure_scratch_symbol(c->device) != ATUS_OK) {
    std::string m = g_err;
    return bail(ATUS_ERR_HIP, m);
  }
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) == hipSuccess && lds > 0)
    c->max_lds = std::min<size_t>((size_t)lds, 163840);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(ATUS_ERR_HIP, "stream creation");
  if (batch && hipMalloc(&c->runs, sizeof(RunParams) * nbatch) != hipSuccess)
    return bail(ATUS_ERR_NOMEM, "per-run parameters");
  if (hipMalloc(&c->psi, sizeof(double2) * c->npts * ncomp) != hipSuccess)
    return bail(ATUS_ERR_NOMEM, "wavefunction allocation (" + std::to_string(sizeof(double2) * c->npts * ncomp) +
                                    " bytes)");
  if (hipMemset(c->psi, 0, sizeof(double2) * c->npts * ncomp) != hipSuccess) return bail(ATUS_ERR_HIP, "memset");
  if (hipMalloc(&c->kin, sizeof(double2) * atus_ctx::KIN_SLOTS_N * 3 * c->kin_nmax) != hipSuccess)
    return bail(ATUS_ERR_NOMEM, "tables");
  if (c->fs && (hipMalloc(&c->kinz, sizeof(double2) * 2 * c->npts) != hipSuccess ||
                ensure_twiddles(c, c->fs_n1) != ATUS_OK || ensure_twiddles(c, c->fs_n2) != ATUS_OK))
    return bail(ATUS_ERR_NOMEM, "four-step tables");
  if (hipMalloc(&c->red_part, sizeof(double) * RED_MAX_BLOCKS * RED_MAX_NV) != hipSuccess ||
      hipMalloc(&c->red_out, sizeof(double) * RED_MAX_NV) != hipSuccess ||
      hipHostMalloc(&c->red_host, sizeof(double) * RED_MAX_NV, hipHostMallocDefault) != hipSuccess)
    return bail(ATUS_ERR_NOMEM, "reduction buffers");
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return bail(ATUS_ERR_HIP, "events");
  for (int a = 0; a < dim; ++a)
    if (ensure_twiddles(c, c->n[a]) != ATUS_OK) {
      std::string m = g_err;
      return bail(ATUS_ERR_NOMEM, m);
    }
  *out = c;
  return ATUS_OK;
}

}  // namespace

extern "C" {

int atus_ctx_create(atus_ctx** out, int dim, const long long n[3], const double dx[3], const double dk[3],
                    int ncomp) {
  return ctx_create(out, dim, n, dx, dk, ncomp, 1, false);
}

int atus_ctx_destroy(atus_ctx* c) {
  if (!c) return ATUS_OK;
  if (c->live_graphs > 0)
    return fail(ATUS_ERR_STATE, "context still has " + std::to_string(c->live_graphs) +
                                    " captured graph(s); destroy them first (atus_graph_destroy)");
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->psi);
  for (double2* p : c->snap) (void)hipFree(p);
  (void)hipFree(c->pot);
  (void)hipFree(c->scratch);
  for (auto& kv : c->tw) (void)hipFree(kv.second);
  (void)hipFree(c->kin);
  (void)hipFree(c->red_part);
  (void)hipFree(c->red_out);
  if (c->red_host) (void)hipHostFree(c->red_host);
  (void)hipFree(c->sob);
  (void)hipFree(c->lap);
  (void)hipFree(c->grad);
  (void)hipFree(c->sob_st);
  (void)hipFree(c->sob_part);
  (void)hipFree(c->sob_log);
  (void)hipFree(c->runs);
  (void)hipFree(c->ftab);
  (void)hipFree(c->kinz);
  (void)hipFree(c->tbuf);
  (void)hipFree(c->kin2);
  (void)hipFree(c->sred_part);
  (void)hipFree(c->sred_out);
  (void)hipFree(c->rbuf);
  (void)hipFree(c->blk_dev);
  if (c->blk_host) (void)hipHostFree(c->blk_host);
  if (c->blk_ev) (void)hipEventDestroy(c->blk_ev);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (hipStream_t x : c->xstream)
    if (x) (void)hipStreamSynchronize(x);
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  for (hipEvent_t e : c->join_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t x : c->xstream)
    if (x) (void)hipStreamDestroy(x);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return ATUS_OK;
}

int atus_synchronize(atus_ctx* c) {
  ATUS_TRY(check_ctx(c));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_upload(atus_ctx* c, int comp, const double* host) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_comp(c, comp));
  if (!host) return fail(ATUS_ERR_ARG, "null host pointer");
  HIP_TRY(hipMemcpyAsync(c->psi + comp * c->npts, host, sizeof(double2) * c->npts, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_download(atus_ctx* c, int comp, double* host) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_comp(c, comp));
  if (!host) return fail(ATUS_ERR_ARG, "null host pointer");
  HIP_TRY(hipMemcpyAsync(host, c->psi + comp * c->npts, sizeof(double2) * c->npts, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_snapshot_save(atus_ctx* c, int slot) {
  ATUS_TRY(check_ctx(c));
  if (slot < 0 || slot >= 4) return fail(ATUS_ERR_ARG, "snapshot slot must be 0..3");
  const size_t bytes = sizeof(double2) * c->npts * c->ncomp;
  if (!c->snap[slot] && hipMalloc(&c->snap[slot], bytes) != hipSuccess) {
    c->snap[slot] = nullptr;
    return fail(ATUS_ERR_NOMEM, "snapshot allocation");
  }
  HIP_TRY(hipMemcpyAsync(c->snap[slot], c->psi, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_snapshot_restore(atus_ctx* c, int slot) {
  ATUS_TRY(check_ctx(c));
  if (slot < 0 || slot >= 4) return fail(ATUS_ERR_ARG, "snapshot slot must be 0..3");
  if (!c->snap[slot]) return fail(ATUS_ERR_STATE, "snapshot slot " + std::to_string(slot) + " holds no snapshot");
  HIP_TRY(hipMemcpyAsync(c->psi, c->snap[slot], sizeof(double2) * c->npts * c->ncomp, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_copy_components(atus_ctx* dst, int dst_comp, atus_ctx* src, int src_comp, int n) {
  ATUS_TRY(check_ctx(dst));
  ATUS_TRY(check_ctx(src));
  if (dst->device != src->device) return fail(ATUS_ERR_ARG, "contexts on different devices");
  if (dst->npts != src->npts) return fail(ATUS_ERR_ARG, "contexts of different grid sizes");
  if (n < 1 || src_comp < 0 || dst_comp < 0 || src_comp + n > src->ncomp || dst_comp + n > dst->ncomp)
    return fail(ATUS_ERR_ARG, "component range out of bounds");
  if (dst == src && dst_comp != src_comp && dst_comp < src_comp + n && src_comp < dst_comp + n)
    return fail(ATUS_ERR_ARG, "overlapping component ranges within one context");
  if (dst == src && dst_comp == src_comp) return ATUS_OK;
  // the source's pending work first; the copy runs on the destination's stream
  HIP_TRY(hipStreamSynchronize(src->stream));
  HIP_TRY(hipMemcpyAsync(dst->psi + (size_t)dst_comp * dst->npts, src->psi + (size_t)src_comp * src->npts,
                         sizeof(double2) * src->npts * n, hipMemcpyDeviceToDevice, dst->stream));
  HIP_TRY(hipStreamSynchronize(dst->stream));
  return ATUS_OK;
}

int atus_set_potential(atus_ctx* c, int comp, const double* host) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_set_potential"));
  if ((!host || !c->pot) && c->live_graphs > 0)
    return fail(ATUS_ERR_STATE, "the potential buffer cannot be freed or allocated while captured graphs of this "
                                "context exist (their V steps read it); destroy them first");
  if (!host) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(c->pot);
    c->pot = nullptr;
    return ATUS_OK;
  }
  ATUS_TRY(check_comp(c, comp));
  if (!c->pot) {
    if (hipMalloc(&c->pot, sizeof(double) * c->npts * c->ncomp) != hipSuccess)
      return fail(ATUS_ERR_NOMEM, "potential allocation");
    HIP_TRY(hipMemsetAsync(c->pot, 0, sizeof(double) * c->npts * c->ncomp, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(c->pot + comp * c->npts, host, sizeof(double) * c->npts, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_set_g(atus_ctx* c, const double* gs) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_set_g"));
  if (!gs) return fail(ATUS_ERR_ARG, "null g");
  for (int i = 0; i < 16; ++i) c->gs[i] = 0;
  for (int i = 0; i < c->ncomp * c->ncomp; ++i) c->gs[i] = gs[i];
  return ATUS_OK;
}

int atus_set_alpha(atus_ctx* c, const double alpha[3]) {
  ATUS_TRY(check_ctx(c));
  for (int a = 0; a < 3; ++a) c->alpha[a] = a < c->dim ? alpha[a] : 0.0;
  c->kin_valid = false;
  return ATUS_OK;
}

int atus_set_alpha2(atus_ctx* c, const double alpha2[3]) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_set_alpha2"));
  if (!alpha2) return fail(ATUS_ERR_ARG, "null alpha2");
  if (c->ncomp < 2 || c->ncomp % 2) return fail(ATUS_ERR_STATE, "two species need an even number of components");
  if (c->fs) return fail(ATUS_ERR_UNSUPPORTED, "two species on 1D grids longer than 4096 are not covered");
  if (!c->alpha2_set && c->live_graphs > 0)
    return fail(ATUS_ERR_STATE, "a second species cannot be added while captured graphs of this context exist");
  for (int a = 0; a < 3; ++a) c->alpha2[a] = a < c->dim ? alpha2[a] : 0.0;
  c->alpha2_set = true;
  c->kin_valid = false;
  return ATUS_OK;
}

int atus_ft(atus_ctx* c, int comp, int isign) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_comp(c, comp));
  if (isign != 1 && isign != -1) return ATUS_OK;
  double2* p = c->psi + comp * c->npts;
  ATUS_TRY(full_ft(c, isign, p, p, 1));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_kinetic(atus_ctx* c, double frac, double dt) {
  ATUS_TRY(check_ctx(c));
  if (frac != 0.5 && frac != 1.0) return fail(ATUS_ERR_ARG, "frac must be 0.5 or 1");
  ATUS_TRY(ensure_kin(c, dt));
  ATUS_TRY(kinetic_unfused(c, frac == 0.5 ? 0 : 1));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_nl_step(atus_ctx* c, double dt) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_nl_step"));
  VArgs va;
  fill_vargs(c, va, dt, nullptr);
  HIP_TRY(launch_point(V_NL, c->ncomp, c->psi, c->grid(), va, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_nl_if_step(atus_ctx* c, double dt, double beta0, const double* deltaL) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_nl_if_step"));
  atus_laser L{};
  L.beta0 = beta0;
  for (int i = 0; i < c->ncomp; ++i) L.deltaL[i] = deltaL ? deltaL[i] : 0.0;
  VArgs va;
  fill_vargs(c, va, dt, &L);
  HIP_TRY(launch_point(V_NL_IF, c->ncomp, c->psi, c->grid(), va, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_bragg_ad(atus_ctx* c, double dt, double t, double F, const atus_laser* laser) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_bragg_ad"));
  if (c->ncomp != 2) return fail(ATUS_ERR_STATE, "bragg_ad needs exactly two components");
  if (!laser) return fail(ATUS_ERR_ARG, "null laser parameters");
  VArgs va;
  fill_vargs(c, va, dt, laser);
  va.t1 = t + 0.5 * dt;
  va.F = F;
  HIP_TRY(launch_point(V_BRAGG, 2, c->psi, c->grid(), va, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_set_momentum(atus_ctx* c, int comp, const double p[3]) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_comp(c, comp));
  hipLaunchKernelGGL(k_momentum, dim3(ew_blocks(c->npts)), dim3(256), 0, c->stream, c->psi + comp * c->npts, c->grid(),
                     p[0], c->dim > 1 ? p[1] : 0.0, c->dim > 2 ? p[2] : 0.0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

}  // extern "C"

namespace {

// per-step time-dependent arguments of the V steps of one block: the
// amplitude callback (single runs), the device amplitude table of a batch
// ([Nk][nbatch]), and for replayed graphs device tables of t1 ([Nk]) and of
// the single-run amplitude ([Nk])
struct StepFeed {
  const atus_step_params* sp = nullptr;
  const atus_ds_params* ds = nullptr;  // pulse gates of the two-species Bragg step
  const double* fbatch = nullptr;
  const double* t1dev = nullptr;
  const double* fdev = nullptr;
};

// the launches of one block of a single run on the context's stream (half
// kinetic, (V, full) x (Nk-1), V, half), without synchronisation; *t
// advances with the reference's bookkeeping
// V-step arguments of step m of a block (t1, amplitude, pulse gates)
int step_vargs(const VArgs& base, int m, double tt, double dt, const StepFeed& fd, VArgs& va) {
  va = base;
  va.t0 = tt;
  va.t1 = tt + 0.5 * dt;
  if (fd.t1dev) {
    va.t1p = fd.t1dev + (m - 1);
    va.Fp = fd.fdev + (m - 1);
  } else {
    va.F = (fd.sp && fd.sp->amp) ? fd.sp->amp(va.t1, fd.sp->amp_user) : 1.0;
    if (!std::isfinite(va.F)) return fail(ATUS_ERR_NUMERIC, "amplitude callback returned a non-finite value");
  }
  if (fd.ds) {
    // bragg_ds.cpp:120-128: a pulse is off once t1 - seq.time exceeds its duration
    const double time = va.t1 - fd.ds->seq_time;
    va.mode1 = time > fd.ds->duration[0] ? 0.0 : 1.0;
    va.mode2 = time > fd.ds->duration[1] ? 0.0 : 1.0;
  }
  return ATUS_OK;
}

// a 3D block with the y and z passes between two kinetic steps chained slab
// by slab of c->slab x-planes (they only couple points of one x-plane).
// Between two kinetic steps the state is kept in x space along x and in k
// space along y and z; a step is
//   per slab [y inv, vfuse (z inv, V, z fwd), y fwd]; kin (x fwd, phase, x inv).
// The half kinetic steps at the block's ends run per axis (the three axes'
// phases commute), so that they fold into the first and last steps:
//   kin_x/2; per slab [kin_y/2, vfuse with z's half step before V, y fwd]; kin
//   ... per slab [y inv, vfuse with z's half step after V, kin_y/2]; kin_x/2
// which is 4 Nk + 1 passes over the grid per block of Nk steps.
// Each slab's passes follow each other while the slab's data is still in
// the device's last-level cache.
int single_block_slabs(atus_ctx* c, int vk, const VArgs& base, int Nk, double dt, double* t, const StepFeed& fd) {
  // without slabs (a grid of at most one slab's size, or slabs switched
  // off) the whole grid is one slab on the context's stream
  const int n0 = c->n[0], N0 = c->n[0];
  const int S = c->slab_planes() > 0 ? c->slab_planes() : n0;
  const long long inner0 = c->inner_of(0);
  // comp0_last: the components in reverse order (component 0's workgroups
  // dispatched last), so that the block's final pass leaves component 0 in
  // the last-level cache for Rabi integrals that follow; one species only
  auto kin = [&](const KinTabs& kt, const VArgs& va, bool comp0_last = false) -> int {
    ProfScope ps(c, ATUS_KFAM_KIN);
    const bool rev = comp0_last && kt.f0 == nullptr && c->ncomp > 1;
    double2* base = rev ? c->psi + (long long)(c->ncomp - 1) * c->npts : c->psi;
    HIP_TRY(launch_kin(N0, c->rm[ATUS_KFAM_KIN], false, V_NONE, 1, base, rev ? -c->npts : c->npts, c->ncomp, inner0,
                       c->tw[N0], kt, va, c->dx[0], lim(c, ATUS_KFAM_KIN), c->stream));
    return ATUS_OK;
  };
  // odd slabs on the second stream (profiled blocks record their events on
  // each launch's stream)
  const int K = S >= n0 ? 1 : std::max(1, std::min(c->slab_streams, atus_ctx::MAX_SLAB_STREAMS));
  if (K > 1 && !c->fork_ev &&
      hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess)
    return fail(ATUS_ERR_HIP, "fork event of the slab chains");
  for (int k = 0; k + 1 < K; ++k) {
    if (c->xstream[k]) continue;
    if (hipStreamCreateWithFlags(&c->xstream[k], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->join_ev[k], hipEventDisableTiming) != hipSuccess)
      return fail(ATUS_ERR_HIP, "extra stream of the slab chains");
  }
  auto st_of = [&](int s0) {
    const int k = (s0 / S) % K;
    return k == 0 ? c->stream : c->xstream[k - 1];
  };
  auto fork = [&]() -> int {
    if (K == 1) return ATUS_OK;
    HIP_TRY(hipEventRecord(c->fork_ev, c->stream));
    for (int k = 0; k + 1 < K; ++k) HIP_TRY(hipStreamWaitEvent(c->xstream[k], c->fork_ev, 0));
    return ATUS_OK;
  };
  auto join = [&]() -> int {
    for (int k = 0; k + 1 < K; ++k) {
      HIP_TRY(hipEventRecord(c->join_ev[k], c->xstream[k]));
      HIP_TRY(hipStreamWaitEvent(c->stream, c->join_ev[k], 0));
    }
    return ATUS_OK;
  };
  double tt = *t;
  ATUS_TRY(kin(axis_tabs(c, 0), base));
  tt += 0.5 * dt;
  for (int m = 1; m <= Nk; ++m) {
    VArgs va;
    ATUS_TRY(step_vargs(base, m, tt, dt, fd, va));
    const bool first = m == 1, last = m == Nk;
    ATUS_TRY(fork());
    for (int s0 = 0; s0 < n0; s0 += S) {
      const int Ss = std::min(S, n0 - s0);
      if (first) ATUS_TRY(kin_y_slab(c, va, s0, Ss, st_of(s0)));
      else ATUS_TRY(pass_slab(c, 1, +1, s0, Ss, st_of(s0)));
      ATUS_TRY(vfuse_slab(c, vk, va, s0, Ss, st_of(s0), (first ? 1 : 0) | (last ? 2 : 0)));
      if (last) ATUS_TRY(kin_y_slab(c, va, s0, Ss, st_of(s0)));
      else ATUS_TRY(pass_slab(c, 1, -1, s0, Ss, st_of(s0)));
    }
    ATUS_TRY(join());
    ATUS_TRY(kin(last ? axis_tabs(c, 0) : kin_tabs(c, 1), va, last));
    tt += last ? 0.5 * dt : dt;
  }
  *t = tt;
  return ATUS_OK;
}

// room for n doubles in the block tables, the host side free to rewrite
int block_tables(atus_ctx* c, size_t n) {
  if (c->blk_ev) HIP_TRY(hipEventSynchronize(c->blk_ev));
  if (!c->blk_ev && hipEventCreateWithFlags(&c->blk_ev, hipEventDisableTiming) != hipSuccess)
    return fail(ATUS_ERR_HIP, "event of the block tables");
  if (c->blk_cap >= n) return ATUS_OK;
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipFree(c->blk_dev);
  if (c->blk_host) (void)hipHostFree(c->blk_host);
  c->blk_dev = nullptr;
  c->blk_host = nullptr;
  c->blk_cap = 0;
  if (hipMalloc(&c->blk_dev, sizeof(double) * n) != hipSuccess ||
      hipHostMalloc(&c->blk_host, sizeof(double) * n, hipHostMallocDefault) != hipSuccess)
    return fail(ATUS_ERR_NOMEM, "block tables");
  c->blk_cap = n;
  return ATUS_OK;
}

// whether a block of this context runs as one k_block1d launch
bool resident_ok(const atus_ctx* c) {
  return c->resident && c->dim == 1 && !c->fs && c->n[0] <= 4096;
}

// a 1D block in one launch: the t1 (and, for a single run, the amplitude)
// of every step are evaluated with the eager block's bookkeeping and
// uploaded first, unless a replayed graph supplies them; batch: the
// amplitudes come from fd.fbatch ([Nk][runs]), else from the callback or
// fd.fdev
int resident_block(atus_ctx* c, int vk, const VArgs& base, int Nk, double dt, double* t, const StepFeed& fd,
                   bool batch) {
  const int N0 = c->n[0];
  double tt = *t + 0.5 * dt;
  const double* t1tab = fd.t1dev;
  const double* Ftab = batch ? fd.fbatch : fd.fdev;
  if (!t1tab) {
    ATUS_TRY(block_tables(c, 2 * (size_t)Nk));
    double tc = tt;
    for (int m = 1; m <= Nk; ++m) {
      if (batch) {
        c->blk_host[m - 1] = tc + 0.5 * dt;
      } else {
        VArgs va;
        ATUS_TRY(step_vargs(base, m, tc, dt, fd, va));
        c->blk_host[m - 1] = va.t1;
        c->blk_host[Nk + m - 1] = va.F;
      }
      tc += (m < Nk) ? dt : 0.5 * dt;
    }
    HIP_TRY(hipMemcpyAsync(c->blk_dev, c->blk_host, sizeof(double) * (batch ? Nk : 2 * Nk), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipEventRecord(c->blk_ev, c->stream));
    t1tab = c->blk_dev;
    if (!batch) Ftab = c->blk_dev + Nk;
  }
  if (!t1tab || !Ftab || (batch && !base.runs)) return fail(ATUS_ERR_STATE, "resident block without its step tables");
  // 8 values per thread: a run's pencil gets the most threads to hide
  // latency with (4096 points, one run: 10.7 us per step against 13.3 at
  // 16), and a batch's runs take the plan of the same run alone
  const int rm = 8;
  {
    ProfScope ps(c, ATUS_KFAM_KIN, Nk + 1.0);
    HIP_TRY(launch_block1d(N0, rm, vk, batch ? 2 : c->ncomp, batch, c->psi, c->npts,
                           batch ? c->nbatch : 1, c->tw[N0], kin_tabs(c, 0).e0, kin_tabs(c, 1).e0, base, t1tab, Ftab,
                           c->dx[0], Nk, lim(c, ATUS_KFAM_KIN), c->stream));
  }
  for (int m = 1; m <= Nk; ++m) tt += (m < Nk) ? dt : 0.5 * dt;
  *t = tt;
  return ATUS_OK;
}

int single_block(atus_ctx* c, int vk, const VArgs& base, int Nk, double dt, double* t, const StepFeed& fd) {
  // the fused kernels cover one species of up to two components; two
  // species (per-component kinetic tables, four-component steps) run the
  // V step and the kinetic step as separate passes
  const bool fused = c->ncomp <= 2 && !c->alpha2_set && (vk == V_NL || vk == V_NL_IF || vk == V_BRAGG);
  if (fused && c->dim == 3) return single_block_slabs(c, vk, base, Nk, dt, t, fd);
  if (fused && !fd.ds && resident_ok(c)) return resident_block(c, vk, base, Nk, dt, t, fd, false);
  const int N0 = c->n[0], NL = c->n[c->dim - 1];
  const long long inner0 = c->inner_of(0);
  double tt = *t;
  auto kin2d = [&](const KinTabs& kt, const VArgs& va) -> int {
    ProfScope ps(c, ATUS_KFAM_KIN);
    HIP_TRY(launch_kin(N0, c->rm[ATUS_KFAM_KIN], false, V_NONE, 1, c->psi, c->npts, c->ncomp, inner0, c->tw[N0], kt,
                       va, c->dx[0], lim(c, ATUS_KFAM_KIN), c->stream));
    return ATUS_OK;
  };
  // half kinetic step; a 2D grid keeps y in k space between two kinetic
  // steps, and the half steps at the block's ends run per axis: x's share
  // here and at the end, y's inside the first and last V steps' passes
  if (fused && c->fs) {
    ATUS_TRY(fs_col(c, FS_ENTER, c->psi, c->ncomp, c->npts, V_NONE, 1, false, base, 1.0));
    ATUS_TRY(fs_row_kin(c, 0, base));
  } else if (fused && c->dim == 2) {
    ATUS_TRY(kin2d(axis_tabs(c, 0), base));
  } else {
    ATUS_TRY(kinetic_unfused(c, 0));
  }
  tt += 0.5 * dt;
  for (int m = 1; m <= Nk; ++m) {
    VArgs va;
    ATUS_TRY(step_vargs(base, m, tt, dt, fd, va));
    const int f = (m < Nk) ? 1 : 0;
    if (!fused) {
      {
        ProfScope ps(c, ATUS_KFAM_POINT);
        HIP_TRY(launch_point(vk, c->ncomp, c->psi, c->grid(), va, c->stream));
      }
      ATUS_TRY(kinetic_unfused(c, f));
    } else if (c->fs) {
      ATUS_TRY(fs_col(c, FS_STEP, c->psi, 1, c->npts, vk, c->ncomp, false, va, 1.0));
      ATUS_TRY(fs_row_kin(c, f, va));
    } else if (c->dim == 1) {
      ProfScope ps(c, ATUS_KFAM_KIN);
      HIP_TRY(launch_kin(N0, c->rm[ATUS_KFAM_KIN], true, vk, c->ncomp, c->psi, c->npts, c->ncomp, 1, c->tw[N0], kin_tabs(c, f), va,
                         c->dx[0], lim(c, ATUS_KFAM_KIN), c->stream));
    } else {
      const int edge = (m == 1 ? 1 : 0) | (m == Nk ? 2 : 0);
      {
        ProfScope ps(c, ATUS_KFAM_VFUSE);
        HIP_TRY(launch_vfuse(NL, c->rm[ATUS_KFAM_VFUSE], vk, c->ncomp, c->psi, c->npts, c->npts / NL, c->tw[NL], va, 1,
                             c->n[0], c->dx[0], lim(c, ATUS_KFAM_VFUSE), c->stream, edge ? axis_tabs(c, 1).e0 : nullptr,
                             edge));
      }
      ATUS_TRY(kin2d(m < Nk ? kin_tabs(c, 1) : axis_tabs(c, 0), va));
    }
    tt += (m < Nk) ? dt : 0.5 * dt;
  }
  if (fused && c->fs) ATUS_TRY(fs_col(c, FS_EXIT, c->psi, c->ncomp, c->npts, V_NONE, 1, false, base, 1.0));
  *t = tt;
  return ATUS_OK;
}

void batch_vargs(atus_ctx* c, VArgs& base, double dt) {
  fill_vargs(c, base, dt, nullptr);
  base.runs = c->runs;
  base.run_shift = (long long)(c->nc_run - 1) * c->npts;
  base.pencils_per_run = c->dim == 1 ? 1 : c->npts / c->n[c->dim - 1];
}

// the launches of one block of every run of a batch, without synchronisation
int batch_block(atus_ctx* c, int vk, const VArgs& base, int Nk, double dt, double* t, const StepFeed& fd) {
  const int B = c->nbatch;
  if (resident_ok(c)) return resident_block(c, vk, base, Nk, dt, t, fd, true);
  const int N0 = c->n[0], NL = c->n[c->dim - 1];
  const long long inner0 = c->inner_of(0);
  double tt = *t;
  auto kin = [&](const KinTabs& kt, const VArgs& va) -> int {
    ProfScope ps(c, ATUS_KFAM_KIN);
    HIP_TRY(launch_kin(N0, c->rm[ATUS_KFAM_KIN], false, V_NONE, 1, c->psi, c->npts, c->ncomp, inner0, c->tw[N0], kt,
                       va, c->dx[0], lim(c, ATUS_KFAM_KIN), c->stream));
    return ATUS_OK;
  };
  // 2D/3D: the order of a single run's block (x in x space, the other axes
  // in k space between two kinetic steps; the half steps at the block's ends
  // per axis, folded into the first and last steps)
  if (c->fs) {
    ATUS_TRY(fs_col(c, FS_ENTER, c->psi, c->ncomp, c->npts, V_NONE, 1, false, base, 1.0));
    ATUS_TRY(fs_row_kin(c, 0, base));
  } else if (c->dim >= 2) {
    ATUS_TRY(kin(axis_tabs(c, 0), base));
  } else {
    ATUS_TRY(kinetic_unfused(c, 0));
  }
  tt += 0.5 * dt;
  for (int m = 1; m <= Nk; ++m) {
    VArgs va = base;
    va.t1 = tt + 0.5 * dt;
    va.Fm = fd.fbatch + (size_t)(m - 1) * B;
    if (fd.t1dev) va.t1p = fd.t1dev + (m - 1);
    const int f = (m < Nk) ? 1 : 0;
    if (c->fs) {
      ATUS_TRY(fs_col(c, FS_STEP, c->psi, B, c->npts, vk, 2, true, va, 1.0));
      ATUS_TRY(fs_row_kin(c, f, va));
    } else if (c->dim == 1) {
      ProfScope ps(c, ATUS_KFAM_KIN);
      HIP_TRY(launch_vstep_batch(N0, c->rm[ATUS_KFAM_KIN], true, vk, c->psi, c->npts, B, c->tw[N0], kin_tabs(c, f), va,
                                 1, N0, c->dx[0], lim(c, ATUS_KFAM_KIN), c->stream));
    } else {
      const bool first = m == 1, last = m == Nk;
      const int edge = (first ? 1 : 0) | (last ? 2 : 0);
      const long long ppx = c->dim == 3 ? c->n[1] : 1;
      if (c->dim == 3) {
        if (first) ATUS_TRY(kin_y_slab(c, va, 0, c->n[0], c->stream));
        else ATUS_TRY(pass(c, 1, +1, c->psi, c->psi, c->ncomp, 1.0));
      }
      {
        ProfScope ps(c, ATUS_KFAM_VFUSE);
        HIP_TRY(launch_vstep_batch(NL, c->rm[ATUS_KFAM_VFUSE], false, vk, c->psi, c->npts,
                                   (long long)B * base.pencils_per_run, c->tw[NL], kin_tabs(c, f), va, ppx, c->n[0],
                                   c->dx[0], lim(c, ATUS_KFAM_VFUSE), c->stream,
                                   edge ? axis_tabs(c, c->dim - 1).e0 : nullptr, edge));
      }
      if (c->dim == 3) {
        if (last) ATUS_TRY(kin_y_slab(c, va, 0, c->n[0], c->stream));
        else ATUS_TRY(pass(c, 1, -1, c->psi, c->psi, c->ncomp, 1.0));
      }
      ATUS_TRY(kin(last ? axis_tabs(c, 0) : kin_tabs(c, 1), va));
    }
    tt += (m < Nk) ? dt : 0.5 * dt;
  }
  if (c->fs) ATUS_TRY(fs_col(c, FS_EXIT, c->psi, c->ncomp, c->npts, V_NONE, 1, false, base, 1.0));
  *t = tt;
  return ATUS_OK;
}

// end of a block run eagerly: event timing or a plain synchronisation
int finish_block(atus_ctx* c) {
  if (c->timing) {
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
  } else {
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return ATUS_OK;
}

}  // namespace

extern "C" {

int atus_run_block(atus_ctx* c, const atus_step_params* sp, int Nk, double* t) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_run_block"));
  if (!sp || !t) return fail(ATUS_ERR_ARG, "null argument");
  if (Nk < 1) return fail(ATUS_ERR_ARG, "Nk must be >= 1");
  const int vk = vkind_of(sp->kind);
  if (vk < 0) return fail(ATUS_ERR_ARG, "unknown step kind " + std::to_string(sp->kind));
  ATUS_TRY(check_vk_ncomp(c, vk));
  ATUS_TRY(ensure_kin(c, sp->dt));
  VArgs base;
  fill_vargs(c, base, sp->dt, &sp->laser);
  if (c->timing) HIP_TRY(hipEventRecord(c->ev0, c->stream));
  StepFeed fd;
  fd.sp = sp;
  ATUS_TRY(single_block(c, vk, base, Nk, sp->dt, t, fd));
  return finish_block(c);
}

int atus_v_step(atus_ctx* c, int kind, double dt, double t, double F, const atus_laser* laser) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_v_step"));
  const int vk = vkind_of(kind);
  if (vk < 0) return fail(ATUS_ERR_ARG, "unknown step kind " + std::to_string(kind));
  ATUS_TRY(check_vk_ncomp(c, vk));
  if (!std::isfinite(F)) return fail(ATUS_ERR_NUMERIC, "non-finite amplitude");
  VArgs va;
  fill_vargs(c, va, dt, laser);
  va.t0 = t;
  va.t1 = t + 0.5 * dt;
  va.F = F;
  HIP_TRY(launch_point(vk, c->ncomp, c->psi, c->grid(), va, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

}  // extern "C"

namespace {

// the V-step kind and the species-independent arguments of a two-species
// step; the per-step t1, F and pulse gates are set by the caller
int ds_vargs(atus_ctx* c, int kind, double dt, const atus_laser* laser1, const atus_ds_params* ds, int* vk, VArgs& base) {
  if (c->ncomp < 2 || c->ncomp % 2) return fail(ATUS_ERR_STATE, "two species need an even number of components");
  if (kind == ATUS_STEP_FREEPROP_IF) {
    *vk = V_NL_IF2;
  } else if (kind == ATUS_STEP_BRAGG_AD) {
    if (c->ncomp != 4) return fail(ATUS_ERR_STATE, "the two-species Bragg step needs four components");
    *vk = V_BRAGG_DS;
  } else {
    return fail(ATUS_ERR_ARG, "two-species steps are freeprop (IF) or bragg_ad steps");
  }
  fill_vargs(c, base, dt, laser1);
  const atus_laser& L2 = ds->laser2;
  base.amp2 = L2.amp0;
  base.laser_k2 = L2.laser_k;
  base.laser_domh2 = L2.laser_domh;
  base.laser_dk2 = L2.laser_dk;
  base.chirp_rate2 = L2.chirp_rate0;
  base.phase2 = L2.phase0;
  base.beta2 = L2.beta0;
  return ATUS_OK;
}

}  // namespace

extern "C" {

int atus_run_block_ds(atus_ctx* c, const atus_step_params* sp, const atus_ds_params* ds, int Nk, double* t) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_run_block_ds"));
  if (!sp || !ds || !t) return fail(ATUS_ERR_ARG, "null argument");
  if (Nk < 1) return fail(ATUS_ERR_ARG, "Nk must be >= 1");
  int vk;
  VArgs base;
  ATUS_TRY(ds_vargs(c, sp->kind, sp->dt, &sp->laser, ds, &vk, base));
  ATUS_TRY(ensure_kin(c, sp->dt));
  if (c->timing) HIP_TRY(hipEventRecord(c->ev0, c->stream));
  StepFeed fd;
  fd.ds = ds;
  ATUS_TRY(single_block(c, vk, base, Nk, sp->dt, t, fd));
  return finish_block(c);
}

int atus_v_step_ds(atus_ctx* c, int kind, double dt, double t, const atus_laser* laser1, const atus_ds_params* ds) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_v_step_ds"));
  if (!laser1 || !ds) return fail(ATUS_ERR_ARG, "null argument");
  int vk;
  VArgs va;
  ATUS_TRY(ds_vargs(c, kind, dt, laser1, ds, &vk, va));
  // bragg_ds.cpp:116-128: t1 = t + dt/2, a pulse is off once t1 - seq.time
  // exceeds its duration
  va.t1 = t + 0.5 * dt;
  va.F = 1.0;
  const double time = va.t1 - ds->seq_time;
  va.mode1 = time > ds->duration[0] ? 0.0 : 1.0;
  va.mode2 = time > ds->duration[1] ? 0.0 : 1.0;
  HIP_TRY(launch_point(vk, c->ncomp, c->psi, c->grid(), va, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_particle_number(atus_ctx* c, int comp, double* out) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_comp(c, comp));
  if (!out) return fail(ATUS_ERR_ARG, "null output");
  RedNorm fn{c->psi + comp * c->npts};
  double s;
  ATUS_TRY(reduce(c, fn, c->npts, &s));
  *out = c->ar() * s;
  if (!std::isfinite(*out)) return fail(ATUS_ERR_NUMERIC, "particle number is not finite");
  return ATUS_OK;
}

int atus_expval_x(atus_ctx* c, int comp, double out[3]) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_comp(c, comp));
  RedX fn{c->psi + comp * c->npts, c->grid()};
  double s[3];
  ATUS_TRY(reduce(c, fn, c->npts, s));
  for (int a = 0; a < 3; ++a) out[a] = a < c->dim ? c->ar() * s[a] : 0.0;
  return ATUS_OK;
}

int atus_expval_k(atus_ctx* c, int comp, double out[3]) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_comp(c, comp));
  ATUS_TRY(ensure_scratch(c));
  ATUS_TRY(full_ft(c, -1, c->psi + comp * c->npts, c->scratch, 1));
  RedK fn{c->scratch, c->grid()};
  double s[3];
  ATUS_TRY(reduce(c, fn, c->npts, s));
  for (int a = 0; a < 3; ++a) out[a] = a < c->dim ? c->ar_k() * s[a] : 0.0;
  return ATUS_OK;
}

int atus_get_moments(atus_ctx* c, int comp, atus_moments* out) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_comp(c, comp));
  if (!out) return fail(ATUS_ERR_ARG, "null output");
  *out = atus_moments{};
  const double2* p = c->psi + comp * c->npts;
  const int D = c->dim;
  double s1[5];
  ATUS_TRY(reduce(c, RedMom1{p, c->grid()}, c->npts, s1));
  if (!(s1[0] > 0.0)) return fail(ATUS_ERR_NUMERIC, "the state is zero");
  // PN, then each moment scaled by ar / PN (ana_expval.cpp:337-341)
  const double PN = c->ar() * s1[0];
  out->N = PN;
  out->phase = s1[4] / s1[0];
  RedMom2<false> f2{p, c->grid(), {0, 0, 0}};
  for (int a = 0; a < D; ++a) {
    out->x[a] = s1[1 + a] * (c->ar() / PN);
    f2.mu[a] = out->x[a];
  }
  double s2[3];
  ATUS_TRY(reduce(c, f2, c->npts, s2));
  for (int a = 0; a < D; ++a) out->x2[a] = s2[a] * (c->ar() / PN);
  // momentum moments of the transformed state (a copy), still over PN
  ATUS_TRY(ensure_scratch(c));
  ATUS_TRY(full_ft(c, -1, p, c->scratch, 1));
  double sk[3];
  ATUS_TRY(reduce(c, RedK{c->scratch, c->grid()}, c->npts, sk));
  RedMom2<true> fk{c->scratch, c->grid(), {0, 0, 0}};
  for (int a = 0; a < D; ++a) {
    out->k[a] = sk[a] * (c->ar_k() / PN);
    fk.mu[a] = out->k[a];
  }
  double sk2[3];
  ATUS_TRY(reduce(c, fk, c->npts, sk2));
  for (int a = 0; a < D; ++a) out->k2[a] = sk2[a] * (c->ar_k() / PN);
  return ATUS_OK;
}

int atus_inflate(atus_ctx* dst, atus_ctx* src, int fak, int fix) {
  ATUS_TRY(check_ctx(dst));
  ATUS_TRY(check_ctx(src));
  if (dst->device != src->device) return fail(ATUS_ERR_ARG, "contexts on different devices");
  if (fak < 1) return fail(ATUS_ERR_ARG, "the inflation factor must be >= 1");
  if (dst->dim != src->dim || dst->ncomp != src->ncomp)
    return fail(ATUS_ERR_ARG, "contexts of different dimension or component count");
  for (int a = 0; a < src->dim; ++a) {
    const long long want = fix ? src->n[a] : (long long)src->n[a] * fak;
    if (dst->n[a] != want)
      return fail(ATUS_ERR_ARG, "destination axis " + std::to_string(a) + " has " + std::to_string(dst->n[a]) +
                                    " points, the inflation needs " + std::to_string(want));
  }
  HIP_TRY(hipStreamSynchronize(src->stream));
  const dim3 grid(ew_blocks(dst->npts), (unsigned)dst->ncomp), block(256);
  if (fix)
    hipLaunchKernelGGL((k_inflate<true>), grid, block, 0, dst->stream, src->psi, dst->psi, src->grid(), dst->grid(), fak,
                       src->npts, dst->npts);
  else
    hipLaunchKernelGGL((k_inflate<false>), grid, block, 0, dst->stream, src->psi, dst->psi, src->grid(), dst->grid(),
                       fak, src->npts, dst->npts);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(dst->stream));
  return ATUS_OK;
}

int atus_rabi_integrals(atus_ctx* c, int comp, const double* k0, int n, double threshold, double* out) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_comp(c, comp));
  if (n < 1 || n > RED_MAX_NV || !k0 || !out) return fail(ATUS_ERR_ARG, "need 1..8 momentum points");
  {
    const RabiReduction rr = rabi_reduction(c, c->psi + comp * c->npts, k0, n, threshold);
    if (rabi_box3_ok(c, rr)) {
      ATUS_TRY(ensure_rabi_box(c, rr, 1));
      ATUS_TRY(rabi_box3(c, c->psi + comp * c->npts, c->npts, 1, rr, c->sred_out));
      HIP_TRY(hipMemcpyAsync(c->red_host, c->sred_out, sizeof(double) * RED_MAX_NV, hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      for (int q = 0; q < n; ++q) out[q] = c->ar_k() * c->red_host[q];
      return ATUS_OK;
    }
  }
  ATUS_TRY(ensure_scratch(c));
  ATUS_TRY(full_ft(c, -1, c->psi + comp * c->npts, c->scratch, 1));
  const RabiReduction rr = rabi_reduction(c, c->scratch, k0, n, threshold);
  double s[RED_MAX_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (rr.use_box) {
    ATUS_TRY(reduce(c, rr.box, rr.npts, s));
  } else {
    ATUS_TRY(reduce(c, rr.full, c->npts, s));
  }
  for (int q = 0; q < n; ++q) out[q] = c->ar_k() * s[q];
  return ATUS_OK;
}

int atus_energy(atus_ctx* c, double* out) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_energy"));
  if (!out) return fail(ATUS_ERR_ARG, "null output");
  ATUS_TRY(ensure_scratch(c));
  double ekin = 0;
  for (int i = 0; i < c->ncomp; ++i) {
    ATUS_TRY(full_ft(c, -1, c->psi + i * c->npts, c->scratch, 1));
    RedEkin fn{};
    fn.psik = c->scratch;
    fn.gr = c->grid();
    const double* al = (c->alpha2_set && i >= c->ncomp / 2) ? c->alpha2 : c->alpha;
    for (int a = 0; a < 3; ++a) fn.alpha[a] = al[a];
    double s;
    ATUS_TRY(reduce(c, fn, c->npts, &s));
    ekin += s;
  }
  RedEpot fp{};
  fp.psi = c->psi;
  fp.pot = c->pot;
  fp.npts = c->npts;
  fp.nc = c->ncomp;
  for (int i = 0; i < 16; ++i) fp.gs[i] = c->gs[i];
  double ep;
  ATUS_TRY(reduce(c, fp, c->npts, &ep));
  *out = c->ar_k() * ekin + c->ar() * ep;
  return ATUS_OK;
}

// ------------------------------------------------------------------ Sobolev

namespace {

// out_sob = F^-1[op F psi], out_lap = F^-1[-alpha k^2 F psi] (optional)
int sob_filter(atus_ctx* c, int comp, const double2* src, double2* out_sob, double2* out_lap) {
  const long long off = comp * c->npts;
  SobArgs sa{};
  sa.alpha = c->sob_alpha[comp];
  sa.dim = c->dim;
  for (int a = 0; a < 3; ++a) {
    sa.dk[a] = c->dk[a];
    sa.n[a] = c->n[a];
  }
  sa.scale = ft_factor(c->dim, c->dx) * ft_factor(c->dim, c->dk);
  const double2* s = src + off;
  double2* w = out_sob + off;
  for (int a = c->dim - 1; a >= 1; --a) {
    ATUS_TRY(pass(c, a, -1, s, w, 1, 1.0));
    s = w;
  }
  const int N0 = c->n[0];
  ATUS_TRY(ensure_twiddles(c, N0));
  HIP_TRY(launch_sobf(N0, c->rm[ATUS_KFAM_KIN], c->dim == 1, s, w, out_lap ? out_lap + off : nullptr, c->npts, 1, c->inner_of(0),
                      c->tw[N0], sa, lim(c, ATUS_KFAM_KIN), c->stream));
  for (int a = 1; a < c->dim; ++a) {
    ATUS_TRY(pass(c, a, +1, w, w, 1, 1.0));
    if (out_lap) ATUS_TRY(pass(c, a, +1, out_lap + off, out_lap + off, 1, 1.0));
  }
  return ATUS_OK;
}

// the launches of one iteration up to the update: psi_sob, the Laplacian,
// the Sobolev gradient and the projection products of every wave function
// (Compute_Sobolev_Psi, Compute_Laplace, Compute_Sobolev_Gradient and the
// dot products of Project_Sobolev_Gradient, CSOB_Base.h:251-367).  On 2D
// and 3D grids the L2 gradient is formed inside the last-axis passes
// (k_sob_gradfuse): the Laplacian's last inverse pass and the gradient
// filter's first forward pass.
int sob_body(atus_ctx* c) {
  const int nw = c->ncomp, dim = c->dim;
  const long long np = c->npts;
  SobState* st = (SobState*)c->sob_st;
  if (dim == 1) {
    for (int i = 0; i < nw; ++i) ATUS_TRY(sob_filter(c, i, c->psi, c->sob, c->lap));
    for (int i = 0; i < nw; ++i) {
      hipLaunchKernelGGL(k_sob_grad_in, dim3(ew_blocks(np)), dim3(256), 0, c->stream, c->psi, c->lap, c->pot, c->grad,
                         np, nw, i, c->gs[0 + i * nw], nw > 1 ? c->gs[1 + i * nw] : 0.0);
      HIP_TRY(hipGetLastError());
      ATUS_TRY(sob_filter(c, i, c->grad, c->grad, nullptr));
    }
  } else {
    const int last = dim - 1, NL = c->n[last], N0 = c->n[0];
    SobArgs sa{};
    sa.dim = dim;
    for (int a = 0; a < 3; ++a) {
      sa.dk[a] = c->dk[a];
      sa.n[a] = c->n[a];
    }
    sa.scale = ft_factor(dim, c->dx) * ft_factor(dim, c->dk);
    ATUS_TRY(ensure_twiddles(c, N0));
    for (int i = 0; i < nw; ++i) {
      // psi_sob complete; the Laplacian short of its last-axis inverse pass
      const long long off = i * np;
      sa.alpha = c->sob_alpha[i];
      const double2* s = c->psi + off;
      double2* w = c->sob + off;
      for (int a = last; a >= 1; --a) {
        ATUS_TRY(pass(c, a, -1, s, w, 1, 1.0));
        s = w;
      }
      HIP_TRY(launch_sobf(N0, c->rm[ATUS_KFAM_KIN], false, s, w, c->lap + off, np, 1, c->inner_of(0), c->tw[N0], sa,
                          lim(c, ATUS_KFAM_KIN), c->stream));
      for (int a = 1; a <= last; ++a) ATUS_TRY(pass(c, a, +1, w, w, 1, 1.0));
      for (int a = 1; a < last; ++a) ATUS_TRY(pass(c, a, +1, c->lap + off, c->lap + off, 1, 1.0));
    }
    for (int i = 0; i < nw; ++i) {
      const long long off = i * np;
      sa.alpha = c->sob_alpha[i];
      double2* w = c->grad + off;
      HIP_TRY(launch_sob_gradfuse(NL, c->rm[ATUS_KFAM_VFUSE], nw, c->lap, c->psi, c->pot, c->grad, np, i,
                                  c->gs[0 + i * nw], nw > 1 ? c->gs[1 + i * nw] : 0.0, c->tw[NL], lim(c, ATUS_KFAM_VFUSE),
                                  c->stream));
      for (int a = last - 1; a >= 1; --a) ATUS_TRY(pass(c, a, -1, w, w, 1, 1.0));
      HIP_TRY(launch_sobf(N0, c->rm[ATUS_KFAM_KIN], false, w, w, nullptr, np, 1, c->inner_of(0), c->tw[N0], sa,
                          lim(c, ATUS_KFAM_KIN), c->stream));
      for (int a = 1; a <= last; ++a) ATUS_TRY(pass(c, a, +1, w, w, 1, 1.0));
    }
  }
  for (int i = 0; i < nw; ++i) {
    RedProj fp{c->psi + i * np, c->grad + i * np, c->sob + i * np};
    const unsigned nb = red_blocks(np);
    hipLaunchKernelGGL((k_reduce<RedProj>), dim3(nb), dim3(RED_THREADS), 0, c->stream, fp, np, c->red_part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_reduce_final<2>), dim3(1), dim3(RED_THREADS), 0, c->stream, c->red_part, (int)nb,
                       &st->proj[i][0]);
    HIP_TRY(hipGetLastError());
  }
  return ATUS_OK;
}

struct RedMu {
  static constexpr int NV = 1;
  const double2* psi;
  const double2* lap;
  const double* pot;
  long long npts;
  int nw, comp;
  double g0, g1;
  __device__ void operator()(long long l, double (&acc)[1]) const {
    double nl = 0.0;
    for (int j = 0; j < nw; ++j) {
      const double2 a = psi[j * npts + l];
      nl += (j == 0 ? g0 : g1) * (a.x * a.x + a.y * a.y);
    }
    const double2 p = psi[comp * npts + l], lp = lap[comp * npts + l];
    acc[0] += -(p.x * lp.x + p.y * lp.y) + (pot[comp * npts + l] + nl) * (p.x * p.x + p.y * p.y);
  }
};

}  // namespace

int atus_sob_setup(atus_ctx* c, const double* alpha_iso, double stepsize) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_sob_setup"));
  if (c->fs) return fail(ATUS_ERR_UNSUPPORTED, "the Sobolev minimiser covers axes up to 4096 points");
  if (c->ncomp > 2) return fail(ATUS_ERR_UNSUPPORTED, "Sobolev minimiser supports 1 or 2 wavefunctions");
  if (!alpha_iso) return fail(ATUS_ERR_ARG, "null alpha");
  if (!c->pot) return fail(ATUS_ERR_STATE, "set the potential before atus_sob_setup");
  for (int i = 0; i < c->ncomp; ++i) c->sob_alpha[i] = alpha_iso[i];
  c->sob_step = stepsize;
  const size_t bytes = sizeof(double2) * c->npts * c->ncomp;
  if (!c->sob && hipMalloc(&c->sob, bytes) != hipSuccess) return fail(ATUS_ERR_NOMEM, "Sobolev buffers");
  if (!c->lap && hipMalloc(&c->lap, bytes) != hipSuccess) return fail(ATUS_ERR_NOMEM, "Sobolev buffers");
  if (!c->grad && hipMalloc(&c->grad, bytes) != hipSuccess) return fail(ATUS_ERR_NOMEM, "Sobolev buffers");
  if ((!c->sob_st && hipMalloc(&c->sob_st, sizeof(SobState)) != hipSuccess) ||
      (!c->sob_part && hipMalloc(&c->sob_part, sizeof(double) * 2 * RED_MAX_BLOCKS * 2) != hipSuccess) ||
      (!c->sob_log && hipMalloc(&c->sob_log, sizeof(double) * SOB_MAX_CHUNK * SOB_LOG_VALS) != hipSuccess))
    return fail(ATUS_ERR_NOMEM, "Sobolev loop state");
  c->sob_ready = true;
  return ATUS_OK;
}

int atus_sob_renormalize(atus_ctx* c) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_sob_renormalize"));
  std::vector<double> N(c->ncomp);
  for (int i = 0; i < c->ncomp; ++i) ATUS_TRY(atus_particle_number(c, i, &N[i]));
  for (int i = 0; i < c->ncomp; ++i) {
    hipLaunchKernelGGL(k_scale, dim3(ew_blocks(c->npts)), dim3(256), 0, c->stream, c->psi + i * c->npts, c->npts,
                       1.0 / sqrt(N[i]));
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ATUS_OK;
}

int atus_sob_iterate(atus_ctx* c, int count, double eps, int* performed, double* res_tot, double* log) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_sob_iterate"));
  if (!c->sob_ready) return fail(ATUS_ERR_STATE, "call atus_sob_setup first");
  if (count < 1 || count > SOB_MAX_CHUNK)
    return fail(ATUS_ERR_ARG, "count must be 1.." + std::to_string(SOB_MAX_CHUNK));
  const int nw = c->ncomp;
  const long long np = c->npts;
  const unsigned nb = red_blocks(np);
  SobState* st = (SobState*)c->sob_st;
  auto finalize = [&](int init, int slot) -> int {
    hipLaunchKernelGGL(k_sob_finalize, dim3(1), dim3(RED_THREADS), 0, c->stream, st, c->sob_part, (int)nb, nw, c->ar(),
                       eps, init, c->sob_log, slot);
    HIP_TRY(hipGetLastError());
    return ATUS_OK;
  };
  // N of the current psi, the renormalisation of the first iteration
  for (int i = 0; i < nw; ++i) {
    hipLaunchKernelGGL(k_sob_update_dev, dim3(nb), dim3(RED_THREADS), 0, c->stream, c->psi + i * np,
                       (const double2*)nullptr, (const double2*)nullptr, np, 0.0, (const SobState*)st, i, 1,
                       c->sob_part + (size_t)i * nb * 2);
    HIP_TRY(hipGetLastError());
  }
  ATUS_TRY(finalize(1, 0));
  for (int k = 0; k < count; ++k) {
    ATUS_TRY(sob_body(c));
    for (int i = 0; i < nw; ++i) {
      hipLaunchKernelGGL(k_sob_update_dev, dim3(nb), dim3(RED_THREADS), 0, c->stream, c->psi + i * np,
                         (const double2*)c->grad + i * np, (const double2*)c->sob + i * np, np, c->sob_step,
                         (const SobState*)st, i, 0, c->sob_part + (size_t)i * nb * 2);
      HIP_TRY(hipGetLastError());
    }
    ATUS_TRY(finalize(0, k));
  }
  SobState h;
  HIP_TRY(hipMemcpyAsync(&h, st, sizeof(SobState), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> lg;
  if (log) {
    lg.resize((size_t)count * SOB_LOG_VALS);
    HIP_TRY(hipMemcpyAsync(lg.data(), c->sob_log, sizeof(double) * lg.size(), hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (performed) *performed = (int)h.iters;
  if (res_tot) *res_tot = h.res_tot;
  if (log) std::copy(lg.begin(), lg.begin() + (size_t)h.iters * SOB_LOG_VALS, log);
  if (!std::isfinite(h.res_tot)) return fail(ATUS_ERR_NUMERIC, "Sobolev residual is not finite");
  return ATUS_OK;
}

int atus_sob_iter(atus_ctx* c, double* res_tot, int* renormalised, double* N_out) {
  double e[SOB_LOG_VALS];
  int done = 0;
  // eps < 0: the iteration is always performed
  ATUS_TRY(atus_sob_iterate(c, 1, -1.0, &done, nullptr, e));
  if (res_tot) *res_tot = e[3];
  if (renormalised) *renormalised = e[2] != 0.0 ? 1 : 0;
  if (N_out)
    for (int i = 0; i < c->ncomp; ++i) N_out[i] = e[i];
  return ATUS_OK;
}

int atus_sob_mu(atus_ctx* c, double* mu) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_sob_mu"));
  if (!c->sob_ready) return fail(ATUS_ERR_STATE, "call atus_sob_setup first");
  if (!mu) return fail(ATUS_ERR_ARG, "null output");
  const int nw = c->ncomp;
  for (int i = 0; i < nw; ++i) ATUS_TRY(sob_filter(c, i, c->psi, c->sob, c->lap));
  for (int i = 0; i < nw; ++i) {
    RedMu fn{c->psi, c->lap, c->pot, c->npts, nw, i, c->gs[0 + i * nw], nw > 1 ? c->gs[1 + i * nw] : 0.0};
    double s;
    ATUS_TRY(reduce(c, fn, c->npts, &s));
    mu[i] = c->ar() * s;
  }
  return ATUS_OK;
}

int atus_sob_run(atus_ctx* c, double eps, long long maxit, long long* iters, double* res_tot) {
  ATUS_TRY(check_ctx(c));
  ATUS_TRY(check_single(c, "atus_sob_run"));
  if (!(eps >= 0.0)) return fail(ATUS_ERR_ARG, "eps must be >= 0");
  long long it = 0;
  double r = 0.0, r_prev = 0.0;
  int chunk = 16;
  for (;;) {
    int n = chunk;
    if (maxit > 0 && maxit - it < n) n = (int)(maxit - it);
    int done = 0;
    ATUS_TRY(atus_sob_iterate(c, n, eps, &done, &r, nullptr));
    it += done;
    if (r <= eps || (maxit > 0 && it >= maxit)) break;
    // the next chunk: about the iterations still needed at the residual's
    // recent rate of decrease (iterations past convergence are enqueued but
    // leave psi unchanged), 16 to 512
    int next = 64;
    if (r_prev > 0 && r < r_prev && r > 0) {
      const double rate = std::log(r_prev / r) / done;
      const double need = std::log(r / eps) / rate;
      next = (int)std::min(512.0, std::max(16.0, 0.9 * need));
    }
    chunk = next;
    r_prev = r;
  }
  if (iters) *iters = it;
  if (res_tot) *res_tot = r;
  return ATUS_OK;
}

}  // extern "C"

namespace {

// runs `block` with every launch bracketed by an event pair on its stream;
// per kernel family: summed ms, launches and grid share
template <class Block>
int profiled(atus_ctx* c, Block block, double* ms, int* count, double* grid) {
  ATUS_TRY(check_ctx(c));
  if (!ms || !count) return fail(ATUS_ERR_ARG, "null output");
  for (int f = 0; f < ATUS_NKFAM; ++f) {
    ms[f] = 0;
    count[f] = 0;
    if (grid) grid[f] = 0;
  }
  c->prof = true;
  c->prof_ev.clear();
  c->prof_fam.clear();
  c->prof_frac.clear();
  int rc = block();
  c->prof = false;
  if (rc == ATUS_OK) {
    if (hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(ATUS_ERR_HIP, "synchronize after profiled block");
  }
  for (size_t i = 0; rc == ATUS_OK && i + 1 < c->prof_ev.size() && i / 2 < c->prof_fam.size(); i += 2) {
    float e = 0;
    if (hipEventElapsedTime(&e, c->prof_ev[i], c->prof_ev[i + 1]) != hipSuccess) {
      rc = fail(ATUS_ERR_HIP, "event timing");
      break;
    }
    const int f = c->prof_fam[i / 2];
    ms[f] += e;
    count[f] += 1;
    if (grid) grid[f] += c->prof_frac[i / 2];
  }
  for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
  c->prof_ev.clear();
  c->prof_fam.clear();
  c->prof_frac.clear();
  return rc;
}

}  // namespace

extern "C" {

int atus_run_block_profiled(atus_ctx* c, const atus_step_params* sp, int Nk, double* t, double* ms, int* count,
                            double* grid) {
  return profiled(c, [&]() { return atus_run_block(c, sp, Nk, t); }, ms, count, grid);
}

int atus_batch_run_block_profiled(atus_ctx* c, int kind, double dt, int Nk, double* t, const double* F, double* ms,
                                  int* count, double* grid) {
  return profiled(c, [&]() { return atus_batch_run_block(c, kind, dt, Nk, t, F); }, ms, count, grid);
}

int atus_copy_bandwidth(atus_ctx* c, int reps, double* gbps) {
  ATUS_TRY(check_ctx(c));
  if (reps < 1 || !gbps) return fail(ATUS_ERR_ARG, "need reps >= 1 and an output");
  const long long n = c->npts * c->ncomp;
  double2* dst = nullptr;
  if (hipMalloc(&dst, sizeof(double2) * n) != hipSuccess) return fail(ATUS_ERR_NOMEM, "copy probe buffer");
  long long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)nb), dim3(256), 0, c->stream, c->psi, dst, n);  // warm up
  int rc = ATUS_OK;
  if (hipEventRecord(c->ev0, c->stream) != hipSuccess) rc = fail(ATUS_ERR_HIP, "event record");
  for (int r = 0; rc == ATUS_OK && r < reps; ++r)
    hipLaunchKernelGGL(k_copy, dim3((unsigned)nb), dim3(256), 0, c->stream, c->psi, dst, n);
  float ms = 0;
  if (rc == ATUS_OK && (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess ||
                        hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess))
    rc = fail(ATUS_ERR_HIP, "copy probe timing");
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(dst);
  if (rc != ATUS_OK) return rc;
  *gbps = 2.0 * sizeof(double2) * (double)n * reps / (ms * 1e-3) / 1e9;
  return ATUS_OK;
}

int atus_set_slab(atus_ctx* c, int planes) {
  ATUS_TRY(check_ctx(c));
  if (planes < -1) return fail(ATUS_ERR_ARG, "slab planes must be -1 (automatic), 0 (off) or positive");
  c->slab = planes;
  return ATUS_OK;
}

int atus_set_resident(atus_ctx* c, int on) {
  ATUS_TRY(check_ctx(c));
  if (on != 0 && on != 1) return fail(ATUS_ERR_ARG, "resident blocks: 0 (off) or 1 (on)");
  c->resident = on;
  return ATUS_OK;
}

int atus_tune(atus_ctx* c, int family, int values_per_thread) {
  ATUS_TRY(check_ctx(c));
  if (family < 0 || family >= ATUS_NKFAM) return fail(ATUS_ERR_ARG, "unknown kernel family");
  if (values_per_thread != 8 && values_per_thread != 16) return fail(ATUS_ERR_ARG, "values per thread must be 8 or 16");
  c->rm[family] = values_per_thread;
  return ATUS_OK;
}

int atus_tune_threads(atus_ctx* c, int family, int threads) {
  ATUS_TRY(check_ctx(c));
  if (family < 0 || family >= ATUS_NKFAM) return fail(ATUS_ERR_ARG, "unknown kernel family");
  if (threads != 0 && threads != 64 && threads != 128 && threads != 256 && threads != 512)
    return fail(ATUS_ERR_ARG, "threads per workgroup must be 0 (automatic), 64, 128, 256 or 512");
  c->wg[family] = threads;
  return ATUS_OK;
}

int atus_set_timing(atus_ctx* c, int enable) {
  ATUS_TRY(check_ctx(c));
  c->timing = enable != 0;
  return ATUS_OK;
}

int atus_last_block_ms(atus_ctx* c, double* ms) {
  ATUS_TRY(check_ctx(c));
  if (!ms) return fail(ATUS_ERR_ARG, "null output");
  *ms = c->timing ? c->last_ms : 0.0;
  return ATUS_OK;
}


// ---------------------------------------------------------------- batches

int atus_batch_create(atus_ctx** out, int dim, const long long n[3], const double dx[3], const double dk[3],
                      int ncomp_per_run, int nbatch) {
  if (ncomp_per_run != 2) return fail(ATUS_ERR_UNSUPPORTED, "batched runs have two components");
  return ctx_create(out, dim, n, dx, dk, ncomp_per_run, nbatch, true);
}

int atus_batch_set_runs(atus_ctx* c, const atus_batch_run* runs, int nbatch) {
  ATUS_TRY(check_ctx(c));
  if (!runs) return fail(ATUS_ERR_ARG, "null run parameters");
  if (!c->runs) return fail(ATUS_ERR_STATE, "not a batch context (atus_batch_create)");
  if (nbatch != c->nbatch)
    return fail(ATUS_ERR_ARG, "expected " + std::to_string(c->nbatch) + " runs, got " + std::to_string(nbatch));
  std::vector<RunParams> h(nbatch);
  for (int r = 0; r < nbatch; ++r) {
    const atus_batch_run& a = runs[r];
    RunParams& p = h[r];
    for (int i = 0; i < 4; ++i) p.gs[i] = a.gs[i];
    p.amp0 = a.laser.amp0;
    p.laser_k = a.laser.laser_k;
    p.laser_domh = a.laser.laser_domh;
    p.laser_dk = a.laser.laser_dk;
    p.chirp = a.laser.chirp;
    p.chirp_rate0 = a.laser.chirp_rate0;
    p.phase0 = a.laser.phase0;
    p.beta0 = a.laser.beta0;
    for (int i = 0; i < 4; ++i) p.deltaL[i] = a.laser.deltaL[i];
  }
  HIP_TRY(hipMemcpyAsync(c->runs, h.data(), sizeof(RunParams) * nbatch, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->runs_set = true;
  return ATUS_OK;
}

int atus_batch_run_block(atus_ctx* c, int kind, double dt, int Nk, double* t, const double* F) {
  ATUS_TRY(check_ctx(c));
  if (!t) return fail(ATUS_ERR_ARG, "null time");
  if (Nk < 1) return fail(ATUS_ERR_ARG, "Nk must be >= 1");
  if (!c->runs || c->nc_run != 2) return fail(ATUS_ERR_STATE, "not a batch context (atus_batch_create)");
  if (!c->runs_set) return fail(ATUS_ERR_STATE, "run parameters not set (atus_batch_set_runs)");
  const int vk = vkind_of(kind);
  if (vk != V_BRAGG && vk != V_NL_IF) return fail(ATUS_ERR_ARG, "batches run bragg_ad or freeprop (IF) steps");
  const int B = c->nbatch;
  const size_t nf = (size_t)Nk * B;
  std::vector<double> ones;
  if (!F) {
    ones.assign(nf, 1.0);
    F = ones.data();
  }
  for (size_t i = 0; i < nf; ++i)
    if (!std::isfinite(F[i])) return fail(ATUS_ERR_NUMERIC, "non-finite amplitude in the table");
  if (c->ftab_cap < nf) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(c->ftab);
    c->ftab = nullptr;
    c->ftab_cap = 0;
    if (hipMalloc(&c->ftab, sizeof(double) * nf) != hipSuccess) return fail(ATUS_ERR_NOMEM, "amplitude table");
    c->ftab_cap = nf;
  }
  HIP_TRY(hipMemcpyAsync(c->ftab, F, sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
  ATUS_TRY(ensure_kin(c, dt));
  VArgs base;
  batch_vargs(c, base, dt);
  if (c->timing) HIP_TRY(hipEventRecord(c->ev0, c->stream));
  StepFeed fd;
  fd.fbatch = c->ftab;
  ATUS_TRY(batch_block(c, vk, base, Nk, dt, t, fd));
  return finish_block(c);
}

int atus_batch_particle_numbers(atus_ctx* c, double* out) {
  ATUS_TRY(check_ctx(c));
  if (!c->runs) return fail(ATUS_ERR_STATE, "not a batch context (atus_batch_create)");
  if (!out) return fail(ATUS_ERR_ARG, "null output");
  RedNorm fn{c->psi};
  ATUS_TRY(reduce_seg(c, fn, c->npts, c->npts, c->ncomp, out));
  for (int i = 0; i < c->ncomp; ++i) {
    out[i] *= c->ar();
    if (!std::isfinite(out[i])) return fail(ATUS_ERR_NUMERIC, "particle number is not finite");
  }
  return ATUS_OK;
}

int atus_batch_particle_numbers_noalloc(atus_ctx* c, double* out) {
  ATUS_TRY(check_ctx(c));
  if (!c->runs) return fail(ATUS_ERR_STATE, "not a batch context (atus_batch_create)");
  if (!out) return fail(ATUS_ERR_ARG, "null output");
  double* ptr = (c->device >= 0 && c->device < MAX_SCRATCH_DEVICES) ? g_scratch_ptr[c->device] : nullptr;
  if (!ptr) return fail(ATUS_ERR_STATE, "no device scratch on this context's device");
  // the same kernels and workgroup count as atus_batch_particle_numbers:
  // nseg * nb partials, then nseg results after them
  const int nseg = c->ncomp;
  const unsigned nb = seg_blocks(c->npts, nseg);
  const size_t need = (size_t)nseg * nb + (size_t)nseg;
  if (need > NOALLOC_SCRATCH_DOUBLES)
    return fail(ATUS_ERR_ARG, "batch of " + std::to_string(nseg) + " components needs " + std::to_string(need) +
                                  " doubles of device scratch; the allocation-free entry point holds " +
                                  std::to_string(NOALLOC_SCRATCH_DOUBLES) + " (use atus_batch_particle_numbers)");
  RedNorm fn{c->psi};
  ATUS_TRY(reduce_seg_into(c, fn, c->npts, c->npts, nseg, nb, ptr, ptr + (size_t)nseg * nb, out));
  for (int i = 0; i < c->ncomp; ++i) {
    out[i] *= c->ar();
    if (!std::isfinite(out[i])) return fail(ATUS_ERR_NUMERIC, "particle number is not finite");
  }
  return ATUS_OK;
}

int atus_batch_rabi_integrals(atus_ctx* c, int comp, const double* k0, int n, double threshold, double* out) {
  ATUS_TRY(check_ctx(c));
  if (!c->runs) return fail(ATUS_ERR_STATE, "not a batch context (atus_batch_create)");
  if (comp < 0 || comp >= c->nc_run) return fail(ATUS_ERR_ARG, "component index within a run out of range");
  if (n < 1 || n > RED_MAX_NV || !k0 || !out) return fail(ATUS_ERR_ARG, "need 1..8 momentum points");
  const long long cs = (long long)c->nc_run * c->npts;
  double2* p = c->psi + (long long)comp * c->npts;
  {
    const RabiReduction rr = rabi_reduction(c, p, k0, n, threshold);
    if (rabi_box3_ok(c, rr)) {
      ATUS_TRY(ensure_rabi_box(c, rr, c->nbatch));
      ATUS_TRY(rabi_box3(c, p, cs, c->nbatch, rr, c->sred_out));
      std::vector<double> s((size_t)c->nbatch * RED_MAX_NV);
      HIP_TRY(hipMemcpyAsync(s.data(), c->sred_out, sizeof(double) * s.size(), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      for (int r = 0; r < c->nbatch; ++r)
        for (int q = 0; q < n; ++q) out[(size_t)r * n + q] = c->ar_k() * s[(size_t)r * RED_MAX_NV + q];
      return ATUS_OK;
    }
  }
  ATUS_TRY(full_ft(c, -1, p, p, c->nbatch, cs));
  const RabiReduction rr = rabi_reduction(c, p, k0, n, threshold);
  std::vector<double> s((size_t)c->nbatch * RED_MAX_NV, 0.0);
  if (rr.use_box) {
    ATUS_TRY(reduce_seg(c, rr.box, rr.npts, cs, c->nbatch, s.data()));
  } else {
    ATUS_TRY(reduce_seg(c, rr.full, c->npts, cs, c->nbatch, s.data()));
  }
  ATUS_TRY(full_ft(c, +1, p, p, c->nbatch, cs));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int r = 0; r < c->nbatch; ++r)
    for (int q = 0; q < n; ++q) out[(size_t)r * n + q] = c->ar_k() * s[(size_t)r * RED_MAX_NV + q];
  return ATUS_OK;
}


// ---------------------------------------------------------------- graphs

}  // extern "C"

struct atus_graph {
  atus_ctx* c = nullptr;
  bool batch = false;
  int vk = 0, Nk = 0;
  double dt = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double* dev = nullptr;   // t1 of each step [Nk], then the amplitudes [Nk][nbatch]
  double* host = nullptr;  // pinned staging of the same
  size_t nf = 0;           // amplitudes per block
  // the context state the captured launches took by value or by pointer:
  // a replay refuses to run once the context no longer matches it
  const double* pot = nullptr;
  double gs[16] = {0};
  bool alpha2_set = false;
};

extern "C" {

int atus_graph_destroy(atus_graph* g) {
  if (!g) return ATUS_OK;
  if (g->c) {
    (void)hipSetDevice(g->c->device);
    (void)hipStreamSynchronize(g->c->stream);
    if (g->exec) --g->c->live_graphs;
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  (void)hipFree(g->dev);
  if (g->host) (void)hipHostFree(g->host);
  delete g;
  return ATUS_OK;
}

int atus_graph_create(atus_ctx* c, int kind, double dt, const atus_laser* laser, int Nk, atus_graph** out) {
  ATUS_TRY(check_ctx(c));
  if (!out) return fail(ATUS_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (Nk < 1) return fail(ATUS_ERR_ARG, "Nk must be >= 1");
  const int vk = vkind_of(kind);
  if (vk < 0) return fail(ATUS_ERR_ARG, "unknown step kind " + std::to_string(kind));
  if (vk != V_NL && vk != V_NL_IF && vk != V_BRAGG)
    return fail(ATUS_ERR_ARG, "graphs replay freeprop, freeprop (IF) and bragg_ad blocks");
  const bool batch = c->runs != nullptr;
  if (batch) {
    if (!c->runs_set) return fail(ATUS_ERR_STATE, "run parameters not set (atus_batch_set_runs)");
    if (vk != V_BRAGG && vk != V_NL_IF) return fail(ATUS_ERR_ARG, "batches run bragg_ad or freeprop (IF) steps");
  } else if (vk == V_BRAGG && c->ncomp != 2) {
    return fail(ATUS_ERR_STATE, "bragg_ad needs exactly two components");
  }
  ATUS_TRY(ensure_kin(c, dt));
  atus_graph* g = new atus_graph();
  g->c = c;
  g->batch = batch;
  g->vk = vk;
  g->Nk = Nk;
  g->dt = dt;
  g->nf = (size_t)Nk * (batch ? c->nbatch : 1);
  auto bail = [&](int code, const std::string& m) {
    atus_graph_destroy(g);
    return fail(code, m);
  };
  if (hipMalloc(&g->dev, sizeof(double) * (Nk + g->nf)) != hipSuccess ||
      hipHostMalloc(&g->host, sizeof(double) * (Nk + g->nf), hipHostMallocDefault) != hipSuccess)
    return bail(ATUS_ERR_NOMEM, "graph step tables");
  VArgs base;
  StepFeed fd;
  fd.t1dev = g->dev;
  if (batch) {
    batch_vargs(c, base, dt);
    fd.fbatch = g->dev + Nk;
  } else {
    fill_vargs(c, base, dt, laser);
    fd.fdev = g->dev + Nk;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(ATUS_ERR_HIP, "synchronize");
  const bool prof = c->prof;
  c->prof = false;
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess)
    return bail(ATUS_ERR_HIP, "stream capture");
  double t = 0.0;
  int rc = batch ? batch_block(c, vk, base, Nk, dt, &t, fd) : single_block(c, vk, base, Nk, dt, &t, fd);
  const std::string err = g_err;
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
  c->prof = prof;
  g->graph = graph;
  if (rc != ATUS_OK) return bail(rc, "capturing the block: " + err);
  if (ec != hipSuccess || !graph) return bail(ATUS_ERR_HIP, "stream capture did not produce a graph");
  if (hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    g->exec = nullptr;
    return bail(ATUS_ERR_HIP, "graph instantiation");
  }
  g->pot = c->pot;
  for (int i = 0; i < 16; ++i) g->gs[i] = c->gs[i];
  g->alpha2_set = c->alpha2_set;
  ++c->live_graphs;
  *out = g;
  return ATUS_OK;
}

int atus_graph_launch(atus_graph* g, double* t, const double* F) {
  if (!g || !t) return fail(ATUS_ERR_ARG, "null argument");
  atus_ctx* c = g->c;
  ATUS_TRY(check_ctx(c));
  if (c->pot != g->pot || c->alpha2_set != g->alpha2_set)
    return fail(ATUS_ERR_STATE, "the context's potential or species changed since the graph was captured; "
                                "capture it again");
  for (int i = 0; i < 16; ++i)
    if (c->gs[i] != g->gs[i])
      return fail(ATUS_ERR_STATE, "the context's couplings changed since the graph was captured; capture it again");
  // the kinetic launches read the context's tables in place: rebuild them
  // for the graph's dt if another block or atus_set_alpha changed them
  ATUS_TRY(ensure_kin(c, g->dt));
  // t1 of each step with the block runner's bookkeeping, then amplitudes
  double tt = *t + 0.5 * g->dt;
  for (int m = 1; m <= g->Nk; ++m) {
    g->host[m - 1] = tt + 0.5 * g->dt;
    tt += (m < g->Nk) ? g->dt : 0.5 * g->dt;
  }
  for (size_t i = 0; i < g->nf; ++i) {
    const double f = F ? F[i] : 1.0;
    if (!std::isfinite(f)) return fail(ATUS_ERR_NUMERIC, "non-finite amplitude in the table");
    g->host[g->Nk + i] = f;
  }
  HIP_TRY(hipMemcpyAsync(g->dev, g->host, sizeof(double) * (g->Nk + g->nf), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipGraphLaunch(g->exec, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *t = tt;
  return ATUS_OK;
}

}  // extern "C"

// ------------------------------------------------------------ scan graphs

struct atus_scan_graph {
  std::vector<atus_ctx*> ctxs;
  int vk = 0, Nk = 0, nk0 = 0;
  double dt = 0;
  int device = 0;
  hipStream_t master = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // device tables: t1 [Nk], then per context its amplitudes [Nk][nbatch]
  double* dev = nullptr;
  double* host = nullptr;
  size_t ntab = 0;
  std::vector<size_t> foff;  // offset of each context's amplitudes in the tables
  // Rabi integrals: per context nbatch x RED_MAX_NV sums (device, pinned host)
  double* rdev = nullptr;
  double* rhost = nullptr;
  std::vector<size_t> roff;
  size_t nr = 0;
  bool counted = false;
};

namespace {

// compute_rabi_integrals of component 0 of every run of a batch
// (CRT_Base_IF.h:208-281), transformed in place and back, the sums left in
// device memory at `res` (nbatch x RED_MAX_NV): the launches of
// atus_batch_rabi_integrals without its host copy
int batch_rabi_launch(atus_ctx* c, const double* k0, int n, double threshold, double* res) {
  const long long cs = (long long)c->nc_run * c->npts;
  double2* p = c->psi;
  {
    const RabiReduction rr = rabi_reduction(c, p, k0, n, threshold);
    if (rabi_box3_ok(c, rr)) return rabi_box3(c, p, cs, c->nbatch, rr, res);
  }
  ATUS_TRY(full_ft(c, -1, p, p, c->nbatch, cs));
  const RabiReduction rr = rabi_reduction(c, p, k0, n, threshold);
  const unsigned nb = seg_blocks(rr.npts, c->nbatch);
  if (c->sred_cap < (size_t)c->nbatch * nb * RED_MAX_NV) return fail(ATUS_ERR_STATE, "reduction buffers not sized");
  if (rr.use_box)
    ATUS_TRY(reduce_seg_launch(c, rr.box, rr.npts, cs, c->nbatch, nb, c->sred_part, res));
  else
    ATUS_TRY(reduce_seg_launch(c, rr.full, c->npts, cs, c->nbatch, nb, c->sred_part, res));
  ATUS_TRY(full_ft(c, +1, p, p, c->nbatch, cs));
  return ATUS_OK;
}

}  // namespace

extern "C" {

int atus_scan_graph_destroy(atus_scan_graph* g) {
  if (!g) return ATUS_OK;
  (void)hipSetDevice(g->device);
  if (g->master) (void)hipStreamSynchronize(g->master);
  if (g->counted)
    for (atus_ctx* c : g->ctxs) --c->live_graphs;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  (void)hipFree(g->dev);
  (void)hipFree(g->rdev);
  if (g->host) (void)hipHostFree(g->host);
  if (g->rhost) (void)hipHostFree(g->rhost);
  for (hipEvent_t e : g->joins)
    if (e) (void)hipEventDestroy(e);
  if (g->fork) (void)hipEventDestroy(g->fork);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  if (g->master) (void)hipStreamDestroy(g->master);
  delete g;
  return ATUS_OK;
}

int atus_scan_graph_create(atus_ctx* const* ctxs, int nctx, int kind, double dt, int Nk, const double* k0, int nk0,
                           double threshold, atus_scan_graph** out) {
  if (!out) return fail(ATUS_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (!ctxs || nctx < 1) return fail(ATUS_ERR_ARG, "need at least one batch context");
  if (Nk < 1) return fail(ATUS_ERR_ARG, "Nk must be >= 1");
  if (nk0 < 0 || nk0 > RED_MAX_NV || (nk0 > 0 && !k0)) return fail(ATUS_ERR_ARG, "need 0..8 momentum points");
  const int vk = vkind_of(kind);
  if (vk != V_BRAGG && vk != V_NL_IF) return fail(ATUS_ERR_ARG, "batches run bragg_ad or freeprop (IF) steps");
  for (int i = 0; i < nctx; ++i) {
    atus_ctx* c = ctxs[i];
    ATUS_TRY(check_ctx(c));
    if (!c->runs || c->nc_run != 2) return fail(ATUS_ERR_STATE, "not a batch context (atus_batch_create)");
    if (!c->runs_set) return fail(ATUS_ERR_STATE, "run parameters not set (atus_batch_set_runs)");
    if (c->device != ctxs[0]->device) return fail(ATUS_ERR_ARG, "contexts on different devices");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == c) return fail(ATUS_ERR_ARG, "a context appears twice");
  }
  atus_scan_graph* g = new atus_scan_graph();
  auto bail = [&](int code, const std::string& m) {
    atus_scan_graph_destroy(g);
    return fail(code, m);
  };
  g->ctxs.assign(ctxs, ctxs + nctx);
  g->vk = vk;
  g->Nk = Nk;
  g->nk0 = nk0;
  g->dt = dt;
  g->device = ctxs[0]->device;
  // tables and buffers, everything the captured launches read, allocated
  // before the capture
  g->ntab = (size_t)Nk;
  for (atus_ctx* c : g->ctxs) {
    g->foff.push_back(g->ntab);
    g->ntab += (size_t)Nk * c->nbatch;
    g->roff.push_back(g->nr);
    g->nr += (size_t)c->nbatch * RED_MAX_NV;
    int rc = ensure_kin(c, dt);
    if (rc == ATUS_OK && c->fs) rc = ensure_tbuf(c);
    if (rc == ATUS_OK && nk0 > 0)
      rc = ensure_sred(c, (size_t)c->nbatch * seg_blocks(c->npts, c->nbatch) * RED_MAX_NV, c->nbatch);
    if (rc == ATUS_OK && nk0 > 0) {
      const RabiReduction rr = rabi_reduction(c, c->psi, k0, nk0, threshold);
      if (rabi_box3_ok(c, rr)) rc = ensure_rabi_box(c, rr, c->nbatch);
    }
    if (rc != ATUS_OK) {
      const std::string m = g_err;
      return bail(rc, m);
    }
  }
  if (hipMalloc(&g->dev, sizeof(double) * g->ntab) != hipSuccess ||
      hipHostMalloc(&g->host, sizeof(double) * g->ntab, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&g->rdev, sizeof(double) * std::max<size_t>(g->nr, 1)) != hipSuccess ||
      hipHostMalloc(&g->rhost, sizeof(double) * std::max<size_t>(g->nr, 1), hipHostMallocDefault) != hipSuccess)
    return bail(ATUS_ERR_NOMEM, "scan graph tables");
  g->joins.assign(nctx, nullptr);
  if (hipStreamCreateWithFlags(&g->master, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess)
    return bail(ATUS_ERR_HIP, "scan graph stream and events");
  for (int i = 0; i < nctx; ++i)
    if (hipEventCreateWithFlags(&g->joins[i], hipEventDisableTiming) != hipSuccess)
      return bail(ATUS_ERR_HIP, "scan graph events");
  for (atus_ctx* c : g->ctxs)
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(ATUS_ERR_HIP, "synchronize");
  // capture: every context's block and Rabi integrals on its own stream,
  // forked from and joined back into the master stream
  std::vector<bool> prof(nctx);
  for (int i = 0; i < nctx; ++i) {
    prof[i] = g->ctxs[i]->prof;
    g->ctxs[i]->prof = false;
  }
  int rc = ATUS_OK;
  std::string err;
  if (hipStreamBeginCapture(g->master, hipStreamCaptureModeRelaxed) != hipSuccess)
    return bail(ATUS_ERR_HIP, "stream capture");
  if (hipEventRecord(g->fork, g->master) != hipSuccess) rc = fail(ATUS_ERR_HIP, "fork event");
  for (int i = 0; rc == ATUS_OK && i < nctx; ++i) {
    atus_ctx* c = g->ctxs[i];
    if (hipStreamWaitEvent(c->stream, g->fork, 0) != hipSuccess) {
      rc = fail(ATUS_ERR_HIP, "fork");
      break;
    }
    VArgs base;
    batch_vargs(c, base, dt);
    StepFeed fd;
    fd.t1dev = g->dev;
    fd.fbatch = g->dev + g->foff[i];
    double t = 0.0;
    rc = batch_block(c, vk, base, Nk, dt, &t, fd);
    if (rc == ATUS_OK && nk0 > 0) rc = batch_rabi_launch(c, k0, nk0, threshold, g->rdev + g->roff[i]);
    if (rc == ATUS_OK && (hipEventRecord(g->joins[i], c->stream) != hipSuccess ||
                          hipStreamWaitEvent(g->master, g->joins[i], 0) != hipSuccess))
      rc = fail(ATUS_ERR_HIP, "join");
  }
  err = g_err;
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(g->master, &graph);
  g->graph = graph;
  for (int i = 0; i < nctx; ++i) g->ctxs[i]->prof = prof[i];
  if (rc != ATUS_OK) return bail(rc, "capturing the scan: " + err);
  if (ec != hipSuccess || !graph) return bail(ATUS_ERR_HIP, "stream capture did not produce a graph");
  if (hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    g->exec = nullptr;
    return bail(ATUS_ERR_HIP, "graph instantiation");
  }
  for (atus_ctx* c : g->ctxs) ++c->live_graphs;
  g->counted = true;
  *out = g;
  return ATUS_OK;
}

int atus_scan_graph_launch(atus_scan_graph* g, double* t, const double* F, double* rabi, double* gpu_ms) {
  if (!g || !t) return fail(ATUS_ERR_ARG, "null argument");
  if (hipSetDevice(g->device) != hipSuccess) return fail(ATUS_ERR_HIP, "hipSetDevice failed");
  for (atus_ctx* c : g->ctxs) ATUS_TRY(ensure_kin(c, g->dt));
  // t1 of each step with the block runner's bookkeeping, then amplitudes
  double tt = *t + 0.5 * g->dt;
  for (int m = 1; m <= g->Nk; ++m) {
    g->host[m - 1] = tt + 0.5 * g->dt;
    tt += (m < g->Nk) ? g->dt : 0.5 * g->dt;
  }
  for (size_t i = (size_t)g->Nk; i < g->ntab; ++i) {
    const double f = F ? F[i - g->Nk] : 1.0;
    if (!std::isfinite(f)) return fail(ATUS_ERR_NUMERIC, "non-finite amplitude in the table");
    g->host[i] = f;
  }
  // the contexts' own pending work first
  for (atus_ctx* c : g->ctxs) HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpyAsync(g->dev, g->host, sizeof(double) * g->ntab, hipMemcpyHostToDevice, g->master));
  HIP_TRY(hipEventRecord(g->ev0, g->master));
  HIP_TRY(hipGraphLaunch(g->exec, g->master));
  HIP_TRY(hipEventRecord(g->ev1, g->master));
  if (g->nk0 > 0)
    HIP_TRY(hipMemcpyAsync(g->rhost, g->rdev, sizeof(double) * g->nr, hipMemcpyDeviceToHost, g->master));
  HIP_TRY(hipStreamSynchronize(g->master));
  if (gpu_ms) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    *gpu_ms = ms;
  }
  if (rabi && g->nk0 > 0) {
    size_t o = 0;
    for (size_t i = 0; i < g->ctxs.size(); ++i) {
      const atus_ctx* c = g->ctxs[i];
      for (int r = 0; r < c->nbatch; ++r)
        for (int q = 0; q < g->nk0; ++q) rabi[o++] = c->ar_k() * g->rhost[g->roff[i] + (size_t)r * RED_MAX_NV + q];
    }
  }
  *t = tt;
  return ATUS_OK;
}

}  // extern "C"
