// kernels.hpp -- the HIP kernels of the split-step / Sobolev path.
//
// Data: complex128 (double2), row-major [n0][n1][n2], axis 2 fastest;
// components are npts apart in one allocation.  A transform along axis a
// sees the array as [outer][N][inner] (inner = product of the faster axes).
//
// Kernels (all launched with blockDim.x = T * P, P = N / R threads per pencil,
// R = min(RM, N) values per thread):
//   k_pass   : one unnormalised DFT direction along one axis, times a scale
//   k_kin    : forward DFT along axis 0, multiply by the kinetic phase, inverse
//              DFT along axis 0 (the whole k-space step of a 1D grid; the
//              innermost step of 2D/3D grids); optionally preceded by the
//              pointwise potential / Bragg step (1D only, all components
//              in one thread)
//   k_vfuse  : inverse DFT along the last (contiguous) axis, pointwise
//              potential / Bragg step, forward DFT along the same axis
//   k_sobf   : forward DFT along axis 0, multiply by 1/(1+alpha k^2) and
//              -alpha k^2, two inverse DFTs (Sobolev psi and Laplacian)
//   k_point  : standalone pointwise steps
//   k_reduce : deterministic two-stage reductions
#pragma once
#include "fft_core.hpp"

namespace atus {

// V_NL_IF2 / V_BRAGG_DS: two species, components [NC/2, NC) are species 2
// V_DBRAGG / V_NUM_BRAGG / V_RAMAN: three-state couplings (standalone
// pointwise kernel only)
enum VKind {
  V_NONE = 0,
  V_NL = 1,
  V_NL_IF = 2,
  V_BRAGG = 3,
  V_NL_IF2 = 4,
  V_BRAGG_DS = 5,
  V_DBRAGG = 6,
  V_NUM_BRAGG = 7,
  V_RAMAN = 8
};

struct Grid {
  int dim;
  int n[3];
  double dx[3];
  double dk[3];
  long long npts;
};

// geometry of one axis pass
struct AxisGeo {
  long long inner;      // element stride of the transformed axis
  long long outer;      // number of [N][inner] slabs
  long long col_tiles;  // strided: inner / T tiles per slab
  long long ntiles;     // gridDim.x
  int T;                // pencils per workgroup
  int ls;               // LDS elements per pencil region (N plus padding)
};

// per-run parameters of a batch of independent two-component runs: the
// members of one Bragg_single / CRT_Base_IF instance that differ per run
struct RunParams {
  double gs[4];
  double amp0, laser_k, laser_domh, laser_dk, chirp, chirp_rate0, phase0, beta0;
  double deltaL[4];
};

// parameters of the pointwise step V (CRT_Base::Do_NL_Step,
// CRT_Base_IF::Do_NL_Step, Bragg_single::Do_Bragg_ad)
struct VArgs {
  double dt;
  double gs[16];            // row-major ncomp x ncomp
  const double* pot;        // NULL: potential not initialised
  long long pot_stride;     // npts
  double beta0;
  double deltaL[4];
  // Bragg
  double t1;                // Get_t() + 0.5*dt
  double F;                 // Amplitude_at_time()
  double amp0, laser_k, laser_domh, laser_dk, chirp, chirp_rate0, phase0;
  // batches (kernels instantiated with BATCH): run r owns the elements
  // [r * run_elems, (r + 1) * run_elems), components npts apart
  const RunParams* runs;    // per run
  const double* Fm;         // Amplitude_at_time() of each run at this step
  long long pencils_per_run;
  long long run_shift;      // run_elems - npts
  // replayed blocks (hipGraph): t1 and, for single runs, F of this step
  // come from device memory the host refills before each replay
  const double* t1p;
  const double* Fp;
  // second species (CRT_Base_IF_2): its laser and gravity, and the pulse
  // gates of bragg_ds (1: on, 0: off) for this step
  double amp2, laser_k2, laser_domh2, laser_dk2, chirp_rate2, phase2, beta2;
  double mode1, mode2;
  // three-state steps: Amp[1], the whole Beta vector (beta.x over all axes)
  // and Get_t() before the step (the numerical Bragg step reads it)
  double amp1;
  double beta[3];
  double t0;
};

// the parameters of run r replace the shared ones
__device__ __forceinline__ void bind_run(VArgs& v, long long r) {
  const RunParams& p = v.runs[r];
#pragma unroll
  for (int i = 0; i < 4; ++i) v.gs[i] = p.gs[i];
  v.amp0 = p.amp0;
  v.laser_k = p.laser_k;
  v.laser_domh = p.laser_domh;
  v.laser_dk = p.laser_dk;
  v.chirp = p.chirp;
  v.chirp_rate0 = p.chirp_rate0;
  v.phase0 = p.phase0;
  v.beta0 = p.beta0;
#pragma unroll
  for (int i = 0; i < 4; ++i) v.deltaL[i] = p.deltaL[i];
  v.F = v.Fm[r];
}

// sin and cos: Taylor series (to x^11 and x^10) for |x| < 2^-5, where the
// truncation error is below 1e-27 relative; the library routine otherwise.
// The phases of the potential and Bragg steps are dt*E, usually small.
__device__ __forceinline__ void sincos_fast(double x, double* s, double* c) {
  if (fabs(x) < 0.03125) {
    const double x2 = x * x;
    *s = x * (1.0 + x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 * (1.0 / 362880.0 +
                                                                                  x2 * (-1.0 / 39916800.0))))));
    *c = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0 + x2 * (-1.0 / 3628800.0)))));
  } else {
    sincos(x, s, c);
  }
}

// lookup tables (twiddles, kinetic phases) of length N are staged in LDS
// behind the pencil regions when they are small enough
template <int N>
constexpr bool tables_in_lds() {
  return N <= 1024;
}
// k_kin's tables in LDS: the twiddles its plan reads (plan_tw_elems, at
// most N - 1 entries) when N <= 1024; the phase table e0 (N entries) as
// well, except for strided columns of 1024 points.  There the 4 pencils of
// a 256-thread workgroup at 16 values per thread already take 64 KiB of
// LDS: with the twiddles (16128 B) two workgroups still fit a CU's 160 KiB,
// with e0 on top only one would.
template <int N, bool CONTIG>
constexpr bool kin_tw_in_lds() {
  return tables_in_lds<N>();
}
template <int N, bool CONTIG>
constexpr bool kin_phase_in_lds() {
  return CONTIG ? tables_in_lds<N>() : N <= 512;
}
template <int N, int RM>
__host__ __device__ constexpr int plan_tw_elems() {
  return Plan<N, RM>::template tw_off<Plan<N, RM>::S>();
}
// LDS elements of k_kin's tables
template <int N, int RM, bool CONTIG>
__host__ __device__ constexpr int kin_lds_table_elems() {
  return (kin_tw_in_lds<N, CONTIG>() ? plan_tw_elems<N, RM>() : 0) + (kin_phase_in_lds<N, CONTIG>() ? N : 0);
}

// copies src[0..N) to lds_tab (then a workgroup barrier) and returns the
// pointer the kernel should read; the global table is returned otherwise
template <int N, bool IN_LDS = tables_in_lds<N>()>
__device__ __forceinline__ const double2* stage_table(double2* lds_tab, const double2* src) {
  if constexpr (IN_LDS) {
    for (int j = threadIdx.x; j < N; j += blockDim.x) lds_tab[j] = src[j];
    return lds_tab;
  } else {
    return src;
  }
}

// The global twiddle table (N > 1024) as a pointer the compiler cannot
// prove equal to earlier ones: the loads of a transform are not merged with
// the identical loads of the transform before it, which would keep all of
// its twiddles live in registers in between.  The table stays a global
// (not a flat) address.
typedef const __attribute__((address_space(1))) double2* global_double2_ptr;
template <int N, bool IN_LDS = tables_in_lds<N>()>
__device__ __forceinline__ auto fresh_tw(const double2* p) {
  if constexpr (!IN_LDS) {
    asm volatile("" : "+s"(p));
    return (global_double2_ptr)p;
  } else {
    return p;
  }
}

// blockIdx.x -> tile so that the workgroups one XCD runs (round-robin
// dispatch: XCD = blockIdx.x % 8) cover one contiguous range of tiles
__device__ __forceinline__ long long xcd_tile(long long b, long long nb) {
  const long long per = nb >> 3, rem = nb & 7;
  const long long xcd = b & 7, idx = b >> 3;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

template <int N, int RM, bool CONTIG>
struct PencilMap {
  int t;
  int p;
  bool valid;
  long long base;    // element offset of pencil element 0
  long long stride;  // element stride along the pencil
  long long pid;     // pencil id: CONTIG -> pencil index; strided -> o*inner + col
  __device__ __forceinline__ PencilMap(const AxisGeo& g) {
    constexpr int P = Plan<N, RM>::P;
    const long long tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x;
    if constexpr (CONTIG) {
      p = tid / P;
      t = tid % P;
      pid = tile * g.T + p;
      valid = pid < g.outer;
      base = pid * (long long)N;
      stride = 1;
    } else {
      p = tid % g.T;
      t = tid / g.T;
      const long long o = tile / g.col_tiles;
      const long long col = (tile % g.col_tiles) * g.T + p;
      valid = col < g.inner;
      base = o * (long long)N * g.inner + col;
      stride = g.inner;
      pid = o * g.inner + col;
    }
  }
  __device__ __forceinline__ long long elem(int j) const { return base + (long long)j * stride; }
};

// element t + P*i of the pencil: one 64-bit multiply for the first, then
// pointer increments of P*stride (64-bit index products per element cost
// several VALU instructions each)
template <int N, int RM>
__device__ __forceinline__ void load_pencil(double2 (&v)[(Plan<N, RM>::R)], const double2* src,
                                            long long base, long long stride, int t, bool valid) {
  constexpr int R = Plan<N, RM>::R, P = Plan<N, RM>::P;
  const double2* p = src + base + (long long)t * stride;
  const long long step = (long long)P * stride;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    v[i] = valid ? p[0] : make_double2(0.0, 0.0);
    p += step;
  }
}

template <int N, int RM>
__device__ __forceinline__ void store_pencil(const double2 (&v)[(Plan<N, RM>::R)], double2* dst,
                                             long long base, long long stride, int t, bool valid) {
  constexpr int R = Plan<N, RM>::R, P = Plan<N, RM>::P;
  if (!valid) return;
  double2* p = dst + base + (long long)t * stride;
  const long long step = (long long)P * stride;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    p[0] = v[i];
    p += step;
  }
}

// ------------------------------------------------------------ pointwise V

// CRT_Base::Do_NL_Step (CRT_Base.h:397-441) / CRT_Base_IF::Do_NL_Step
// (CRT_Base_IF.h:381-419) on one point of NC components.  `extra[i]` is the
// potential V_i (NL, added first as in CRT_Base.h:416) or beta0*x0 - DeltaL_i
// (NL_IF, added last as in CRT_Base_IF.h:405).
template <int NC, bool EXTRA_FIRST>
__device__ __forceinline__ void nl_point(double2 (&a)[NC], const double (&extra)[NC], const VArgs& p) {
  double phi[NC];
  double n[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) n[j] = a[j].x * a[j].x + a[j].y * a[j].y;
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    double ph = EXTRA_FIRST ? extra[i] : 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) ph += p.gs[j + NC * i] * n[j];
    if (!EXTRA_FIRST) ph += extra[i];
    phi[i] = ph * (-p.dt);
  }
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    double s, c;
    sincos_fast(phi[i], &s, &c);
    const double re = a[i].x;
    a[i].x = re * c - a[i].y * s;
    a[i].y = a[i].y * c + re * s;
  }
}

// Rabi frequency Omega(x0) and eta(x0) of Do_Bragg_ad (bragg.cpp:224,240)
__device__ __forceinline__ void bragg_field(double x0, const VArgs& p, double& Om, double2& eta) {
  const double t1 = p.t1p ? *p.t1p : p.t1;
  const double F = p.Fp ? *p.Fp : p.F;
  Om = F * p.amp0 *
       cos(p.laser_k * x0 - (p.laser_domh + p.chirp * t1 + p.chirp_rate0 * t1) * t1 + p.phase0 / 2);
  const double arg = -0.5 * (p.laser_dk * x0);
  if (p.laser_dk == 0.0) {
    // arg is +-0: cos = 1 and sin = arg exactly (the sign of zero kept)
    eta = make_double2(1.0, arg);
  } else {
    double s, c;
    sincos(arg, &s, &c);
    eta = make_double2(c, s);
  }
}

// per-point analytic 2x2 exponential of Do_Bragg_ad (bragg.cpp:226-301),
// same operation order as the reference
__device__ __forceinline__ void bragg_point(double2& a, double2& b, double bx, double Om, double2 eta,
                                            const VArgs& p) {
  const double dt = p.dt;
  const double n1 = a.x * a.x + a.y * a.y;
  const double n2 = b.x * b.x + b.y * b.y;
  const double V11 = p.gs[0] * n1 + p.gs[1] * n2 + bx;
  const double V22 = p.gs[2] * n1 + p.gs[3] * n2 - p.deltaL[1] + bx;
  double s, c;
  if (Om == 0.0) {
    sincos_fast(-dt * V11, &s, &c);
    double re = a.x;
    a.x = re * c - a.y * s;
    a.y = re * s + a.y * c;
    sincos_fast(-dt * V22, &s, &c);
    re = b.x;
    b.x = re * c - b.y * s;
    b.y = re * s + b.y * c;
    return;
  }
  const double r = sqrt((V11 - V22) * (V11 - V22) + 4.0 * Om * Om);
  const double Ep = 0.5 * (V11 + V22 + r);
  const double Em = 0.5 * (V11 + V22 - r);
  sincos_fast(-dt * Ep, &s, &c);
  double nrm = 1.0 / fabs((V11 - Ep) * (V11 - Ep) + Om * Om);
  const double gpr = nrm * c, gpi = nrm * s;
  sincos_fast(-dt * Em, &s, &c);
  nrm = 1.0 / fabs((V11 - Em) * (V11 - Em) + Om * Om);
  const double gmr = nrm * c, gmi = nrm * s;
  const double o2 = Om * Om;
  const double O11r = o2 * (gpr + gmr), O11i = o2 * (gpi + gmi);
  const double dp = V11 - Ep, dm = V11 - Em;
  const double O22r = dp * dp * gpr + dm * dm * gmr;
  const double O22i = dp * dp * gpi + dm * dm * gmi;
  const double Xr = -Om * (dp * gpr + dm * gmr);
  const double Xi = -Om * (dp * gpi + dm * gmi);
  const double O12r = eta.x * Xr - eta.y * Xi, O12i = eta.y * Xr + eta.x * Xi;
  const double O21r = eta.x * Xr + eta.y * Xi, O21i = eta.x * Xi - eta.y * Xr;
  const double ar = a.x, ai = a.y, br = b.x, bi = b.y;
  a.x = (O11r * ar - O11i * ai) + (O12r * br - O12i * bi);
  a.y = (O11r * ai + O11i * ar) + (O12r * bi + O12i * br);
  b.x = (O21r * ar - O21i * ai) + (O22r * br - O22i * bi);
  b.y = (O21r * ai + O21i * ar) + (O22r * bi + O22i * br);
}

// the 2x2 exponential of one species of bragg_ds (bragg_ds.cpp:174-206):
// std::complex arithmetic of the reference written out; gamma divides by
// |d^2 + Omega^2| as the reference does
__device__ __forceinline__ void ds_pair(double2& a, double2& b, double V11, double V22, double Om, double theta,
                                        double dt) {
  double ei, er;
  sincos(theta, &ei, &er);
  const double tmp1 = sqrt((V11 - V22) * (V11 - V22) + 4.0 * Om * Om);
  const double Ep = 0.5 * (V11 + V22 + tmp1);
  const double Em = 0.5 * (V11 + V22 - tmp1);
  const double d0 = V11 - Ep, d1 = V11 - Em;
  const double den0 = fabs(d0 * d0 + Om * Om), den1 = fabs(d1 * d1 + Om * Om);
  double s0, c0, s1, c1;
  sincos(-dt * Ep, &s0, &c0);
  sincos(-dt * Em, &s1, &c1);
  const double g0r = c0 / den0, g0i = s0 / den0, g1r = c1 / den1, g1i = s1 / den1;
  const double tr = d0 * g0r + d1 * g1r, ti = d0 * g0i + d1 * g1i;
  const double o2 = Om * Om;
  const double M00r = o2 * (g0r + g1r), M00i = o2 * (g0i + g1i);
  const double ar = -Om * er, ai = -Om * ei;
  const double M01r = ar * tr - ai * ti, M01i = ar * ti + ai * tr;
  const double cr = -Om * er, ci = -Om * (-ei);
  const double M10r = cr * tr - ci * ti, M10i = cr * ti + ci * tr;
  const double dd0 = d0 * d0, dd1 = d1 * d1;
  const double M11r = dd0 * g0r + dd1 * g1r, M11i = dd0 * g0i + dd1 * g1i;
  const double p0r = a.x, p0i = a.y, p1r = b.x, p1i = b.y;
  a.x = (M00r * p0r - M00i * p0i) + (M01r * p1r - M01i * p1i);
  a.y = (M00r * p0i + M00i * p0r) + (M01r * p1i + M01i * p1r);
  b.x = (M10r * p0r - M10i * p0i) + (M11r * p1r - M11i * p1i);
  b.y = (M10r * p0i + M10i * p0r) + (M11r * p1i + M11i * p1r);
}

__device__ __forceinline__ void ds_diag(double2& a, double phi, double dt) {
  double im1, re1;
  sincos(-phi * dt, &im1, &re1);
  const double tmp1 = a.x;
  a.x = tmp1 * re1 - a.y * im1;
  a.y = tmp1 * im1 + a.y * re1;
}

// Bragg_single::Do_Bragg_ad of bragg_ds (bragg_ds.cpp:112-300): two species
// of two components, couplings through all four densities, one laser per
// species switched off by its gate
__device__ __forceinline__ void bragg_ds_point(double2 (&a)[4], double x0, const VArgs& p) {
  double n[4], phi[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) n[j] = a[j].x * a[j].x + a[j].y * a[j].y;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double ph = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) ph += p.gs[j + 4 * i] * n[j];
    ph += (i < 2 ? p.beta0 : p.beta2) * x0 - p.deltaL[i];
    phi[i] = ph;
  }
  const double t1 = p.t1p ? *p.t1p : p.t1;
  const double Om = p.mode1 * p.amp0 * cos(p.laser_k * x0 - (p.laser_domh + p.chirp_rate0 * t1) * t1 + p.phase0 / 2);
  const double Om2 = p.mode2 * p.amp2 * cos(p.laser_k2 * x0 - (p.laser_domh2 + p.chirp_rate2 * t1) * t1 + p.phase2 / 2);
  const double th1 = -0.5 * (p.laser_dk * x0 + p.phase0);
  const double th2 = -0.5 * (p.laser_dk2 * x0 + p.phase2);
  if (Om == 0.0 && Om2 == 0.0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) ds_diag(a[i], phi[i], p.dt);
  } else if (Om == 0.0) {
    ds_diag(a[0], phi[0], p.dt);
    ds_diag(a[1], phi[1], p.dt);
    ds_pair(a[2], a[3], phi[2], phi[3], Om2, th2, p.dt);
  } else if (Om2 == 0.0) {
    ds_diag(a[2], phi[2], p.dt);
    ds_diag(a[3], phi[3], p.dt);
    ds_pair(a[0], a[1], phi[0], phi[1], Om, th1, p.dt);
  } else {
    ds_pair(a[0], a[1], phi[0], phi[1], Om, th1, p.dt);
    ds_pair(a[2], a[3], phi[2], phi[3], Om2, th2, p.dt);
  }
}

// V step on the registers of NC component pencils at one element.
// x0: first-axis coordinate of the element; l: linear point index.
template <int VK, int NC>
__device__ __forceinline__ void v_point(double2 (&a)[NC], double x0, long long l, const VArgs& p,
                                        double Om, double2 eta) {
  if constexpr (VK == V_NL) {
    double extra[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) extra[c] = p.pot ? p.pot[c * p.pot_stride + l] : 0.0;
    nl_point<NC, true>(a, extra, p);
  } else if constexpr (VK == V_NL_IF) {
    double extra[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) extra[c] = p.beta0 * x0 - p.deltaL[c];
    nl_point<NC, false>(a, extra, p);
  } else if constexpr (VK == V_BRAGG) {
    static_assert(NC == 2, "Bragg coupling needs two components");
    bragg_point(a[0], a[1], p.beta0 * x0, Om, eta, p);
  } else if constexpr (VK == V_NL_IF2) {
    // CRT_Base_IF_2::Do_NL_Step (CRT_Base_IF_2.h:324-366)
    double extra[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) extra[c] = (c < NC / 2 ? p.beta0 : p.beta2) * x0 - p.deltaL[c];
    nl_point<NC, false>(a, extra, p);
  } else if constexpr (VK == V_BRAGG_DS) {
    static_assert(NC == 4, "the two-species Bragg step has four components");
    bragg_ds_point(a, x0, p);
  }
}

// ------------------------------------------------------------ three states

// psi <- exp(-i dt H) psi for a Hermitian N x N matrix H (H is overwritten).
// Cyclic complex Jacobi rotations diagonalise H = V diag(lambda) V^H; then
// psi <- V diag(exp(-i dt lambda)) V^H psi.  The reference forms the same
// exponential from GSL's Hermitian eigensolver (CRT_Base_IF.h:491-512);
// Jacobi keeps the eigenvalues accurate to eps*|H| also when they are
// (nearly) degenerate.  A rotation on (p, q) is J = D R with
// D = diag(1, e^{-i phi}) (phi = arg H_pq) and R the real Jacobi rotation
// [[c, s], [-s, c]] that zeroes the then real off-diagonal element.
template <int N>
__device__ __forceinline__ void herm_expm_apply(double2 (&H)[N][N], double dt, double2 (&psi)[N]) {
  double2 V[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) V[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  for (int sweep = 0; sweep < 16; ++sweep) {
    double off = 0.0, dg = 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p) {
      dg += H[p][p].x * H[p][p].x;
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += H[p][q].x * H[p][q].x + H[p][q].y * H[p][q].y;
    }
    if (off == 0.0 || off <= 1e-36 * dg) break;
#pragma unroll
    for (int p = 0; p < N - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const double2 apq = H[p][q];
        const double r = hypot(apq.x, apq.y);
        if (r == 0.0) continue;
        const double d = (H[q][q].x - H[p][p].x) / (2.0 * r);
        const double t = (d >= 0.0 ? 1.0 : -1.0) / (fabs(d) + sqrt(1.0 + d * d));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
        const double2 em = make_double2(apq.x / r, -apq.y / r);  // e^{-i phi}
        const double2 ep = make_double2(em.x, -em.y);            // e^{+i phi}
        // H J (columns p, q)
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double2 hp = H[k][p], hq = cmul(em, H[k][q]);
          H[k][p] = make_double2(c * hp.x - sn * hq.x, c * hp.y - sn * hq.y);
          H[k][q] = make_double2(sn * hp.x + c * hq.x, sn * hp.y + c * hq.y);
        }
        // J^H (H J) (rows p, q)
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double2 bp = H[p][k], bq = cmul(ep, H[q][k]);
          H[p][k] = make_double2(c * bp.x - sn * bq.x, c * bp.y - sn * bq.y);
          H[q][k] = make_double2(sn * bp.x + c * bq.x, sn * bp.y + c * bq.y);
        }
        H[p][q] = make_double2(0.0, 0.0);
        H[q][p] = make_double2(0.0, 0.0);
        H[p][p].y = 0.0;
        H[q][q].y = 0.0;
        // V J
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double2 vp = V[k][p], vq = cmul(em, V[k][q]);
          V[k][p] = make_double2(c * vp.x - sn * vq.x, c * vp.y - sn * vq.y);
          V[k][q] = make_double2(sn * vp.x + c * vq.x, sn * vp.y + c * vq.y);
        }
      }
    }
  }
  double2 y[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < N; ++k) {
      // conj(V[k][i]) * psi[k]
      acc.x += V[k][i].x * psi[k].x + V[k][i].y * psi[k].y;
      acc.y += V[k][i].x * psi[k].y - V[k][i].y * psi[k].x;
    }
    double sn, cs;
    sincos(-dt * H[i][i].x, &sn, &cs);
    y[i] = cmul(acc, make_double2(cs, sn));
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < N; ++i) acc = cadd(acc, cmul(V[k][i], y[i]));
    psi[k] = acc;
  }
}

// diagonal of the three-state steps: sum_j g_ij |psi_j|^2 + beta.x - DeltaL_i
// (CRT_Base_IF.h:455-466, 569-580)
template <int NC>
__device__ __forceinline__ void three_state_diag(double2 (&H)[NC][NC], const double2 (&a)[NC], double bx,
                                                 const VArgs& p) {
  double n[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) n[j] = a[j].x * a[j].x + a[j].y * a[j].y;
#pragma unroll
  for (int i = 0; i < NC; ++i) {
#pragma unroll
    for (int j = 0; j < NC; ++j) H[i][j] = make_double2(0.0, 0.0);
    double ph = 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) ph += p.gs[j + NC * i] * n[j];
    ph += bx - p.deltaL[i];
    H[i][i].x = ph;
  }
}

// CRT_Base_IF::Numerical_Bragg (CRT_Base_IF.h:426-527): state 0 coupled to
// state i+1 by the lasers of wave number +-laser_k and chirp +-chirp_rate,
// at t = Get_t() before the step
template <int NC>
__device__ __forceinline__ void num_bragg_point(double2 (&a)[NC], double x0, double bx, const VArgs& p) {
  static_assert(NC == 2 || NC == 3, "the numerical Bragg step has two or three states");
  double2 H[NC][NC];
  three_state_diag<NC>(H, a, bx, p);
  const double t = p.t0;
#pragma unroll
  for (int i = 0; i < NC - 1; ++i) {
    const double k = i == 0 ? p.laser_k : -p.laser_k;
    const double cr = i == 0 ? p.chirp_rate0 : -p.chirp_rate0;
    double sn, cs;
    sincos((-p.laser_domh + cr * t) * t + k * x0 - 0.5 * p.phase0, &sn, &cs);
    const double er = p.amp0 * cs / 2 + p.amp1 * cs / 2;
    const double ei = p.amp0 * sn / 2 - p.amp1 * sn / 2;
    H[i + 1][0] = make_double2(er, ei);
    H[0][i + 1] = make_double2(er, -ei);
  }
  herm_expm_apply<NC>(H, p.dt, a);
}

// CRT_Base_IF::Numerical_Raman (CRT_Base_IF.h:534-632): states 0 and 1
// coupled to state 2, which carries the detuning laser_domh
template <int NC>
__device__ __forceinline__ void raman_point(double2 (&a)[NC], double x0, double bx, const VArgs& p) {
  static_assert(NC >= 3, "the Raman step couples three states");
  double2 H[NC][NC];
  three_state_diag<NC>(H, a, bx, p);
  H[2][2].x = H[2][2].x + p.laser_domh;
  double sn, cs;
  sincos(p.laser_k * x0, &sn, &cs);
  H[0][2] = make_double2(p.amp0 / 2 * cs, p.amp0 / 2 * sn);
  H[2][0] = make_double2(p.amp0 / 2 * cs, -(p.amp0 / 2 * sn));
  H[1][2] = make_double2(p.amp1 / 2 * cs, -(p.amp1 / 2 * sn));
  H[2][1] = make_double2(p.amp1 / 2 * cs, p.amp1 / 2 * sn);
  herm_expm_apply<NC>(H, p.dt, a);
}

// Bragg_double::Do_Double_Bragg_ad (double_bragg.cpp:108-249): the
// reference's closed-form 3x3 exponential, its operations in its order
// (O23 and O32 take their phases as written there)
__device__ __forceinline__ void dbragg_point(double2 (&a)[3], double x0, double bx, const VArgs& p) {
  const double dt = p.dt, t1 = p.t1;
  const double F = p.F < 0.0 ? 0.0 : p.F;
  const double n1 = a[0].x * a[0].x + a[0].y * a[0].y;
  const double n2 = a[1].x * a[1].x + a[1].y * a[1].y;
  const double n3 = a[2].x * a[2].x + a[2].y * a[2].y;
  const double V11 = p.gs[0] * n1 + p.gs[1] * n2 + p.gs[2] * n3 + bx;
  const double V22 = p.gs[3] * n1 + p.gs[4] * n2 + p.gs[5] * n3 - p.deltaL[1] + bx;
  double dw = p.laser_domh + p.chirp_rate0 * t1;
  const double Op = F * p.amp0 * cos(p.laser_k * x0 - dw * t1 - p.phase0 / 2);
  dw = p.laser_domh - p.chirp_rate0 * t1;
  const double Om = F * p.amp0 * cos(-p.laser_k * x0 - dw * t1 - p.phase0 / 2);
  double im1, re1, tmp1, tmp2;
  if (Om == 0.0 && Op == 0.0) {
    sincos(-dt * V11, &im1, &re1);
    tmp1 = a[0].x;
    a[0].x = tmp1 * re1 - a[0].y * im1;
    a[0].y = tmp1 * im1 + a[0].y * re1;
    sincos(-dt * V22, &im1, &re1);
#pragma unroll
    for (int c = 1; c < 3; ++c) {
      tmp1 = a[c].x;
      a[c].x = tmp1 * re1 - a[c].y * im1;
      a[c].y = tmp1 * im1 + a[c].y * re1;
    }
    return;
  }
  sincos(-0.5 * p.laser_dk * x0, &im1, &re1);
  const double e0 = re1, e1 = im1;
  tmp1 = sqrt((V11 - V22) * (V11 - V22) + 4.0 * (Op * Op + Om * Om));
  const double E1 = 0.5 * (V11 + V22 + tmp1);
  const double E2 = 0.5 * (V11 + V22 - tmp1);
  sincos(-dt * E1, &im1, &re1);
  tmp1 = 1.0 / fabs((V22 - E1) * (V22 - E1) + Op * Op + Om * Om);
  const double g1r = tmp1 * re1, g1i = tmp1 * im1;
  sincos(-dt * E2, &im1, &re1);
  tmp1 = 1.0 / fabs((V22 - E2) * (V22 - E2) + Op * Op + Om * Om);
  const double g2r = tmp1 * re1, g2i = tmp1 * im1;
  sincos(-dt * V22, &im1, &re1);
  tmp1 = 1.0 / fabs(Op * Op + Om * Om);
  const double g3r = tmp1 * re1, g3i = tmp1 * im1;
  tmp1 = (V22 - E2) * (V22 - E2);
  tmp2 = (V22 - E1) * (V22 - E1);
  const double O11r = tmp1 * g2r + tmp2 * g1r, O11i = tmp1 * g2i + tmp2 * g1i;
  tmp1 = (V22 - E2) * g2r + (V22 - E1) * g1r;
  tmp2 = (V22 - E2) * g2i + (V22 - E1) * g1i;
  double O12r = -Op * tmp1, O12i = -Op * tmp2;
  double O13r = -Om * tmp1, O13i = -Om * tmp2;
  double O21r = O12r, O21i = O12i;
  double O31r = O13r, O31i = O13i;
  tmp1 = O12r;
  O12r = e0 * tmp1 - e1 * O12i;
  O12i = e1 * tmp1 + e0 * O12i;
  tmp1 = O21r;
  O21r = e0 * tmp1 + e1 * O21i;
  O21i = e0 * O21i - e1 * tmp1;
  tmp1 = O31r;
  O31r = e0 * tmp1 - e1 * O31i;
  O31i = e1 * tmp1 + e0 * O31i;
  tmp1 = O13r;
  O13r = e0 * tmp1 + e1 * O13i;
  O13i = e0 * O13i - e1 * tmp1;
  tmp1 = Op * Op;
  tmp2 = Om * Om;
  const double O22r = tmp1 * (g2r + g1r) + tmp2 * g3r, O22i = tmp1 * (g2i + g1i) + tmp2 * g3i;
  const double O33r = tmp2 * (g2r + g1r) + tmp1 * g3r, O33i = tmp2 * (g2i + g1i) + tmp1 * g3i;
  tmp1 = Om * Op;
  double O23r = tmp1 * (g1r + g2r - g3r), O23i = tmp1 * (g1i + g2i - g3i);
  double O32r = O23r, O32i = O23i;
  tmp1 = O23r;
  O23r = (e0 * e0 + e1 * e1) * tmp1 + 2 * e0 * e1 * O23i;
  O23i = (e0 * e0 + e1 * e1) * O23i - 2 * e0 * e1 * tmp1;
  tmp1 = O32r;
  O32r = (e0 * e0 - e1 * e1) * tmp1 - 2 * e0 * e1 * O32i;
  O32i = (e0 * e0 - e1 * e1) * O23i + 2 * e0 * e1 * tmp1;
  const double2 q1 = a[0], q2 = a[1], q3 = a[2];
  a[0].x = (O11r * q1.x - O11i * q1.y) + (O12r * q2.x - O12i * q2.y) + (O13r * q3.x - O13i * q3.y);
  a[0].y = (O11r * q1.y + O11i * q1.x) + (O12r * q2.y + O12i * q2.x) + (O13r * q3.y + O13i * q3.x);
  a[1].x = (O21r * q1.x - O21i * q1.y) + (O22r * q2.x - O22i * q2.y) + (O23r * q3.x - O23i * q3.y);
  a[1].y = (O21r * q1.y + O21i * q1.x) + (O22r * q2.y + O22i * q2.x) + (O23r * q3.y + O23i * q3.x);
  a[2].x = (O31r * q1.x - O31i * q1.y) + (O32r * q2.x - O32i * q2.y) + (O33r * q3.x - O33i * q3.y);
  a[2].y = (O31r * q1.y + O31i * q1.x) + (O32r * q2.y + O32i * q2.x) + (O33r * q3.y + O33i * q3.x);
}

// ------------------------------------------------------------ k_pass

template <int N, int RM, bool CONTIG, int DIR>
__global__ __launch_bounds__(256) void k_pass(const double2* src, double2* dst,
                                              long long comp_stride, AxisGeo g, const double2* __restrict__ tw,
                                              double scale) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N, RM>::R;
  PencilMap<N, RM, CONTIG> m(g);
  double2* lds = lds_all + m.p * g.ls;
  const double2* twl = stage_table<N>(lds_all + (size_t)g.T * g.ls, fft_tw<N, RM>(tw));
  const long long coff = (long long)blockIdx.y * comp_stride;
  double2 v[R];
  load_pencil<N, RM>(v, src + coff, m.base, m.stride, m.t, m.valid);
  if constexpr (tables_in_lds<N>()) __syncthreads();
  fft_pencil<N, RM, DIR, CONTIG && (Plan<N, RM>::P <= 64)>(v, m.t, lds, twl);
  if (scale != 1.0) {
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = cscale(v[i], scale);
  }
  store_pencil<N, RM>(v, dst + coff, m.base, m.stride, m.t, m.valid);
}

// ------------------------------------------------------------ k_kin
// Kinetic step along axis 0: forward DFT, multiply by
//   E0[j] * E1[i1] * E2[i2]  (per-axis phase tables, E0 includes the
//   ft(-1)*ft(+1) scale), inverse DFT.
// NC > 1 or VK != V_NONE: 1D grids only (the pencil is the whole axis 0 run
// and all components are held by the same thread).
struct KinTabs {
  const double2* e0;  // n0
  const double2* e1;  // n1 (NULL in 1D)
  const double2* e2;  // n2 (NULL in 1D/2D)
  int n2;
  // contiguous pencils only: a phase per pencil element and pencil row,
  // ez[(pid % ez_rows) * N + j] (the four-step 1D spectrum), else NULL
  const double2* ez;
  long long ez_rows;
  // second species (CRT_Base_2): components blockIdx.y >= split use these
  // tables instead (one component per blockIdx.y); f0 NULL: one species
  const double2* f0;
  const double2* f1;
  const double2* f2;
  int split;
};

template <int N, int RM, bool CONTIG, int VK, int NC, bool BATCH = false, int LB = 256>
__global__ __launch_bounds__(LB) void k_kin(double2* __restrict__ psi, long long comp_stride, AxisGeo g,
                                             const double2* __restrict__ tw, KinTabs kt, VArgs va, double x0_dx) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N, RM>::R, P = Plan<N, RM>::P;
  constexpr bool WAVE = CONTIG && (P <= 64);
  constexpr bool KT = kin_tw_in_lds<N, CONTIG>(), KE = kin_phase_in_lds<N, CONTIG>();
  constexpr int TWN = KT ? plan_tw_elems<N, RM>() : 0;
  PencilMap<N, RM, CONTIG> m(g);
  double2* lds = lds_all + m.p * g.ls;
  const double2* twl = stage_table<TWN, KT>(lds_all + (size_t)g.T * g.ls, fft_tw<N, RM>(tw));
  // the component of this block (one per blockIdx.y) may belong to the
  // second species, with its own tables (workgroup-uniform choice)
  const bool sp2 = NC == 1 && kt.f0 != nullptr && (int)blockIdx.y >= kt.split;
  const double2* t0 = sp2 ? kt.f0 : kt.e0;
  const double2* t1t = sp2 ? kt.f1 : kt.e1;
  const double2* t2t = sp2 ? kt.f2 : kt.e2;
  const double2* e0l = stage_table<N, KE>(lds_all + (size_t)g.T * g.ls + TWN, t0);
  // pencil constant of the phase (axes 1, 2)
  double2 cp = make_double2(1.0, 0.0);
  if constexpr (!CONTIG) {
    const long long col = m.pid % g.inner;
    if (t1t) {
      const int i1 = (int)(col / kt.n2), i2 = (int)(col % kt.n2);
      cp = t1t[i1];
      if (t2t) cp = cmul(cp, t2t[i2]);
    }
  }
  double2 v[NC][R];
  long long c0 = (NC == 1) ? (long long)blockIdx.y * comp_stride : 0;
  if constexpr (BATCH) {
    // 1D batch: pencil m.pid is run m.pid
    static_assert(CONTIG && VK != V_NONE, "batched kinetic steps carry the V step of their run");
    if (m.valid) bind_run(va, m.pid);
    c0 = m.pid * va.run_shift;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) load_pencil<N, RM>(v[c], psi + c0 + c * comp_stride, m.base, m.stride, m.t, m.valid);
  if constexpr (VK != V_NONE) {
    // 1D: the pencil is axis 0 itself, x0 = (j - n0/2) dx
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int j = m.t + P * i;
      const double x0 = (double)(j - N / 2) * x0_dx;
      double2 a[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) a[c] = v[c][i];
      double Om = 0.0;
      double2 eta = make_double2(1.0, 0.0);
      if constexpr (VK == V_BRAGG) bragg_field(x0, va, Om, eta);
      if (m.valid) v_point<VK, NC>(a, x0, m.base + j, va, Om, eta);
#pragma unroll
      for (int c = 0; c < NC; ++c) v[c][i] = a[c];
    }
  }
  if constexpr (KT || KE) __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    fft_pencil<N, RM, -1, WAVE>(v[c], m.t, lds, fresh_tw<N, KT>(twl));
    const double2* ezr = nullptr;
    if constexpr (CONTIG) {
      if (kt.ez) ezr = kt.ez + (m.pid % kt.ez_rows) * N;
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
      double2 e;
      if constexpr (CONTIG) {
        e = ezr ? ezr[m.t + P * i] : e0l[m.t + P * i];
      } else {
        e = cmul(e0l[m.t + P * i], cp);
      }
      v[c][i] = cmul(v[c][i], e);
    }
    fft_pencil<N, RM, +1, WAVE>(v[c], m.t, lds, fresh_tw<N, KT>(twl));
    store_pencil<N, RM>(v[c], psi + c0 + c * comp_stride, m.base, m.stride, m.t, m.valid);
  }
}

// ------------------------------------------------------------ k_block1d
// A whole block of 1D runs in one launch (N <= 4096: each run's pencil
// stays in the registers of its threads for all Nk steps): the half
// kinetic step, then Nk times the V step followed by the full kinetic step
// (the last one half), as the eager block launches them, with k_kin's
// operations (transform, V step, phase multiply).  The results agree with
// a launch per step to rounding: the plan may differ (RM), and multiply-adds
// are fused differently inside the step loop.  eh / ef: phase tables of the
// half and the full step.
// Step m reads its t1 from t1tab[m-1] and its amplitude from Ftab[m-1]
// (single run) or Ftab[(m-1)*runs + run] (BATCH: pencil r is run r).
template <int N, int RM, int VK, int NC, bool BATCH, int LB = 256>
__global__ __launch_bounds__(LB) void k_block1d(double2* __restrict__ psi, long long comp_stride, AxisGeo g,
                                                 const double2* __restrict__ tw, const double2* __restrict__ eh,
                                                 const double2* __restrict__ ef, VArgs va,
                                                 const double* __restrict__ t1tab, const double* __restrict__ Ftab,
                                                 double x0_dx, int Nk) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N, RM>::R, P = Plan<N, RM>::P;
  constexpr bool WAVE = P <= 64;
  PencilMap<N, RM, true> m(g);
  double2* lds = lds_all + m.p * g.ls;
  double2* tab = lds_all + (size_t)g.T * g.ls;
  // the twiddles live in LDS at every length (the block reads them
  // 2 (Nk + 1) times); the phase tables when N <= 1024
  {
    const double2* src = fft_tw<N, RM>(tw);
    for (int j = threadIdx.x; j < N; j += blockDim.x) tab[j] = src[j];
  }
  const double2* twl = tab;
  const double2* ehl = stage_table<N>(tab + N, eh);
  const double2* efl = stage_table<N>(tab + 2 * N, ef);
  long long c0 = 0;
  if constexpr (BATCH) {
    // bind_run takes the amplitude from the row of Fm: the first step's
    // (every step sets its own below)
    va.Fm = Ftab;
    if (m.valid) bind_run(va, m.pid);
    c0 = m.pid * va.run_shift;
  }
  double2 v[NC][R];
#pragma unroll
  for (int c = 0; c < NC; ++c) load_pencil<N, RM>(v[c], psi + c0 + c * comp_stride, m.base, 1, m.t, m.valid);
  __syncthreads();
  auto kinetic = [&](const double2* e) {
    if constexpr (NC == 1 && !tables_in_lds<N>()) {
      // phases from global memory: loaded before the forward transform,
      // whose barriers the loads could not otherwise move across
      double2 ev[R];
#pragma unroll
      for (int i = 0; i < R; ++i) ev[i] = e[m.t + P * i];
      fft_pencil<N, RM, -1, WAVE>(v[0], m.t, lds, twl);
#pragma unroll
      for (int i = 0; i < R; ++i) v[0][i] = cmul(v[0][i], ev[i]);
      fft_pencil<N, RM, +1, WAVE>(v[0], m.t, lds, twl);
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        fft_pencil<N, RM, -1, WAVE>(v[c], m.t, lds, twl);
#pragma unroll
        for (int i = 0; i < R; ++i) v[c][i] = cmul(v[c][i], e[m.t + P * i]);
        fft_pencil<N, RM, +1, WAVE>(v[c], m.t, lds, twl);
      }
    }
  };
  // the potential is the same for every step of the block: read once into
  // registers rather than after every step's last barrier
  double potv[VK == V_NL ? NC : 1][VK == V_NL ? R : 1];
  if constexpr (VK == V_NL) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int i = 0; i < R; ++i)
        potv[c][i] = (va.pot && m.valid) ? va.pot[c * va.pot_stride + m.base + m.t + P * i] : 0.0;
    }
  }
  kinetic(ehl);
  for (int s = 1; s <= Nk; ++s) {
    va.t1p = t1tab + (s - 1);
    if constexpr (BATCH) {
      va.Fp = nullptr;
      if (m.valid) va.F = Ftab[(size_t)(s - 1) * g.outer + m.pid];
    } else {
      va.Fp = Ftab + (s - 1);
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int j = m.t + P * i;
      const double x0 = (double)(j - N / 2) * x0_dx;
      double2 a[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) a[c] = v[c][i];
      double Om = 0.0;
      double2 eta = make_double2(1.0, 0.0);
      if constexpr (VK == V_BRAGG) bragg_field(x0, va, Om, eta);
      if constexpr (VK == V_NL) {
        // v_point's V_NL branch with the preloaded potential
        double extra[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) extra[c] = potv[c][i];
        if (m.valid) nl_point<NC, true>(a, extra, va);
      } else {
        if (m.valid) v_point<VK, NC>(a, x0, m.base + j, va, Om, eta);
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) v[c][i] = a[c];
    }
    kinetic(s < Nk ? efl : ehl);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) store_pencil<N, RM>(v[c], psi + c0 + c * comp_stride, m.base, 1, m.t, m.valid);
}

// ------------------------------------------------------------ k_vfuse
// inverse DFT along the last axis, V step, forward DFT along the last axis.
// The pencil (contiguous) has fixed coordinates on the other axes, so
// x0 and the Bragg field are per-pencil constants.
// EDGE: the first and last steps of a block also carry the last axis's
// share of the half kinetic step (forward DFT, times kz[j], inverse DFT;
// kz includes this axis's 1/N): before the V step when EDGE & 1 (the pencil
// arrives in x space), after it when EDGE & 2 (the pencil leaves in x
// space).  kz is staged in LDS behind the twiddles when N <= 1024.
template <int N, int RM, int VK, int NC, bool BATCH = false, int EDGE = 0>
__global__ __launch_bounds__(256) void k_vfuse(double2* __restrict__ psi, long long comp_stride, AxisGeo g,
                                               const double2* __restrict__ tw, VArgs va, long long pencils_per_x0,
                                               int n0, double dx0, const double2* __restrict__ kz) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N, RM>::R, P = Plan<N, RM>::P;
  constexpr bool WAVE = P <= 64;
  PencilMap<N, RM, true> m(g);
  double2* lds = lds_all + m.p * g.ls;
  const double2* twl = stage_table<N>(lds_all + (size_t)g.T * g.ls, fft_tw<N, RM>(tw));
  long long pid = m.pid;
  if constexpr (BATCH) {
    const long long r = m.pid / va.pencils_per_run;
    pid = m.pid - r * va.pencils_per_run;
    if (m.valid) bind_run(va, r);
    psi += r * va.run_shift;
  }
  const int i0 = (int)(pid / pencils_per_x0);
  const double x0 = (double)(i0 - n0 / 2) * dx0;
  double Om = 0.0;
  double2 eta = make_double2(1.0, 0.0);
  if constexpr (VK == V_BRAGG) bragg_field(x0, va, Om, eta);
  const double2* kzl = nullptr;
  if constexpr (EDGE != 0) kzl = stage_table<N>(lds_all + (size_t)g.T * g.ls + N, kz);
  double2 v[NC][R];
#pragma unroll
  for (int c = 0; c < NC; ++c) load_pencil<N, RM>(v[c], psi + c * comp_stride, m.base, 1, m.t, m.valid);
  // one component's potential is loaded with the pencil, before the
  // transforms' barriers, which the loads could not otherwise move across
  constexpr bool POT_PRE = VK == V_NL && NC == 1;
  double potv[POT_PRE ? R : 1];
  if constexpr (POT_PRE) {
#pragma unroll
    for (int i = 0; i < R; ++i) potv[i] = (va.pot && m.valid) ? va.pot[m.base + m.t + P * i] : 0.0;
  }
  if constexpr (tables_in_lds<N>()) __syncthreads();
  // the half kinetic phase along this axis, on the spectrum of component c
  auto kinetic_phase = [&](int c) {
#pragma unroll
    for (int i = 0; i < R; ++i) v[c][i] = cmul(v[c][i], kzl[m.t + P * i]);
  };
  if constexpr ((EDGE & 1) != 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      fft_pencil<N, RM, -1, WAVE>(v[c], m.t, lds, fresh_tw<N>(twl));
      kinetic_phase(c);
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) fft_pencil<N, RM, +1, WAVE>(v[c], m.t, lds, fresh_tw<N>(twl));
  if constexpr (NC == 2) {
    // the second component waits in this pencil's LDS region (each thread
    // at its own positions) while the V step runs point by point, so only
    // one component's values stay in registers through the V step's
    // temporaries
    lds_sync<WAVE>();  // the last exchange's reads of the region are done
#pragma unroll
    for (int i = 0; i < R; ++i) lds[lds_pos(m.t + P * i)] = v[1][i];
    asm volatile("" ::: "memory");
#pragma unroll
    for (int i = 0; i < R; ++i) {
      double2 a[2];
      a[0] = v[0][i];
      a[1] = lds[lds_pos(m.t + P * i)];
      if (m.valid) v_point<VK, 2>(a, x0, m.base + m.t + P * i, va, Om, eta);
      v[0][i] = a[0];
      lds[lds_pos(m.t + P * i)] = a[1];
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int i = 0; i < R; ++i) v[1][i] = lds[lds_pos(m.t + P * i)];
  } else {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      double2 a[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) a[c] = v[c][i];
      if constexpr (POT_PRE) {
        // v_point's V_NL branch with the preloaded potential
        double extra[1] = {potv[i]};
        if (m.valid) nl_point<NC, true>(a, extra, va);
      } else {
        if (m.valid) v_point<VK, NC>(a, x0, m.base + m.t + P * i, va, Om, eta);
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) v[c][i] = a[c];
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    fft_pencil<N, RM, -1, WAVE>(v[c], m.t, lds, fresh_tw<N>(twl));
    if constexpr ((EDGE & 2) != 0) {
      kinetic_phase(c);
      fft_pencil<N, RM, +1, WAVE>(v[c], m.t, lds, fresh_tw<N>(twl));
    }
    store_pencil<N, RM>(v[c], psi + c * comp_stride, m.base, 1, m.t, m.valid);
  }
}

// ------------------------------------------------------------ four-step 1D
// A 1D grid of N = N1 * N2 points is the matrix [N1][N2], j = j1 N2 + j2.
// Forward transform: DFT_N1 down every column j2, times W_N^(j2 k1)
// (W_N = exp(-2 pi i / N); j2 k1 < N), then DFT_N2 along every row k1, after
// which row k1, column k2 holds frequency k1 + N1 k2.  The inverse runs
// the same steps backwards with conjugate factors.  k_fs_col is the column
// half; the row half is a contiguous pass.
//   FS_ENTER: x -> Y (DFT_N1, twiddle)
//   FS_EXIT:  Y -> x (conjugate twiddle, inverse DFT_N1, times scale)
//   FS_STEP:  Y -> Y with the V step on the x-space values in between
// Components: NC == 1 and no batch: blockIdx.y is the component; otherwise
// blockIdx.y is the run (batch) or 0, and each thread holds NC components.
enum { FS_ENTER = 0, FS_EXIT = 1, FS_STEP = 2 };

template <int N1, int RM, int MODE, int VK, int NC, bool BATCH>
__global__ __launch_bounds__(256) void k_fs_col(double2* __restrict__ psi, long long comp_stride, AxisGeo g,
                                                const double2* __restrict__ tw1, const double2* __restrict__ wN,
                                                VArgs va, double dx, double scale) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N1, RM>::R, P = Plan<N1, RM>::P;
  PencilMap<N1, RM, false> m(g);
  double2* lds = lds_all + m.p * g.ls;
  const double2* twl = stage_table<N1>(lds_all + (size_t)g.T * g.ls, fft_tw<N1, RM>(tw1));
  const long long N2 = g.inner;
  const long long N = N2 * N1;
  const long long j2 = m.pid % N2;
  long long c0 = 0;
  if constexpr (BATCH) {
    static_assert(MODE == FS_STEP && NC == 2, "batches carry the V step of their run");
    bind_run(va, blockIdx.y);
    c0 = (long long)blockIdx.y * (va.run_shift + comp_stride);
  } else if constexpr (NC == 1) {
    c0 = (long long)blockIdx.y * comp_stride;
  }
  double2 v[NC][R];
#pragma unroll
  for (int c = 0; c < NC; ++c) load_pencil<N1, RM>(v[c], psi + c0 + c * comp_stride, m.base, m.stride, m.t, m.valid);
  if constexpr (tables_in_lds<N1>()) __syncthreads();
  if constexpr (MODE != FS_ENTER) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int i = 0; i < R; ++i) v[c][i] = cmulc(v[c][i], wN[j2 * (m.t + P * i)]);
      fft_pencil<N1, RM, +1, false>(v[c], m.t, lds, twl);
    }
  }
  if constexpr (MODE == FS_EXIT) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (scale != 1.0) {
#pragma unroll
        for (int i = 0; i < R; ++i) v[c][i] = cscale(v[c][i], scale);
      }
      store_pencil<N1, RM>(v[c], psi + c0 + c * comp_stride, m.base, m.stride, m.t, m.valid);
    }
    return;
  }
  if constexpr (MODE == FS_STEP) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const long long j = (long long)(m.t + P * i) * N2 + j2;
      const double x0 = (double)(j - N / 2) * dx;
      double2 a[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) a[c] = v[c][i];
      double Om = 0.0;
      double2 eta = make_double2(1.0, 0.0);
      if constexpr (VK == V_BRAGG) bragg_field(x0, va, Om, eta);
      if (m.valid) v_point<VK, NC>(a, x0, j, va, Om, eta);
#pragma unroll
      for (int c = 0; c < NC; ++c) v[c][i] = a[c];
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    fft_pencil<N1, RM, -1, false>(v[c], m.t, lds, twl);
#pragma unroll
    for (int i = 0; i < R; ++i) v[c][i] = cmul(v[c][i], wN[j2 * (m.t + P * i)]);
    store_pencil<N1, RM>(v[c], psi + c0 + c * comp_stride, m.base, m.stride, m.t, m.valid);
  }
}

// ------------------------------------------------------------ k_sobf
// Sobolev filters along axis 0 (CSOB_Base.h:251-301 fused): from the
// transform F of psi along the other axes, writes
//   out_sob = inverse0( op(k) * forward0(F) ) * scale,  op = 1/(1+alpha|k|^2)
//   out_lap = inverse0( -alpha|k|^2 * forward0(F) ) * scale  (if non-NULL)
// k is the native-order wave vector (cft_3d.cpp:145-166).
struct SobArgs {
  double alpha;
  double dk[3];
  int n[3];
  int dim;
  double scale;
};

__device__ __forceinline__ double kval(int i, int n, double dk) {
  const int s = n / 2;
  return dk * (double)(((i + s) % n) - s);
}

template <int N, int RM, bool CONTIG>
__global__ __launch_bounds__(256) void k_sobf(const double2* src, double2* out_sob, double2* out_lap, long long comp_stride, AxisGeo g,
                                              const double2* __restrict__ tw, SobArgs sa) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N, RM>::R, P = Plan<N, RM>::P;
  constexpr bool WAVE = CONTIG && (P <= 64);
  PencilMap<N, RM, CONTIG> m(g);
  double2* lds = lds_all + m.p * g.ls;
  const double2* twl = stage_table<N>(lds_all + (size_t)g.T * g.ls, fft_tw<N, RM>(tw));
  const long long coff = (long long)blockIdx.y * comp_stride;
  double k12 = 0.0;  // sum over axes 1, 2 of k^2 (k.scale(alpha).k with isotropic alpha)
  if constexpr (!CONTIG) {
    const long long col = m.pid % g.inner;
    if (sa.dim >= 2) {
      const int i1 = (sa.dim == 3) ? (int)(col / sa.n[2]) : (int)col;
      const double k1 = kval(i1, sa.n[1], sa.dk[1]);
      k12 += (k1 * sa.alpha) * k1;
    }
    if (sa.dim == 3) {
      const int i2 = (int)(col % sa.n[2]);
      const double k2 = kval(i2, sa.n[2], sa.dk[2]);
      k12 += (k2 * sa.alpha) * k2;
    }
  }
  double2 v[R];
  load_pencil<N, RM>(v, src + coff, m.base, m.stride, m.t, m.valid);
  if constexpr (tables_in_lds<N>()) __syncthreads();
  fft_pencil<N, RM, -1, WAVE>(v, m.t, lds, twl);
  double2 w[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const double k0 = kval(m.t + P * i, N, sa.dk[0]);
    const double phi = (k0 * sa.alpha) * k0 + k12;
    w[i] = cscale(v[i], -phi * sa.scale);
    v[i] = cscale(v[i], sa.scale / (1.0 + phi));
  }
  fft_pencil<N, RM, +1, WAVE>(v, m.t, lds, twl);
  store_pencil<N, RM>(v, out_sob + coff, m.base, m.stride, m.t, m.valid);
  if (out_lap) {
    fft_pencil<N, RM, +1, WAVE>(w, m.t, lds, twl);
    store_pencil<N, RM>(w, out_lap + coff, m.base, m.stride, m.t, m.valid);
  }
}

// ------------------------------------------------------------ k_sob_gradfuse
// The L2 gradient between the two Sobolev filters of one iteration
// (Compute_Sobolev_Gradient, CSOB_Base.h:304-337), fused into the passes
// along the last (contiguous) axis of a 2D/3D grid: the inverse DFT that
// finishes the Laplacian of component `comp` (lap holds it transformed
// along that axis only), then per point
//   grad_in = -lap + (V + sum_j g_j |psi_j|^2) psi_comp
// (the operations of k_sob_grad_in, in its order), then the forward DFT
// that starts the filter of grad_in.  out: the forward-transformed grad_in.
template <int N, int RM, int NW>
__global__ __launch_bounds__(256) void k_sob_gradfuse(const double2* __restrict__ lap, const double2* __restrict__ psi,
                                                      const double* __restrict__ pot, double2* __restrict__ out,
                                                      long long npts, AxisGeo g, const double2* __restrict__ tw,
                                                      int comp, double g0, double g1) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N, RM>::R, P = Plan<N, RM>::P;
  constexpr bool WAVE = P <= 64;
  PencilMap<N, RM, true> m(g);
  double2* lds = lds_all + m.p * g.ls;
  const double2* twl = stage_table<N>(lds_all + (size_t)g.T * g.ls, fft_tw<N, RM>(tw));
  const long long off = (long long)comp * npts;
  double2 v[R];
  load_pencil<N, RM>(v, lap + off, m.base, 1, m.t, m.valid);
  if constexpr (tables_in_lds<N>()) __syncthreads();
  fft_pencil<N, RM, +1, WAVE>(v, m.t, lds, twl);
  const double gj[2] = {g0, g1};
  if (m.valid) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const long long l = m.base + m.t + P * i;
      double nl = 0.0;
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        const double2 a = psi[j * npts + l];
        nl += gj[j] * (a.x * a.x + a.y * a.y);
      }
      const double vv = pot[off + l] + nl;
      const double2 p = psi[off + l];
      v[i] = make_double2(-v[i].x + vv * p.x, -v[i].y + vv * p.y);
    }
  }
  fft_pencil<N, RM, -1, WAVE>(v, m.t, lds, twl);
  store_pencil<N, RM>(v, out + off, m.base, 1, m.t, m.valid);
}

// ------------------------------------------------------------ pointwise

__device__ __forceinline__ double x0_of(long long l, const Grid& gr) {
  const long long inner = gr.npts / gr.n[0];
  const int i0 = (int)(l / inner);
  return (double)(i0 - gr.n[0] / 2) * gr.dx[0];
}

// beta.x over the grid's axes, summed as CPoint::operator* does
// (CPoint.h:145-151)
__device__ __forceinline__ double beta_dot_x(long long l, const Grid& gr, const double (&beta)[3]) {
  long long rest = l;
  long long idx[3] = {0, 0, 0};
  for (int a = gr.dim - 1; a >= 0; --a) {
    idx[a] = rest % gr.n[a];
    rest /= gr.n[a];
  }
  double bx = 0.0;
  for (int a = 0; a < gr.dim; ++a) bx += beta[a] * ((double)(idx[a] - gr.n[a] / 2) * gr.dx[a]);
  return bx;
}

template <int VK, int NC>
__global__ __launch_bounds__(256) void k_point(double2* __restrict__ psi, Grid gr, VArgs va) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < gr.npts; l += stride) {
    double2 a[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) a[c] = psi[c * gr.npts + l];
    const double x0 = x0_of(l, gr);
    double Om = 0.0;
    double2 eta = make_double2(1.0, 0.0);
    if constexpr (VK == V_BRAGG) bragg_field(x0, va, Om, eta);
    if constexpr (VK == V_DBRAGG) {
      dbragg_point(a, x0, beta_dot_x(l, gr, va.beta), va);
    } else if constexpr (VK == V_NUM_BRAGG) {
      num_bragg_point<NC>(a, x0, beta_dot_x(l, gr, va.beta), va);
    } else if constexpr (VK == V_RAMAN) {
      raman_point<NC>(a, x0, beta_dot_x(l, gr, va.beta), va);
    } else {
      v_point<VK, NC>(a, x0, l, va, Om, eta);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) psi[c * gr.npts + l] = a[c];
  }
}

// ------------------------------------------------------------ reductions

constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 1024;
constexpr int RED_MAX_NV = 8;

template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NV], double* out) {
  __shared__ double sh[RED_THREADS / 64][NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double x = acc[v];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    acc[v] = x;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) sh[wave][v] = acc[v];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

template <class Fn>
__global__ __launch_bounds__(RED_THREADS) void k_reduce(Fn fn, long long n, double* __restrict__ partials) {
  constexpr int NV = Fn::NV;
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += stride) fn(l, acc);
  block_reduce_store<NV>(acc, partials + (long long)blockIdx.x * NV);
}

// segmented form: blockIdx.y selects segment y, whose data start
// y * seg_stride elements after the functor's; partials[y][blockIdx.x]
template <class Fn>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_seg(Fn fn, long long n, long long seg_stride,
                                                            double* __restrict__ partials) {
  constexpr int NV = Fn::NV;
  fn.shift((long long)blockIdx.y * seg_stride);
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += stride) fn(l, acc);
  block_reduce_store<NV>(acc, partials + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * NV);
}

// one workgroup per segment: out[blockIdx.x][v]
template <int NV>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_seg_final(const double* __restrict__ partials, int nblocks,
                                                                  double* __restrict__ out) {
  partials += (long long)blockIdx.x * nblocks * NV;
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += partials[(long long)b * NV + v];
  }
  block_reduce_store<NV>(acc, out + (long long)blockIdx.x * NV);
}

template <int NV>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double* __restrict__ partials, int nblocks,
                                                              double* __restrict__ out) {
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += partials[(long long)b * NV + v];
  }
  block_reduce_store<NV>(acc, out);
}

// sum |psi|^2
struct RedNorm {
  static constexpr int NV = 1;
  const double2* psi;
  __device__ void shift(long long off) { psi += off; }
  __device__ void operator()(long long l, double (&acc)[1]) const {
    const double2 a = psi[l];
    acc[0] += a.x * a.x + a.y * a.y;
  }
};

// sum x_a |psi|^2 (CRT_Base::Expval_Position)
struct RedX {
  static constexpr int NV = 3;
  const double2* psi;
  Grid gr;
  __device__ void operator()(long long l, double (&acc)[3]) const {
    const double2 a = psi[l];
    const double d = a.x * a.x + a.y * a.y;
    const long long n12 = (long long)gr.n[1] * gr.n[2];
    const int i0 = (int)(l / n12);
    const int i1 = (int)((l - i0 * n12) / gr.n[2]);
    const int i2 = (int)(l - i0 * n12 - (long long)i1 * gr.n[2]);
    acc[0] += ((double)(i0 - gr.n[0] / 2) * gr.dx[0]) * d;
    acc[1] += ((double)(i1 - gr.n[1] / 2) * gr.dx[1]) * d;
    acc[2] += ((double)(i2 - gr.n[2] / 2) * gr.dx[2]) * d;
  }
};

__device__ __forceinline__ void k_of(long long l, const Grid& gr, double (&k)[3]) {
  const long long n12 = (long long)gr.n[1] * gr.n[2];
  const int i0 = (int)(l / n12);
  const int i1 = (int)((l - i0 * n12) / gr.n[2]);
  const int i2 = (int)(l - i0 * n12 - (long long)i1 * gr.n[2]);
  k[0] = kval(i0, gr.n[0], gr.dk[0]);
  k[1] = gr.dim > 1 ? kval(i1, gr.n[1], gr.dk[1]) : 0.0;
  k[2] = gr.dim > 2 ? kval(i2, gr.n[2], gr.dk[2]) : 0.0;
}

// sum k_a |psi^|^2 (CRT_Base::Expval_Momentum) on a transformed buffer
struct RedK {
  static constexpr int NV = 3;
  const double2* psik;
  Grid gr;
  __device__ void operator()(long long l, double (&acc)[3]) const {
    const double2 a = psik[l];
    const double d = a.x * a.x + a.y * a.y;
    double k[3];
    k_of(l, gr, k);
    acc[0] += k[0] * d;
    acc[1] += k[1] * d;
    acc[2] += k[2] * d;
  }
};

// ana_expval (ana_expval.cpp:50-107): sum |psi|^2, sum x_a |psi|^2 and
// sum atan2(Im, Re) |psi|^2
struct RedMom1 {
  static constexpr int NV = 5;
  const double2* psi;
  Grid gr;
  __device__ void operator()(long long l, double (&acc)[5]) const {
    const double2 a = psi[l];
    const double d = a.x * a.x + a.y * a.y;
    const long long n12 = (long long)gr.n[1] * gr.n[2];
    const int i0 = (int)(l / n12);
    const int i1 = (int)((l - i0 * n12) / gr.n[2]);
    const int i2 = (int)(l - i0 * n12 - (long long)i1 * gr.n[2]);
    acc[0] += d;
    acc[1] += ((double)(i0 - gr.n[0] / 2) * gr.dx[0]) * d;
    acc[2] += ((double)(i1 - gr.n[1] / 2) * gr.dx[1]) * d;
    acc[3] += ((double)(i2 - gr.n[2] / 2) * gr.dx[2]) * d;
    acc[4] += atan2(a.y, a.x) * d;
  }
};

// sum (x_a - mu_a)^2 |psi|^2, or of k_a for a transformed state
// (ana_expval.cpp:109-185, Expval_2nd_moment / _Momentum)
template <bool KSPACE>
struct RedMom2 {
  static constexpr int NV = 3;
  const double2* psi;
  Grid gr;
  double mu[3];
  __device__ void operator()(long long l, double (&acc)[3]) const {
    const double2 a = psi[l];
    const double d = a.x * a.x + a.y * a.y;
    const long long n12 = (long long)gr.n[1] * gr.n[2];
    const int i0 = (int)(l / n12);
    const int i1 = (int)((l - i0 * n12) / gr.n[2]);
    const int i2 = (int)(l - i0 * n12 - (long long)i1 * gr.n[2]);
    const int id[3] = {i0, i1, i2};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const double c = KSPACE ? kval(id[ax], gr.n[ax], gr.dk[ax]) : (double)(id[ax] - gr.n[ax] / 2) * gr.dx[ax];
      acc[ax] += (c - mu[ax]) * (c - mu[ax]) * d;
    }
  }
};

// Rabi momentum bins (CRT_Base_IF::compute_rabi_integrals, CRT_Base_IF.h:
// 240-256), including the slot counter that only advances on a miss
// one point of compute_rabi_integrals (the reference's accumulation): |psi^|^2
// = d at wave vector k goes to acc[j], j counting the balls k0_q (radius
// thr) passed over before the first one that holds k
__device__ __forceinline__ void rabi_accumulate(const double (&k)[3], double d, int dim, int nk0,
                                                const double (&k0)[RED_MAX_NV][3], double thr,
                                                double (&acc)[RED_MAX_NV]) {
  int j = 0;
  for (int q = 0; q < nk0; ++q) {
    double dd = 0.0;
    for (int ax = 0; ax < dim; ++ax) {
      const double e = k0[q][ax] - k[ax];
      dd += e * e;
    }
    if (sqrt(dd) < thr) {
#pragma unroll
      for (int s = 0; s < RED_MAX_NV; ++s)
        if (s == j) acc[s] += d;
      continue;
    }
    ++j;
  }
}

struct RedRabi {
  static constexpr int NV = RED_MAX_NV;
  const double2* psik;
  Grid gr;
  int nk0;
  double k0[RED_MAX_NV][3];
  double thr;
  __device__ void shift(long long off) { psik += off; }
  __device__ void operator()(long long l, double (&acc)[RED_MAX_NV]) const {
    const double2 a = psik[l];
    const double d = a.x * a.x + a.y * a.y;
    double k[3];
    k_of(l, gr, k);
    rabi_accumulate(k, d, gr.dim, nk0, k0, thr, acc);
  }
};

// RedRabi over the index box that holds every ball |k - k0_q| < thr: box
// point l has centred indices m_a = lo[a] + (its offset along axis a), so
// k_a = dk_a m_a is kval's value at grid index m_a mod n_a; points outside
// the box lie outside every ball and add nothing.  The membership test and
// the accumulation are RedRabi's.
struct RedRabiBox {
  static constexpr int NV = RED_MAX_NV;
  const double2* psik;
  Grid gr;
  int lo[3];
  int len[3];  // box extent per axis (1 beyond the grid's dimension)
  int nk0;
  double k0[RED_MAX_NV][3];
  double thr;
  __device__ void shift(long long off) { psik += off; }
  __device__ void operator()(long long l, double (&acc)[RED_MAX_NV]) const {
    const unsigned u = (unsigned)l;  // boxes hold fewer than 2^31 points
    const unsigned q12 = (unsigned)len[1] * (unsigned)len[2];
    const unsigned b0 = u / q12, r = u - b0 * q12;
    const unsigned b1 = r / (unsigned)len[2], b2 = r - b1 * (unsigned)len[2];
    int m[3] = {lo[0] + (int)b0, lo[1] + (int)b1, lo[2] + (int)b2};
    int i[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) i[a] = m[a] < 0 ? m[a] + gr.n[a] : m[a];
    const double2 v = psik[((long long)i[0] * gr.n[1] + i[1]) * gr.n[2] + i[2]];
    const double d = v.x * v.x + v.y * v.y;
    double k[3];
    k[0] = gr.dk[0] * (double)m[0];
    k[1] = gr.dim > 1 ? gr.dk[1] * (double)m[1] : 0.0;
    k[2] = gr.dim > 2 ? gr.dk[2] * (double)m[2] : 0.0;
    rabi_accumulate(k, d, gr.dim, nk0, k0, thr, acc);
  }
};

// ---------------------------------------------- Rabi integrals of a 3D box
// compute_rabi_integrals on a 3D grid only needs the spectrum inside the
// index box of the balls (RedRabiBox's box, len[a] of the n[a] indices
// along axis a).  Three passes transform only what reaches the box:
//   k_rabi_z: forward DFT along z of every (x, y) pencil of the grid,
//             keeping the len2 outputs inside the box: zb [n0][n1][len2]
//   k_rabi_y: forward DFT along y of zb's (x, kz) pencils, keeping the len1
//             outputs inside the box: yb [n0][len1][len2]
//   k_rabi_x: forward DFT along x of yb's pencils times the transform's
//             scale, then RedRabi's sums over the whole kx range: one
//             partial per workgroup and segment
// blockIdx.y is the segment (component or run), `seg` elements apart in
// the input and compact in the buffers.  Each pencil's transform is the
// full pass's (same plan, twiddles and operation order).
struct RabiBox {
  int n[3];
  int lo[3];
  int len[3];
  double dk[3];
  int nk0;
  double k0[RED_MAX_NV][3];
  double thr;
  double scale;
};

// grid index of box index b along axis a (centred index lo + b, wrapped)
__device__ __forceinline__ int rabi_box_index(const RabiBox& rb, int a, int b) {
  const int m = rb.lo[a] + b;
  return m < 0 ? m + rb.n[a] : m;
}

// box index of grid index i along axis a, or -1 outside the box
__device__ __forceinline__ int rabi_box_of(const RabiBox& rb, int a, int i) {
  const int m = i < rb.n[a] / 2 ? i : i - rb.n[a];
  const int b = m - rb.lo[a];
  return (b >= 0 && b < rb.len[a]) ? b : -1;
}

template <int N2, int RM>
__global__ __launch_bounds__(256) void k_rabi_z(const double2* __restrict__ src, long long seg, double2* zb,
                                                AxisGeo g, const double2* __restrict__ tw, RabiBox rb) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N2, RM>::R, P = Plan<N2, RM>::P;
  PencilMap<N2, RM, true> m(g);
  double2* lds = lds_all + m.p * g.ls;
  const double2* twl = stage_table<N2>(lds_all + (size_t)g.T * g.ls, fft_tw<N2, RM>(tw));
  double2 v[R];
  load_pencil<N2, RM>(v, src + (long long)blockIdx.y * seg, m.base, 1, m.t, m.valid);
  if constexpr (tables_in_lds<N2>()) __syncthreads();
  fft_pencil<N2, RM, -1, (P <= 64)>(v, m.t, lds, twl);
  if (!m.valid) return;
  double2* out = zb + ((long long)blockIdx.y * g.outer + m.pid) * rb.len[2];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int b = rabi_box_of(rb, 2, m.t + P * i);
    if (b >= 0) out[b] = v[i];
  }
}

template <int N1, int RM>
__global__ __launch_bounds__(256) void k_rabi_y(const double2* __restrict__ zb, double2* yb, AxisGeo g,
                                                const double2* __restrict__ tw, RabiBox rb) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N1, RM>::R, P = Plan<N1, RM>::P;
  PencilMap<N1, RM, false> m(g);  // columns: kz box index; slabs: x
  double2* lds = lds_all + m.p * g.ls;
  const double2* twl = stage_table<N1>(lds_all + (size_t)g.T * g.ls, fft_tw<N1, RM>(tw));
  const long long len12 = (long long)rb.len[1] * rb.len[2];
  double2 v[R];
  load_pencil<N1, RM>(v, zb + (long long)blockIdx.y * g.outer * N1 * g.inner, m.base, m.stride, m.t, m.valid);
  if constexpr (tables_in_lds<N1>()) __syncthreads();
  fft_pencil<N1, RM, -1, false>(v, m.t, lds, twl);
  if (!m.valid) return;
  const long long o = m.pid / g.inner, b2 = m.pid - o * g.inner;
  double2* out = yb + (long long)blockIdx.y * g.outer * len12 + o * len12 + b2;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int b = rabi_box_of(rb, 1, m.t + P * i);
    if (b >= 0) out[(long long)b * rb.len[2]] = v[i];
  }
}

template <int N0, int RM>
__global__ __launch_bounds__(256) void k_rabi_x(const double2* __restrict__ yb, AxisGeo g,
                                                const double2* __restrict__ tw, RabiBox rb,
                                                double* __restrict__ partials) {
  extern __shared__ double2 lds_all[];
  constexpr int R = Plan<N0, RM>::R, P = Plan<N0, RM>::P;
  PencilMap<N0, RM, false> m(g);  // columns: (ky, kz) box index pairs
  double2* lds = lds_all + m.p * g.ls;
  const double2* twl = stage_table<N0>(lds_all + (size_t)g.T * g.ls, fft_tw<N0, RM>(tw));
  double2 v[R];
  load_pencil<N0, RM>(v, yb + (long long)blockIdx.y * N0 * g.inner, m.base, m.stride, m.t, m.valid);
  if constexpr (tables_in_lds<N0>()) __syncthreads();
  fft_pencil<N0, RM, -1, false>(v, m.t, lds, twl);
  double acc[RED_MAX_NV];
#pragma unroll
  for (int q = 0; q < RED_MAX_NV; ++q) acc[q] = 0.0;
  if (m.valid) {
    const int b1 = (int)(m.pid / rb.len[2]), b2 = (int)(m.pid - (long long)b1 * rb.len[2]);
    double k[3];
    k[1] = rb.dk[1] * (double)(rb.lo[1] + b1);
    k[2] = rb.dk[2] * (double)(rb.lo[2] + b2);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int j = m.t + P * i;
      k[0] = rb.dk[0] * (double)(j < N0 / 2 ? j : j - N0);
      const double2 a = cscale(v[i], rb.scale);
      rabi_accumulate(k, a.x * a.x + a.y * a.y, 3, rb.nk0, rb.k0, rb.thr, acc);
    }
  }
  block_reduce_store<RED_MAX_NV>(acc, partials + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * RED_MAX_NV);
}

// kinetic energy density sum_a alpha_a k_a^2 |psi^|^2 on a transformed buffer
struct RedEkin {
  static constexpr int NV = 1;
  const double2* psik;
  Grid gr;
  double alpha[3];
  __device__ void operator()(long long l, double (&acc)[1]) const {
    const double2 a = psik[l];
    double k[3];
    k_of(l, gr, k);
    double ks = 0.0;
    for (int ax = 0; ax < gr.dim; ++ax) ks += alpha[ax] * k[ax] * k[ax];
    acc[0] += ks * (a.x * a.x + a.y * a.y);
  }
};

// potential + interaction energy sum_i (V_i + 1/2 sum_j g_ij n_j) n_i
struct RedEpot {
  static constexpr int NV = 1;
  const double2* psi;
  const double* pot;
  long long npts;
  int nc;
  double gs[16];
  __device__ void operator()(long long l, double (&acc)[1]) const {
    double n[4];
    for (int i = 0; i < nc; ++i) {
      const double2 a = psi[i * npts + l];
      n[i] = a.x * a.x + a.y * a.y;
    }
    double e = 0.0;
    for (int i = 0; i < nc; ++i) {
      const double v = pot ? pot[i * npts + l] : 0.0;
      double nl = 0.0;
      for (int j = 0; j < nc; ++j) nl += gs[j + nc * i] * n[j];
      e += (v + 0.5 * nl) * n[i];
    }
    acc[0] += e;
  }
};

// <psi, g> and <psi, psi_sob> (real parts, CSOB_Base.h:350-356)
struct RedProj {
  static constexpr int NV = 2;
  const double2* psi;
  const double2* grad;
  const double2* sob;
  __device__ void operator()(long long l, double (&acc)[2]) const {
    const double2 p = psi[l], g = grad[l], s = sob[l];
    acc[0] += p.x * g.x + p.y * g.y;
    acc[1] += p.x * s.x + p.y * s.y;
  }
};

}  // namespace atus
