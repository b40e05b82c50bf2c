// kernels_misc.hpp -- non-template kernels, launched only by atus_abi.hip
// (kept out of kernels.hpp so each has one definition).
#pragma once
#include "kernels.hpp"

namespace atus {

// psi *= exp(i p.x)  (CRT_Base::Setup_Momentum, CRT_Base.h:449-473)
__global__ __launch_bounds__(256) void k_momentum(double2* __restrict__ psi, Grid gr, double p0, double p1,
                                                  double p2) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long n12 = (long long)gr.n[1] * gr.n[2];
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < gr.npts; l += stride) {
    const int i0 = (int)(l / n12);
    const int i1 = (int)((l - i0 * n12) / gr.n[2]);
    const int i2 = (int)(l - i0 * n12 - (long long)i1 * gr.n[2]);
    double px = p0 * ((double)(i0 - gr.n[0] / 2) * gr.dx[0]);
    if (gr.dim > 1) px += p1 * ((double)(i1 - gr.n[1] / 2) * gr.dx[1]);
    if (gr.dim > 2) px += p2 * ((double)(i2 - gr.n[2] / 2) * gr.dx[2]);
    double s, c;
    sincos(px, &s, &c);
    const double2 a = psi[l];
    psi[l] = make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// multiply a component by a complex table per point (kinetic step of the
// unfused path)
__global__ __launch_bounds__(256) void k_mul_phase(double2* __restrict__ psi, long long npts, KinTabs kt, int n1,
                                                   int n2, int dim) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long n12 = (long long)n1 * n2;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < npts; l += stride) {
    const int i0 = (int)(l / n12);
    const int i1 = (int)((l - i0 * n12) / n2);
    const int i2 = (int)(l - i0 * n12 - (long long)i1 * n2);
    double2 e = kt.e0[i0];
    if (dim > 1) e = cmul(e, kt.e1[i1]);
    if (dim > 2) e = cmul(e, kt.e2[i2]);
    psi[blockIdx.y * npts + l] = cmul(psi[blockIdx.y * npts + l], e);
  }
}

// inflate_domain (inflate_domain.cpp:41-223) on the device, one thread per
// destination point of one component (blockIdx.y): FIX == false pads the
// source grid, fak times smaller per axis, with zeros around the centre
// (offset (n_dst - n_src)/2); FIX == true keeps the number of points and
// places every fak-th source point around the centre (offset
// n*(fak-1)/2/fak); everything else is zero
template <bool FIX>
__global__ __launch_bounds__(256) void k_inflate(const double2* __restrict__ src, double2* __restrict__ dst, Grid gs,
                                                 Grid gd, int fak, long long src_stride, long long dst_stride) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const double2* s = src + (long long)blockIdx.y * src_stride;
  double2* d = dst + (long long)blockIdx.y * dst_stride;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < gd.npts; l += stride) {
    long long rest = l;
    int id[3] = {0, 0, 0};
    for (int a = 2; a >= 0; --a) {
      id[a] = (int)(rest % gd.n[a]);
      rest /= gd.n[a];
    }
    bool inside = true;
    long long ls = 0;
    for (int a = 0; a < 3; ++a) {
      int is;
      if (!FIX) {
        is = id[a] - (gd.n[a] - gs.n[a]) / 2;
      } else {
        // destination off + i/fak <- source i for i = 0, fak, 2 fak, ...
        const int off = gs.n[a] * (fak - 1) / 2 / fak;
        const int q = id[a] - off;
        is = q * fak;
        if (gs.n[a] == 1) is = q;  // unused axis
      }
      if (is < 0 || is >= gs.n[a]) inside = false;
      ls = ls * gs.n[a] + (inside ? is : 0);
    }
    d[l] = inside ? s[ls] : make_double2(0.0, 0.0);
  }
}

// streaming copy (bandwidth probe)
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += stride) dst[l] = src[l];
}

// ------------------------------------------------------------ Sobolev pointwise

// L2 gradient -Lap psi + (V + sum_j g_ij |psi_j|^2) psi  (CSOB_Base.h:322-327)
__global__ __launch_bounds__(256) void k_sob_grad_in(const double2* __restrict__ psi, const double2* __restrict__ lap,
                                                     const double* __restrict__ pot, double2* __restrict__ out,
                                                     long long npts, int nw, int comp, double g0, double g1) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const double g[2] = {g0, g1};
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < npts; l += stride) {
    double nl = 0.0;
    for (int j = 0; j < nw; ++j) {
      const double2 a = psi[j * npts + l];
      nl += g[j] * (a.x * a.x + a.y * a.y);
    }
    const double v = pot[comp * npts + l] + nl;
    const double2 p = psi[comp * npts + l];
    const double2 lp = lap[comp * npts + l];
    out[comp * npts + l] = make_double2(-lp.x + v * p.x, -lp.y + v * p.y);
  }
}

// g -= fak * psi_sob (CSOB_Base.h:360-365), then sum |g|^2 and |psi|^2
__global__ __launch_bounds__(RED_THREADS) void k_sob_project_res(double2* __restrict__ grad,
                                                                 const double2* __restrict__ sob,
                                                                 const double2* __restrict__ psi, long long n,
                                                                 double fak, double* __restrict__ partials) {
  double acc[2] = {0.0, 0.0};
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += stride) {
    double2 g = grad[l];
    const double2 s = sob[l];
    g.x -= fak * s.x;
    g.y -= fak * s.y;
    grad[l] = g;
    const double2 p = psi[l];
    acc[0] += g.x * g.x + g.y * g.y;
    acc[1] += p.x * p.x + p.y * p.y;
  }
  block_reduce_store<2>(acc, partials + (long long)blockIdx.x * 2);
}

// psi <- f * psi - step * g  (Renormalize_All_Psi then the update,
// CSOB_Base.h:481-492; f = 1 when no renormalisation is due)
__global__ __launch_bounds__(256) void k_sob_update(double2* __restrict__ psi, const double2* __restrict__ grad,
                                                    long long n, double f, double step) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += stride) {
    double2 p = psi[l];
    const double2 g = grad[l];
    if (f != 1.0) {
      p.x *= f;
      p.y *= f;
    }
    p.x -= step * g.x;
    p.y -= step * g.y;
    psi[l] = p;
  }
}

// ---- the Sobolev loop on the device (CSOB_Base::run, CSOB_Base.h:427-505),
// iterations enqueued back to back without a host round trip.  An iteration
// that ends with res_tot <= eps sets `done`; the iterations enqueued after it
// leave psi and the state as they are, so the loop stops where the
// reference's do-while stops.
struct SobState {
  double proj[2][2];  // <psi,grad>, <psi,psi_sob> of each wave function
  double f[2];        // scale of psi before this iteration's update (1 or 1/sqrt(N))
  double N[2];        // particle numbers of the current psi
  double res_tot;     // residual of the last iteration performed
  long long iters;    // iterations performed
  int done;           // an iteration ended with res_tot <= eps
  int renorm;         // this iteration renormalises (f != 1)
};
// per performed iteration: N[0], N[1] (checked before its update), whether
// it renormalised, res_tot
constexpr int SOB_LOG_VALS = 4;
constexpr int SOB_MAX_CHUNK = 1024;

// psi <- f psi - step (grad - fak psi_sob) unless the loop is done (the
// projection, CSOB_Base.h:360-365; Renormalize_All_Psi and the update,
// :470-492), with fak = proj[0] / proj[1]; partials of |grad - fak psi_sob|^2
// and of |psi|^2 after the update.  measure_only: the |psi|^2 partials of
// the current psi alone (grad and psi_sob are not read).
__global__ __launch_bounds__(RED_THREADS) void k_sob_update_dev(double2* __restrict__ psi,
                                                                const double2* __restrict__ grad,
                                                                const double2* __restrict__ sob, long long n,
                                                                double step, const SobState* __restrict__ st, int comp,
                                                                int measure_only, double* __restrict__ partials) {
  double acc[2] = {0.0, 0.0};
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (measure_only) {
    for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += stride) {
      const double2 p = psi[l];
      acc[1] += p.x * p.x + p.y * p.y;
    }
  } else {
    const double fak = st->proj[comp][0] / st->proj[comp][1];
    const double f = st->f[comp];
    const bool frozen = st->done != 0;
    for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += stride) {
      double2 g = grad[l];
      const double2 s = sob[l];
      g.x -= fak * s.x;
      g.y -= fak * s.y;
      double2 p = psi[l];
      if (!frozen) {
        if (f != 1.0) {
          p.x *= f;
          p.y *= f;
        }
        p.x -= step * g.x;
        p.y -= step * g.y;
        psi[l] = p;
      }
      acc[0] += g.x * g.x + g.y * g.y;
      acc[1] += p.x * p.x + p.y * p.y;
    }
  }
  block_reduce_store<2>(acc, partials + (long long)blockIdx.x * 2);
}

// one workgroup: sums the update partials of the nw wave functions
// (partials[i][nb][2], fixed order) and advances the state.  init: N of the
// current psi only (start of a run of iterations).  Otherwise, unless the
// loop is done: res_i = ar sum |g|^2 (Compute_res, CSOB_Base.h:370-397),
// res_tot = sqrt(sum res_i^2), the iteration's log entry, iters + 1, done
// if res_tot <= eps; then N of the updated psi and the renormalisation of
// the next iteration (|N_i - 1| > 1e-5 for any i, CSOB_Base.h:470-479).
__global__ __launch_bounds__(RED_THREADS) void k_sob_finalize(SobState* __restrict__ st,
                                                              const double* __restrict__ partials, int nb, int nw,
                                                              double ar, double eps, int init, double* __restrict__ log,
                                                              int slot) {
  __shared__ double tot[2][2];
  for (int i = 0; i < nw; ++i) {
    double acc[2] = {0.0, 0.0};
    const double* p = partials + (long long)i * nb * 2;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      acc[0] += p[(long long)b * 2];
      acc[1] += p[(long long)b * 2 + 1];
    }
    block_reduce_store<2>(acc, tot[i]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (init) {
    st->done = 0;
    st->iters = 0;
    st->res_tot = 0.0;
  } else {
    if (st->done) return;
    double rsq = 0.0;
    for (int i = 0; i < nw; ++i) {
      const double r = tot[i][0] * ar;
      rsq += r * r;
    }
    const double rt = sqrt(rsq);
    if (log) {
      double* e = log + (long long)slot * SOB_LOG_VALS;
      e[0] = st->N[0];
      e[1] = nw > 1 ? st->N[1] : 0.0;
      e[2] = st->renorm;
      e[3] = rt;
    }
    st->res_tot = rt;
    st->iters += 1;
    if (rt <= eps) st->done = 1;
  }
  bool renorm = false;
  for (int i = 0; i < nw; ++i) {
    st->N[i] = ar * tot[i][1];
    if (fabs(st->N[i] - 1) > 1e-5) renorm = true;
  }
  for (int i = 0; i < nw; ++i) st->f[i] = renorm ? 1.0 / sqrt(st->N[i]) : 1.0;
  st->renorm = renorm ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_scale(double2* __restrict__ psi, long long n, double f) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += stride) psi[l] = cscale(psi[l], f);
}


// dst = scale * transpose(src) for `ncomp` matrices of rows x cols, each
// comp_stride elements apart in both buffers; 32 x 32 tiles through LDS,
// one tile per workgroup (blockIdx.x = (comp * row_tiles + i) * col_tiles + j)
__global__ __launch_bounds__(256) void k_transpose(const double2* __restrict__ src, double2* __restrict__ dst,
                                                   long long rows, long long cols, long long comp_stride,
                                                   double scale) {
  __shared__ double2 tile[32][33];
  const long long ct = (cols + 31) / 32, rt = (rows + 31) / 32;
  const long long b = blockIdx.x;
  const long long comp = b / (rt * ct), rem = b % (rt * ct);
  const long long i0 = (rem / ct) * 32, j0 = (rem % ct) * 32;
  src += comp * comp_stride;
  dst += comp * comp_stride;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const long long i = i0 + k, j = j0 + tx;
    if (i < rows && j < cols) tile[k][tx] = src[i * cols + j];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const long long j = j0 + k, i = i0 + tx;
    if (i < rows && j < cols) {
      const double2 a = tile[tx][k];
      dst[j * rows + i] = make_double2(a.x * scale, a.y * scale);
    }
  }
}

}  // namespace atus
