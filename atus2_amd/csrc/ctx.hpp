// ctx.hpp -- the context object behind the C-ABI handle and the error
// plumbing shared by the library's host sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "../../include/atus_hip.h"
#include "launch.hpp"

namespace atus_host {
// message of the last failure on this thread (atus_last_error)
extern thread_local std::string g_err;
inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace atus_host

#define HIP_TRY(expr)                                                                                        \
  do {                                                                                                       \
    hipError_t e_ = (expr);                                                                                  \
    if (e_ != hipSuccess) return atus_host::fail(ATUS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define ATUS_TRY(expr)            \
  do {                            \
    int r_ = (expr);              \
    if (r_ != ATUS_OK) return r_; \
  } while (0)

using atus::Grid;
using atus::RunParams;

struct atus_ctx {
  int dim = 0;
  int n[3] = {1, 1, 1};
  double dx[3] = {0, 0, 0};
  double dk[3] = {0, 0, 0};
  int ncomp = 0;            // all components held (nbatch * nc_run)
  long long npts = 0;
  int nbatch = 1;           // independent runs (atus_batch_create)
  int nc_run = 0;           // components per run
  // 1D grids longer than 4096: four-step transform on [fs_n1][fs_n2]
  bool fs = false;
  int fs_n1 = 0, fs_n2 = 0;
  double2* kinz = nullptr;  // [2 fracs][N] kinetic phases in [k1][k2] order
  double2* tbuf = nullptr;  // ncomp * npts, the transposes of full transforms
  int device = 0;
  hipStream_t stream = nullptr;
  // extra streams of the slab chains (created on first use, default
  // priority), the fork event on `stream` and one join event per extra
  // stream; slab i runs on stream i % slab_streams (0: `stream`)
  static constexpr int MAX_SLAB_STREAMS = 4;
  hipStream_t xstream[MAX_SLAB_STREAMS - 1] = {nullptr, nullptr, nullptr};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[MAX_SLAB_STREAMS - 1] = {nullptr, nullptr, nullptr};
  int slab_streams = 2;
  size_t max_lds = 65536;

  double2* psi = nullptr;   // ncomp * npts
  double2* snap[4] = {nullptr, nullptr, nullptr, nullptr};  // device snapshots of psi
  double* pot = nullptr;    // ncomp * npts, or null
  double2* scratch = nullptr;
  std::map<int, double2*> tw;

  // kinetic phase tables: half step, full step, half step of one axis alone
  static constexpr int KIN_SLOTS_N = 3;
  double2* kin = nullptr;   // [KIN_SLOTS_N][3 axes][nmax]
  int kin_nmax = 0;
  double kin_dt = NAN;
  bool kin_valid = false;
  double alpha[3] = {0, 0, 0};
  // second species (CRT_Base_2): components [ncomp/2, ncomp) use alpha2
  bool alpha2_set = false;
  double alpha2[3] = {0, 0, 0};
  double2* kin2 = nullptr;  // same layout as kin
  double gs[16] = {0};

  double* red_part = nullptr;
  double* red_out = nullptr;
  double* red_host = nullptr;

  // batches: per-run parameters, the amplitude table of one block and
  // the segmented-reduction buffers
  RunParams* runs = nullptr;
  bool runs_set = false;
  double* ftab = nullptr;
  size_t ftab_cap = 0;
  double* sred_part = nullptr;
  double* sred_out = nullptr;
  size_t sred_cap = 0;
  int sred_nseg = 0;
  // compact spectra of the Rabi integrals of 3D grids (k_rabi_z / _y)
  double2* rbuf = nullptr;
  size_t rbuf_cap = 0;

  bool sob_ready = false;
  double sob_alpha[2] = {0, 0};
  double sob_step = 0;
  double2 *sob = nullptr, *lap = nullptr, *grad = nullptr;
  // the device-side loop: its state (atus::SobState), the update partials
  // [2][RED_MAX_BLOCKS][2] and the per-iteration log [SOB_MAX_CHUNK][4]
  void* sob_st = nullptr;
  double* sob_part = nullptr;
  double* sob_log = nullptr;

  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;

  // values per thread of the FFT kernels, per kernel family (ATUS_KFAM_*);
  // defaults: what the autotuner picks at 256^3 (8 for the strided passes
  // and for the register-heavy V pass, which spills at 16)
  int rm[ATUS_NKFAM] = {16, 8, 16, 8, 16};
  // threads per workgroup at most, per kernel family (0: as many as the
  // kernel's launch bounds allow); sets the pencils per workgroup
  int wg[ATUS_NKFAM] = {0, 0, 0, 0, 0};
  // 3D blocks: the y and z passes between two kinetic steps run slab by
  // slab of this many x-planes, alternate slabs on two streams (0: over the
  // whole grid each; -1: chosen from the size of the device's last-level
  // cache)
  int slab = -1;
  int slab_planes() const {
    if (dim != 3 || fs) return 0;
    int s = slab;
    if (s < 0) {
      // two slabs in flight hold about 160 MiB of the 256 MiB Infinity
      // Cache of MI355X (measured best at 256^3: 32 to 40 planes)
      const double plane = 16.0 * ncomp * n[1] * n[2];
      s = (int)std::max(1.0, std::floor(83886080.0 / plane));
    }
    return s < n[0] ? s : 0;
  }
  // 1D blocks of up to 4096 points in one launch, each run's pencil
  // resident in registers for all steps (k_block1d); 0: a launch per step
  int resident = 1;
  // per-step t1 and amplitudes of a resident block: pinned host staging,
  // the device table, and the event after its upload (the host side is
  // rewritten only once that upload has completed)
  double* blk_host = nullptr;
  double* blk_dev = nullptr;
  size_t blk_cap = 0;
  hipEvent_t blk_ev = nullptr;

  // captured graphs (atus_graph_create) not yet destroyed: they hold raw
  // pointers into this context, so it refuses to free or replace the
  // buffers they read while any is alive
  int live_graphs = 0;

  // per-launch event pairs while profiling
  bool prof = false;
  std::vector<hipEvent_t> prof_ev;
  std::vector<int> prof_fam;
  std::vector<double> prof_frac;  // share of the whole grid each launch covers

  Grid grid() const {
    Grid g{};
    g.dim = dim;
    for (int a = 0; a < 3; ++a) {
      g.n[a] = n[a];
      g.dx[a] = dx[a];
      g.dk[a] = dk[a];
    }
    g.npts = npts;
    return g;
  }
  double ar() const {
    double a = dx[0];
    if (dim > 1) a *= dx[1];
    if (dim > 2) a *= dx[2];
    return a;
  }
  double ar_k() const {
    double a = dk[0];
    if (dim > 1) a *= dk[1];
    if (dim > 2) a *= dk[2];
    return a;
  }
  long long inner_of(int a) const {
    long long v = 1;
    for (int b = a + 1; b < dim; ++b) v *= n[b];
    return v;
  }
  long long outer_of(int a) const {
    long long v = 1;
    for (int b = 0; b < a; ++b) v *= n[b];
    return v;
  }
};
