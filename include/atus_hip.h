/*
 * atus_hip.h -- C-ABI of libatus_hip.so, the MI355X (gfx950) implementation
 * of the ATUS-2 split-step Fourier propagator and Sobolev-gradient
 * minimiser.
 *
 * Every entry point returns ATUS_OK (0) or a negative status; the message of
 * the last failure on the calling thread is returned by atus_last_error().
 * No C++ exception crosses this boundary.  Pointers named `host` are host
 * memory in the fftw_complex layout of the reference: interleaved (re, im)
 * doubles, row-major [nx][ny][nz], z fastest (cft_base.h:209).
 *
 * A context owns one grid, its components (device resident, complex128),
 * optional potentials and one HIP stream.  Calls on one context must come
 * from one thread at a time (as the reference's solver objects, which are
 * not thread-safe).  Grid axes must be powers of two between 2 and 4096
 * (1 for unused axes); a 1D grid may have up to 2^20 points (four-step
 * transform).  The Sobolev minimiser covers axes up to 4096.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef ATUS_HIP_H
#define ATUS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ATUS_OK 0
#define ATUS_ERR_ARG (-1)          /* invalid argument (sizes, comp index, ...) */
#define ATUS_ERR_HIP (-2)          /* HIP runtime failure (message has the HIP error) */
#define ATUS_ERR_UNSUPPORTED (-3)  /* grid shape outside what the kernels cover */
#define ATUS_ERR_NOMEM (-4)        /* device allocation failed */
#define ATUS_ERR_STATE (-5)        /* call not valid in the context's state */
#define ATUS_ERR_NUMERIC (-6)      /* a reduction returned NaN or Inf */

typedef struct atus_ctx atus_ctx;

/* message of the last failing call on this thread ("" if none) */
const char *atus_last_error(void);
/* library version string and the device it runs on */
const char *atus_version(void);
int atus_device_info(char *name, int len, int *cu_count, double *hbm_gbps_peak);

/* ---- context: replaces cft_base<dim> construction + CRT_shared::Read_header
 * (cft_base.h:47-88, CRT_shared.h:131-172).  n[a] = nDimX/Y/Z, dx/dk as in
 * generic_header (my_structs.h:60-65).  ncomp = no_int_states (1..4). */
int atus_ctx_create(atus_ctx **out, int dim, const long long n[3], const double dx[3], const double dk[3],
                    int ncomp);
int atus_ctx_destroy(atus_ctx *ctx);
int atus_synchronize(atus_ctx *ctx);

/* ---- state: replaces the host buffers behind cft_base::Getp2In()
 * (cft_base.h:175) and CRT_Base::LoadFiles / Save_Phi (CRT_Base.h:207-240,
 * 607-618): copies of component `comp` between host and device. */
int atus_upload(atus_ctx *ctx, int comp, const double *host);
int atus_download(atus_ctx *ctx, int comp, double *host);
/* device snapshots of all components in slots 0..3 (allocated on first
 * save): the chirp scan of CRT_Base_IF::run_sequence saves the start state
 * to "%.3f_%d.bin" files and reloads them before every chirp
 * (CRT_Base_IF.h:757-766, 893-903, 913-921); a snapshot restores the same
 * bits without the file and host round trip */
int atus_snapshot_save(atus_ctx *ctx, int slot);
int atus_snapshot_restore(atus_ctx *ctx, int slot);
/* device copy of src components [src_comp, src_comp + n) into dst
 * components [dst_comp, dst_comp + n); same device and grid size (e.g. one
 * run's state into the runs of a batch).  Within one context the two
 * ranges must not overlap (ATUS_ERR_ARG); identical ranges are a no-op. */
int atus_copy_components(atus_ctx *dst, int dst_comp, atus_ctx *src, int src_comp, int n);
/* CRT_Base::Init_Potential + Setup_Potential (CRT_Base.h:277-298):
 * npts doubles for component comp; host == NULL removes all potentials
 * (m_potenial_initialized == false) */
int atus_set_potential(atus_ctx *ctx, int comp, const double *host);
/* m_gs (CRT_Base.h:114, read at :162-169): ncomp*ncomp, row-major */
int atus_set_g(atus_ctx *ctx, const double *gs);
/* m_alpha (CRT_Base.h:89, read at :172-173) */
int atus_set_alpha(atus_ctx *ctx, const double alpha[3]);

/* ---- spectral field: cft_Nd::ft(isign) (cft_3d.cpp:58-74) with the
 * scale() normalisation, in place on one component; isign not +-1 is a
 * no-op as in the reference */
int atus_ft(atus_ctx *ctx, int comp, int isign);

/* ---- step functions of the registry m_map_stepfcts (CRT_Base.h:152-155,
 * CRT_Base_IF.h:124-126, bragg.cpp:82-83) */
/* Do_FT_Step_half (frac 0.5) / Do_FT_Step_full (frac 1) (CRT_Base.h:337-392) */
int atus_kinetic(atus_ctx *ctx, double frac, double dt);
/* CRT_Base::Do_NL_Step (CRT_Base.h:397-441), "freeprop" of rt_solver */
int atus_nl_step(atus_ctx *ctx, double dt);
/* CRT_Base_IF::Do_NL_Step (CRT_Base_IF.h:381-419), "freeprop" of bragg:
 * deltaL has ncomp entries */
int atus_nl_if_step(atus_ctx *ctx, double dt, double beta0, const double *deltaL);

/* laser parameters of CRT_Base_IF (CRT_Base_IF.h:63-72, UpdateParams :145-168) */
typedef struct {
  double amp0;        /* Amp[0] */
  double laser_k;     /* laser_k[0] */
  double laser_domh;  /* laser_domh[0] */
  double laser_dk;    /* laser_dk[0] */
  double chirp;       /* chirp */
  double chirp_rate0; /* chirp_rate[0] */
  double phase0;      /* phase[0] */
  double beta0;       /* beta[0] */
  double deltaL[4];   /* DeltaL */
  double amp1;        /* Amp[1]: numerical Bragg and Raman steps */
  double beta[3];     /* Beta: the three-state steps add beta.x over all axes */
} atus_laser;

/* Bragg_single::Do_Bragg_ad (bragg.cpp:186-304), "bragg_ad": t = Get_t()
 * before the step, F = Amplitude_at_time() (CRT_Base_IF.h:171-198) */
int atus_bragg_ad(atus_ctx *ctx, double dt, double t, double F, const atus_laser *laser);
/* CRT_Base::Setup_Momentum (CRT_Base.h:449-473): psi_comp *= exp(i p.x) */
int atus_set_momentum(atus_ctx *ctx, int comp, const double p[3]);

/* ---- one outer iteration of the sequence hot loop (CRT_Base.h:776-786,
 * CRT_Base_IF.h:779-788): half kinetic, (V, full kinetic) x (Nk-1), V, half
 * kinetic, with the reference's time bookkeeping on *t.  The kinetic steps
 * and the V steps are fused across FFT passes on the device. */
#define ATUS_STEP_FREEPROP 0    /* CRT_Base::Do_NL_Step */
#define ATUS_STEP_FREEPROP_IF 1 /* CRT_Base_IF::Do_NL_Step */
#define ATUS_STEP_BRAGG_AD 2    /* Bragg_single::Do_Bragg_ad */
/* three internal states (ncomp 3; the numerical Bragg step also takes 2
 * and the Raman step 4); these run the V step as its own pass */
#define ATUS_STEP_DOUBLE_BRAGG 3 /* Bragg_double::Do_Double_Bragg_ad (double_bragg.cpp:108-249) */
#define ATUS_STEP_NUM_BRAGG 4    /* CRT_Base_IF::Numerical_Bragg (CRT_Base_IF.h:426-527), "bragg" */
#define ATUS_STEP_RAMAN 5        /* CRT_Base_IF::Numerical_Raman (CRT_Base_IF.h:534-632), "raman" */

/* AMP_T as a callback: returns Amplitude_at_time() for the time t1 that
 * CRT_Base_IF::Amplitude_at_time would evaluate (Get_t()+0.5*Get_dt()) */
typedef double (*atus_amp_fn)(double t1, void *user);

typedef struct {
  int kind;            /* ATUS_STEP_* */
  double dt;
  atus_laser laser;    /* FREEPROP_IF uses beta0 and deltaL; BRAGG_AD all */
  atus_amp_fn amp;     /* NULL: Amplitude_at_time() == 1 */
  void *amp_user;
} atus_step_params;

int atus_run_block(atus_ctx *ctx, const atus_step_params *sp, int Nk, double *t);
/* one V step of a single run (no kinetic step), any kind of atus_run_block:
 * t = Get_t() before the step, F = Amplitude_at_time() (read by BRAGG_AD
 * and DOUBLE_BRAGG, which clamps it at 0 as double_bragg.cpp:117 does) */
int atus_v_step(atus_ctx *ctx, int kind, double dt, double t, double F, const atus_laser *laser);

/* ---- two species (CRT_Base_2, CRT_Base_IF_2, the bragg_ds driver):
 * components [ncomp/2, ncomp) are species 2 */
/* m_alpha2 (CRT_Base_2.h:195-221, Alpha_2): kinetic coefficients of species 2 */
int atus_set_alpha2(atus_ctx *ctx, const double alpha2[3]);
typedef struct {
  atus_laser laser2;  /* species 2 (CRT_Base_IF_2.h:104-123): amp0 = Amp_2[0], laser_k =
                         laser_k_2, laser_dk = laser_dk_2, laser_domh = laser_domh_2,
                         chirp_rate0 = chirp_rate2[0], phase0 = phase2[0], beta0 =
                         Beta_2[0]; chirp and deltaL are not read */
  double seq_time;    /* sequence_item::time, start of the pulses */
  double duration[2]; /* sequence_item::duration[0..1]: pulse lengths of species 1 and 2 */
} atus_ds_params;
/* one block of CRT_Base_IF_2::run_sequence (CRT_Base_IF_2.h:439-600) with
 * kind ATUS_STEP_FREEPROP_IF (CRT_Base_IF_2::Do_NL_Step, :324-366) or
 * ATUS_STEP_BRAGG_AD (Do_Bragg_ad of bragg_ds.cpp:112-300, four
 * components); sp->laser is species 1 (bragg_ds reads neither its chirp
 * nor AMP_T); the kinetic steps use alpha and alpha2 per species */
int atus_run_block_ds(atus_ctx *ctx, const atus_step_params *sp, const atus_ds_params *ds, int Nk, double *t);
/* one V step of a two-species run, no kinetic step: kind
 * ATUS_STEP_FREEPROP_IF (CRT_Base_IF_2::Do_NL_Step, CRT_Base_IF_2.h:324-366)
 * or ATUS_STEP_BRAGG_AD (Bragg_single::Do_Bragg_ad, bragg_ds.cpp:112-300)
 * at t1 = t + dt/2, pulse gates from ds; laser1 as sp->laser above */
int atus_v_step_ds(atus_ctx *ctx, int kind, double dt, double t, const atus_laser *laser1, const atus_ds_params *ds);

/* ---- reductions (CRT_Base.h:481-599, CRT_Base_IF.h:208-281) */
/* Get_Particle_Number: ar * sum |psi_comp|^2 */
int atus_particle_number(atus_ctx *ctx, int comp, double *out);
/* Expval_Position: ar * sum x |psi|^2 (dim entries) */
int atus_expval_x(atus_ctx *ctx, int comp, double out[3]);
/* Expval_Momentum: ar_k * sum k |psi^|^2 (psi is left untouched) */
int atus_expval_k(atus_ctx *ctx, int comp, double out[3]);
/* compute_rabi_integrals: k0 holds n points of 3 doubles, n <= 8;
 * out[q] = ar_k * sum of |psi^|^2 in slot q with the reference's slot rule */
int atus_rabi_integrals(atus_ctx *ctx, int comp, const double *k0, int n, double threshold, double *out);
/* the moments ana_expval prints (ana_expval.cpp:50-185, 281-394) for
 * component comp: N = ar sum |psi|^2; <x_a> = ar sum x_a |psi|^2 / N and
 * <(x_a - <x_a>)^2> likewise; <k_a>, <(k_a - <k_a>)^2> from the transformed
 * state with ar_k, also divided by N; phase = sum atan2(Im, Re) |psi|^2 /
 * sum |psi|^2 (ana_expval prints it for 1D states).  psi is unchanged. */
typedef struct {
  double N;
  double x[3], x2[3];
  double k[3], k2[3];
  double phase;
} atus_moments;
int atus_get_moments(atus_ctx *ctx, int comp, atus_moments *out);
/* inflate_domain (inflate_domain.cpp:41-223) between two contexts of the
 * same dimension and component count: fix == 0 copies src into the centre
 * of dst, whose axes are fak times longer (zeros around it); fix == 1 keeps
 * the number of points and places every fak-th point of src around the
 * centre of dst.  The header changes (domain, dx, dk) are the caller's. */
int atus_inflate(atus_ctx *dst, atus_ctx *src, int fak, int fix);
/* energy functional (no reference counterpart):
 * ar_k sum_i sum_k alpha.k^2 |psi_i^|^2 + ar sum_i sum_x (V_i + 1/2 sum_j g_ij |psi_j|^2) |psi_i|^2 */
int atus_energy(atus_ctx *ctx, double *out);

/* ---- Sobolev gradient minimiser (CSOB_Base.h) on the context's
 * components (no_wf = ncomp <= 2).  gs must already carry the
 * sqrt(N_i N_j) factor (CSOB_Base.h:177); alpha_iso[i] = Alpha_i[0]. */
/* CSOB_Base ctor + Init (CSOB_Base.h:97-193, 229-248) */
int atus_sob_setup(atus_ctx *ctx, const double *alpha_iso, double stepsize);
/* Renormalize_All_Psi (CSOB_Base.h:523-543) */
int atus_sob_renormalize(atus_ctx *ctx);
/* one body of the do-while loop of CSOB_Base::run (CSOB_Base.h:431-503);
 * optional outputs: res_tot, renormalised (0/1) and N (ncomp particle
 * numbers checked before the update, CSOB_Base.h:470-474) */
int atus_sob_iter(atus_ctx *ctx, double *res_tot, int *renormalised, double *N);
/* up to `count` (1..1024) bodies of that loop enqueued back to back, with
 * no host round trip between them: the projection factor, the residual,
 * the particle numbers and the renormalisation decision stay on the
 * device.  The first body whose res_tot <= eps ends the loop as the
 * reference's do-while does: the bodies enqueued after it leave psi
 * unchanged.  Outputs: *performed (bodies that updated psi), *res_tot of
 * the last of them, and if log != NULL four values per performed body:
 * N[0], N[1] (checked before its update), renormalised (0/1), res_tot --
 * what CSOB_Base::run prints per iteration.  eps < 0: all count bodies. */
int atus_sob_iterate(atus_ctx *ctx, int count, double eps, int *performed, double *res_tot, double *log);
/* Compute_mu (CSOB_Base.h:400-424, with Get_Ar() = m_ar): ncomp values */
int atus_sob_mu(atus_ctx *ctx, double *mu);
/* CSOB_Base::run (CSOB_Base.h:427-505) after Renormalize_All_Psi: loop
 * bodies until res_tot <= eps; maxit > 0 caps the count (the reference
 * loops without a cap).  Outputs: iterations done, last res_tot.  Runs
 * atus_sob_iterate in chunks sized from the residual's rate of decrease. */
int atus_sob_run(atus_ctx *ctx, double eps, long long maxit, long long *iters, double *res_tot);

/* ---- timing helper: average device time (ms) of the last atus_run_block
 * call's kernels, measured with HIP events on the context's stream; 0 if
 * timing is off.  enable: 1 turns per-block event timing on. */
int atus_set_timing(atus_ctx *ctx, int enable);
int atus_last_block_ms(atus_ctx *ctx, double *ms);

/* ---- per-kernel device times: runs one block exactly as atus_run_block,
 * with a HIP event pair on the context's stream around every launch.
 * ms[f] is the summed duration and count[f] the number of launches of
 * kernel family f; grid[f] (optional, may be NULL) sums the share of the
 * whole grid each launch covered (1 for a whole-grid pass, S/nx for a pass
 * over a slab of S x-planes). */
#define ATUS_KFAM_PASS_CONTIG 0  /* DFT pass along the contiguous axis */
#define ATUS_KFAM_PASS_STRIDED 1 /* DFT pass along a strided axis */
#define ATUS_KFAM_KIN 2          /* fused kinetic step along axis 0 */
#define ATUS_KFAM_VFUSE 3        /* fused inverse / V step / forward along the last axis */
#define ATUS_KFAM_POINT 4        /* standalone pointwise V step */
#define ATUS_NKFAM 5
int atus_run_block_profiled(atus_ctx *ctx, const atus_step_params *sp, int Nk, double *t, double *ms, int *count,
                            double *grid);

/* tuning: values per thread (8 or 16) of the FFT kernels of one kernel
 * family (ATUS_KFAM_*; the Sobolev filter follows ATUS_KFAM_KIN).  16 halves
 * the LDS exchanges, 8 halves the registers per thread.  Defaults: 16 for
 * the contiguous passes and the kinetic step, 8 for the strided passes and
 * the fused V pass. */
int atus_tune(atus_ctx *ctx, int family, int values_per_thread);
/* tile choice of a kernel family (no reference counterpart): at most
 * `threads` threads per workgroup (64, 128, 256 or 512; 0 = as many as the
 * kernel allows, the default), i.e. threads / (N / values_per_thread)
 * pencils per workgroup, at least one.  Fewer pencils per workgroup need
 * less LDS, so more workgroups fit a CU, but strided passes then read rows
 * in shorter segments.  The results do not change. */
int atus_tune_threads(atus_ctx *ctx, int family, int threads);
/* scheduling of 3D blocks (no reference counterpart): the y and z passes
 * between two kinetic steps only couple points of one x-plane, so they run
 * slab by slab of `planes` x-planes, each slab's passes back to back while
 * its data is still in the last-level cache, alternate slabs on a second
 * stream of the context (joined before each kinetic step).  -1 (default):
 * slabs of about 80 MiB of state; 0: whole-grid passes on one stream.
 * The results do not change. */
int atus_set_slab(atus_ctx *ctx, int planes);
/* scheduling of 1D blocks (no reference counterpart): grids of up to 4096
 * points (single runs of one or two components, and batches) run a whole
 * block of atus_run_block / atus_batch_run_block / a graph in one launch,
 * each run's pencil resident in registers from the first half kinetic step
 * to the last; the per-step t1 and amplitudes are evaluated on the host
 * with the per-step bookkeeping and uploaded before the launch.  on = 1
 * (default) or 0 (a launch per step).  The two agree to rounding: the
 * arithmetic is the same, but multiply-adds are fused differently inside
 * the kernel's step loop. */
int atus_set_resident(atus_ctx *ctx, int on);

/* HBM copy probe: copies all components (ncomp * npts complex128) into a
 * scratch buffer `reps` times with a plain streaming kernel and returns the
 * achieved GB/s (bytes read + written per second, HIP events).  Measures
 * the copy peak the kernels are compared with; no reference counterpart. */
int atus_copy_bandwidth(atus_ctx *ctx, int reps, double *gbps);

/* ---- batches of independent runs of one grid.  The reference scans
 * parameters with one process per case (tammy/src/
 * batch_case_generation_4_tammy.cpp:45-291, each case a bragg run); here B
 * runs share one context and every launch advances all of them.  Run r owns
 * components r*ncomp_per_run .. (r+1)*ncomp_per_run - 1, which atus_upload,
 * atus_download and the single-component reductions address by that index.
 * Runs have two components (the Bragg pair of Bragg_single). */
typedef struct {
  double gs[4];     /* m_gs of the run, row-major 2x2 */
  atus_laser laser; /* the run's CRT_Base_IF laser parameters */
} atus_batch_run;

int atus_batch_create(atus_ctx **out, int dim, const long long n[3], const double dx[3], const double dk[3],
                      int ncomp_per_run, int nbatch);
int atus_batch_set_runs(atus_ctx *ctx, const atus_batch_run *runs, int nbatch);
/* one block of every run, as atus_run_block with kind ATUS_STEP_BRAGG_AD or
 * ATUS_STEP_FREEPROP_IF; F[(m-1)*nbatch + r] is Amplitude_at_time() of run
 * r for V step m = 1..Nk (NULL: all 1).  The t1 of step m is accumulated as
 * in atus_run_block: tt = *t + dt/2, then per step t1 = tt + dt/2, tt += dt
 * (mathematically *t + m*dt). */
int atus_batch_run_block(atus_ctx *ctx, int kind, double dt, int Nk, double *t, const double *F);
/* the same with the per-family device times of atus_run_block_profiled
 * (no reference counterpart; the autotuner's measurement) */
int atus_batch_run_block_profiled(atus_ctx *ctx, int kind, double dt, int Nk, double *t, const double *F, double *ms,
                                  int *count, double *grid);
/* Get_Particle_Number of every component: out[r*ncomp_per_run + c] */
int atus_batch_particle_numbers(atus_ctx *ctx, double *out);
/* No counterpart in the reference.  atus_batch_particle_numbers for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same kernels and the
 * same results, with the partial sums in a fixed 256 KiB device array of
 * the library (one per GPU, held until the process ends) instead of buffers
 * the context allocates on first use.  The call allocates nothing.  A batch
 * needing more than 32768 doubles of partials and results (more than 16384
 * components, i.e. 8192 runs) is refused with ATUS_ERR_ARG.  Calls on one
 * GPU share that array: calls that overlap in time on one GPU (several host
 * threads, or contexts driven concurrently) race. */
int atus_batch_particle_numbers_noalloc(atus_ctx *ctx, double *out);
/* compute_rabi_integrals of component `comp` of every run, transformed in
 * place and back as in CRT_Base_IF.h:208-281: out[r*n + q] */
int atus_batch_rabi_integrals(atus_ctx *ctx, int comp, const double *k0, int n, double threshold, double *out);

/* ---- one block of the sequence hot loop (CRT_Base.h:776-786,
 * CRT_Base_IF.h:779-788) captured once as a HIP graph and replayed: the
 * launches of atus_run_block (single runs; the laser parameters are fixed at
 * capture) or atus_batch_run_block (batches; the run parameters stay those
 * of atus_batch_set_runs).  Each launch takes the amplitudes of the block,
 * F[m-1] (single) or F[(m-1)*nbatch + r] (batch), for V step m = 1..Nk at
 * the t1 the block runner would use (NULL: all 1), and advances *t.
 *
 * What a graph freezes at capture, and what it checks at each launch:
 *  - dt and Nk, the step kind and (single runs) the laser parameters are
 *    the graph's own; they never change.
 *  - The kinetic launches read the context's kinetic tables in place.  A
 *    launch rebuilds them for the graph's dt first, so blocks of another dt
 *    (or atus_set_alpha) in between do not leak into a replay; a replay
 *    uses the context's current alpha.
 *  - The V launches took the couplings (atus_set_g) by value and the
 *    potential buffer by address.  A launch returns ATUS_ERR_STATE if the
 *    couplings, the potential buffer or the species split (atus_set_alpha2)
 *    differ from capture; capture again.
 *  - While a graph of a context exists, atus_ctx_destroy, freeing or first
 *    allocating its potential (atus_set_potential) and atus_set_alpha2
 *    return ATUS_ERR_STATE: destroy the context's graphs first. */
typedef struct atus_graph atus_graph;
int atus_graph_create(atus_ctx *ctx, int kind, double dt, const atus_laser *laser, int Nk, atus_graph **out);
int atus_graph_launch(atus_graph *graph, double *t, const double *F);
int atus_graph_destroy(atus_graph *graph);

/* ---- a parameter scan of runs on different grids as one HIP graph (no
 * reference counterpart: the reference runs one process per case,
 * tammy/src/batch_case_generation_4_tammy.cpp:45-291).  The block of every
 * batch context in ctxs (atus_batch_run_block) and the Rabi integrals of
 * component 0 of all its runs (atus_batch_rabi_integrals; none if nk0 ==
 * 0) are captured once, each context's launches on its own stream, forked
 * from and joined into the graph's stream: the device runs the contexts'
 * launches side by side, and a block of the whole scan is one graph
 * launch.  Launch: F holds each context's amplitudes [Nk][nbatch], the
 * contexts in order (NULL: all 1); rabi receives sum_i nbatch_i x nk0
 * populations in the same order; gpu_ms (optional) the device time of the
 * graph (HIP events around it).  The contexts' state, the run parameters
 * (atus_batch_set_runs) and alpha are read at each launch; dt, Nk and the
 * momentum points are fixed.  While the graph exists its contexts refuse
 * atus_ctx_destroy. */
typedef struct atus_scan_graph atus_scan_graph;
int atus_scan_graph_create(atus_ctx *const *ctxs, int nctx, int kind, double dt, int Nk, const double *k0, int nk0,
                           double threshold, atus_scan_graph **out);
int atus_scan_graph_launch(atus_scan_graph *graph, double *t, const double *F, double *rabi, double *gpu_ms);
int atus_scan_graph_destroy(atus_scan_graph *graph);

/* ---- learned warm start of the Sobolev minimiser (no reference
 * counterpart: the network only proposes the start state, which
 * atus_sob_run then polishes to the same residual).  A perceptron
 * n_in -> h1 -> h2 -> h3 -> n_out, tanh hidden layers, linear output, fp32
 * parameters laid out as [W1 (n_in x h1, row-major), b1, W2, b2, W3, b3,
 * W4, b4]; matrix products take bf16 operands (atus_net_set_precision) and
 * accumulate in fp32 on the matrix cores; training minimises the mean squared error with
 * Adam (beta1 0.9, beta2 0.999, eps 1e-8). */
typedef struct atus_net atus_net;
int atus_net_create(atus_net **out, int n_in, int h1, int h2, int h3, int n_out, int max_batch,
                    unsigned long long seed);
int atus_net_destroy(atus_net *net);
/* operands of every matrix product: 1 = one bf16 value each; 3 (the
 * default) = a bf16 pair hi + lo each and the three products hi*hi +
 * hi*lo + lo*hi, about 16 significant bits, still on the bf16 matrix cores */
int atus_net_set_precision(atus_net *net, int bf16_terms);
int atus_net_param_count(atus_net *net, long long *count);
int atus_net_get_params(atus_net *net, float *host);
/* also resets the optimiser state */
int atus_net_set_params(atus_net *net, const float *host);
/* y (batch x n_out) from x (batch x n_in), host arrays */
int atus_net_forward(atus_net *net, const float *x, int batch, float *y);
/* loss and, if grads != NULL, its gradient w.r.t. every parameter */
int atus_net_gradients(atus_net *net, const float *x, const float *t, int batch, float *loss, float *grads);
/* device-resident training set: count samples of inputs and targets */
int atus_net_set_data(atus_net *net, const float *x, const float *t, int count);
/* `steps` Adam steps on shuffled batches of the resident set, learning
 * rate decaying from lr to lr_final on a cosine; losses[step] if non-NULL */
int atus_net_train(atus_net *net, int steps, int batch, float lr, float lr_final, unsigned long long seed,
                   float *losses);
/* component comp of a 3D context <- the network's 32^3 output for input x,
 * read as periodic samples at x_j = (j - 16) half_width[a] / 16 per axis
 * and interpolated onto the context's grid with the trigonometric
 * interpolant of each axis (0 outside [-half_width, half_width)) */
int atus_net_warm_start(atus_net *net, const float *x, const double half_width[3], atus_ctx *ctx, int comp);

#ifdef __cplusplus
}
#endif
#endif
