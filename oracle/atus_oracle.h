/*
 * atus_oracle.h -- CPU restatement of the ATUS-2 split-step / Sobolev hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check and the cpu_baseline leg of bench.py call it to produce the expected
 * results and the CPU timing.  The product (libatus_hip.so and the host API
 * under atus2_amd/host) never links, loads or calls it.
 *
 * Conventions (all follow the reference; file:line cited per function in
 * atus_oracle.c):
 *   - complex fields are interleaved double (re, im), row-major [nx][ny][nz],
 *     z fastest, exactly the fftw_complex layout of cft_base.h:209;
 *   - x_i = (i - n/2) * dx  (cft_3d.cpp:128-138);
 *   - native FFT order, k_i = dk * (((i + n/2) mod n) - n/2) (cft_3d.cpp:145-166);
 *   - ft(-1): unnormalised e^{-i} DFT, then * prod(dx)/(2pi)^{d/2};
 *     ft(+1): unnormalised e^{+i} DFT, then * prod(dk)/(2pi)^{d/2}
 *     (cft_1d.cpp:191-201, cft_2d.cpp:449, cft_3d.cpp:627-637).
 */
#ifndef ATUS_ORACLE_H
#define ATUS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int dim;          /* 1, 2 or 3 */
  long long n[3];   /* nDimX, nDimY, nDimZ; unused axes are 1 */
  double dx[3];     /* dx, dy, dz */
  double dk[3];     /* dkx, dky, dkz */
} orc_grid;

/* ---- spectral field (cft_base / cft_Nd) ---- */
long long orc_npts(const orc_grid *g);
int orc_threads(void);  /* OpenMP threads the oracle runs with */
double orc_ar(const orc_grid *g);
double orc_ar_k(const orc_grid *g);
void orc_get_x(const orc_grid *g, long long l, double x[3]);
void orc_get_k(const orc_grid *g, long long l, double k[3]);
/* unnormalised in-place DFT of the whole grid; sign -1 = e^{-i}, +1 = e^{+i} */
void orc_fft(const orc_grid *g, double *data, int sign);
/* cft_Nd::ft(isign) with the scale() normalisation */
void orc_ft(const orc_grid *g, double *data, int isign);

/* ---- real-time propagator (CRT_Base / CRT_Base_IF / Bragg_single) ---- */
/* cos/sin(s * (-dt) * sum_i alpha_i k_i^2) per point (CRT_Base::Init) */
void orc_kinetic_table(const orc_grid *g, const double alpha[3], double dt,
                       double s, double *tab);
/* Do_FT_Step_*: ft(-1), multiply by table, ft(+1), per component */
void orc_ft_step(const orc_grid *g, int ncomp, double **psi, const double *tab);
/* CRT_Base::Do_NL_Step; pot may be NULL (potential not initialised) */
void orc_nl_step(const orc_grid *g, int ncomp, double **psi, const double *gs,
                 double dt, double **pot);
/* CRT_Base_IF::Do_NL_Step (gravity beta0 * x0 and detunings DeltaL) */
void orc_nl_if_step(const orc_grid *g, int ncomp, double **psi,
                    const double *gs, double dt, double beta0,
                    const double *deltaL);

typedef struct {
  double dt;          /* Get_dt() */
  double t;           /* Get_t(), time before the step */
  double F;           /* Amplitude_at_time() evaluated at t + dt/2 */
  double amp0;        /* Amp[0] */
  double laser_k;     /* laser_k[0] */
  double laser_domh;  /* laser_domh[0] */
  double laser_dk;    /* laser_dk[0] */
  double chirp;       /* chirp */
  double chirp_rate0; /* chirp_rate[0] */
  double phase0;      /* phase[0] */
  double beta0;       /* beta[0] */
  double deltaL1;     /* DeltaL[1] */
  double gs[4];       /* m_gs, row-major 2x2 */
} orc_bragg_params;
/* Bragg_single::Do_Bragg_ad */
void orc_bragg_ad(const orc_grid *g, double *psi1, double *psi2,
                  const orc_bragg_params *p);

/* ---- reductions ---- */
double orc_particle_number(const orc_grid *g, const double *psi);
void orc_expval_x(const orc_grid *g, const double *psi, double out[3]);
/* CRT_Base::Expval_Momentum: in-place ft(-1) / ft(+1) round trip of psi */
void orc_expval_k(const orc_grid *g, double *psi, double out[3]);
/* CRT_Base_IF::compute_rabi_integrals on one component (round trip) */
void orc_rabi_integrals(const orc_grid *g, double *psi, const double *k0,
                        int nk0, double threshold, double *out);
/* energy functional (no reference counterpart; definition in the .c) */
double orc_energy(const orc_grid *g, int ncomp, double **psi,
                  const double *gs, const double alpha[3], double **pot);

/* ---- sequence block (CRT_Base::run_sequence / CRT_Base_IF::run_sequence) */
typedef double (*orc_amp_fn)(double t, void *user);

enum {
  ORC_FREEPROP = 0,
  ORC_FREEPROP_IF = 1,
  ORC_BRAGG_AD = 2,
  ORC_FREEPROP_IF2 = 3,
  ORC_BRAGG_DS = 4,
  ORC_DOUBLE_BRAGG = 5,
  ORC_NUM_BRAGG = 6,
  ORC_RAMAN = 7
};

typedef struct {
  orc_grid g;
  int ncomp;
  double *psi[4];
  double *pot[4];      /* NULL entries: potential not initialised */
  double gs[16];
  double alpha[3];
  double dt;
  double t;            /* m_header.t, advanced by the block */
  double *half_tab;    /* npts complex, kinetic half step */
  double *full_tab;    /* npts complex, kinetic full step */
  double tab_dt;       /* dt the tables were built for */
  /* interferometry (CRT_Base_IF::UpdateParams) */
  double beta0;
  double deltaL[4];
  double amp0, laser_k, laser_domh, laser_dk, chirp, chirp_rate0, phase0;
  orc_amp_fn amp;      /* NULL => Amplitude_at_time() == 1 */
  void *amp_user;
  /* two species (CRT_Base_2 / CRT_Base_IF_2 / bragg_ds): components
   * [ncomp/2, ncomp) are species 2 with their own kinetic tables and laser */
  int two_species;
  double alpha2[3];
  double *half_tab2, *full_tab2;
  double beta2;
  double amp2, laser_k2, laser_domh2, laser_dk2, chirp_rate2, phase2;
  double seq_time;     /* sequence_item::time, start of the pulses */
  double duration[2];  /* sequence_item::duration[0..1] (pulse gating of bragg_ds) */
  /* three-state steps: Amp[1] and the whole Beta vector */
  double amp1;
  double beta[3];
  /* rounding model (not part of the reference): if jitter > 0, every
   * element of every component is multiplied by (1 + jitter * u), u
   * uniform in [-1, 1) from a counter-based generator keyed by jitter_seed
   * and the number of perturbations so far, after each kinetic and each V
   * application of orc_rt_block.  Two runs that differ only by it show how
   * far last-bit rounding differences at every step carry. */
  double jitter;
  unsigned long long jitter_seed;
  long long jitter_count;
} orc_rt;

/* CRT_Base_2::Do_FT_Step_*: table 1 on components [0, ncomp/2), table 2 on
 * the rest (CRT_Base_2.h:274-348) */
void orc_ft_step2(const orc_grid *g, int ncomp, double **psi, const double *tab1, const double *tab2);
/* CRT_Base_IF_2::Do_NL_Step (CRT_Base_IF_2.h:324-366): species 1 with beta,
 * species 2 with beta2 */
void orc_nl_if2_step(const orc_grid *g, int ncomp, double **psi, const double *gs, double dt, double beta,
                     double beta2, const double *deltaL);
/* bragg_ds Bragg_single::Do_Bragg_ad (bragg_ds.cpp:112-300), four components;
 * t1 = Get_t() + dt/2, mode1/mode2 the pulse gates */
void orc_bragg_ds(const orc_grid *g, double **psi, const orc_rt *s, double t1, int mode1, int mode2);

/* U = exp(-i dt H) of a Hermitian n x n matrix (n <= 4; H and U row-major,
 * interleaved complex): eigen-decomposition by Jacobi rotations in long
 * double, each rotation on the largest remaining off-diagonal element */
void orc_herm_expm(int n, const double *H, double dt, double *U);
/* Bragg_double::Do_Double_Bragg_ad (double_bragg.cpp:108-249), three
 * components; t1 = Get_t() + dt/2, F = Amplitude_at_time() before the clamp */
void orc_double_bragg(const orc_grid *g, double **psi, const orc_rt *s, double t1, double F);
/* CRT_Base_IF::Numerical_Bragg (CRT_Base_IF.h:426-527), ncomp 2 or 3;
 * t = Get_t() before the step */
void orc_num_bragg(const orc_grid *g, int ncomp, double **psi, const orc_rt *s, double t);
/* CRT_Base_IF::Numerical_Raman (CRT_Base_IF.h:534-632), ncomp >= 3 */
void orc_raman(const orc_grid *g, int ncomp, double **psi, const orc_rt *s);

/* one outer iteration of the hot loop: half, (V, full) x (Nk-1), V, half
 * (CRT_Base.h:776-786, CRT_Base_IF.h:779-788) */
void orc_rt_block(orc_rt *s, int kind, int Nk);
/* rebuild tables when dt changed (CRT_shared::Set_dt -> CRT_Base::Init) */
void orc_rt_set_dt(orc_rt *s, double dt);

/* ---- Sobolev gradient minimiser (CSOB_Base) ---- */
typedef struct {
  orc_grid g;
  int nwf;
  double *psi[2], *lap[2], *sob[2], *grad[2];  /* complex, npts each */
  double *op[2], *lapop[2], *pot[2];           /* real, npts each */
  double gs[4];        /* already scaled by sqrt(N_i N_j) */
  double alpha[2];     /* isotropic Alpha_i[0] */
  double stepsize;
  double res[2];
  double res_tot;
  double N[2];
  int renormalised;    /* set when the last iteration renormalised */
} orc_sob;

void orc_sob_init_ops(orc_sob *s);            /* CSOB_Base::Init */
void orc_sob_renormalize(orc_sob *s);         /* Renormalize_All_Psi */
void orc_sob_iter(orc_sob *s);                /* one body of CSOB_Base::run */
void orc_sob_mu(orc_sob *s, double *mu);      /* Compute_mu definition */

#ifdef __cplusplus
}
#endif
#endif
