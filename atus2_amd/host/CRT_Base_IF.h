// CRT_Base_IF.h -- interferometry layer over CRT_Base (the reference's
// CRT_Base_IF<T,dim,no_int_states>, CRT_Base_IF.h:48-946): laser
// parameters, Amplitude_at_time (AMP_T), the gravity/detuning free step,
// Rabi momentum-bin integrals, chirp / phase scans and their text outputs.
//
// Chirp scans restart every chirp from the sequence's start state and hand
// the end state of chirp 0 to the next sequence.  The restart files are
// written as in the reference, and the states are also kept as device
// snapshots, restored without reading the files back unless a later write
// (an output of the same "%.3f" name) has replaced a file; then the file is
// read, which is what the reference's reload sees.
//
// "bragg" and "raman" (CRT_Base_IF.h:426-632: the light couplings of two or
// three states exponentiated through GSL's Hermitian eigensolver) run as
// device steps whose per-point exponential diagonalises the same matrix.
// phase[0] starts at 0 (the reference leaves it uninitialised unless
// chirp_mode == 1).
#ifndef ATUS_AMD_CRT_BASE_IF_H
#define ATUS_AMD_CRT_BASE_IF_H

#include <cstdlib>
#include <list>
#include <sstream>

#include "CRT_Base.h"

template <class T, int dim, int no_int_states>
class CRT_Base_IF : public CRT_Base<T, dim, no_int_states> {
 public:
  explicit CRT_Base_IF(ParameterHandler*);
  virtual ~CRT_Base_IF() = default;
  void run_sequence();

 protected:
  using CRT_Base<T, dim, no_int_states>::m_header;
  using CRT_Base<T, dim, no_int_states>::m_params;
  using CRT_Base<T, dim, no_int_states>::m_fields;
  using CRT_Base<T, dim, no_int_states>::m_custom_fct;
  using CRT_Base<T, dim, no_int_states>::m_set;
  using CRT_shared::m_no_of_pts;

  CPoint<dim> beta;
  std::array<double, no_int_states> DeltaL;
  std::array<double, 2> Amp, laser_k, laser_dk, laser_domh, phase, chirp_rate;
  double chirp;
  bool amp_is_t;

  static void Do_NL_Step_Wrapper(void*, sequence_item&);
  static void Numerical_Bragg_Wrapper(void*, sequence_item&);
  static void Numerical_Raman_Wrapper(void*, sequence_item&);
  void Do_NL_Step();
  void Numerical_Bragg();
  void Numerical_Raman();

  void UpdateParams();
  void Output_rabi_freq_list(std::string, const long long);
  void Output_chirps_list(std::string);

  virtual bool run_custom_sequence(const sequence_item&) override = 0;

  double m_rabi_threshold;
  std::vector<CPoint<dim>> m_rabi_momentum_list;
  std::list<std::list<double>> m_rabi_freq_list;
  std::list<std::list<double>> m_chirps_list;

  void compute_rabi_integrals();
  double Amplitude_at_time();

  // device snapshots beside the chirp scan's restart files "%.3f_<k>.bin"
  // (CRT_Base_IF.h:757-766 and 835-843 write them, 893-903 and 913-921
  // read them back): slot 0 the sequence's start, slot 1 the end of chirp 0
  struct Snapshot {
    std::vector<std::string> names;
    std::vector<long long> writes;
    bool valid = false;
  };
  std::array<Snapshot, 2> m_snap;
  std::vector<std::string> restart_names(double t) const;
  void take_snapshot(int slot, double t);
  // the state held by the restart files of time t: the device snapshot
  // while no write has touched those files since it was taken, else the
  // files themselves, as the reference reads them
  void load_restart_files(int slot, double t);

  // phase scans (chirp_mode == 1) as one batch of independent runs, one
  // per chirp, all started from the sequence's start state; returns the end
  // time of the chirps
  bool phase_scan_batchable(const sequence_item& seq, StepFunction step_fct, int Na, int Nk) const;
  double run_phase_scan_batched(sequence_item& seq, int Na, int Nk, int seq_counter);

  // AMP_T evaluator
  atus_host::Expr m_amp_expr;
  double m_amp_time = 0.0;
  double amplitude_at(double t1) override;
  // ATUS_DEVICE_STEPS (comma-separated sequence names): the step function a
  // driver registered under "bragg_ad" (the reference's Do_Bragg_ad,
  // bragg.cpp:186-304, written on the host mirrors) runs as the device step
  // of the same formula, so an unchanged driver gets the fused path
  void device_step_by_name(const std::string& name, StepFunction step);
  bool has_amplitude() const override { return amp_is_t; }
  void fill_step_params(atus_step_params& sp, const sequence_item& seq) override;
  atus_laser laser_params() const;
};

template <class T, int dim, int no_int_states>
CRT_Base_IF<T, dim, no_int_states>::CRT_Base_IF(ParameterHandler* params) : CRT_Base<T, dim, no_int_states>(params) {
  this->m_map_stepfcts["freeprop"] = &Do_NL_Step_Wrapper;
  this->m_map_stepfcts["bragg"] = &Numerical_Bragg_Wrapper;
  this->m_map_stepfcts["raman"] = &Numerical_Raman_Wrapper;
  this->m_device_steps.erase(&CRT_Base<T, dim, no_int_states>::Do_NL_Step_Wrapper);
  this->m_device_steps[&Do_NL_Step_Wrapper] = ATUS_STEP_FREEPROP_IF;
  if (no_int_states == 2 || no_int_states == 3) this->m_device_steps[&Numerical_Bragg_Wrapper] = ATUS_STEP_NUM_BRAGG;
  if (no_int_states >= 3) this->m_device_steps[&Numerical_Raman_Wrapper] = ATUS_STEP_RAMAN;
  phase.fill(0.0);
  chirp_rate.fill(0.0);
  Amp.fill(0.0);
  laser_k.fill(0.0);
  laser_dk.fill(0.0);
  laser_domh.fill(0.0);
  amp_is_t = false;
  UpdateParams();
}

// CRT_Base_IF::UpdateParams (CRT_Base_IF.h:145-168)
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::UpdateParams() {
  for (int i = 0; i < dim; i++) beta[i] = m_params->Get_VConstant("Beta", i);
  for (int i = 0; i < no_int_states; i++) DeltaL[i] = m_params->Get_VConstant("Delta_L", i);
  Amp[0] = m_params->Get_VConstant("Amp_1", 0);
  Amp[1] = m_params->Get_VConstant("Amp_1", 1);
  laser_k[0] = m_params->Get_Constant("laser_k");
  laser_dk[0] = m_params->Get_Constant("laser_dk");
  laser_domh[0] = m_params->Get_Constant("laser_domh");
  m_rabi_threshold = m_params->Get_Constant("rabi_threshold");
  chirp = m_params->Get_Constant("chirp");
}

template <class T, int dim, int no_int_states>
double CRT_Base_IF<T, dim, no_int_states>::amplitude_at(double t1) {
  m_amp_time = t1;
  return m_amp_expr.Eval();
}

template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::device_step_by_name(const std::string& name, StepFunction step) {
  if (this->m_device_steps.count(step) || no_int_states != 2 || name != "bragg_ad") return;
  const char* e = std::getenv("ATUS_DEVICE_STEPS");
  if (!e) return;
  std::stringstream ss(e);
  std::string item;
  while (std::getline(ss, item, ','))
    if (item == name) {
      this->m_device_steps[step] = ATUS_STEP_BRAGG_AD;
      std::cout << "FYI: " << name << " runs as the device step (ATUS_DEVICE_STEPS)\n";
      return;
    }
}

// CRT_Base_IF::Amplitude_at_time (CRT_Base_IF.h:171-198)
template <class T, int dim, int no_int_states>
double CRT_Base_IF<T, dim, no_int_states>::Amplitude_at_time() {
  if (!amp_is_t) return 1;
  // host steps written for the reference call this from every thread of a
  // parallel region (bragg.cpp:201); the evaluator's time variable is shared
  double F;
#pragma omp critical(atus_amplitude)
  F = amplitude_at(this->Get_t() + 0.5 * this->Get_dt());
  return F;
}

template <class T, int dim, int no_int_states>
atus_laser CRT_Base_IF<T, dim, no_int_states>::laser_params() const {
  atus_laser L{};
  L.amp0 = Amp[0];
  L.laser_k = laser_k[0];
  L.laser_domh = laser_domh[0];
  L.laser_dk = laser_dk[0];
  L.chirp = chirp;
  L.chirp_rate0 = chirp_rate[0];
  L.phase0 = phase[0];
  L.beta0 = beta[0];
  for (int i = 0; i < no_int_states && i < 4; ++i) L.deltaL[i] = DeltaL[i];
  L.amp1 = Amp[1];
  for (int a = 0; a < dim; ++a) L.beta[a] = beta[a];
  return L;
}

template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::fill_step_params(atus_step_params& sp, const sequence_item&) {
  sp.laser = laser_params();
  sp.amp = nullptr;
  sp.amp_user = nullptr;
}

template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::Do_NL_Step_Wrapper(void* ptr, sequence_item&) {
  static_cast<CRT_Base_IF*>(ptr)->Do_NL_Step();
}
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::Numerical_Bragg_Wrapper(void* ptr, sequence_item&) {
  static_cast<CRT_Base_IF*>(ptr)->Numerical_Bragg();
}
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::Numerical_Raman_Wrapper(void* ptr, sequence_item&) {
  static_cast<CRT_Base_IF*>(ptr)->Numerical_Raman();
}

// CRT_Base_IF::Numerical_Bragg (CRT_Base_IF.h:426-527): state 0 coupled to
// states 1 and 2 by lasers of wave numbers +-laser_k and chirps
// +-chirp_rate[0] at t = Get_t(); two or three states
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::Numerical_Bragg() {
  laser_k[1] = -laser_k[0];
  chirp_rate[1] = -chirp_rate[0];
  const atus_laser L = laser_params();
  atus_host::check(atus_v_step(m_set->device(), ATUS_STEP_NUM_BRAGG, m_header.dt, m_header.t, 1.0, &L),
                   "atus_v_step");
}

// CRT_Base_IF::Numerical_Raman (CRT_Base_IF.h:534-632): states 0 and 1
// coupled to state 2 (three or four states)
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::Numerical_Raman() {
  const atus_laser L = laser_params();
  atus_host::check(atus_v_step(m_set->device(), ATUS_STEP_RAMAN, m_header.dt, m_header.t, 1.0, &L), "atus_v_step");
}

// CRT_Base_IF::Do_NL_Step (CRT_Base_IF.h:381-419)
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::Do_NL_Step() {
  double dl[4] = {0, 0, 0, 0};
  for (int i = 0; i < no_int_states && i < 4; ++i) dl[i] = DeltaL[i];
  atus_host::check(atus_nl_if_step(m_set->device(), m_header.dt, beta[0], dl), "atus_nl_if_step");
}

// CRT_Base_IF::compute_rabi_integrals (CRT_Base_IF.h:208-281); the device
// transforms a copy, so psi is left as it was
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::compute_rabi_integrals() {
  const int n = (int)m_rabi_momentum_list.size();
  if (n == 0) {
    std::cerr << "WARNING: Rabi momentum list is empty. Cannot compute rabi frquencies." << std::endl;
    return;
  }
  std::vector<double> k0(3 * (size_t)n, 0.0), out(n);
  for (int q = 0; q < n; ++q)
    for (int a = 0; a < dim; ++a) k0[3 * q + a] = m_rabi_momentum_list[q][a];
  atus_host::check(atus_rabi_integrals(m_set->device_readonly(), 0, k0.data(), n, m_rabi_threshold, out.data()),
                   "atus_rabi_integrals");
  m_rabi_freq_list.push_back(std::list<double>(out.begin(), out.end()));
}

// CRT_Base_IF::Output_rabi_freq_list (CRT_Base_IF.h:289-314)
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::Output_rabi_freq_list(std::string filename, const long long Nk) {
  std::ofstream txtfile(filename);
  double t = 0;
  txtfile << "# time \t";
  for (auto& i : m_rabi_momentum_list) txtfile << i << "\t";
  txtfile << std::endl;
  for (auto& i : m_rabi_freq_list) {
    txtfile << t << "\t";
    for (auto j : i) txtfile << j << "\t";
    txtfile << std::endl;
    t += m_header.dt * double(Nk);
  }
}

// CRT_Base_IF::Output_chirps_list (CRT_Base_IF.h:320-342)
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::Output_chirps_list(std::string filename) {
  std::ofstream txtfile(filename);
  txtfile << "# Step \t Chirp \t ";
  for (auto& i : m_rabi_momentum_list) txtfile << i << "\t";
  txtfile << std::endl;
  for (auto& i : m_chirps_list) {
    for (auto j : i) {
      txtfile.precision(8);
      txtfile << j << "\t";
    }
    txtfile << std::endl;
  }
}

template <class T, int dim, int no_int_states>
std::vector<std::string> CRT_Base_IF<T, dim, no_int_states>::restart_names(double t) const {
  std::vector<std::string> out;
  char filename[1024];
  for (int k = 0; k < no_int_states; k++) {
    snprintf(filename, sizeof(filename), "%.3f_%d.bin", t, k + 1);
    out.emplace_back(filename);
  }
  return out;
}

template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::take_snapshot(int slot, double t) {
  Snapshot& s = m_snap[slot];
  s.names = restart_names(t);
  s.writes.clear();
  for (auto& n : s.names) s.writes.push_back(this->m_file_writes[n]);
  atus_host::check(atus_snapshot_save(m_set->device_readonly(), slot), "atus_snapshot_save");
  s.valid = true;
}

template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::load_restart_files(int slot, double t) {
  const Snapshot& s = m_snap[slot];
  const std::vector<std::string> names = restart_names(t);
  bool current = s.valid && s.names == names;
  for (size_t k = 0; current && k < names.size(); ++k) current = this->m_file_writes[names[k]] == s.writes[k];
  if (current) {
    atus_host::check(atus_snapshot_restore(m_set->device(), slot), "atus_snapshot_restore");
    return;
  }
  for (int k = 0; k < no_int_states; k++) {
    std::ifstream file1(names[k], std::ifstream::binary);
    file1.seekg(sizeof(generic_header), std::ifstream::beg);
    file1.read((char*)m_fields[k]->Getp2In(), sizeof(fftw_complex) * m_no_of_pts);
  }
}

// A phase scan records no dphi, so the bisection of the chirp rates
// (CRT_Base_IF.h:905-912) never swaps and every chirp's rate and phase are
// known before the first chirp: the chirps are independent runs from the
// same start state.  They run as one batch when nothing needs the state of
// a chirp other than 0 between blocks (no custom function, no per-block
// snapshots whose names could overwrite the restart files) and the batch
// fits in 64 GiB.  ATUS_CHIRP_BATCH=0 turns this off.
template <class T, int dim, int no_int_states>
bool CRT_Base_IF<T, dim, no_int_states>::phase_scan_batchable(const sequence_item& seq, StepFunction step_fct, int Na,
                                                               int Nk) const {
  if (no_int_states != 2 || seq.no_of_chirps < 2 || seq.chirp_mode != 1 || Na < 1) return false;
  if (m_rabi_momentum_list.empty() || m_rabi_momentum_list.size() > 8) return false;
  auto it = this->m_device_steps.find(step_fct);
  if (it == this->m_device_steps.end() || it->second != ATUS_STEP_BRAGG_AD) return false;
  if (m_custom_fct != nullptr && seq.custom_freq != freq::none) return false;
  if (seq.output_freq == freq::each) return false;
  // restart files (start) and end files carry different "%.3f" names
  if (double(Na) * double(Nk) * seq.dt < 0.0011) return false;
  if (const char* e = std::getenv("ATUS_CHIRP_BATCH"))
    if (std::string(e) == "0") return false;
  const double bytes = double(seq.no_of_chirps) * no_int_states * double(m_no_of_pts) * sizeof(fftw_complex);
  return bytes <= 64.0 * double(1ull << 30);
}

// the chirp loop of CRT_Base_IF::run_sequence (CRT_Base_IF.h:769-921) for a
// phase scan, all chirps advancing together (atus_batch_*); what each chirp
// prints is collected and printed chirp after chirp, as the loop prints it
template <class T, int dim, int no_int_states>
double CRT_Base_IF<T, dim, no_int_states>::run_phase_scan_batched(sequence_item& seq, int Na, int Nk, int seq_counter) {
  const int nch = seq.no_of_chirps;
  const int nrm = (int)m_rabi_momentum_list.size();
  const double dt = m_header.dt;
  std::vector<double> dw(nch, 0.0);
  dw[0] = seq.chirp_min;
  dw[1] = seq.chirp_max;
  for (int s = 1; s < nch - 1; ++s) dw[s + 1] = (dw[s] + dw[s - 1]) / 2;
  std::vector<atus_batch_run> runs(nch);
  for (int s = 0; s < nch; ++s) {
    chirp_rate[0] = dw[s];
    phase[0] = s * 2.0 * M_PI / (1.0 * nch);
    runs[s].laser = laser_params();
    for (int i = 0; i < 4; ++i) runs[s].gs[i] = this->m_gs[i];
  }
  struct BatchCtx {
    atus_ctx* c;
    ~BatchCtx() { atus_ctx_destroy(c); }
  } b{m_set->make_batch(nch)};
  atus_ctx* single = m_set->device_readonly();
  for (int s = 0; s < nch; ++s)
    atus_host::check(atus_copy_components(b.c, 2 * s, single, 0, 2), "atus_copy_components");
  double al[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) al[a] = this->m_alpha[a];
  atus_host::check(atus_set_alpha(b.c, al), "atus_set_alpha");
  atus_host::check(atus_batch_set_runs(b.c, runs.data(), nch), "atus_batch_set_runs");

  std::vector<std::ostringstream> out(nch);
  std::vector<std::list<std::list<double>>> rabi(nch);
  std::vector<double> F(this->has_amplitude() ? (size_t)Nk * nch : 0), k0(3 * (size_t)nrm, 0.0);
  std::vector<double> rv((size_t)nrm * nch), pn((size_t)2 * nch);
  for (int q = 0; q < nrm; ++q)
    for (int a = 0; a < dim; ++a) k0[3 * q + a] = m_rabi_momentum_list[q][a];
  auto batch_rabi = [&]() {
    atus_host::check(atus_batch_rabi_integrals(b.c, 0, k0.data(), nrm, m_rabi_threshold, rv.data()),
                     "atus_batch_rabi_integrals");
    for (int s = 0; s < nch; ++s)
      rabi[s].push_back(std::list<double>(rv.begin() + (size_t)s * nrm, rv.begin() + (size_t)(s + 1) * nrm));
  };
  auto batch_pn = [&]() {
    atus_host::check(atus_batch_particle_numbers(b.c, pn.data()), "atus_batch_particle_numbers");
    for (int s = 0; s < nch; ++s)
      for (int c = 0; c < 2; ++c) out[s] << "N[" << c << "] = " << pn[2 * s + c] << "\n";
  };
  // chirp 0's state into the single-run context, for its file outputs
  auto chirp0_to_single = [&](double t) {
    atus_host::check(atus_copy_components(m_set->device(), 0, b.c, 0, 2), "atus_copy_components");
    m_header.t = t;
  };
  char filename[1024];
  double t = m_header.t;
  for (int i = 1; i <= Na; i++) {
    if (this->has_amplitude()) {
      // the t1 of the block's V steps, accumulated as the block runner does
      double tt = t + 0.5 * dt;
      for (int m = 1; m <= Nk; ++m) {
        const double f = this->amplitude_at(tt + 0.5 * dt);
        for (int s = 0; s < nch; ++s) F[(size_t)(m - 1) * nch + s] = f;
        tt += m < Nk ? dt : 0.5 * dt;
      }
    }
    atus_host::check(atus_batch_run_block(b.c, ATUS_STEP_BRAGG_AD, dt, Nk, &t, F.empty() ? nullptr : F.data()),
                     "atus_batch_run_block");
    for (int s = 0; s < nch; ++s) out[s] << "t = " << std::to_string(t) << "\n";
    if (seq.output_freq == freq::packed) {
      chirp0_to_single(t);
      for (int k = 0; k < no_int_states; k++) {
        snprintf(filename, sizeof(filename), "Seq_%d_%d.bin", seq_counter, k + 1);
        this->Append_Phi(filename, k);
      }
    }
    if (seq.compute_pn_freq == freq::each) batch_pn();
    if (seq.rabi_output_freq == freq::each) batch_rabi();
  }
  // chirp 0's end: outputs, restart files and snapshot (CRT_Base_IF.h:826-843)
  chirp0_to_single(t);
  if (seq.output_freq == freq::last)
    for (int k = 0; k < no_int_states; k++) {
      snprintf(filename, sizeof(filename), "%.3f_%d.bin", t, k + 1);
      this->Save_Phi(filename, k);
    }
  for (int k = 0; k < no_int_states; k++) {
    snprintf(filename, sizeof(filename), "%.3f_%d.bin", t, k + 1);
    this->Save_Phi(filename, k);
  }
  take_snapshot(1, t);
  if (seq.compute_pn_freq == freq::last) batch_pn();
  if (seq.rabi_output_freq == freq::last) {
    batch_rabi();
    for (int s = 0; s < nch; ++s) {
      out[s] << "Rabi output for timestep ";
      for (auto& row : rabi[s]) {
        out[s] << t << "\n";
        for (auto j : row) out[s] << j << "\t";
        out[s] << "\n";
      }
    }
  }
  for (int s = 0; s < nch; ++s) {
    std::cout << out[s].str() << std::flush;
    if (seq.rabi_output_freq == freq::each) {
      m_rabi_freq_list = rabi[s];
      snprintf(filename, sizeof(filename), "Rabi_%d_%d.txt", seq_counter, s);
      Output_rabi_freq_list(filename, seq.Nk);
    }
  }
  // the Rabi populations at the end of every chirp (CRT_Base_IF.h:884-899)
  atus_host::check(atus_batch_rabi_integrals(b.c, 0, k0.data(), nrm, m_rabi_threshold, rv.data()),
                   "atus_batch_rabi_integrals");
  for (int s = 0; s < nch; ++s) {
    std::list<double> row(rv.begin() + (size_t)s * nrm, rv.begin() + (size_t)(s + 1) * nrm);
    row.push_front(s * 2.0 * M_PI / (1.0 * nch));
    row.push_front(s);
    m_chirps_list.push_back(row);
  }
  m_rabi_freq_list = rabi[nch - 1];
  m_rabi_freq_list.push_back(std::list<double>(rv.end() - nrm, rv.end()));
  chirp_rate[0] = dw[nch - 1];
  phase[0] = (nch - 1) * 2.0 * M_PI / (1.0 * nch);
  return t;
}

// CRT_Base_IF::run_sequence (CRT_Base_IF.h:639-946)
template <class T, int dim, int no_int_states>
void CRT_Base_IF<T, dim, no_int_states>::run_sequence() {
  StepFunction step_fct = nullptr, half_step_fct = nullptr, full_step_fct = nullptr;
  char filename[1024];
  std::cout << "FYI: Found " << m_params->m_sequence.size() << " sequences." << std::endl;
  const int nrm = (int)m_rabi_momentum_list.size();
  if (nrm == 0) std::cerr << "WARNING: Rabi momentum list is empty. Cannot compute rabi frquencies." << std::endl;
  try {
    half_step_fct = this->m_map_stepfcts.at("half_step");
    full_step_fct = this->m_map_stepfcts.at("full_step");
  } catch (const std::out_of_range&) {
    throw std::string("Critical Error: Invalid fct ptr to ft_half_step or ft_full_step");
  }
  this->sync_params();
  int seq_counter = 1;
  for (auto seq : m_params->m_sequence) {
    if (run_custom_sequence(seq)) {
      seq_counter++;
      continue;
    }
    if (seq.name == "set_momentum") {
      std::vector<std::string> vec;
      {
        std::string cur;
        for (char ch : seq.content) {
          if (ch == ',') {
            vec.push_back(cur);
            cur.clear();
          } else
            cur += ch;
        }
        vec.push_back(cur);
      }
      if (!((int)vec.size() > dim)) throw std::string("set_momentum needs more than dim comma-separated values");
      if (!(seq.comp <= dim)) throw std::string("set_momentum: component index out of range");
      CPoint<dim> P;
      for (int i = 0; i < dim; i++) P[i] = std::stod(vec[i]);
      this->Setup_Momentum(P, seq.comp);
      std::cout << "FYI: started new sequence " << seq.name << "\n";
      std::cout << "FYI: momentum set for component " << seq.comp << "\n";
      continue;
    }
    double max_duration = 0;
    for (double d : seq.duration)
      if (d > max_duration) max_duration = d;
    int subN = int(max_duration / seq.dt);
    int Nk = seq.Nk;
    int Na = subN / seq.Nk;
    std::cout << "FYI: started new sequence " << seq.name << "\n";
    std::cout << "FYI: sequence no : " << seq_counter << "\n";
    std::cout << "FYI: duration    : " << max_duration << "\n";
    std::cout << "FYI: dt          : " << seq.dt << "\n";
    std::cout << "FYI: Na          : " << Na << "\n";
    std::cout << "FYI: Nk          : " << Nk << "\n";
    std::cout << "FYI: Na*Nk*dt    : " << double(Na * Nk) * seq.dt << "\n";
    try {
      std::string amp = m_params->Get_simulation("AMP_T");
      std::cout << "FYI: Amp is      : " << amp << "\n";
      amp_is_t = true;
      m_params->Setup_muParser(m_amp_expr);
      m_amp_expr.SetExpr(amp);
      m_amp_expr.DefineVar("t", &m_amp_time);
    } catch (std::string&) {
      std::cout << "FYI: Amp is      : 1" << std::endl;
      amp_is_t = false;
    }
    if (double(Na * Nk) * seq.dt != max_duration) std::cout << "FYI: double(Na*Nk)*seq.dt != max_duration\n";
    if (this->Get_dt() != seq.dt) this->Set_dt(seq.dt);
    auto it = this->m_map_stepfcts.find(seq.name);
    if (it == this->m_map_stepfcts.end()) throw std::string("Critical Error: Invalid squence name " + seq.name);
    step_fct = it->second;
    device_step_by_name(seq.name, step_fct);
    if (seq.name == "freeprop") seq.no_of_chirps = 1;
    const double backup_t = m_header.t;
    double backup_end_t = m_header.t;
    for (int k = 0; k < no_int_states; k++) {
      snprintf(filename, sizeof(filename), "Seq_%d_%d.bin", seq_counter, k + 1);
      std::remove(filename);
    }
    m_chirps_list.clear();
    const int nch = seq.no_of_chirps > 0 ? seq.no_of_chirps : 1;
    std::vector<double> dw(std::max(nch, 2), 0.0), dphi(std::max(nch, 2), 0.0);
    if (seq.no_of_chirps > 1) {
      for (int k = 0; k < no_int_states; k++) {
        snprintf(filename, sizeof(filename), "%.3f_%d.bin", this->Get_t(), k + 1);
        this->Save_Phi(filename, k);
      }
      take_snapshot(0, this->Get_t());
      dw[0] = seq.chirp_min;
      dw[1] = seq.chirp_max;
    } else {
      dw[0] = 0;
    }
    const bool fused = this->device_block_possible(step_fct);
    const bool batched = fused && phase_scan_batchable(seq, step_fct, Na, Nk);
    if (batched) backup_end_t = run_phase_scan_batched(seq, Na, Nk, seq_counter);
    for (int s = 0; s < (batched ? 0 : nch); ++s) {
      m_rabi_freq_list.clear();
      chirp_rate[0] = dw[s];
      if (seq.chirp_mode == 1) phase[0] = s * 2.0 * M_PI / (1.0 * seq.no_of_chirps);
      for (int i = 1; i <= Na; i++) {
        if (fused) {
          this->run_device_block(step_fct, seq, Nk);
        } else {
          (*half_step_fct)(this, seq);
          for (int j = 2; j <= Nk; j++) {
            (*step_fct)(this, seq);
            (*full_step_fct)(this, seq);
          }
          (*step_fct)(this, seq);
          (*half_step_fct)(this, seq);
        }
        std::cout << "t = " << std::to_string(m_header.t) << std::endl;
        if (seq.output_freq == freq::each && s == 0)
          for (int k = 0; k < no_int_states; k++) {
            snprintf(filename, sizeof(filename), "%.3f_%d.bin", this->Get_t(), k + 1);
            this->Save_Phi(filename, k);
          }
        if (seq.output_freq == freq::packed && s == 0)
          for (int k = 0; k < no_int_states; k++) {
            snprintf(filename, sizeof(filename), "Seq_%d_%d.bin", seq_counter, k + 1);
            this->Append_Phi(filename, k);
          }
        if (seq.compute_pn_freq == freq::each)
          for (int c = 0; c < no_int_states; c++)
            std::cout << "N[" << c << "] = " << this->Get_Particle_Number(c) << std::endl;
        if (seq.rabi_output_freq == freq::each) compute_rabi_integrals();
        if (seq.custom_freq == freq::each && m_custom_fct != nullptr) (*m_custom_fct)(this, seq);
      }
      if (seq.output_freq == freq::last && s == 0)
        for (int k = 0; k < no_int_states; k++) {
          snprintf(filename, sizeof(filename), "%.3f_%d.bin", this->Get_t(), k + 1);
          this->Save_Phi(filename, k);
        }
      if (seq.no_of_chirps > 1 && s == 0) {
        backup_end_t = this->Get_t();
        for (int k = 0; k < no_int_states; k++) {
          snprintf(filename, sizeof(filename), "%.3f_%d.bin", this->Get_t(), k + 1);
          this->Save_Phi(filename, k);
        }
        take_snapshot(1, this->Get_t());
      }
      if (seq.compute_pn_freq == freq::last)
        for (int c = 0; c < no_int_states; c++)
          std::cout << "N[" << c << "] = " << this->Get_Particle_Number(c) << std::endl;
      if (seq.custom_freq == freq::last && m_custom_fct != nullptr) (*m_custom_fct)(this, seq);
      if (seq.rabi_output_freq == freq::last) {
        compute_rabi_integrals();
        std::cout << "Rabi output for timestep ";
        for (auto& i : m_rabi_freq_list) {
          std::cout << this->Get_t() << "\n";
          for (auto j : i) std::cout << j << "\t";
          std::cout << std::endl;
        }
      }
      if (seq.rabi_output_freq == freq::each) {
        snprintf(filename, sizeof(filename), "Rabi_%d_%d.txt", seq_counter, s);
        Output_rabi_freq_list(filename, seq.Nk);
      }
      if (seq.no_of_chirps > 1) {
        if (nrm > 0) {
          compute_rabi_integrals();
          std::list<double> tmp = m_rabi_freq_list.back();
          if (seq.chirp_mode == 1) {
            tmp.push_front(phase[0]);
          } else {
            dphi[s] = tmp.front();
            tmp.push_front(chirp_rate[0]);
          }
          tmp.push_front(s);
          m_chirps_list.push_back(tmp);
        }
        if (s < seq.no_of_chirps - 1) {
          load_restart_files(0, backup_t);
          m_header.t = backup_t;
        }
        if (s > 0 && s < seq.no_of_chirps - 1) {
          dw[s + 1] = (dw[s] + dw[s - 1]) / 2;
          if (dphi[s] - dphi[s - 1] < 0) {
            dw[s] = dw[s - 1];
            dphi[s] = dphi[s - 1];
          }
        }
      }
    }
    if (seq.no_of_chirps > 1) {
      snprintf(filename, sizeof(filename), "Chirp_%d.txt", seq_counter);
      Output_chirps_list(filename);
      load_restart_files(1, backup_end_t);
      m_header.t = backup_end_t;
    }
    seq_counter++;
  }
}

#endif
