// CRT_Base_IF_2.h -- interferometry layer of the two-species propagator (the
// reference's CRT_Base_IF_2<T,dim,no_int_states>, CRT_Base_IF_2.h:40-595):
// per-species laser parameters (Beta / Beta_2, Amp_1 / Amp_2, laser_k /
// laser_k_2, laser_dk / laser_dk_2, laser_domh / laser_domh_2), the
// two-species gravity/detuning free step, Rabi momentum-bin integrals of
// both species and their Spezie_*_Rabi_*.txt outputs.
//
// The free step and the derived classes' device steps run inside fused
// device blocks (atus_run_block_ds); called one by one (custom kinetic
// steps) they are single device V steps (atus_v_step_ds).
//
// Differences from the reference, all deliberate:
//   * phase, phase2, chirp_rate and chirp_rate2 start at 0 (the reference
//     leaves them uninitialised; only chirp_rate[0] is set, to 0, before
//     each chirp);
//   * parameter errors in UpdateParams propagate (the reference prints the
//     message and continues with the fields unset).
#ifndef ATUS_AMD_CRT_BASE_IF_2_H
#define ATUS_AMD_CRT_BASE_IF_2_H

#include <list>

#include "CRT_Base_2.h"

template <class T, int dim, int no_int_states>
class CRT_Base_IF_2 : public CRT_Base_2<T, dim, no_int_states> {
 public:
  explicit CRT_Base_IF_2(ParameterHandler*);
  virtual ~CRT_Base_IF_2() = default;
  void run_sequence();

 protected:
  using CRT_Base_2<T, dim, no_int_states>::m_header;
  using CRT_Base_2<T, dim, no_int_states>::m_params;
  using CRT_Base_2<T, dim, no_int_states>::m_fields;
  using CRT_Base_2<T, dim, no_int_states>::m_custom_fct;
  using CRT_Base_2<T, dim, no_int_states>::m_set;
  using CRT_shared::m_no_of_pts;

  CPoint<dim> beta, beta2;
  std::array<double, no_int_states> DeltaL;
  std::array<double, 2> Amp, Amp2, laser_k, laser_k2, laser_dk, laser_dk2, laser_domh, laser_domh2, phase, phase2,
      chirp_rate, chirp_rate2;

  static void Do_NL_Step_Wrapper(void*, sequence_item&);
  void Do_NL_Step();

  void UpdateParams();
  void Output_rabi_freq_list(std::string, std::string, const long long);
  void Output_chirps_list(std::string);

  virtual bool run_custom_sequence(const sequence_item&) override = 0;

  double m_rabi_threshold;
  std::vector<CPoint<dim>> m_rabi_momentum_list;   // species 1, component 0
  std::vector<CPoint<dim>> m_rabi_momentum_list2;  // species 2, component no_int_states/2
  std::list<std::list<double>> m_rabi_freq_list;
  std::list<std::list<double>> m_rabi_freq_list2;
  std::list<std::list<double>> m_chirps_list;

  void compute_rabi_integrals();

  // device parameters: laser1 carries species 1 and the detunings, ds
  // species 2, the sequence start and the pulse lengths
  atus_laser laser1_params() const;
  atus_ds_params ds_params(const sequence_item& seq) const;
  bool device_block(StepFunction step, sequence_item& seq, int Nk) override;
};

template <class T, int dim, int no_int_states>
CRT_Base_IF_2<T, dim, no_int_states>::CRT_Base_IF_2(ParameterHandler* params)
    : CRT_Base_2<T, dim, no_int_states>(params) {
  this->m_map_stepfcts["freeprop"] = &Do_NL_Step_Wrapper;
  this->m_device_steps.erase(&CRT_Base<T, dim, no_int_states>::Do_NL_Step_Wrapper);
  this->m_device_steps[&Do_NL_Step_Wrapper] = ATUS_STEP_FREEPROP_IF;
  for (auto* a : {&Amp, &Amp2, &laser_k, &laser_k2, &laser_dk, &laser_dk2, &laser_domh, &laser_domh2, &phase, &phase2,
                  &chirp_rate, &chirp_rate2})
    a->fill(0.0);
  DeltaL.fill(0.0);
  UpdateParams();
}

// CRT_Base_IF_2::UpdateParams (CRT_Base_IF_2.h:101-137)
template <class T, int dim, int no_int_states>
void CRT_Base_IF_2<T, dim, no_int_states>::UpdateParams() {
  for (int i = 0; i < dim; i++) {
    beta[i] = m_params->Get_VConstant("Beta", i);
    beta2[i] = m_params->Get_VConstant("Beta_2", i);
  }
  for (int i = 0; i < no_int_states; i++) DeltaL[i] = m_params->Get_VConstant("Delta_L", i);
  Amp[0] = m_params->Get_VConstant("Amp_1", 0);
  Amp[1] = m_params->Get_VConstant("Amp_1", 1);
  Amp2[0] = m_params->Get_VConstant("Amp_2", 0);
  Amp2[1] = m_params->Get_VConstant("Amp_2", 1);
  laser_k[0] = m_params->Get_Constant("laser_k");
  laser_k2[0] = m_params->Get_Constant("laser_k_2");
  laser_dk[0] = m_params->Get_Constant("laser_dk");
  laser_dk2[0] = m_params->Get_Constant("laser_dk_2");
  laser_domh[0] = m_params->Get_Constant("laser_domh");
  laser_domh2[0] = m_params->Get_Constant("laser_domh_2");
  m_rabi_threshold = m_params->Get_Constant("rabi_threshold");
}

template <class T, int dim, int no_int_states>
atus_laser CRT_Base_IF_2<T, dim, no_int_states>::laser1_params() const {
  atus_laser L{};
  L.amp0 = Amp[0];
  L.laser_k = laser_k[0];
  L.laser_domh = laser_domh[0];
  L.laser_dk = laser_dk[0];
  L.chirp_rate0 = chirp_rate[0];
  L.phase0 = phase[0];
  L.beta0 = beta[0];
  for (int i = 0; i < no_int_states && i < 4; ++i) L.deltaL[i] = DeltaL[i];
  return L;
}

template <class T, int dim, int no_int_states>
atus_ds_params CRT_Base_IF_2<T, dim, no_int_states>::ds_params(const sequence_item& seq) const {
  atus_ds_params ds{};
  ds.laser2.amp0 = Amp2[0];
  ds.laser2.laser_k = laser_k2[0];
  ds.laser2.laser_domh = laser_domh2[0];
  ds.laser2.laser_dk = laser_dk2[0];
  ds.laser2.chirp_rate0 = chirp_rate2[0];
  ds.laser2.phase0 = phase2[0];
  ds.laser2.beta0 = beta2[0];
  ds.seq_time = seq.time;
  // the pulse gates read duration[0] and duration[1] (bragg_ds.cpp:124-128);
  // a sequence with one duration leaves the second pulse on throughout (the
  // reference reads past the end of the vector there)
  ds.duration[0] = seq.duration.size() > 0 ? seq.duration[0] : 1e300;
  ds.duration[1] = seq.duration.size() > 1 ? seq.duration[1] : 1e300;
  return ds;
}

template <class T, int dim, int no_int_states>
bool CRT_Base_IF_2<T, dim, no_int_states>::device_block(StepFunction step, sequence_item& seq, int Nk) {
  if (!this->device_block_possible(step)) return false;
  const int kind = this->m_device_steps.at(step);
  if (kind == ATUS_STEP_FREEPROP) return CRT_Base_2<T, dim, no_int_states>::device_block(step, seq, Nk);
  atus_step_params sp{};
  sp.kind = kind;
  sp.dt = m_header.dt;
  sp.laser = laser1_params();
  const atus_ds_params ds = ds_params(seq);
  atus_host::check(atus_run_block_ds(m_set->device(), &sp, &ds, Nk, &m_header.t), "atus_run_block_ds");
  return true;
}

template <class T, int dim, int no_int_states>
void CRT_Base_IF_2<T, dim, no_int_states>::Do_NL_Step_Wrapper(void* ptr, sequence_item&) {
  static_cast<CRT_Base_IF_2*>(ptr)->Do_NL_Step();
}

// CRT_Base_IF_2::Do_NL_Step (CRT_Base_IF_2.h:324-366): species s sees
// beta_s[0] x - DeltaL[i] besides the contact terms; no trap potential
template <class T, int dim, int no_int_states>
void CRT_Base_IF_2<T, dim, no_int_states>::Do_NL_Step() {
  const atus_laser L = laser1_params();
  sequence_item none{};
  const atus_ds_params ds = ds_params(none);
  atus_host::check(atus_v_step_ds(m_set->device(), ATUS_STEP_FREEPROP_IF, m_header.dt, m_header.t, &L, &ds),
                   "atus_v_step_ds");
}

// CRT_Base_IF_2::compute_rabi_integrals (CRT_Base_IF_2.h:140-243): list 1 on
// component 0, list 2 on component no_int_states/2; psi is left as it was
template <class T, int dim, int no_int_states>
void CRT_Base_IF_2<T, dim, no_int_states>::compute_rabi_integrals() {
  const int n = (int)m_rabi_momentum_list.size(), m = (int)m_rabi_momentum_list2.size();
  if (n == 0 || m == 0) {
    std::cerr << "WARNING: Rabi momentum list is empty. Cannot compute rabi frquencies." << std::endl;
    return;
  }
  auto integrals = [&](int comp, const std::vector<CPoint<dim>>& list) {
    const int q = (int)list.size();
    std::vector<double> k0(3 * (size_t)q, 0.0), out(q);
    for (int i = 0; i < q; ++i)
      for (int a = 0; a < dim; ++a) k0[3 * i + a] = list[i][a];
    atus_host::check(atus_rabi_integrals(m_set->device_readonly(), comp, k0.data(), q, m_rabi_threshold, out.data()),
                     "atus_rabi_integrals");
    return std::list<double>(out.begin(), out.end());
  };
  m_rabi_freq_list.push_back(integrals(0, m_rabi_momentum_list));
  m_rabi_freq_list2.push_back(integrals(no_int_states / 2, m_rabi_momentum_list2));
}

// CRT_Base_IF_2::Output_rabi_freq_list (CRT_Base_IF_2.h:245-292)
template <class T, int dim, int no_int_states>
void CRT_Base_IF_2<T, dim, no_int_states>::Output_rabi_freq_list(std::string filename, std::string filename2,
                                                                 const long long Nk) {
  auto write = [&](const std::string& name, const std::vector<CPoint<dim>>& klist,
                   const std::list<std::list<double>>& rows) {
    std::ofstream txtfile(name);
    double t = 0;
    txtfile << "# time \t";
    for (auto& i : klist) txtfile << i << "\t";
    txtfile << std::endl;
    for (auto& i : rows) {
      txtfile << t << "\t";
      for (auto j : i) txtfile << j << "\t";
      txtfile << std::endl;
      t += m_header.dt * double(Nk);
    }
  };
  write(filename, m_rabi_momentum_list, m_rabi_freq_list);
  write(filename2, m_rabi_momentum_list2, m_rabi_freq_list2);
}

// CRT_Base_IF_2::Output_chirps_list (CRT_Base_IF_2.h:294-316)
template <class T, int dim, int no_int_states>
void CRT_Base_IF_2<T, dim, no_int_states>::Output_chirps_list(std::string filename) {
  std::ofstream txtfile(filename);
  txtfile << "# Step \t Chirp \t ";
  for (auto& i : m_rabi_momentum_list) txtfile << i << "\t";
  txtfile << std::endl;
  for (auto& i : m_chirps_list) {
    for (auto j : i) txtfile << j << "\t";
    txtfile << std::endl;
  }
}

// CRT_Base_IF_2::run_sequence (CRT_Base_IF_2.h:369-593)
template <class T, int dim, int no_int_states>
void CRT_Base_IF_2<T, dim, no_int_states>::run_sequence() {
  StepFunction step_fct = nullptr, half_step_fct = nullptr, full_step_fct = nullptr;
  char filename[1024], filename2[1024];
  std::cout << "FYI: Found " << m_params->m_sequence.size() << " sequences." << std::endl;
  const int nrm = (int)m_rabi_momentum_list.size();
  if (nrm == 0) std::cerr << "WARNING: Rabi momentum list is empty. Cannot compute rabi frquencies." << std::endl;
  try {
    half_step_fct = this->m_map_stepfcts.at("half_step");
    full_step_fct = this->m_map_stepfcts.at("full_step");
  } catch (const std::out_of_range&) {
    throw std::string("Critical Error: Invalid fct ptr to half_step or full_step");
  }
  this->sync_params();
  int seq_counter = 1;
  for (auto seq : m_params->m_sequence) {
    if (this->run_custom_sequence(seq)) continue;
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
    seq.time = this->Get_t();
    std::cout << "FYI: started new sequence " << seq.name << "\n";
    std::cout << "FYI: sequence no : " << seq_counter << "\n";
    std::cout << "FYI: duration    : " << max_duration << "\n";
    std::cout << "FYI: dt          : " << seq.dt << "\n";
    std::cout << "FYI: Na          : " << Na << "\n";
    std::cout << "FYI: Nk          : " << Nk << "\n";
    std::cout << "FYI: Na*Nk*dt    : " << double(Na * Nk) * seq.dt << "\n";
    if (double(Na * Nk) * seq.dt != max_duration) std::cout << "FYI: double(Na*Nk)*seq.dt != max_duration\n";
    if (this->Get_dt() != seq.dt) this->Set_dt(seq.dt);
    auto it = this->m_map_stepfcts.find(seq.name);
    if (it == this->m_map_stepfcts.end()) throw std::string("Critical Error: Invalid squence name " + seq.name);
    step_fct = it->second;
    if (seq.name == "freeprop") seq.no_of_chirps = 1;
    const double backup_t = m_header.t;
    m_chirps_list.clear();
    for (int s = 0; s < seq.no_of_chirps; ++s) {
      m_rabi_freq_list.clear();
      m_rabi_freq_list2.clear();
      chirp_rate[0] = 0.0;
      for (int i = 1; i <= Na; i++) {
        this->run_outer_block(step_fct, half_step_fct, full_step_fct, seq, Nk);
        std::cout << "t = " << std::to_string(m_header.t) << std::endl;
        if (seq.output_freq == freq::each)
          for (int k = 0; k < no_int_states; k++) {
            snprintf(filename, sizeof(filename), "%.3f_%d.bin", this->Get_t(), k + 1);
            this->Save_Phi(filename, k);
          }
        if (seq.compute_pn_freq == freq::each)
          for (int c = 0; c < no_int_states; c++)
            std::cout << "N[" << c << "] = " << this->Get_Particle_Number(c) << std::endl;
        if (seq.rabi_output_freq == freq::each) compute_rabi_integrals();
        if (seq.custom_freq == freq::each && m_custom_fct != nullptr) (*m_custom_fct)(this, seq);
      }
      if (seq.output_freq == freq::last)
        for (int k = 0; k < no_int_states; k++) {
          snprintf(filename, sizeof(filename), "%.3f_%d.bin", this->Get_t(), k + 1);
          this->Save_Phi(filename, k);
        }
      if (seq.compute_pn_freq == freq::last)
        for (int c = 0; c < no_int_states; c++)
          std::cout << "N[" << c << "] = " << this->Get_Particle_Number(c) << std::endl;
      if (seq.custom_freq == freq::last && m_custom_fct != nullptr) (*m_custom_fct)(this, seq);
      if (seq.rabi_output_freq == freq::last) {
        compute_rabi_integrals();
        std::cout << "Spezie 1" << std::endl;
        std::cout << "Rabi output for timestep ";
        for (auto& i : m_rabi_freq_list) {
          std::cout << this->Get_t() << "\n";
          for (auto j : i) std::cout << j << "\t";
          std::cout << std::endl;
        }
        std::cout << "Spezie 2" << std::endl;
        for (auto& i : m_rabi_freq_list2) {
          std::cout << this->Get_t() << "\n";
          for (auto j : i) std::cout << j << "\t";
          std::cout << std::endl;
        }
      }
      if (seq.rabi_output_freq == freq::each) {
        snprintf(filename, sizeof(filename), "Spezie_%d_Rabi_%d_%d.txt", 1, seq_counter, s);
        snprintf(filename2, sizeof(filename2), "Spezie_%d_Rabi_%d_%d.txt", 2, seq_counter, s);
        Output_rabi_freq_list(filename, filename2, seq.Nk);
      }
      if (seq.no_of_chirps > 1 && nrm > 0) {
        compute_rabi_integrals();
        std::list<double> tmp = m_rabi_freq_list.back();
        tmp.push_front(chirp_rate[0]);
        tmp.push_front(s);
        m_chirps_list.push_back(tmp);
      }
      // the next chirp restarts from the files of the sequence's start time,
      // where an earlier output wrote them (CRT_Base_IF_2.h:571-582); a file
      // that is not there leaves the component as it is
      if (seq.no_of_chirps > 1 && s < seq.no_of_chirps - 1) {
        for (int k = 0; k < no_int_states; k++) {
          snprintf(filename, sizeof(filename), "%.3f_%d.bin", backup_t, k + 1);
          std::ifstream file1(filename, std::ifstream::binary);
          if (!file1.is_open()) continue;
          file1.seekg(sizeof(generic_header), std::ifstream::beg);
          file1.read((char*)m_fields[k]->Getp2In(), sizeof(fftw_complex) * m_no_of_pts);
        }
        m_header.t = backup_t;
      }
    }
    if (seq.no_of_chirps > 1) {
      snprintf(filename, sizeof(filename), "Chirp_%d.txt", seq_counter);
      Output_chirps_list(filename);
    }
    seq_counter++;
  }
}

#endif
