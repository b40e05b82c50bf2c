// CRT_Base_2.h -- two-species real-time propagator (the reference's
// CRT_Base_2<T,dim,no_int_states>, CRT_Base_2.h:39-735) on libatus_hip.so.
//
// Components [0, no_int_states/2) are species 1 with kinetic coefficients
// Alpha_1, components [no_int_states/2, no_int_states) species 2 with
// Alpha_2.  The device keeps one kinetic table set per species
// (atus_set_alpha2) and picks the set per component, so the built-in
// half_step / full_step are the species-aware steps of CRT_Base_2.h:282-347
// and run_sequence's fused blocks use them too.
//
// Shares CRT_Base's device plumbing (field set, step registry, Save_Phi,
// reductions).  run_sequence follows CRT_Base_2::run_sequence
// (CRT_Base_2.h:526-661): sequence_item::time is set to Get_t() at the start
// of each sequence and there is no packed output.
#ifndef ATUS_AMD_CRT_BASE_2_H
#define ATUS_AMD_CRT_BASE_2_H

#include "CRT_Base.h"

template <class T, int dim, int no_int_states>
class CRT_Base_2 : public CRT_Base<T, dim, no_int_states> {
  static_assert(no_int_states >= 2 && no_int_states % 2 == 0, "two species need an even number of components");

 public:
  explicit CRT_Base_2(ParameterHandler*);
  virtual ~CRT_Base_2() = default;
  void run_sequence();

 protected:
  using CRT_Base<T, dim, no_int_states>::m_header;
  using CRT_Base<T, dim, no_int_states>::m_params;
  using CRT_Base<T, dim, no_int_states>::m_set;
  using CRT_Base<T, dim, no_int_states>::m_custom_fct;

  CPoint<dim> m_alpha2;
  bool m_alpha2_read = false;

  // CRT_Base_2::Init (CRT_Base_2.h:195-221): both species' kinetic
  // coefficients; the phases are evaluated on the device per dt
  void Init() override;
  // one outer block (half, (V, full) x (Nk-1), V, half) of the sequence's
  // step on the device, or the registered host steps one by one
  void run_outer_block(StepFunction step, StepFunction half, StepFunction full, sequence_item& seq, int Nk);
  virtual bool device_block(StepFunction step, sequence_item& seq, int Nk);
};

// CRT_Base_2::CRT_Base_2 (CRT_Base_2.h:96-130)
template <class T, int dim, int no_int_states>
CRT_Base_2<T, dim, no_int_states>::CRT_Base_2(ParameterHandler* params) : CRT_Base<T, dim, no_int_states>(params) {
  for (int i = 0; i < dim; i++) m_alpha2[i] = m_params->Get_VConstant("Alpha_2", i);
  m_alpha2_read = true;
  Init();
}

template <class T, int dim, int no_int_states>
void CRT_Base_2<T, dim, no_int_states>::Init() {
  CRT_Base<T, dim, no_int_states>::Init();
  if (!m_set || !m_alpha2_read) return;
  double al[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) al[a] = m_alpha2[a];
  atus_host::check(atus_set_alpha2(m_set->device_readonly(), al), "atus_set_alpha2");
}

template <class T, int dim, int no_int_states>
bool CRT_Base_2<T, dim, no_int_states>::device_block(StepFunction step, sequence_item& seq, int Nk) {
  if (!this->device_block_possible(step)) return false;
  this->run_device_block(step, seq, Nk);
  return true;
}

template <class T, int dim, int no_int_states>
void CRT_Base_2<T, dim, no_int_states>::run_outer_block(StepFunction step_fct, StepFunction half_step_fct,
                                                        StepFunction full_step_fct, sequence_item& seq, int Nk) {
  if (device_block(step_fct, seq, Nk)) return;
  (*half_step_fct)(this, seq);
  for (int j = 2; j <= Nk; j++) {
    (*step_fct)(this, seq);
    (*full_step_fct)(this, seq);
  }
  (*step_fct)(this, seq);
  (*half_step_fct)(this, seq);
}

// CRT_Base_2::run_sequence (CRT_Base_2.h:526-661)
template <class T, int dim, int no_int_states>
void CRT_Base_2<T, dim, no_int_states>::run_sequence() {
  StepFunction step_fct = nullptr, half_step_fct = nullptr, full_step_fct = nullptr;
  char filename[1024];
  std::cout << "FYI: Found " << m_params->m_sequence.size() << " sequences." << std::endl;
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
    for (int i = 1; i <= Na; i++) {
      run_outer_block(step_fct, half_step_fct, full_step_fct, seq, Nk);
      std::cout << "t = " << std::to_string(m_header.t) << std::endl;
      if (seq.output_freq == freq::each)
        for (int k = 0; k < no_int_states; k++) {
          snprintf(filename, sizeof(filename), "%.3f_%d.bin", this->Get_t(), k + 1);
          this->Save_Phi(filename, k);
        }
      if (seq.compute_pn_freq == freq::each)
        for (int c = 0; c < no_int_states; c++)
          std::cout << "N[" << c << "] = " << this->Get_Particle_Number(c) << std::endl;
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
    seq_counter++;
  }
}

#endif
