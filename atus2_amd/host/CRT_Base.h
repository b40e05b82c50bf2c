// CRT_Base.h -- N-component real-time split-step propagator (the reference's
// CRT_Base<T,dim,no_int_states>, CRT_Base.h:47-861) on libatus_hip.so.
//
// Same public API, step registry and run_sequence semantics.  The wave
// functions live on the device; when the sequence's step function is one of
// the registered device steps and the kinetic steps are the built-in ones,
// run_sequence executes each outer iteration (half, (V, full) x (Nk-1), V,
// half) as one fused device block (atus_run_block).  Any other registered
// step function is called per step as in the reference, on the host mirror
// when it touches Getp2In().
//
// Differences from the reference, all deliberate:
//   * the kinetic phase is computed from the current dt at every step; the
//     reference keeps tables built for the header's dt until the first
//     Set_dt (CRT_Base.h:149,174,757-758), so a file dt different from
//     ALGORITHM/DT with an equal sequence dt would propagate with the file dt
//     there;
//   * errors thrown as std::string propagate to main() as in the reference;
//     unknown sequence names also throw instead of calling exit().
#ifndef ATUS_AMD_CRT_BASE_H
#define ATUS_AMD_CRT_BASE_H

#include <array>
#include <cstdio>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "CRT_shared.h"

template <class T, int dim, int no_int_states>
class CRT_Base : public CRT_shared {
 public:
  explicit CRT_Base(ParameterHandler*);
  virtual ~CRT_Base();

  double Get_Particle_Number(const int comp = 0);
  void run_sequence();

  void Setup_Momentum(CPoint<dim>, const int comp = 0);
  void Expval_Position(CPoint<dim>&, const int comp = 0);
  void Expval_Momentum(CPoint<dim>&, const int comp = 0);
  void Save(fftw_complex*, std::string);
  void Save_Phi(std::string, const int comp = 0);
  void Append_Phi(std::string, const int comp = 0);

  void Set_custom_fct(StepFunction& fct) { m_custom_fct = fct; }

  void Init_Potential();
  void Setup_Potential(const int, const int, const double);
  // energy functional on the device (no reference counterpart)
  double Get_Energy();

 protected:
  static void Do_FT_Step_full_Wrapper(void*, sequence_item&);
  static void Do_FT_Step_half_Wrapper(void*, sequence_item&);
  static void Do_NL_Step_Wrapper(void*, sequence_item&);
  static void Do_NL_Step_Wrapper_one(void*, sequence_item&);

  void Do_FT_Step_full();
  void Do_FT_Step_half();
  void Do_NL_Step();

  ParameterHandler* m_params;
  CPoint<dim> m_alpha;

  void Init() override;
  void Allocate();
  void LoadFiles();

  bool m_potenial_initialized;
  virtual bool run_custom_sequence(const sequence_item&) = 0;

  std::array<double, no_int_states * no_int_states> m_gs;
  std::array<T*, no_int_states> m_fields;
  std::array<std::vector<double>, no_int_states> m_Potential;

  std::map<std::string, StepFunction> m_map_stepfcts;
  StepFunction m_custom_fct;
  // writes per file name by Save / Save_Phi / Append_Phi: a restart file
  // whose count has not changed since it was written still holds that state
  std::map<std::string, long long> m_file_writes;

  // --- device side
  std::shared_ptr<atus_host::FieldSet> m_set;
  // step functions that have a fused device block, with their kind
  std::map<StepFunction, int> m_device_steps;
  // device parameters of the step kinds (IF layer fills the laser part)
  virtual void fill_step_params(atus_step_params& sp, const sequence_item& seq);
  // uploads g, alpha and (if initialised) the potentials
  void sync_params();
  bool device_block_possible(StepFunction step) const;
  virtual void run_device_block(StepFunction step, sequence_item& seq, int Nk);
  static double amp_callback(double t1, void* self);
  virtual double amplitude_at(double t1) { return 1.0; }
  virtual bool has_amplitude() const { return false; }
};

template <class T, int dim, int no_int_states>
CRT_Base<T, dim, no_int_states>::CRT_Base(ParameterHandler* params) {
  m_params = params;
  Read_header(params->Get_simulation("FILENAME"), dim);
  if (m_header.nDims != dim)
    throw std::string("Critical error: header.nDims of " + params->Get_simulation("FILENAME") +
                      " does not match the solver dimension");
  Allocate();
  LoadFiles();
  Init();

  m_map_stepfcts["half_step"] = &Do_FT_Step_half_Wrapper;
  m_map_stepfcts["full_step"] = &Do_FT_Step_full_Wrapper;
  m_map_stepfcts["freeprop"] = &Do_NL_Step_Wrapper;
  m_map_stepfcts["freeprop_lin"] = &Do_NL_Step_Wrapper_one;
  m_device_steps[&Do_NL_Step_Wrapper] = ATUS_STEP_FREEPROP;
  m_custom_fct = nullptr;
  m_potenial_initialized = false;

  for (int i = 0; i < no_int_states; i++) {
    std::string tmpstr = "GS_" + std::to_string(i + 1);
    for (int j = 0; j < no_int_states; j++) m_gs[j + no_int_states * i] = m_params->Get_VConstant(tmpstr, j);
  }
  for (int i = 0; i < dim; i++) m_alpha[i] = m_params->Get_VConstant("Alpha_1", i);
  m_header.dt = params->Get_dt();
  sync_params();
}

template <class T, int dim, int no_int_states>
CRT_Base<T, dim, no_int_states>::~CRT_Base() {
  for (int i = 0; i < no_int_states; i++) delete m_fields[i];
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Allocate() {
  m_set = std::make_shared<atus_host::FieldSet>(m_header, no_int_states);
  for (int i = 0; i < no_int_states; i++) {
    m_fields[i] = new T(m_set, i, m_header);
    m_fields[i]->SetFix(false);
  }
}

// CRT_Base::LoadFiles (CRT_Base.h:207-240): FILENAME, FILENAME_2, ...
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::LoadFiles() {
  for (int i = 0; i < no_int_states; i++) {
    std::string key = i == 0 ? std::string("FILENAME") : "FILENAME_" + std::to_string(i + 1);
    std::string name = m_params->Get_simulation(key);
    std::ifstream in(name, std::ifstream::binary);
    if (!in.is_open()) throw std::string("Could not open file " + name + "\n");
    in.seekg(sizeof(generic_header), std::ifstream::beg);
    in.read((char*)m_fields[i]->Getp2In(), sizeof(fftw_complex) * m_no_of_pts);
    if (!in) throw std::string("File " + name + " holds fewer points than its header announces\n");
  }
}

// CRT_Base::Init (CRT_Base.h:253-274): the kinetic phase is evaluated on the
// device from alpha and the dt of each call, so only alpha is pushed here
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Init() {
  if (!m_set) return;
  double al[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) al[a] = m_alpha[a];
  atus_host::check(atus_set_alpha(m_set->device_readonly(), al), "atus_set_alpha");
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::sync_params() {
  atus_ctx* c = m_set->device_readonly();
  atus_host::check(atus_set_g(c, m_gs.data()), "atus_set_g");
  Init();
  if (m_potenial_initialized) {
    for (int i = 0; i < no_int_states; ++i)
      atus_host::check(atus_set_potential(c, i, m_Potential[i].data()), "atus_set_potential");
  } else {
    atus_host::check(atus_set_potential(c, 0, nullptr), "atus_set_potential");
  }
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Init_Potential() {
  for (auto& it : m_Potential) it.assign(m_no_of_pts, 0.0);
  m_potenial_initialized = true;
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Setup_Potential(const int comp, const int index, const double val) {
  if (!m_potenial_initialized) {
    std::cerr << "Potential not initialized. You forgot to invoke Init_Potential." << std::endl;
    throw std::string("Potential not initialized. You forgot to invoke Init_Potential.");
  }
  if (comp < 0 || comp >= no_int_states || index < 0 || index >= m_no_of_pts)
    throw std::string("Error in Setup_Potential: index out of bounds\n");
  m_Potential[comp][index] = val;
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Do_FT_Step_full_Wrapper(void* ptr, sequence_item&) {
  static_cast<CRT_Base*>(ptr)->Do_FT_Step_full();
}
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Do_FT_Step_half_Wrapper(void* ptr, sequence_item&) {
  static_cast<CRT_Base*>(ptr)->Do_FT_Step_half();
}
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Do_NL_Step_Wrapper(void* ptr, sequence_item&) {
  static_cast<CRT_Base*>(ptr)->Do_NL_Step();
}
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Do_NL_Step_Wrapper_one(void*, sequence_item&) {}

// Do_FT_Step_full / half (CRT_Base.h:337-392)
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Do_FT_Step_full() {
  atus_host::check(atus_kinetic(m_set->device(), 1.0, m_header.dt), "atus_kinetic");
  m_header.t += m_header.dt;
}
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Do_FT_Step_half() {
  atus_host::check(atus_kinetic(m_set->device(), 0.5, m_header.dt), "atus_kinetic");
  m_header.t += 0.5 * m_header.dt;
}
// CRT_Base::Do_NL_Step (CRT_Base.h:397-441)
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Do_NL_Step() {
  atus_host::check(atus_nl_step(m_set->device(), m_header.dt), "atus_nl_step");
}

// CRT_Base::Setup_Momentum (CRT_Base.h:449-473)
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Setup_Momentum(CPoint<dim> px, const int comp) {
  if (comp < 0 || comp >= no_int_states)
    throw std::string("Error in " + std::string(__func__) + ": comp out of bounds\n");
  double p[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) p[a] = px[a];
  atus_host::check(atus_set_momentum(m_set->device(), comp, p), "atus_set_momentum");
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Expval_Position(CPoint<dim>& retval, const int comp) {
  if (comp < 0 || comp >= no_int_states)
    throw std::string("Error in " + std::string(__func__) + ": comp out of bounds\n");
  double o[3];
  atus_host::check(atus_expval_x(m_set->device_readonly(), comp, o), "atus_expval_x");
  for (int a = 0; a < dim; ++a) retval[a] = o[a];
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Expval_Momentum(CPoint<dim>& retval, const int comp) {
  if (comp < 0 || comp >= no_int_states)
    throw std::string("Error in " + std::string(__func__) + ": comp out of bounds\n");
  double o[3];
  atus_host::check(atus_expval_k(m_set->device_readonly(), comp, o), "atus_expval_k");
  for (int a = 0; a < dim; ++a) retval[a] = o[a];
}

template <class T, int dim, int no_int_states>
double CRT_Base<T, dim, no_int_states>::Get_Particle_Number(const int comp) {
  if (comp < 0 || comp >= no_int_states)
    throw std::string("Error in " + std::string(__func__) + ": comp out of bounds\n");
  double n = 0;
  atus_host::check(atus_particle_number(m_set->device_readonly(), comp, &n), "atus_particle_number");
  return n;
}

template <class T, int dim, int no_int_states>
double CRT_Base<T, dim, no_int_states>::Get_Energy() {
  double e = 0;
  atus_host::check(atus_energy(m_set->device_readonly(), &e), "atus_energy");
  return e;
}

// Save_Phi (CRT_Base.h:607-618)
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Save_Phi(std::string filename, const int comp) {
  if (comp < 0 || comp >= no_int_states)
    throw std::string("Error in " + std::string(__func__) + ": comp out of bounds\n");
  ++m_file_writes[filename];
  std::ofstream f(filename, std::ofstream::binary);
  f.write(reinterpret_cast<char*>(&m_header), sizeof(generic_header));
  f.write(reinterpret_cast<const char*>(m_set->host_view(comp)), (size_t)m_no_of_pts * sizeof(fftw_complex));
}

// Append_Phi (CRT_Base.h:626-641)
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Append_Phi(std::string filename, const int comp) {
  if (comp < 0 || comp >= no_int_states)
    throw std::string("Error in " + std::string(__func__) + ": comp out of bounds\n");
  ++m_file_writes[filename];
  std::ofstream f(filename, std::ofstream::binary | std::ofstream::app);
  if (f.fail()) throw std::string("File " + filename + " could not be opened.");
  f.write(reinterpret_cast<char*>(&m_header), sizeof(generic_header));
  f.write(reinterpret_cast<const char*>(m_set->host_view(comp)), (size_t)m_no_of_pts * sizeof(fftw_complex));
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::Save(fftw_complex* data, std::string filename) {
  ++m_file_writes[filename];
  std::ofstream f(filename, std::ofstream::binary);
  f.write(reinterpret_cast<char*>(&m_header), sizeof(generic_header));
  f.write(reinterpret_cast<char*>(data), (size_t)m_no_of_pts * sizeof(fftw_complex));
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::fill_step_params(atus_step_params& sp, const sequence_item&) {
  sp.laser = atus_laser{};
  sp.amp = nullptr;
  sp.amp_user = nullptr;
}

template <class T, int dim, int no_int_states>
double CRT_Base<T, dim, no_int_states>::amp_callback(double t1, void* self) {
  return static_cast<CRT_Base*>(self)->amplitude_at(t1);
}

template <class T, int dim, int no_int_states>
bool CRT_Base<T, dim, no_int_states>::device_block_possible(StepFunction step) const {
  auto h = m_map_stepfcts.find("half_step");
  auto f = m_map_stepfcts.find("full_step");
  return m_device_steps.count(step) && h != m_map_stepfcts.end() && f != m_map_stepfcts.end() &&
         h->second == &Do_FT_Step_half_Wrapper && f->second == &Do_FT_Step_full_Wrapper;
}

template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::run_device_block(StepFunction step, sequence_item& seq, int Nk) {
  atus_step_params sp{};
  fill_step_params(sp, seq);
  sp.kind = m_device_steps.at(step);
  sp.dt = m_header.dt;
  if (has_amplitude()) {
    sp.amp = &CRT_Base::amp_callback;
    sp.amp_user = this;
  }
  atus_host::check(atus_run_block(m_set->device(), &sp, Nk, &m_header.t), "atus_run_block");
}

// CRT_Base::run_sequence (CRT_Base.h:685-841)
template <class T, int dim, int no_int_states>
void CRT_Base<T, dim, no_int_states>::run_sequence() {
  StepFunction step_fct = nullptr, half_step_fct = nullptr, full_step_fct = nullptr;
  char filename[1024];
  std::cout << "FYI: Found " << m_params->m_sequence.size() << " sequences." << std::endl;
  try {
    half_step_fct = m_map_stepfcts.at("half_step");
    full_step_fct = m_map_stepfcts.at("full_step");
  } catch (const std::out_of_range& oor) {
    throw std::string("Critical Error: Invalid fct ptr to ft_half_step or ft_full_step");
  }
  sync_params();
  int seq_counter = 1;
  for (auto seq : m_params->m_sequence) {
    if (run_custom_sequence(seq)) continue;
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
      // the reference asserts vec.size() > dim and seq.comp <= dim (CRT_Base.h:724-725)
      if (!((int)vec.size() > dim)) throw std::string("set_momentum needs more than dim comma-separated values");
      if (!(seq.comp <= dim)) throw std::string("set_momentum: component index out of range");
      CPoint<dim> P;
      for (int i = 0; i < dim; i++) P[i] = std::stod(vec[i]);
      Setup_Momentum(P, seq.comp);
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
    if (double(Na * Nk) * seq.dt != max_duration) std::cout << "FYI: double(Na*Nk)*seq.dt != max_duration\n";
    if (Get_dt() != seq.dt) Set_dt(seq.dt);
    auto it = m_map_stepfcts.find(seq.name);
    if (it == m_map_stepfcts.end()) throw std::string("Critical Error: Invalid squence name " + seq.name);
    step_fct = it->second;
    for (int k = 0; k < no_int_states; k++) {
      snprintf(filename, sizeof(filename), "Seq_%d_%d.bin", seq_counter, k + 1);
      std::remove(filename);
    }
    const bool fused = device_block_possible(step_fct);
    for (int i = 1; i <= Na; i++) {
      if (fused) {
        run_device_block(step_fct, seq, Nk);
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
      if (seq.output_freq == freq::each)
        for (int k = 0; k < no_int_states; k++) {
          snprintf(filename, sizeof(filename), "%.3f_%d.bin", Get_t(), k + 1);
          Save_Phi(filename, k);
        }
      if (seq.output_freq == freq::packed)
        for (int k = 0; k < no_int_states; k++) {
          snprintf(filename, sizeof(filename), "Seq_%d_%d.bin", seq_counter, k + 1);
          Append_Phi(filename, k);
        }
      if (seq.compute_pn_freq == freq::each)
        for (int c = 0; c < no_int_states; c++)
          std::cout << "N[" << c << "] = " << Get_Particle_Number(c) << std::endl;
      if (seq.custom_freq == freq::each && m_custom_fct != nullptr) (*m_custom_fct)(this, seq);
    }
    if (seq.output_freq == freq::last)
      for (int k = 0; k < no_int_states; k++) {
        snprintf(filename, sizeof(filename), "%.3f_%d.bin", Get_t(), k + 1);
        Save_Phi(filename, k);
      }
    if (seq.compute_pn_freq == freq::last)
      for (int c = 0; c < no_int_states; c++) std::cout << "N[" << c << "] = " << Get_Particle_Number(c) << std::endl;
    if (seq.custom_freq == freq::last && m_custom_fct != nullptr) (*m_custom_fct)(this, seq);
    seq_counter++;
  }
}

#endif
