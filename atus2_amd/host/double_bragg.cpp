// double_bragg -- symmetric double Bragg beam splitter of three states (the
// reference's double_bragg, double_bragg.cpp:42-329): CRT_Base_IF<T,dim,3>
// with "bragg_ad" = Do_Double_Bragg_ad, state 1 coupled to states 2 and 3 by
// counter-propagating lasers (+laser_k and -laser_k, chirps of opposite
// sign), the reference's closed-form 3x3 exponential on the device, fused
// into the split-step blocks.  Rabi momentum list {0, +2 laser_k,
// -2 laser_k} along the first axis.
//
// The step prints the amplitude F of every V step, clamped at 0 as the
// reference does (double_bragg.cpp:117-120); the reference prints it once
// per OpenMP thread, this build once.  Errors print their message and exit
// with status 1 (the reference exits 0).
#include <cstdlib>
#include <iostream>

#include "CRT_Base_IF.h"
#include "driver_util.h"

namespace RT_Solver {

template <class T, int dim>
class Bragg_double : public CRT_Base_IF<T, dim, 3> {
 public:
  explicit Bragg_double(ParameterHandler* p) : CRT_Base_IF<T, dim, 3>(p) {
    this->m_map_stepfcts["bragg_ad"] = &Do_Double_Bragg_ad_Wrapper;
    this->m_device_steps[&Do_Double_Bragg_ad_Wrapper] = ATUS_STEP_DOUBLE_BRAGG;
    CPoint<dim> pt1, pt2, pt3;
    pt2[0] = 2 * this->laser_k[0];
    pt3[0] = -2 * this->laser_k[0];
    this->m_rabi_momentum_list = {pt1, pt2, pt3};
  }

 protected:
  bool m_print_amplitude = false;

  bool run_custom_sequence(const sequence_item&) override { return false; }

  static void Do_Double_Bragg_ad_Wrapper(void* ptr, sequence_item&) {
    static_cast<Bragg_double*>(ptr)->Do_Double_Bragg_ad();
  }

  // Amplitude_at_time(), clamped at 0 and printed (double_bragg.cpp:116-120)
  double step_amplitude(double t1) {
    double F = this->amp_is_t ? CRT_Base_IF<T, dim, 3>::amplitude_at(t1) : 1.0;
    if (F < 0) F = 0;
    std::cout << F << std::endl;
    return F;
  }

  // Bragg_double::Do_Double_Bragg_ad (double_bragg.cpp:108-249), one step
  void Do_Double_Bragg_ad() {
    const atus_laser L = this->laser_params();
    const double F = step_amplitude(this->Get_t() + 0.5 * this->Get_dt());
    atus_host::check(atus_v_step(this->m_set->device(), ATUS_STEP_DOUBLE_BRAGG, this->Get_dt(), this->Get_t(), F, &L),
                     "atus_v_step");
  }

  // fused blocks of the double Bragg step evaluate (and print) the
  // amplitude of every V step through the callback
  void run_device_block(StepFunction step, sequence_item& seq, int Nk) override {
    m_print_amplitude = this->m_device_steps.at(step) == ATUS_STEP_DOUBLE_BRAGG;
    CRT_Base_IF<T, dim, 3>::run_device_block(step, seq, Nk);
    m_print_amplitude = false;
  }
  bool has_amplitude() const override { return this->amp_is_t || m_print_amplitude; }
  double amplitude_at(double t1) override {
    if (m_print_amplitude) return step_amplitude(t1);
    return CRT_Base_IF<T, dim, 3>::amplitude_at(t1);
  }
};

}  // namespace RT_Solver

int main(int argc, char* argv[]) {
  if (argc != 2) {
    printf("No parameter xml file specified.\n");
    return EXIT_FAILURE;
  }
  try {
    ParameterHandler params(argv[1]);
    const int dim = std::stoi(params.Get_simulation("DIM"));
    atus_host::set_threads_from_env();
    if (dim == 1) {
      RT_Solver::Bragg_double<Fourier::cft_1d, 1> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 2) {
      RT_Solver::Bragg_double<Fourier::cft_2d, 2> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 3) {
      RT_Solver::Bragg_double<Fourier::cft_3d, 3> rtsol(&params);
      rtsol.run_sequence();
    } else {
      std::cout << "You have found a new dimension!" << std::endl;
      return EXIT_FAILURE;
    }
  } catch (std::string& str) {
    std::cout << str << std::endl;
    return EXIT_FAILURE;
  } catch (std::exception& e) {
    std::cout << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
