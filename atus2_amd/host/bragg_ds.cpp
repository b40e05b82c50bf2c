// bragg_ds -- Bragg beam splitter for two species of two internal states
// each (the reference's bragg_ds, bragg_ds.cpp:42-346): CRT_Base_IF_2<T,dim,4>
// with the analytic 2x2 step "bragg_ad" per species (device, fused into the
// split-step blocks) and the Rabi momentum lists {0, +2 laser_k, -2 laser_k}
// for species 1 and {0, +2 laser_k_2, -2 laser_k_2} for species 2 along the
// first axis.
//
// Components: 0, 1 species 1 (FILENAME, FILENAME_2), 2, 3 species 2
// (FILENAME_3, FILENAME_4).  Pulse s is on while t + dt/2 - (start of the
// sequence) <= duration[s], so <bragg_ad>0.04,0.02</bragg_ad> gives species 2
// a pulse half as long.
//
// Errors print their message and exit with status 1 (the reference exits 0).
#include <cstdlib>
#include <iostream>

#include "CRT_Base_IF_2.h"
#include "driver_util.h"

namespace RT_Solver_2 {

template <class T, int dim>
class Bragg_single : public CRT_Base_IF_2<T, dim, 4> {
 public:
  explicit Bragg_single(ParameterHandler* p) : CRT_Base_IF_2<T, dim, 4>(p) {
    CPoint<dim> pt1, pt2, pt3;
    pt2[0] = 2 * this->laser_k[0];
    pt3[0] = -2 * this->laser_k[0];
    this->m_rabi_momentum_list = {pt1, pt2, pt3};
    pt2[0] = 2 * this->laser_k2[0];
    pt3[0] = -2 * this->laser_k2[0];
    this->m_rabi_momentum_list2 = {pt1, pt2, pt3};
    this->m_map_stepfcts["bragg_ad"] = &Do_Bragg_ad_Wrapper;
    this->m_device_steps[&Do_Bragg_ad_Wrapper] = ATUS_STEP_BRAGG_AD;
  }

 protected:
  bool run_custom_sequence(const sequence_item&) override { return false; }

  static void Do_Bragg_ad_Wrapper(void* ptr, sequence_item& seq) { static_cast<Bragg_single*>(ptr)->Do_Bragg_ad(seq); }

  // Bragg_single::Do_Bragg_ad (bragg_ds.cpp:112-300), one step on the device
  void Do_Bragg_ad(sequence_item& seq) {
    const atus_laser L = this->laser1_params();
    const atus_ds_params ds = this->ds_params(seq);
    atus_host::check(atus_v_step_ds(this->m_set->device(), ATUS_STEP_BRAGG_AD, this->Get_dt(), this->Get_t(), &L, &ds),
                     "atus_v_step_ds");
  }
};

}  // namespace RT_Solver_2

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
      RT_Solver_2::Bragg_single<Fourier::cft_1d, 1> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 2) {
      RT_Solver_2::Bragg_single<Fourier::cft_2d, 2> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 3) {
      RT_Solver_2::Bragg_single<Fourier::cft_3d, 3> rtsol(&params);
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
