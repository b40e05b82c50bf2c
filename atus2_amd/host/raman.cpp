// raman -- Raman transitions of three states (the reference's raman,
// raman.cpp:42-155): CRT_Base_IF<T,dim,3> whose "raman" step (states 1 and
// 2 coupled to state 3, CRT_Base_IF.h:534-632) runs on the device, fused
// into the split-step blocks.  Rabi momentum list {0, +2 laser_k,
// -2 laser_k} along the first axis.
//
// Errors print their message and exit with status 1 (the reference exits 0).
#include <cstdlib>
#include <iostream>

#include "CRT_Base_IF.h"
#include "driver_util.h"

namespace RT_Solver {

template <class T, int dim>
class Raman_single : public CRT_Base_IF<T, dim, 3> {
 public:
  explicit Raman_single(ParameterHandler* p) : CRT_Base_IF<T, dim, 3>(p) {
    CPoint<dim> pt1, pt2, pt3;
    pt2[0] = 2 * this->laser_k[0];
    pt3[0] = -2 * this->laser_k[0];
    this->m_rabi_momentum_list = {pt1, pt2, pt3};
  }

 protected:
  bool run_custom_sequence(const sequence_item&) override { return false; }
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
      RT_Solver::Raman_single<Fourier::cft_1d, 1> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 2) {
      RT_Solver::Raman_single<Fourier::cft_2d, 2> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 3) {
      RT_Solver::Raman_single<Fourier::cft_3d, 3> rtsol(&params);
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
