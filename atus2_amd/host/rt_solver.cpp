// rt_solver -- real-time GPE propagation driven by one XML file (the
// reference's rt_solver, rt_solver.cpp:46-204): CRT_Base<T,dim,1> with the
// potential POTENTIAL_nD evaluated on the grid, then run_sequence().
//
// Errors print their message and exit with status 1 (the reference exits 0).
#include <cstdlib>
#include <iostream>

#include "CRT_Base.h"
#include "driver_util.h"

namespace RT_Solver {

template <class T, int dim>
class CRT_Propagation : public CRT_Base<T, dim, 1> {
 public:
  explicit CRT_Propagation(ParameterHandler* p) : CRT_Base<T, dim, 1>(p) {
    this->Init_Potential();
    Setup_Potential();
  }

 protected:
  bool run_custom_sequence(const sequence_item&) override { return false; }

  // CRT_Propagation::Setup_Potential (rt_solver.cpp:74-131)
  void Setup_Potential() {
    const std::string expr = this->m_params->Get_simulation("POTENTIAL_" + std::to_string(dim) + "D");
    atus_host::eval_on_grid<dim>(*this->m_params, expr, this->m_no_of_pts,
                                 [this](long long l) { return this->m_fields[0]->Get_x(l); },
                                 this->m_Potential[0].data());
    if (dim == 2) {
      std::ofstream outf("potential.txt");
      for (int i = 0; i < this->Get_dimX(); i++) {
        double x = double(i - this->m_shift_x) * this->Get_dx();
        for (int j = 0; j < this->Get_dimY(); j++) {
          double y = double(j - this->m_shift_y) * this->Get_dy();
          outf << x << "\t" << y << "\t" << this->m_Potential[0][j + this->Get_dimY() * i] << "\n";
        }
        outf << "\n";
      }
    }
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
      RT_Solver::CRT_Propagation<Fourier::cft_1d, 1> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 2) {
      RT_Solver::CRT_Propagation<Fourier::cft_2d, 2> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 3) {
      RT_Solver::CRT_Propagation<Fourier::cft_3d, 3> rtsol(&params);
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
