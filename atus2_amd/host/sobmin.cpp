// sobmin -- ground state by Sobolev-gradient minimisation (the reference's
// sob_solver, sob_solver.cpp:41-232): CSOB_Base<T,dim,1> with the real
// initial guess GUESS_nD and the potential POTENTIAL_nD; after run() the
// state is rescaled to N and written to FILENAME, plus zero.bin.
//
// Optional environment: ATUS_SOB_MAXITER caps the iterations (the reference
// has no cap).  Errors print their message and exit with status 1.
#include <cstdlib>
#include <iostream>

#include "CSOB_Base.h"
#include "driver_util.h"

namespace SOB_Solver {

template <class T, int dim>
class CSOB_Min : public CSOB_Base<T, dim, 1> {
 public:
  explicit CSOB_Min(ParameterHandler* p) : CSOB_Base<T, dim, 1>(p) {
    Setup_Guess();
    Setup_Potential();
  }

  // CSOB_Min::Setup_Potential (sob_solver.cpp:64-104)
  void Setup_Potential() {
    const std::string expr = this->m_params->Get_simulation("POTENTIAL_" + std::to_string(dim) + "D");
    atus_host::eval_on_grid<dim>(*this->m_params, expr, this->m_no_of_pts,
                                 [this](long long l) { return this->m_fields[0]->Get_x(l); }, this->m_Potential[0]);
  }
  // CSOB_Min::Setup_Guess (sob_solver.cpp:106-150): real guess
  void Setup_Guess() {
    const std::string expr = this->m_params->Get_simulation("GUESS_" + std::to_string(dim) + "D");
    std::vector<double> g(this->m_no_of_pts);
    atus_host::eval_on_grid<dim>(*this->m_params, expr, this->m_no_of_pts,
                                 [this](long long l) { return this->m_fields[0]->Get_x(l); }, g.data());
    fftw_complex* psi = this->m_fields[0]->Getp2In();
    for (int l = 0; l < this->m_no_of_pts; ++l) {
      psi[l][0] = g[l];
      psi[l][1] = 0;
    }
  }
};

}  // namespace SOB_Solver

template <class S>
static void solve(ParameterHandler& params) {
  S sol(&params);
  if (const char* m = std::getenv("ATUS_SOB_MAXITER")) sol.Set_max_iterations(std::atol(m));
  sol.run();
  std::vector<std::string> filenames{params.Get_simulation("FILENAME")};
  sol.Save(filenames, true);
  sol.Save_Zero();
}

int main(int argc, char* argv[]) {
  if (argc != 2) {
    printf("No parameter xml file specified.\n");
    return EXIT_FAILURE;
  }
  try {
    ParameterHandler params(argv[1]);
    const int dim = std::stoi(params.Get_simulation("DIM"));
    atus_host::set_threads_from_env();
    if (dim == 1) solve<SOB_Solver::CSOB_Min<Fourier::cft_1d, 1>>(params);
    else if (dim == 2) solve<SOB_Solver::CSOB_Min<Fourier::cft_2d, 2>>(params);
    else if (dim == 3) solve<SOB_Solver::CSOB_Min<Fourier::cft_3d, 3>>(params);
    else {
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
