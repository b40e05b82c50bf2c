// sobmin_2 -- two-component ground state by Sobolev-gradient minimisation
// (the reference's sob_solver_2, sob_solver_2.cpp:37-253): CSOB_Base<T,dim,2>
// with the real guesses GUESS_nD / GUESS_nD_2 and the potentials
// POTENTIAL_nD / POTENTIAL_nD_2; after run() the states are written, not
// rescaled, to FILENAME and FILENAME_3 (sob_solver_2.cpp:202-219), plus
// zero.bin.
//
// Optional environment: ATUS_SOB_MAXITER caps the iterations (the reference
// has no cap).  Errors print their message and exit with status 1.
#include <cstdlib>
#include <iostream>

#include "CSOB_Base.h"
#include "driver_util.h"

namespace SOB_Solver {

template <class T, int dim>
class CSOB_Min_2 : public CSOB_Base<T, dim, 2> {
 public:
  explicit CSOB_Min_2(ParameterHandler* p) : CSOB_Base<T, dim, 2>(p) {
    Setup_Guess();
    Setup_Potential();
  }

  // CSOB_Min_2::Setup_Potential (sob_solver_2.cpp:61-111)
  void Setup_Potential() {
    const std::string d = std::to_string(dim);
    for (int c = 0; c < 2; ++c) {
      const std::string expr = this->m_params->Get_simulation("POTENTIAL_" + d + "D" + (c ? "_2" : ""));
      atus_host::eval_on_grid<dim>(*this->m_params, expr, this->m_no_of_pts,
                                   [this](long long l) { return this->m_fields[0]->Get_x(l); }, this->m_Potential[c]);
    }
  }
  // CSOB_Min_2::Setup_Guess (sob_solver_2.cpp:113-167): real guesses
  void Setup_Guess() {
    const std::string d = std::to_string(dim);
    std::vector<double> g(this->m_no_of_pts);
    for (int c = 0; c < 2; ++c) {
      const std::string expr = this->m_params->Get_simulation("GUESS_" + d + "D" + (c ? "_2" : ""));
      atus_host::eval_on_grid<dim>(*this->m_params, expr, this->m_no_of_pts,
                                   [this](long long l) { return this->m_fields[0]->Get_x(l); }, g.data());
      fftw_complex* psi = this->m_fields[c]->Getp2In();
      for (int l = 0; l < this->m_no_of_pts; ++l) {
        psi[l][0] = g[l];
        psi[l][1] = 0;
      }
    }
  }
};

}  // namespace SOB_Solver

template <class S>
static void solve(ParameterHandler& params) {
  S sol(&params);
  if (const char* m = std::getenv("ATUS_SOB_MAXITER")) sol.Set_max_iterations(std::atol(m));
  sol.run();
  std::vector<std::string> filenames{params.Get_simulation("FILENAME"), params.Get_simulation("FILENAME_3")};
  sol.Save(filenames, false);
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
    if (dim == 1) solve<SOB_Solver::CSOB_Min_2<Fourier::cft_1d, 1>>(params);
    else if (dim == 2) solve<SOB_Solver::CSOB_Min_2<Fourier::cft_2d, 2>>(params);
    else if (dim == 3) solve<SOB_Solver::CSOB_Min_2<Fourier::cft_3d, 3>>(params);
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
