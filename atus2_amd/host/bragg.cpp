// bragg -- two-state Bragg beam splitter (the reference's bragg,
// bragg.cpp:48-445): CRT_Base_IF<T,dim,2> with the analytic 2x2 step
// "bragg_ad" (device, fused into the split-step blocks), the host step
// "bragg_simple", the custom sequence "set_mirror" and the Rabi momentum
// list {0, +2 laser_k, -2 laser_k} along the first axis.
//
// Errors print their message and exit with status 1 (the reference exits 0).
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "CRT_Base_IF.h"
#include "driver_util.h"

namespace RT_Solver {

template <class T, int dim>
class Bragg_single : public CRT_Base_IF<T, dim, 2> {
 public:
  explicit Bragg_single(ParameterHandler* p) : CRT_Base_IF<T, dim, 2>(p) {
    this->m_map_stepfcts["bragg_ad"] = &Do_Bragg_ad_Wrapper;
    this->m_map_stepfcts["bragg_simple"] = &Do_Bragg_simple_Wrapper;
    this->m_device_steps[&Do_Bragg_ad_Wrapper] = ATUS_STEP_BRAGG_AD;
    CPoint<dim> pt1, pt2, pt3;
    pt2[0] = 2 * this->laser_k[0];
    pt3[0] = -2 * this->laser_k[0];
    this->m_rabi_momentum_list.push_back(pt1);
    this->m_rabi_momentum_list.push_back(pt2);
    this->m_rabi_momentum_list.push_back(pt3);
    m_no_of_Mirror_pts = (int)(this->Get_dimY() * this->Get_dimZ());
    m_Mirror.assign(m_no_of_Mirror_pts, 0.0);
  }

 protected:
  std::vector<double> m_Mirror;
  int m_no_of_Mirror_pts;

  bool check_consistency(const generic_header& head) {
    switch (dim) {
      case 2: return head.nDimX == this->Get_dimY();
      case 3: return head.nDimX == this->Get_dimY() && head.nDimY == this->Get_dimZ();
    }
    return false;
  }

  // Bragg_single::run_custom_sequence (bragg.cpp:131-159)
  bool run_custom_sequence(const sequence_item& item) override {
    if (item.name != "set_mirror") return false;
    generic_header header_tmp{};
    std::ifstream file1(item.content, std::ifstream::binary);
    file1.read((char*)&header_tmp, sizeof(generic_header));
    if (!check_consistency(header_tmp)) throw std::string("Change number of points for mirror.");
    file1.seekg(sizeof(generic_header), std::ifstream::beg);
    file1.read((char*)m_Mirror.data(), sizeof(double) * m_no_of_Mirror_pts);
    return true;
  }

  static void Do_Bragg_ad_Wrapper(void* ptr, sequence_item&) { static_cast<Bragg_single*>(ptr)->Do_Bragg_ad(); }
  static void Do_Bragg_simple_Wrapper(void* ptr, sequence_item&) {
    static_cast<Bragg_single*>(ptr)->Do_Bragg_simple();
  }

  // Bragg_single::Do_Bragg_ad (bragg.cpp:186-304), one step on the device
  void Do_Bragg_ad() {
    atus_laser L = this->laser_params();
    const double F = this->Amplitude_at_time();
    atus_host::check(atus_bragg_ad(this->m_set->device(), this->Get_dt(), this->Get_t(), F, &L), "atus_bragg_ad");
  }

  // Bragg_single::Do_Bragg_simple (bragg.cpp:306-371), on the host mirrors
  void Do_Bragg_simple() {
    const double dt = this->Get_dt();
    fftw_complex* Psi_1 = this->m_fields[0]->Getp2In();
    fftw_complex* Psi_2 = this->m_fields[1]->Getp2In();
    const double Omega = -this->Amp[0] * this->Amp[0] / (4 * this->DeltaL[1]);
    const double phase1 = 0;
    const double om[2] = {std::cos(2 * Omega * dt), std::sin(2 * Omega * dt)};
    const double ph[2] = {std::cos(phase1), std::sin(phase1)};
    const double O11[2] = {0.5 * (1 + om[0]), 0.5 * om[1]};
    const double O22[2] = {O11[0], O11[1]};
    const double O12[2] = {0.5 * (ph[0] * om[0] - ph[0] - ph[1] * om[1]), 0.5 * (ph[0] * om[1] + ph[1] * om[0] - ph[1])};
    const double O21[2] = {0.5 * (ph[0] * om[0] - ph[0] + ph[1] * om[1]), 0.5 * (ph[0] * om[1] - ph[1] * om[0] + ph[1])};
    for (int l = 0; l < this->m_no_of_pts; l++) {
      const double a0 = Psi_1[l][0], a1 = Psi_1[l][1], b0 = Psi_2[l][0], b1 = Psi_2[l][1];
      Psi_1[l][0] = (O11[0] * a0 - O11[1] * a1) + (O12[0] * b0 - O12[1] * b1);
      Psi_1[l][1] = (O11[0] * a1 + O11[1] * a0) + (O12[0] * b1 + O12[1] * b0);
      Psi_2[l][0] = (O21[0] * a0 - O21[1] * a1) + (O22[0] * b0 - O22[1] * b1);
      Psi_2[l][1] = (O21[0] * a1 + O21[1] * a0) + (O22[0] * b1 + O22[1] * b0);
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
      RT_Solver::Bragg_single<Fourier::cft_1d, 1> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 2) {
      RT_Solver::Bragg_single<Fourier::cft_2d, 2> rtsol(&params);
      rtsol.run_sequence();
    } else if (dim == 3) {
      RT_Solver::Bragg_single<Fourier::cft_3d, 3> rtsol(&params);
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
