// gen_psi_0 -- initial .bin states (the reference's gen_psi_0,
// gen_psi_0.cpp:37-266): psi = GUESS_nD(x) (real), normalised so that
// ar * sum psi^2 = CONSTANTS/N, header dt = 0.001, t = 0, written to
// FILENAME.  Host only.
//
// Optional: GEN_PSI_NOISE=<amplitude> and GEN_PSI_SEED=<seed> add complex
// noise amplitude*(xi_r + i xi_i), xi ~ N(0,1) from SplitMix64 + Box-Muller,
// before normalising (the seeded inputs of the benchmark configurations).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "ParameterHandler.h"
#include "driver_util.h"

namespace {

// SplitMix64 (Steele, Lea & Flood 2014) -> uniform (0,1) -> Box-Muller
struct Normal {
  uint64_t s;
  bool have = false;
  double spare = 0;
  explicit Normal(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double operator()() {
    if (have) {
      have = false;
      return spare;
    }
    double u1 = uniform(), u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(u1));
    spare = r * std::sin(2 * M_PI * u2);
    have = true;
    return r * std::cos(2 * M_PI * u2);
  }
};

template <int dim>
void generate(ParameterHandler& ph) {
  generic_header h{};
  h.nDims = dim;
  h.nself = sizeof(generic_header);
  h.bAtom = 1;
  h.bComplex = 1;
  h.nDatatyp = 2 * sizeof(double);
  h.xMax = ph.Get_xMax();
  h.xMin = ph.Get_xMin();
  h.yMax = ph.Get_yMax();
  h.yMin = ph.Get_yMin();
  h.zMax = ph.Get_zMax();
  h.zMin = ph.Get_zMin();
  h.dt = 0.001;
  h.nDimX = ph.Get_NX();
  h.nDimY = dim > 1 ? ph.Get_NY() : 1;
  h.nDimZ = dim > 2 ? ph.Get_NZ() : 1;
  h.dx = std::fabs(h.xMax - h.xMin) / double(h.nDimX);
  h.dkx = 2 * M_PI / std::fabs(h.xMax - h.xMin);
  double ar = h.dx;
  if (dim > 1) {
    h.dy = std::fabs(h.yMax - h.yMin) / double(h.nDimY);
    h.dky = 2 * M_PI / std::fabs(h.yMax - h.yMin);
    ar *= h.dy;
  }
  if (dim > 2) {
    h.dz = std::fabs(h.zMax - h.zMin) / double(h.nDimZ);
    h.dkz = 2 * M_PI / std::fabs(h.zMax - h.zMin);
    ar *= h.dz;
  }
  const std::string filename = ph.Get_simulation("FILENAME");
  const std::string guess = ph.Get_simulation("GUESS_" + std::to_string(dim) + "D");
  const double n_target = ph.Get_Constant("N");
  const long long Ntot = h.nDimX * h.nDimY * h.nDimZ;
  std::vector<double> re(Ntot), psi(2 * Ntot, 0.0);
  // gen_psi_0.cpp:199-220: x = (i - n/2) dx, independent of xMin
  auto coord = [&](long long l) {
    CPoint<dim> x;
    long long i = l / (h.nDimY * h.nDimZ);
    long long j = (l - i * h.nDimY * h.nDimZ) / h.nDimZ;
    long long k = l - i * h.nDimY * h.nDimZ - j * h.nDimZ;
    x[0] = double(i - (h.nDimX >> 1)) * h.dx;
    if (dim > 1) x[dim > 1 ? 1 : 0] = double(j - (h.nDimY >> 1)) * h.dy;
    if (dim > 2) x[dim > 2 ? 2 : 0] = double(k - (h.nDimZ >> 1)) * h.dz;
    return x;
  };
  atus_host::eval_on_grid<dim>(ph, guess, Ntot, coord, re.data());
  for (long long l = 0; l < Ntot; ++l) psi[2 * l] = re[l];
  if (const char* na = std::getenv("GEN_PSI_NOISE")) {
    const double amp = std::atof(na);
    const char* sd = std::getenv("GEN_PSI_SEED");
    Normal rng(sd ? std::strtoull(sd, nullptr, 10) : 1ULL);
    for (long long l = 0; l < Ntot; ++l) {
      psi[2 * l] += amp * rng();
      psi[2 * l + 1] += amp * rng();
    }
  }
  double N = 0;
  for (long long l = 0; l < Ntot; ++l) N += psi[2 * l] * psi[2 * l] + psi[2 * l + 1] * psi[2 * l + 1];
  N *= ar;
  if (N > 0.0) {
    const double f = std::sqrt(n_target / N);
    for (auto& v : psi) v *= f;
  }
  std::ofstream ofs(filename, std::ios::binary);
  ofs.write(reinterpret_cast<char*>(&h), sizeof(generic_header));
  ofs.write(reinterpret_cast<char*>(psi.data()), sizeof(double) * psi.size());
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc != 2) {
    printf("No xml file specified.\n");
    return EXIT_FAILURE;
  }
  try {
    ParameterHandler params(argv[1]);
    const int dim = std::stoi(params.Get_simulation("DIM"));
    std::cout << "DIM = " << dim << std::endl;
    atus_host::set_threads_from_env();
    if (dim == 1) generate<1>(params);
    else if (dim == 2) generate<2>(params);
    else if (dim == 3) generate<3>(params);
    else return EXIT_FAILURE;
  } catch (std::string& s) {
    std::cout << s << std::endl;
    return EXIT_FAILURE;
  } catch (std::exception& e) {
    std::cout << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
