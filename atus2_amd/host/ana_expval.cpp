// ana_expval -- moments of a .bin wave function (the reference's ana_expval,
// ana_expval.cpp:232-399), computed on the device (atus_get_moments): the
// particle number N, <x>, <(x-<x>)^2>, <k>, <(k-<k>)^2> (each normalised by
// N) and, for 1D states, the density-weighted mean phase.  One line,
// tab-separated, CPoint values as "a,b,c":
//   1D:     phase  N  <x>  <x^2>  <k>  <k^2>
//   2D, 3D: N  <x>  <x^2>  <k>  <k^2>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "CPoint.h"
#include "cft.h"
#include "driver_util.h"

template <int dim>
static CPoint<dim> point(const double* v) {
  CPoint<dim> p;
  for (int a = 0; a < dim; ++a) p[a] = v[a];
  return p;
}

template <int dim>
static void print(const atus_moments& m) {
  if (dim == 1) std::cout << m.phase << "\t";
  std::cout << m.N << "\t" << point<dim>(m.x) << "\t" << point<dim>(m.x2) << "\t" << point<dim>(m.k) << "\t"
            << point<dim>(m.k2) << std::endl;
}

int main(int argc, char* argv[]) {
  atus_host::set_threads_from_env();
  generic_header header{};
  if (argc > 1) {
    std::ifstream f(argv[1], std::ifstream::binary);
    if (!f) {
      printf("Could not open file %s.\n", argv[1]);
      return 0;
    }
    f.read(reinterpret_cast<char*>(&header), sizeof(generic_header));
  }
  if (header.nDims == 1) header.nDimY = header.nDimZ = 1;
  if (header.nDims == 2) header.nDimZ = 1;
  if (header.nDims < 1 || header.nDims > 3 || header.bComplex != 1 || header.nDatatyp != sizeof(fftw_complex))
    return 0;
  try {
    atus_host::FieldSet set(header, 1);
    const long long n = set.npts();
    std::vector<double> data(2 * (size_t)n, 0.0);
    {
      std::ifstream f(argv[1], std::ifstream::binary);
      f.seekg(sizeof(generic_header), std::ifstream::beg);
      f.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(sizeof(fftw_complex) * n));
    }
    atus_host::check(atus_upload(set.device(), 0, data.data()), "atus_upload");
    atus_moments m{};
    atus_host::check(atus_get_moments(set.device_readonly(), 0, &m), "atus_get_moments");
    if (header.nDims == 1) print<1>(m);
    if (header.nDims == 2) print<2>(m);
    if (header.nDims == 3) print<3>(m);
  } catch (std::string& s) {
    std::cout << s << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
