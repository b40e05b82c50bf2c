// inflate_domain -- embeds a .bin wave function in a larger domain (the
// reference's inflate_domain, inflate_domain.cpp:41-461), on the device
// (atus_inflate):
//   inflate_domain --fak N file.bin   axes N times longer (domain N times
//                                     larger, same dx), the state in the
//                                     centre, zeros around it
//   inflate_domain --fix N file.bin   the same number of points, dx N times
//                                     larger: every N-th point of the state
//                                     around the centre, zeros around it
// (also -f N / -g N, --fak=N / --fix=N).  Writes inf_<file.bin> with the new
// header (dt = 0.001, as the reference sets it) and prints both headers.
//
// The reference's option check (result.count("") == 0, inflate_domain.cpp:
// 241-245) is always true there, so it prints its help and stops; this tool
// runs the inflation the rest of that file implements.  The new grid must
// have power-of-two axes (the device's FFT sizes); --fix in 1D, which the
// reference leaves unimplemented, is an error.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cft.h"
#include "driver_util.h"

static void print_header(const generic_header& h) {
  printf("# nDims    == %lld\n", h.nDims);
  printf("# nDimX    == %lld\n", h.nDimX);
  printf("# nDimY    == %lld\n", h.nDimY);
  printf("# nDimZ    == %lld\n", h.nDimZ);
  printf("# nDatatyp == %lld\n", h.nDatatyp);
  printf("# bAtom    == %d\n", h.bAtom);
  printf("# bComplex == %d\n", h.bComplex);
  printf("# t        == %g\n", h.t);
  printf("# dt       == %g\n", h.dt);
  printf("# xMin     == %g\n", h.xMin);
  printf("# xMax     == %g\n", h.xMax);
  printf("# yMin     == %g\n", h.yMin);
  printf("# yMax     == %g\n", h.yMax);
  printf("# zMin     == %g\n", h.zMin);
  printf("# zMax     == %g\n", h.zMax);
  printf("# dx       == %g\n", h.dx);
  printf("# dy       == %g\n", h.dy);
  printf("# dz       == %g\n", h.dz);
  printf("# dkx      == %g\n", h.dkx);
  printf("# dky      == %g\n", h.dky);
  printf("# dkz      == %g\n", h.dkz);
}

// the header changes of inflate_{1,2,3}d[_fix] (inflate_domain.cpp:41-223)
static generic_header new_header(const generic_header& o, int fak, bool fix) {
  generic_header h = o;
  const int dim = (int)o.nDims;
  const double f = double(fak);
  double* mins[3] = {&h.xMin, &h.yMin, &h.zMin};
  double* maxs[3] = {&h.xMax, &h.yMax, &h.zMax};
  double* dxs[3] = {&h.dx, &h.dy, &h.dz};
  double* dks[3] = {&h.dkx, &h.dky, &h.dkz};
  long long* ns[3] = {&h.nDimX, &h.nDimY, &h.nDimZ};
  for (int a = 0; a < dim; ++a) {
    *mins[a] *= f;
    *maxs[a] *= f;
    *dks[a] /= f;
    if (fix)
      *dxs[a] *= f;
    else
      *ns[a] *= fak;
  }
  if (!fix) {
    if (dim < 3) h.nDimZ = 1;
    if (dim < 2) h.nDimY = 1;
  }
  h.dt = 0.001;
  return h;
}

int main(int argc, char* argv[]) {
  int fak = -1, fix = -1;
  std::string filename;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto value = [&](const std::string& opt) -> int {
      if (a.size() > opt.size() && a.compare(0, opt.size() + 1, opt + "=") == 0) return std::stoi(a.substr(opt.size() + 1));
      if (i + 1 >= argc) throw std::string("option " + opt + " needs a value");
      return std::stoi(argv[++i]);
    };
    try {
      if (a == "-h" || a == "--help") {
        std::cout << "inflate_domain [--fak N | --fix N] file.bin\n";
        return EXIT_FAILURE;
      } else if (a == "-f" || a.rfind("--fak", 0) == 0) {
        fak = value(a == "-f" ? "-f" : "--fak");
      } else if (a == "-g" || a.rfind("--fix", 0) == 0) {
        fix = value(a == "-g" ? "-g" : "--fix");
      } else if (filename.empty()) {
        filename = a;
      }
    } catch (std::string& s) {
      std::cout << "error parsing options: " << s << std::endl;
      return EXIT_FAILURE;
    } catch (std::exception&) {
      std::cout << "error parsing options: " << a << std::endl;
      return EXIT_FAILURE;
    }
  }
  if (filename.empty()) {
    std::cout << "error parsing options: missing file name" << std::endl;
    return EXIT_FAILURE;
  }
  if (fak == fix) {
    std::cout << "Hmmm" << std::endl;
    return EXIT_FAILURE;
  }
  const bool bfix = fix != -1;
  const int factor = bfix ? fix : fak;
  atus_host::set_threads_from_env();

  generic_header old_h{}, new_h{};
  std::ifstream in(filename, std::ifstream::binary);
  if (!in) {
    printf("Cannot open file: %s\n", filename.c_str());
    return EXIT_FAILURE;
  }
  in.read(reinterpret_cast<char*>(&old_h), sizeof(generic_header));
  printf("### %s\n", filename.c_str());
  print_header(old_h);
  if (old_h.nDims >= 1 && old_h.nDims <= 3 && old_h.bComplex == 1 && old_h.bAtom == 1 &&
      old_h.nDatatyp == sizeof(fftw_complex)) {
    try {
      if (factor < 1) throw std::string("the inflation factor must be >= 1");
      if (bfix && old_h.nDims == 1) throw std::string("--fix is not available for 1D states");
      generic_header src_h = old_h;
      if (src_h.nDims < 3) src_h.nDimZ = 1;
      if (src_h.nDims < 2) src_h.nDimY = 1;
      new_h = new_header(src_h, factor, bfix);
      atus_host::FieldSet src(src_h, 1), dst(new_h, 1);
      std::vector<double> data(2 * (size_t)src.npts(), 0.0);
      in.seekg(sizeof(generic_header), std::ifstream::beg);
      in.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(sizeof(fftw_complex) * src.npts()));
      atus_host::check(atus_upload(src.device(), 0, data.data()), "atus_upload");
      atus_host::check(atus_inflate(dst.device(), src.device_readonly(), factor, bfix ? 1 : 0), "atus_inflate");
      std::vector<double> out(2 * (size_t)dst.npts());
      atus_host::check(atus_download(dst.device_readonly(), 0, out.data()), "atus_download");
      std::ofstream o("inf_" + filename, std::ofstream::binary);
      o.write(reinterpret_cast<const char*>(&new_h), sizeof(generic_header));
      o.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)(sizeof(double) * out.size()));
    } catch (std::string& s) {
      std::cout << s << std::endl;
      return EXIT_FAILURE;
    }
  }
  printf("# new header\n");
  print_header(new_h);
  return EXIT_SUCCESS;
}
