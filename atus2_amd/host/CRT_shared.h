// CRT_shared.h -- common base of the real-time solvers (the reference's
// CRT_shared, CRT_shared.h:39-188): header bookkeeping, the StepFunction
// type of the step registry, Set_dt.
#ifndef ATUS_AMD_CRT_SHARED_H
#define ATUS_AMD_CRT_SHARED_H

#include <cmath>
#include <fstream>
#include <string>

#include "CPoint.h"
#include "ParameterHandler.h"
#include "cft.h"
#include "my_structs.h"

/** Function pointer with a sequence_item (CRT_shared.h:39) */
typedef void (*StepFunction)(void*, sequence_item&);

class CRT_shared {
 public:
  CRT_shared() : m_no_of_pts(0), m_no_of_pts_red(0), m_shift_x(0), m_shift_y(0), m_shift_z(0), m_ar(0), m_ar_k(0) {
    m_header = {};
  }
  virtual ~CRT_shared() = default;

  double& Get_t() { return m_header.t; }
  const double& Get_dt() const { return m_header.dt; }
  const double& Get_dx() const { return m_header.dx; }
  const double& Get_dy() const { return m_header.dy; }
  const double& Get_dz() const { return m_header.dz; }
  const double& Get_dkx() const { return m_header.dkx; }
  const double& Get_dky() const { return m_header.dky; }
  const double& Get_dkz() const { return m_header.dkz; }
  long long& Get_dimX() { return m_header.nDimX; }
  long long& Get_dimY() { return m_header.nDimY; }
  long long& Get_dimZ() { return m_header.nDimZ; }
  const int& Get_No_Points() const { return m_no_of_pts; }

  /// change dt and rebuild what depends on it (CRT_shared.h:117-121)
  void Set_dt(const double dt) {
    m_header.dt = dt;
    Init();
  }

 protected:
  /// reads the 1380-byte header of `filename` (CRT_shared.h:131-172)
  void Read_header(const std::string& filename, const int dim) {
    std::ifstream file1(filename, std::ifstream::binary);
    if (!file1.is_open()) throw std::string("Could not open file " + filename + ".\n");
    file1.read((char*)&m_header, sizeof(generic_header));
    if (!file1) throw std::string("Could not read the header of " + filename + ".\n");
    Set_geometry(dim);
  }
  void Set_geometry(const int dim) {
    switch (dim) {
      case 1:
        m_ar = m_header.dx;
        m_ar_k = m_header.dkx;
        m_no_of_pts = (int)m_header.nDimX;
        m_shift_x = (int)m_header.nDimX / 2;
        m_no_of_pts_red = m_shift_x + 1;
        m_header.nDimY = 1;
        m_header.nDimZ = 1;
        break;
      case 2:
        m_ar = m_header.dx * m_header.dy;
        m_ar_k = m_header.dkx * m_header.dky;
        m_no_of_pts = (int)(m_header.nDimX * m_header.nDimY);
        m_shift_x = (int)m_header.nDimX / 2;
        m_shift_y = (int)m_header.nDimY / 2;
        m_no_of_pts_red = (int)m_header.nDimX * (m_shift_y + 1);
        m_header.nDimZ = 1;
        break;
      case 3:
        m_ar = m_header.dx * m_header.dy * m_header.dz;
        m_ar_k = m_header.dkx * m_header.dky * m_header.dkz;
        m_no_of_pts = (int)(m_header.nDimX * m_header.nDimY * m_header.nDimZ);
        m_shift_x = (int)m_header.nDimX / 2;
        m_shift_y = (int)m_header.nDimY / 2;
        m_shift_z = (int)m_header.nDimZ / 2;
        m_no_of_pts_red = (int)(m_header.nDimX * m_header.nDimY) * (m_shift_z + 1);
        break;
    }
  }

  generic_header m_header;
  int m_no_of_pts;
  // points of the half spectrum of a real transform along the last axis
  // (CRT_shared.h:147-169).  No solver here reads it; it stays because it
  // belongs to the protected members that drivers derived from the solver
  // templates compile against.
  int m_no_of_pts_red;
  int m_shift_x;
  int m_shift_y;
  int m_shift_z;
  double m_ar;
  double m_ar_k;
  virtual void Init() = 0;
};

#endif
