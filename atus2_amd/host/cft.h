// cft.h -- spectral field objects of the solver API (the reference's
// Fourier::cft_base<dim> and cft_1d/2d/3d, cft_base.h:38-262), backed by
// device memory of libatus_hip.so.
//
// Fields of one solver share one device context (a FieldSet).  Each field
// keeps a host mirror in the fftw_complex layout; Getp2In() returns it after
// downloading the device copy if that is newer, and marks the mirror as
// possibly modified, so the next device operation uploads it first.  Host
// code written against the reference (potential setup, custom steps) keeps
// working; the device path never touches the mirrors.
#ifndef ATUS_AMD_CFT_H
#define ATUS_AMD_CFT_H

#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/atus_hip.h"
#include "CPoint.h"
#include "my_structs.h"

typedef double fftw_complex[2];

namespace atus_host {

// throws std::string with the library's message when rc != ATUS_OK
void check(int rc, const char* what);

class FieldSet {
 public:
  FieldSet(const generic_header& h, int ncomp);
  ~FieldSet();
  FieldSet(const FieldSet&) = delete;
  FieldSet& operator=(const FieldSet&) = delete;

  int ncomp() const { return m_ncomp; }
  long long npts() const { return m_npts; }
  int dim() const { return m_dim; }
  // the context with every host-modified mirror uploaded; the caller is
  // about to run device work that may change all components
  atus_ctx* device();
  // same, for work that only reads the device copies
  atus_ctx* device_readonly();
  // host mirror of component c, current; may be modified by the caller
  fftw_complex* host(int c);
  // host mirror of component c, current, for reading only (no upload
  // before the next device operation)
  const fftw_complex* host_view(int c);
  // mark all components as changed on the device
  void device_changed();
  // a new batch context (atus_batch_create) of `nbatch` runs of this set's
  // grid, each with this set's number of components; the caller destroys it
  atus_ctx* make_batch(int nbatch) const;

 private:
  enum State { IN_SYNC = 0, HOST_NEWER = 1, DEVICE_NEWER = 2 };
  atus_ctx* m_ctx = nullptr;
  int m_ncomp = 0;
  int m_dim = 0;
  long long m_npts = 0;
  long long m_n[3] = {1, 1, 1};
  double m_dx[3] = {0, 0, 0}, m_dk[3] = {0, 0, 0};
  std::vector<std::vector<double>> m_mirror;
  std::vector<int> m_state;
  // host code written for the reference calls Getp2In() from every thread
  // of a parallel region (bragg.cpp:187-194): the mirror states change
  // under this lock
  std::mutex m_mutex;
  void flush();
};

}  // namespace atus_host

namespace Fourier {

enum TYPE { REAL, COMPLEX };

template <int dim>
class cft_base {
 public:
  // standalone field with its own one-component device context
  // (cft_base.h:47-88); only complex in-place fields exist on the device
  cft_base(const generic_header& header, bool inplace = true, bool fix = false);
  // component `comp` of a shared context
  cft_base(std::shared_ptr<atus_host::FieldSet> set, int comp, const generic_header& header);
  virtual ~cft_base() = default;

  void SetFix(bool b) { m_bfix = b; }
  // cft_Nd::ft: -1 forward (e^{-i}) then * prod(dx)/(2pi)^{d/2}; +1 backward
  // then * prod(dk)/(2pi)^{d/2}; other values do nothing.  With SetFix(true)
  // the result is also reordered to centred k order with the (-1)^index
  // factor of cft_Nd::fix (cft_1d.cpp:158-181).
  virtual void ft(int isign);
  virtual CPoint<dim> Get_k(const int64_t l);
  virtual CPoint<dim> Get_x(const int64_t l);

  fftw_complex* Getp2In() { return m_set->host(m_comp); }
  fftw_complex* Getp2Out() { return m_set->host(m_comp); }

  int Get_Dim_X() { return m_dim_x; }
  int Get_Dim_Y() { return m_dim_y; }
  int Get_Dim_Z() { return m_dim_z; }
  int64_t Get_Dim_RS() { return m_dim; }
  int64_t Get_Dim_FS() { return m_dim; }
  // writes the header and this field (cft_base.h:119-168, rs = real space)
  void save(const std::string& filename, bool rs = true);

  atus_host::FieldSet& fieldset() { return *m_set; }
  int comp() const { return m_comp; }

 protected:
  std::shared_ptr<atus_host::FieldSet> m_set;
  int m_comp = 0;
  int m_dim_x = 1, m_dim_y = 1, m_dim_z = 1;
  int m_shift_x = 0, m_shift_y = 0, m_shift_z = 0;
  int64_t m_dim = 0;
  bool m_bfix = false;
  double m_dx = 0, m_dy = 0, m_dz = 0, m_dkx = 0, m_dky = 0, m_dkz = 0;
  generic_header m_header{};
  void setup(const generic_header& h);
  void apply_fix();
};

class cft_1d : public cft_base<1> {
 public:
  // the reference's 1D field starts in fix() ordering (cft_1d.cpp:42)
  cft_1d(const generic_header& h, bool b = true, bool f = false) : cft_base<1>(h, b, f) { m_bfix = true; }
  cft_1d(std::shared_ptr<atus_host::FieldSet> s, int c, const generic_header& h) : cft_base<1>(s, c, h) {
    m_bfix = true;
  }
};
class cft_2d : public cft_base<2> {
 public:
  using cft_base<2>::cft_base;
};
class cft_3d : public cft_base<3> {
 public:
  using cft_base<3>::cft_base;
};

}  // namespace Fourier

#endif
