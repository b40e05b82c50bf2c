// cft.cpp -- see cft.h.
#include "cft.h"

#include <cstring>
#include <fstream>

namespace atus_host {

void check(int rc, const char* what) {
  if (rc != ATUS_OK) throw std::string(std::string("Error in ") + what + ": " + atus_last_error());
}

FieldSet::FieldSet(const generic_header& h, int ncomp) : m_ncomp(ncomp) {
  m_dim = (int)h.nDims;
  long long n[3] = {h.nDimX, m_dim > 1 ? h.nDimY : 1, m_dim > 2 ? h.nDimZ : 1};
  double dx[3] = {h.dx, h.dy, h.dz};
  double dk[3] = {h.dkx, h.dky, h.dkz};
  for (int a = m_dim; a < 3; ++a) dx[a] = dk[a] = 0.0;
  m_npts = n[0] * n[1] * n[2];
  for (int a = 0; a < 3; ++a) {
    m_n[a] = n[a];
    m_dx[a] = dx[a];
    m_dk[a] = dk[a];
  }
  check(atus_ctx_create(&m_ctx, m_dim, n, dx, dk, ncomp), "atus_ctx_create");
  m_mirror.resize(ncomp);
  m_state.assign(ncomp, IN_SYNC);
  for (int c = 0; c < ncomp; ++c) m_mirror[c].assign((size_t)2 * m_npts, 0.0);
}

FieldSet::~FieldSet() { atus_ctx_destroy(m_ctx); }

void FieldSet::flush() {
  for (int c = 0; c < m_ncomp; ++c)
    if (m_state[c] == HOST_NEWER) {
      check(atus_upload(m_ctx, c, m_mirror[c].data()), "atus_upload");
      m_state[c] = IN_SYNC;
    }
}

atus_ctx* FieldSet::device() {
  std::lock_guard<std::mutex> lk(m_mutex);
  flush();
  for (int c = 0; c < m_ncomp; ++c) m_state[c] = DEVICE_NEWER;
  return m_ctx;
}

atus_ctx* FieldSet::device_readonly() {
  std::lock_guard<std::mutex> lk(m_mutex);
  flush();
  return m_ctx;
}

fftw_complex* FieldSet::host(int c) {
  std::lock_guard<std::mutex> lk(m_mutex);
  if (m_state[c] == DEVICE_NEWER) check(atus_download(m_ctx, c, m_mirror[c].data()), "atus_download");
  m_state[c] = HOST_NEWER;
  return reinterpret_cast<fftw_complex*>(m_mirror[c].data());
}

const fftw_complex* FieldSet::host_view(int c) {
  std::lock_guard<std::mutex> lk(m_mutex);
  if (m_state[c] == DEVICE_NEWER) {
    check(atus_download(m_ctx, c, m_mirror[c].data()), "atus_download");
    m_state[c] = IN_SYNC;
  }
  return reinterpret_cast<const fftw_complex*>(m_mirror[c].data());
}

atus_ctx* FieldSet::make_batch(int nbatch) const {
  atus_ctx* b = nullptr;
  check(atus_batch_create(&b, m_dim, m_n, m_dx, m_dk, m_ncomp, nbatch), "atus_batch_create");
  return b;
}

void FieldSet::device_changed() {
  std::lock_guard<std::mutex> lk(m_mutex);
  for (int c = 0; c < m_ncomp; ++c) m_state[c] = DEVICE_NEWER;
}

}  // namespace atus_host

namespace Fourier {

template <int dim>
cft_base<dim>::cft_base(const generic_header& header, bool /*inplace*/, bool fix) : m_bfix(fix) {
  if (header.nDims != dim) throw std::string("Critical error: header.nDims does not match template parameter dim");
  m_set = std::make_shared<atus_host::FieldSet>(header, 1);
  m_comp = 0;
  setup(header);
}

template <int dim>
cft_base<dim>::cft_base(std::shared_ptr<atus_host::FieldSet> set, int comp, const generic_header& header)
    : m_set(std::move(set)), m_comp(comp) {
  if (header.nDims != dim) throw std::string("Critical error: header.nDims does not match template parameter dim");
  setup(header);
}

template <int dim>
void cft_base<dim>::setup(const generic_header& h) {
  m_header = h;
  m_dim_x = (int)h.nDimX;
  m_dim_y = (dim > 1 && h.nDimY > 0) ? (int)h.nDimY : 1;
  m_dim_z = (dim > 2 && h.nDimZ > 0) ? (int)h.nDimZ : 1;
  m_dx = h.dx;
  m_dy = h.dy;
  m_dz = h.dz;
  m_dkx = h.dkx;
  m_dky = h.dky;
  m_dkz = h.dkz;
  m_shift_x = m_dim_x / 2;
  m_shift_y = m_dim_y / 2;
  m_shift_z = m_dim_z / 2;
  m_dim = (int64_t)m_dim_x * m_dim_y * m_dim_z;
}

// cft_3d::Get_x (cft_3d.cpp:128-138)
template <int dim>
CPoint<dim> cft_base<dim>::Get_x(const int64_t l) {
  CPoint<dim> r;
  const int64_t i = l / m_dim_y / m_dim_z;
  const int64_t j = (l - i * m_dim_y * m_dim_z) / m_dim_z;
  const int64_t k = l - i * m_dim_y * m_dim_z - j * m_dim_z;
  r[0] = double(i - m_shift_x) * m_dx;
  if (dim > 1) r[dim > 1 ? 1 : 0] = double(j - m_shift_y) * m_dy;
  if (dim > 2) r[dim > 2 ? 2 : 0] = double(k - m_shift_z) * m_dz;
  return r;
}

// cft_3d::Get_k (cft_3d.cpp:145-166): native order unless fixed
template <int dim>
CPoint<dim> cft_base<dim>::Get_k(const int64_t l) {
  CPoint<dim> r;
  int64_t i = l / m_dim_y / m_dim_z;
  int64_t j = (l - i * m_dim_y * m_dim_z) / m_dim_z;
  int64_t k = l - i * m_dim_y * m_dim_z - j * m_dim_z;
  if (!m_bfix) {
    i = (i + m_shift_x) % m_dim_x;
    j = (j + m_shift_y) % m_dim_y;
    k = (k + m_shift_z) % m_dim_z;
  }
  r[0] = m_dkx * double(i - m_shift_x);
  if (dim > 1) r[dim > 1 ? 1 : 0] = m_dky * double(j - m_shift_y);
  if (dim > 2) r[dim > 2 ? 2 : 0] = m_dkz * double(k - m_shift_z);
  return r;
}

template <int dim>
void cft_base<dim>::ft(int isign) {
  if (isign != 1 && isign != -1) return;
  atus_host::check(atus_ft(m_set->device(), m_comp, isign), "atus_ft");
  if (m_bfix) apply_fix();
}

// centred reordering (swap of half-spaces along every axis) with the sign
// (-1)^(sum of new indices), the reordering part of cft_Nd::fix
template <int dim>
void cft_base<dim>::apply_fix() {
  fftw_complex* d = m_set->host(m_comp);
  std::vector<double> tmp((size_t)2 * m_dim);
  for (int64_t i = 0; i < m_dim_x; ++i)
    for (int64_t j = 0; j < m_dim_y; ++j)
      for (int64_t k = 0; k < m_dim_z; ++k) {
        const int64_t src = ((i + m_shift_x) % m_dim_x) * m_dim_y * m_dim_z + ((j + m_shift_y) % m_dim_y) * m_dim_z +
                            ((k + m_shift_z) % m_dim_z);
        const int64_t dst = i * m_dim_y * m_dim_z + j * m_dim_z + k;
        const double s = ((i + j + k) & 1) ? -1.0 : 1.0;
        tmp[2 * dst] = s * d[src][0];
        tmp[2 * dst + 1] = s * d[src][1];
      }
  std::memcpy(d, tmp.data(), sizeof(double) * tmp.size());
}

template <int dim>
void cft_base<dim>::save(const std::string& filename, bool rs) {
  generic_header h = m_header;
  h.nDatatyp = 2 * sizeof(double);
  h.bComplex = 1;
  h.fs = rs ? 0 : 1;
  std::ofstream ofs(filename, std::ios::binary);
  ofs.write(reinterpret_cast<const char*>(&h), sizeof(generic_header));
  ofs.write(reinterpret_cast<const char*>(m_set->host(m_comp)), sizeof(fftw_complex) * m_dim);
}

template class cft_base<1>;
template class cft_base<2>;
template class cft_base<3>;

}  // namespace Fourier
