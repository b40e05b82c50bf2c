// CPoint.h -- fixed-size coordinate vector of the solver API (the reference's
// CPoint<dim>, CPoint.h:17-197): element access, assignment from a scalar
// (all elements) or a vector, difference, element-wise scale and dot
// product.  The reference's operator*(double), which assigns instead of
// scaling (CPoint.h:161-167), is not reproduced: here it scales.
#ifndef ATUS_AMD_CPOINT_H
#define ATUS_AMD_CPOINT_H

#include <initializer_list>
#include <ostream>
#include <string>
#include <vector>

template <int dim>
class CPoint {
 public:
  CPoint() {
    for (int i = 0; i < dim; ++i) m_data[i] = 0.0;
  }
  CPoint(std::initializer_list<double> list) {
    if ((int)list.size() != dim) throw std::string("Error: initializer_list size mismatch\n");
    int i = 0;
    for (double v : list) m_data[i++] = v;
  }
  CPoint& operator=(double s) {
    for (int i = 0; i < dim; ++i) m_data[i] = s;
    return *this;
  }
  CPoint& operator=(const std::vector<double>& v) {
    for (int i = 0; i < dim; ++i) m_data[i] = v[i];
    return *this;
  }
  double& operator[](int i) { return m_data[i]; }
  const double& operator[](int i) const { return m_data[i]; }

  CPoint operator+(const CPoint& r) const {
    CPoint o;
    for (int i = 0; i < dim; ++i) o[i] = m_data[i] + r[i];
    return o;
  }
  CPoint operator-(const CPoint& r) const {
    CPoint o;
    for (int i = 0; i < dim; ++i) o[i] = m_data[i] - r[i];
    return o;
  }
  // element-wise product
  CPoint scale(const CPoint& b) const {
    CPoint o;
    for (int i = 0; i < dim; ++i) o[i] = m_data[i] * b[i];
    return o;
  }
  // dot product, summed from the first element on
  double operator*(const CPoint& r) const {
    double s = 0.0;
    for (int i = 0; i < dim; ++i) s += m_data[i] * r[i];
    return s;
  }
  CPoint operator*(double s) const {
    CPoint o;
    for (int i = 0; i < dim; ++i) o[i] = m_data[i] * s;
    return o;
  }
  bool operator==(double s) const {
    for (int i = 0; i < dim; ++i)
      if (m_data[i] != s) return false;
    return true;
  }
  // elements separated by commas (CPoint::Dump, CPoint.h:171-179)
  friend std::ostream& operator<<(std::ostream& os, const CPoint& p) {
    for (int i = 0; i < dim; ++i) os << p.m_data[i] << (i + 1 < dim ? "," : "");
    return os;
  }

 private:
  double m_data[dim];
};

#endif
