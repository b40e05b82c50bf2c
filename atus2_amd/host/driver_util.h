// driver_util.h -- helpers shared by the drivers: thread count from
// MY_NO_OF_THREADS (default 4, as rt_solver.cpp:159-165) and evaluation of
// an XML formula at every grid point, one evaluator per OpenMP thread.
#ifndef ATUS_AMD_DRIVER_UTIL_H
#define ATUS_AMD_DRIVER_UTIL_H

#include <cstdlib>
#include <functional>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "CPoint.h"
#include "ParameterHandler.h"
#include "expr.h"

namespace atus_host {

inline int set_threads_from_env() {
  int n = 4;
  if (const char* s = std::getenv("MY_NO_OF_THREADS")) n = std::atoi(s);
  if (n < 1) n = 1;
#ifdef _OPENMP
  omp_set_num_threads(n);
#endif
  return n;
}

// out[l] = formula(x(l)) with x, y, z bound to the coordinates of point l
template <int dim, class Coord>
void eval_on_grid(ParameterHandler& params, const std::string& formula, long long npts, Coord coord, double* out) {
  std::string err;
#pragma omp parallel
  {
    CPoint<dim> x;
    Expr e;
    params.Setup_muParser(e);
    e.SetExpr(formula);
    e.DefineVar("x", &x[0]);
    if (dim > 1) e.DefineVar("y", &x[dim > 1 ? 1 : 0]);
    if (dim > 2) e.DefineVar("z", &x[dim > 2 ? 2 : 0]);
#pragma omp for schedule(static)
    for (long long l = 0; l < npts; ++l) {
      try {
        x = coord(l);
        out[l] = e.Eval();
      } catch (std::string& s) {
#pragma omp critical
        err = s;
      }
    }
  }
  if (!err.empty()) throw err;
}

}  // namespace atus_host

#endif
