// fftw_threads.h -- the FFTW 3 thread set-up names the reference's drivers
// call around their solvers (rt_solver.cpp:163-165, 202; bragg.cpp;
// sob_solver.cpp).  The transforms run on the device here, so the threads
// they name are the host threads of the host loops (formula evaluation,
// custom steps on the host mirrors): fftw_plan_with_nthreads sets the
// OpenMP thread count.
#ifndef ATUS_AMD_FFTW_THREADS_H
#define ATUS_AMD_FFTW_THREADS_H

#ifdef _OPENMP
#include <omp.h>
#endif

inline int fftw_init_threads(void) { return 1; }
inline void fftw_plan_with_nthreads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}
inline void fftw_cleanup_threads(void) {}
inline void fftw_cleanup(void) {}

#endif
