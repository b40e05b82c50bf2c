// cft_3d.h -- the reference's header name of Fourier::cft_3d
// (cft_3d.h), whose device-backed form is declared in cft.h together with
// the FFTW thread names the drivers call.
#ifndef ATUS_AMD_CFT_3D_H
#define ATUS_AMD_CFT_3D_H

#include "cft.h"
#include "fftw_threads.h"

#endif
