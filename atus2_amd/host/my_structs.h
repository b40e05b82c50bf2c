// my_structs.h -- the .bin wavefunction header of ATUS-2 (my_structs.h:40-72
// of the reference), byte for byte: 1380 bytes, 4-byte packing, followed by
// nDimX*nDimY*nDimZ complex128 values (row-major, z fastest).
#ifndef ATUS_AMD_MY_STRUCTS_H
#define ATUS_AMD_MY_STRUCTS_H

#define DIM1D 1
#define DIM2D 2
#define DIM3D 3

#pragma pack(push)
#pragma pack(4)
struct generic_header {
  long long nself;           // size of this struct (1380)
  long long nDatatyp;        // size of one value (16 for complex128)
  long long nself_and_data;
  long long nDims;
  long long nDimX;
  long long nDimY;
  long long nDimZ;
  int bAtom;
  int bComplex;
  double t;
  double xMin;
  double xMax;
  double yMin;
  double yMax;
  double zMin;
  double zMax;
  double dx;
  double dy;
  double dz;
  double dkx;
  double dky;
  double dkz;
  double dt;
  int ks;  // coordinate system
  int fs;  // 1: Fourier space, 0: real space
  int nFuture[99];
  double dFuture[100];
};
#pragma pack(pop)

static_assert(sizeof(generic_header) == 1380, "generic_header must be 1380 bytes");

#endif
