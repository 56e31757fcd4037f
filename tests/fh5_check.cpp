// Writes one .f5 file of fixed content through a TACSFH5File: the
// reference's own writer (compiled with -DREF_FH5 against
// src/io/TACSFH5.cpp by oracle/Makefile, target ref) or this library's
// (host/TACSShellGPU.h).  The CPU tests compare the two files byte by byte.
// Usage: fh5_check <out.f5>
#ifdef REF_FH5
#include "TACSFH5.h"
#define NEW_FILE() new TACSFH5File(MPI_COMM_WORLD)
#else
#include "TACSShellGPU.h"
using namespace shellgpu;
#define NEW_FILE() new TACSFH5File()
#endif

int main(int argc, char **argv) {
  if (argc != 2) return 2;
#ifdef REF_FH5
  MPI_Init(&argc, &argv);
#endif
  TACSFH5File *f = NEW_FILE();
  f->incref();
  char c0[] = "Component 0", c1[] = "Skin panels";
  char *names[2] = {c0, c1};
  if (f->createFile(argv[1], 2, names)) return 1;
  int comp[5] = {0, 1, 1, 0, 1};
  char zn[] = "components";
  f->writeZoneData(zn, zn, TACSFH5File::FH5_INT, 5, 1, comp);
  float cont[27];
  for (int i = 0; i < 27; i++) cont[i] = 0.25f * i - 1.0f;
  char cn[] = "continuous data t=0.0000000000e+00", cv[] = "X,Y,Z,u,v,w,rotx,roty,rotz";
  f->writeZoneData(cn, cv, TACSFH5File::FH5_FLOAT, 3, 9, cont);
  double dd[6] = {1.5, -2.25, 3e-7, 4e7, -0.0, 6.125};
  char dn[] = "element data t=0.0000000000e+00", dv[] = "a,b";
  f->writeZoneData(dn, dv, TACSFH5File::FH5_DOUBLE, 3, 2, dd);
  f->close();
  f->decref();
#ifdef REF_FH5
  MPI_Finalize();
#endif
  return 0;
}
