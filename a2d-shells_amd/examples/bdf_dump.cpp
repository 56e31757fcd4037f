// bdf_dump.cpp -- scan a Nastran bulk-data file with TACSMeshLoader and
// print what it read (no GPU use): counts, nodes in id order, tensor-order
// connectivity with components, and the merged BC entries.
//
// Usage: bdf_dump file.bdf
#include <cstdio>

#include "TACSShellGPU.h"

using namespace shellgpu;

int main(int argc, char *argv[]) {
  if (argc < 2) {
    fprintf(stderr, "usage: bdf_dump file.bdf\n");
    return 2;
  }
  TACSMeshLoader *mesh = new TACSMeshLoader();
  mesh->incref();
  if (mesh->scanBDFFile(argv[1])) return 1;
  const int nn = mesh->getNumNodes(), ne = mesh->getNumElements();
  printf("nodes %d elements %d components %d\n", nn, ne, mesh->getNumComponents());
  for (int c = 0; c < mesh->getNumComponents(); c++) printf("C %d %s\n", c, mesh->getElementDescript(c));
  const std::vector<TacsScalar> &X = mesh->getNodes();
  for (int k = 0; k < nn; k++) printf("X %d %.17g %.17g %.17g\n", k, X[3 * k], X[3 * k + 1], X[3 * k + 2]);
  const std::vector<int> &conn = mesh->getConnectivity();
  const std::vector<int> &comp = mesh->getComponents();
  for (int e = 0; e < ne; e++)
    printf("E %d %d %d %d %d %d\n", e, conn[4 * e], conn[4 * e + 1], conn[4 * e + 2], conn[4 * e + 3], comp[e]);
  std::vector<int> bn, bm;
  std::vector<TacsScalar> bv;
  mesh->getBCs(bn, bm, bv);
  for (size_t k = 0; k < bn.size(); k++) {
    printf("B %d %d", bn[k], bm[k]);
    for (int j = 0; j < 6; j++) printf(" %.17g", bv[6 * k + j]);
    printf("\n");
  }
  mesh->decref();
  return 0;
}
