// Harness for oracle/_ref: calls the reference's gaussianQuadrature
// (quadrature.cpp:43-75) and prints every node component as a hex float, so a
// test can compare them bit for bit with the oracle's rule.
#include <cstdio>
#include <vector>
#include "Vector3D.h"

void gaussianQuadrature(int dim, std::vector<std::vector<Vec3D> >& QP);

int main() {
  std::vector<std::vector<Vec3D> > QP;
  gaussianQuadrature(6, QP);
  std::printf("%zu\n", QP.size());
  for (size_t k = 0; k < QP.size(); ++k) {
    for (size_t s = 0; s < QP[k].size(); ++s)
      for (int c = 0; c < 3; ++c) std::printf("%a ", QP[k][s][c]);
    std::printf("\n");
  }
  return 0;
}
