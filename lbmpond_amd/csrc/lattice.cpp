// lattice.cpp -- host evaluation of the lattice weights with the reference's formulas.
// Compiled with -ffp-contract=off: every operation is rounded separately, as in the
// reference's constant initialisers (consts.h:34-37, 127-129, 200-229).
#include "lattice.h"

namespace pond {

bool make_lattice(int lattice, LatticeTables* t) {
  *t = LatticeTables();
  t->lattice = lattice;
  if (lattice == kD3Q27) {
    t->Qn = 27;
    t->TLat = 1. / 3.;
    t->Dim = 3;
    const double T = t->TLat;
    const double W[4] = {(1 - T) * (1 - T) * (1 - T), (1 - T) * (1 - T) * T / 2.0, (1 - T) * T * T / 4.0,
                         T * T * T / 8.};
    for (int i = 0; i < 27; i++) {
      int n2 = 0;
      for (int d = 0; d < 3; d++) {
        t->e[i][d] = kE27[i][d];
        t->ef[i][d] = static_cast<double>(kE27[i][d]);
        n2 += kE27[i][d] * kE27[i][d];
      }
      t->w[i] = W[n2];  // consts.h:38-47: weight by |e|^2 class
    }
    return true;
  }
  if (lattice != kD1Q5 && lattice != kD3Q125) return false;
  const double ec1d = kEc1 * kEc1, ec2d = kEc2 * kEc2;
  const double T = 1;
  const double W0 = (3 * T * T - (ec1d + ec2d) * T + ec1d * ec2d) / (ec1d * ec2d);
  const double W1 = T * (ec2d - 3 * T) / (2 * ec1d * (ec2d - ec1d));
  const double W2 = T * (ec1d - 3 * T) / (2 * ec2d * (ec1d - ec2d));
  const double wa[5] = {W0, W1, W1, W2, W2};
  const int ai[5] = {0, static_cast<int>(kEc1), -static_cast<int>(kEc1), static_cast<int>(kEc2),
                     -static_cast<int>(kEc2)};
  for (int k = 0; k < 5; k++) t->wa[k] = wa[k];
  t->TLat = T;
  if (lattice == kD1Q5) {
    t->Qn = 5;
    t->Dim = 1;
    for (int i = 0; i < 5; i++) {
      t->ef[i][0] = kA5[i];
      t->e[i][0] = ai[i];
      t->w[i] = wa[i];
    }
    return true;
  }
  t->Qn = 125;
  t->Dim = 3;
  for (int ix = 0; ix < 5; ix++)
    for (int iy = 0; iy < 5; iy++)
      for (int iz = 0; iz < 5; iz++) {
        const int i = 25 * ix + 5 * iy + iz;
        t->ef[i][0] = kA5[ix];
        t->ef[i][1] = kA5[iy];
        t->ef[i][2] = kA5[iz];
        t->e[i][0] = ai[ix];
        t->e[i][1] = ai[iy];
        t->e[i][2] = ai[iz];
        t->w[i] = wa[ix] * wa[iy] * wa[iz];  // (Wx*Wy)*Wz, consts.h:203-229
      }
  return true;
}

}  // namespace pond
