// lattice.h -- lattice tables of the PonD step, shared by host and device code.
//
// D1Q5  consts.h:111-134   Gauss-Hermite 5-point, TLat=1, Dim=1
// D3Q27 consts.h:14-65     integer velocities in the reference's (irregular) order, TLat=1/3
// D3Q125 consts.h:136-245  tensor product of D1Q5, index 25*ix+5*iy+iz over the per-axis
//                          order {0,+ec1,-ec1,+ec2,-ec2}
//
// The abscissae ec1/ec2 are the reference's decimal literals (consts.h:117-118).  Weights are
// evaluated at run time on the host with the reference's formulas, operation by operation,
// because they differ from the closed forms by a few ulp and the plain path must reproduce
// the reference bit for bit.
#pragma once
#include <cstdint>

namespace pond {

constexpr int kD1Q5 = 0, kD3Q27 = 1, kD3Q125 = 2;
constexpr int kQmax = 125;

constexpr double kEc1 = 1.3556261799742657;
constexpr double kEc2 = 2.8569700138728056;

// D3Q27 velocity table in the reference order (consts.h:20-30).
constexpr int kE27[27][3] = {
    {0, 0, 0},
    {1, 0, 0},   {0, 1, 0},   {0, 0, 1},  {1, 1, 0},  {1, 0, 1}, {0, 1, 1}, {1, 1, 1},
    {-1, 0, 0},  {-1, 1, 0},  {-1, 0, 1}, {-1, 1, 1},
    {0, -1, 0},  {0, -1, 1},  {1, -1, 0}, {1, -1, 1},
    {0, 0, -1},  {1, 0, -1},  {0, 1, -1}, {1, 1, -1},
    {-1, -1, 0}, {-1, -1, 1},
    {-1, 0, -1}, {-1, 1, -1},
    {0, -1, -1}, {1, -1, -1},
    {-1, -1, -1}};

// population index of D3Q27 velocity (jx,jy,jz), components in {-1,0,1} mapped to 0..2
struct Idx27 {
  int v[3][3][3];
  constexpr Idx27() : v() {
    for (int i = 0; i < 27; i++) v[kE27[i][0] + 1][kE27[i][1] + 1][kE27[i][2] + 1] = i;
  }
};
constexpr Idx27 kIdx27{};

// per-axis abscissae of D1Q5 / D3Q125 in table order
constexpr double kA5[5] = {0.0, kEc1, -kEc1, kEc2, -kEc2};

constexpr int qn_of(int lattice) {
  return lattice == kD1Q5 ? 5 : lattice == kD3Q27 ? 27 : lattice == kD3Q125 ? 125 : 0;
}

struct LatticeTables {
  int lattice = 0, Qn = 0;
  double TLat = 0, Dim = 0;
  int e[kQmax][3] = {};
  double ef[kQmax][3] = {};
  double w[kQmax] = {};
  double wa[5] = {};  // per-axis D1Q5 weights (D1Q5/D3Q125)
};

// fills t for `lattice`; returns false for an unknown lattice
bool make_lattice(int lattice, LatticeTables* t);

}  // namespace pond
