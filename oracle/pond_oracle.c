/*
 * pond_oracle.c -- CPU restatement of the reference particles-on-demand (PonD) step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP step kernels.  Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark load it; the
 * product library (lbmpond_amd) never links, loads or calls anything in oracle/.
 *
 * It restates the reference's per-cell arithmetic operation by operation: the same loop
 * orders, the same expression trees, IEEE per-operation rounding (build with
 * -ffp-contract=off, no fast-math), and no memoisation: every stencil node is re-loaded and
 * re-transferred exactly as often as the reference does it, so its running time is also a
 * fair CPU baseline for the reference's algorithm.
 *
 * Citations are file:line in engrab6/lbmPonD.
 *   lattice tables ............ consts.h:14-65 (D3Q27), 111-134 (D1Q5), 136-245 (D3Q125)
 *   sign / sqrts .............. params.h:101-112
 *   Cell::set_base ............ lbpack.cu:15-21
 *   Cell::calcMoments ......... structs.cuh:315-342
 *   Cell::transfer_gauge ...... structs.cuh:343-551
 *   isConv .................... steps.cu:58-75
 *   LagrPol<N> ................ steps.cu:595-601
 *   propagateWENO<3,5> ........ steps.cu:774-825  (+ LagrEval<3>, steps.cu:924-937)
 *   propagate<4> .............. steps.cu:986-1036
 *   LBMpull (D1Q5/D3Q27) ...... steps.cu:356-365, 421-512, 513-564, 582
 *   PonDstepD3Q125 ............ steps.cu:151-250
 *   mirror boundaries ......... steps.cu:167-190 (PonDstepD3Q125), 1004-1022 (propagate<N>),
 *                               structs.cuh:552-578 (Cell::mirrX/Y/Z), consts.h:62-64 and
 *                               lbm.cu:206-244 (reverse tables)
 *   calcTotMoments ............ update.cu:52-69
 *   diagn ..................... diagn.cu:3-22
 *   fill/ind_zip layout ....... structs.cuh:153-166 (x fastest, then y, then z)
 *
 * Dead code of the reference that does not change results is not restated: the probe-cell
 * printfs and H-function logs (steps.cu:480-481, 516-531, 571-581), the discarded per-cell
 * atomics (steps.cu:583-585) and the unused M4 neighbour loads (steps.cu:436-440).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_D1Q5 0
#define ORC_D3Q27 1
#define ORC_D3Q125 2
#define ORC_WENO5 0
#define ORC_LAGR4 1
#define ORC_WENO5C 2
#define ORC_WENO5G 3
#define QMAX 125

typedef struct {
  int lattice, Qn;
  double TLat, Dim;
  int e[QMAX][3];     /* integer velocities, e_host */
  double ef[QMAX][3]; /* float velocities ef_host; for D3Q27 the integer table as doubles */
  double w[QMAX];
  double ec1, ec2;
  int rev[3][QMAX];   /* reverseX/Y/Z: the population whose velocity has axis d negated */
} olat;

typedef struct {
  int lattice, interp;
  long nx, ny, nz;
  double dtau, tol_abs, tol_rel;
  int max_iter;
  int mirror;  /* bit d: mirror boundaries on axis d (the reference's outflow switch) */
  int amr;     /* 1: cell activation and removal along x after every step (amr.cu) */
} orc_params;

typedef struct {
  orc_params p;
  olat L;
  long ncell;
  int rec;          /* doubles per cell record: Qn populations, rho, ux, uy, uz, T */
  double* ld;       /* load buffer (state at the start of the step) */
  double* st;       /* store buffer */
  int* niter;
  int* set;         /* Cinfo::set: 1 = the cell is stepped (structs.cuh:92) */
  int* setnew;      /* Cinfo::setnew */
} orc_state;

/* ---------------------------------------------------------------- lattice tables */

/* D3Q27 velocity order of consts.h:20-30 (irregular; kept because the population index is
   part of the interface and of the moment summation order). */
static const int D3Q27_E[27][3] = {
  {0, 0, 0},
  {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1},
  {-1, 0, 0}, {-1, 1, 0}, {-1, 0, 1}, {-1, 1, 1},
  {0, -1, 0}, {0, -1, 1}, {1, -1, 0}, {1, -1, 1},
  {0, 0, -1}, {1, 0, -1}, {0, 1, -1}, {1, 1, -1},
  {-1, -1, 0}, {-1, -1, 1},
  {-1, 0, -1}, {-1, 1, -1},
  {0, -1, -1}, {1, -1, -1},
  {-1, -1, -1}};

static void lat_init_tables(olat* L, int lattice);
static void lat_reverse(olat* L);
static void lat_init(olat* L, int lattice) {
  lat_init_tables(L, lattice);
  lat_reverse(L);
}

static void lat_init_tables(olat* L, int lattice) {
  memset(L, 0, sizeof(*L));
  L->lattice = lattice;
  if (lattice == ORC_D3Q27) {
    /* consts.h:15-47: TLat=1/3, Dim=3, weights by |e|^2 class */
    L->Qn = 27; L->TLat = 1. / 3.; L->Dim = 3;
    const double T = L->TLat;
    const double W0 = (1 - T) * (1 - T) * (1 - T);
    const double W1 = (1 - T) * (1 - T) * T / 2.0;
    const double W2 = (1 - T) * T * T / 4.0;
    const double W3 = T * T * T / 8.;
    const double Wc[4] = {W0, W1, W2, W3};
    for (int i = 0; i < 27; i++) {
      int n2 = 0;
      for (int d = 0; d < 3; d++) {
        L->e[i][d] = D3Q27_E[i][d];
        L->ef[i][d] = (double)D3Q27_E[i][d];
        n2 += D3Q27_E[i][d] * D3Q27_E[i][d];
      }
      L->w[i] = Wc[n2];
    }
  } else {
    /* consts.h:114-130 (D1Q5) and 139-229 (D3Q125): Gauss-Hermite 5-point abscissae given
       as decimal literals, weights by the reference's formulas. */
    const double ec1 = 1.3556261799742657, ec2 = 2.8569700138728056;
    const double ec1d = ec1 * ec1, ec2d = ec2 * ec2;
    const double T = 1;
    const double W0 = (3 * T * T - (ec1d + ec2d) * T + ec1d * ec2d) / (ec1d * ec2d);
    const double W1 = T * (ec2d - 3 * T) / (2 * ec1d * (ec2d - ec1d));
    const double W2 = T * (ec1d - 3 * T) / (2 * ec2d * (ec1d - ec2d));
    const double a[5] = {0, ec1, -ec1, ec2, -ec2};
    const int ai[5] = {0, (int)ec1, -(int)ec1, (int)ec2, -(int)ec2};
    const double wa[5] = {W0, W1, W1, W2, W2};
    L->ec1 = ec1; L->ec2 = ec2; L->TLat = T;
    if (lattice == ORC_D1Q5) {
      L->Qn = 5; L->Dim = 1;
      for (int i = 0; i < 5; i++) {
        L->ef[i][0] = a[i]; L->e[i][0] = ai[i];
        L->w[i] = wa[i];
      }
    } else {
      /* index 25*ix + 5*iy + iz over the per-axis order {0,+ec1,-ec1,+ec2,-ec2},
         weight (Wx*Wy)*Wz as written in consts.h:203-229 */
      L->Qn = 125; L->Dim = 3;
      for (int ix = 0; ix < 5; ix++)
        for (int iy = 0; iy < 5; iy++)
          for (int iz = 0; iz < 5; iz++) {
            int i = 25 * ix + 5 * iy + iz;
            L->ef[i][0] = a[ix]; L->ef[i][1] = a[iy]; L->ef[i][2] = a[iz];
            L->e[i][0] = ai[ix]; L->e[i][1] = ai[iy]; L->e[i][2] = ai[iz];
            L->w[i] = wa[ix] * wa[iy] * wa[iz];
          }
    }
  }
}

/* reverseX/Y/Z (consts.h:62-64 for D3Q27, consts.h:131-133 for D1Q5, recomputed for D3Q125 in
   lbm.cu:206-244): for a velocity with a non-zero component on axis d, the population whose
   velocity is its mirror image in that axis; the identity otherwise (Cell::mirrX/Y/Z,
   structs.cuh:552-578, leave those populations in place) */
static void lat_reverse(olat* L) {
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < L->Qn; i++) {
      L->rev[d][i] = i;
      if (L->ef[i][d] == 0) continue;
      for (int j = 0; j < L->Qn; j++) {
        int same = 1;
        for (int c = 0; c < 3; c++) same = same && L->ef[j][c] == (c == d ? -L->ef[i][c] : L->ef[i][c]);
        if (same) { L->rev[d][i] = j; break; }
      }
    }
}

/* ---------------------------------------------------------------- scalar helpers */

static int isign(double v) { return (0.0 < v) - (v < 0.0); } /* params.h:101-103 */
static double sqrts(double v) { return sqrt(fabs(v)) * isign(v); } /* params.h:110-112 */

/* ---------------------------------------------------------------- transfer_gauge */

/* One target population `inew` of the source cell (f, gauge uT0) re-expressed in gauge
   `base` (structs.cuh:343-551).  NaN/Inf terms are dropped (structs.cuh:390,516,542). */
static double transfer(const olat* L, const double* f, const double* uT0, const double* base, int inew) {
  const double Tbase = L->TLat;
  const double u0x = uT0[0], u0y = uT0[1], u0z = uT0[2], T0 = uT0[3];
  const double u1x = base[0], u1y = base[1], u1z = base[2], T1 = base[3];
  double fnew = 0;
  double wk = 1. / (2 * T1), wl = wk, wm = wk;
  const int Qn = L->Qn;
  if (L->lattice == ORC_D1Q5) {
    const double e1_to = L->ef[1][0] * sqrts(T1 / Tbase);
    const double e2_to = L->ef[3][0] * sqrts(T1 / Tbase);
    double e_fr[5], e_to[5];
    for (int i = 0; i < Qn; i++) {
      e_fr[i] = L->ef[i][0] * sqrts(T0 / Tbase);
      e_to[i] = L->ef[i][0] * sqrts(T1 / Tbase);
    }
    if (L->ef[inew][0] == 0) wk = 1. / (e1_to * e1_to * e2_to * e2_to);
    else if (fabs(L->ef[inew][0]) == L->ec1) wk = 1. / (2 * e1_to * e1_to * (e1_to * e1_to - e2_to * e2_to));
    else wk = 1. / (2 * e2_to * e2_to * (e2_to * e2_to - e1_to * e1_to));
    const double M4moment = 0;
    fnew = wk * M4moment; /* structs.cuh:365-368: NaN when wk is infinite */
    for (int iold = 0; iold < Qn; iold++) {
      const double ef_ = e_fr[iold], et = e_to[inew];
      double K0, K1, K2, K3, K4;
      K4 = 1;
      K3 = -4 * ef_ - et;
      if (inew == 0) {
        K2 = 6 * ef_ * ef_ - (e1_to * e1_to + e2_to * e2_to);
        K1 = 2 * ef_ * (e1_to * e1_to + e2_to * e2_to - 2 * ef_ * ef_);
        K0 = (e1_to * e1_to - ef_ * ef_) * (e2_to * e2_to - ef_ * ef_);
      } else {
        K2 = 6 * ef_ * ef_ + 3 * ef_ * et - (e1_to * e1_to + e2_to * e2_to - et * et);
        K1 = (2 * ef_ + et) * (e1_to * e1_to + e2_to * e2_to - et * et) - ef_ * ef_ * (4 * ef_ + 3 * et);
        K0 = ef_ * (ef_ * ef_ - (e1_to * e1_to + e2_to * e2_to) + et * et) * (ef_ + et);
      }
      const double difu = u1x - u0x;
      double dm4 = u0x + ef_;
      dm4 = dm4 * dm4 * dm4 * dm4;
      dm4 *= 0;
      const double gx = K4 * difu * difu * difu * difu + K3 * difu * difu * difu + K2 * difu * difu + K1 * difu + K0 - dm4;
      const double fadd = wk * gx * f[iold];
      if (!isnan(fadd) && !isinf(fadd)) fnew += fadd;
    }
  } else if (L->lattice == ORC_D3Q125) {
    const double S0 = sqrts(T0 / Tbase);
    const double S1 = sqrts(T1 / Tbase);
    const double e1_to = L->ec1 * S1;
    const double e2_to = L->ec2 * S1;
    double wdir[3];
    int idx[3];
    for (int d = 0; d < 3; d++) {
      const double en = L->ef[inew][d];
      double wv;
      if (en == 0) wv = 1. / (e1_to * e1_to * e2_to * e2_to);
      else if (fabs(en) == L->ec1) wv = 1. / (2 * e1_to * e1_to * (e1_to * e1_to - e2_to * e2_to));
      else wv = 1. / (2 * e2_to * e2_to * (e2_to * e2_to - e1_to * e1_to));
      wdir[d] = wv;
      int k = 0;
      if (en == +L->ec1) k = 1; else if (en == -L->ec1) k = 2; else if (en == +L->ec2) k = 3; else if (en == -L->ec2) k = 4;
      idx[d] = k;
    }
    wk = wdir[0]; wl = wdir[1]; wm = wdir[2];
    const double du[3] = {u0x - u1x, u0y - u1y, u0z - u1z};
    for (int iold = 0; iold < Qn; iold++) {
      double g[3];
      for (int d = 0; d < 3; d++) {
        const double efr = L->ef[iold][d] * S0;
        const double fact[5] = {du[d] + efr, du[d] + efr - e1_to, du[d] + efr + e1_to, du[d] + efr - e2_to, du[d] + efr + e2_to};
        const int X = idx[d];
        g[d] = fact[(X + 1) % 5] * fact[(X + 2) % 5] * fact[(X + 3) % 5] * fact[(X + 4) % 5];
      }
      const double gxyz = g[0] * g[1] * g[2];
      const double fadd = wk * wl * wm * gxyz * f[iold];
      if (!isnan(fadd) && !isinf(fadd)) fnew += fadd;
    }
  } else {
    if (L->e[inew][0] == 0) wk = -1. / T1;
    if (L->e[inew][1] == 0) wl = -1. / T1;
    if (L->e[inew][2] == 0) wm = -1. / T1;
    for (int iold = 0; iold < Qn; iold++) {
      const double Ax = (u1x - u0x) * sqrt(Tbase) - L->e[iold][0] * sqrt(fabs(T0)) * isign(T0);
      const double Ay = (u1y - u0y) * sqrt(Tbase) - L->e[iold][1] * sqrt(fabs(T0)) * isign(T0);
      const double Az = (u1z - u0z) * sqrt(Tbase) - L->e[iold][2] * sqrt(fabs(T0)) * isign(T0);
      double Bx = T1, By = T1, Bz = T1;
      if (L->e[inew][0] != 0) Bx = L->e[inew][0] * sqrt(fabs(T1)) * isign(T1) * Ax;
      if (L->e[inew][1] != 0) By = L->e[inew][1] * sqrt(fabs(T1)) * isign(T1) * Ay;
      if (L->e[inew][2] != 0) Bz = L->e[inew][2] * sqrt(fabs(T1)) * isign(T1) * Az;
      const double gx = Ax * Ax - Bx;
      const double gy = Ay * Ay - By;
      const double gz = Az * Az - Bz;
      const double fadd = wk * wl * wm * gx * gy * gz * f[iold];
      if (!isnan(fadd) && !isinf(fadd)) fnew += fadd;
    }
  }
  return fnew;
}

/* ---------------------------------------------------------------- interpolation */

/* tensor Lagrange basis on integer nodes 0..N-1, product in loop order (steps.cu:595-601) */
static double lagrpol(int N, int ix, int iy, int iz, double sx, double sy, double sz) {
  double a = 1;
  for (int p = 0; p < N; p++) if (p != ix) a *= (sx - p) / (ix - p);
  for (int p = 0; p < N; p++) if (p != iy) a *= (sy - p) / (iy - p);
  for (int p = 0; p < N; p++) if (p != iz) a *= (sz - p) / (iz - p);
  return a;
}

static long wrap(long v, long n) { return (v % n + n) % n; }

/* floor(v) as a stencil coordinate: the reference converts with make_int3(floor(...)), whose
   CUDA double->int conversion maps NaN to 0; values beyond +-2^30 cells (reachable only from
   non-finite or absurd velocities) are clamped, as the HIP kernels do. */
static int floor_int(double v) {
  const double f = floor(v);
  if (!(f == f)) return 0;
  if (f > 1073741824.0) return 1073741824;
  if (f < -1073741824.0) return -1073741824;
  return (int)f;
}

static const double* node(const orc_state* S, long x, long y, long z) {
  x = wrap(x, S->p.nx); y = wrap(y, S->p.ny); z = wrap(z, S->p.nz);
  return S->ld + (size_t)(x + S->p.nx * (y + S->p.ny * z)) * S->rec;
}

/* transfer of a stencil node: Cell::load (full record) then transfer_gauge */
static double node_transfer(const orc_state* S, long x, long y, long z, const double* base, int i) {
  const double* c = node(S, x, y, z);
  const int Qn = S->L.Qn;
  double f[QMAX], uT[4];
  for (int k = 0; k < Qn; k++) f[k] = c[k];
  for (int k = 0; k < 4; k++) uT[k] = c[Qn + 1 + k];
  return transfer(&S->L, f, uT, base, i);
}

/* A stencil node under mirror boundaries (propagate<N>, steps.cu:1004-1022): a coordinate
   below 0 or at or above N on a mirrored axis is reflected about that face (-1-p, or
   N-1-p+N), then wrapped; the node's cell is then mirrored on each such axis (Cell::mirrX/Y/Z,
   structs.cuh:552-578: the axis velocity negated, the populations with a non-zero velocity
   component on the axis exchanged with their mirror images), in the order x, y, z. */
static long mirror_coord(long p, long n, int* rev) {
  if (p < 0) { p = -1 - p; *rev = 1; }
  if (p >= n) { p = n - 1 - p + n; *rev = 1; }
  return p;
}

static void mirror_cell(const olat* L, int d, double* f, double* uT) {
  double fn[QMAX];
  uT[d] = -uT[d];
  for (int i = 0; i < L->Qn; i++) fn[i] = L->ef[i][d] == 0 ? f[i] : f[L->rev[d][i]];
  for (int i = 0; i < L->Qn; i++) f[i] = fn[i];
}

/* the populations and gauge of node (x, y, z) as the interpolation sees it */
static void node_mirrored(const orc_state* S, long x, long y, long z, double* f, double* uT) {
  int rv[3] = {0, 0, 0};
  long p[3] = {x, y, z};
  const long n[3] = {S->p.nx, S->p.ny, S->p.nz};
  for (int d = 0; d < 3; d++)
    if (S->p.mirror & (1 << d)) p[d] = mirror_coord(p[d], n[d], &rv[d]);
  const double* c = node(S, p[0], p[1], p[2]);
  const int Qn = S->L.Qn;
  for (int k = 0; k < Qn; k++) f[k] = c[k];
  for (int k = 0; k < 4; k++) uT[k] = c[Qn + 1 + k];
  for (int d = 0; d < 3; d++)
    if (rv[d]) mirror_cell(&S->L, d, f, uT);
}

static double node_transfer_mirrored(const orc_state* S, long x, long y, long z, const double* base, int i) {
  double f[QMAX], uT[4];
  node_mirrored(S, x, y, z, f, uT);
  return transfer(&S->L, f, uT, base, i);
}

/* LagrEval<3> (steps.cu:924-937) */
static double lagreval3(const orc_state* S, const double* crd, int lx, int ly, int lz, int i, const double* base) {
  const double sx = crd[0] - lx, sy = crd[1] - ly, sz = crd[2] - lz;
  double fpr = 0;
  for (int ix = 0; ix < 3; ix++)
    for (int iy = 0; iy < 3; iy++)
      for (int iz = 0; iz < 3; iz++)
        fpr += lagrpol(3, ix, iy, iz, sx, sy, sz) * node_transfer(S, (long)lx + ix, (long)ly + iy, (long)lz + iz, base, i);
  return fpr;
}

/* propagateWENO<3,5> (steps.cu:774-825), including its quirks: y/z sub-stencil nodes
   floor(y+0.5)+0..2 (extrapolating) and smoothness samples taken on the row
   (y,z) = (r mod Ny, r mod Nz) with r = floor(x+0.5) (steps.cu:783).
   centred != 0 selects the corrected variant that is not in the reference: y/z sub-stencil
   nodes floor(y+0.5)-1..+1 and smoothness samples on the departure point's own row
   (floor(y+0.5), floor(z+0.5)).  Everything else, including the operation order, is shared. */
static double interp_weno5_impl(const orc_state* S, const double* crd, int i, const double* base, int centred) {
  const int r = floor_int(crd[0] + 0.5);
  const int ry = floor_int(crd[1] + 0.5), rz = floor_int(crd[2] + 0.5);
  const int ly = centred ? ry - 1 : ry, lz = centred ? rz - 1 : rz;
  const double S0 = lagreval3(S, crd, r - 2, ly, lz, i, base);
  const double S1 = lagreval3(S, crd, r - 1, ly, lz, i, base);
  const double S2 = lagreval3(S, crd, r, ly, lz, i, base);
  const double cdist = crd[0] - r;
  const double d0 = (cdist * cdist - 3 * cdist + 2) / 12.;
  const double d2 = (cdist * cdist + 3 * cdist + 2) / 12.;
  const double d1 = 1. - d0 - d2;
  const long by = centred ? wrap(ry, S->p.ny) : wrap(r, S->p.ny);
  const long bz = centred ? wrap(rz, S->p.nz) : wrap(r, S->p.nz);
  const double u1 = node_transfer(S, (long)r - 2, by, bz, base, i);
  const double u2 = node_transfer(S, (long)r - 1, by, bz, base, i);
  const double u3 = node_transfer(S, (long)r + 0, by, bz, base, i);
  const double u4 = node_transfer(S, (long)r + 1, by, bz, base, i);
  const double u5 = node_transfer(S, (long)r + 2, by, bz, base, i);
  const double beta0 = (10 * u3 * u3 + (11 * u1 - 31 * u2) * u3 + 25 * u2 * u2 - 19 * u1 * u2 + 4 * u1 * u1) / 3.0;
  const double beta1 = (4 * u4 * u4 + (5 * u2 - 13 * u3) * u4 + 13 * u3 * u3 - 13 * u2 * u3 + 4 * u2 * u2) / 3.0;
  const double beta2 = (4 * u5 * u5 + (11 * u3 - 19 * u4) * u5 + 25 * u4 * u4 - 31 * u3 * u4 + 10 * u3 * u3) / 3.0;
  double w0 = d0 * 1, w1 = d1 * 1, w2 = d2 * 1;
  const double sc = 1e-6;
  w0 /= (sc + beta0) * (sc + beta0);
  w1 /= (sc + beta1) * (sc + beta1);
  w2 /= (sc + beta2) * (sc + beta2);
  const double wsum = w0 + w1 + w2;
  w0 /= wsum; w1 /= wsum; w2 /= wsum;
  return w0 * S0 + w1 * S1 + w2 * S2;
}

/* the generic propagateWENO<5,M> (steps.cu:610-654) with propagate3 (steps.cu:909-923; the
   same sum as LagrEval<3>).  lshift = 1 (steps.cu:615), so S0, S1, S2 have x-left nodes r, r-1,
   r-2: the mirror image of propagateWENO<3,5>'s numbering, with the linear weight
   (c^2+3c+2)/12 on S0.  y/z nodes floor(y+0.5)+0..2 and the smoothness row (r mod Ny, r mod Nz)
   as in propagateWENO<3,5> (steps.cu:616-624).  The indicators are written expanded
   (steps.cu:642-644) and each weight is d / (1e-6 + beta)^2 in one expression (645-647).  It is
   the same interpolant as propagateWENO<3,5>; the operation order, and so the rounding,
   differs. */
static double interp_weno5g(const orc_state* S, const double* crd, int i, const double* base) {
  const int lshift = 1;
  const int r = floor_int(crd[0] + 0.5);
  const int ly = floor_int(crd[1] + 0.5), lz = floor_int(crd[2] + 0.5);
  const double S0 = lagreval3(S, crd, r - lshift + 1, ly, lz, i, base);
  const double S1 = lagreval3(S, crd, r - lshift, ly, lz, i, base);
  const double S2 = lagreval3(S, crd, r - lshift - 1, ly, lz, i, base);
  const double cdist = crd[0] - r;
  const double d0 = (cdist * cdist + 3 * cdist + 2) / 12.;
  const double d2 = (cdist * cdist - 3 * cdist + 2) / 12.;
  const double d1 = 1. - d0 - d2;
  /* crd_left = (r - 2, r, r): the samples sit on row y = r, z = r (wrapped) */
  const long by = wrap(r, S->p.ny), bz = wrap(r, S->p.nz);
  const double u1 = node_transfer(S, (long)r - lshift - 1, by, bz, base, i);
  const double u2 = node_transfer(S, (long)r - lshift, by, bz, base, i);
  const double u3 = node_transfer(S, (long)r - lshift + 1, by, bz, base, i);
  const double u4 = node_transfer(S, (long)r - lshift + 2, by, bz, base, i);
  const double u5 = node_transfer(S, (long)r - lshift + 3, by, bz, base, i);
  const double beta0 = (10 * u3 * u3 - 31 * u3 * u4 + 25 * u4 * u4 + 11 * u3 * u5 - 19 * u4 * u5 + 4 * u5 * u5) / 3.0;
  const double beta1 = (4 * u2 * u2 - 13 * u2 * u3 + 13 * u3 * u3 + 5 * u2 * u4 - 13 * u3 * u4 + 4 * u4 * u4) / 3.0;
  const double beta2 = (4 * u1 * u1 - 19 * u1 * u2 + 25 * u2 * u2 + 11 * u1 * u3 - 31 * u2 * u3 + 10 * u3 * u3) / 3.0;
  double w0 = d0 / ((1e-6 + beta0) * (1e-6 + beta0));
  double w1 = d1 / ((1e-6 + beta1) * (1e-6 + beta1));
  double w2 = d2 / ((1e-6 + beta2) * (1e-6 + beta2));
  const double wsum = w0 + w1 + w2;
  w0 /= wsum; w1 /= wsum; w2 /= wsum;
  return w0 * S0 + w1 * S1 + w2 * S2;
}

/* propagate<4> (steps.cu:986-1036): tricubic Lagrange, nodes floor(x)-1 .. floor(x)+2 */
static double interp_lagr4(const orc_state* S, const double* crd, int i, const double* base) {
  const int lx = floor_int(crd[0]) - 4 / 2 + 1;
  const int ly = floor_int(crd[1]) - 4 / 2 + 1;
  const int lz = floor_int(crd[2]) - 4 / 2 + 1;
  const double sx = crd[0] - lx, sy = crd[1] - ly, sz = crd[2] - lz;
  double fpr = 0;
  for (int ix = 0; ix < 4; ix++)
    for (int iy = 0; iy < 4; iy++)
      for (int iz = 0; iz < 4; iz++)
        fpr += lagrpol(4, ix, iy, iz, sx, sy, sz) *
               (S->p.mirror ? node_transfer_mirrored(S, (long)lx + ix, (long)ly + iy, (long)lz + iz, base, i)
                            : node_transfer(S, (long)lx + ix, (long)ly + iy, (long)lz + iz, base, i));
  return fpr;
}

static double interp_any(const orc_state* S, const double* crd, int i, const double* base) {
  if (S->p.interp == ORC_LAGR4) return interp_lagr4(S, crd, i, base);
  if (S->p.interp == ORC_WENO5G) return interp_weno5g(S, crd, i, base);
  return interp_weno5_impl(S, crd, i, base, S->p.interp == ORC_WENO5C);
}

/* ---------------------------------------------------------------- per-cell pieces */

typedef struct { double f[QMAX]; double rho; double uT[4]; } ocell;

/* Cell::calcMoments (structs.cuh:315-342): serial sums in population order, division by rho */
static void moments(const olat* L, ocell* c, double (*vi)[3]) {
  double rho = 0, u0 = 0, u1 = 0, u2 = 0, u3 = 0;
  for (int i = 0; i < L->Qn; i++) {
    rho += c->f[i];
    u0 += c->f[i] * vi[i][0];
    u1 += c->f[i] * vi[i][1];
    u2 += c->f[i] * vi[i][2];
    u3 += c->f[i] * (vi[i][0] * vi[i][0] + vi[i][1] * vi[i][1] + vi[i][2] * vi[i][2]);
  }
  if (rho != 0) { u0 /= rho; u1 /= rho; u2 /= rho; u3 /= rho; }
  u3 = (u3 - u0 * u0 - u1 * u1 - u2 * u2) * (1. / L->Dim);
  c->rho = rho;
  c->uT[0] = u0; c->uT[1] = u1; c->uT[2] = u2; c->uT[3] = u3;
}

/* isConv (steps.cu:58-75) */
static int is_conv(const orc_state* S, const ocell* cn, const ocell* co) {
  const double ea = S->p.tol_abs, er = S->p.tol_rel;
  for (int k = 0; k < 4; k++)
    if (fabs(cn->uT[k] - co->uT[k]) >= ea + er * fabs(cn->uT[k])) return 0;
  for (int i = 0; i < S->L.Qn; i++)
    if (fabs(cn->f[i] - co->f[i]) >= ea + er * fabs(cn->f[i])) return 0;
  return 1;
}

static void load_cell(const orc_state* S, const double* buf, long idx, ocell* c) {
  const double* r = buf + (size_t)idx * S->rec;
  const int Qn = S->L.Qn;
  for (int i = 0; i < Qn; i++) c->f[i] = r[i];
  c->rho = r[Qn];
  for (int k = 0; k < 4; k++) c->uT[k] = r[Qn + 1 + k];
}

static void save_cell(const orc_state* S, double* buf, long idx, const ocell* c) {
  double* r = buf + (size_t)idx * S->rec;
  const int Qn = S->L.Qn;
  for (int i = 0; i < Qn; i++) r[i] = c->f[i];
  r[Qn] = c->rho;
  for (int k = 0; k < 4; k++) r[Qn + 1 + k] = c->uT[k];
}

/* Cell::set_base (lbpack.cu:15-21): f_i = w_i*rho */
static void set_base(const olat* L, ocell* c, double rho, const double* uT) {
  c->rho = rho;
  for (int k = 0; k < 4; k++) c->uT[k] = uT[k];
  for (int i = 0; i < L->Qn; i++) c->f[i] = L->w[i] * rho;
}

static void velocities(const olat* L, const double* uT, double (*vi)[3]) {
  const double s = sqrt(uT[3] / L->TLat);
  for (int i = 0; i < L->Qn; i++)
    for (int d = 0; d < 3; d++) vi[i][d] = s * L->ef[i][d] + uT[d];
}

static void zero_reset(const olat* L, ocell* c) {
  if (c->rho == 0 && c->uT[0] == 0 && c->uT[1] == 0 && c->uT[2] == 0 && c->uT[3] == 0) {
    const double g[4] = {0, 0, 0, L->TLat};
    set_base(L, c, 1, g);
  }
}

static void collide(const orc_state* S, ocell* c) {
  const double dtau = S->p.dtau;
  for (int i = 0; i < S->L.Qn; i++) {
    const double feq = c->rho * S->L.w[i];
    c->f[i] += dtau * (feq - c->f[i]);
  }
}

/* LBMpull for D1Q5 / D3Q27 (steps.cu:356-365, 421-512, 557-564, 582) */
static void cell_step(orc_state* S, long x, long y, long z) {
  const olat* L = &S->L;
  const long idx = x + S->p.nx * (y + S->p.ny * z);
  /* a switched-off cell is not stepped: its store record and iteration count keep what they
     held (LBMpull, steps.cu:359) */
  if (!S->set[idx]) return;
  ocell c0, cnew;
  load_cell(S, S->ld, idx, &c0);
  zero_reset(L, &c0);
  double vi[QMAX][3];
  int niter = 0;
  while (niter < S->p.max_iter) {
    niter++;
    double uT[4] = {c0.uT[0], c0.uT[1], c0.uT[2], c0.uT[3]};
    velocities(L, uT, vi);
    for (int i = 0; i < L->Qn; i++) {
      const double crd[3] = {(double)x - vi[i][0], (double)y - vi[i][1], (double)z - vi[i][2]};
      cnew.f[i] = interp_any(S, crd, i, uT);
    }
    moments(L, &cnew, vi);
    if (cnew.uT[3] < 0) cnew.uT[3] = c0.uT[3];
    const int conv = is_conv(S, &cnew, &c0);
    c0 = cnew;
    if (conv) break;
  }
  collide(S, &c0);
  save_cell(S, S->st, idx, &c0);
  S->niter[idx] = niter;
}

/* PonDstepD3Q125 (steps.cu:151-250): fixed 3x3x3 stencil coord-1..coord+1 */
static void cell_step_q125(orc_state* S, long x, long y, long z) {
  const olat* L = &S->L;
  const int Qn = L->Qn;
  const long idx = x + S->p.nx * (y + S->p.ny * z);
  if (!S->set[idx]) return; /* steps.cu:156 */
  ocell c0, cnew;
  load_cell(S, S->ld, idx, &c0);
  zero_reset(L, &c0);
  /* cfix: the 3x3x3 neighbourhood, mirrored at mirror boundaries (steps.cu:163-190) */
  static __thread double fixf[27][QMAX], fixg[27][4];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int c = 0; c < 3; c++) node_mirrored(S, x - 1 + a, y - 1 + b, z - 1 + c, fixf[9 * a + 3 * b + c], fixg[9 * a + 3 * b + c]);
  double vi[QMAX][3];
  int niter = 0;
  while (niter < S->p.max_iter) {
    niter++;
    double uT[4] = {c0.uT[0], c0.uT[1], c0.uT[2], c0.uT[3]};
    velocities(L, uT, vi);
    for (int inew = 0; inew < Qn; inew++) {
      double nv[27];
      for (int n = 0; n < 27; n++) {
        nv[n] = transfer(L, fixf[n], fixg[n], uT, inew);
      }
      const double crd[3] = {(double)x - vi[inew][0], (double)y - vi[inew][1], (double)z - vi[inew][2]};
      const double sh[3] = {crd[0] - (double)(x - 1), crd[1] - (double)(y - 1), crd[2] - (double)(z - 1)};
      double fpr = 0;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          for (int c = 0; c < 3; c++) fpr += lagrpol(3, a, b, c, sh[0], sh[1], sh[2]) * nv[9 * a + 3 * b + c];
      cnew.f[inew] = fpr;
    }
    moments(L, &cnew, vi);
    if (cnew.uT[3] < 0) cnew.uT[3] = c0.uT[3];
    const int conv = is_conv(S, &cnew, &c0);
    c0 = cnew;
    if (conv) break;
  }
  collide(S, &c0);
  save_cell(S, S->st, idx, &c0);
  S->niter[idx] = niter;
}

/* ---------------------------------------------------------------- exported API */

int orc_qn(int lattice) { olat L; lat_init(&L, lattice); return L.Qn; }

void orc_lattice(int lattice, int* e, double* ef, double* w, double* tlat_dim) {
  olat L; lat_init(&L, lattice);
  for (int i = 0; i < L.Qn; i++) {
    for (int d = 0; d < 3; d++) { e[3 * i + d] = L.e[i][d]; ef[3 * i + d] = L.ef[i][d]; }
    w[i] = L.w[i];
  }
  tlat_dim[0] = L.TLat; tlat_dim[1] = L.Dim;
}

/* reverseX/Y/Z tables of a lattice: rev[d*Qn + i] */
void orc_reverse(int lattice, int* rev) {
  olat L; lat_init(&L, lattice);
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < L.Qn; i++) rev[d * L.Qn + i] = L.rev[d][i];
}

double orc_transfer(int lattice, const double* f, const double* uT0, const double* base, int inew) {
  olat L; lat_init(&L, lattice);
  return transfer(&L, f, uT0, base, inew);
}

double orc_lagrpol(int N, int ix, int iy, int iz, double sx, double sy, double sz) {
  return lagrpol(N, ix, iy, iz, sx, sy, sz);
}

/* moments of populations f taken in gauge g (velocities v_i = sqrt(T/TLat) e_i + u) */
void orc_moments(int lattice, const double* f, const double* g, double* out5) {
  olat L; lat_init(&L, lattice);
  ocell c; double vi[QMAX][3];
  for (int i = 0; i < L.Qn; i++) c.f[i] = f[i];
  velocities(&L, g, vi);
  moments(&L, &c, vi);
  out5[0] = c.rho;
  for (int k = 0; k < 4; k++) out5[1 + k] = c.uT[k];
}

void* orc_create(const orc_params* p) {
  orc_state* S = (orc_state*)calloc(1, sizeof(orc_state));
  if (!S) return NULL;
  S->p = *p;
  lat_init(&S->L, p->lattice);
  S->ncell = p->nx * p->ny * p->nz;
  S->rec = S->L.Qn + 5;
  S->ld = (double*)calloc((size_t)S->ncell * S->rec, sizeof(double));
  S->st = (double*)calloc((size_t)S->ncell * S->rec, sizeof(double));
  S->niter = (int*)calloc((size_t)S->ncell, sizeof(int));
  S->set = (int*)calloc((size_t)S->ncell, sizeof(int));
  S->setnew = (int*)calloc((size_t)S->ncell, sizeof(int));
  if (!S->ld || !S->st || !S->niter || !S->set || !S->setnew) {
    free(S->ld); free(S->st); free(S->niter); free(S->set); free(S->setnew); free(S);
    return NULL;
  }
  for (long n = 0; n < S->ncell; n++) S->set[n] = 1; /* fill() activates every cell (init.cu:106) */
  return S;
}

void orc_destroy(void* h) {
  orc_state* S = (orc_state*)h;
  if (!S) return;
  free(S->ld); free(S->st); free(S->niter); free(S->set); free(S->setnew); free(S);
}

/* init()/fill() analogue (init.cu:96-106): set_base of the given fields into both buffers */
void orc_set_state(void* h, const double* rho, const double* ux, const double* uy, const double* uz, const double* T) {
  orc_state* S = (orc_state*)h;
  for (long n = 0; n < S->ncell; n++) {
    ocell c;
    const double g[4] = {ux[n], uy[n], uz[n], T[n]};
    set_base(&S->L, &c, rho[n], g);
    save_cell(S, S->ld, n, &c);
    save_cell(S, S->st, n, &c);
    S->niter[n] = 0;
  }
}

/* full state: Qn population planes (plane-major) plus the five moment planes */
void orc_set_populations(void* h, const double* f, const double* rho, const double* ux, const double* uy,
                         const double* uz, const double* T) {
  orc_state* S = (orc_state*)h;
  const int Qn = S->L.Qn;
  for (long n = 0; n < S->ncell; n++) {
    ocell c;
    for (int i = 0; i < Qn; i++) c.f[i] = f[(size_t)i * S->ncell + n];
    c.rho = rho[n]; c.uT[0] = ux[n]; c.uT[1] = uy[n]; c.uT[2] = uz[n]; c.uT[3] = T[n];
    save_cell(S, S->ld, n, &c);
    save_cell(S, S->st, n, &c);
  }
}

/* ---------------------------------------------------------------- 1D cell activation (amr.cu) */

static double rec_rho(const orc_state* S, const double* buf, long idx) { return buf[(size_t)idx * S->rec + S->L.Qn]; }
static double rec_u(const orc_state* S, const double* buf, long idx, int k) {
  return buf[(size_t)idx * S->rec + S->L.Qn + 1 + k];
}

/* AMR_add (amr.cu:1-27), for one active cell: the next active cell along x (wrapping); where
   their new states differ (|du_x|^2 / min(T) > 0.01 or |1 - rho_m/rho_p| > 0.01) and the gap is
   even, the cell half way receives the average of the two records in both buffers (save_12)
   and is switched on from the next step (setnew).  Reads only active cells' store records and
   `set`, writes only cells strictly inside the gap, so the cells can be taken in any order. */
static void amr_add_cell(orc_state* S, long x, long y, long z) {
  const long nx = S->p.nx;
  const long row = S->p.nx * (y + S->p.ny * z);
  if (!S->set[row + x]) return;
  long xp = x, dist = 0;
  do {
    xp = (xp + 1) % nx;
    dist++;
  } while (!S->set[row + xp]);
  const long im = row + x, ip = row + xp;
  const double vdif = rec_u(S, S->st, ip, 0) - rec_u(S, S->st, im, 0);
  const double Tm = rec_u(S, S->st, im, 3), Tp = rec_u(S, S->st, ip, 3);
  if (vdif * vdif / fmin(Tm, Tp) > 0.01 || fabs(1 - rec_rho(S, S->st, im) / rec_rho(S, S->st, ip)) > 0.01) {
    if (dist % 2 == 0) {
      const long imid = row + (x + dist / 2) % nx;
      ocell cm, cp, c0;
      load_cell(S, S->st, im, &cm);
      load_cell(S, S->st, ip, &cp);
      c0.rho = 0.5 * (cm.rho + cp.rho);
      for (int k = 0; k < 4; k++) c0.uT[k] = 0.5 * (cm.uT[k] + cp.uT[k]);
      for (int i = 0; i < S->L.Qn; i++) c0.f[i] = 0.5 * (cm.f[i] + cp.f[i]);
      save_cell(S, S->ld, imid, &c0);
      save_cell(S, S->st, imid, &c0);
      S->setnew[imid] = 1;
    }
  }
}

/* AMR_remove (amr.cu:28-55): one pass along the row y = z = 0 in increasing x, each decision
   seeing the ones before it: an active cell whose previous and next active cells lie within 3
   cells and match it (|du_x|^2 / T < 0.01, |rho ratio - 1| < 0.01 both ways, its own T and rho)
   is switched off. */
static void amr_remove_row(orc_state* S) {
  const long nx = S->p.nx;
  for (long x = 0; x < nx; x++) {
    if (!S->set[x]) continue;
    long xm = x, xp = x, distM = 0, distP = 0;
    do { xm = (xm - 1 + nx) % nx; distM++; } while (!S->set[xm]);
    do { xp = (xp + 1) % nx; distP++; } while (!S->set[xp]);
    const double u0 = rec_u(S, S->st, x, 0), T0 = rec_u(S, S->st, x, 3), r0 = rec_rho(S, S->st, x);
    const double vdifm = rec_u(S, S->st, xm, 0) - u0;
    const double vdifp = rec_u(S, S->st, xp, 0) - u0;
    if (vdifm * vdifm / T0 < 0.01 && vdifp * vdifp / T0 < 0.01 && fabs(rec_rho(S, S->st, xp) / r0 - 1) < 0.01 &&
        fabs(rec_rho(S, S->st, xm) / r0 - 1) < 0.01 && (distM < 4 && distP < 4))
      S->set[x] = 0;
  }
}

/* calcStep() without I/O: one PonD step per iteration, then (with amr) the activation passes of
   update.cu:102-105 as they stand with their launches enabled, then swap load/store
   (update.cu:86-142, structs.cuh:148) */
void orc_step(void* h, int nsteps) {
  orc_state* S = (orc_state*)h;
  const long nx = S->p.nx, ny = S->p.ny, nz = S->p.nz;
  for (int s = 0; s < nsteps; s++) {
#pragma omp parallel for schedule(dynamic, 16)
    for (long n = 0; n < S->ncell; n++) {
      const long x = n % nx, y = (n / nx) % ny, z = n / (nx * ny);
      if (S->L.lattice == ORC_D3Q125) cell_step_q125(S, x, y, z);
      else cell_step(S, x, y, z);
    }
    (void)nz;
    if (S->p.amr) {
      for (long n = 0; n < S->ncell; n++) S->setnew[n] = S->set[n]; /* cinfo_setnew_reset */
      for (long n = 0; n < S->ncell; n++) amr_add_cell(S, n % nx, (n / nx) % ny, n / (nx * ny));
      for (long n = 0; n < S->ncell; n++) S->set[n] = S->setnew[n]; /* cinfo_set_new */
      amr_remove_row(S);
    }
    double* t = S->ld; S->ld = S->st; S->st = t;
  }
}

/* Cinfo::set of every cell */
void orc_get_active(void* h, int* set) {
  orc_state* S = (orc_state*)h;
  for (long n = 0; n < S->ncell; n++) set[n] = S->set[n];
}

void orc_set_active(void* h, const int* set) {
  orc_state* S = (orc_state*)h;
  for (long n = 0; n < S->ncell; n++) S->set[n] = set[n] != 0;
}

/* fields of the current state (the last step's output, i.e. the load buffer after the swap) */
void orc_get_fields(void* h, double* rho, double* ux, double* uy, double* uz, double* T, int* niter) {
  orc_state* S = (orc_state*)h;
  const int Qn = S->L.Qn;
  for (long n = 0; n < S->ncell; n++) {
    const double* r = S->ld + (size_t)n * S->rec;
    if (rho) rho[n] = r[Qn];
    if (ux) ux[n] = r[Qn + 1];
    if (uy) uy[n] = r[Qn + 2];
    if (uz) uz[n] = r[Qn + 3];
    if (T) T[n] = r[Qn + 4];
    if (niter) niter[n] = S->niter[n];
  }
}

void orc_get_populations(void* h, double* f) {
  orc_state* S = (orc_state*)h;
  const int Qn = S->L.Qn;
  for (long n = 0; n < S->ncell; n++)
    for (int i = 0; i < Qn; i++) f[(size_t)i * S->ncell + n] = S->ld[(size_t)n * S->rec + i];
}

/* one interpolated population of the current state at departure point crd, in gauge g
   (the value LBMpull stores in cnew.f[i]) */
double orc_interp(void* h, int i, const double* crd, const double* g) {
  orc_state* S = (orc_state*)h;
  return interp_any(S, crd, i, g);
}

/* calcTotMoments (update.cu:52-69, active cells), summed serially in cell order; diagn k
   (diagn.cu:3-22, every cell) */
void orc_totals(void* h, double* out4) {
  orc_state* S = (orc_state*)h;
  const int Qn = S->L.Qn;
  const double Dim = S->L.Dim;
  double mass = 0, mom = 0, en = 0, k = 0;
  const double d8pi3 = 1.0 / (8 * M_PI * M_PI * M_PI);
  const double dx = 2 * M_PI / S->p.nx, dy = 2 * M_PI / S->p.ny, dz = 2 * M_PI / S->p.nz;
  for (long n = 0; n < S->ncell; n++) {
    const double* r = S->ld + (size_t)n * S->rec;
    const double rho = r[Qn], ux = r[Qn + 1], uy = r[Qn + 2], uz = r[Qn + 3], T = r[Qn + 4];
    if (S->set[n]) {
      mass += rho * 1.0;
      mom += rho * ux * 1.0;
      en += rho * (ux * ux + uy * uy + uz * uz + Dim * T) * 1.0;
    }
    k += 0.5 * d8pi3 * (ux * ux + uy * uy + uz * uz) * dx * dy * dz;
  }
  out4[0] = mass; out4[1] = mom; out4[2] = en; out4[3] = k;
}

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
