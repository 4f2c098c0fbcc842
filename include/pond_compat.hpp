/*
 * pond_compat.hpp -- C++ host facade with the reference's names, over the C ABI of pond.h.
 *
 * A host program written against the reference's driver keeps its calls:
 *
 *   reference                                           here (namespace lbmpond)
 *   --------------------------------------------------  ----------------------------------------
 *   struct PhysPars { dtau; set(); }  structs.cuh:96-99  PhysPars (same fields)
 *   struct LBMParams / LBMParamsHost  structs.cuh:100-149 LBMParams / LBMParamsHost: Nx, Ny, Nz,
 *     (Nx, Ny, Nz, iStep, curtime, dt, NiterMax, mass,     iStep, curtime, dt, pc, and the values
 *      enrg, moment, pc; set/reset/clear/swapLS)           behind the managed pointers NiterMax,
 *                                                          mass, enrg, moment (host values here)
 *   LBMParamsHost parsHost            lbm.cu:14          parsHost (one inline global)
 *   LBMParamsHost::set()              lbm.cu:134-168     set(): sizes kept, dt = 1, pc.set()
 *   init() + fill<<<>>>               init.cu:2-128      init(): allocates the state and loads the
 *                                                        initial condition from parsHost.fill, or
 *                                                        init(rho, ux, uy, uz, T) from arrays
 *   calcStep()                        update.cu:86-142   calcStep(): one step, NiterMax, mass,
 *                                                        moment, enrg, curtime, iStep; with drop_dir
 *                                                        set it drops every step, and the files named
 *                                                        N hold the state after N steps, as the
 *                                                        reference's (which drops the step's input)
 *   drop(std::string*)                drop.cu:2-34       drop(std::string*): same .arr files
 *   diagnostics()                     diagn.cu:24-30     diagnostics(): k (printed as "k= %g")
 *   Cinfo                             structs.cuh:89-94  Cinfo, get_cinfo(): niter of the last step
 *   e / ef / w / TLat / Dim tables    consts.h, lbm.cu:16-22  lattice_tables()
 *
 * The reference picks lattice and interpolator at compile time (consts.h:6-12, steps.cu:451/458);
 * here they are fields of LBMParamsHost (lattice, interp, precision) read by init().  Errors
 * throw lbmpond::Error carrying pond_last_error() instead of exit() (params.h:74-81).
 * Header only; link with libpond.so.  Not thread-safe, like the reference's globals.
 */
#ifndef POND_COMPAT_HPP
#define POND_COMPAT_HPP

#include <cstdint>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "pond.h"

namespace lbmpond {

struct Error : std::runtime_error {
  pond_status status;
  Error(pond_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(pond_status s, const pond_ctx* ctx, const char* what) {
  if (s != POND_OK) throw Error(s, std::string(what) + ": " + pond_last_error(ctx));
}

struct double3v { double x, y, z; };

struct PhysPars {
  double dtau;
  void set() { dtau = 1.0 / 0.9; }
};

struct Cinfo {
  int set, niter, setnew, ctype;
};

struct LBMParams {
  long int Nx = 1, Ny = 1, Nz = 1;
  int iStep = 0;
  double curtime = 0;
  double dt = 1;
  /* values the reference keeps in managed memory (lbm.cu:161-164), filled by calcStep */
  int NiterMax = 0;
  double mass = 0, enrg = 0;
  double3v moment = {0, 0, 0};
  PhysPars pc = {1.0 / 0.9};
};

/* the initial condition of one cell: the reference's fill kernel (init.cu:2-108) evaluated per
   cell on the host; T is the temperature of Cell::uT.w */
using FillFunc = std::function<void(long x, long y, long z, double& rho, double& ux, double& uy, double& uz,
                                    double& T)>;

struct LBMParamsHost : LBMParams {
  int lattice = POND_D3Q27;
  int interp = POND_LAGR4;   /* the reference's active propagate<4> (steps.cu:451) */
  int precision = POND_FP64;
  int max_iter = 100;
  double k = 0;              /* diagn's kinetic-energy sum (diagn.cu:2) */
  std::string drop_dir;      /* im3DHost.drop_dir: calcStep drops each step when set */
  FillFunc fill;
  pond_ctx* ctx = nullptr;

  void set() {
    dt = 1;
    pc.set();
  }
  /* the reference clears Cinfo (init.cu:130-132); here the per-cell counts restart with the
     next initial condition */
  void reset() {}
  void clear() {
    if (ctx) pond_destroy(ctx);
    ctx = nullptr;
  }
  /* ping-pong buffers are swapped inside pond_step; kept for source compatibility */
  void swapLS() {}
  long int cells() const { return Nx * Ny * Nz; }
};

inline LBMParamsHost parsHost;

inline void init(const double* rho, const double* ux, const double* uy, const double* uz, const double* T) {
  parsHost.clear();
  pond_params p;
  pond_params_default(&p, parsHost.lattice);
  p.interp = parsHost.interp;
  p.precision = parsHost.precision;
  p.nx = parsHost.Nx;
  p.ny = parsHost.Ny;
  p.nz = parsHost.Nz;
  p.dtau = parsHost.pc.dtau;
  p.max_iter = parsHost.max_iter;
  check(pond_create(&p, &parsHost.ctx), nullptr, "init");
  check(pond_set_state(parsHost.ctx, rho, ux, uy, uz, T), parsHost.ctx, "init");
  parsHost.iStep = 0;
  parsHost.curtime = 0;
}

inline void init() {
  if (!parsHost.fill) throw Error(POND_ERR_STATE, "init: parsHost.fill is not set");
  const size_t n = (size_t)parsHost.cells();
  std::vector<double> rho(n), ux(n), uy(n), uz(n), T(n);
  for (long z = 0; z < parsHost.Nz; z++)
    for (long y = 0; y < parsHost.Ny; y++)
      for (long x = 0; x < parsHost.Nx; x++) {
        const size_t c = (size_t)x + (size_t)parsHost.Nx * ((size_t)y + (size_t)parsHost.Ny * (size_t)z);
        parsHost.fill(x, y, z, rho[c], ux[c], uy[c], uz[c], T[c]);
      }
  init(rho.data(), ux.data(), uy.data(), uz.data(), T.data());
}

inline void drop(std::string* fprefix) {
  check(pond_drop(parsHost.ctx, fprefix->c_str()), parsHost.ctx, "drop");
}

inline void diagnostics() {
  pond_diag d;
  check(pond_diagnostics(parsHost.ctx, &d), parsHost.ctx, "diagnostics");
  parsHost.k = d.k;
  std::printf("k= %g\n", d.k);
}

inline void calcStep() {
  if (!parsHost.ctx) throw Error(POND_ERR_STATE, "calcStep before init");
  check(pond_step(parsHost.ctx, 1), parsHost.ctx, "calcStep: pond_step");
  pond_diag d;
  check(pond_diagnostics(parsHost.ctx, &d), parsHost.ctx, "calcStep: pond_diagnostics");
  if (!parsHost.drop_dir.empty()) drop(&parsHost.drop_dir);
  parsHost.NiterMax = d.niter_max;
  parsHost.mass = d.mass;
  parsHost.moment = {d.momentum_x, 0, 0};
  parsHost.enrg = d.energy;
  parsHost.k = d.k;
  parsHost.curtime += parsHost.dt;
  parsHost.iStep++;
}

inline void get_cinfo(std::vector<Cinfo>& out) {
  const size_t n = (size_t)parsHost.cells();
  std::vector<int32_t> it(n);
  check(pond_get_fields(parsHost.ctx, nullptr, nullptr, nullptr, nullptr, nullptr, it.data()), parsHost.ctx,
        "get_cinfo");
  out.resize(n);
  for (size_t c = 0; c < n; c++) out[c] = Cinfo{1, it[c], 0, 0};
}

struct LatticeTables {
  std::vector<int32_t> e;   /* Qn x 3 */
  std::vector<double> ef;   /* Qn x 3 */
  std::vector<double> w;    /* Qn */
  double TLat = 0, Dim = 0;
  int Qn = 0;
};

inline LatticeTables lattice_tables(int lattice) {
  LatticeTables t;
  t.Qn = pond_qn(lattice);
  if (t.Qn == 0) throw Error(POND_ERR_INVALID, "lattice_tables: unknown lattice");
  t.e.resize((size_t)t.Qn * 3);
  t.ef.resize((size_t)t.Qn * 3);
  t.w.resize((size_t)t.Qn);
  double td[2];
  check(pond_lattice_tables(lattice, t.e.data(), t.ef.data(), t.w.data(), td), nullptr, "lattice_tables");
  t.TLat = td[0];
  t.Dim = td[1];
  return t;
}

}  // namespace lbmpond

#endif /* POND_COMPAT_HPP */
