/*
 * pond.h -- C ABI of the MI355X-native particles-on-demand (PonD) step.
 *
 * This is the drop-in boundary for the reference's step API (engrab6/lbmPonD):
 *
 *   reference                                         here
 *   ------------------------------------------------  ---------------------------------------
 *   LBMParamsHost::set / PhysPars::set                pond_params_default, pond_create
 *     (lbm.cu:134-168, structs.cuh:96-149)
 *   init() + fill<<<>>>  (init.cu:96-128)             pond_set_state (set_base per cell)
 *   copy2dev / Group::pack of a full state            pond_set_populations
 *     (params.h:136-146, structs.cuh:269-288)
 *   calcStep()  (update.cu:86-142, minus I/O)         pond_step
 *   PonDstep / PonDstepD3Q125 kernels                 (launched inside pond_step)
 *     (update.cu:11-14, steps.cu:151-250)
 *   drop(std::string*)  (drop.cu:2-34)                pond_drop (same .arr format)
 *   diagnostics() + calcTotMoments + NiterMax         pond_diagnostics
 *     (diagn.cu:24-30, update.cu:52-69,124-135, steps.cu:513)
 *   reading loadGroups back (drop.cu:3-4)             pond_get_fields, pond_get_populations
 *   CHECK_ERROR -> exit  (params.h:74-81)             pond_status codes + pond_last_error
 *   global parsHost / pars singletons (lbm.cu:14-15)  opaque pond_ctx, one per simulation
 *
 * All arrays are host arrays in the reference's cell order: linear index x + Nx*(y + Ny*z)
 * (structs.cuh:153-166).  Population arrays are plane-major: f[i*N + cell].
 * A context owns its device memory and one HIP stream; it is not thread-safe.
 */
#ifndef POND_H
#define POND_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  POND_OK = 0,
  POND_ERR_INVALID = 1, /* bad argument or unsupported combination */
  POND_ERR_HIP = 2,     /* HIP runtime error (launch, memcpy, fault) */
  POND_ERR_NOMEM = 3,   /* device or host allocation failed */
  POND_ERR_IO = 4,      /* file I/O failed (pond_drop, pond_read_arr) */
  POND_ERR_STATE = 5    /* call out of order (e.g. step before any state was set) */
} pond_status;

/* lattices: consts.h:111-134 (D1Q5), 14-65 (D3Q27), 136-245 (D3Q125) */
typedef enum { POND_D1Q5 = 0, POND_D3Q27 = 1, POND_D3Q125 = 2 } pond_lattice;

/* interpolators for D1Q5/D3Q27; D3Q125 always uses its fixed 3x3x3 Lagrange stencil
   (steps.cu:204-212).
   POND_WENO5  propagateWENO<3,5> as written (steps.cu:458, 774-825), including its y/z
               sub-stencil at floor(y+0.5)+0..2 and smoothness row (r mod Ny, r mod Nz).  The
               y/z stencil extrapolates downwind, so in 3D grid-scale modes grow every step.
   POND_LAGR4  propagate<4>, the reference's active interpolator (steps.cu:451, 986-1036).
   POND_WENO5C the corrected WENO5 (no reference counterpart): y/z nodes floor(y+0.5)-1..+1,
               smoothness row through the departure point.  Identical to POND_WENO5 on a line
               (Ny = Nz = 1).
   POND_WENO5G the generic propagateWENO<5,M> (steps.cu:457, 610-654) with propagate3
               (909-923): the nodes, smoothness samples and interpolant of POND_WENO5 with the
               sub-stencils numbered from the right (lshift = 1) and the indicators and weights
               in another operation order, so POND_FP64 results differ from POND_WENO5 in the
               last bits.  POND_SPLIT runs it as POND_WENO5. */
typedef enum { POND_WENO5 = 0, POND_LAGR4 = 1, POND_WENO5C = 2, POND_WENO5G = 3 } pond_interp;

/* POND_FP64: plain path, reference operation order, FP64 everywhere.
   POND_SPLIT: FP32 populations/streaming/interpolation/collide, compensated moment sums,
   FP64 gauge residual, temperature/energy moment and moment solve.  D3Q27 (any
   interpolator, up to 134,217,664 cells) and D3Q125 (its fixed 3x3x3 stencil). */
typedef enum { POND_FP64 = 0, POND_SPLIT = 1 } pond_precision;

typedef struct {
  int lattice;      /* pond_lattice */
  int interp;       /* pond_interp */
  int precision;    /* pond_precision */
  int64_t nx, ny, nz;
  double dtau;      /* BGK relaxation (PhysPars::set: 1/0.9) */
  double tol_abs;   /* isConv absolute tolerance (steps.cu:62: 1e-12) */
  double tol_rel;   /* isConv relative tolerance (steps.cu:63: 1e-10) */
  int max_iter;     /* fixed-point cap (steps.cu:422: 100) */
  int use_mlp;      /* 1: learned starting gauge (POND_SPLIT only) */
  double split_tol_rel; /* POND_SPLIT: relative tolerance on the FP64 gauge residual */
  double split_tol_abs; /* POND_SPLIT: absolute tolerance on the FP64 gauge residual */
  /* Work distribution of the D3Q27 step kernels.  None of these changes a result (each is
     tested bit for bit against the defaults); pond_params_default sets the measured defaults.
     tile_patch    cells are taken in 64-cell x-row segments whose (y, z) rows are visited in
                   tile_patch x tile_patch patches, so the waves resident on one XCD share
                   their stencil halo in its L2; used where nx % 64 == 0 and tile_patch divides
                   ny and nz.  0: plain row order.  -1 (default): 8 for the POND_SPLIT
                   D3Q27 WENO5C step, 4 otherwise.
     refill_chunk  64-cell tiles owned by one wave, whose lanes take the next unstarted cell
                   of the chunk when theirs converges.  0: one tile per wave.  -1 (default):
                   FP64 2 tiles; split the whole x row with at least 4 tiles.
     pass1_iters   iterations of the first pass of the two-pass step; cells not converged by
                   then are finished by a second kernel over a compacted list.  -1 (default):
                   FP64 8, split one pass.
     global_loads  1: FP64 D3Q27 node loads through 64-bit global addresses, the path of grids
                   whose population planes do not fit 32-bit buffer offsets (above 67,108,800
                   cells).  0 (default): buffer loads where they fit. */
  int tile_patch;
  int refill_chunk;
  int pass1_iters;
  int global_loads;
  /* POND_SPLIT with use_mlp: largest predicted change (|du|/sqrt(T), |dT|/T) taken as a
     start; larger predictions start from the previous gauge.  Default 0.3. */
  double learned_trust;
  /* Mirror boundaries, bit d for axis d (1: x, 2: y, 4: z); 0 (default): periodic.  The
     reference's compile-time outflow switch (steps.cu:167-190, 1004-1022): a stencil node
     beyond a mirrored face is the mirror image of the cell inside (position reflected about the
     face, axis velocity negated, populations exchanged with their mirror images, Cell::mirrX/Y/Z,
     structs.cuh:552-578).  The reference has it in propagate<N> and PonDstepD3Q125 only, so it
     applies to POND_LAGR4 (D1Q5, D3Q27) and to D3Q125; WENO5 interpolation with a mirror is
     rejected (POND_ERR_INVALID). */
  int mirror;
  /* 1: the 1D cell activation of amr.cu after every step, as calcStep would run it with its
     launches enabled (update.cu:102-105): a gap between active cells across which the new states
     differ (|du_x|^2/min T > 0.01 or |rho ratio - 1| > 0.01) and that is an even number of cells
     wide gets the cell half way switched on with the average record (AMR_add, amr.cu:1-27);
     then one pass along the line in increasing x switches off each active cell whose previous
     and next active cells are within 3 cells and match it (AMR_remove, amr.cu:28-55).  A
     switched-off cell is not stepped (its record is left as it was, LBMpull, steps.cu:359) but
     stays an interpolation node.  D1Q5 lines (ny = nz = 1) of up to 32768 cells, POND_FP64
     only (POND_ERR_INVALID otherwise).  Default 0. */
  int amr;
} pond_params;

typedef struct {
  double mass;       /* sum rho                       (update.cu:65) */
  double momentum_x; /* sum rho*ux                    (update.cu:66) */
  double energy;     /* sum rho*(|u|^2 + Dim*T)       (update.cu:67) */
  double k;          /* kinetic energy of diagn.cu:18 (every cell) */
  int32_t niter_max; /* max gauge iterations, last step (steps.cu:513) */
  double niter_mean; /* mean gauge iterations per cell, last step
                        (mass, momentum, energy, the iteration figures and max_departure cover
                        the active cells, pond_get_active) */
  int64_t step;      /* steps taken (LBMParams::iStep) */
  double max_departure; /* largest population displacement per step along an axis, in cells:
                           max over cells and axes of |u_d| + sqrt(T/TLat)*max|e_d| (the
                           interpolation stencils reach this far plus their own width) */
} pond_diag;

typedef struct {
  double last_ms;    /* device time of the step kernel(s) of the last pond_step call (HIP events
                        on the ctx stream around each step kernel launch) */
  double total_ms;   /* accumulated step-kernel time of all pond_step calls */
  int64_t steps;     /* steps covered by total_ms */
  double last_mlp_ms;/* device time of the learned initializer's inference in the last step
                        (outside last_ms) */
} pond_timing;

typedef struct pond_ctx pond_ctx;

/* defaults of the reference: dtau=1/0.9, tol 1e-12 + 1e-10*|x|, cap 100, WENO5, FP64 */
void pond_params_default(pond_params* p, int lattice);

pond_status pond_create(const pond_params* p, pond_ctx** out);
void pond_destroy(pond_ctx* ctx);

/* last error message of ctx (or of the last failed pond_create when ctx is NULL) */
const char* pond_last_error(const pond_ctx* ctx);

/* number of populations of a lattice (5, 27, 125); 0 for an unknown lattice */
int pond_qn(int lattice);

/* lattice tables as evaluated by this library: e (int, Qn*3), ef (Qn*3), w (Qn),
   tlat_dim[2] = {TLat, Dim} */
pond_status pond_lattice_tables(int lattice, int32_t* e, double* ef, double* w, double* tlat_dim);

/* equilibrium state f_i = w_i*rho in each cell's own gauge (Cell::set_base, lbpack.cu:15-21) */
pond_status pond_set_state(pond_ctx* ctx, const double* rho, const double* ux, const double* uy,
                           const double* uz, const double* T);

/* arbitrary full state: f plane-major (Qn*N) plus the five moment planes */
pond_status pond_set_populations(pond_ctx* ctx, const double* f, const double* rho, const double* ux,
                                 const double* uy, const double* uz, const double* T);

/* the current state of src (populations, moments, per-cell iteration counts, step count) into
   dst, on the device, converting between FP64 and split storage (FP64 -> FP32 rounds to
   nearest).  Same lattice and grid; the interpolator and parameters of dst are its own.  No
   counterpart in the reference (one global state); it lets a run switch precision, and lets one
   step of two paths start from the same state. */
pond_status pond_copy_state(pond_ctx* dst, pond_ctx* src);

/* Cinfo::set of every cell (structs.cuh:92): 1 = stepped, 0 = switched off (kept as it is, still
   an interpolation node).  Every cell starts switched on (fill, init.cu:106); pond_params.amr
   switches cells on and off after each step.  pond_set_active applies to D1Q5 POND_FP64
   contexts (the cells of other contexts are always on; POND_ERR_INVALID). */
pond_status pond_get_active(pond_ctx* ctx, int32_t* set);
pond_status pond_set_active(pond_ctx* ctx, const int32_t* set);

/* nsteps PonD steps (calcStep without I/O).  Asynchronous on the ctx stream; returns after
   the launches are queued.  pond_sync / any get call waits. */
pond_status pond_step(pond_ctx* ctx, int nsteps);
pond_status pond_sync(pond_ctx* ctx);

/* current state (output of the last step); any pointer may be NULL */
pond_status pond_get_fields(pond_ctx* ctx, double* rho, double* ux, double* uy, double* uz, double* T,
                            int32_t* niter);
pond_status pond_get_populations(pond_ctx* ctx, double* f);

pond_status pond_diagnostics(pond_ctx* ctx, pond_diag* out);
pond_status pond_timing_info(pond_ctx* ctx, pond_timing* out);

/* write Rho/Vx/Vy/Vz/T_%07d.arr of the current state into directory dir (drop.cu:2-34) */
pond_status pond_drop(pond_ctx* ctx, const char* dir);

/* checkpoint of the current state: populations, moments, per-cell iteration counts and the step
   count, in the device layout, so that pond_checkpoint_read into a context with the same
   lattice, precision and grid restarts bit for bit.  No counterpart in the reference, whose
   drop() (drop.cu:2-34) writes the moments only.  The learned initializer's network is not
   part of it (pond_mlp_get_params / pond_mlp_set_params). */
pond_status pond_checkpoint_write(pond_ctx* ctx, const char* path);
pond_status pond_checkpoint_read(pond_ctx* ctx, const char* path);

/* read one .arr file written by pond_drop (or by the reference's drop()) into out
   (nx*ny*nz doubles); dims[3] receives Nx,Ny,Nz.  out may be NULL to query dims. */
pond_status pond_read_arr(const char* path, int64_t* dims, double* out);

/* ---- learned gauge initializer (no counterpart in the reference) ----
   A 135-64-64-4 ReLU network in BF16 (FP32 accumulate) predicts each cell's converged gauge of
   the next step from the previous step's (rho,u,T) on its 3x3x3 neighbourhood.  It is trained
   on the GPU from samples the solver itself produces.  A POND_SPLIT context created with
   use_mlp=1 uses it for the starting gauge once it has been trained or set. */

/* number of FP32 parameters, packed W1[135][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4] */
int64_t pond_mlp_num_params(void);

/* seeded He-uniform initialisation (also resets the Adam state) */
pond_status pond_mlp_init(pond_ctx* ctx, uint64_t seed);

/* append nsamples seeded-random cells of the last step: features from the state before the
   step, targets from the converged gauge after it (POND_SPLIT contexts) */
pond_status pond_mlp_collect(pond_ctx* ctx, int64_t nsamples, uint64_t seed);
pond_status pond_mlp_dataset_size(pond_ctx* ctx, int64_t* n);

/* train on the collected samples: epochs of shuffled minibatches (batch a multiple of 16),
   Adam with learning rate lr; epoch_loss[epochs] (may be NULL) receives each epoch's mean
   normalized squared error over every 64th batch */
pond_status pond_mlp_train(pond_ctx* ctx, int epochs, int batch, float lr, uint64_t seed, float* epoch_loss);

/* parameters (kMlp params) and scales[9] = {1/scale of (rho,ux,uy,uz,T) features, output
   scales of (dux,duy,duz,dT)} */
pond_status pond_mlp_get_params(pond_ctx* ctx, float* params, float* scales);
pond_status pond_mlp_set_params(pond_ctx* ctx, const float* params, const float* scales);

/* network on a host batch of already-normalized features X[B][135] -> Y[B][4] (output
   scales applied), through the fused MFMA inference kernel */
pond_status pond_mlp_forward(pond_ctx* ctx, const float* X, int64_t B, float* Y);

/* pond_mlp_forward that allocates nothing, for callers that cannot allocate device memory
   (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
   counterpart in the reference.  The device staging of X and Y is a static array of the library,
   one per GPU, held until the process ends: B is at most 16384 rows (larger B returns
   POND_ERR_INVALID).  Every call on one GPU uses the same array, so calls that overlap on one
   GPU (from several threads or contexts) race.  Call with the context's device current. */
pond_status pond_mlp_forward_noalloc(pond_ctx* ctx, const float* X, int64_t B, float* Y);

/* the network's predicted start for every cell of the current state, as the step uses it:
   out[4][n] (x fastest) = (du_x/sqrt(T), du_y/sqrt(T), du_z/sqrt(T), dT/T) relative to the
   cell's gauge, computed by the fused inference kernel from the 3x3x3 neighbourhoods.  No
   counterpart in the reference (the learned initializer is new); for inspecting the learned
   start. */
pond_status pond_mlp_predict(pond_ctx* ctx, float* out);

/* gradient of the mean-squared loss of a host batch (X[B][135], targets T[B][4], B a multiple
   of 16) through the MFMA training kernels; grad has pond_mlp_num_params() entries */
pond_status pond_mlp_grad(pond_ctx* ctx, const float* X, const float* T, int64_t B, float* grad, float* loss);

/* device bytes held by the ctx (state buffers, scratch, network) */
int64_t pond_device_bytes(const pond_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* POND_H */
