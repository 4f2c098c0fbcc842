// pond_ctx.h -- the opaque context behind the C ABI (include/pond.h).
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../../include/pond.h"
#include "lattice.h"
#include "pond_device.h"
#include "step_split.h"
#include "mlp.h"

namespace pond {

// device buffers of one state (load or store) of the plain FP64 path
struct F64Buf {
  double* f = nullptr;  // populations (f64_fidx layout)
  double* m = nullptr;  // rho, ux, uy, uz, T planes
};

// device buffers of one state of the split path: FP32 populations and (rho, ux, uy, uz),
// FP64 temperature
struct SplitBuf {
  float* f = nullptr;   // populations (split_fidx layout)
  float* m = nullptr;   // rho, ux, uy, uz planes
  double* T = nullptr;  // T plane
};

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
};

// learned gauge initializer: network, training state and dataset
struct MlpCtx {
  bool initialized = false, trained = false;
  float* params = nullptr;  // FP32 master parameters (kMlpParams, packed)
  float* grad = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  int64_t adam_t = 0;
  unsigned short* W1b = nullptr;  // bf16 copies: W1 144x64, W2 64x64, W3 64x16
  unsigned short* W2b = nullptr;
  unsigned short* W3b = nullptr;
  float* gw3tmp = nullptr;
  double* mom = nullptr;
  float in_inv[5] = {1, 1, 1, 1, 1};
  float out_scale[4] = {1, 1, 1, 1};
  float out_inv[4] = {1, 1, 1, 1};
  float* X = nullptr;  // dataset: raw features (cap x 135) and targets (cap x 4)
  float* Y = nullptr;
  int64_t count = 0, cap = 0;
  int bcap = 0;  // minibatch buffers
  float *bx = nullptr, *bt = nullptr, *z1acc = nullptr, *z1 = nullptr, *z2acc = nullptr, *z2 = nullptr,
        *yacc = nullptr, *y = nullptr, *dy = nullptr, *g2 = nullptr, *dz2 = nullptr, *g1 = nullptr, *dz1 = nullptr,
        *sqerr = nullptr, *loss = nullptr;
  unsigned short *xb = nullptr, *a1b = nullptr, *a2b = nullptr, *dyb = nullptr, *dz2b = nullptr, *dz1b = nullptr;
  int* idx = nullptr;
};

hipError_t launch_node_prep(int lattice, const double* f, const double* m, double* P, unsigned char* ok, int64_t n,
                            int64_t plane, hipStream_t st);
hipError_t launch_step_f64(int lattice, int interp, const double* fin, const double* min, const double* Pin,
                           const unsigned char* okin, double* fout, double* mout, int* niter, const Pending& pend,
                           const StepArgs& a, hipStream_t st);
// the cell activation passes after a D1Q5 FP64 step (amr.hip): st = the step's output buffer,
// ld = its input buffer
hipError_t launch_amr(const F64Buf& st, const F64Buf& ld, int* set, int* setnew, const Grid& g, int qn,
                      hipStream_t stream);
int amr_max_line();

}  // namespace pond

struct pond_ctx {
  pond_params p{};
  pond::LatticeTables L{};
  pond::Grid g{};
  pond::StepArgs args{};
  int Qn = 0;
  hipStream_t stream = nullptr;
  int cur = 0;              // index of the buffer holding the current state
  bool has_state = false;
  bool split = false;       // POND_SPLIT state layout
  pond::F64Buf f64[2];
  pond::SplitBuf sp[2];
  pond::SplitArgs sargs{};
  double* nodeP = nullptr;  // plain path: per cell of the load buffer, D3Q27 the record (ux, uy, uz, P),
                            // else P, with P = sqrt(|T|)*sign(T)
  unsigned char* nodeok = nullptr;  // plain path: populations finite and below 2^200
  float* sq = nullptr;      // split path: node gauge records (ux, uy, uz, sqrt(|T|)*sign(T)) of the load buffer
  float* guess = nullptr;   // split path: learned starting gauge increments (4 planes)
  // cells left unconverged by the first pass of a two-pass step: count, then per entry the
  // cell index, 4 gauge doubles, one double and one int (48 B per cell of the grid)
  char* pend = nullptr;
  int64_t pend_cap = 0;
  int* niter = nullptr;     // per-cell gauge iterations of the last step
  int* cell_set = nullptr;  // D1Q5 FP64: Cinfo::set per cell (1 = stepped) and its next value
  int* cell_setnew = nullptr;
  double* red = nullptr;    // reduction partials
  int64_t red_len = 0;
  int64_t step = 0;
  int64_t bytes = 0;
  // timing of the step kernel launches (HIP events on `stream`)
  std::vector<pond::EventPair> ev_pool;
  std::vector<pond::EventPair> ev_pending;
  double last_ms = 0, total_ms = 0, last_mlp_ms = 0;
  int64_t timed_steps = 0;
  pond::MlpCtx mlp;
  hipEvent_t mlp_ev[2] = {nullptr, nullptr};  // around the last inference launch
  bool mlp_pending = false;
  std::string err;
};

pond::MlpWeightsBF pond_mlp_weights(const pond_ctx* ctx);
void pond_mlp_free(pond_ctx* ctx);
