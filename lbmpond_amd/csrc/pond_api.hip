// pond_api.hip -- C ABI of the PonD step (include/pond.h): context lifetime, state upload and
// download, the step driver (calcStep, update.cu:86-142, without the I/O it times), global
// diagnostics and the .arr field dump.
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "pond_ctx.h"

using namespace pond;

namespace {

thread_local std::string g_create_error;

#define POND_HIP(ctx, expr)                                                                  \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                        \
      return POND_ERR_HIP;                                                                   \
    }                                                                                        \
  } while (0)

pond_status fail(pond_ctx* ctx, pond_status s, const std::string& msg) {
  ctx->err = msg;
  return s;
}

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// ---------------------------------------------------------------------------------- kernels

// Cell::set_base (lbpack.cu:15-21) from the moment planes into the FP64 populations
__global__ void k_set_base(double* __restrict__ f, const double* __restrict__ rho, int64_t n, int64_t plane,
                           int lattice, int qn, const StepArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double r = rho[c];
  for (int i = 0; i < qn; i++) f[f64_fidx(lattice, i, c, plane)] = a.w[i] * r;
}

// split layout from FP64 moment planes: f_i = w_i*rho in FP32, (rho,u) in FP32, T in FP64
__global__ void k_set_split(const double* __restrict__ m, float* __restrict__ f, float* __restrict__ mf,
                            double* __restrict__ T, int64_t n, int64_t plane, int lattice, int qn, const StepArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double r = m[c];
  for (int i = 0; i < qn; i++) f[split_fidx(lattice, i, c, plane)] = (float)(a.w[i] * r);
  for (int k = 0; k < 4; k++) mf[(int64_t)k * plane + c] = (float)m[(int64_t)k * plane + c];
  T[c] = m[(int64_t)4 * plane + c];
}

// one state in another context's layout (pond_copy_state): populations and moments of every
// cell from an FP64 or split source into an FP64 or split destination (FP64 -> FP32 rounds to
// nearest, FP32 -> FP64 is exact)
struct StateView {
  const double* f64;  // FP64 populations (f64_fidx layout) and moment planes ...
  const double* m64;
  const float* f32;   // ... or split populations (split_fidx layout), (rho, u) planes and T
  const float* m32;
  const double* T;
};
struct StateOut {
  double* f64;
  double* m64;
  float* f32;
  float* m32;
  double* T;
};

__global__ void k_copy_state(StateView in, StateOut out, int64_t n, int64_t plane, int lattice, int qn) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (int i = 0; i < qn; i++) {
    const double v = in.f64 ? in.f64[f64_fidx(lattice, i, c, plane)] : (double)in.f32[split_fidx(lattice, i, c, plane)];
    if (out.f64) out.f64[f64_fidx(lattice, i, c, plane)] = v;
    else out.f32[split_fidx(lattice, i, c, plane)] = (float)v;
  }
  for (int k = 0; k < 5; k++) {
    const double v = in.m64 ? in.m64[(int64_t)k * plane + c] : k < 4 ? (double)in.m32[(int64_t)k * plane + c] : in.T[c];
    if (out.m64) out.m64[(int64_t)k * plane + c] = v;
    else if (k < 4) out.m32[(int64_t)k * plane + c] = (float)v;
    else out.T[c] = v;
  }
}

// block-level deterministic partial sums: mass, momentum_x, energy, k, niter sum; niter max
struct RedOut {
  double s[6];  // mass, momentum_x, energy, k, sum of iteration counts, active cells
  double vmax;  // largest |u_d| + sqrt(|T|/TLat)*emax
  int imax;
};

__device__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += sh[i];
  return t;
}

__device__ int block_max(int v, int* sh) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  int t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t = max(t, sh[i]);
  return t;
}

__device__ double block_maxd(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t = fmax(t, sh[i]);
  return t;
}

// the displacement bound of one cell: max_d |u_d| + sqrt(|T|/TLat) * emax
__device__ __forceinline__ double departure(double ux, double uy, double uz, double T, double inv_tlat, double emax) {
  return fmax(fmax(fabs(ux), fabs(uy)), fabs(uz)) + sqrt(fabs(T) * inv_tlat) * emax;
}

// calcTotMoments (update.cu:52-69, active cells) + diagn (diagn.cu:3-22, every cell) + NiterMax
// (steps.cu:513, stepped cells), for the FP64 state; set: Cinfo::set or null (every cell on)
__global__ __launch_bounds__(256) void k_reduce_f64(const double* __restrict__ m, const int* __restrict__ niter,
                                                    const int* __restrict__ set, int64_t n, int64_t plane, double dim,
                                                    double cx, double cy, double cz, double inv_tlat, double emax,
                                                    RedOut* __restrict__ out) {
  __shared__ double shd[4];
  __shared__ int shi[4];
  double s[6] = {0, 0, 0, 0, 0, 0};
  double vmax = 0;
  int imax = 0;
  const double d8pi3 = 1.0 / (8 * M_PI * M_PI * M_PI);
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const double rho = m[c], ux = m[plane + c], uy = m[2 * plane + c], uz = m[3 * plane + c], T = m[4 * plane + c];
    s[3] += 0.5 * d8pi3 * (ux * ux + uy * uy + uz * uz) * cx * cy * cz;
    if (set && !set[c]) continue;
    s[0] += rho * 1.0;
    s[1] += rho * ux * 1.0;
    s[2] += rho * (ux * ux + uy * uy + uz * uz + dim * T) * 1.0;
    s[4] += (double)niter[c];
    s[5] += 1.0;
    imax = max(imax, niter[c]);
    vmax = fmax(vmax, departure(ux, uy, uz, T, inv_tlat, emax));
  }
  RedOut r;
  for (int q = 0; q < 6; q++) r.s[q] = block_sum(s[q], shd);
  r.imax = block_max(imax, shi);
  r.vmax = block_maxd(vmax, shd);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// the same reductions for the split layout (FP32 rho,u and FP64 T)
__global__ __launch_bounds__(256) void k_reduce_split(const float* __restrict__ m, const double* __restrict__ Tp,
                                                      const int* __restrict__ niter, int64_t n, int64_t plane,
                                                      double dim, double cx, double cy, double cz, double inv_tlat,
                                                      double emax, RedOut* __restrict__ out) {
  __shared__ double shd[4];
  __shared__ int shi[4];
  double s[6] = {0, 0, 0, 0, 0, 0};
  double vmax = 0;
  int imax = 0;
  const double d8pi3 = 1.0 / (8 * M_PI * M_PI * M_PI);
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const double rho = m[c], ux = m[plane + c], uy = m[2 * plane + c], uz = m[3 * plane + c], T = Tp[c];
    s[5] += 1.0;
    s[0] += rho;
    s[1] += rho * ux;
    s[2] += rho * (ux * ux + uy * uy + uz * uz + dim * T);
    s[3] += 0.5 * d8pi3 * (ux * ux + uy * uy + uz * uz) * cx * cy * cz;
    s[4] += (double)niter[c];
    imax = max(imax, niter[c]);
    vmax = fmax(vmax, departure(ux, uy, uz, T, inv_tlat, emax));
  }
  RedOut r;
  for (int q = 0; q < 6; q++) r.s[q] = block_sum(s[q], shd);
  r.imax = block_max(imax, shi);
  r.vmax = block_maxd(vmax, shd);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

__global__ void k_reduce_final(const RedOut* __restrict__ part, int nb, RedOut* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  RedOut r{};
  for (int b = 0; b < nb; b++) {
    for (int q = 0; q < 6; q++) r.s[q] += part[b].s[q];
    r.imax = max(r.imax, part[b].imax);
    r.vmax = fmax(r.vmax, part[b].vmax);
  }
  *out = r;
}

constexpr int kRedBlocks = 1024;

pond_status ensure_events(pond_ctx* ctx) {
  if (ctx->ev_pool.empty()) {
    for (int i = 0; i < 64; i++) {
      EventPair e;
      POND_HIP(ctx, hipEventCreate(&e.a));
      POND_HIP(ctx, hipEventCreate(&e.b));
      ctx->ev_pool.push_back(e);
    }
  }
  return POND_OK;
}

// fold the finished event pairs into the timing totals
pond_status collect_events(pond_ctx* ctx) {
  if (ctx->ev_pending.empty()) return POND_OK;
  double ms_sum = 0;
  for (auto& e : ctx->ev_pending) {
    POND_HIP(ctx, hipEventSynchronize(e.b));
    float ms = 0;
    POND_HIP(ctx, hipEventElapsedTime(&ms, e.a, e.b));
    ms_sum += ms;
    ctx->ev_pool.push_back(e);
  }
  ctx->last_ms = ms_sum;
  ctx->total_ms += ms_sum;
  if (ctx->mlp_pending) {
    float mm = 0;
    POND_HIP(ctx, hipEventElapsedTime(&mm, ctx->mlp_ev[0], ctx->mlp_ev[1]));
    ctx->last_mlp_ms = mm;
    ctx->mlp_pending = false;
  }
  ctx->timed_steps += (int64_t)ctx->ev_pending.size();
  ctx->ev_pending.clear();
  return POND_OK;
}

}  // namespace

extern "C" {

void pond_params_default(pond_params* p, int lattice) {
  std::memset(p, 0, sizeof(*p));
  p->lattice = lattice;
  p->interp = POND_WENO5;
  p->precision = POND_FP64;
  p->nx = p->ny = p->nz = 1;
  p->dtau = 1.0 / 0.9;  // PhysPars::set (structs.cuh:98)
  p->tol_abs = 1e-12;   // isConv (steps.cu:62-63)
  p->tol_rel = 1e-10;
  p->max_iter = 100;    // steps.cu:422
  p->use_mlp = 0;
  // FP32 populations put a noise floor near 3e-8 (relative to sqrt(T) and T) under the gauge
  // residual; 1e-7 converges reliably above it
  p->split_tol_rel = 1e-7;
  p->split_tol_abs = 1e-12;
  // Row patches (-1: chosen in pond_create).  384^3 shock-bubble D3Q27: the FP64 step (pair
  // planes, lane refill) 4.10 s in row order, 3.13 s with 4x4 patches, 3.14 s with 8x8; the split
  // step with its node loads one step ahead, WENO5C 261 ms in row order, 161.4 ms with 4x4
  // patches, 157.0 with 8x8, 157.8 with 6x6, 157.2 with 12x12, 158.2 with 16x16; Lagrange-4
  // 422, 225.2, 226.2 and 232.2 ms (row order, 4, 8, 16).
  p->tile_patch = -1;
  p->refill_chunk = -1;
  p->pass1_iters = -1;
  p->global_loads = 0;
  // 384^3 shock-bubble, learned path: trust radius 0.1 1.290 mean iterations, 307 ms step
  // kernel; 0.3 and 1.0 1.284, 300 ms
  p->learned_trust = 0.3;
}

int pond_qn(int lattice) { return qn_of(lattice); }

pond_status pond_lattice_tables(int lattice, int32_t* e, double* ef, double* w, double* tlat_dim) {
  LatticeTables t;
  if (!make_lattice(lattice, &t)) return POND_ERR_INVALID;
  for (int i = 0; i < t.Qn; i++) {
    for (int d = 0; d < 3; d++) {
      if (e) e[3 * i + d] = t.e[i][d];
      if (ef) ef[3 * i + d] = t.ef[i][d];
    }
    if (w) w[i] = t.w[i];
  }
  if (tlat_dim) { tlat_dim[0] = t.TLat; tlat_dim[1] = t.Dim; }
  return POND_OK;
}

const char* pond_last_error(const pond_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

pond_status pond_create(const pond_params* p, pond_ctx** out) {
  if (!p || !out) { g_create_error = "pond_create: null argument"; return POND_ERR_INVALID; }
  *out = nullptr;
  auto* ctx = new (std::nothrow) pond_ctx();
  if (!ctx) { g_create_error = "pond_create: out of host memory"; return POND_ERR_NOMEM; }
  auto bail = [&](pond_status s, const std::string& m) {
    g_create_error = m;
    pond_destroy(ctx);
    return s;
  };
  ctx->p = *p;
  if (!make_lattice(p->lattice, &ctx->L)) return bail(POND_ERR_INVALID, "pond_create: unknown lattice");
  if (p->nx < 1 || p->ny < 1 || p->nz < 1 || p->nx > (1 << 20) || p->ny > (1 << 20) || p->nz > (1 << 20))
    return bail(POND_ERR_INVALID, "pond_create: grid sizes must be in [1, 2^20]");
  // node data are addressed with 32-bit byte offsets into 8-byte planes
  if (p->nx * p->ny * p->nz > ((int64_t)1 << 29) - 64)
    return bail(POND_ERR_INVALID, "pond_create: at most 2^29 - 64 cells per context");
  if (p->interp != POND_WENO5 && p->interp != POND_LAGR4 && p->interp != POND_WENO5C && p->interp != POND_WENO5G)
    return bail(POND_ERR_INVALID, "pond_create: unknown interp");
  if (p->precision != POND_FP64 && p->precision != POND_SPLIT)
    return bail(POND_ERR_INVALID, "pond_create: unknown precision");
  if (p->precision == POND_SPLIT && p->lattice == POND_D1Q5)
    return bail(POND_ERR_INVALID, "pond_create: POND_SPLIT is built for D3Q27 and D3Q125");
  if (p->precision == POND_SPLIT && !(p->split_tol_rel >= 0 && p->split_tol_abs >= 0))
    return bail(POND_ERR_INVALID, "pond_create: split tolerances must be >= 0");
  if (p->max_iter < 1) return bail(POND_ERR_INVALID, "pond_create: max_iter must be >= 1");
  if (p->use_mlp && p->precision != POND_SPLIT)
    return bail(POND_ERR_INVALID, "pond_create: the learned initializer requires POND_SPLIT");
  ctx->Qn = ctx->L.Qn;
  Grid& g = ctx->g;
  g.nx = (int)p->nx; g.ny = (int)p->ny; g.nz = (int)p->nz;
  g.n = p->nx * p->ny * p->nz;
  g.plane = round_up(g.n, 64);
  if (p->tile_patch < -1) return bail(POND_ERR_INVALID, "pond_create: tile_patch must be >= -1");
  if (p->refill_chunk < -1) return bail(POND_ERR_INVALID, "pond_create: refill_chunk must be >= -1");
  if (p->pass1_iters < -1 || p->pass1_iters == 0)
    return bail(POND_ERR_INVALID, "pond_create: pass1_iters must be -1 or >= 1");
  if (!(p->learned_trust >= 0)) return bail(POND_ERR_INVALID, "pond_create: learned_trust must be >= 0");
  if (p->mirror < 0 || p->mirror > 7) return bail(POND_ERR_INVALID, "pond_create: mirror must be a mask of axes 0..7");
  if (p->mirror && p->lattice != POND_D3Q125 && p->interp != POND_LAGR4)
    return bail(POND_ERR_INVALID,
                "pond_create: mirror boundaries apply to POND_LAGR4 and D3Q125 (the reference's WENO5 has no mirror path)");
  if (p->amr != 0 && p->amr != 1) return bail(POND_ERR_INVALID, "pond_create: amr must be 0 or 1");
  if (p->amr && (p->lattice != POND_D1Q5 || p->precision != POND_FP64 || p->ny != 1 || p->nz != 1 ||
                 p->nx > amr_max_line()))
    return bail(POND_ERR_INVALID, "pond_create: amr applies to POND_FP64 D1Q5 lines (ny = nz = 1) of up to " +
                                      std::to_string(amr_max_line()) + " cells");
  // row patches where the grid divides into them (ignored elsewhere); by default 8x8 for the
  // split D3Q27 WENO5C step, 4x4 for the others
  const int rp = p->tile_patch >= 0 ? p->tile_patch
                 : (p->precision == POND_SPLIT && p->lattice == POND_D3Q27 && p->interp == POND_WENO5C) ? 8
                                                                                                        : 4;
  g.rp = (rp > 0 && g.nx % 64 == 0 && g.ny % rp == 0 && g.nz % rp == 0) ? rp : 0;
  g.ntiles = g.rp > 0 ? (int64_t)(g.nx / 64) * g.ny * g.nz : (g.n + 63) / 64;
  g.mirror = p->mirror;
  // the split step addresses nine FP32 population planes through one 32-bit buffer offset
  // the split step addresses two population quad planes (16 bytes per cell each) through one
  // 32-bit buffer offset
  if (p->precision == POND_SPLIT && p->lattice == POND_D3Q27 && g.plane * 32 > (int64_t)UINT32_MAX)
    return bail(POND_ERR_INVALID, "pond_create: POND_SPLIT D3Q27 supports up to 134,217,664 cells (about 512^3)");
  StepArgs& a = ctx->args;
  a.g = g;
  a.dtau = p->dtau;
  a.tol_abs = p->tol_abs;
  a.tol_rel = p->tol_rel;
  a.max_iter = p->max_iter;
  a.sqrt_tbase = std::sqrt(ctx->L.TLat);
  a.tlat = ctx->L.TLat;
  a.inv_dim = 1. / ctx->L.Dim;
  a.g = g;
  a.global_loads = p->global_loads != 0;
  // lane refill in the first pass of the FP64 D3Q27 step.  384^3 shock-bubble FP64 step: one
  // tile per wave 8.07 s, chunks of 2 tiles 6.38 s, 4 tiles 7.05 s (4 tiles in a single pass
  // 6.90 s); with 4x4 row patches 2 tiles 2.94 s, 3 3.06 s, 6 3.09 s.
  a.refill_chunk = p->refill_chunk >= 0 ? p->refill_chunk : 2;
  // Iterations of the first pass of the two-pass step; cells still unconverged are finished by
  // the second pass.  Measured on the shock-bubble: FP64 D3Q27, 128^3: K=3 2.98, K=5 3.19, K=8
  // 3.35, single pass 2.86 MLUPS; split D3Q27, 192^3: K=3 7.6, K=8 8.9, single pass 11.5 MLUPS
  // (its second pass runs alone after the first and a few slowly converging cells make a serial
  // tail), so the split step runs in one pass.
  a.pass1_iters = p->pass1_iters > 0 ? p->pass1_iters : p->precision == POND_SPLIT ? (1 << 30) : 8;
  for (int i = 0; i < ctx->Qn; i++) a.w[i] = ctx->L.w[i];

  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(POND_ERR_HIP, "pond_create: hipStreamCreate failed");
  ctx->split = p->precision == POND_SPLIT;
  if (!ctx->split) {
    const size_t fbytes = (size_t)f64_fcount(p->lattice, ctx->Qn, g.plane) * sizeof(double);
    const size_t mbytes = (size_t)5 * g.plane * sizeof(double);
    for (int b = 0; b < 2; b++) {
      if (hipMalloc(&ctx->f64[b].f, fbytes) != hipSuccess || hipMalloc(&ctx->f64[b].m, mbytes) != hipSuccess)
        return bail(POND_ERR_NOMEM, "pond_create: hipMalloc of the state buffers failed");
      ctx->bytes += (int64_t)(fbytes + mbytes);
    }
    // D3Q27: 32-byte gauge records (ux, uy, uz, P); otherwise one P plane
    if (hipMalloc(&ctx->nodeP, (size_t)(p->lattice == POND_D3Q27 ? 4 : 1) * g.plane * sizeof(double)) != hipSuccess ||
        hipMalloc(&ctx->nodeok, (size_t)g.plane) != hipSuccess)
      return bail(POND_ERR_NOMEM, "pond_create: hipMalloc of the node planes failed");
    ctx->bytes += (int64_t)g.plane * ((p->lattice == POND_D3Q27 ? 32 : 8) + 1);
  } else {
    const size_t fbytes = (size_t)split_fcount(p->lattice, ctx->Qn, g.plane) * sizeof(float);
    const size_t mbytes = (size_t)4 * g.plane * sizeof(float);
    const size_t tbytes = (size_t)g.plane * sizeof(double);
    for (int b = 0; b < 2; b++) {
      if (hipMalloc(&ctx->sp[b].f, fbytes) != hipSuccess || hipMalloc(&ctx->sp[b].m, mbytes) != hipSuccess ||
          hipMalloc(&ctx->sp[b].T, tbytes) != hipSuccess)
        return bail(POND_ERR_NOMEM, "pond_create: hipMalloc of the split state buffers failed");
      ctx->bytes += (int64_t)(fbytes + mbytes + tbytes);
    }
    if (hipMalloc(&ctx->sq, (size_t)g.plane * 4 * sizeof(float)) != hipSuccess)
      return bail(POND_ERR_NOMEM, "pond_create: hipMalloc of the node gauge records failed");
    ctx->bytes += (int64_t)g.plane * 4 * sizeof(float);
    if (p->use_mlp) {
      if (hipMalloc(&ctx->guess, (size_t)4 * g.plane * sizeof(float)) != hipSuccess)
        return bail(POND_ERR_NOMEM, "pond_create: hipMalloc of the predicted-gauge planes failed");
      ctx->bytes += (int64_t)4 * g.plane * sizeof(float);
      if (hipEventCreate(&ctx->mlp_ev[0]) != hipSuccess || hipEventCreate(&ctx->mlp_ev[1]) != hipSuccess)
        return bail(POND_ERR_HIP, "pond_create: hipEventCreate failed");
    }
    SplitArgs& sa = ctx->sargs;
    sa.pass1_iters = a.pass1_iters;
    // lane refill over chunks of 64-cell tiles.  Measured on the 384^3 shock-bubble, split step
    // ms, with paired x-slice loads: one tile 1065, chunks of 2 tiles 805, 3 976, 4 775-778, 6
    // 763, 8 840, 12 910 (one slice per load and one tile per wave: 1025).  Larger chunks keep
    // more lanes busy but spread the active cells of a wave over more rows, so its loads touch
    // more cache lines.  With the current node loop (gather-bound): 3 319, 4 293, 5 305, 6 279,
    // 7 321, 8 317, 9 328, 12 321.  Six tiles are one whole x row at 384^3: with row patches the
    // tiles are 64-cell row segments, nx/64 per row, so chunks of whole rows let the waves
    // resident on an XCD cover a compact patch of rows (traffic from beyond L2 per launch:
    // 1.23 TB with chunks of 4 tiles, 0.50 TB with 6; at 256^3, 4 tiles per row: chunks of 4
    // 113 ms, 3 128, 6 127, 8 124).  Default: whole rows, at least 4 tiles.
    sa.refill_chunk = 4;
    if (g.rp > 0) {
      const int wpr = g.nx / 64;
      sa.refill_chunk = wpr * ((4 + wpr - 1) / wpr);
    }
    if (p->refill_chunk >= 0) sa.refill_chunk = p->refill_chunk;
    sa.learned_trust = (float)p->learned_trust;
    sa.g = g;
    sa.dtau = p->dtau;
    sa.tol_abs = p->split_tol_abs;
    sa.tol_rel = p->split_tol_rel;
    sa.max_iter = p->max_iter;
    sa.sqrt_tbase = a.sqrt_tbase;
    sa.tlat = a.tlat;
    sa.inv_dim = a.inv_dim;
    for (int i = 0; i < ctx->Qn; i++) sa.w[i] = (float)ctx->L.w[i];
  }
  if (hipMalloc(&ctx->niter, (size_t)g.plane * sizeof(int)) != hipSuccess)
    return bail(POND_ERR_NOMEM, "pond_create: hipMalloc of niter failed");
  ctx->pend_cap = g.plane;
  const size_t pend_bytes = 256 + (size_t)g.plane * (4 + 4 * 8 + 8 + 4);
  if (hipMalloc(&ctx->pend, pend_bytes) != hipSuccess)
    return bail(POND_ERR_NOMEM, "pond_create: hipMalloc of the pending-cell list failed");
  ctx->bytes += (int64_t)pend_bytes;
  ctx->red_len = kRedBlocks + 1;
  if (hipMalloc(&ctx->red, (size_t)ctx->red_len * sizeof(RedOut)) != hipSuccess)
    return bail(POND_ERR_NOMEM, "pond_create: hipMalloc of the reduction buffer failed");
  ctx->bytes += (int64_t)g.plane * sizeof(int) + ctx->red_len * (int64_t)sizeof(RedOut);
  if (hipMemsetAsync(ctx->niter, 0, (size_t)g.plane * sizeof(int), ctx->stream) != hipSuccess)
    return bail(POND_ERR_HIP, "pond_create: hipMemset failed");
  if (p->lattice == POND_D1Q5 && p->precision == POND_FP64) {
    // Cinfo::set, every cell on (fill, init.cu:106)
    if (hipMalloc(&ctx->cell_set, (size_t)g.plane * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->cell_setnew, (size_t)g.plane * sizeof(int)) != hipSuccess)
      return bail(POND_ERR_NOMEM, "pond_create: hipMalloc of the cell flags failed");
    ctx->bytes += (int64_t)g.plane * 2 * sizeof(int);
    if (hipMemsetD32Async((hipDeviceptr_t)ctx->cell_set, 1, (size_t)g.plane, ctx->stream) != hipSuccess)
      return bail(POND_ERR_HIP, "pond_create: hipMemset failed");
    ctx->args.cell_set = ctx->cell_set;
  }
  if (ensure_events(ctx) != POND_OK) return bail(POND_ERR_HIP, "pond_create: " + ctx->err);
  *out = ctx;
  return POND_OK;
}

void pond_destroy(pond_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (int b = 0; b < 2; b++) {
    if (ctx->f64[b].f) (void)hipFree(ctx->f64[b].f);
    if (ctx->f64[b].m) (void)hipFree(ctx->f64[b].m);
    if (ctx->sp[b].f) (void)hipFree(ctx->sp[b].f);
    if (ctx->sp[b].m) (void)hipFree(ctx->sp[b].m);
    if (ctx->sp[b].T) (void)hipFree(ctx->sp[b].T);
  }
  pond_mlp_free(ctx);
  for (auto& e : ctx->mlp_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->niter) (void)hipFree(ctx->niter);
  if (ctx->cell_set) (void)hipFree(ctx->cell_set);
  if (ctx->cell_setnew) (void)hipFree(ctx->cell_setnew);
  if (ctx->sq) (void)hipFree(ctx->sq);
  if (ctx->nodeP) (void)hipFree(ctx->nodeP);
  if (ctx->nodeok) (void)hipFree(ctx->nodeok);
  if (ctx->guess) (void)hipFree(ctx->guess);
  if (ctx->pend) (void)hipFree(ctx->pend);
  if (ctx->red) (void)hipFree(ctx->red);
  for (auto& e : ctx->ev_pool) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
  for (auto& e : ctx->ev_pending) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

static pond_status upload_moments(pond_ctx* ctx, double* mdev, const double* rho, const double* ux, const double* uy,
                                  const double* uz, const double* T) {
  const double* src[5] = {rho, ux, uy, uz, T};
  for (int k = 0; k < 5; k++) {
    if (!src[k]) return fail(ctx, POND_ERR_INVALID, "null field pointer");
    POND_HIP(ctx, hipMemcpyAsync(mdev + (size_t)k * ctx->g.plane, src[k], (size_t)ctx->g.n * sizeof(double),
                                 hipMemcpyHostToDevice, ctx->stream));
  }
  return POND_OK;
}

pond_status pond_set_state(pond_ctx* ctx, const double* rho, const double* ux, const double* uy, const double* uz,
                           const double* T) {
  if (!ctx) return POND_ERR_INVALID;
  const Grid& g = ctx->g;
  const unsigned blocks = (unsigned)((g.n + 255) / 256);
  if (ctx->split) {
    double* tmp = nullptr;
    POND_HIP(ctx, hipMalloc(&tmp, (size_t)5 * g.plane * sizeof(double)));
    pond_status s = upload_moments(ctx, tmp, rho, ux, uy, uz, T);
    if (s == POND_OK) {
      hipLaunchKernelGGL(k_set_split, dim3(blocks), dim3(256), 0, ctx->stream, tmp, ctx->sp[0].f, ctx->sp[0].m,
                         ctx->sp[0].T, g.n, g.plane, ctx->p.lattice, ctx->Qn, ctx->args);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) s = fail(ctx, POND_ERR_HIP, std::string("k_set_split: ") + hipGetErrorString(e));
    }
    (void)hipFree(tmp);
    if (s != POND_OK) return s;
    POND_HIP(ctx, hipMemcpyAsync(ctx->sp[1].f, ctx->sp[0].f,
                                 (size_t)split_fcount(ctx->p.lattice, ctx->Qn, g.plane) * sizeof(float),
                                 hipMemcpyDeviceToDevice, ctx->stream));
    POND_HIP(ctx, hipMemcpyAsync(ctx->sp[1].m, ctx->sp[0].m, (size_t)4 * g.plane * sizeof(float),
                                 hipMemcpyDeviceToDevice, ctx->stream));
    POND_HIP(ctx, hipMemcpyAsync(ctx->sp[1].T, ctx->sp[0].T, (size_t)g.plane * sizeof(double),
                                 hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    pond_status s = upload_moments(ctx, ctx->f64[0].m, rho, ux, uy, uz, T);
    if (s != POND_OK) return s;
    hipLaunchKernelGGL(k_set_base, dim3(blocks), dim3(256), 0, ctx->stream, ctx->f64[0].f, ctx->f64[0].m, g.n,
                       g.plane, ctx->p.lattice, ctx->Qn, ctx->args);
    POND_HIP(ctx, hipGetLastError());
    // both buffers hold the initial state, as fill() does (init.cu:104-105)
    POND_HIP(ctx, hipMemcpyAsync(ctx->f64[1].f, ctx->f64[0].f,
                                 (size_t)f64_fcount(ctx->p.lattice, ctx->Qn, g.plane) * sizeof(double),
                                 hipMemcpyDeviceToDevice, ctx->stream));
    POND_HIP(ctx, hipMemcpyAsync(ctx->f64[1].m, ctx->f64[0].m, (size_t)5 * g.plane * sizeof(double),
                                 hipMemcpyDeviceToDevice, ctx->stream));
  }
  POND_HIP(ctx, hipMemsetAsync(ctx->niter, 0, (size_t)g.plane * sizeof(int), ctx->stream));
  POND_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->cur = 0;
  ctx->has_state = true;
  ctx->step = 0;
  return POND_OK;
}

pond_status pond_set_populations(pond_ctx* ctx, const double* f, const double* rho, const double* ux, const double* uy,
                                 const double* uz, const double* T) {
  if (!ctx) return POND_ERR_INVALID;
  if (!f) return fail(ctx, POND_ERR_INVALID, "pond_set_populations: null f");
  // the context stream is non-blocking and the copies below use the null stream: no queued
  // step may still read or write the state buffers
  POND_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const Grid& g = ctx->g;
  if (ctx->split) {
    const double* src[5] = {rho, ux, uy, uz, T};
    for (int k = 0; k < 5; k++)
      if (!src[k]) return fail(ctx, POND_ERR_INVALID, "null field pointer");
    const int64_t fcount = split_fcount(ctx->p.lattice, ctx->Qn, g.plane);
    std::vector<float> fs((size_t)fcount, 0.0f);
    for (int i = 0; i < ctx->Qn; i++)
      for (int64_t c = 0; c < g.n; c++) fs[split_fidx(ctx->p.lattice, i, c, g.plane)] = (float)f[(size_t)i * g.n + c];
    std::vector<float> tmp((size_t)g.n);
    for (int b = 0; b < 2; b++) {
      POND_HIP(ctx, hipMemcpy(ctx->sp[b].f, fs.data(), (size_t)fcount * sizeof(float), hipMemcpyHostToDevice));
      for (int k = 0; k < 4; k++) {
        for (int64_t c = 0; c < g.n; c++) tmp[c] = (float)src[k][c];
        POND_HIP(ctx, hipMemcpy(ctx->sp[b].m + (size_t)k * g.plane, tmp.data(), (size_t)g.n * sizeof(float),
                                hipMemcpyHostToDevice));
      }
      POND_HIP(ctx, hipMemcpy(ctx->sp[b].T, T, (size_t)g.n * sizeof(double), hipMemcpyHostToDevice));
    }
  } else {
    const int64_t fcount = f64_fcount(ctx->p.lattice, ctx->Qn, g.plane);
    std::vector<double> fs((size_t)fcount, 0.0);
    for (int i = 0; i < ctx->Qn; i++)
      for (int64_t c = 0; c < g.n; c++) fs[f64_fidx(ctx->p.lattice, i, c, g.plane)] = f[(size_t)i * g.n + c];
    for (int b = 0; b < 2; b++) {
      pond_status s = upload_moments(ctx, ctx->f64[b].m, rho, ux, uy, uz, T);
      if (s != POND_OK) return s;
      // a blocking copy: fs is released at the end of this block
      POND_HIP(ctx, hipMemcpy(ctx->f64[b].f, fs.data(), (size_t)fcount * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  POND_HIP(ctx, hipMemsetAsync(ctx->niter, 0, (size_t)g.plane * sizeof(int), ctx->stream));
  POND_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->cur = 0;
  ctx->has_state = true;
  ctx->step = 0;
  return POND_OK;
}

pond_status pond_get_active(pond_ctx* ctx, int32_t* set) {
  if (!ctx || !set) return POND_ERR_INVALID;
  pond_status s = pond_sync(ctx);
  if (s != POND_OK) return s;
  const int64_t n = ctx->g.n;
  if (!ctx->cell_set) {
    for (int64_t c = 0; c < n; c++) set[c] = 1;
    return POND_OK;
  }
  POND_HIP(ctx, hipMemcpy(set, ctx->cell_set, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
  return POND_OK;
}

pond_status pond_set_active(pond_ctx* ctx, const int32_t* set) {
  if (!ctx || !set) return POND_ERR_INVALID;
  if (!ctx->cell_set)
    return fail(ctx, POND_ERR_INVALID, "pond_set_active: cells can be switched off in POND_FP64 D1Q5 contexts only");
  const int64_t n = ctx->g.n;
  std::vector<int32_t> h((size_t)n);
  for (int64_t c = 0; c < n; c++) h[(size_t)c] = set[c] != 0;
  POND_HIP(ctx, hipStreamSynchronize(ctx->stream));
  POND_HIP(ctx, hipMemcpy(ctx->cell_set, h.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  return POND_OK;
}

pond_status pond_copy_state(pond_ctx* dst, pond_ctx* src) {
  if (!dst || !src) return POND_ERR_INVALID;
  if (!src->has_state) return fail(dst, POND_ERR_STATE, "pond_copy_state: the source has no state");
  if (dst->p.lattice != src->p.lattice || dst->g.nx != src->g.nx || dst->g.ny != src->g.ny || dst->g.nz != src->g.nz)
    return fail(dst, POND_ERR_INVALID, "pond_copy_state: lattice and grid must match");
  if (dst == src) return POND_OK;
  const Grid& g = dst->g;
  POND_HIP(src, hipStreamSynchronize(src->stream));
  POND_HIP(dst, hipStreamSynchronize(dst->stream));
  StateView in{};
  if (src->split) { in.f32 = src->sp[src->cur].f; in.m32 = src->sp[src->cur].m; in.T = src->sp[src->cur].T; }
  else { in.f64 = src->f64[src->cur].f; in.m64 = src->f64[src->cur].m; }
  const unsigned blocks = (unsigned)((g.n + 255) / 256);
  // both buffers of the destination hold the state, as after pond_set_populations
  for (int b = 0; b < 2; b++) {
    StateOut out{};
    if (dst->split) { out.f32 = dst->sp[b].f; out.m32 = dst->sp[b].m; out.T = dst->sp[b].T; }
    else { out.f64 = dst->f64[b].f; out.m64 = dst->f64[b].m; }
    hipLaunchKernelGGL(k_copy_state, dim3(blocks), dim3(256), 0, dst->stream, in, out, g.n, g.plane, dst->p.lattice,
                       dst->Qn);
    POND_HIP(dst, hipGetLastError());
  }
  POND_HIP(dst, hipMemcpyAsync(dst->niter, src->niter, (size_t)g.plane * sizeof(int), hipMemcpyDeviceToDevice,
                               dst->stream));
  if (dst->cell_set && src->cell_set)
    POND_HIP(dst, hipMemcpyAsync(dst->cell_set, src->cell_set, (size_t)g.plane * sizeof(int),
                                 hipMemcpyDeviceToDevice, dst->stream));
  POND_HIP(dst, hipStreamSynchronize(dst->stream));
  dst->cur = 0;
  dst->has_state = true;
  dst->step = src->step;
  return POND_OK;
}

// the pending-cell list as the step kernels see it (pond_ctx::pend layout)
static SplitPending pending_view(pond_ctx* ctx) {
  const int64_t cap = ctx->pend_cap;
  char* b = ctx->pend;
  SplitPending v;
  v.count = (int*)b;
  v.g = (double*)(b + 256);
  v.res = v.g + 4 * cap;
  v.cell = (int*)(v.res + cap);
  v.meta = v.cell + cap;
  v.cap = cap;
  return v;
}

pond_status pond_step(pond_ctx* ctx, int nsteps) {
  if (!ctx) return POND_ERR_INVALID;
  if (!ctx->has_state) return fail(ctx, POND_ERR_STATE, "pond_step: no state set");
  if (nsteps < 0) return fail(ctx, POND_ERR_INVALID, "pond_step: nsteps < 0");
  pond_status s = collect_events(ctx);
  if (s != POND_OK) return s;
  for (int k = 0; k < nsteps; k++) {
    if (ctx->ev_pool.empty()) {
      s = collect_events(ctx);
      if (s != POND_OK) return s;
    }
    EventPair e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    const int in = ctx->cur, outb = 1 - ctx->cur;
    hipError_t he = hipSuccess;
    if (ctx->split && ctx->p.lattice == POND_D3Q27) {
      he = launch_node_gauge(ctx->sp[in].f, ctx->sp[in].m, ctx->sp[in].T, ctx->sq, ctx->g.n, ctx->g.plane,
                             ctx->stream);
      if (he != hipSuccess) return fail(ctx, POND_ERR_HIP, std::string("node gauge kernel: ") + hipGetErrorString(he));
    }
    const bool learned = ctx->split && ctx->p.use_mlp && ctx->mlp.trained;
    if (learned) {
      // predicted starting gauge from the load state (part of the step's work)
      POND_HIP(ctx, hipEventRecord(ctx->mlp_ev[0], ctx->stream));
      // only the tiles holding a cell that will take the prediction (kLearnedPrevMin ..
      // kLearnedPrevMax iterations in the previous step) are evaluated
      MlpState ms{ctx->g, ctx->sp[in].m, ctx->sp[in].T, ctx->niter, kLearnedPrevMin, kLearnedPrevMax};
      he = mlp_launch_infer(pond_mlp_weights(ctx), ms, ctx->guess, ctx->g.plane, ctx->stream);
      if (he != hipSuccess) return fail(ctx, POND_ERR_HIP, std::string("mlp inference: ") + hipGetErrorString(he));
      POND_HIP(ctx, hipEventRecord(ctx->mlp_ev[1], ctx->stream));
      ctx->mlp_pending = true;
    }
    if (!ctx->split) {
      he = launch_node_prep(ctx->p.lattice, ctx->f64[in].f, ctx->f64[in].m, ctx->nodeP, ctx->nodeok, ctx->g.n,
                            ctx->g.plane, ctx->stream);
      if (he != hipSuccess) return fail(ctx, POND_ERR_HIP, std::string("node prep kernel: ") + hipGetErrorString(he));
    }
    // the event pair brackets the step kernel alone
    POND_HIP(ctx, hipEventRecord(e.a, ctx->stream));
    if (ctx->split) {
      const SplitPtrs pin{ctx->sp[in].f, ctx->sp[in].m, ctx->sp[in].T};
      const SplitPtrs pout{ctx->sp[outb].f, ctx->sp[outb].m, ctx->sp[outb].T};
      // the generic WENO5 is the as-written interpolant in another FP64 operation order; the
      // split path takes the as-written kernel for it
      he = launch_step_split(ctx->p.lattice, ctx->p.interp == POND_WENO5G ? POND_WENO5 : ctx->p.interp, pin, pout,
                             ctx->sq, ctx->niter,
                             learned ? ctx->guess : nullptr, pending_view(ctx), ctx->sargs, ctx->stream);
    } else {
      he = launch_step_f64(ctx->p.lattice, ctx->p.interp, ctx->f64[in].f, ctx->f64[in].m, ctx->nodeP, ctx->nodeok,
                           ctx->f64[outb].f, ctx->f64[outb].m, ctx->niter, pending_view(ctx), ctx->args, ctx->stream);
    }
    if (he != hipSuccess) return fail(ctx, POND_ERR_HIP, std::string("step kernel launch: ") + hipGetErrorString(he));
    POND_HIP(ctx, hipEventRecord(e.b, ctx->stream));
    ctx->ev_pending.push_back(e);
    if (ctx->p.amr) {
      he = launch_amr(ctx->f64[outb], ctx->f64[in], ctx->cell_set, ctx->cell_setnew, ctx->g, ctx->Qn, ctx->stream);
      if (he != hipSuccess) return fail(ctx, POND_ERR_HIP, std::string("amr kernels: ") + hipGetErrorString(he));
    }
    ctx->cur = outb;  // swapLS (structs.cuh:148)
    ctx->step++;
  }
  return POND_OK;
}

pond_status pond_sync(pond_ctx* ctx) {
  if (!ctx) return POND_ERR_INVALID;
  POND_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return collect_events(ctx);
}

pond_status pond_get_fields(pond_ctx* ctx, double* rho, double* ux, double* uy, double* uz, double* T, int32_t* niter) {
  if (!ctx) return POND_ERR_INVALID;
  if (!ctx->has_state) return fail(ctx, POND_ERR_STATE, "pond_get_fields: no state set");
  pond_status s = pond_sync(ctx);
  if (s != POND_OK) return s;
  double* dst[5] = {rho, ux, uy, uz, T};
  const Grid& g = ctx->g;
  if (ctx->split) {
    std::vector<float> tmp((size_t)g.n);
    for (int k = 0; k < 4; k++)
      if (dst[k]) {
        POND_HIP(ctx, hipMemcpy(tmp.data(), ctx->sp[ctx->cur].m + (size_t)k * g.plane, (size_t)g.n * sizeof(float),
                                hipMemcpyDeviceToHost));
        for (int64_t c = 0; c < g.n; c++) dst[k][c] = tmp[c];
      }
    if (T) POND_HIP(ctx, hipMemcpy(T, ctx->sp[ctx->cur].T, (size_t)g.n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    for (int k = 0; k < 5; k++)
      if (dst[k])
        POND_HIP(ctx, hipMemcpy(dst[k], ctx->f64[ctx->cur].m + (size_t)k * g.plane, (size_t)g.n * sizeof(double),
                                hipMemcpyDeviceToHost));
  }
  if (niter) POND_HIP(ctx, hipMemcpy(niter, ctx->niter, (size_t)g.n * sizeof(int), hipMemcpyDeviceToHost));
  return POND_OK;
}

pond_status pond_get_populations(pond_ctx* ctx, double* f) {
  if (!ctx) return POND_ERR_INVALID;
  if (!ctx->has_state) return fail(ctx, POND_ERR_STATE, "pond_get_populations: no state set");
  if (!f) return fail(ctx, POND_ERR_INVALID, "pond_get_populations: null f");
  pond_status s = pond_sync(ctx);
  if (s != POND_OK) return s;
  const Grid& g = ctx->g;
  if (ctx->split) {
    const int64_t fcount = split_fcount(ctx->p.lattice, ctx->Qn, g.plane);
    std::vector<float> fs((size_t)fcount);
    POND_HIP(ctx, hipMemcpy(fs.data(), ctx->sp[ctx->cur].f, (size_t)fcount * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < ctx->Qn; i++)
      for (int64_t c = 0; c < g.n; c++) f[(size_t)i * g.n + c] = fs[split_fidx(ctx->p.lattice, i, c, g.plane)];
    return POND_OK;
  }
  const int64_t fcount = f64_fcount(ctx->p.lattice, ctx->Qn, g.plane);
  std::vector<double> fs((size_t)fcount);
  POND_HIP(ctx, hipMemcpy(fs.data(), ctx->f64[ctx->cur].f, (size_t)fcount * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < ctx->Qn; i++)
    for (int64_t c = 0; c < g.n; c++) f[(size_t)i * g.n + c] = fs[f64_fidx(ctx->p.lattice, i, c, g.plane)];
  return POND_OK;
}

pond_status pond_diagnostics(pond_ctx* ctx, pond_diag* out) {
  if (!ctx || !out) return POND_ERR_INVALID;
  if (!ctx->has_state) return fail(ctx, POND_ERR_STATE, "pond_diagnostics: no state set");
  const Grid& g = ctx->g;
  auto* part = reinterpret_cast<RedOut*>(ctx->red);
  const int nb = (int)std::min<int64_t>(kRedBlocks, (g.n + 255) / 256);
  // largest velocity component of the lattice, in thermal units
  double emax = 0;
  for (int i = 0; i < ctx->Qn; i++)
    for (int d = 0; d < 3; d++) emax = std::max(emax, std::fabs(ctx->L.ef[i][d]));
  if (ctx->split)
    hipLaunchKernelGGL(k_reduce_split, dim3(nb), dim3(256), 0, ctx->stream, ctx->sp[ctx->cur].m, ctx->sp[ctx->cur].T,
                       ctx->niter, g.n, g.plane, ctx->L.Dim, 2 * M_PI / g.nx, 2 * M_PI / g.ny, 2 * M_PI / g.nz,
                       1.0 / ctx->L.TLat, emax, part);
  else
    hipLaunchKernelGGL(k_reduce_f64, dim3(nb), dim3(256), 0, ctx->stream, ctx->f64[ctx->cur].m, ctx->niter,
                       (const int*)ctx->cell_set, g.n, g.plane, ctx->L.Dim, 2 * M_PI / g.nx, 2 * M_PI / g.ny,
                       2 * M_PI / g.nz, 1.0 / ctx->L.TLat, emax, part);
  POND_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, ctx->stream, part, nb, part + kRedBlocks);
  POND_HIP(ctx, hipGetLastError());
  RedOut r;
  POND_HIP(ctx, hipMemcpyAsync(&r, part + kRedBlocks, sizeof(RedOut), hipMemcpyDeviceToHost, ctx->stream));
  pond_status s = pond_sync(ctx);
  if (s != POND_OK) return s;
  out->mass = r.s[0];
  out->momentum_x = r.s[1];
  out->energy = r.s[2];
  out->k = r.s[3];
  out->niter_mean = ctx->step > 0 && r.s[5] > 0 ? r.s[4] / r.s[5] : 0.0;
  out->niter_max = ctx->step > 0 ? r.imax : 0;
  out->step = ctx->step;
  out->max_departure = r.vmax;
  return POND_OK;
}

pond_status pond_timing_info(pond_ctx* ctx, pond_timing* out) {
  if (!ctx || !out) return POND_ERR_INVALID;
  pond_status s = pond_sync(ctx);
  if (s != POND_OK) return s;
  out->last_ms = ctx->last_ms;
  out->total_ms = ctx->total_ms;
  out->steps = ctx->timed_steps;
  out->last_mlp_ms = ctx->last_mlp_ms;
  return POND_OK;
}

int64_t pond_device_bytes(const pond_ctx* ctx) { return ctx ? ctx->bytes : 0; }

// drop() (drop.cu:2-34): five files Rho/Vx/Vy/Vz/T_%07d.arr, header
// {int comment_len=0, int dim=3, int sizeof(double), int Nx, int Ny, int Nz}, then the field
// z-outer / x-fastest
pond_status pond_drop(pond_ctx* ctx, const char* dir) {
  if (!ctx || !dir) return POND_ERR_INVALID;
  const Grid& g = ctx->g;
  std::vector<double> buf((size_t)g.n);
  static const char* names[5] = {"Rho", "Vx", "Vy", "Vz", "T"};
  for (int k = 0; k < 5; k++) {
    double* dst[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    dst[k] = buf.data();
    pond_status s = pond_get_fields(ctx, dst[0], dst[1], dst[2], dst[3], dst[4], nullptr);
    if (s != POND_OK) return s;
    char path[4096];
    std::snprintf(path, sizeof(path), "%s/%s_%07d.arr", dir, names[k], (int)ctx->step);
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(ctx, POND_ERR_IO, std::string("pond_drop: cannot open ") + path);
    const int32_t hdr[6] = {0, 3, (int32_t)sizeof(double), g.nx, g.ny, g.nz};
    bool ok = std::fwrite(hdr, sizeof(int32_t), 6, fp) == 6;
    ok = ok && std::fwrite(buf.data(), sizeof(double), (size_t)g.n, fp) == (size_t)g.n;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return fail(ctx, POND_ERR_IO, std::string("pond_drop: write failed: ") + path);
  }
  return POND_OK;
}

// ---- checkpoint: the full state in the device layout, for an exact restart
namespace {
struct CkptHeader {
  char magic[8];  // "PONDCKPT"
  int32_t version, lattice, precision, interp;
  int64_t nx, ny, nz, plane, step;
  int64_t fbytes, mbytes, tbytes, nbytes;  // section sizes: populations, moments, T, niter
  // contexts with switchable cells (D1Q5 FP64): the cell flags, and the other buffer's
  // populations and moments, whose switched-off cells hold their own (older) records
  int64_t sbytes;
};

struct CkptSections {
  void* f; void* m; void* T;
  int64_t fbytes, mbytes, tbytes, nbytes;
};

CkptSections ckpt_sections(pond_ctx* ctx, int b) {
  const Grid& g = ctx->g;
  CkptSections c{};
  if (ctx->split) {
    c.f = ctx->sp[b].f; c.m = ctx->sp[b].m; c.T = ctx->sp[b].T;
    c.fbytes = split_fcount(ctx->p.lattice, ctx->Qn, g.plane) * (int64_t)sizeof(float);
    c.mbytes = 4 * g.plane * (int64_t)sizeof(float);
    c.tbytes = g.plane * (int64_t)sizeof(double);
  } else {
    c.f = ctx->f64[b].f; c.m = ctx->f64[b].m; c.T = nullptr;
    c.fbytes = f64_fcount(ctx->p.lattice, ctx->Qn, g.plane) * (int64_t)sizeof(double);
    c.mbytes = 5 * g.plane * (int64_t)sizeof(double);
    c.tbytes = 0;
  }
  c.nbytes = g.plane * (int64_t)sizeof(int);
  return c;
}
}  // namespace

pond_status pond_checkpoint_write(pond_ctx* ctx, const char* path) {
  if (!ctx || !path) return POND_ERR_INVALID;
  if (!ctx->has_state) return fail(ctx, POND_ERR_STATE, "pond_checkpoint_write: no state set");
  pond_status s = pond_sync(ctx);
  if (s != POND_OK) return s;
  const Grid& g = ctx->g;
  const CkptSections c = ckpt_sections(ctx, ctx->cur);
  CkptHeader h{};
  std::memcpy(h.magic, "PONDCKPT", 8);
  h.version = 2; h.lattice = ctx->p.lattice; h.precision = ctx->p.precision; h.interp = ctx->p.interp;
  h.nx = g.nx; h.ny = g.ny; h.nz = g.nz; h.plane = g.plane; h.step = ctx->step;
  h.fbytes = c.fbytes; h.mbytes = c.mbytes; h.tbytes = c.tbytes; h.nbytes = c.nbytes;
  h.sbytes = ctx->cell_set ? g.plane * (int64_t)sizeof(int) : 0;
  const CkptSections o = ckpt_sections(ctx, 1 - ctx->cur);
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return fail(ctx, POND_ERR_IO, std::string("pond_checkpoint_write: cannot open ") + path);
  bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1;
  const void* src[7] = {c.f, c.m, c.T, ctx->niter, o.f, o.m, ctx->cell_set};
  const int64_t len[7] = {c.fbytes, c.mbytes, c.tbytes, c.nbytes, h.sbytes ? c.fbytes : 0, h.sbytes ? c.mbytes : 0,
                          h.sbytes};
  std::vector<char> buf;
  for (int k = 0; k < 7 && ok; k++) {
    if (len[k] == 0) continue;
    buf.resize((size_t)len[k]);
    if (hipMemcpy(buf.data(), src[k], (size_t)len[k], hipMemcpyDeviceToHost) != hipSuccess) {
      std::fclose(fp);
      return fail(ctx, POND_ERR_HIP, "pond_checkpoint_write: device copy failed");
    }
    ok = std::fwrite(buf.data(), 1, (size_t)len[k], fp) == (size_t)len[k];
  }
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) return fail(ctx, POND_ERR_IO, std::string("pond_checkpoint_write: write failed: ") + path);
  return POND_OK;
}

pond_status pond_checkpoint_read(pond_ctx* ctx, const char* path) {
  if (!ctx || !path) return POND_ERR_INVALID;
  POND_HIP(ctx, hipStreamSynchronize(ctx->stream));
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return fail(ctx, POND_ERR_IO, std::string("pond_checkpoint_read: cannot open ") + path);
  const Grid& g = ctx->g;
  CkptHeader h{};
  const CkptSections c = ckpt_sections(ctx, 0);
  bool ok = std::fread(&h, sizeof h, 1, fp) == 1;
  if (!ok || std::memcmp(h.magic, "PONDCKPT", 8) != 0 || h.version != 2) {
    std::fclose(fp);
    return fail(ctx, POND_ERR_IO, std::string("pond_checkpoint_read: not a checkpoint: ") + path);
  }
  if (h.lattice != ctx->p.lattice || h.precision != ctx->p.precision || h.nx != g.nx || h.ny != g.ny ||
      h.nz != g.nz || h.plane != g.plane || h.fbytes != c.fbytes || h.mbytes != c.mbytes || h.tbytes != c.tbytes ||
      h.nbytes != c.nbytes || h.sbytes != (ctx->cell_set ? g.plane * (int64_t)sizeof(int) : 0)) {
    std::fclose(fp);
    return fail(ctx, POND_ERR_INVALID, "pond_checkpoint_read: lattice, precision or grid differ from the context");
  }
  const int64_t len[7] = {c.fbytes, c.mbytes, c.tbytes, c.nbytes, h.sbytes ? c.fbytes : 0, h.sbytes ? c.mbytes : 0,
                          h.sbytes};
  std::vector<char> sec[7];
  for (int k = 0; k < 7 && ok; k++) {
    sec[k].resize((size_t)len[k]);
    if (len[k]) ok = std::fread(sec[k].data(), 1, (size_t)len[k], fp) == (size_t)len[k];
  }
  std::fclose(fp);
  if (!ok) return fail(ctx, POND_ERR_IO, std::string("pond_checkpoint_read: truncated file: ") + path);
  // buffer 0 holds the state; buffer 1 the other buffer's records where the file has them, else
  // the state again, as after pond_set_state
  for (int b = 0; b < 2; b++) {
    const CkptSections d = ckpt_sections(ctx, b);
    void* dst[3] = {d.f, d.m, d.T};
    for (int k = 0; k < 3; k++) {
      const int from = (b == 1 && h.sbytes && k < 2) ? 4 + k : k;
      if (len[k]) POND_HIP(ctx, hipMemcpy(dst[k], sec[from].data(), (size_t)len[k], hipMemcpyHostToDevice));
    }
  }
  POND_HIP(ctx, hipMemcpy(ctx->niter, sec[3].data(), (size_t)len[3], hipMemcpyHostToDevice));
  if (h.sbytes) POND_HIP(ctx, hipMemcpy(ctx->cell_set, sec[6].data(), (size_t)len[6], hipMemcpyHostToDevice));
  ctx->cur = 0;
  ctx->step = h.step;
  ctx->has_state = true;
  return POND_OK;
}

pond_status pond_read_arr(const char* path, int64_t* dims, double* out) {
  if (!path) return POND_ERR_INVALID;
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return POND_ERR_IO;
  int32_t clen = 0;
  pond_status st = POND_OK;
  if (std::fread(&clen, sizeof(int32_t), 1, fp) != 1 || clen < 0 || clen % 4 != 0) st = POND_ERR_IO;
  if (st == POND_OK && clen > 0 && std::fseek(fp, clen, SEEK_CUR) != 0) st = POND_ERR_IO;
  int32_t h[5] = {0, 0, 0, 0, 0};
  if (st == POND_OK && std::fread(h, sizeof(int32_t), 5, fp) != 5) st = POND_ERR_IO;
  if (st == POND_OK && (h[0] != 3 || h[1] != (int32_t)sizeof(double) || h[2] < 1 || h[3] < 1 || h[4] < 1))
    st = POND_ERR_IO;
  if (st == POND_OK) {
    if (dims) { dims[0] = h[2]; dims[1] = h[3]; dims[2] = h[4]; }
    const size_t n = (size_t)h[2] * h[3] * h[4];
    if (out && std::fread(out, sizeof(double), n, fp) != n) st = POND_ERR_IO;
  }
  std::fclose(fp);
  return st;
}

}  // extern "C"
