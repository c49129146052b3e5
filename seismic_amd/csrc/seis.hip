// libseis.so - gfx950 2-D acoustic FD propagator (forward / tangent / adjoint),
// batched over (model, shot) wavefields of uneven size.  C-ABI in include/seis.h.
//
// Physics follows core/seis_forward2.py of the reference:
//   p[it+2] = temp1*p[it+1] - temp2*p[it] + alpha*lapg(p[it+1]) (+ source)   :453-460
//   temp1 = 2 + 2*c1*alpha - kappa, temp2 = 1 - kappa, kappa = min_vel*damp*dt :341-344
// The kernels advance the algebraically identical increment form
//   q[n+1] = (1-kappa)*q[n] + alpha*Lf(p[n]) (+ source);  p[n+1] = p[n] + q[n+1]
// with q[n] = p[n] - p[n-1] and Lf(p) = lapg(p) + 2*c1*p evaluated as a sum of
// neighbour-minus-centre differences (4*c2 + 4*c3 = -2*c1).  In float32 this keeps
// the seismogram within ~3e-7 relative L2 of the float64 recursion over 999 steps,
// where the reference's own three-level form loses ~2e-5.
//
// The adjoint (update_p_adjoint, :517-579) is run in gather form:
//   mu[m] = (1-kappa)*mu[m+1] + Lf(alpha*lam[m+1]) + d[m];  lam[m] = lam[m+1] + mu[m]
// with the imaging terms reduced to two per-cell accumulators
//   GA = sum lam[m]*Lf(p[m-1])         (alpha_adjoint + 2*c1*temp1_adjoint)
//   GK = sum lam[m]*(p[m-2] - p[m-1])  (kappa_adjoint = -temp1_adjoint - temp2_adjoint)
// and prep_run_adjoint (:361-377) + unpad_edge_padded_gradient
// (core/seis_numerics.py:14-79) folded into one finishing kernel.
//
// All kernels are compiled with -ffp-contract=off; fused multiply-adds are
// written explicitly so that forward, recompute and fused-tangent passes
// produce bit-identical wavefields.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "seis.h"

namespace {

// ---------------------------------------------------------------- constants
constexpr float kC2 = 4.0f / 3.0f;    // core/seis_forward2.py:80
constexpr float kC3 = -1.0f / 12.0f;  // core/seis_forward2.py:81

constexpr int TX = 64;           // tile width (contiguous x), one wavefront row
constexpr int TZ = 16;           // tile height (z)
constexpr int NTHR = 256;        // 4 wavefronts
constexpr int RZ = TZ / (NTHR / TX);  // rows per thread (4)
constexpr int HALO = 2;
constexpr int LW = TX + 2 * HALO;  // 68
constexpr int LH = TZ + 2 * HALO;  // 20
constexpr int NXCD = 8;            // MI355X: 8 XCDs, blocks dealt round-robin

// Per-wavefield state regions in the float arena.
enum Region {
  R_HIST = 0,  // full history (nt+2 slots) or recompute segment (K+1 slots)
  R_PP,        // forward ping-pong p (2 slots) in checkpoint mode
  R_Q,         // q = p[n] - p[n-1]
  R_DP,        // tangent dp ping-pong (2 slots)
  R_DQ,        // tangent dq
  R_LAM,       // adjoint lambda ping-pong (2 slots)
  R_MU,        // adjoint mu = lam[m] - lam[m+1]
  R_GA,        // imaging accumulator for alpha
  R_GK,        // imaging accumulator for kappa
  R_CKP,       // checkpoints of p[jK] (J slots)
  R_CKQ,       // checkpoints of q[jK+1] (J slots)
  R_COUNT
};

struct ModelDev {
  int nz, nx, NZ, NX;
  int nbc, n_shots, first_wave, active;
  int64_t vec_off;    // into the concatenated parameter vectors
  int64_t cell_off;   // into per-model padded coefficient fields
  int64_t seis_off;   // into the concatenated seismograms
  int64_t seis_len;   // n_shots * nt * n_rcv
  int64_t fold_off;   // into the float64 fold scratch (NZ*nx + NZ per model)
  int rcv_row, rcv_x0, rcv_dx, n_rcv;  // padded row / first padded column
  double dt;          // time step [s]
  double dtdx;        // dt / dx
  double kunit;       // dt * kappa0 / (nbc-1)^2  ->  kappa = min_vel*kunit*j^2
  double kappa0;      // 3 ln(1e7) / (2 (nbc-1) dx)
};

struct WaveDev {
  int model, shot, NZ, NX;
  int tiles_x, src_cell, src_z, src_x;  // src_z/src_x physical
  int64_t seis_off;                      // (t, r) block of this shot
  int64_t off[R_COUNT];
};

struct ParamDev {       // per model, written by the prep kernel
  float kscale;         // min_vel * kunit
  float dkscale;        // dmin_vel * kunit
};

struct WaveParam {      // per wavefield
  float amp;            // (v_src*dt)^2                  core/seis_forward2.py:206
  float amp_d;          // 2*v_src*dv_src*dt^2           core/seis_forward2.py:253
  double ssum;          // sum_t lam[t+2][src]*s[t]      core/seis_forward2.py:314
};

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(SEIS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------- device helpers
__device__ __forceinline__ int wrapi(int i, int n) {
  while (i < 0) i += n;
  while (i >= n) i -= n;
  return i;
}

// Damping index j of padded cell (z, x): AbcCoef2D (:46-59).  Left/right
// strips span all rows, top/bottom strips only interior columns, so corners
// take the x profile; damp = kappa0 * (j/(nbc-1))^2, zero inside.
__device__ __forceinline__ int damp_j(int z, int x, int nz, int nx, int nbc) {
  if (x < nbc) return nbc - 1 - x;
  if (x >= nx + nbc) return x - nx - nbc;
  if (z < nbc) return nbc - 1 - z;
  if (z >= nz + nbc) return z - nz - nbc;
  return 0;
}

// Lf(p) at tile position (r, c) (centre at T[r][c]): lapg(p) + 2*c1*p written
// as neighbour-minus-centre differences, periodic wrap handled by the loader.
__device__ __forceinline__ float lap_full(const float (*T)[LW], int r, int c) {
  const float pc = T[r][c];
  float t1 = (T[r][c + 1] - pc) + (T[r][c - 1] - pc);
  t1 = t1 + (T[r + 1][c] - pc);
  t1 = t1 + (T[r - 1][c] - pc);
  float t2 = (T[r][c + 2] - pc) + (T[r][c - 2] - pc);
  t2 = t2 + (T[r + 2][c] - pc);
  t2 = t2 + (T[r - 2][c] - pc);
  return fmaf(kC2, t1, kC3 * t2);
}

// Stage the (TZ+4) x (TX+4) window around tile (z0, x0) of field F into LDS,
// wrapping periodically (core/seis_forward2.py:395-403).
__device__ __forceinline__ void load_tile(float (*T)[LW], const float* __restrict__ F,
                                          int NZ, int NX, int z0, int x0) {
  for (int i = threadIdx.x; i < LH * LW; i += NTHR) {
    const int r = i / LW, c = i - (i / LW) * LW;
    const int gz = wrapi(z0 + r - HALO, NZ);
    const int gx = wrapi(x0 + c - HALO, NX);
    T[r][c] = F[(int64_t)gz * NX + gx];
  }
}

// Same, for the product alpha*lam (adjoint stencil, :567-575).
__device__ __forceinline__ void load_tile_prod(float (*T)[LW], const float* __restrict__ A,
                                               const float* __restrict__ L, int NZ, int NX,
                                               int z0, int x0) {
  for (int i = threadIdx.x; i < LH * LW; i += NTHR) {
    const int r = i / LW, c = i - (i / LW) * LW;
    const int gz = wrapi(z0 + r - HALO, NZ);
    const int gx = wrapi(x0 + c - HALO, NX);
    const int64_t g = (int64_t)gz * NX + gx;
    T[r][c] = A[g] * L[g];
  }
}

struct BlockInfo {
  int w, z0, x0;
};

__device__ __forceinline__ bool block_info(const int2* __restrict__ blk, const WaveDev* waves,
                                           const ModelDev* models, BlockInfo& bi) {
  const int2 e = blk[blockIdx.x];
  if (e.x < 0) return false;
  const WaveDev& w = waves[e.x];
  if (!models[w.model].active) return false;
  const int tz = e.y / w.tiles_x;
  bi.w = e.x;
  bi.z0 = tz * TZ;
  bi.x0 = (e.y - tz * w.tiles_x) * TX;
  return true;
}

// ---------------------------------------------------------------- kernels
struct FwdArgs {
  const int2* blk;
  const WaveDev* waves;
  const ModelDev* models;
  const ParamDev* params;
  const WaveParam* wparam;
  const float* alpha;    // per-model padded alpha
  const float* dalpha;   // per-model padded dalpha (tangent)
  const float* wavelet;  // s[t], float32
  float* arena;
  float* seis;
  float* dseis;
  int it;                // time step: produces p[it+2]
  int in_reg, in_slot, out_reg, out_slot;  // where p[n] is read / p[n+1] written
  int dp_in_slot, dp_out_slot;             // tangent ping-pong slots
  int ck_p_slot, ck_q_slot;                // checkpoint saves (-1 = none)
  int write_seis;
};

// One time step of the forward (and, with TAN, the fused tangent) recursion
// for every (model, shot) tile of the batch.  update_p :428-462 and
// update_p_diff :470-509 with lapg :381-421 fused in.
template <bool TAN>
__global__ __launch_bounds__(NTHR) void k_fwd_step(FwdArgs a) {
  __shared__ float Tp[LH][LW];
  __shared__ float Td[TAN ? LH : 1][TAN ? LW : 1];
  BlockInfo bi;
  if (!block_info(a.blk, a.waves, a.models, bi)) return;
  const WaveDev& w = a.waves[bi.w];
  const ModelDev& m = a.models[w.model];
  const int NZ = w.NZ, NX = w.NX;
  const int64_t P = (int64_t)NZ * NX;
  const float* pin = a.arena + w.off[a.in_reg] + a.in_slot * P;
  float* pout = a.arena + w.off[a.out_reg] + a.out_slot * P;
  float* q = a.arena + w.off[R_Q];
  load_tile(Tp, pin, NZ, NX, bi.z0, bi.x0);
  const float* dpin = nullptr;
  float* dpout = nullptr;
  float* dq = nullptr;
  if constexpr (TAN) {
    dpin = a.arena + w.off[R_DP] + a.dp_in_slot * P;
    dpout = a.arena + w.off[R_DP] + a.dp_out_slot * P;
    dq = a.arena + w.off[R_DQ];
    load_tile(Td, dpin, NZ, NX, bi.z0, bi.x0);
  }
  __syncthreads();

  const float kscale = a.params[w.model].kscale;
  const float dkscale = TAN ? a.params[w.model].dkscale : 0.f;
  const float s_t = a.wavelet[a.it];
  const float src_add = a.wparam[bi.w].amp * s_t;
  const float src_add_d = TAN ? a.wparam[bi.w].amp_d * s_t : 0.f;
  const float* alpha = a.alpha + m.cell_off;
  const float* dalpha = TAN ? a.dalpha + m.cell_off : nullptr;
  float* ckp = a.ck_p_slot >= 0 ? a.arena + w.off[R_CKP] + a.ck_p_slot * P : nullptr;
  float* ckq = a.ck_q_slot >= 0 ? a.arena + w.off[R_CKQ] + a.ck_q_slot * P : nullptr;

  const int tx = threadIdx.x % TX;
  const int ty = threadIdx.x / TX;
  const int x = bi.x0 + tx;
  if (x >= NX) return;
#pragma unroll
  for (int k = 0; k < RZ; ++k) {
    const int r = ty * RZ + k;
    const int z = bi.z0 + r;
    if (z >= NZ) break;
    const int64_t g = (int64_t)z * NX + x;
    const int j = damp_j(z, x, m.nz, m.nx, m.nbc);
    const float jj = (float)(j * j);
    const float k1 = 1.0f - kscale * jj;
    const float L = lap_full(Tp, r + HALO, tx + HALO);
    const float pc = Tp[r + HALO][tx + HALO];
    const float qo = q[g];
    float qn = fmaf(k1, qo, alpha[g] * L);
    if (g == w.src_cell) qn = qn + src_add;
    const float pn = pc + qn;
    q[g] = qn;
    pout[g] = pn;
    if (ckp) ckp[g] = pn;
    if (ckq) ckq[g] = qn;
    if constexpr (TAN) {
      const float Ld = lap_full(Td, r + HALO, tx + HALO);
      const float dpc = Td[r + HALO][tx + HALO];
      const float dqo = dq[g];
      // dq' = (1-kappa) dq + alpha Lf(dp) + dalpha Lf(p) - dkappa q
      float dqn = fmaf(k1, dqo, alpha[g] * Ld);
      dqn = fmaf(dalpha[g], L, dqn);
      dqn = fmaf(-(dkscale * jj), qo, dqn);
      if (g == w.src_cell) dqn = dqn + src_add_d;
      const float dpn = dpc + dqn;
      dq[g] = dqn;
      dpout[g] = dpn;
      if (a.write_seis && z == m.rcv_row) {
        const int d = x - m.rcv_x0;
        if (d >= 0 && d % m.rcv_dx == 0 && d / m.rcv_dx < m.n_rcv)
          a.dseis[w.seis_off + (int64_t)a.it * m.n_rcv + d / m.rcv_dx] = dpn;
      }
    }
    if (a.write_seis && z == m.rcv_row) {
      const int d = x - m.rcv_x0;
      if (d >= 0 && d % m.rcv_dx == 0 && d / m.rcv_dx < m.n_rcv)
        a.seis[w.seis_off + (int64_t)a.it * m.n_rcv + d / m.rcv_dx] = pn;
    }
  }
}

struct AdjArgs {
  const int2* blk;
  const WaveDev* waves;
  const ModelDev* models;
  const ParamDev* params;
  WaveParam* wparam;
  const float* alpha;
  const float* wavelet;
  float* arena;
  const float* inj_a;   // adjoint source = inj_a - inj_b (inj_b may be null)
  const float* inj_b;
  int m;                // produces lam[m] (m = nt+1 .. 2)
  int lam_in_slot, lam_out_slot;
  int p_reg, p1_slot, p0_slot;  // p[m-1], p[m-2]
};

// One reverse step: update_p_adjoint (:517-579) in gather form, with the
// imaging accumulators and the source-wavelet correlation (:314) fused.
__global__ __launch_bounds__(NTHR) void k_adj_step(AdjArgs a) {
  __shared__ float Ta[LH][LW];  // alpha*lam[m+1]
  __shared__ float Tp[LH][LW];  // p[m-1]
  BlockInfo bi;
  if (!block_info(a.blk, a.waves, a.models, bi)) return;
  const WaveDev& w = a.waves[bi.w];
  const ModelDev& m = a.models[w.model];
  const int NZ = w.NZ, NX = w.NX;
  const int64_t P = (int64_t)NZ * NX;
  const float* alpha = a.alpha + m.cell_off;
  const float* lin = a.arena + w.off[R_LAM] + a.lam_in_slot * P;
  float* lout = a.arena + w.off[R_LAM] + a.lam_out_slot * P;
  float* mu = a.arena + w.off[R_MU];
  float* GA = a.arena + w.off[R_GA];
  float* GK = a.arena + w.off[R_GK];
  const float* p1 = a.arena + w.off[a.p_reg] + a.p1_slot * P;
  const float* p0 = a.arena + w.off[a.p_reg] + a.p0_slot * P;
  load_tile_prod(Ta, alpha, lin, NZ, NX, bi.z0, bi.x0);
  load_tile(Tp, p1, NZ, NX, bi.z0, bi.x0);
  __syncthreads();

  const float kscale = a.params[w.model].kscale;
  const int tsrc = a.m - 2;
  const int tx = threadIdx.x % TX;
  const int ty = threadIdx.x / TX;
  const int x = bi.x0 + tx;
  if (x >= NX) return;
#pragma unroll
  for (int k = 0; k < RZ; ++k) {
    const int r = ty * RZ + k;
    const int z = bi.z0 + r;
    if (z >= NZ) break;
    const int64_t g = (int64_t)z * NX + x;
    const int j = damp_j(z, x, m.nz, m.nx, m.nbc);
    const float k1 = 1.0f - kscale * (float)(j * j);
    const float La = lap_full(Ta, r + HALO, tx + HALO);
    float mun = fmaf(k1, mu[g], La);
    if (z == m.rcv_row) {
      const int d = x - m.rcv_x0;
      if (d >= 0 && d % m.rcv_dx == 0 && d / m.rcv_dx < m.n_rcv) {
        const int64_t si = w.seis_off + (int64_t)tsrc * m.n_rcv + d / m.rcv_dx;
        const float inj = a.inj_b ? a.inj_a[si] - a.inj_b[si] : a.inj_a[si];
        mun = mun + inj;
      }
    }
    const float ln = lin[g] + mun;
    mu[g] = mun;
    lout[g] = ln;
    const float Lp = lap_full(Tp, r + HALO, tx + HALO);
    const float pc = Tp[r + HALO][tx + HALO];
    GA[g] = fmaf(ln, Lp, GA[g]);
    GK[g] = fmaf(ln, p0[g] - pc, GK[g]);
    if (g == w.src_cell)
      a.wparam[bi.w].ssum += (double)ln * (double)a.wavelet[tsrc];
  }
}

// Zero `nslots` slots of region `reg` of every wavefield (graph-safe fill;
// no hipMemset nodes are recorded in captured graphs).
__global__ __launch_bounds__(NTHR) void k_fill(const WaveDev* waves, const ModelDev* models,
                                                float* arena, int reg, int slot0, int nslots) {
  const WaveDev& w = waves[blockIdx.y];
  if (!models[w.model].active) return;
  const int64_t P = (int64_t)w.NZ * w.NX;
  float* base = arena + w.off[reg] + slot0 * P;
  const int64_t n = P * nslots;
  for (int64_t i = (int64_t)blockIdx.x * NTHR + threadIdx.x; i < n; i += (int64_t)gridDim.x * NTHR)
    base[i] = 0.f;
}

// Copy slot `src_slot` of region `src_reg` into `dst_reg`/`dst_slot`, and
// optionally dst2 = dst + add (restores p[jK+1] = p[jK] + q[jK+1]).
__global__ __launch_bounds__(NTHR) void k_restore(const WaveDev* waves, const ModelDev* models,
                                                   float* arena, int j) {
  const WaveDev& w = waves[blockIdx.y];
  if (!models[w.model].active) return;
  const int64_t P = (int64_t)w.NZ * w.NX;
  const float* ckp = arena + w.off[R_CKP] + j * P;
  const float* ckq = arena + w.off[R_CKQ] + j * P;
  float* s0 = arena + w.off[R_HIST];
  float* s1 = s0 + P;
  float* q = arena + w.off[R_Q];
  for (int64_t i = (int64_t)blockIdx.x * NTHR + threadIdx.x; i < P; i += (int64_t)gridDim.x * NTHR) {
    const float pv = ckp[i];
    const float qv = ckq[i];
    s0[i] = pv;
    q[i] = qv;
    s1[i] = pv + qv;
  }
}

// prep_run (:333-344) and prep_run_diff (:346-359): padded alpha / dalpha
// from the edge-padded velocity (np.pad mode='edge' == index clamping).
__global__ __launch_bounds__(NTHR) void k_prep(const ModelDev* models, const float* vec,
                                                const float* dvec, float* alpha, float* dalpha) {
  const ModelDev& m = models[blockIdx.y];
  if (!m.active) return;
  const float* v = vec + m.vec_off;
  const float* dv = dvec ? dvec + m.vec_off : nullptr;
  const int64_t P = (int64_t)m.NZ * m.NX;
  const double dtdx2 = m.dtdx * m.dtdx;
  for (int64_t i = (int64_t)blockIdx.x * NTHR + threadIdx.x; i < P; i += (int64_t)gridDim.x * NTHR) {
    const int z = (int)(i / m.NX);
    const int x = (int)(i - (int64_t)z * m.NX);
    const int iz = min(max(z - m.nbc, 0), m.nz - 1);
    const int ix = min(max(x - m.nbc, 0), m.nx - 1);
    const double vv = v[iz * m.nx + ix];
    alpha[m.cell_off + i] = (float)(vv * vv * dtdx2);
    if (dv) dalpha[m.cell_off + i] = (float)(2.0 * vv * (double)dv[iz * m.nx + ix] * dtdx2);
  }
}

// Per-model damping scale (min_vel * damp * dt, :339-342) and per-shot source
// amplitudes (:206-207, :253-254); also clears the adjoint source correlation.
__global__ void k_params(const ModelDev* models, const WaveDev* waves, int n_waves,
                         ParamDev* params, WaveParam* wparam, const float* vec,
                         const float* dvec) {
  const int wi = blockIdx.x * blockDim.x + threadIdx.x;
  if (wi >= n_waves) return;
  const WaveDev& w = waves[wi];
  const ModelDev& m = models[w.model];
  if (!m.active) return;
  const float* v = vec + m.vec_off;
  const double vs = v[w.src_z * m.nx + w.src_x];
  WaveParam& wp = wparam[wi];
  wp.amp = (float)(vs * m.dt * vs * m.dt);
  wp.amp_d = 0.f;
  wp.ssum = 0.0;
  if (dvec) {
    const float* dv = dvec + m.vec_off;
    wp.amp_d = (float)(2.0 * vs * (double)dv[w.src_z * m.nx + w.src_x] * m.dt * m.dt);
  }
  if (w.shot == 0) {
    params[w.model].kscale = (float)((double)v[m.nz * m.nx] * m.kunit);
    params[w.model].dkscale = dvec ? (float)((double)dvec[m.vec_off + m.nz * m.nx] * m.kunit) : 0.f;
  }
}

constexpr int kRedBlocks = 32;  // partial-sum blocks per model (misfit)

// Block-wide fixed-order tree sum of NTHR doubles (deterministic).
__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int k = NTHR / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// First half of prep_run_adjoint (:361-377) + unpad_edge_padded_gradient
// (seis_numerics.py:14-79): for padded row Z, fold the columns that edge
// padding maps onto each physical column (left strip -> column 0, right strip
// -> column nx-1), summing the alpha accumulator over shots in float64.  Also
// the row's share of sum(j^2 * kappa accumulator) for d/d min_vel (:366-370).
__global__ __launch_bounds__(NTHR) void k_fold_rows(const ModelDev* models, const WaveDev* waves,
                                                     const float* arena, double* fold) {
  __shared__ double red[NTHR];
  const ModelDev& m = models[blockIdx.y];
  if (!m.active || (int)blockIdx.x >= m.NZ) return;
  const int Z = blockIdx.x;
  double* R = fold + m.fold_off + (int64_t)Z * m.nx;
  double* rowK = fold + m.fold_off + (int64_t)m.NZ * m.nx;
  double sl = 0.0, sr = 0.0, sk = 0.0;
  for (int X = threadIdx.x; X < m.NX; X += NTHR) {
    const int64_t g = (int64_t)Z * m.NX + X;
    double a = 0.0, k = 0.0;
    for (int s = 0; s < m.n_shots; ++s) {
      const WaveDev& w = waves[m.first_wave + s];
      a += (double)arena[w.off[R_GA] + g];
      k += (double)arena[w.off[R_GK] + g];
    }
    const int j = damp_j(Z, X, m.nz, m.nx, m.nbc);
    sk += (double)(j * j) * k;
    const int ix = X - m.nbc;
    if (ix <= 0) sl += a;
    else if (ix >= m.nx - 1) sr += a;
    else R[ix] = a;
  }
  sl = block_sum(sl, red);
  sr = block_sum(sr, red);
  sk = block_sum(sk, red);
  if (threadIdx.x == 0) {
    R[0] = sl;
    R[m.nx - 1] = sr;
    rowK[Z] = sk;
  }
}

// Second half: fold rows (top strip -> row 0, bottom strip -> row nz-1),
// apply v_adj = 2 v (dt/dx)^2 alpha_adj (:367-369), add the source-amplitude
// term 2 dt^2 v_src sum_t s_adj[t] s[t] (:314) and d/d min_vel (:370-374).
__global__ __launch_bounds__(NTHR) void k_fold_cols(const ModelDev* models, const WaveDev* waves,
                                                     const WaveParam* wparam, const double* fold,
                                                     const float* vec, double* grad) {
  const ModelDev& m = models[blockIdx.y];
  if (!m.active) return;
  const double* R = fold + m.fold_off;
  const int64_t go = m.vec_off;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double* rowK = R + (int64_t)m.NZ * m.nx;
    double sk = 0.0;
    for (int Z = 0; Z < m.NZ; ++Z) sk += rowK[Z];
    const double nb1 = (double)(m.nbc - 1);
    grad[go + (int64_t)m.nz * m.nx] = m.dt * m.kappa0 / (nb1 * nb1) * sk;
  }
  const int n = m.nz * m.nx;
  const int i = blockIdx.x * NTHR + threadIdx.x;
  if (i >= n) return;
  const int iz = i / m.nx, ix = i - (i / m.nx) * m.nx;
  const int z_lo = iz == 0 ? 0 : iz + m.nbc;
  const int z_hi = iz == m.nz - 1 ? m.NZ - 1 : iz + m.nbc;
  double acc = 0.0;
  for (int Z = z_lo; Z <= z_hi; ++Z) acc += R[(int64_t)Z * m.nx + ix];
  const double v = vec[m.vec_off + i];
  double gv = 2.0 * v * m.dtdx * m.dtdx * acc;
  for (int s = 0; s < m.n_shots; ++s) {
    const WaveDev& w = waves[m.first_wave + s];
    if (w.src_z == iz && w.src_x == ix) gv += 2.0 * m.dt * m.dt * v * wparam[m.first_wave + s].ssum;
  }
  grad[go + i] = gv;
}

// sum((seis - dd)^2) per model (cost_and_gradient, core/seis_invert.py:41,
// before the mean): kRedBlocks fixed-order partials, then one final sum.
__global__ __launch_bounds__(NTHR) void k_misfit_part(const ModelDev* models, const float* seis,
                                                       const float* dd, double* part) {
  __shared__ double red[NTHR];
  const ModelDev& m = models[blockIdx.y];
  if (!m.active) return;
  const float* s = seis + m.seis_off;
  const float* d = dd + m.seis_off;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * NTHR + threadIdx.x; i < m.seis_len;
       i += (int64_t)kRedBlocks * NTHR) {
    const double r = (double)s[i] - (double)d[i];
    acc += r * r;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * kRedBlocks + blockIdx.x] = acc;
}

__global__ void k_misfit_final(const ModelDev* models, int n_models, const double* part, double* out) {
  const int mi = blockIdx.x * blockDim.x + threadIdx.x;
  if (mi >= n_models || !models[mi].active) return;
  double acc = 0.0;
  for (int b = 0; b < kRedBlocks; ++b) acc += part[(int64_t)mi * kRedBlocks + b];
  out[mi] = acc;
}

}  // namespace

// ======================================================================== host
namespace {

constexpr int kFillBlocks = 128;

__global__ __launch_bounds__(NTHR) void k_copy_active(const ModelDev* models, const float* src,
                                                       float* dst) {
  const ModelDev& m = models[blockIdx.y];
  if (!m.active) return;
  for (int64_t i = (int64_t)blockIdx.x * NTHR + threadIdx.x; i < m.seis_len;
       i += (int64_t)gridDim.x * NTHR)
    dst[m.seis_off + i] = src[m.seis_off + i];
}

int64_t align64(int64_t n) { return (n + 63) & ~int64_t(63); }

struct Layout {
  std::vector<ModelDev> models;
  std::vector<WaveDev> waves;
  std::vector<int64_t> vec_off, seis_off;
  int64_t cells = 0, arena = 0, seis = 0, vec = 0, fold = 0;
  int K = 0, J = 0;
  bool full = true, fwd_hist = true;
  int n_blk = 0, max_P = 0, max_cells_phys = 0;
  std::vector<int2> blk;
  int64_t bytes = 0;
};

int validate_and_layout(const seis_model_desc* md, int n_models, const seis_time_desc* td,
                        int ckpt_every, int caps, Layout& L) {
  if (!md || n_models <= 0) return fail(SEIS_EINVAL, "n_models must be positive");
  if (!td || td->nt < 1) return fail(SEIS_EINVAL, "nt must be >= 1");
  if (!(td->dt > 0.f)) return fail(SEIS_EINVAL, "dt must be positive");
  if (td->nbc < 2) return fail(SEIS_EINVAL, "nbc must be >= 2");
  if (ckpt_every < 0 || ckpt_every == 1) return fail(SEIS_EINVAL, "ckpt_every must be 0 or >= 2");
  if (!(caps & SEIS_CAP_FORWARD)) return fail(SEIS_EINVAL, "caps must include SEIS_CAP_FORWARD");
  const bool tan = caps & SEIS_CAP_TANGENT, adj = caps & SEIS_CAP_ADJOINT;
  const int nt = td->nt, nbc = td->nbc;
  L.full = ckpt_every == 0 || ckpt_every >= nt;
  L.K = L.full ? 0 : ckpt_every;
  L.J = L.full ? 0 : (nt + L.K - 1) / L.K;
  L.fwd_hist = L.full && adj;
  std::vector<int2> logical;
  for (int mi = 0; mi < n_models; ++mi) {
    const seis_model_desc& d = md[mi];
    if (d.nz < 2 || d.nx < 2) return fail(SEIS_EINVAL, "model " + std::to_string(mi) + ": nz, nx must be >= 2");
    if (!(d.dx > 0.f)) return fail(SEIS_EINVAL, "model " + std::to_string(mi) + ": dx must be positive");
    if (d.n_shots < 1 || d.n_shots > SEIS_MAX_SHOTS)
      return fail(SEIS_EINVAL, "model " + std::to_string(mi) + ": n_shots out of range");
    if (d.src_z < 0 || d.src_z >= d.nz || d.rcv_z < 0 || d.rcv_z >= d.nz)
      return fail(SEIS_EINVAL, "model " + std::to_string(mi) + ": source/receiver row outside grid");
    if (d.n_rcv < 1 || d.rcv_dx < 1 || d.rcv_x0 < 0 || d.rcv_x0 + (int64_t)(d.n_rcv - 1) * d.rcv_dx >= d.nx)
      return fail(SEIS_EINVAL, "model " + std::to_string(mi) + ": receivers outside grid");
    ModelDev m{};
    m.nz = d.nz; m.nx = d.nx; m.NZ = d.nz + 2 * nbc; m.NX = d.nx + 2 * nbc;
    m.nbc = nbc; m.n_shots = d.n_shots; m.first_wave = (int)L.waves.size(); m.active = 1;
    m.vec_off = L.vec; m.cell_off = L.cells; m.seis_off = L.seis;
    m.seis_len = (int64_t)d.n_shots * nt * d.n_rcv;
    m.rcv_row = d.rcv_z + nbc; m.rcv_x0 = d.rcv_x0 + nbc; m.rcv_dx = d.rcv_dx; m.n_rcv = d.n_rcv;
    m.dt = td->dt; m.dtdx = (double)td->dt / (double)d.dx;
    // AbcCoef2D (:49-51): a = (nbc-1) dx, kappa0 = 3 ln(1e7) / (2a)
    m.kappa0 = 3.0 * std::log(1e7) / (2.0 * (nbc - 1) * (double)d.dx);
    m.kunit = (double)td->dt * m.kappa0 / ((double)(nbc - 1) * (nbc - 1));
    const int64_t P = (int64_t)m.NZ * m.NX;
    if (P > INT32_MAX / 2) return fail(SEIS_EINVAL, "model too large");
    L.max_P = std::max<int>(L.max_P, (int)P);
    L.max_cells_phys = std::max(L.max_cells_phys, d.nz * d.nx);
    L.vec_off.push_back(L.vec);
    L.seis_off.push_back(L.seis);
    L.vec += (int64_t)d.nz * d.nx + 1;
    m.fold_off = L.fold;
    L.fold += align64((int64_t)m.NZ * m.nx + m.NZ);
    L.cells += align64(P);
    L.seis += m.seis_len;
    const int tiles_x = (m.NX + TX - 1) / TX, tiles_z = (m.NZ + TZ - 1) / TZ;
    for (int s = 0; s < d.n_shots; ++s) {
      if (d.src_x[s] < 0 || d.src_x[s] >= d.nx)
        return fail(SEIS_EINVAL, "model " + std::to_string(mi) + ": source column outside grid");
      WaveDev w{};
      w.model = mi; w.shot = s; w.NZ = m.NZ; w.NX = m.NX; w.tiles_x = tiles_x;
      w.src_z = d.src_z; w.src_x = d.src_x[s];
      w.src_cell = (d.src_z + nbc) * m.NX + (d.src_x[s] + nbc);
      w.seis_off = m.seis_off + (int64_t)s * nt * d.n_rcv;
      auto region = [&](int r, int64_t slots) {
        w.off[r] = L.arena;
        L.arena += align64(P) * slots;
      };
      region(R_HIST, L.fwd_hist ? nt + 2 : (adj ? L.K + 1 : 0));
      region(R_PP, L.fwd_hist ? 0 : 2);
      region(R_Q, 1);
      region(R_DP, tan ? 2 : 0);
      region(R_DQ, tan ? 1 : 0);
      region(R_LAM, adj ? 2 : 0);
      region(R_MU, adj ? 1 : 0);
      region(R_GA, adj ? 1 : 0);
      region(R_GK, adj ? 1 : 0);
      region(R_CKP, adj && !L.full ? L.J : 0);
      region(R_CKQ, adj && !L.full ? L.J : 0);
      const int wi = (int)L.waves.size();
      for (int t = 0; t < tiles_x * tiles_z; ++t) logical.push_back(make_int2(wi, t));
      L.waves.push_back(w);
    }
    L.models.push_back(m);
  }
  L.vec_off.push_back(L.vec);
  L.seis_off.push_back(L.seis);
  // XCD-aware block order: hardware deals block b to XCD b % 8, so physical
  // block b takes logical tile (b % 8) * (n/8) + b / 8 and consecutive tiles of
  // one wavefield (which share halos) stay in one XCD's L2.
  const int nlog = (int)logical.size();
  L.n_blk = (nlog + NXCD - 1) / NXCD * NXCD;
  const int per = L.n_blk / NXCD;
  L.blk.assign(L.n_blk, make_int2(-1, 0));
  for (int b = 0; b < L.n_blk; ++b) {
    const int lg = (b % NXCD) * per + b / NXCD;
    if (lg < nlog) L.blk[b] = logical[lg];
  }
  const int64_t nw = (int64_t)L.waves.size();
  L.bytes = (int64_t)sizeof(float) * (L.arena + L.cells * (tan ? 2 : 1) + L.seis * (1 + (tan ? 1 : 0) + (adj ? 1 : 0)) + nt) +
            (int64_t)sizeof(double) * (adj ? L.fold : 0) + (int64_t)sizeof(double) * 32 * n_models +
            (int64_t)(sizeof(ModelDev) + 2 * sizeof(int64_t) + sizeof(ParamDev)) * n_models +
            (int64_t)(sizeof(WaveDev) + sizeof(WaveParam)) * nw + (int64_t)sizeof(int2) * L.n_blk;
  return SEIS_OK;
}

}  // namespace

struct seis_plan {
  Layout L;
  int n_models = 0, n_waves = 0, nt = 0, caps = 0;
  ModelDev* d_models = nullptr;
  WaveDev* d_waves = nullptr;
  ParamDev* d_params = nullptr;
  WaveParam* d_wparam = nullptr;
  int2* d_blk = nullptr;
  double* d_fold = nullptr;
  double* d_part = nullptr;
  float* d_alpha = nullptr;
  float* d_dalpha = nullptr;
  float* d_wavelet = nullptr;
  float* d_arena = nullptr;
  float* d_seis = nullptr;
  float* d_dseis = nullptr;
  float* d_dd = nullptr;
  hipStream_t cap = nullptr;
  hipGraphExec_t g_fwd[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipGraphExec_t g_adj[3] = {nullptr, nullptr, nullptr};
};

namespace {

void plan_free(seis_plan* p) {
  for (auto& a : p->g_fwd)
    for (auto& g : a)
      if (g) (void)hipGraphExecDestroy(g);
  for (auto& g : p->g_adj)
    if (g) (void)hipGraphExecDestroy(g);
  void* bufs[] = {p->d_models, p->d_waves, p->d_params, p->d_wparam, p->d_blk, p->d_fold, p->d_part,
                  p->d_alpha, p->d_dalpha, p->d_wavelet, p->d_arena, p->d_seis, p->d_dseis, p->d_dd};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  delete p;
}

template <typename T>
int dev_alloc(T** ptr, int64_t count) {
  if (count <= 0) { *ptr = nullptr; return SEIS_OK; }
  hipError_t e = hipMalloc((void**)ptr, sizeof(T) * count);
  if (e != hipSuccess) {
    *ptr = nullptr;
    (void)hipGetLastError();
    return fail(SEIS_ENOMEM, "hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed: " +
                                 hipGetErrorString(e));
  }
  return SEIS_OK;
}

FwdArgs fwd_args(seis_plan* p) {
  FwdArgs a{};
  a.blk = p->d_blk; a.waves = p->d_waves; a.models = p->d_models; a.params = p->d_params;
  a.wparam = p->d_wparam; a.alpha = p->d_alpha; a.dalpha = p->d_dalpha; a.wavelet = p->d_wavelet;
  a.arena = p->d_arena; a.seis = p->d_seis; a.dseis = p->d_dseis;
  a.ck_p_slot = a.ck_q_slot = -1;
  return a;
}

void fill(seis_plan* p, hipStream_t s, int reg, int slot0, int nslots) {
  hipLaunchKernelGGL(k_fill, dim3(kFillBlocks, p->n_waves), dim3(NTHR), 0, s, p->d_waves,
                     p->d_models, p->d_arena, reg, slot0, nslots);
}

// Forward time loop (with the fused tangent when `tan`), saving checkpoints in
// checkpoint mode when `save`.  core/seis_forward2.py:211-247 / :255-283.
void enqueue_forward(seis_plan* p, hipStream_t s, bool tan, bool save) {
  const Layout& L = p->L;
  fill(p, s, R_Q, 0, 1);
  if (L.fwd_hist) {
    fill(p, s, R_HIST, 0, 2);
  } else {
    fill(p, s, R_PP, 0, 2);
    if (save && (p->caps & SEIS_CAP_ADJOINT)) {
      fill(p, s, R_CKP, 0, 1);
      fill(p, s, R_CKQ, 0, 1);
    }
  }
  if (tan) {
    fill(p, s, R_DP, 0, 2);
    fill(p, s, R_DQ, 0, 1);
  }
  FwdArgs a = fwd_args(p);
  a.write_seis = 1;
  const bool ck = save && !L.full && (p->caps & SEIS_CAP_ADJOINT);
  for (int it = 0; it < p->nt; ++it) {
    const int n = it + 1;
    a.it = it;
    if (L.fwd_hist) {
      a.in_reg = R_HIST; a.in_slot = n; a.out_reg = R_HIST; a.out_slot = n + 1;
    } else {
      a.in_reg = R_PP; a.in_slot = n & 1; a.out_reg = R_PP; a.out_slot = (n + 1) & 1;
    }
    a.dp_in_slot = n & 1;
    a.dp_out_slot = (n + 1) & 1;
    a.ck_p_slot = a.ck_q_slot = -1;
    if (ck) {
      if ((n + 1) % L.K == 0 && (n + 1) / L.K < L.J) a.ck_p_slot = (n + 1) / L.K;
      if (n % L.K == 0 && n / L.K >= 1 && n / L.K < L.J) a.ck_q_slot = n / L.K;
    }
    if (tan)
      hipLaunchKernelGGL(k_fwd_step<true>, dim3(L.n_blk), dim3(NTHR), 0, s, a);
    else
      hipLaunchKernelGGL(k_fwd_step<false>, dim3(L.n_blk), dim3(NTHR), 0, s, a);
  }
}

// Reverse time loop: update_p_adjoint (:292-313) for m = nt+1 .. 2, with
// segment recompute from checkpoints in checkpoint mode.
void enqueue_adjoint(seis_plan* p, hipStream_t s, int kind) {
  const Layout& L = p->L;
  fill(p, s, R_LAM, 0, 2);
  fill(p, s, R_MU, 0, 1);
  fill(p, s, R_GA, 0, 1);
  fill(p, s, R_GK, 0, 1);
  AdjArgs a{};
  a.blk = p->d_blk; a.waves = p->d_waves; a.models = p->d_models; a.params = p->d_params;
  a.wparam = p->d_wparam; a.alpha = p->d_alpha; a.wavelet = p->d_wavelet; a.arena = p->d_arena;
  if (kind == 0) { a.inj_a = p->d_seis; a.inj_b = p->d_dd; }
  else if (kind == 1) { a.inj_a = p->d_dd; a.inj_b = nullptr; }
  else { a.inj_a = p->d_dseis; a.inj_b = nullptr; }
  const int nt = p->nt;
  auto adj_step = [&](int m, int reg, int base) {
    a.m = m;
    a.lam_in_slot = (m + 1) & 1;
    a.lam_out_slot = m & 1;
    a.p_reg = reg;
    a.p1_slot = m - 1 - base;
    a.p0_slot = m - 2 - base;
    hipLaunchKernelGGL(k_adj_step, dim3(L.n_blk), dim3(NTHR), 0, s, a);
  };
  if (L.full) {
    for (int m = nt + 1; m >= 2; --m) adj_step(m, R_HIST, 0);
    return;
  }
  FwdArgs f = fwd_args(p);
  f.write_seis = 0;
  f.in_reg = f.out_reg = R_HIST;
  for (int j = L.J - 1; j >= 0; --j) {
    const int lo = j * L.K;
    const int hi = std::min(lo + L.K, nt);  // largest p index this segment needs
    hipLaunchKernelGGL(k_restore, dim3(kFillBlocks, p->n_waves), dim3(NTHR), 0, s, p->d_waves,
                       p->d_models, p->d_arena, j);
    for (int n = lo + 1; n <= hi - 1; ++n) {
      f.it = n - 1;
      f.in_slot = n - lo;
      f.out_slot = n + 1 - lo;
      hipLaunchKernelGGL(k_fwd_step<false>, dim3(L.n_blk), dim3(NTHR), 0, s, f);
    }
    const int m_hi = std::min(lo + L.K + 1, nt + 1);
    for (int m = m_hi; m >= lo + 2; --m) adj_step(m, R_HIST, lo);
  }
}

template <typename F>
int get_graph(seis_plan* p, hipGraphExec_t* slot, F&& body) {
  if (*slot) return SEIS_OK;
  HIP_TRY(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
  body(p->cap);
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamEndCapture(p->cap, &g));
  hipError_t e = hipGraphInstantiate(slot, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    *slot = nullptr;
    return fail(SEIS_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return SEIS_OK;
}

int run_prep(seis_plan* p, hipStream_t s, const float* vec, const float* dvec) {
  const Layout& L = p->L;
  const int bx = std::min(256, (L.max_P + NTHR - 1) / NTHR);
  hipLaunchKernelGGL(k_prep, dim3(bx, p->n_models), dim3(NTHR), 0, s, p->d_models, vec, dvec,
                     p->d_alpha, p->d_dalpha);
  hipLaunchKernelGGL(k_params, dim3((p->n_waves + 63) / 64), dim3(64), 0, s, p->d_models,
                     p->d_waves, p->n_waves, p->d_params, p->d_wparam, vec, dvec);
  HIP_TRY(hipGetLastError());
  return SEIS_OK;
}

int finish_grad(seis_plan* p, hipStream_t s, const float* vec, double* grad) {
  const Layout& L = p->L;
  int max_nz_pad = 0;
  for (const auto& m : L.models) max_nz_pad = std::max(max_nz_pad, m.NZ);
  hipLaunchKernelGGL(k_fold_rows, dim3(max_nz_pad, p->n_models), dim3(NTHR), 0, s, p->d_models,
                     p->d_waves, p->d_arena, p->d_fold);
  hipLaunchKernelGGL(k_fold_cols, dim3((L.max_cells_phys + NTHR - 1) / NTHR, p->n_models),
                     dim3(NTHR), 0, s, p->d_models, p->d_waves, p->d_wparam, p->d_fold, vec, grad);
  HIP_TRY(hipGetLastError());
  return SEIS_OK;
}

void launch_misfit(seis_plan* p, hipStream_t s, double* out) {
  hipLaunchKernelGGL(k_misfit_part, dim3(kRedBlocks, p->n_models), dim3(NTHR), 0, s, p->d_models,
                     p->d_seis, p->d_dd, p->d_part);
  hipLaunchKernelGGL(k_misfit_final, dim3((p->n_models + 63) / 64), dim3(64), 0, s, p->d_models,
                     p->n_models, p->d_part, out);
}

int copy_out(seis_plan* p, hipStream_t s, const float* src, float* dst) {
  if (!dst) return SEIS_OK;
  hipLaunchKernelGGL(k_copy_active, dim3(64, p->n_models), dim3(NTHR), 0, s, p->d_models, src, dst);
  HIP_TRY(hipGetLastError());
  return SEIS_OK;
}

#define TRY(expr)                 \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != SEIS_OK) return rc_; \
  } while (0)

}  // namespace

extern "C" {

const char* seis_last_error(void) { return g_last_error.c_str(); }

int64_t seis_plan_bytes(const seis_model_desc* models, int32_t n_models,
                        const seis_time_desc* time, int32_t ckpt_every, int32_t caps) {
  Layout L;
  int rc = validate_and_layout(models, n_models, time, ckpt_every, caps, L);
  return rc == SEIS_OK ? L.bytes : (int64_t)rc;
}

int seis_plan_create(const seis_model_desc* models, int32_t n_models, const seis_time_desc* time,
                     int32_t ckpt_every, int32_t caps, seis_plan_t** plan_out) {
  if (!plan_out) return fail(SEIS_EINVAL, "plan_out is NULL");
  *plan_out = nullptr;
  if (time && !time->wavelet) return fail(SEIS_EINVAL, "wavelet is NULL");
  seis_plan* p = new (std::nothrow) seis_plan();
  if (!p) return fail(SEIS_ENOMEM, "host allocation failed");
  int rc = validate_and_layout(models, n_models, time, ckpt_every, caps, p->L);
  if (rc != SEIS_OK) { delete p; return rc; }
  const Layout& L = p->L;
  p->n_models = n_models;
  p->n_waves = (int)L.waves.size();
  p->nt = time->nt;
  p->caps = caps;
  const bool tan = caps & SEIS_CAP_TANGENT, adj = caps & SEIS_CAP_ADJOINT;
  auto bail = [&](int code) { plan_free(p); return code; };
  if ((rc = dev_alloc(&p->d_models, n_models)) || (rc = dev_alloc(&p->d_waves, p->n_waves)) ||
      (rc = dev_alloc(&p->d_params, n_models)) || (rc = dev_alloc(&p->d_wparam, p->n_waves)) ||
      (rc = dev_alloc(&p->d_blk, L.n_blk)) || (rc = dev_alloc(&p->d_fold, adj ? L.fold : 0)) ||
      (rc = dev_alloc(&p->d_part, (int64_t)kRedBlocks * n_models)) ||
      (rc = dev_alloc(&p->d_alpha, L.cells)) || (rc = dev_alloc(&p->d_dalpha, tan ? L.cells : 0)) ||
      (rc = dev_alloc(&p->d_wavelet, p->nt)) || (rc = dev_alloc(&p->d_arena, L.arena)) ||
      (rc = dev_alloc(&p->d_seis, L.seis)) || (rc = dev_alloc(&p->d_dseis, tan ? L.seis : 0)) ||
      (rc = dev_alloc(&p->d_dd, adj ? L.seis : 0)))
    return bail(rc);
  std::vector<float> wl(p->nt);
  for (int i = 0; i < p->nt; ++i) wl[i] = (float)time->wavelet[i];
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMemcpy(p->d_models, L.models.data(), sizeof(ModelDev) * n_models, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_waves, L.waves.data(), sizeof(WaveDev) * p->n_waves, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_blk, L.blk.data(), sizeof(int2) * L.n_blk, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_wavelet, wl.data(), sizeof(float) * p->nt, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(p->d_wparam, 0, sizeof(WaveParam) * p->n_waves);
  if (e == hipSuccess) e = hipMemset(p->d_params, 0, sizeof(ParamDev) * n_models);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    fail(SEIS_EHIP, std::string("plan setup: ") + hipGetErrorString(e));
    return bail(SEIS_EHIP);
  }
  *plan_out = p;
  return SEIS_OK;
}

int seis_plan_destroy(seis_plan_t* plan) {
  if (!plan) return SEIS_OK;
  (void)hipDeviceSynchronize();
  plan_free(plan);
  return SEIS_OK;
}

int64_t seis_plan_device_bytes(const seis_plan_t* plan) { return plan ? plan->L.bytes : 0; }

int seis_plan_layout(const seis_plan_t* plan, int64_t* vec_off, int64_t* seis_off) {
  if (!plan) return fail(SEIS_EINVAL, "plan is NULL");
  if (vec_off) std::memcpy(vec_off, plan->L.vec_off.data(), sizeof(int64_t) * (plan->n_models + 1));
  if (seis_off) std::memcpy(seis_off, plan->L.seis_off.data(), sizeof(int64_t) * (plan->n_models + 1));
  return SEIS_OK;
}

int seis_set_active(seis_plan_t* p, const int32_t* active) {
  if (!p || !active) return fail(SEIS_EINVAL, "NULL argument");
  for (int i = 0; i < p->n_models; ++i) p->L.models[i].active = active[i] ? 1 : 0;
  HIP_TRY(hipMemcpy(p->d_models, p->L.models.data(), sizeof(ModelDev) * p->n_models,
                    hipMemcpyHostToDevice));
  return SEIS_OK;
}

int seis_forward(seis_plan_t* p, const float* vec, float* seis, seis_stream_t stream) {
  if (!p || !vec) return fail(SEIS_EINVAL, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, nullptr));
  TRY(get_graph(p, &p->g_fwd[0][0], [&](hipStream_t c) { enqueue_forward(p, c, false, false); }));
  HIP_TRY(hipGraphLaunch(p->g_fwd[0][0], s));
  return copy_out(p, s, p->d_seis, seis);
}

int seis_tangent(seis_plan_t* p, const float* vec, const float* dvec, float* seis, float* dseis,
                 seis_stream_t stream) {
  if (!p || !vec || !dvec) return fail(SEIS_EINVAL, "NULL argument");
  if (!(p->caps & SEIS_CAP_TANGENT)) return fail(SEIS_ESTATE, "plan lacks SEIS_CAP_TANGENT");
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, dvec));
  TRY(get_graph(p, &p->g_fwd[1][0], [&](hipStream_t c) { enqueue_forward(p, c, true, false); }));
  HIP_TRY(hipGraphLaunch(p->g_fwd[1][0], s));
  TRY(copy_out(p, s, p->d_seis, seis));
  return copy_out(p, s, p->d_dseis, dseis);
}

int seis_gradient(seis_plan_t* p, const float* vec, const float* dd, int32_t on_residual, float* seis,
                  double* misfit, double* grad, seis_stream_t stream) {
  if (!p || !vec || !dd || !grad) return fail(SEIS_EINVAL, "NULL argument");
  if (!(p->caps & SEIS_CAP_ADJOINT)) return fail(SEIS_ESTATE, "plan lacks SEIS_CAP_ADJOINT");
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, nullptr));
  HIP_TRY(hipMemcpyAsync(p->d_dd, dd, sizeof(float) * p->L.seis, hipMemcpyDeviceToDevice, s));
  TRY(get_graph(p, &p->g_fwd[0][1], [&](hipStream_t c) { enqueue_forward(p, c, false, true); }));
  HIP_TRY(hipGraphLaunch(p->g_fwd[0][1], s));
  if (on_residual && misfit)
    launch_misfit(p, s, misfit);
  const int kind = on_residual ? 0 : 1;
  TRY(get_graph(p, &p->g_adj[kind], [&](hipStream_t c) { enqueue_adjoint(p, c, kind); }));
  HIP_TRY(hipGraphLaunch(p->g_adj[kind], s));
  TRY(finish_grad(p, s, vec, grad));
  return copy_out(p, s, p->d_seis, seis);
}

int seis_gn_product(seis_plan_t* p, const float* vec, const float* dvec, float* dseis, double* grad,
                    seis_stream_t stream) {
  if (!p || !vec || !dvec || !grad) return fail(SEIS_EINVAL, "NULL argument");
  if ((p->caps & (SEIS_CAP_TANGENT | SEIS_CAP_ADJOINT)) != (SEIS_CAP_TANGENT | SEIS_CAP_ADJOINT))
    return fail(SEIS_ESTATE, "plan lacks SEIS_CAP_TANGENT | SEIS_CAP_ADJOINT");
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, dvec));
  TRY(get_graph(p, &p->g_fwd[1][1], [&](hipStream_t c) { enqueue_forward(p, c, true, true); }));
  HIP_TRY(hipGraphLaunch(p->g_fwd[1][1], s));
  TRY(get_graph(p, &p->g_adj[2], [&](hipStream_t c) { enqueue_adjoint(p, c, 2); }));
  HIP_TRY(hipGraphLaunch(p->g_adj[2], s));
  TRY(finish_grad(p, s, vec, grad));
  return copy_out(p, s, p->d_dseis, dseis);
}

}  // extern "C"
