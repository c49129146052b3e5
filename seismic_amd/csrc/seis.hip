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
// A saving forward stores q[n] per cell and step.  The adjoint's imaging terms
// and the tangent's source terms dalpha*Lf(p[n]) and dkappa*q[n] are
// pointwise in q[n] and in alpha*Lf(p[n]), which the increment form gives
// back from two consecutive levels: alpha*Lf(p[n]) = q[n+1] - (1-kappa)*q[n]
// (minus the source term at the source cell), evaluated as one fused
// multiply-add in every kernel that needs it, so the tangent and the adjoint
// each run one stencil per step and read 4 bytes of history per cell-update.
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
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "seis.h"

namespace {

// ---------------------------------------------------------------- constants
constexpr float kC2 = 4.0f / 3.0f;    // core/seis_forward2.py:80
constexpr float kC3 = -1.0f / 12.0f;  // core/seis_forward2.py:81

constexpr int NTHR = 256;        // 4 wavefronts per workgroup
constexpr int NXCD = 8;          // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr int kKzPad = 64;       // rows of periodic apron on each side of a model's 1 - kappa row table
// Work lists: one per kernel kind, plus an internal one with a fixed tiled
// window that runs the launches of streamed passes shorter than their k.
constexpr int kTail = SEIS_KIND_COUNT, kLists = SEIS_KIND_COUNT + 1;

// Per-wavefield state regions in the float arena.
enum Region {
  R_HIST = 0,  // q[n] per cell: full history n = 0..nt+1 (nt+2 slots, slot 0 the
               // zero initial state), or a recompute segment n = lo..lo+Kc+1
               // (Kc+2 slots; slot 0 unused); one float per cell
  R_PP,        // forward p ping-pong (2 slots)
  R_Q,         // q = p[n] - p[n-1] ping-pong (2 slots)
  R_DP,        // tangent dp ping-pong (2 slots)
  R_DQ,        // tangent dq
  R_LAM,       // adjoint lambda ping-pong (2 slots)
  R_MU,        // adjoint mu = lam[m] - lam[m+1]
  R_GA,        // imaging accumulator for alpha
  R_GK,        // imaging accumulator for kappa
  R_CKP,       // checkpoints of p[jK] (J slots)
  R_CKQ,       // checkpoints of q[jK] (J slots)
  R_COUNT
};

struct ModelDev {
  int nz, nx, NZ, NX;
  int nbc, n_shots, first_wave, active;
  int64_t vec_off;    // into the concatenated parameter vectors
  int64_t cell_off;   // into per-model padded coefficient fields
  int64_t kz_off;     // into the alpha array: 1 - kappa of the padded rows (kKzPad apron each side)
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

// Raw buffer resource over device memory starting at p.  A load or store
// through it takes the thread's byte offset in a vector register and a
// wave-uniform byte offset in a scalar register (buffer_load_dword ... offen),
// so a row access costs no vector address arithmetic and row offsets advance
// as 32-bit scalar adds.  Offsets must stay below 4 GiB: the plan refuses a
// streamed config when one wavefield's state region is larger.  Word 3 is the
// gfx9-family raw 32-bit buffer format.  Every resource here is built from
// a wave-uniform base; the base is read from the first lane so that the
// compiler never treats it as divergent (a descriptor in vector registers
// makes it loop over the distinct values, several instructions per access).
typedef __amdgpu_buffer_rsrc_t brsrc_t;
__device__ __forceinline__ brsrc_t brsrc(const void* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  void* const u = (void*)(((unsigned long long)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(u, 0, -1, 0x00020000);
}
__device__ __forceinline__ float bld(brsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void bst(brsrc_t r, unsigned voff, unsigned soff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
}
// Byte offset of the next row (one padded row on, wrapping at the slot end
// `end`, a slot being `slot` bytes): a compare and a subtract on the scalar unit.
__device__ __forceinline__ unsigned next_row(unsigned o, unsigned rowb, unsigned end, unsigned slot) {
  o += rowb;
  return o >= end ? o - slot : o;
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

// Lf(p) at window row r (LDS row r + 2), column cc (LDS column cc + 2):
// lapg(p) + 2*c1*p written as neighbour-minus-centre differences (exact
// cancellation of the centre weight, 4*c2 + 4*c3 = -2*c1).  The vertical
// neighbours come from `col` (col[t] = window row r0 - 2 + t of this column,
// held in registers for the thread's own rows), the horizontal ones from the
// LDS window T.  Fixed evaluation order: every kernel that advances a
// wavefield shares it, so results do not depend on tiling.
template <int W>
__device__ __forceinline__ float lap_col(const float (*T)[W], const float* col, int j, int r, int cc) {
  const float pc = col[j + 2];
  float t1 = (T[r + 2][cc + 3] - pc) + (T[r + 2][cc + 1] - pc);
  t1 = t1 + (col[j + 3] - pc);
  t1 = t1 + (col[j + 1] - pc);
  float t2 = (T[r + 2][cc + 4] - pc) + (T[r + 2][cc] - pc);
  t2 = t2 + (col[j + 4] - pc);
  t2 = t2 + (col[j] - pc);
  return fmaf(kC2, t1, kC3 * t2);
}

// Same stencil with every neighbour read from the LDS window and the centre
// value pc (the thread's own cell) from a register.
template <int W>
__device__ __forceinline__ float lap_lds(const float (*T)[W], float pc, int r, int cc) {
  float t1 = (T[r + 2][cc + 3] - pc) + (T[r + 2][cc + 1] - pc);
  t1 = t1 + (T[r + 3][cc + 2] - pc);
  t1 = t1 + (T[r + 1][cc + 2] - pc);
  float t2 = (T[r + 2][cc + 4] - pc) + (T[r + 2][cc] - pc);
  t2 = t2 + (T[r + 4][cc + 2] - pc);
  t2 = t2 + (T[r][cc + 2] - pc);
  return fmaf(kC2, t1, kC3 * t2);
}

// LDS floats addressed by their 32-bit LDS address.  A row base passed
// through an empty asm is opaque to the compiler, so it stays in one
// register and the four horizontal neighbours of a row load as two ds_read2
// with immediate offsets, with no address arithmetic per step.
typedef __attribute__((address_space(3))) float lds_float;
__device__ __forceinline__ unsigned lds_row(const float* p) {
  unsigned a = (unsigned)(size_t)(const lds_float*)p;
  asm("" : "+v"(a));
  return a;
}
__device__ __forceinline__ lds_float* lds_at(unsigned a) { return (lds_float*)(size_t)a; }

// lap_col with the horizontal neighbours read through rp = &T[r + 2][cc]
// (same operations in the same order).
__device__ __forceinline__ float lap_row(const lds_float* rp, const float* col, int j) {
  const float pc = col[j + 2];
  float t1 = (rp[3] - pc) + (rp[1] - pc);
  t1 = t1 + (col[j + 3] - pc);
  t1 = t1 + (col[j + 1] - pc);
  float t2 = (rp[4] - pc) + (rp[0] - pc);
  t2 = t2 + (col[j + 4] - pc);
  t2 = t2 + (col[j] - pc);
  return fmaf(kC2, t1, kC3 * t2);
}

// Window row of a thread's j-th owned row.  STRIP: each wavefront owns RPW
// consecutive rows (vertical neighbours of inner rows stay in registers);
// otherwise rows are dealt round-robin over the NW wavefronts, which keeps
// the shrinking valid band and the interior evenly spread over them.
template <bool STRIP, int NW, int RPW>
__device__ __forceinline__ int own_row(int wv, int j) {
  return STRIP ? wv * RPW + j : wv + NW * j;
}

// Vertical strip of column cc for a thread owning window rows r0 .. r0+RPW-1:
// two halo rows above and below from the LDS window, the own rows from `own`
// (stride CPR, starting at chunk c).
template <int W, int RPW, int CPR>
__device__ __forceinline__ void load_col(const float (*T)[W], const float* own, int c, int r0, int cc,
                                         float* col) {
  col[0] = T[r0][cc + 2];
  col[1] = T[r0 + 1][cc + 2];
#pragma unroll
  for (int j = 0; j < RPW; ++j) col[j + 2] = own[j * CPR + c];
  col[RPW + 2] = T[r0 + RPW + 2][cc + 2];
  col[RPW + 3] = T[r0 + RPW + 3][cc + 2];
}


// Tile of the compacted work list: one (wavefield, interior origin) pair.
// Interior origins are unwrapped padded-grid coordinates; the window staged
// in LDS extends 2K cells beyond the interior on every side and wraps
// periodically (core/seis_forward2.py:395-403).
struct TileEnt {
  int w, z0, x0, pad;
};

// ---------------------------------------------------------------- forward (+ tangent)
struct FwdArgs {
  const TileEnt* tiles;
  const int* n_tiles;
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
  int nt;                // time samples
  int n0, K;             // advance state n0 -> n0+K; step n produces p[n+1] with s[n-1]
  int Kt;                // steps the tile list was built for (halo 2*Kt); K <= Kt
  int p_in_slot, p_out_slot;  // R_PP slots holding p[n0] and receiving p[n0+K]
  int hist_slot0;        // >= 0: q[n0+s+1] (the output of step n0+s) -> R_HIST slot hist_slot0 + s
  int q_in_slot, q_out_slot;
  int dp_in_slot, dp_out_slot, dq_in_slot, dq_out_slot;
  int ck_slot;           // >= 0: also store the final (p, q) as checkpoint ck_slot
  int write_seis;
};

constexpr int KMAX = 8;  // most time steps per launch

// Receiver column index of window column cc (-1 if none).
__device__ __forceinline__ int rcv_index(int gx, const ModelDev& m) {
  const int d = gx - m.rcv_x0;
  if (d < 0) return -1;
  const int r = d / m.rcv_dx;
  return (r * m.rcv_dx == d && r < m.n_rcv) ? r : -1;
}

// K forward time steps (update_p :428-462 with lapg :381-421 and the source
// injection fused), and with TAN the Born recursion (update_p_diff :470-509),
// for every tile of the work list.  Each thread owns RPW window rows (see
// own_row) and one column per 64-column chunk, and keeps p, q, alpha and
// (1-kappa) of its cells in registers; with STRIP only the horizontal
// neighbours and two halo rows above and below the strip are read from LDS.
// The LDS window is double-buffered: step s reads buffer s&1 and writes the
// other, one barrier per step.  Every owned cell is advanced each step (cells
// outside the valid band, which shrinks by 2 per step, compute values no
// committed cell depends on), so the row loop has no branches and the LDS
// reads of different rows can be in flight together; seismogram samples are
// stored after it.  TAN (two stencils per cell) skips rows outside the band
// instead, which keeps its register count down.  No global load is issued inside the step loop: the
// launch's wavelet samples are staged in LDS with the window.  With HIST (a
// saving forward) the kernel also stores, per interior cell and step, the q
// it produced, which the tangent and adjoint recursions consume.  With TAN the
// source term dalpha*Lf(p[n]) is (dalpha/alpha) times alpha*Lf(p[n]) recovered
// from q[n] and q[n+1], the expression the Born kernel evaluates on the saved
// history, so the two give identical bits.
template <int RX, int RZ, int NT, bool TAN, bool STRIP, bool HIST>
__global__ __launch_bounds__(NT) void k_fwd(FwdArgs a) {
  constexpr int NW = NT / 64, CPR = RX / 64, RPW = RZ / NW, NC = CPR * RPW;
  static_assert(RZ % NW == 0, "rows must split evenly over the wavefronts");
  static_assert(NC <= 32 && RPW <= 32, "tile too large for the bit masks");
  // Window cell (r, c) lives at [r + 2][c + 2]: a 2-cell apron keeps every
  // stencil inside the array.  The apron and cells outside the valid band
  // hold stale values that no committed cell depends on.
  constexpr int PW = RX + 4, PH = RZ + 4;
  // the two p windows interleaved by rows (row r of window b at PP[r][b]):
  // a register holding the address of PP[r][0][lane] reaches the
  // neighbours of both windows with ds_read2 offsets
  __shared__ float PP[PH][2][PW];
  constexpr int PS = 2 * PW;   // row stride of either window
  __shared__ float D0[TAN ? PH : 1][TAN ? PW : 1], D1[TAN ? PH : 1][TAN ? PW : 1];
  __shared__ float SW[KMAX];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform rows
  const int r0 = wv * RPW;
  const int ntiles = *a.n_tiles;
  const int H = 2 * a.Kt;  // window halo of the tile list; a.K <= a.Kt steps run
  for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
    const TileEnt te = a.tiles[e];
    if (te.w < 0) continue;
    const WaveDev& w = a.waves[te.w];
    const ModelDev& m = a.models[w.model];
    const int NZ = w.NZ, NX = w.NX;
    const int64_t Pn = (int64_t)NZ * NX;
    const float* pin = a.arena + w.off[R_PP] + a.p_in_slot * Pn;
    const float* qin = a.arena + w.off[R_Q] + a.q_in_slot * Pn;
    float* hist = HIST ? a.arena + w.off[R_HIST] + a.hist_slot0 * Pn : nullptr;
    const float* alpha = a.alpha + m.cell_off;
    const float kscale = a.params[w.model].kscale;
    const float amp = a.wparam[te.w].amp;
    const int oz = te.z0 - H, ox = te.x0 - H;
    const int gint0 = oz * NX + ox;  // unwrapped index of window cell (0, 0)
    float pv[NC], q[NC], al[NC], k1[NC];
    float dpv[TAN ? NC : 1], dq[TAN ? NC : 1], dal[TAN ? NC : 1], dk[TAN ? NC : 1];
    float dkscale = 0.f, amp_d = 0.f;
    const float* dpin = nullptr;
    const float* dqin = nullptr;
    if constexpr (TAN) {
      dpin = a.arena + w.off[R_DP] + a.dp_in_slot * Pn;
      dqin = a.arena + w.off[R_DQ] + a.dq_in_slot * Pn;
      dkscale = a.params[w.model].dkscale;
      amp_d = a.wparam[te.w].amp_d;
    }
    if (threadIdx.x < KMAX) {
      const int it = a.n0 + (int)threadIdx.x - 1;
      SW[threadIdx.x] = ((int)threadIdx.x < a.K && it >= 0 && it < a.nt) ? a.wavelet[it] : 0.f;
    }
    // row bits (srcrow, introw, rcvmask) are wave-uniform; cell bits per lane
    unsigned srcmask = 0, rcvmask = 0, intmask = 0, srcrow = 0, introw = 0;
    const int srow = w.src_cell / NX;
    // per own column, once per tile: wrapped x, receiver index, the x damping
    // profile (used in the left/right strips), interior flag
    int gxc[CPR], rcol[CPR];
    bool isxc[CPR], cin[CPR];
    float jjxc[CPR];
#pragma unroll
    for (int c = 0; c < CPR; ++c) {
      const int cc = lane + 64 * c;
      const int gx = wrapi(ox + cc, NX);
      gxc[c] = gx;
      rcol[c] = rcv_index(gx, m);
      isxc[c] = gx < m.nbc || gx >= m.nx + m.nbc;
      const int jx = damp_j(m.nbc, gx, m.nz, m.nx, m.nbc);
      jjxc[c] = (float)(jx * jx);
      cin[c] = cc >= H && cc < RX - H && ox + cc < NX;
    }
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
      const int r = own_row<STRIP, NW, RPW>(wv, j);
      const int gz = wrapi(oz + r, NZ);
      if (gz == m.rcv_row) rcvmask |= 1u << j;
      if (gz == srow) srcrow |= 1u << j;
      const bool rin = r >= H && r < RZ - H && oz + r < NZ;
      if (rin) introw |= 1u << j;
      const int jz = damp_j(gz, m.nbc, m.nz, m.nx, m.nbc);   // the z profile (outside the strips)
      const float jjz = (float)(jz * jz);
      const int grow = gz * NX;
#pragma unroll
      for (int c = 0; c < CPR; ++c) {
        const int i = j * CPR + c;
        const int cc = lane + 64 * c;
        const int g = grow + gxc[c];
        pv[i] = pin[g];
        PP[r + 2][0][cc + 2] = pv[i];
        q[i] = qin[g];
        al[i] = alpha[g];
        const float jj = isxc[c] ? jjxc[c] : jjz;
        k1[i] = 1.0f - kscale * jj;
        if (g == w.src_cell) srcmask |= 1u << i;
        if (rin && cin[c]) intmask |= 1u << i;
        if constexpr (TAN) {
          dpv[i] = dpin[g];
          D0[r + 2][cc + 2] = dpv[i];
          dq[i] = dqin[g];
          dal[i] = a.dalpha[m.cell_off + g];
          dk[i] = dkscale * jj;
        }
      }
    }
    // LDS addresses of the own rows, &PP[r + 2][0][lane] (strips without
    // the tangent, whose register budget is tighter)
    constexpr bool RB = STRIP && !TAN;
    unsigned rb[RB ? RPW : 1];
    if constexpr (RB) {
#pragma unroll
      for (int j = 0; j < RPW; ++j) rb[j] = lds_row(&PP[r0 + j + 2][0][lane]);
    }
    __syncthreads();
    // Step s reads the window buffer s&1 and writes the other; the buffer
    // is a compile-time choice so that reads and writes never alias.
    auto step = [&](auto cb, const int s) {
      constexpr int CUR = decltype(cb)::value;
      float (*const Prd)[PS] = reinterpret_cast<float (*)[PS]>(&PP[0][CUR][0]);
      float (*const Pwr)[PS] = reinterpret_cast<float (*)[PS]>(&PP[0][1 - CUR][0]);
      float (*const Drd)[TAN ? PW : 1] = CUR ? D1 : D0;
      float (*const Dwr)[TAN ? PW : 1] = CUR ? D0 : D1;
      const int it = a.n0 + s - 1;
      const bool live = it >= 0 && it < a.nt;
      const float sv = SW[s];
      const float src = amp * sv;
      const float srcd = amp_d * sv;
      const int lo = 2 * (s + 1);
      // history slot of this step, through a buffer resource (see k_adj)
      const brsrc_t VH = brsrc(HIST ? hist + s * Pn : a.arena);
      if (!STRIP || (r0 + RPW > lo && r0 < RZ - lo)) {  // the strip meets the valid band (wave-uniform)
#pragma unroll
        for (int c = 0; c < CPR; ++c) {
          const int cc = lane + 64 * c;
          float col[STRIP ? RPW + 4 : 1];
          if constexpr (STRIP) load_col<PS, RPW, CPR>(Prd, pv, c, r0, cc, col);
          float dcol[TAN && STRIP ? RPW + 4 : 1];
          if constexpr (TAN && STRIP) load_col<PW, RPW, CPR>(Drd, dpv, c, r0, cc, dcol);
#pragma unroll
          for (int j = 0; j < RPW; ++j) {
            const int r = own_row<STRIP, NW, RPW>(wv, j);
            if constexpr (TAN) {  // two stencils: skip rows outside the band to save registers
              if (r < lo || r >= RZ - lo) continue;
            }
            const int i = j * CPR + c;
            float L;
            if constexpr (RB) L = lap_row(lds_at(rb[j]) + CUR * PW + 64 * c, col, j);
            else if constexpr (STRIP) L = lap_col<PS>(Prd, col, j, r, cc);
            else L = lap_lds<PS>(Prd, pv[i], r, cc);
            float qn = fmaf(k1[i], q[i], al[i] * L);
            if ((srcrow >> j) & 1u) qn = ((srcmask >> i) & 1u) ? qn + src : qn;
            if constexpr (HIST) {
              if ((introw >> j) & 1u) {
                if ((intmask >> i) & 1u) bst(VH, (unsigned)(gint0 + r * NX + cc) * 4u, 0u, qn);
              }
            }
            if constexpr (TAN) {
              // dq' = (1-kappa) dq + alpha Lf(dp) + dalpha Lf(p) - dkappa q, with
              // dalpha Lf(p) = (dalpha/alpha) (alpha Lf(p)) and alpha Lf(p)
              // recovered from q and q' as the Born kernel recovers it
              float Ld;
              if constexpr (STRIP) Ld = lap_col<PW>(Drd, dcol, j, r, cc);
              else Ld = lap_lds<PW>(Drd, dpv[i], r, cc);
              float aL = fmaf(-k1[i], q[i], qn);
              if ((srcrow >> j) & 1u) aL = ((srcmask >> i) & 1u) ? aL - src : aL;
              float dqn = fmaf(k1[i], dq[i], al[i] * Ld);
              dqn = fmaf(dal[i], aL, dqn);
              dqn = fmaf(-dk[i], q[i], dqn);
              if ((srcrow >> j) & 1u) dqn = ((srcmask >> i) & 1u) ? dqn + srcd : dqn;
              const float dp1 = dpv[i] + dqn;
              Dwr[r + 2][cc + 2] = dp1;
              dpv[i] = dp1;
              dq[i] = dqn;
            }
            const float p1 = pv[i] + qn;
            if constexpr (RB) lds_at(rb[j])[(1 - CUR) * PW + 64 * c + 2] = p1;
            else Pwr[r + 2][cc + 2] = p1;
            pv[i] = p1;
            q[i] = qn;
          }
        }
      }
      if (a.write_seis && live && rcvmask) {  // receiver samples of the interior
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
          if (!((rcvmask >> j) & 1u)) continue;
#pragma unroll
          for (int c = 0; c < CPR; ++c) {
            const int i = j * CPR + c;
            if (!((intmask >> i) & 1u) || rcol[c] < 0) continue;
            const int64_t si = w.seis_off + (int64_t)it * m.n_rcv + rcol[c];
            a.seis[si] = pv[i];
            if constexpr (TAN) a.dseis[si] = dpv[i];
          }
        }
      }
      __syncthreads();
    };
    for (int s = 0; s < a.K; s += 2) {
      step(std::integral_constant<int, 0>(), s);
      if (s + 1 < a.K) step(std::integral_constant<int, 1>(), s + 1);
    }
    // state (p, q[, dp, dq]) at n0+K on the interior
    float* qout = a.arena + w.off[R_Q] + a.q_out_slot * Pn;
    float* pout = a.arena + w.off[R_PP] + a.p_out_slot * Pn;
    float* ckp = a.ck_slot >= 0 ? a.arena + w.off[R_CKP] + a.ck_slot * Pn : nullptr;
    float* ckq = a.ck_slot >= 0 ? a.arena + w.off[R_CKQ] + a.ck_slot * Pn : nullptr;
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
#pragma unroll
      for (int c = 0; c < CPR; ++c) {
        const int i = j * CPR + c;
        if (!((intmask >> i) & 1u)) continue;
        const int g = gint0 + own_row<STRIP, NW, RPW>(wv, j) * NX + lane + 64 * c;
        qout[g] = q[i];
        pout[g] = pv[i];
        if (ckp) {
          ckp[g] = pv[i];
          ckq[g] = q[i];
        }
        if constexpr (TAN) {
          a.arena[w.off[R_DP] + a.dp_out_slot * Pn + g] = dpv[i];
          a.arena[w.off[R_DQ] + a.dq_out_slot * Pn + g] = dq[i];
        }
      }
    }
  }
}

// ---------------------------------------------------------------- forward, streamed down the rows
constexpr int SUN = 12;  // pipeline steps per unrolled block: a multiple of every ring length
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {  // f(integral_constant<I>) .. f(<N-1>)
  if constexpr (I < N) {
    f(std::integral_constant<int, I>());
    static_for<I + 1, N>(f);
  }
}

constexpr int SPD = 3;   // pipeline steps between a row's load and its intake


// 1 - kappa of padded row gz in a column outside the left/right damping
// strips (the same expression as the per-cell value of k_fwd).
__device__ __forceinline__ float kappa_row(int gz, const ModelDev& m, float kscale) {
  const int j = damp_j(gz, m.nbc, m.nz, m.nx, m.nbc);
  return 1.0f - kscale * (float)(j * j);
}

// KL forward time steps (update_p :428-462, lapg :381-421, source injection
// fused) over a column strip of NT padded columns, one per thread, marched
// down the rows of a segment in a software pipeline.  Pipeline step t takes
// in window row t of p[n0], q[n0], alpha and the row's 1 - kappa (requested
// SPD steps earlier) as level 0, and level l = 1..KL computes window row
// t - 2l of p[n0+l], q[n0+l] from level l-1; levels run in increasing order,
// so the lower neighbour two rows down is the value level l-1 produced in
// the same step.  Each thread keeps the last rows of its own column of every
// level in registers (centre, vertical neighbours, q) and reads the
// horizontal neighbours from a three-row LDS ring per level, with one
// barrier per pipeline step; the (alpha, 1 - kappa) pairs of its column's
// rows in flight sit in a private LDS ring.  1 - kappa outside the
// left/right damping strips, one value per padded row, comes from the
// model's row table.  The window is the segment's rows (TileEnt.pad of them)
// plus 2*KL rows above and below and 2*KL columns either side, wrapped
// periodically like k_fwd's tiles; cells outside the valid band compute
// values no committed cell depends on.  Blocks of SUN pipeline steps in
// which some level meets the source row or the receiver row run a separate
// copy of the block, so the others carry no per-cell source or receiver
// work; receiver samples are kept in registers and stored after the march.
// With TAN every level also advances the Born recursion (update_p_diff
// :470-509) on (dp, dq), from its own LDS ring and registers, with the
// source term dalpha*Lf(p) taken as (dalpha/alpha) times alpha*Lf(p)
// recovered from the level's q and q' as k_fwd does; with HIST each level
// stores its q of the committed rows into the history (saving forward).
// Same stencil order and fused multiply-adds as k_fwd, so wavefields,
// seismograms and histories are bit-identical to it.  Launches always
// advance KL steps (shorter ones run k_fwd).
template <int NT, int KL, bool TAN, bool HIST>
__global__ __launch_bounds__(NT) void k_fws(FwdArgs a) {
  // pipeline steps per unrolled block: 12, or 18 when the rows in flight
  // (2*KL + SPD) need a longer (alpha, 1 - kappa) ring (5 levels)
  constexpr int SUNF = 2 * KL + SPD < SUN ? SUN : 18;
  static_assert(2 * KL + SPD < SUNF && SUNF % 6 == 0 && SUNF % SPD == 0, "pipeline rings too short");
  constexpr int H = 2 * KL;
  // rows of the level rings, column-major: window column x (two apron
  // columns either side) holds its KL*3 ring slots at [(x + 2) * RS], so a
  // thread reaches every slot of its own and its neighbours' columns from one
  // base address with immediate offsets; RS (at least the 3*KL slots of
  // levels 0..KL-1) is odd, so the 64 lanes of a wavefront (stride RS) fall
  // in distinct banks (an even stride would put them in a few banks)
  constexpr int RS = (3 * KL) | 1;
  __shared__ float RG[(NT + 4) * RS];
  __shared__ float2 AK[SUNF][NT];  // (alpha, 1 - kappa) of window row u at [u % SUNF], own column
  __shared__ float DG[TAN ? (NT + 4) * RS : 1];      // dp rings
  __shared__ float2 DK[TAN ? SUNF : 1][TAN ? NT : 1];  // (dalpha/alpha, dkappa), like AK
  const unsigned c = threadIdx.x;
  const int ntiles = *a.n_tiles;
  const int n0 = a.n0, nt = a.nt;
  for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
    const TileEnt te = a.tiles[e];
    if (te.w < 0) continue;
    const WaveDev& w = a.waves[te.w];
    const ModelDev& m = a.models[w.model];
    const int NZ = w.NZ, NX = w.NX, nz = m.nz, nx = m.nx, nbc = m.nbc;
    const unsigned rowb = (unsigned)NX * 4u, P4 = (unsigned)NZ * rowb;   // bytes of a row, of a slot
    // the wavefield's state through one buffer resource, alpha (and dalpha)
    // through their own; byte offsets of the slots this launch reads and writes
    const brsrc_t VW = brsrc(a.arena + w.off[R_HIST]);
    const brsrc_t VA = brsrc(a.alpha + m.cell_off);
    const brsrc_t VD = brsrc(TAN ? a.dalpha + m.cell_off : a.alpha);
    auto reg_off = [&](int r, int slot) {
      return (unsigned)(w.off[r] - w.off[R_HIST]) * 4u + (unsigned)slot * P4;
    };
    const unsigned o_pi = reg_off(R_PP, a.p_in_slot), o_qi = reg_off(R_Q, a.q_in_slot);
    const unsigned o_po = reg_off(R_PP, a.p_out_slot), o_qo = reg_off(R_Q, a.q_out_slot);
    const bool ck = a.ck_slot >= 0;
    const unsigned o_cp = ck ? reg_off(R_CKP, a.ck_slot) : 0u, o_cq = ck ? reg_off(R_CKQ, a.ck_slot) : 0u;
    const unsigned o_dpi = TAN ? reg_off(R_DP, a.dp_in_slot) : 0u, o_dqi = TAN ? reg_off(R_DQ, a.dq_in_slot) : 0u;
    const unsigned o_dpo = TAN ? reg_off(R_DP, a.dp_out_slot) : 0u, o_dqo = TAN ? reg_off(R_DQ, a.dq_out_slot) : 0u;
    const unsigned o_h = HIST ? reg_off(R_HIST, a.hist_slot0) : 0u;   // q[n0+1]; q[n0+l] one slot per level on
    float* seis = a.write_seis ? a.seis + w.seis_off : nullptr;
    float* dseis = TAN && a.write_seis ? a.dseis + w.seis_off : nullptr;
    const float kscale = a.params[w.model].kscale;
    const float dkscale = TAN ? a.params[w.model].dkscale : 0.f;
    const float amp = a.wparam[te.w].amp;
    const float amp_d = TAN ? a.wparam[te.w].amp_d : 0.f;
    const int n_rcv = m.n_rcv, rcv_row = m.rcv_row;
    const float* kzt = a.alpha + m.kz_off + kKzPad;   // kzt[z]: padded row z, -kKzPad <= z < NZ + kKzPad
    const int IZ = te.pad;              // rows of the segment (all inside the grid)
    const int U = IZ + 2 * H;           // window rows
    const int oz = te.z0 - H, ox = te.x0 - H;
    const unsigned gx = wrapi(ox + (int)c, NX);
    const unsigned vx = gx * 4u;                  // the thread's column, bytes into a row
    const bool colint = (int)c >= H && (int)c < NT - H && ox + (int)c < NX;
    const bool isx = (int)gx < nbc || (int)gx >= nx + nbc;
    const int jx = damp_j(0, gx, nz, nx, nbc);
    const float jjx = (float)(jx * jx);
    const float kx1 = 1.0f - kscale * jjx;
    const int rc = rcv_index(gx, m);
    const int srow = w.src_cell / NX;
    const bool srcc = (int)gx == w.src_cell - srow * NX;
    // window rows holding the source row (twice when the window wraps), and
    // the segment row holding the receiver row (none: far out of range)
    const int su0 = wrapi(srow - oz, NZ), su1 = su0 + NZ;
    const int ruz = rcv_row - te.z0;
    const int rui = ruz >= 0 && ruz < IZ ? ruz + H : -(1 << 20);
    float srcv[KL], srcd[TAN ? KL : 1], rec[KL], drec[TAN ? KL : 1];
#pragma unroll
    for (int l = 1; l <= KL; ++l) {
      const int it = n0 + l - 2;
      const float wl = (it >= 0 && it < nt) ? a.wavelet[it] : 0.f;
      srcv[l - 1] = amp * wl;
      rec[l - 1] = 0.f;
      if constexpr (TAN) {
        srcd[l - 1] = amp_d * wl;
        drec[l - 1] = 0.f;
      }
    }
    float pr[KL][6], qr[KL][3], pfp[SPD], pfq[SPD], pfa[SPD], pfk[SPD];
    float dpr[TAN ? KL : 1][6], dqr[TAN ? KL : 1][3], pfdp[TAN ? SPD : 1], pfdq[TAN ? SPD : 1], pfd[TAN ? SPD : 1];
#pragma unroll
    for (int l = 0; l < KL; ++l) {
#pragma unroll
      for (int s = 0; s < 6; ++s) pr[l][s] = 0.f;
#pragma unroll
      for (int s = 0; s < 3; ++s) qr[l][s] = 0.f;
      if constexpr (TAN) {
#pragma unroll
        for (int s = 0; s < 6; ++s) dpr[l][s] = 0.f;
#pragma unroll
        for (int s = 0; s < 3; ++s) dqr[l][s] = 0.f;
      }
    }
#pragma unroll
    for (int d = 0; d < SPD; ++d) {
      const unsigned ro = (unsigned)wrapi(oz + d, NZ) * rowb;
      pfp[d] = bld(VW, vx, o_pi + ro);
      pfq[d] = bld(VW, vx, o_qi + ro);
      pfa[d] = bld(VA, vx, ro);
      pfk[d] = kzt[oz + d];
      if constexpr (TAN) {
        pfdp[d] = bld(VW, vx, o_dpi + ro);
        pfdq[d] = bld(VW, vx, o_dqi + ro);
        pfd[d] = bld(VD, vx, ro);
      }
    }
    unsigned rq = (unsigned)wrapi(oz + SPD, NZ) * rowb;   // row requested next
    // (oz + u) * rowb of level KL's row u = t - 2 KL (used on committed rows only)
    unsigned so = (unsigned)(te.z0 - 2 * H) * rowb;
    int gzc = wrapi(oz, NZ);            // padded row of window row t (dkappa of the row)
    __syncthreads();
    // one pipeline step; SPECIAL: some level meets the source or receiver row
    auto body = [&](auto ic, auto sc, const int t) {
      constexpr int i = decltype(ic)::value;
      constexpr bool SPECIAL = decltype(sc)::value;
      {  // level 0: intake of window row t
        const float p0 = pfp[i % SPD];
        RG[(c + 2) * RS + i % 3] = p0;
        pr[0][i % 6] = p0;
        qr[0][i % 3] = pfq[i % SPD];
        AK[i][c] = make_float2(pfa[i % SPD], isx ? kx1 : pfk[i % SPD]);
        if constexpr (TAN) {
          const float d0 = pfdp[i % SPD];
          DG[(c + 2) * RS + i % 3] = d0;
          dpr[0][i % 6] = d0;
          dqr[0][i % 3] = pfdq[i % SPD];
          const int jz = damp_j(gzc, nbc, nz, nx, nbc);
          DK[i][c] = make_float2(pfd[i % SPD], dkscale * (isx ? jjx : (float)(jz * jz)));
        }
      }
      static_for<1, KL + 1>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const int u = t - 2 * l;                      // window row of level l
        constexpr int B = 4 * SUNF;                    // keeps the ring indices non-negative
        constexpr int s6 = (i - 2 * l + B) % 6, s3 = (i - 2 * l + B) % 3, sr = (i - 2 * l + B) % SUNF;
        const float pc = pr[l - 1][s6];
        const float* T = RG + c * RS + (l - 1) * 3 + s3;   // T[k * RS]: window column c + k - 2
        float t1 = (T[3 * RS] - pc) + (T[RS] - pc);
        t1 = t1 + (pr[l - 1][(i - 2 * l + 1 + B) % 6] - pc);
        t1 = t1 + (pr[l - 1][(i - 2 * l - 1 + B) % 6] - pc);
        float t2 = (T[4 * RS] - pc) + (T[0] - pc);
        t2 = t2 + (pr[l - 1][(i - 2 * l + 2 + B) % 6] - pc);
        t2 = t2 + (pr[l - 1][(i - 2 * l - 2 + B) % 6] - pc);
        const float L = fmaf(kC2, t1, kC3 * t2);
        const float2 ak = AK[sr][c];
        const float qo = qr[l - 1][s3];
        float qn = fmaf(ak.y, qo, ak.x * L);
        if constexpr (SPECIAL) {
          if (u == su0 || u == su1) qn = srcc ? qn + srcv[l - 1] : qn;
        }
        const float p1 = pc + qn;
        if constexpr (SPECIAL) {
          if (u == rui) rec[l - 1] = p1;
        }
        float dqn = 0.f, dp1 = 0.f;
        if constexpr (TAN) {
          // dq' = (1-kappa) dq + alpha Lf(dp) + dalpha Lf(p) - dkappa q, with
          // dalpha Lf(p) = (dalpha/alpha) (alpha Lf(p)) and alpha Lf(p)
          // recovered from q and q' (k_fwd, k_bns)
          const float dc = dpr[l - 1][s6];
          const float* D = DG + c * RS + (l - 1) * 3 + s3;
          float d1 = (D[3 * RS] - dc) + (D[RS] - dc);
          d1 = d1 + (dpr[l - 1][(i - 2 * l + 1 + B) % 6] - dc);
          d1 = d1 + (dpr[l - 1][(i - 2 * l - 1 + B) % 6] - dc);
          float d2 = (D[4 * RS] - dc) + (D[0] - dc);
          d2 = d2 + (dpr[l - 1][(i - 2 * l + 2 + B) % 6] - dc);
          d2 = d2 + (dpr[l - 1][(i - 2 * l - 2 + B) % 6] - dc);
          const float Ld = fmaf(kC2, d1, kC3 * d2);
          const float2 dk = DK[sr][c];
          float aL = fmaf(-ak.y, qo, qn);
          if constexpr (SPECIAL) {
            if (u == su0 || u == su1) aL = srcc ? aL - srcv[l - 1] : aL;
          }
          dqn = fmaf(ak.y, dqr[l - 1][s3], ak.x * Ld);
          dqn = fmaf(dk.x, aL, dqn);
          dqn = fmaf(-dk.y, qo, dqn);
          if constexpr (SPECIAL) {
            if (u == su0 || u == su1) dqn = srcc ? dqn + srcd[l - 1] : dqn;
          }
          dp1 = dc + dqn;
          if constexpr (SPECIAL) {
            if (u == rui) drec[l - 1] = dp1;
          }
        }
        if constexpr (HIST) {   // q[n0+l] of committed rows into history slot hist_slot0 + l - 1
          if (u >= H && u < H + IZ && colint)
            bst(VW, vx, o_h + (unsigned)(l - 1) * P4 + so + (unsigned)(2 * (KL - l)) * rowb, qn);
        }
        if constexpr (l < KL) {
          RG[(c + 2) * RS + l * 3 + s3] = p1;
          pr[l][s6] = p1;
          qr[l][s3] = qn;
          if constexpr (TAN) {
            DG[(c + 2) * RS + l * 3 + s3] = dp1;
            dpr[l][s6] = dp1;
            dqr[l][s3] = dqn;
          }
        } else if (u >= H && u < H + IZ) {  // committed row of p[n0+KL], q[n0+KL] (and dp, dq)
          if (colint) {
            bst(VW, vx, o_po + so, p1);
            bst(VW, vx, o_qo + so, qn);
            if (ck) {
              bst(VW, vx, o_cp + so, p1);
              bst(VW, vx, o_cq + so, qn);
            }
            if constexpr (TAN) {
              bst(VW, vx, o_dpo + so, dp1);
              bst(VW, vx, o_dqo + so, dqn);
            }
          }
        }
      });
      so += rowb;
      if constexpr (TAN) gzc = gzc + 1 == NZ ? 0 : gzc + 1;
      {  // request of window row t + SPD
        pfp[i % SPD] = bld(VW, vx, o_pi + rq);
        pfq[i % SPD] = bld(VW, vx, o_qi + rq);
        pfa[i % SPD] = bld(VA, vx, rq);
        pfk[i % SPD] = kzt[oz + t + SPD];             // oz + t + SPD in [-H, NZ + H + SUNF)
        if constexpr (TAN) {
          pfdp[i % SPD] = bld(VW, vx, o_dpi + rq);
          pfdq[i % SPD] = bld(VW, vx, o_dqi + rq);
          pfd[i % SPD] = bld(VD, vx, rq);
        }
        rq = next_row(rq, rowb, P4, P4);
      }
      __syncthreads();
    };
    // does some level 1..KL compute window row r in steps tb .. tb+SUNF-1
    auto meets = [&](int tb, int r) { return r + 2 <= tb + SUNF - 1 && r + 2 * KL >= tb; };
    for (int tb = 0; tb < U; tb += SUNF) {
      if (meets(tb, su0) || meets(tb, su1) || meets(tb, rui))
        static_for<0, SUNF>([&](auto ic) { body(ic, std::true_type(), tb + decltype(ic)::value); });
      else
        static_for<0, SUNF>([&](auto ic) { body(ic, std::false_type(), tb + decltype(ic)::value); });
    }
    if (seis && rui >= H && colint && rc >= 0) {
#pragma unroll
      for (int l = 1; l <= KL; ++l) {
        const int it = n0 + l - 2;
        if (it >= 0 && it < nt) {
          seis[(int64_t)it * n_rcv + rc] = rec[l - 1];
          if constexpr (TAN) dseis[(int64_t)it * n_rcv + rc] = drec[l - 1];
        }
      }
    }
  }
}

// ---------------------------------------------------------------- tangent on the saved history
struct BornArgs {
  const TileEnt* tiles;
  const int* n_tiles;
  const WaveDev* waves;
  const ModelDev* models;
  const ParamDev* params;
  const WaveParam* wparam;
  const float* alpha;
  const float* dalpha;
  const float* wavelet;
  float* arena;
  float* dseis;
  int nt;
  int n0, K;             // advance (dp, dq) n0 -> n0+K
  int Kt;                // steps the tile list was built for (halo 2*Kt); K <= Kt
  int dp_in_slot, dp_out_slot, dq_in_slot, dq_out_slot;
};

// K steps of the Born recursion (update_p_diff :470-509) with the forward
// wavefield taken from the saved history: q[n+1] of every band cell is read
// from HBM (q[n] is the previous step's) instead of advancing p, and
// alpha*Lf(p[n]) = q[n+1] - (1-kappa) q[n] (minus the source term at the
// source cell), which leaves one stencil per step.  The next step's q is
// requested as soon as the current step's last use of the registers it lands
// in is issued, so the barrier covers part of its latency.  Same window layout and double-buffering as k_fwd, and its
// fused multiply-add sequence, so dseis and (dp, dq) are bit-identical to the
// fused forward + tangent pass (whatever the row ownership of either).
template <int RX, int RZ, int NT, bool STRIP>
__global__ __launch_bounds__(NT) void k_born(BornArgs a) {
  constexpr int NW = NT / 64, CPR = RX / 64, RPW = RZ / NW, NC = CPR * RPW;
  static_assert(RZ % NW == 0, "rows must split evenly over the wavefronts");
  static_assert(NC <= 32 && RPW <= 32, "tile too large for the bit masks");
  constexpr int PW = RX + 4, PH = RZ + 4;
  __shared__ float D0[PH][PW], D1[PH][PW];
  __shared__ float SW[KMAX];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform rows
  const int r0 = wv * RPW;
  const int ntiles = *a.n_tiles;
  const int H = 2 * a.Kt;
  for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
    const TileEnt te = a.tiles[e];
    if (te.w < 0) continue;
    const WaveDev& w = a.waves[te.w];
    const ModelDev& m = a.models[w.model];
    const int NZ = w.NZ, NX = w.NX;
    const int64_t Pn = (int64_t)NZ * NX;
    const float* dpin = a.arena + w.off[R_DP] + a.dp_in_slot * Pn;
    const float* dqin = a.arena + w.off[R_DQ] + a.dq_in_slot * Pn;
    const float* hist = a.arena + w.off[R_HIST];   // q[n] at slot n
    const float* alpha = a.alpha + m.cell_off;
    const float* dalpha = a.dalpha + m.cell_off;
    const float kscale = a.params[w.model].kscale;
    const float dkscale = a.params[w.model].dkscale;
    const float amp = a.wparam[te.w].amp;
    const float amp_d = a.wparam[te.w].amp_d;
    const int oz = te.z0 - H, ox = te.x0 - H;
    const int gint0 = oz * NX + ox;
    float dpv[NC], dq[NC], al[NC], dal[NC], jj[NC];
    float qo[NC], qn[NC];   // q[n], q[n+1] of the step
    int gw[NC];
    if (threadIdx.x < KMAX) {
      const int it = a.n0 + (int)threadIdx.x - 1;
      SW[threadIdx.x] = ((int)threadIdx.x < a.K && it >= 0 && it < a.nt) ? a.wavelet[it] : 0.f;
    }
    unsigned srcmask = 0, rcvmask = 0, intmask = 0, srcrow = 0;
    const int srow = w.src_cell / NX;
    // per own column, once per tile: wrapped x, receiver index, the x damping
    // profile (used in the left/right strips), interior flag
    int gxc[CPR], rcol[CPR];
    bool isxc[CPR], cin[CPR];
    float jjxc[CPR];
#pragma unroll
    for (int c = 0; c < CPR; ++c) {
      const int cc = lane + 64 * c;
      const int gx = wrapi(ox + cc, NX);
      gxc[c] = gx;
      rcol[c] = rcv_index(gx, m);
      isxc[c] = gx < m.nbc || gx >= m.nx + m.nbc;
      const int jx = damp_j(m.nbc, gx, m.nz, m.nx, m.nbc);
      jjxc[c] = (float)(jx * jx);
      cin[c] = cc >= H && cc < RX - H && ox + cc < NX;
    }
    const float* h0 = hist + (int64_t)a.n0 * Pn;
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
      const int r = own_row<STRIP, NW, RPW>(wv, j);
      const int gz = wrapi(oz + r, NZ);
      if (gz == m.rcv_row) rcvmask |= 1u << j;
      if (gz == srow) srcrow |= 1u << j;
      const bool rin = r >= H && r < RZ - H && oz + r < NZ;
      const int jz = damp_j(gz, m.nbc, m.nz, m.nx, m.nbc);   // the z profile (outside the strips)
      const float jjz = (float)(jz * jz);
      const int grow = gz * NX;
#pragma unroll
      for (int c = 0; c < CPR; ++c) {
        const int i = j * CPR + c;
        const int cc = lane + 64 * c;
        const int g = grow + gxc[c];
        gw[i] = g;
        qo[i] = h0[g];
        qn[i] = h0[Pn + g];
        dpv[i] = dpin[g];
        D0[r + 2][cc + 2] = dpv[i];
        dq[i] = dqin[g];
        al[i] = alpha[g];
        dal[i] = dalpha[g];
        jj[i] = isxc[c] ? jjxc[c] : jjz;
        if (g == w.src_cell) srcmask |= 1u << i;
        if (rin && cin[c]) intmask |= 1u << i;
      }
    }
    __syncthreads();
    // Step s reads the window buffer s&1 and writes the other; the buffer
    // is a compile-time choice so that reads and writes never alias.
    auto step = [&](auto cb, const int s) {
      constexpr int CUR = decltype(cb)::value;
      float (*const Drd)[PW] = CUR ? D1 : D0;
      float (*const Dwr)[PW] = CUR ? D0 : D1;
      const int it = a.n0 + s - 1;
      const bool live = it >= 0 && it < a.nt;
      const float src = amp * SW[s];
      const float srcd = amp_d * SW[s];
      const int lo = 2 * (s + 1);
      if (!STRIP || (r0 + RPW > lo && r0 < RZ - lo)) {
#pragma unroll
        for (int c = 0; c < CPR; ++c) {
          const int cc = lane + 64 * c;
          float dcol[STRIP ? RPW + 4 : 1];
          if constexpr (STRIP) load_col<PW, RPW, CPR>(Drd, dpv, c, r0, cc, dcol);
#pragma unroll
          for (int j = 0; j < RPW; ++j) {
            const int r = own_row<STRIP, NW, RPW>(wv, j);
            const int i = j * CPR + c;
            float Ld;
            if constexpr (STRIP) Ld = lap_col<PW>(Drd, dcol, j, r, cc);
            else Ld = lap_lds<PW>(Drd, dpv[i], r, cc);
            const float k1 = 1.0f - kscale * jj[i];
            const float dk = dkscale * jj[i];
            float aL = fmaf(-k1, qo[i], qn[i]);
            if ((srcrow >> j) & 1u) aL = ((srcmask >> i) & 1u) ? aL - src : aL;
            float dqn = fmaf(k1, dq[i], al[i] * Ld);
            dqn = fmaf(dal[i], aL, dqn);
            dqn = fmaf(-dk, qo[i], dqn);
            if ((srcrow >> j) & 1u) dqn = ((srcmask >> i) & 1u) ? dqn + srcd : dqn;
            const float dp1 = dpv[i] + dqn;
            Dwr[r + 2][cc + 2] = dp1;
            dpv[i] = dp1;
            dq[i] = dqn;
          }
        }
      }
      if (live && rcvmask) {  // receiver samples of the interior
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
          if (!((rcvmask >> j) & 1u)) continue;
#pragma unroll
          for (int c = 0; c < CPR; ++c) {
            const int i = j * CPR + c;
            if (!((intmask >> i) & 1u) || rcol[c] < 0) continue;
            a.dseis[w.seis_off + (int64_t)it * m.n_rcv + rcol[c]] = dpv[i];
          }
        }
      }
      if (s + 1 < a.K) {  // q of the next step (every owned cell, back to back)
        const brsrc_t VH = brsrc(hist + (int64_t)(a.n0 + s + 2) * Pn);   // the slot (see k_adj)
#pragma unroll
        for (int i = 0; i < NC; ++i) {
          qo[i] = qn[i];
          qn[i] = bld(VH, (unsigned)gw[i] * 4u, 0u);
        }
      }
      __syncthreads();
    };
    for (int s = 0; s < a.K; s += 2) {
      step(std::integral_constant<int, 0>(), s);
      if (s + 1 < a.K) step(std::integral_constant<int, 1>(), s + 1);
    }
    float* dpout = a.arena + w.off[R_DP] + a.dp_out_slot * Pn;
    float* dqout = a.arena + w.off[R_DQ] + a.dq_out_slot * Pn;
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
#pragma unroll
      for (int c = 0; c < CPR; ++c) {
        const int i = j * CPR + c;
        if (!((intmask >> i) & 1u)) continue;
        const int g = gint0 + own_row<STRIP, NW, RPW>(wv, j) * NX + lane + 64 * c;
        dpout[g] = dpv[i];
        dqout[g] = dq[i];
      }
    }
  }
}

// KL steps of the Born recursion (update_p_diff :470-509) on the saved
// history, streamed like k_fws: a column strip of NT padded columns (one per
// thread) marched down the rows of a segment; pipeline step t takes in
// window row t of dp[n0], dq[n0], q[n0] and the row's coefficients (alpha,
// dalpha/alpha, 1 - kappa, dkappa) as level 0, and level l = 1..KL computes
// window row t - 2l of (dp, dq)[n0+l] from level l-1, from q[n0+l] of that
// row (requested SPH steps ahead) and from q[n0+l-1] of that row, which
// level l-1 used two steps earlier and left in a three-slot register ring.  The coefficients of a
// row sit in a private LDS ring of SUNB slots, written after the step's
// levels have read the slot's previous row (the last level reads row t-2KL).
// Same stencil order and fused multiply-add sequence as k_born, so dseis and
// (dp, dq) are bit-identical to it (and to the fused tangent).  Launches
// always advance KL steps (a shorter last launch runs k_born).
constexpr int SPB = 2;    // steps between a level-0 row's request and its intake
constexpr int SPH = 2;    // steps between a history row's request and its use
template <int NT, int KL>
__global__ __launch_bounds__(NT) void k_bns(BornArgs a) {
  // pipeline steps per unrolled block, also the coefficient ring length
  constexpr int SUNB = 2 * KL <= 6 ? 6 : 12;
  static_assert(2 * KL <= SUNB && SUNB % 6 == 0 && SUNB % SPB == 0 && SUNB % SPH == 0, "pipeline rings");
  constexpr int H = 2 * KL;
  constexpr int RS = (3 * KL) | 1;        // column-major level rings, odd stride (see k_fws)
  __shared__ float RG[(NT + 4) * RS];
  __shared__ float4 CF[SUNB][NT];         // (alpha, dalpha, 1 - kappa, dkappa) of window row u at [u % SUNB]
  const unsigned c = threadIdx.x;
  const int ntiles = *a.n_tiles;
  const int n0 = a.n0, nt = a.nt;
  for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
    const TileEnt te = a.tiles[e];
    if (te.w < 0) continue;
    const WaveDev& w = a.waves[te.w];
    const ModelDev& m = a.models[w.model];
    const int NZ = w.NZ, NX = w.NX, nz = m.nz, nx = m.nx, nbc = m.nbc;
    const unsigned rowb = (unsigned)NX * 4u, P4 = (unsigned)NZ * rowb;   // bytes of a row, of a slot
    // the wavefield's state through one buffer resource (history region
    // first: q[n] at slot n), each coefficient field through its own; byte
    // offsets of the slots this launch reads and writes
    const brsrc_t VW = brsrc(a.arena + w.off[R_HIST]);
    const brsrc_t VA = brsrc(a.alpha + m.cell_off), VD = brsrc(a.dalpha + m.cell_off);
    const unsigned rdp = (unsigned)(w.off[R_DP] - w.off[R_HIST]) * 4u;
    const unsigned rdq = (unsigned)(w.off[R_DQ] - w.off[R_HIST]) * 4u;
    const unsigned o_dpi = rdp + (unsigned)a.dp_in_slot * P4, o_dqi = rdq + (unsigned)a.dq_in_slot * P4;
    const unsigned o_dpo = rdp + (unsigned)a.dp_out_slot * P4, o_dqo = rdq + (unsigned)a.dq_out_slot * P4;
    const unsigned o_h0 = (unsigned)n0 * P4;      // q[n0]
    float* dseis = a.dseis + w.seis_off;
    const float kscale = a.params[w.model].kscale;
    const float dkscale = a.params[w.model].dkscale;
    const float amp = a.wparam[te.w].amp;
    const float amp_d = a.wparam[te.w].amp_d;
    const int n_rcv = m.n_rcv, rcv_row = m.rcv_row;
    const int IZ = te.pad;
    const int U = IZ + 2 * H;
    const int oz = te.z0 - H, ox = te.x0 - H;
    const unsigned gx = wrapi(ox + (int)c, NX);
    const unsigned vx = gx * 4u;                  // the thread's column, bytes into a row
    const bool colint = (int)c >= H && (int)c < NT - H && ox + (int)c < NX;
    const bool isx = (int)gx < nbc || (int)gx >= nx + nbc;
    const int jx = damp_j(0, gx, nz, nx, nbc);
    const float jjx = (float)(jx * jx);
    const int rc = rcv_index(gx, m);
    const int srow = w.src_cell / NX;
    const bool srcc = (int)gx == w.src_cell - srow * NX;
    const int su0 = wrapi(srow - oz, NZ), su1 = su0 + NZ;
    const int ruz = rcv_row - te.z0;
    const int rui = ruz >= 0 && ruz < IZ ? ruz + H : -(1 << 20);
    float srcv[KL], srcp[KL], rec[KL];
#pragma unroll
    for (int l = 1; l <= KL; ++l) {
      const int it = n0 + l - 2;
      const float wl = (it >= 0 && it < nt) ? a.wavelet[it] : 0.f;
      srcv[l - 1] = amp_d * wl;
      srcp[l - 1] = amp * wl;   // the forward's source term in q[n0+l]
      rec[l - 1] = 0.f;
    }
    float pr[KL][6], qr[KL][3], pfp[SPB], pfq[SPB], pfa[SPB], pfd[SPB], pfh[SPB];
    float hq[KL][SPH];   // q[n0+l] of the row level l meets SPH steps ahead
    float qps[KL][3];    // q[n0+l] of the row level l met at step t, at [t % 3]
#pragma unroll
    for (int l = 0; l < KL; ++l) {
#pragma unroll
      for (int s = 0; s < 6; ++s) pr[l][s] = 0.f;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        qr[l][s] = 0.f;
        qps[l][s] = 0.f;
      }
#pragma unroll
      for (int s = 0; s < SPH; ++s) hq[l][s] = 0.f;
    }
    // coefficients of padded row gz into the ring slot (window row u at u % SUNB)
    auto put_cf = [&](int slot, int gz, float al, float dal) {
      const int jz = damp_j(gz, nbc, nz, nx, nbc);
      const float jj = isx ? jjx : (float)(jz * jz);
      CF[slot][c] = make_float4(al, dal, 1.0f - kscale * jj, dkscale * jj);
    };
    auto row_off = [&](int u) { return (unsigned)wrapi(oz + u, NZ) * rowb; };  // window row u
#pragma unroll
    for (int d = 0; d < SPB; ++d) {
      const unsigned ro = row_off(d);
      pfp[d] = bld(VW, vx, o_dpi + ro);
      pfq[d] = bld(VW, vx, o_dqi + ro);
      pfa[d] = bld(VA, vx, ro);
      pfd[d] = bld(VD, vx, ro);
      pfh[d] = bld(VW, vx, o_h0 + ro);
    }
    // history of the first steps: level l meets window row t - 2l at step t;
    // hb[l-1] is the offset of the row level l requests next (q[n0+l] of
    // window row t - 2l + SPH), he[l-1] the end of that slot
    unsigned hb[KL], he[KL];
#pragma unroll
    for (int l = 1; l <= KL; ++l) {
      const unsigned slot = o_h0 + (unsigned)l * P4;
#pragma unroll
      for (int d = 0; d < SPH; ++d) hq[l - 1][d] = bld(VW, vx, slot + row_off(d - 2 * l));
      hb[l - 1] = slot + row_off(SPH - 2 * l);
      he[l - 1] = slot + P4;
    }
    unsigned rq = row_off(SPB);            // next level-0 request
    int gzc = wrapi(oz, NZ);               // padded row of window row t
    // (oz + u) * rowb of level KL's row u = t - 2 KL (used on committed rows only)
    unsigned so = (unsigned)(te.z0 - 2 * H) * rowb;
    __syncthreads();
    auto body = [&](auto ic, auto sc, const int t) {
      constexpr int i = decltype(ic)::value;
      constexpr bool SPECIAL = decltype(sc)::value;
      float al0 = 0.f, dal0 = 0.f;
      {  // level 0: intake of window row t (its coefficients go to the ring after the levels)
        const float p0 = pfp[i % SPB];
        RG[(c + 2) * RS + i % 3] = p0;
        pr[0][i % 6] = p0;
        qr[0][i % 3] = pfq[i % SPB];
        al0 = pfa[i % SPB];
        dal0 = pfd[i % SPB];
      }
      const float qin0 = pfh[i % SPB];   // q[n0] of row t: level 1 reads it two steps on
      static_for<1, KL + 1>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const int u = t - 2 * l;
        constexpr int B = 4 * SUNB;
        constexpr int s6 = (i - 2 * l + B) % 6, s3 = (i - 2 * l + B) % 3, sr = (i - 2 * l + B) % SUNB;
        const float pc = pr[l - 1][s6];
        const float* T = RG + c * RS + (l - 1) * 3 + s3;
        float t1 = (T[3 * RS] - pc) + (T[RS] - pc);
        t1 = t1 + (pr[l - 1][(i - 2 * l + 1 + B) % 6] - pc);
        t1 = t1 + (pr[l - 1][(i - 2 * l - 1 + B) % 6] - pc);
        float t2 = (T[4 * RS] - pc) + (T[0] - pc);
        t2 = t2 + (pr[l - 1][(i - 2 * l + 2 + B) % 6] - pc);
        t2 = t2 + (pr[l - 1][(i - 2 * l - 2 + B) % 6] - pc);
        const float Ld = fmaf(kC2, t1, kC3 * t2);
        const float4 cf = CF[sr][c];
        const float qnh = hq[l - 1][i % SPH];                  // q[n0+l] of row u
        const float qoh = qps[l - 1][(i - 2 + B) % 3];         // q[n0+l-1] of row u
        float aL = fmaf(-cf.z, qoh, qnh);
        if constexpr (SPECIAL) {
          if (u == su0 || u == su1) aL = srcc ? aL - srcp[l - 1] : aL;
        }
        float dqn = fmaf(cf.z, qr[l - 1][s3], cf.x * Ld);
        dqn = fmaf(cf.y, aL, dqn);
        dqn = fmaf(-cf.w, qoh, dqn);
        if constexpr (SPECIAL) {
          if (u == su0 || u == su1) dqn = srcc ? dqn + srcv[l - 1] : dqn;
        }
        const float dp1 = pc + dqn;
        if constexpr (SPECIAL) {
          if (u == rui) rec[l - 1] = dp1;
        }
        if constexpr (l < KL) {
          RG[(c + 2) * RS + l * 3 + s3] = dp1;
          pr[l][s6] = dp1;
          qr[l][s3] = dqn;
        } else if (u >= H && u < H + IZ) {  // committed row of (dp, dq)[n0+KL]
          if (colint) {
            bst(VW, vx, o_dpo + so, dp1);
            bst(VW, vx, o_dqo + so, dqn);
          }
        }
        if constexpr (l < KL) qps[l][i % 3] = qnh;      // for level l+1, two steps on
        // this level's q for its row SPH steps ahead
        hq[l - 1][i % SPH] = bld(VW, vx, hb[l - 1]);
        hb[l - 1] = next_row(hb[l - 1], rowb, he[l - 1], P4);
      });
      qps[0][i % 3] = qin0;
      put_cf(i % SUNB, gzc, al0, dal0);
      gzc = gzc + 1 == NZ ? 0 : gzc + 1;
      so += rowb;
      {  // request of window row t + SPB
        pfp[i % SPB] = bld(VW, vx, o_dpi + rq);
        pfq[i % SPB] = bld(VW, vx, o_dqi + rq);
        pfa[i % SPB] = bld(VA, vx, rq);
        pfd[i % SPB] = bld(VD, vx, rq);
        pfh[i % SPB] = bld(VW, vx, o_h0 + rq);
        rq = next_row(rq, rowb, P4, P4);
      }
      __syncthreads();
    };
    auto meets = [&](int tb, int r) { return r + 2 <= tb + SUNB - 1 && r + 2 * KL >= tb; };
    for (int tb = 0; tb < U; tb += SUNB) {
      if (meets(tb, su0) || meets(tb, su1) || meets(tb, rui))
        static_for<0, SUNB>([&](auto ic) { body(ic, std::true_type(), tb + decltype(ic)::value); });
      else
        static_for<0, SUNB>([&](auto ic) { body(ic, std::false_type(), tb + decltype(ic)::value); });
    }
    if (rui >= H && colint && rc >= 0) {
#pragma unroll
      for (int l = 1; l <= KL; ++l) {
        const int it = n0 + l - 2;
        if (it >= 0 && it < nt) dseis[(int64_t)it * n_rcv + rc] = rec[l - 1];
      }
    }
  }
}

// ---------------------------------------------------------------- adjoint
struct AdjArgs {
  const TileEnt* tiles;
  const int* n_tiles;
  const WaveDev* waves;
  const ModelDev* models;
  const ParamDev* params;
  WaveParam* wparam;
  const float* alpha;
  const float* wavelet;
  float* arena;
  const float* inj_a;   // adjoint source = inj_a - inj_b (inj_b may be null)
  const float* inj_b;
  int nt;
  int M0, K;            // produces lam[M0-1] .. lam[M0-K]
  int Kt;               // steps the tile list was built for (halo 2*Kt); K <= Kt
  int lam_in_slot, lam_out_slot, mu_in_slot, mu_out_slot;
  int h_base;           // the forward's q[n] sits at R_HIST slot n - h_base
};

// K reverse steps of update_p_adjoint (:517-579) in gather form for every
// tile, with the imaging accumulators (alpha / kappa adjoints) and the
// source-wavelet correlation (:314) fused on the tile interior.  lam, mu,
// alpha, (1-kappa) and the imaging accumulators of owned cells stay in
// registers; alpha*lam is staged in a double-buffered LDS window for the
// stencil (row ownership as k_fwd, one barrier per step).  The forward
// enters only through the saved q of interior cells: step m uses q[m-1] and
// q[m] (the previous step's), alpha*Lf(p[m-1]) = q[m] - (1-kappa) q[m-1]
// (minus the source term at the source cell), and the alpha accumulator
// collects lam * alpha*Lf(p), divided by alpha when the gradient is folded;
// the next step's q is requested right after the current step's last use.
// The accumulators see the same fused multiply-add sequence as a per-step
// read-modify-write, so results do not depend on K.
template <int RX, int RZ, int NT, bool STRIP>
__global__ __launch_bounds__(NT) void k_adj(AdjArgs a) {
  constexpr int NW = NT / 64, CPR = RX / 64, RPW = RZ / NW, NC = CPR * RPW;
  static_assert(RZ % NW == 0, "rows must split evenly over the wavefronts");
  static_assert(NC <= 32 && RPW <= 32, "tile too large for the bit masks");
  constexpr int PW = RX + 4, PH = RZ + 4;     // alpha*lam windows with a 2-cell apron
  // the two alpha*lam windows interleaved by rows (row r of window b at
  // AA[r][b]): one register holding the address of AA[r][0][lane] reaches
  // the neighbours of both windows with ds_read2 offsets
  __shared__ float AA[PH][2][PW];
  constexpr int PS = 2 * PW;   // row stride of either window
  __shared__ float INJ[KMAX][RX];
  __shared__ float SW[KMAX];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform rows
  const int r0 = wv * RPW;
  const int ntiles = *a.n_tiles;
  const int H = 2 * a.Kt;  // window halo of the tile list; a.K <= a.Kt steps run
  for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
    const TileEnt te = a.tiles[e];
    if (te.w < 0) continue;
    const WaveDev& w = a.waves[te.w];
    const ModelDev& m = a.models[w.model];
    const int NZ = w.NZ, NX = w.NX;
    const int64_t Pn = (int64_t)NZ * NX;
    const float* lin = a.arena + w.off[R_LAM] + a.lam_in_slot * Pn;
    const float* min_ = a.arena + w.off[R_MU] + a.mu_in_slot * Pn;
    float* GA = a.arena + w.off[R_GA];
    float* GK = a.arena + w.off[R_GK];
    const float* hist = a.arena + w.off[R_HIST];   // q[n] at slot n - h_base
    const float* alpha = a.alpha + m.cell_off;
    const float kscale = a.params[w.model].kscale;
    const float amp = a.wparam[te.w].amp;
    const int oz = te.z0 - H, ox = te.x0 - H;
    const int gint0 = oz * NX + ox;
    float lam[NC], mu[NC], al[NC], k1[NC], ga[NC], gk[NC];
    float alm[NC];   // alpha * lam, the value the stencil sees
    // q of interior cells: at even steps qx holds q[m] and qy q[m-1], at odd
    // steps the roles swap, so the history needs no register copies; the
    // array holding q[m] takes alpha*Lf(p[m-1]) and then the next load
    float qx[NC], qy[NC];
    int hidx[NC];  // history index of an interior cell; 0 (a dummy read) elsewhere
    unsigned srcmask = 0, rcvmask = 0, intmask = 0;
    // per own column, once per tile: wrapped x, receiver index, 1 - kappa in
    // the left/right damping strips, interior flag
    int gxc[CPR], rcol[CPR];
    bool isxc[CPR], cin[CPR];
    float kxc[CPR];
#pragma unroll
    for (int c = 0; c < CPR; ++c) {
      const int cc = lane + 64 * c;
      const int gx = wrapi(ox + cc, NX);
      gxc[c] = gx;
      rcol[c] = rcv_index(gx, m);
      isxc[c] = gx < m.nbc || gx >= m.nx + m.nbc;
      const int jx = damp_j(m.nbc, gx, m.nz, m.nx, m.nbc);   // the x profile (used in the strips)
      kxc[c] = 1.0f - kscale * (float)(jx * jx);
      cin[c] = cc >= H && cc < RX - H && ox + cc < NX;
    }
    if (threadIdx.x < KMAX) {
      const int t = a.M0 - 3 - (int)threadIdx.x;
      SW[threadIdx.x] = ((int)threadIdx.x < a.K && t >= 0 && t < a.nt) ? a.wavelet[t] : 0.f;
    }
    // adjoint sources of this launch's steps along the window's columns (each
    // wavefront fills steps wv, wv + NW, ... of its lanes' columns)
    for (int s = wv; s < a.K; s += NW) {
#pragma unroll
      for (int c = 0; c < CPR; ++c) {
        float v = 0.f;
        if (rcol[c] >= 0) {
          const int64_t si = w.seis_off + (int64_t)(a.M0 - 3 - s) * m.n_rcv + rcol[c];
          v = a.inj_b ? a.inj_a[si] - a.inj_b[si] : a.inj_a[si];
        }
        INJ[s][lane + 64 * c] = v;
      }
    }
    const float* h0 = hist + (int64_t)(a.M0 - 2 - a.h_base) * Pn;  // q[M0-2]; q[M0-1] one slot on
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
      const int r = own_row<STRIP, NW, RPW>(wv, j);
      const int gz = wrapi(oz + r, NZ);
      if (gz == m.rcv_row) rcvmask |= 1u << j;
      const bool rin = r >= H && r < RZ - H && oz + r < NZ;        // wave-uniform
      const int jz = damp_j(gz, m.nbc, m.nz, m.nx, m.nbc);        // the z profile (outside the strips)
      const float kz = 1.0f - kscale * (float)(jz * jz);
      const int grow = gz * NX, gir = gint0 + r * NX;
#pragma unroll
      for (int c = 0; c < CPR; ++c) {
        const int i = j * CPR + c;
        const int cc = lane + 64 * c;
        const int g = grow + gxc[c];
        lam[i] = lin[g];
        mu[i] = min_[g];
        al[i] = alpha[g];
        alm[i] = al[i] * lam[i];
        AA[r + 2][0][cc + 2] = alm[i];
        k1[i] = isxc[c] ? kxc[c] : kz;
        if (g == w.src_cell) srcmask |= 1u << i;
        // interior cells load their accumulators and history; the others read
        // one dummy address (their sums are never stored)
        const bool inner = rin && cin[c];
        if (inner) intmask |= 1u << i;
        hidx[i] = inner ? gir + cc : 0;
        ga[i] = GA[hidx[i]];
        gk[i] = GK[hidx[i]];
        qy[i] = h0[hidx[i]];
        qx[i] = h0[Pn + hidx[i]];
      }
    }
    double ssum = 0.0;
    const bool src_here = (srcmask & intmask) != 0;
    if (src_here) ssum = a.wparam[te.w].ssum;
    // LDS addresses of the own rows, &AA[r + 2][0][lane]
    unsigned rb[STRIP ? RPW : 1];
    if constexpr (STRIP) {
#pragma unroll
      for (int j = 0; j < RPW; ++j) rb[j] = lds_row(&AA[r0 + j + 2][0][lane]);
    }
    __syncthreads();
    // Step s reads the window buffer s&1 and writes the other; the buffer
    // is a compile-time choice so that reads and writes never alias.
    auto step = [&](auto cb, const int s) {
      constexpr int CUR = decltype(cb)::value;
      float (*const Ard)[PS] = reinterpret_cast<float (*)[PS]>(&AA[0][CUR][0]);
      float (*const Awr)[PS] = reinterpret_cast<float (*)[PS]>(&AA[0][1 - CUR][0]);
      float (&qm)[NC] = CUR ? qy : qx;   // q[m]; becomes alpha*Lf(p[m-1]), then q[m-2]
      float (&ql)[NC] = CUR ? qx : qy;   // q[m-1]
      const int mm = a.M0 - 1 - s;
      const int lo = 2 * (s + 1);
      const float sw = SW[s];
      if (!STRIP || (r0 + RPW > lo && r0 < RZ - lo)) {
#pragma unroll
        for (int c = 0; c < CPR; ++c) {
          const int cc = lane + 64 * c;
          float acol[STRIP ? RPW + 4 : 1];
          if constexpr (STRIP) {
            acol[0] = Ard[r0][cc + 2];
            acol[1] = Ard[r0 + 1][cc + 2];
#pragma unroll
            for (int j = 0; j < RPW; ++j) acol[j + 2] = alm[j * CPR + c];
            acol[RPW + 2] = Ard[r0 + RPW + 2][cc + 2];
            acol[RPW + 3] = Ard[r0 + RPW + 3][cc + 2];
          }
          const float inj = INJ[s][cc];
          const bool rc = rcol[c] >= 0;
#pragma unroll
          for (int j = 0; j < RPW; ++j) {
            const int r = own_row<STRIP, NW, RPW>(wv, j);
            const int i = j * CPR + c;
            float La;
            if constexpr (STRIP) La = lap_row(lds_at(rb[j]) + CUR * PW + 64 * c, acol, j);
            else La = lap_lds<PS>(Ard, alm[i], r, cc);
            float mun = fmaf(k1[i], mu[i], La);
            if ((rcvmask >> j) & 1u) mun = rc ? mun + inj : mun;
            const float ln = lam[i] + mun;
            alm[i] = al[i] * ln;
            if constexpr (STRIP) lds_at(rb[j])[(1 - CUR) * PW + 64 * c + 2] = alm[i];
            else Awr[r + 2][cc + 2] = alm[i];
            lam[i] = ln;
            mu[i] = mun;
          }
        }
      }
      // imaging on the interior (always inside the band): lam now holds lam[m]
      // GA += lam[m] alpha Lf(p[m-1]);  GK += lam[m] (p[m-2] - p[m-1]) = -lam[m] q[m-1]
      // every cell accumulates; only interior cells' sums are stored.
      // alpha Lf(p[m-1]) = q[m] - (1-kappa) q[m-1], less the forward's source
      // term at the source cell (a divergent branch taken by one lane only)
#pragma unroll
      for (int i = 0; i < NC; ++i) qm[i] = fmaf(-k1[i], ql[i], qm[i]);
      if (src_here) {  // only the thread owning the source cell
        const float srcf = amp * sw;   // the forward's source term in q[mm]
#pragma unroll
        for (int i = 0; i < NC; ++i)
          if ((srcmask & intmask) >> i & 1u) {
            qm[i] = qm[i] - srcf;
            ssum += (double)lam[i] * (double)sw;
          }
      }
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        ga[i] = fmaf(lam[i], qm[i], ga[i]);
        gk[i] = fmaf(lam[i], -ql[i], gk[i]);
      }
      // q[mm - 2] for the next step into the array that held q[m], issued
      // unconditionally (cells outside the interior all read one dummy
      // address; the last step re-reads q[mm - 1], just read) so that the
      // loads go out back to back, through a buffer resource over the slot
      // (built on the scalar unit) so no load carries vector address arithmetic
      {
        const int slot = (s + 1 < a.K ? mm - 2 : mm - 1) - a.h_base;
        const brsrc_t VH = brsrc(hist + (int64_t)slot * Pn);
#pragma unroll
        for (int i = 0; i < NC; ++i) qm[i] = bld(VH, (unsigned)hidx[i] * 4u, 0u);
      }
      __syncthreads();
    };
    for (int s = 0; s < a.K; s += 2) {
      step(std::integral_constant<int, 0>(), s);
      if (s + 1 < a.K) step(std::integral_constant<int, 1>(), s + 1);
    }
    float* lout = a.arena + w.off[R_LAM] + a.lam_out_slot * Pn;
    float* mout = a.arena + w.off[R_MU] + a.mu_out_slot * Pn;
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
#pragma unroll
      for (int c = 0; c < CPR; ++c) {
        const int i = j * CPR + c;
        if (!((intmask >> i) & 1u)) continue;
        const int g = gint0 + own_row<STRIP, NW, RPW>(wv, j) * NX + lane + 64 * c;
        lout[g] = lam[i];
        mout[g] = mu[i];
        GA[g] = ga[i];
        GK[g] = gk[i];
      }
    }
    if (src_here) a.wparam[te.w].ssum = ssum;
    __syncthreads();
  }
}

// KL reverse steps of update_p_adjoint (:517-579) with the imaging terms,
// streamed like k_bns: a column strip of NT padded columns (one per thread)
// marched down the rows of a segment.  Pipeline step t takes in window row t
// of lam[M0], mu[M0], alpha, the imaging accumulators and q[M0-1] as level 0,
// and level l = 1..KL computes window row t - 2l of (lam, mu)[M0-l] from level
// l-1 and adds its imaging terms to that row's accumulators, which level l-1
// passed on two steps earlier; level KL stores the committed rows.  Level l
// reads q[M0-l-1] of its row (requested SPH steps ahead) and takes q[M0-l]
// from level l-1, which read it two steps earlier.  alpha*lam of every level
// sits in registers (own column) and in a three-row LDS ring (neighbours);
// (alpha, 1 - kappa) of the rows in flight in a private LDS ring, 1 - kappa
// outside the left/right damping strips from the model's row table.  Same
// stencil order and fused multiply-add sequence as k_adj, and each cell's
// accumulators see the same sequence of updates, so lam, mu, the imaging
// accumulators and the source-wavelet correlation are bit-identical to it.
// Launches always advance KL steps (shorter ones run k_adj).
template <int NT, int KL>
__global__ __launch_bounds__(NT) void k_ads(AdjArgs a) {
  constexpr int SUNA = 2 * KL <= 12 ? 12 : 24;   // pipeline steps per unrolled block; coefficient ring length
  static_assert(2 * KL <= SUNA && SUNA % 6 == 0 && SUNA % SPB == 0 && SUNA % SPH == 0, "pipeline rings");
  constexpr int H = 2 * KL;
  constexpr int RS = (3 * KL) | 1;        // column-major level rings, odd stride (see k_fws)
  __shared__ float RG[(NT + 4) * RS];     // alpha*lam of levels 0..KL-1
  __shared__ float2 AK[SUNA][NT];         // (alpha, 1 - kappa) of window row u at [u % SUNA], own column
  const unsigned c = threadIdx.x;
  const int ntiles = *a.n_tiles;
  const int M0 = a.M0, nt = a.nt;
  for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
    const TileEnt te = a.tiles[e];
    if (te.w < 0) continue;
    const WaveDev& w = a.waves[te.w];
    const ModelDev& m = a.models[w.model];
    const int NZ = w.NZ, NX = w.NX, nz = m.nz, nx = m.nx, nbc = m.nbc;
    const unsigned rowb = (unsigned)NX * 4u, P4 = (unsigned)NZ * rowb;   // bytes of a row, of a slot
    const brsrc_t VW = brsrc(a.arena + w.off[R_HIST]);
    const brsrc_t VA = brsrc(a.alpha + m.cell_off);
    auto reg_off = [&](int r, int slot) {
      return (unsigned)(w.off[r] - w.off[R_HIST]) * 4u + (unsigned)slot * P4;
    };
    const unsigned o_li = reg_off(R_LAM, a.lam_in_slot), o_mi = reg_off(R_MU, a.mu_in_slot);
    const unsigned o_lo = reg_off(R_LAM, a.lam_out_slot), o_mo = reg_off(R_MU, a.mu_out_slot);
    const unsigned o_ga = reg_off(R_GA, 0), o_gk = reg_off(R_GK, 0);
    auto o_q = [&](int n) { return (unsigned)(n - a.h_base) * P4; };   // q[n]
    const unsigned o_q1 = o_q(M0 - 1);
    const float kscale = a.params[w.model].kscale;
    const float amp = a.wparam[te.w].amp;
    const float* kzt = a.alpha + m.kz_off + kKzPad;   // kzt[z]: padded row z, -kKzPad <= z < NZ + kKzPad
    const int IZ = te.pad;
    const int U = IZ + 2 * H;
    const int oz = te.z0 - H, ox = te.x0 - H;
    const unsigned gx = wrapi(ox + (int)c, NX);
    const unsigned vx = gx * 4u;
    const bool colint = (int)c >= H && (int)c < NT - H && ox + (int)c < NX;
    const bool isx = (int)gx < nbc || (int)gx >= nx + nbc;
    const int jx = damp_j(0, gx, nz, nx, nbc);
    const float kx1 = 1.0f - kscale * (float)(jx * jx);
    const int rc = rcv_index(gx, m);
    const bool rcc = rc >= 0;
    const int srow = w.src_cell / NX;
    const bool srcc = (int)gx == w.src_cell - srow * NX;
    // window rows holding the source row and the receiver row (twice when
    // the window wraps); the committed instance of the source row, if any
    const int su0 = wrapi(srow - oz, NZ), su1 = su0 + NZ;
    const int ru0 = wrapi(m.rcv_row - oz, NZ), ru1 = ru0 + NZ;
    const int suz = srow - te.z0;
    const int sui = suz >= 0 && suz < IZ ? suz + H : -(1 << 20);
    const bool src_here = srcc && colint && sui >= 0;
    // per level: adjoint source sample of the thread's column, wavelet sample
    float injv[KL], sw[KL];
#pragma unroll
    for (int l = 1; l <= KL; ++l) {
      const int it = M0 - 2 - l;
      sw[l - 1] = (it >= 0 && it < nt) ? a.wavelet[it] : 0.f;
      float v = 0.f;
      if (rcc) {
        const int64_t si = w.seis_off + (int64_t)it * m.n_rcv + rc;
        v = a.inj_b ? a.inj_a[si] - a.inj_b[si] : a.inj_a[si];
      }
      injv[l - 1] = v;
    }
    double ssum = src_here ? a.wparam[te.w].ssum : 0.0;
    // rings of levels 0..KL-1 (own column): alpha*lam by row [u % 6]; lam, mu,
    // the accumulators after the level, and q[M0-l] by row [u % 3]
    float ar[KL][6], lr[KL][3], mr[KL][3], gar[KL][3], gkr[KL][3], qpr[KL][3];
    float pfl[SPB], pfm[SPB], pfa[SPB], pfg[SPB], pfk[SPB], pfq[SPB], pfz[SPB];
    float hq[KL][SPH];   // q[M0-l-1] of the row level l meets SPH steps ahead
#pragma unroll
    for (int l = 0; l < KL; ++l) {
#pragma unroll
      for (int s = 0; s < 6; ++s) ar[l][s] = 0.f;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        lr[l][s] = 0.f;
        mr[l][s] = 0.f;
        gar[l][s] = 0.f;
        gkr[l][s] = 0.f;
        qpr[l][s] = 0.f;
      }
    }
    auto row_off = [&](int u) { return (unsigned)wrapi(oz + u, NZ) * rowb; };  // window row u
#pragma unroll
    for (int d = 0; d < SPB; ++d) {
      const unsigned ro = row_off(d);
      pfl[d] = bld(VW, vx, o_li + ro);
      pfm[d] = bld(VW, vx, o_mi + ro);
      pfa[d] = bld(VA, vx, ro);
      pfg[d] = bld(VW, vx, o_ga + ro);
      pfk[d] = bld(VW, vx, o_gk + ro);
      pfq[d] = bld(VW, vx, o_q1 + ro);
      pfz[d] = kzt[oz + d];
    }
    // hb[l-1]: offset of the row level l requests next (q[M0-l-1] of window
    // row t - 2l + SPH), he[l-1] the end of that slot
    unsigned hb[KL], he[KL];
#pragma unroll
    for (int l = 1; l <= KL; ++l) {
      const unsigned slot = o_q(M0 - l - 1);
#pragma unroll
      for (int d = 0; d < SPH; ++d) hq[l - 1][d] = bld(VW, vx, slot + row_off(d - 2 * l));
      hb[l - 1] = slot + row_off(SPH - 2 * l);
      he[l - 1] = slot + P4;
    }
    unsigned rq = row_off(SPB);                     // next level-0 request
    unsigned so = (unsigned)(te.z0 - 2 * H) * rowb;  // (oz + u) * rowb of level KL's row u = t - 2 KL
    __syncthreads();
    unsigned sm[KL];   // per level: bit i set if step tb + i meets a source or receiver row
    auto body = [&](auto ic, const int t) {
      constexpr int i = decltype(ic)::value;
      float al0 = 0.f, k10 = 0.f;
      {  // level 0: intake of window row t (its coefficients go to the ring after the levels)
        al0 = pfa[i % SPB];
        k10 = isx ? kx1 : pfz[i % SPB];
        const float l0 = pfl[i % SPB];
        const float a0 = al0 * l0;
        RG[(c + 2) * RS + i % 3] = a0;
        ar[0][i % 6] = a0;
        lr[0][i % 3] = l0;
        mr[0][i % 3] = pfm[i % SPB];
        gar[0][i % 3] = pfg[i % SPB];
        gkr[0][i % 3] = pfk[i % SPB];
        qpr[0][i % 3] = pfq[i % SPB];
      }
      static_for<1, KL + 1>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const int u = t - 2 * l;
        constexpr int B = 4 * SUNA;
        constexpr int s6 = (i - 2 * l + B) % 6, s3 = (i - 2 * l + B) % 3, sr = (i - 2 * l + B) % SUNA;
        const float pc = ar[l - 1][s6];
        const float* T = RG + c * RS + (l - 1) * 3 + s3;
        float t1 = (T[3 * RS] - pc) + (T[RS] - pc);
        t1 = t1 + (ar[l - 1][(i - 2 * l + 1 + B) % 6] - pc);
        t1 = t1 + (ar[l - 1][(i - 2 * l - 1 + B) % 6] - pc);
        float t2 = (T[4 * RS] - pc) + (T[0] - pc);
        t2 = t2 + (ar[l - 1][(i - 2 * l + 2 + B) % 6] - pc);
        t2 = t2 + (ar[l - 1][(i - 2 * l - 2 + B) % 6] - pc);
        const float La = fmaf(kC2, t1, kC3 * t2);
        const float2 ak = AK[sr][c];
        float mun = fmaf(ak.y, mr[l - 1][s3], La);
        const bool sp = (sm[l - 1] >> i) & 1u;
        if (sp) {
          if (u == ru0 || u == ru1) mun = rcc ? mun + injv[l - 1] : mun;
        }
        const float ln = lr[l - 1][s3] + mun;
        // imaging: GA += lam[m] alpha Lf(p[m-1]), GK -= lam[m] q[m-1], m = M0 - l
        const float qlo = hq[l - 1][i % SPH];      // q[m-1] of row u
        const float qhi = qpr[l - 1][s3];          // q[m] of row u
        float aL = fmaf(-ak.y, qlo, qhi);
        if (sp) {
          if (u == su0 || u == su1) aL = srcc ? aL - amp * sw[l - 1] : aL;
          if (src_here && u == sui) ssum += (double)ln * (double)sw[l - 1];
        }
        const float gan = fmaf(ln, aL, gar[l - 1][s3]);
        const float gkn = fmaf(ln, -qlo, gkr[l - 1][s3]);
        if constexpr (l < KL) {
          const float an = ak.x * ln;
          RG[(c + 2) * RS + l * 3 + s3] = an;
          ar[l][s6] = an;
          lr[l][s3] = ln;
          mr[l][s3] = mun;
          gar[l][s3] = gan;
          gkr[l][s3] = gkn;
          qpr[l][s3] = qlo;
        } else if (u >= H && u < H + IZ) {  // committed row of (lam, mu)[M0-KL] and the accumulators
          if (colint) {
            bst(VW, vx, o_lo + so, ln);
            bst(VW, vx, o_mo + so, mun);
            bst(VW, vx, o_ga + so, gan);
            bst(VW, vx, o_gk + so, gkn);
          }
        }
        hq[l - 1][i % SPH] = bld(VW, vx, hb[l - 1]);   // this level's q for its row SPH steps ahead
        hb[l - 1] = next_row(hb[l - 1], rowb, he[l - 1], P4);
      });
      AK[i % SUNA][c] = make_float2(al0, k10);
      so += rowb;
      {  // request of window row t + SPB
        pfl[i % SPB] = bld(VW, vx, o_li + rq);
        pfm[i % SPB] = bld(VW, vx, o_mi + rq);
        pfa[i % SPB] = bld(VA, vx, rq);
        pfg[i % SPB] = bld(VW, vx, o_ga + rq);
        pfk[i % SPB] = bld(VW, vx, o_gk + rq);
        pfq[i % SPB] = bld(VW, vx, o_q1 + rq);
        pfz[i % SPB] = kzt[oz + t + SPB];             // oz + t + SPB in [-H, NZ + H + SUNA)
        rq = next_row(rq, rowb, P4, P4);
      }
      __syncthreads();
    };
    // One copy of the unrolled block: a wave-uniform bit per level and step
    // guards the source and receiver terms.  (Two copies of the block, picked
    // per block, make the register allocator keep both copies' values apart
    // across the loop: about 90 more VGPRs at 6 levels.)
    for (int tb = 0; tb < U; tb += SUNA) {
#pragma unroll
      for (int l = 1; l <= KL; ++l) {
        unsigned msk = 0;
        const int rows[4] = {su0, su1, ru0, ru1};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int d = rows[q] + 2 * l - tb;
          msk |= (d >= 0 && d < SUNA) ? 1u << d : 0u;
        }
        sm[l - 1] = msk;
      }
      static_for<0, SUNA>([&](auto ic) { body(ic, tb + decltype(ic)::value); });
    }
    if (src_here) a.wparam[te.w].ssum = ssum;
  }
}

// Zero `nslots` slots of region `reg` of every active wavefield (graph-safe
// fill: no hipMemset nodes are recorded in captured graphs).
__global__ __launch_bounds__(NTHR) void k_fill(const WaveDev* waves, const ModelDev* models,
                                                int n_waves, float* arena, int reg, int slot0,
                                                int nslots) {
  for (int wi = blockIdx.y; wi < n_waves; wi += gridDim.y) {
    const WaveDev& w = waves[wi];
    if (!models[w.model].active) continue;
    const int64_t P = (int64_t)w.NZ * w.NX;
    float* base = arena + w.off[reg] + slot0 * P;
    const int64_t n = P * nslots;
    for (int64_t i = (int64_t)blockIdx.x * NTHR + threadIdx.x; i < n; i += (int64_t)gridDim.x * NTHR)
      base[i] = 0.f;
  }
}

// Restore checkpoint j (state p[jK], q[jK]) into p slot 0 and q slot 0
// before the segment is recomputed.
__global__ __launch_bounds__(NTHR) void k_restore(const WaveDev* waves, const ModelDev* models,
                                                   int n_waves, float* arena, int j) {
  for (int wi = blockIdx.y; wi < n_waves; wi += gridDim.y) {
    const WaveDev& w = waves[wi];
    if (!models[w.model].active) continue;
    const int64_t P = (int64_t)w.NZ * w.NX;
    const float* ckp = arena + w.off[R_CKP] + j * P;
    const float* ckq = arena + w.off[R_CKQ] + j * P;
    float* s0 = arena + w.off[R_PP];
    float* q = arena + w.off[R_Q];
    for (int64_t i = (int64_t)blockIdx.x * NTHR + threadIdx.x; i < P; i += (int64_t)gridDim.x * NTHR) {
      s0[i] = ckp[i];
      q[i] = ckq[i];
    }
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
  // 1 - kappa of every padded row outside the left/right damping strips, with
  // a periodic apron (entry e is padded row e - kKzPad modulo NZ); kappa
  // scale and expression as in k_params and k_fwd
  const float kscale = (float)((double)v[m.nz * m.nx] * m.kunit);
  for (int e = blockIdx.x * NTHR + threadIdx.x; e < m.NZ + 2 * kKzPad; e += gridDim.x * NTHR) {
    int z = (e - kKzPad) % m.NZ;
    if (z < 0) z += m.NZ;
    const int j = damp_j(z, m.nbc, m.nz, m.nx, m.nbc);
    alpha[m.kz_off + e] = 1.0f - kscale * (float)(j * j);
  }
  for (int64_t i = (int64_t)blockIdx.x * NTHR + threadIdx.x; i < P; i += (int64_t)gridDim.x * NTHR) {
    const int z = (int)(i / m.NX);
    const int x = (int)(i - (int64_t)z * m.NX);
    const int iz = min(max(z - m.nbc, 0), m.nz - 1);
    const int ix = min(max(x - m.nbc, 0), m.nx - 1);
    const double vv = v[iz * m.nx + ix];
    alpha[m.cell_off + i] = (float)(vv * vv * dtdx2);
    // the tangent kernels take dalpha / alpha = 2 dv / v (prep_run_diff,
    // :352-355), the factor that turns the recovered alpha Lf(p) into dalpha Lf(p)
    if (dv) dalpha[m.cell_off + i] = vv != 0.0 ? (float)(2.0 * (double)dv[iz * m.nx + ix] / vv) : 0.f;
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

// Fixed-order tree sum over the 64 lanes of a wavefront (cross-lane moves,
// no LDS); lane 0 ends with the sum.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-wide sum of NTHR doubles: each wavefront reduces its lanes by
// shuffles, then the first wavefront reduces the NTHR/64 partials.  The
// order is fixed, so results are deterministic and independent of what
// else the launch processes.
__device__ __forceinline__ double block_sum(double v, double* red) {
  static_assert(NTHR % 64 == 0 && NTHR <= 64 * 64, "block_sum: whole wavefronts, one partial per lane");
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x < 64) {
    double w = (int)threadIdx.x < NTHR / 64 ? red[threadIdx.x] : 0.0;
    w = wave_sum(w);
    if (threadIdx.x == 0) red[NTHR / 64] = w;
  }
  __syncthreads();
  const double r = red[NTHR / 64];
  __syncthreads();
  return r;
}

// First half of prep_run_adjoint (:361-377) + unpad_edge_padded_gradient
// (seis_numerics.py:14-79): for padded row Z, fold the columns that edge
// padding maps onto each physical column (left strip -> column 0, right strip
// -> column nx-1), summing the alpha accumulator over shots in float64.  Also
// the row's share of sum(j^2 * kappa accumulator) for d/d min_vel (:366-370).
__global__ __launch_bounds__(NTHR) void k_fold_rows(const ModelDev* models, const WaveDev* waves,
                                                     const float* arena, const float* alpha, double* fold) {
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
    // the adjoint accumulated lam * alpha Lf(p): the alpha adjoint is that over alpha
    const double al = (double)alpha[m.cell_off + g];
    a = al != 0.0 ? a / al : 0.0;
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

// ---------------------------------------------------------------- LBFGS direction
constexpr int kLbfgsThreads = 1024;
constexpr int kLbfgsChunk = 16;       // requests per launch (kernel-argument descriptors)
constexpr int kLbfgsMaxPer = 32;      // entries per thread: n <= 32768

struct LbfgsBatch {
  seis_lbfgs_req r[kLbfgsChunk];
};

// One LBFGS direction per workgroup (bfgs, core/seis_numerics.py:350-392):
// the history update (y = g - g_prev, s = t d_prev, kept when y.s > 1e-10,
// H_diag = y.s / y.y) and the two-loop recursion of the lbfgs_backward /
// lbfgs_forward kernels (:450-565), newest pair first, then oldest first.
// The pairs live in a ring of cap + 1 rows: the new pair is written into the
// free row and joins the ring when accepted; a full ring (history_size
// reached) drops its oldest pair.  Each thread keeps its E entries of the
// vector q / r in registers and only reads back history entries it wrote or
// that earlier launches wrote, so the one cross-thread exchange is the
// fixed-order block sum: results do not depend on what else the launch
// carries.  Also returns g.d, max|d|, sum|g| and max|g| for the line search
// and stopping tests.
template <int E>
__global__ __launch_bounds__(kLbfgsThreads) void k_lbfgs(LbfgsBatch b) {
  constexpr int NWV = kLbfgsThreads / 64;
  __shared__ double red[2][NWV];
  const seis_lbfgs_req& q = b.r[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int phase = 0;
  // Block reduction with alternating partial buffers: one barrier each (a
  // buffer is rewritten two reductions later, after every thread has passed
  // the barrier of the reduction in between, i.e. finished reading it).
  // max that propagates NaN, as torch's max does
  auto nmax = [](double a, double b) { return (a > b || a != a) ? a : b; };
  auto reduce = [&](double v, bool is_max) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off, 64);
      v = is_max ? nmax(v, o) : v + o;
    }
    double* r = red[phase];
    phase ^= 1;
    if (lane == 0) r[wv] = v;
    __syncthreads();
    double acc = r[0];
    for (int w = 1; w < NWV; ++w) acc = is_max ? nmax(acc, r[w]) : acc + r[w];
    return acc;
  };
  const int n = q.n, R = q.cap + 1;
  double x[E];
  int head = 0, cnt = 0;
  double hd = 1.0;
  if (q.first) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + e * kLbfgsThreads;
      x[e] = i < n ? -q.g[i] : 0.0;
    }
  } else {
    head = q.ring[0];
    cnt = q.ring[1];
    hd = q.hdiag[0];
    const int fr = (head + cnt) % R;   // the free row
    double* Sf = q.S + (int64_t)fr * n;
    double* Yf = q.Y + (int64_t)fr * n;
    double ys = 0.0, yy = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + e * kLbfgsThreads;
      if (i < n) {
        const double y = q.g[i] - q.g_prev[i];
        const double sv = q.d_prev[i] * q.t;
        Yf[i] = y;
        Sf[i] = sv;
        ys += y * sv;
        yy += y * y;
      }
    }
    ys = reduce(ys, false);
    yy = reduce(yy, false);
    double rho_new = 0.0;
    const bool accept = ys > 1e-10;
    if (accept) {
      rho_new = 1.0 / ys;
      if (tid == 0) q.rho[fr] = rho_new;
      if (cnt == q.cap) head = (head + 1) % R;
      else ++cnt;
      hd = ys / yy;
    }
    // rho of the ring's entry i: the new pair's from the registers (its
    // stored copy is for later launches)
    auto rho_of = [&](int i, int slot) { return (accept && i == cnt - 1) ? rho_new : q.rho[slot]; };
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + e * kLbfgsThreads;
      x[e] = i < n ? -q.g[i] : 0.0;
    }
    for (int i = cnt - 1; i >= 0; --i) {          // lbfgs_backward (:451-501)
      const int slot = (head + i) % R;
      const double* Si = q.S + (int64_t)slot * n;
      const double* Yi = q.Y + (int64_t)slot * n;
      double p = 0.0;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int j = tid + e * kLbfgsThreads;
        if (j < n) p += Si[j] * x[e];
      }
      const double a = reduce(p, false) * rho_of(i, slot);
      if (tid == 0) q.al[slot] = a;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int j = tid + e * kLbfgsThreads;
        if (j < n) x[e] -= a * Yi[j];
      }
    }
    __threadfence_block();
    __syncthreads();                               // al[] written by thread 0 is read by all below
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] *= hd;
    for (int i = 0; i < cnt; ++i) {                // lbfgs_forward (:511-559)
      const int slot = (head + i) % R;
      const double* Si = q.S + (int64_t)slot * n;
      const double* Yi = q.Y + (int64_t)slot * n;
      double p = 0.0;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int j = tid + e * kLbfgsThreads;
        if (j < n) p += Yi[j] * x[e];
      }
      const double be = reduce(p, false) * rho_of(i, slot);
      const double c = q.al[slot] - be;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int j = tid + e * kLbfgsThreads;
        if (j < n) x[e] += c * Si[j];
      }
    }
  }
  double gd = 0.0, dmax = 0.0, gsum = 0.0, gmax = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = tid + e * kLbfgsThreads;
    if (i < n) {
      const double gv = q.g[i];
      q.d[i] = x[e];
      gd += gv * x[e];
      dmax = nmax(dmax, fabs(x[e]));
      gsum += fabs(gv);
      gmax = nmax(gmax, fabs(gv));
    }
  }
  gd = reduce(gd, false);
  dmax = reduce(dmax, true);
  gsum = reduce(gsum, false);
  gmax = reduce(gmax, true);
  if (tid == 0) {
    q.scal[0] = gd;
    q.scal[1] = dmax;
    q.scal[2] = gsum;
    q.scal[3] = gmax;
    q.ring[0] = head;
    q.ring[1] = cnt;
    q.hdiag[0] = hd;
  }
}

// ---------------------------------------------------------------- flat-area labelling
constexpr int kLabelMax = 24576;    // most cells of one labelled grid (LDS: 5 bytes per cell)
constexpr int kLabelThreads = 1024;

// Connected components of one velocity grid per workgroup, where 4-neighbours
// are joined when |v[n] - v[p]| <= thr (label_thresholded_components,
// core/seis_numerics.py:81-124, used by RestrictFlatAreas,
// core/seis_prior.py:256-262).  Edges are evaluated once in float64 (bit 0:
// right neighbour, bit 1: lower neighbour; the test is symmetric).  Every
// cell starts with its own index as label and repeatedly takes the minimum
// over its joined neighbours and over its label's label (pointer jumping);
// labels only decrease and stay indices of cells of the same component, so
// the sweep ends (no change in a full pass) with every cell holding its
// component's smallest index.  Roots are then numbered in index order by a
// workgroup prefix sum: the numbering of scipy's connected_components on
// the same graph (components in order of their smallest node).
__global__ __launch_bounds__(kLabelThreads) void k_label(const int* dims, const int64_t* offs, const double* vel,
                                                          double thr, int* labels, int* counts) {
  __shared__ int L[kLabelMax];
  __shared__ unsigned char E[kLabelMax];
  __shared__ int scan[kLabelThreads];
  __shared__ int changed;
  const int m = blockIdx.x, t = threadIdx.x;
  const int H = dims[2 * m], W = dims[2 * m + 1], N = H * W;
  const double* A = vel + offs[m];
  for (int i = t; i < N; i += kLabelThreads) {
    const int r = i / W, c = i - r * W;
    const double a = A[i];
    unsigned char e = 0;
    if (c + 1 < W && fabs(A[i + 1] - a) <= thr) e |= 1;
    if (r + 1 < H && fabs(A[i + W] - a) <= thr) e |= 2;
    E[i] = e;
    L[i] = i;
  }
  __syncthreads();
  for (;;) {
    if (t == 0) changed = 0;
    __syncthreads();
    bool mine = false;
    for (int i = t; i < N; i += kLabelThreads) {
      const int r = i / W, c = i - r * W;
      const unsigned char e = E[i];
      int v = L[i];
      if (e & 1) v = min(v, L[i + 1]);
      if (e & 2) v = min(v, L[i + W]);
      if (c > 0 && (E[i - 1] & 1)) v = min(v, L[i - 1]);
      if (r > 0 && (E[i - W] & 2)) v = min(v, L[i - W]);
      v = min(v, L[v]);
      if (v < L[i]) {
        L[i] = v;
        mine = true;
      }
    }
    if (mine) changed = 1;
    __syncthreads();
    const bool again = changed != 0;
    __syncthreads();   // every thread has read `changed` before it is reset
    if (!again) break;
  }
  // number the roots (L[i] == i) in index order: contiguous chunks per thread
  const int C = (N + kLabelThreads - 1) / kLabelThreads;
  const int b0 = min(N, t * C), b1 = min(N, b0 + C);
  int cnt = 0;
  for (int i = b0; i < b1; ++i) cnt += L[i] == i;
  scan[t] = cnt;
  __syncthreads();
  for (int d = 1; d < kLabelThreads; d <<= 1) {  // inclusive Hillis-Steele scan
    const int add = t >= d ? scan[t - d] : 0;
    __syncthreads();
    scan[t] += add;
    __syncthreads();
  }
  int rank = scan[t] - cnt;
  for (int i = b0; i < b1; ++i)
    if (L[i] == i) L[i] = -(rank++) - 1;   // a root holds -(its number) - 1
  __syncthreads();
  int* out = labels + offs[m];
  for (int i = t; i < N; i += kLabelThreads) {
    const int v = L[i];
    out[i] = v < 0 ? -v - 1 : -L[v] - 1;
  }
  if (t == 0) counts[m] = scan[kLabelThreads - 1];
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

__global__ __launch_bounds__(NTHR) void k_copy_active(const ModelDev* models, const float* src,
                                                       float* dst) {
  const ModelDev& m = models[blockIdx.y];
  if (!m.active) return;
  for (int64_t i = (int64_t)blockIdx.x * NTHR + threadIdx.x; i < m.seis_len;
       i += (int64_t)gridDim.x * NTHR)
    dst[m.seis_off + i] = src[m.seis_off + i];
}

// Activity mask for k_compact, passed by value: one bit per model.
constexpr int kMaskWords = 768;   // 24576 models
struct CompactArgs {
  const TileEnt* all[kLists];
  TileEnt* tmp[kLists];
  TileEnt* out[kLists];
  int* count[kLists];
  int n_all[kLists];
  uint32_t bits[kMaskWords];
};

// Stream-ordered seis_set_active: workgroup k compacts kind k's work list to
// the tiles of active models (the sequence order of make_tiles, then its XCD
// interleave: the block on XCD e % 8 takes a contiguous eighth), and
// workgroup 0 also writes the models' activity flags.  Replaces the host
// rebuild and upload, so no device synchronisation is needed.
__global__ __launch_bounds__(1024) void k_compact(CompactArgs a, const WaveDev* waves, ModelDev* models,
                                                  int n_models) {
  __shared__ int scan[1024];
  __shared__ int base;
  const int k = blockIdx.x, tid = threadIdx.x;
  auto active = [&](int m) { return (a.bits[m >> 5] >> (m & 31)) & 1u; };
  if (k == 0)
    for (int i = tid; i < n_models; i += 1024) models[i].active = (int)active(i);
  const int n = a.n_all[k];
  if (tid == 0) base = 0;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += 1024) {
    const int i = c0 + tid;
    TileEnt te{-1, 0, 0, 0};
    int f = 0;
    if (i < n) {
      te = a.all[k][i];
      f = (int)active(waves[te.w].model);
    }
    scan[tid] = f;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {   // inclusive Hillis-Steele scan
      const int add = tid >= d ? scan[tid - d] : 0;
      __syncthreads();
      scan[tid] += add;
      __syncthreads();
    }
    if (f) a.tmp[k][base + scan[tid] - 1] = te;
    __syncthreads();
    if (tid == 1023) base += scan[1023];
    __syncthreads();
  }
  const int total = base;
  const int per = (total + NXCD - 1) / NXCD;
  for (int e = tid; e < per * NXCD; e += 1024) {
    const int lg = (e % NXCD) * per + e / NXCD;
    a.out[k][e] = lg < total ? a.tmp[k][lg] : TileEnt{-1, 0, 0, 0};
  }
  if (tid == 0) *a.count[k] = per * NXCD;
}

int64_t align64(int64_t n) { return (n + 63) & ~int64_t(63); }

// Kernel launch configuration: LDS window (rx x rz cells, halo included),
// time steps per launch k and row ownership (strips: 1 = consecutive rows per
// wavefront, 0 = round-robin rows).  The interior of a window is
// (rz-4k) x (rx-4k).
struct KCfg {
  int rx, rz, k, strips;
};

// Instantiated windows per kernel kind: rx x rz cells on `nt` threads.  The
// forward keeps 8-16 cells per thread; the fused tangent and the adjoint carry
// about twice the per-cell state and run 4 cells per thread (5-6 wavefronts
// per SIMD), plus one 8-cell adjoint window (128x64) whose larger interior
// outweighs its lower occupancy; the Born kernel has both.
struct Shape {
  int rx, rz, nt;
};
const Shape kShapes[SEIS_KIND_COUNT][7] = {
    {{64, 32, 256}, {128, 32, 512}, {64, 64, 512}, {128, 64, 1024}, {256, 32, 1024}, {128, 96, 1024},
     {128, 128, 1024}},
    {{64, 32, 512}, {128, 32, 1024}, {64, 64, 1024}, {0, 0, 0}, {0, 0, 0}},
    {{64, 32, 512}, {128, 32, 1024}, {64, 64, 1024}, {128, 64, 1024}, {0, 0, 0}},
    {{64, 32, 512}, {128, 32, 1024}, {64, 64, 1024}, {128, 64, 1024}, {256, 32, 1024}},
    {{64, 32, 256}, {128, 32, 512}, {64, 64, 512}, {128, 64, 1024}, {256, 32, 1024}, {128, 96, 1024},
     {128, 128, 1024}}};

// Streamed forward (strips = 2): rx-column strips, one thread per column, k
// levels per launch, rz rows per segment (0: the whole column); the plain
// forward also at 5 levels, the saving forward and fused tangent at 3 or 4.
constexpr int kStreamNT[4] = {64, 128, 192, 256};
constexpr int kStreamKMin = 3, kStreamKMax = 4, kStreamFwdKMax = 5;

// Streamed Born (strips = 2): the same strips, 3 or 4 levels per launch.
constexpr int kStreamBornKMin = 3, kStreamBornKMax = 4;

// Streamed adjoint (strips = 2): 64- or 128-column strips, 4 or 6 levels.
bool stream_adj_ok(const KCfg& c) { return (c.rx == 64 || c.rx == 128) && (c.k == 4 || c.k == 6); }

bool stream_cfg(int kind, const KCfg& c) {
  if (c.strips != 2) return false;
  bool nt_ok = false;
  for (int nt : kStreamNT) nt_ok = nt_ok || c.rx == nt;
  if (!nt_ok || !(c.rz == 0 || c.rz >= 8)) return false;
  if (kind == SEIS_KIND_FORWARD) return c.k >= kStreamKMin && c.k <= kStreamFwdKMax;
  // saving forward and fused forward + tangent: 128- or 192-column strips
  if (kind == SEIS_KIND_RECORD || kind == SEIS_KIND_TANGENT)
    return (c.rx == 128 || c.rx == 192) && c.k >= kStreamKMin && c.k <= kStreamKMax;
  if (kind == SEIS_KIND_BORN) return c.k >= kStreamBornKMin && c.k <= kStreamBornKMax;
  if (kind == SEIS_KIND_ADJOINT) return stream_adj_ok(c);
  return false;
}

int shape_threads(int kind, int rx, int rz) {
  for (const Shape& sh : kShapes[kind])
    if (sh.rx == rx && sh.rz == rz && sh.nt) return sh.nt;
  return 0;
}

bool cfg_ok(int kind, const KCfg& c) {
  if (c.strips == 2) return stream_cfg(kind, c);
  return shape_threads(kind, c.rx, c.rz) && c.k >= 1 && c.k <= 8 && c.rz - 4 * c.k >= 4 &&
         c.rx - 4 * c.k >= 8 && (c.strips == 0 || c.strips == 1);
}

const KCfg kDefaultCfg[SEIS_KIND_COUNT] = {
    {64, 32, 3, 1}, {64, 32, 3, 0}, {64, 32, 3, 0}, {64, 32, 2, 1}, {64, 32, 3, 1}};
// Window of the tail list: 64x64 is a shape of every kind (forward and record
// on 512 threads, the others on 1024), round-robin rows, 3 steps per launch:
// the config the tuner picks for tiled fused-tangent passes, and close to the
// single-model picks of the others, since the list also runs whole passes of
// streamed kinds when few models are active.
const KCfg kTailCfg = {64, 64, 3, 0};

struct Layout {
  std::vector<ModelDev> models;
  std::vector<WaveDev> waves;
  std::vector<int64_t> vec_off, seis_off;
  int64_t cells = 0, arena = 0, seis = 0, vec = 0, fold = 0;
  int Kc = 0, J = 0;          // checkpoint interval and count (checkpoint mode)
  bool full = true, fwd_hist = true;
  int max_P = 0, max_cells_phys = 0, max_NZ = 0;
  int64_t max_wave_bytes = 0;   // largest state region of one wavefield (all its slots)
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
  L.Kc = L.full ? 0 : ckpt_every;
  L.J = L.full ? 0 : (nt + L.Kc - 1) / L.Kc;
  L.fwd_hist = L.full && adj;
  for (int mi = 0; mi < n_models; ++mi) {
    const seis_model_desc& d = md[mi];
    const std::string who = "model " + std::to_string(mi) + ": ";
    if (d.nz < 2 || d.nx < 2) return fail(SEIS_EINVAL, who + "nz, nx must be >= 2");
    if (!(d.dx > 0.f)) return fail(SEIS_EINVAL, who + "dx must be positive");
    if (d.n_shots < 1 || d.n_shots > SEIS_MAX_SHOTS) return fail(SEIS_EINVAL, who + "n_shots out of range");
    if (d.src_z < 0 || d.src_z >= d.nz || d.rcv_z < 0 || d.rcv_z >= d.nz)
      return fail(SEIS_EINVAL, who + "source/receiver row outside grid");
    if (d.n_rcv < 1 || d.rcv_dx < 1 || d.rcv_x0 < 0 || d.rcv_x0 + (int64_t)(d.n_rcv - 1) * d.rcv_dx >= d.nx)
      return fail(SEIS_EINVAL, who + "receivers outside grid");
    for (int s = 0; s < d.n_shots; ++s)
      if (d.src_x[s] < 0 || d.src_x[s] >= d.nx) return fail(SEIS_EINVAL, who + "source column outside grid");
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
    if (P > INT32_MAX / 4) return fail(SEIS_EINVAL, who + "model too large");
    m.kz_off = L.cells + align64(P);
    L.max_P = std::max<int>(L.max_P, (int)P);
    L.max_NZ = std::max(L.max_NZ, m.NZ);
    L.max_cells_phys = std::max(L.max_cells_phys, d.nz * d.nx);
    L.vec_off.push_back(L.vec);
    L.seis_off.push_back(L.seis);
    L.vec += (int64_t)d.nz * d.nx + 1;
    m.fold_off = L.fold;
    L.fold += align64((int64_t)m.NZ * m.nx + m.NZ);
    L.cells += align64(P) + align64(m.NZ + 2 * kKzPad);
    L.seis += m.seis_len;
    for (int s = 0; s < d.n_shots; ++s) {
      WaveDev w{};
      w.model = mi; w.shot = s; w.NZ = m.NZ; w.NX = m.NX; w.tiles_x = 0;
      w.src_z = d.src_z; w.src_x = d.src_x[s];
      w.src_cell = (d.src_z + nbc) * m.NX + (d.src_x[s] + nbc);
      w.seis_off = m.seis_off + (int64_t)s * nt * d.n_rcv;
      auto region = [&](int r, int64_t slots) {
        w.off[r] = L.arena;
        L.arena += align64(P) * slots;
      };
      region(R_HIST, L.fwd_hist ? nt + 2 : (adj ? L.Kc + 2 : 0));
      region(R_PP, 2);
      region(R_Q, 2);
      region(R_DP, tan ? 2 : 0);
      region(R_DQ, tan ? 2 : 0);
      region(R_LAM, adj ? 2 : 0);
      region(R_MU, adj ? 2 : 0);
      region(R_GA, adj ? 1 : 0);
      region(R_GK, adj ? 1 : 0);
      region(R_CKP, adj && !L.full ? L.J : 0);
      region(R_CKQ, adj && !L.full ? L.J : 0);
      L.max_wave_bytes = std::max<int64_t>(L.max_wave_bytes, (L.arena - w.off[R_HIST]) * (int64_t)sizeof(float));
      L.waves.push_back(w);
    }
    L.models.push_back(m);
  }
  L.vec_off.push_back(L.vec);
  L.seis_off.push_back(L.seis);
  const int64_t nw = (int64_t)L.waves.size();
  L.bytes = (int64_t)sizeof(float) * (L.arena + L.cells * (tan ? 2 : 1) + L.seis * (2 + (tan ? 1 : 0)) + nt) +
            (int64_t)sizeof(double) * (adj ? L.fold : 0) + (int64_t)sizeof(double) * 32 * n_models +
            (int64_t)(sizeof(ModelDev) + sizeof(ParamDev)) * n_models +
            (int64_t)(sizeof(WaveDev) + sizeof(WaveParam)) * nw;
  return SEIS_OK;
}

// Work list of one kernel kind on the current active set.
struct TileList {
  TileEnt* d = nullptr;
  int* d_count = nullptr;
  int cap = 0;     // entries allocated (all models active)
  int grid = 0;    // persistent grid (multiple of 8), fixed while the config is
  TileEnt* d_all = nullptr;  // every model's tiles in sequence order (compaction input)
  TileEnt* d_tmp = nullptr;  // compacted sequence (before the XCD interleave)
  int n_all = 0;
  std::vector<int> per_model;  // tiles of each model
  int n_active = 0;            // tiles of the active models
};

}  // namespace

struct seis_plan {
  Layout L;
  int n_models = 0, n_waves = 0, nt = 0, caps = 0, num_cu = 0;
  ModelDev* d_models = nullptr;
  WaveDev* d_waves = nullptr;
  ParamDev* d_params = nullptr;
  WaveParam* d_wparam = nullptr;
  float* d_alpha = nullptr;
  float* d_dalpha = nullptr;
  float* d_wavelet = nullptr;
  float* d_arena = nullptr;
  float* d_seis = nullptr;
  float* d_dseis = nullptr;
  float* d_dd = nullptr;
  double* d_fold = nullptr;
  double* d_part = nullptr;
  int* d_mask = nullptr;      // activity mask staged for the device-side compaction
  KCfg cfg[kLists];
  TileList tl[kLists];
  hipStream_t cap = nullptr;
  // captured passes, by pass and by whether their streamed kinds ran on the
  // tail list (few active models): forward [tan][save][tail], adjoint
  // [injection][adjoint on tail + 2 * record on tail], Born [tail]
  hipGraphExec_t g_fwd[2][2][2] = {};
  hipGraphExec_t g_adj[3][4] = {};
  hipGraphExec_t g_born[2] = {};
  std::vector<char> hist_ok;  // per model: a saving forward has filled R_HIST
  int stream_min = -1;        // active tiles a streamed config needs (< 0: the CU count)
};

namespace {

void drop_graphs(seis_plan* p) {
  for (auto& a : p->g_fwd)
    for (auto& b : a)
      for (auto& g : b)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  for (auto& a : p->g_adj)
    for (auto& g : a)
      if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  for (auto& g : p->g_born)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

void plan_free(seis_plan* p) {
  drop_graphs(p);
  void* bufs[] = {p->d_models, p->d_waves, p->d_params, p->d_wparam, p->d_fold, p->d_part,
                  p->d_alpha, p->d_dalpha, p->d_wavelet, p->d_arena, p->d_seis, p->d_dseis, p->d_dd,
                  p->d_mask};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto& t : p->tl) {
    if (t.d) (void)hipFree(t.d);
    if (t.d_count) (void)hipFree(t.d_count);
    if (t.d_all) (void)hipFree(t.d_all);
    if (t.d_tmp) (void)hipFree(t.d_tmp);
  }
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

// ---- kernel dispatch over the instantiated window shapes
#define SEIS_FWD_SHAPES(X) X(64, 32, 256) X(128, 32, 512) X(64, 64, 512) X(128, 64, 1024) X(256, 32, 1024) \
  X(128, 96, 1024) X(128, 128, 1024)
#define SEIS_TAN_SHAPES(X) X(64, 32, 512) X(128, 32, 1024) X(64, 64, 1024)
#define SEIS_ADJ_SHAPES(X) X(64, 32, 512) X(128, 32, 1024) X(64, 64, 1024) X(128, 64, 1024)
#define SEIS_BORN_SHAPES(X) X(64, 32, 512) X(128, 32, 1024) X(64, 64, 1024) X(128, 64, 1024) X(256, 32, 1024)

// Instantiation of a step kernel for config c (row ownership, and for the
// forward whether it stores the history pairs).
template <int RX, int RZ, int NT, bool TAN, bool HIST>
const void* fwd_ptr(const KCfg& c) {
  return c.strips ? (const void*)k_fwd<RX, RZ, NT, TAN, true, HIST> : (const void*)k_fwd<RX, RZ, NT, TAN, false, HIST>;
}

template <int NT>
const void* fws_ptr_nt(int k) {
  return k == 3   ? (const void*)k_fws<NT, 3, false, false>
         : k == 5 ? (const void*)k_fws<NT, 5, false, false>
                  : (const void*)k_fws<NT, 4, false, false>;
}

// the streamed forward variants: plain (64-256 columns), saving forward
// (history) and fused forward + tangent (with or without history; 128 or 192)
template <int NT, bool TAN, bool HIST>
const void* fws_ptr_th(int k) {
  return k == 3 ? (const void*)k_fws<NT, 3, TAN, HIST> : (const void*)k_fws<NT, 4, TAN, HIST>;
}

template <bool TAN, bool HIST>
const void* fws_ptr_w(const KCfg& c) {
  return c.rx == 128 ? fws_ptr_th<128, TAN, HIST>(c.k) : fws_ptr_th<192, TAN, HIST>(c.k);
}

const void* fws_ptr(const KCfg& c, bool tan, bool hist) {
  if (tan) return hist ? fws_ptr_w<true, true>(c) : fws_ptr_w<true, false>(c);
  if (hist) return fws_ptr_w<false, true>(c);
  if (c.rx == 64) return fws_ptr_nt<64>(c.k);
  if (c.rx == 128) return fws_ptr_nt<128>(c.k);
  if (c.rx == 192) return fws_ptr_nt<192>(c.k);
  return fws_ptr_nt<256>(c.k);
}

template <int NT>
const void* bns_ptr_nt(int k) {
  return k == 3 ? (const void*)k_bns<NT, 3> : (const void*)k_bns<NT, 4>;
}

const void* bns_ptr(const KCfg& c) {
  if (c.rx == 64) return bns_ptr_nt<64>(c.k);
  if (c.rx == 128) return bns_ptr_nt<128>(c.k);
  if (c.rx == 192) return bns_ptr_nt<192>(c.k);
  return bns_ptr_nt<256>(c.k);
}

template <int NT>
const void* ads_ptr_nt(int k) {
  return k == 4 ? (const void*)k_ads<NT, 4> : (const void*)k_ads<NT, 6>;
}

const void* ads_ptr(const KCfg& c) { return c.rx == 64 ? ads_ptr_nt<64>(c.k) : ads_ptr_nt<128>(c.k); }

const void* kernel_ptr(int kind, const KCfg& c) {
  if (c.strips == 2)
    return kind == SEIS_KIND_BORN ? bns_ptr(c)
           : kind == SEIS_KIND_ADJOINT ? ads_ptr(c)
                                       : fws_ptr(c, kind == SEIS_KIND_TANGENT, kind == SEIS_KIND_RECORD);
#define PTR_F(RX, RZ, NT) if (c.rx == RX && c.rz == RZ) return fwd_ptr<RX, RZ, NT, false, false>(c);
#define PTR_R(RX, RZ, NT) if (c.rx == RX && c.rz == RZ) return fwd_ptr<RX, RZ, NT, false, true>(c);
#define PTR_T(RX, RZ, NT) if (c.rx == RX && c.rz == RZ) return fwd_ptr<RX, RZ, NT, true, false>(c);
#define PTR_A(RX, RZ, NT) \
  if (c.rx == RX && c.rz == RZ) return c.strips ? (const void*)k_adj<RX, RZ, NT, true> : (const void*)k_adj<RX, RZ, NT, false>;
#define PTR_B(RX, RZ, NT) \
  if (c.rx == RX && c.rz == RZ) return c.strips ? (const void*)k_born<RX, RZ, NT, true> : (const void*)k_born<RX, RZ, NT, false>;
  if (kind == SEIS_KIND_FORWARD) { SEIS_FWD_SHAPES(PTR_F) }
  else if (kind == SEIS_KIND_TANGENT) { SEIS_TAN_SHAPES(PTR_T) }
  else if (kind == SEIS_KIND_ADJOINT) { SEIS_ADJ_SHAPES(PTR_A) }
  else if (kind == SEIS_KIND_BORN) { SEIS_BORN_SHAPES(PTR_B) }
  else { SEIS_FWD_SHAPES(PTR_R) }
#undef PTR_F
#undef PTR_R
#undef PTR_T
#undef PTR_A
#undef PTR_B
  return nullptr;
}

template <int RX, int RZ, int NT, bool TAN>
void launch_fwd_t(const KCfg& c, int grid, hipStream_t s, const FwdArgs& a) {
  const bool hist = a.hist_slot0 >= 0;
  if (c.strips) {
    if (hist) hipLaunchKernelGGL((k_fwd<RX, RZ, NT, TAN, true, true>), dim3(grid), dim3(NT), 0, s, a);
    else hipLaunchKernelGGL((k_fwd<RX, RZ, NT, TAN, true, false>), dim3(grid), dim3(NT), 0, s, a);
  } else {
    if (hist) hipLaunchKernelGGL((k_fwd<RX, RZ, NT, TAN, false, true>), dim3(grid), dim3(NT), 0, s, a);
    else hipLaunchKernelGGL((k_fwd<RX, RZ, NT, TAN, false, false>), dim3(grid), dim3(NT), 0, s, a);
  }
}

template <int NT>
void launch_fws_t(const KCfg& c, int grid, hipStream_t s, const FwdArgs& a) {
  if (c.k == 3) hipLaunchKernelGGL((k_fws<NT, 3, false, false>), dim3(grid), dim3(NT), 0, s, a);
  else if (c.k == 5) hipLaunchKernelGGL((k_fws<NT, 5, false, false>), dim3(grid), dim3(NT), 0, s, a);
  else hipLaunchKernelGGL((k_fws<NT, 4, false, false>), dim3(grid), dim3(NT), 0, s, a);
}

template <int NT, bool TAN, bool HIST>
void launch_fws_th(const KCfg& c, int grid, hipStream_t s, const FwdArgs& a) {
  if (c.k == 3) hipLaunchKernelGGL((k_fws<NT, 3, TAN, HIST>), dim3(grid), dim3(NT), 0, s, a);
  else hipLaunchKernelGGL((k_fws<NT, 4, TAN, HIST>), dim3(grid), dim3(NT), 0, s, a);
}

template <bool TAN, bool HIST>
void launch_fws_w(const KCfg& c, int grid, hipStream_t s, const FwdArgs& a) {
  if (c.rx == 128) launch_fws_th<128, TAN, HIST>(c, grid, s, a);
  else launch_fws_th<192, TAN, HIST>(c, grid, s, a);
}

// One launch of the forward (with tan: the fused forward + tangent) on
// config c; a launch that stores the history (hist_slot0 >= 0) runs the
// history variant.  Streamed configs of the kinds that take them: plain
// forward 64-256 columns, history or tangent 128 or 192 (stream_cfg).
void launch_fwd(const KCfg& c, bool tan, int grid, hipStream_t s, const FwdArgs& a) {
  if (c.strips == 2) {
    const bool hist = a.hist_slot0 >= 0;
    if (tan) {
      if (hist) launch_fws_w<true, true>(c, grid, s, a);
      else launch_fws_w<true, false>(c, grid, s, a);
    } else if (hist) {
      launch_fws_w<false, true>(c, grid, s, a);
    } else if (c.rx == 64) {
      launch_fws_t<64>(c, grid, s, a);
    } else if (c.rx == 128) {
      launch_fws_t<128>(c, grid, s, a);
    } else if (c.rx == 192) {
      launch_fws_t<192>(c, grid, s, a);
    } else {
      launch_fws_t<256>(c, grid, s, a);
    }
    return;
  }
#define LF(RX, RZ, NT) if (c.rx == RX && c.rz == RZ) { launch_fwd_t<RX, RZ, NT, false>(c, grid, s, a); return; }
#define LT(RX, RZ, NT) if (c.rx == RX && c.rz == RZ) { launch_fwd_t<RX, RZ, NT, true>(c, grid, s, a); return; }
  if (tan) { SEIS_TAN_SHAPES(LT) }
  else { SEIS_FWD_SHAPES(LF) }
#undef LF
#undef LT
}

template <int RX, int RZ, int NT>
void launch_adj_t(const KCfg& c, int grid, hipStream_t s, const AdjArgs& a) {
  if (c.strips) hipLaunchKernelGGL((k_adj<RX, RZ, NT, true>), dim3(grid), dim3(NT), 0, s, a);
  else hipLaunchKernelGGL((k_adj<RX, RZ, NT, false>), dim3(grid), dim3(NT), 0, s, a);
}

template <int NT>
void launch_ads_t(const KCfg& c, int grid, hipStream_t s, const AdjArgs& a) {
  if (c.k == 4) hipLaunchKernelGGL((k_ads<NT, 4>), dim3(grid), dim3(NT), 0, s, a);
  else hipLaunchKernelGGL((k_ads<NT, 6>), dim3(grid), dim3(NT), 0, s, a);
}

void launch_adj(const KCfg& c, int grid, hipStream_t s, const AdjArgs& a) {
  if (c.strips == 2) {
    if (c.rx == 64) launch_ads_t<64>(c, grid, s, a);
    else launch_ads_t<128>(c, grid, s, a);
    return;
  }
#define LA(RX, RZ, NT) if (c.rx == RX && c.rz == RZ) { launch_adj_t<RX, RZ, NT>(c, grid, s, a); return; }
  SEIS_ADJ_SHAPES(LA)
#undef LA
}

template <int RX, int RZ, int NT>
void launch_born_t(const KCfg& c, int grid, hipStream_t s, const BornArgs& a) {
  if (c.strips) hipLaunchKernelGGL((k_born<RX, RZ, NT, true>), dim3(grid), dim3(NT), 0, s, a);
  else hipLaunchKernelGGL((k_born<RX, RZ, NT, false>), dim3(grid), dim3(NT), 0, s, a);
}

template <int NT>
void launch_bns_t(const KCfg& c, int grid, hipStream_t s, const BornArgs& a) {
  if (c.k == 3) hipLaunchKernelGGL((k_bns<NT, 3>), dim3(grid), dim3(NT), 0, s, a);
  else hipLaunchKernelGGL((k_bns<NT, 4>), dim3(grid), dim3(NT), 0, s, a);
}

void launch_born(const KCfg& c, int grid, hipStream_t s, const BornArgs& a) {
  if (c.strips == 2) {
    if (c.rx == 64) launch_bns_t<64>(c, grid, s, a);
    else if (c.rx == 128) launch_bns_t<128>(c, grid, s, a);
    else if (c.rx == 192) launch_bns_t<192>(c, grid, s, a);
    else launch_bns_t<256>(c, grid, s, a);
    return;
  }
#define LB(RX, RZ, NT) if (c.rx == RX && c.rz == RZ) { launch_born_t<RX, RZ, NT>(c, grid, s, a); return; }
  SEIS_BORN_SHAPES(LB)
#undef LB
}

// Tiles of the active wavefields for window config c, laid out so that the
// persistent block handling list entry e (on XCD e % 8) receives a contiguous
// eighth of the tile sequence: neighbouring tiles of a wavefield, which share
// halo rows and columns, stay in one XCD's L2.
std::vector<TileEnt> tile_sequence(const Layout& L, const KCfg& c, bool all_active) {
  std::vector<TileEnt> seq;
  const bool stream = c.strips == 2;
  const int IX = c.rx - 4 * c.k;
  for (int wi = 0; wi < (int)L.waves.size(); ++wi) {
    const WaveDev& w = L.waves[wi];
    if (!all_active && !L.models[w.model].active) continue;
    // streamed strips carry their segment's row count in `pad`
    const int IZ = stream ? (c.rz > 0 ? c.rz : w.NZ) : c.rz - 4 * c.k;
    for (int z0 = 0; z0 < w.NZ; z0 += IZ)
      for (int x0 = 0; x0 < w.NX; x0 += IX)
        seq.push_back(TileEnt{wi, z0, x0, stream ? std::min(IZ, w.NZ - z0) : 0});
  }
  return seq;
}

std::vector<TileEnt> make_tiles(const Layout& L, const KCfg& c, bool all_active) {
  const std::vector<TileEnt> seq = tile_sequence(L, c, all_active);
  const int n = (int)seq.size();
  const int per = (n + NXCD - 1) / NXCD;
  std::vector<TileEnt> out((size_t)per * NXCD, TileEnt{-1, 0, 0, 0});
  for (int e = 0; e < per * NXCD; ++e) {
    const int lg = (e % NXCD) * per + e / NXCD;
    if (lg < n) out[e] = seq[lg];
  }
  return out;
}

// Active tiles of every work list, from the host's activity flags.
void count_active(seis_plan* p) {
  for (TileList& t : p->tl) {
    int n = 0;
    for (int m = 0; m < p->n_models && m < (int)t.per_model.size(); ++m)
      if (p->L.models[m].active) n += t.per_model[m];
    t.n_active = n;
  }
}

// A streamed config needs at least one strip segment per CU to fill the
// chip: with fewer active (few models of a large plan active, as in serving
// or chained schedules) the pass runs on the tail list's tiled window.
bool use_tail(const seis_plan* p, int kind) {
  return p->cfg[kind].strips == 2 && p->tl[kind].n_active < (p->stream_min >= 0 ? p->stream_min : p->num_cu);
}

int upload_tiles(seis_plan* p, int kind) {
  TileList& t = p->tl[kind];
  const std::vector<TileEnt> v = make_tiles(p->L, p->cfg[kind], false);
  const int n = (int)v.size();
  if (n > t.cap) return fail(SEIS_EINVAL, "tile list overflow");
  if (n) HIP_TRY(hipMemcpy(t.d, v.data(), sizeof(TileEnt) * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(t.d_count, &n, sizeof(int), hipMemcpyHostToDevice));
  return SEIS_OK;
}

// (Re)build the work list of `kind` for its current config: capacity for the
// all-active case, persistent grid from the kernel's occupancy.
int setup_kind(seis_plan* p, int kind) {
  TileList& t = p->tl[kind];
  const int cap = (int)make_tiles(p->L, p->cfg[kind], true).size();
  if (cap > t.cap) {
    if (t.d) (void)hipFree(t.d);
    t.d = nullptr;
    t.cap = 0;
    int rc = dev_alloc(&t.d, cap);
    if (rc) return rc;
    t.cap = cap;
  }
  if (!t.d_count) {
    int rc = dev_alloc(&t.d_count, 1);
    if (rc) return rc;
  }
  const std::vector<TileEnt> all = tile_sequence(p->L, p->cfg[kind], true);
  if ((int)all.size() > t.n_all || !t.d_all) {
    if (t.d_all) (void)hipFree(t.d_all);
    if (t.d_tmp) (void)hipFree(t.d_tmp);
    t.d_all = t.d_tmp = nullptr;
    t.n_all = 0;
    int rc = dev_alloc(&t.d_all, std::max<int64_t>(1, (int64_t)all.size()));
    if (!rc) rc = dev_alloc(&t.d_tmp, std::max<int64_t>(1, (int64_t)all.size()));
    if (rc) return rc;
  }
  t.n_all = (int)all.size();
  if (t.n_all) HIP_TRY(hipMemcpy(t.d_all, all.data(), sizeof(TileEnt) * t.n_all, hipMemcpyHostToDevice));
  t.per_model.assign(p->n_models, 0);
  for (const TileEnt& te : all) ++t.per_model[p->L.waves[te.w].model];
  count_active(p);
  int per_cu = 0;
  // the tail list's grid: the occupancy of the adjoint kernel on its window
  // (the fewest blocks per CU of the kernels that run it)
  const int kk = kind == kTail ? (int)SEIS_KIND_ADJOINT : kind;
  const int nthr = p->cfg[kind].strips == 2 ? p->cfg[kind].rx : shape_threads(kk, p->cfg[kind].rx, p->cfg[kind].rz);
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_ptr(kk, p->cfg[kind]), nthr, 0));
  int gmax = std::max(1, per_cu) * p->num_cu;
  gmax = std::max(NXCD, gmax / NXCD * NXCD);
  t.grid = std::max(NXCD, std::min(cap, gmax));
  return upload_tiles(p, kind);
}

void fill(seis_plan* p, hipStream_t s, int reg, int slot0, int nslots) {
  hipLaunchKernelGGL(k_fill, dim3(64, std::min(p->n_waves, 256)), dim3(NTHR), 0, s, p->d_waves,
                     p->d_models, p->n_waves, p->d_arena, reg, slot0, nslots);
}

FwdArgs fwd_args(seis_plan* p, int kind) {
  FwdArgs a{};
  a.tiles = p->tl[kind].d; a.n_tiles = p->tl[kind].d_count;
  a.waves = p->d_waves; a.models = p->d_models; a.params = p->d_params; a.wparam = p->d_wparam;
  a.alpha = p->d_alpha; a.dalpha = p->d_dalpha; a.wavelet = p->d_wavelet;
  a.arena = p->d_arena; a.seis = p->d_seis; a.dseis = p->d_dseis;
  a.nt = p->nt;
  a.Kt = p->cfg[kind].k;
  a.ck_slot = -1;
  return a;
}

// Forward time loop from the zero state at n = 0 through n = nt (p[1] .. p[nt+1],
// core/seis_forward2.py:211-247), fused with the tangent when `tan`
// (:255-283).  A saving forward (`save`) on a full-history plan stores the
// pairs (Lf(p[n]), q[n]) of n = 0..nt (the RECORD kind's config); in
// checkpoint mode its launches stop at every checkpoint boundary n = j*Kc and
// store the state there.
void enqueue_forward(seis_plan* p, hipStream_t s, bool tan, bool save) {
  const Layout& L = p->L;
  const bool hist = save && L.fwd_hist;
  const int kind0 = tan ? SEIS_KIND_TANGENT : (hist ? SEIS_KIND_RECORD : SEIS_KIND_FORWARD);
  const int kind = use_tail(p, kind0) ? kTail : kind0;
  const KCfg& c = p->cfg[kind];
  const bool ck = save && !L.full && (p->caps & SEIS_CAP_ADJOINT);
  fill(p, s, R_Q, 0, 1);
  fill(p, s, R_PP, 0, 1);
  if (ck) {
    fill(p, s, R_CKP, 0, 1);
    fill(p, s, R_CKQ, 0, 1);
  }
  if (tan) {
    fill(p, s, R_DP, 0, 1);
    fill(p, s, R_DQ, 0, 1);
  }
  FwdArgs a = fwd_args(p, kind);
  a.write_seis = 1;
  // a streamed forward always advances c.k steps; shorter launches (the last
  // one, or up to a checkpoint) run k_fwd on the tail list
  const KCfg& cr = p->cfg[kTail];
  FwdArgs ar = fwd_args(p, kTail);
  int n = 0, li = 0;
  while (n < p->nt + 1) {
    int K = std::min(c.k, p->nt + 1 - n);
    if (ck) K = std::min(K, (n / L.Kc + 1) * L.Kc - n);
    const bool tail = c.strips == 2 && K < c.k;
    if (tail) K = std::min(K, cr.k);
    FwdArgs& f = tail ? ar : a;
    f.write_seis = 1;
    f.n0 = n;
    f.K = K;
    f.p_in_slot = f.q_in_slot = f.dp_in_slot = f.dq_in_slot = li & 1;
    f.p_out_slot = f.q_out_slot = f.dp_out_slot = f.dq_out_slot = (li + 1) & 1;
    f.hist_slot0 = hist ? n + 1 : -1;   // step n stores q[n+1] at slot n+1
    f.ck_slot = (ck && (n + K) % L.Kc == 0 && (n + K) / L.Kc < L.J) ? (n + K) / L.Kc : -1;
    if (tail) launch_fwd(cr, tan, p->tl[kTail].grid, s, f);
    else launch_fwd(c, tan, p->tl[kind].grid, s, f);
    n += K;
    ++li;
  }
}

// Tangent time loop n = 0..nt on the saved history of a full-history plan
// (same recursion and outputs as the fused pass of enqueue_forward).
void enqueue_born(seis_plan* p, hipStream_t s) {
  fill(p, s, R_DP, 0, 1);
  fill(p, s, R_DQ, 0, 1);
  const int kb = use_tail(p, SEIS_KIND_BORN) ? kTail : (int)SEIS_KIND_BORN;
  const KCfg& c = p->cfg[kb];
  BornArgs a{};
  a.tiles = p->tl[kb].d; a.n_tiles = p->tl[kb].d_count;
  a.waves = p->d_waves; a.models = p->d_models; a.params = p->d_params; a.wparam = p->d_wparam;
  a.alpha = p->d_alpha; a.dalpha = p->d_dalpha; a.wavelet = p->d_wavelet;
  a.arena = p->d_arena; a.dseis = p->d_dseis; a.nt = p->nt;
  a.Kt = c.k;
  // a streamed Born pass always advances c.k steps; the shorter last launch
  // runs k_born on the tail list
  const KCfg& ct = p->cfg[kTail];
  int n = 0, li = 0;
  while (n < p->nt + 1) {
    int K = std::min(c.k, p->nt + 1 - n);
    const bool tail = c.strips == 2 && K < c.k;
    if (tail) K = std::min(K, ct.k);
    a.n0 = n;
    a.K = K;
    a.dp_in_slot = a.dq_in_slot = li & 1;
    a.dp_out_slot = a.dq_out_slot = (li + 1) & 1;
    if (tail) {
      BornArgs at = a;
      at.tiles = p->tl[kTail].d;
      at.n_tiles = p->tl[kTail].d_count;
      at.Kt = ct.k;
      launch_born(ct, p->tl[kTail].grid, s, at);
    } else {
      launch_born(c, p->tl[kb].grid, s, a);
    }
    n += K;
    ++li;
  }
}

// Reverse time loop: lam[m] for m = nt+1 .. 2 (update_p_adjoint graph,
// core/seis_forward2.py:292-313), reading the pairs of n = m-1.  In checkpoint
// mode each segment is first recomputed from its checkpoint: steps
// n = j*Kc .. min(j*Kc + Kc, nt) store their pairs into the segment buffer.
void enqueue_adjoint(seis_plan* p, hipStream_t s, int kind) {
  const Layout& L = p->L;
  fill(p, s, R_LAM, 0, 1);
  fill(p, s, R_MU, 0, 1);
  fill(p, s, R_GA, 0, 1);
  fill(p, s, R_GK, 0, 1);
  const int ka = use_tail(p, SEIS_KIND_ADJOINT) ? kTail : (int)SEIS_KIND_ADJOINT;
  const KCfg& ca = p->cfg[ka];
  AdjArgs a{};
  a.tiles = p->tl[ka].d; a.n_tiles = p->tl[ka].d_count;
  a.waves = p->d_waves; a.models = p->d_models; a.params = p->d_params; a.wparam = p->d_wparam;
  a.alpha = p->d_alpha; a.wavelet = p->d_wavelet; a.arena = p->d_arena; a.nt = p->nt;
  a.Kt = ca.k;
  if (kind == 0) { a.inj_a = p->d_seis; a.inj_b = p->d_dd; }
  else if (kind == 1) { a.inj_a = p->d_dd; a.inj_b = nullptr; }
  else { a.inj_a = p->d_dseis; a.inj_b = nullptr; }
  const int grid = p->tl[ka].grid;
  // a streamed adjoint always advances ca.k steps; shorter launches run k_adj
  // on the tail list
  const KCfg& ct = p->cfg[kTail];
  AdjArgs at = a;
  at.tiles = p->tl[kTail].d;
  at.n_tiles = p->tl[kTail].d_count;
  at.Kt = ct.k;
  int li = 0;
  auto run_range = [&](int m_hi, int m_lo, int base) {
    int M0 = m_hi + 1;
    while (M0 - 1 >= m_lo) {
      int K = std::min(ca.k, M0 - m_lo);
      const bool tail = ca.strips == 2 && K < ca.k;
      if (tail) K = std::min(K, ct.k);
      AdjArgs& f = tail ? at : a;
      f.M0 = M0;
      f.K = K;
      f.lam_in_slot = f.mu_in_slot = li & 1;
      f.lam_out_slot = f.mu_out_slot = (li + 1) & 1;
      f.h_base = base;
      if (tail) launch_adj(ct, p->tl[kTail].grid, s, f);
      else launch_adj(ca, grid, s, f);
      M0 -= K;
      ++li;
    }
  };
  if (L.full) {
    run_range(p->nt + 1, 2, 0);
    return;
  }
  const int kr = use_tail(p, SEIS_KIND_RECORD) ? kTail : (int)SEIS_KIND_RECORD;
  const KCfg& cf = p->cfg[kr];
  FwdArgs f = fwd_args(p, kr);
  f.write_seis = 0;
  FwdArgs ft = fwd_args(p, kTail);   // launches shorter than a streamed record config's k
  ft.write_seis = 0;
  for (int j = L.J - 1; j >= 0; --j) {
    const int lo = j * L.Kc;
    const int hi = std::min(lo + L.Kc, p->nt);  // largest n whose pair the segment needs
    hipLaunchKernelGGL(k_restore, dim3(64, std::min(p->n_waves, 256)), dim3(NTHR), 0, s, p->d_waves,
                       p->d_models, p->n_waves, p->d_arena, j);
    int n = lo, fi = 0;
    while (n <= hi) {
      int K = std::min(cf.k, hi + 1 - n);
      const bool tail = cf.strips == 2 && K < cf.k;
      if (tail) K = std::min(K, p->cfg[kTail].k);
      FwdArgs& g = tail ? ft : f;
      g.n0 = n;
      g.K = K;
      g.p_in_slot = g.q_in_slot = fi & 1;
      g.p_out_slot = g.q_out_slot = (fi + 1) & 1;
      g.hist_slot0 = n + 1 - lo;        // q[n+1] at segment slot n+1-lo
      if (tail) launch_fwd(p->cfg[kTail], false, p->tl[kTail].grid, s, g);
      else launch_fwd(cf, false, p->tl[kr].grid, s, g);
      n += K;
      ++fi;
    }
    run_range(std::min(lo + L.Kc + 1, p->nt + 1), lo + 2, lo);
  }
}

// Record that the active models' R_HIST now holds a complete saved forward.
void mark_saved(seis_plan* p) {
  if (!p->L.fwd_hist) return;
  for (int i = 0; i < p->n_models; ++i)
    if (p->L.models[i].active) p->hist_ok[i] = 1;
}

int check_saved(seis_plan* p) {
  if ((p->caps & (SEIS_CAP_TANGENT | SEIS_CAP_ADJOINT)) != (SEIS_CAP_TANGENT | SEIS_CAP_ADJOINT))
    return fail(SEIS_ESTATE, "plan lacks SEIS_CAP_TANGENT | SEIS_CAP_ADJOINT");
  if (!p->L.fwd_hist) return fail(SEIS_ESTATE, "saved-history passes need a full-history plan (ckpt_every 0)");
  for (int i = 0; i < p->n_models; ++i)
    if (p->L.models[i].active && !p->hist_ok[i])
      return fail(SEIS_ESTATE, "model " + std::to_string(i) +
                                   " has no saved forward: run seis_gradient, seis_forward_save or "
                                   "seis_gn_product on it first");
  return SEIS_OK;
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
  hipLaunchKernelGGL(k_fold_rows, dim3(L.max_NZ, p->n_models), dim3(NTHR), 0, s, p->d_models,
                     p->d_waves, p->d_arena, p->d_alpha, p->d_fold);
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

#define TRY(expr)                   \
  do {                              \
    int rc_ = (expr);               \
    if (rc_ != SEIS_OK) return rc_; \
  } while (0)

// Launch a pass's captured graph (captured on first use), keyed by whether
// its streamed kinds run on the tail list for the current active set.
int launch_pass_forward(seis_plan* p, hipStream_t s, bool tan, bool save) {
  const int kind = tan ? SEIS_KIND_TANGENT : (save && p->L.fwd_hist ? SEIS_KIND_RECORD : SEIS_KIND_FORWARD);
  hipGraphExec_t* g = &p->g_fwd[tan][save][use_tail(p, kind)];
  TRY(get_graph(p, g, [&](hipStream_t c) { enqueue_forward(p, c, tan, save); }));
  HIP_TRY(hipGraphLaunch(*g, s));
  return SEIS_OK;
}

int launch_pass_adjoint(seis_plan* p, hipStream_t s, int kind) {
  hipGraphExec_t* g = &p->g_adj[kind][use_tail(p, SEIS_KIND_ADJOINT) + 2 * use_tail(p, SEIS_KIND_RECORD)];
  TRY(get_graph(p, g, [&](hipStream_t c) { enqueue_adjoint(p, c, kind); }));
  HIP_TRY(hipGraphLaunch(*g, s));
  return SEIS_OK;
}

int launch_pass_born(seis_plan* p, hipStream_t s) {
  hipGraphExec_t* g = &p->g_born[use_tail(p, SEIS_KIND_BORN)];
  TRY(get_graph(p, g, [&](hipStream_t c) { enqueue_born(p, c); }));
  HIP_TRY(hipGraphLaunch(*g, s));
  return SEIS_OK;
}


// Grid table of a labelling call: validated sizes, per-grid offsets into the
// concatenated cells (int64) and (nz, nx) pairs, staged on the device as
// offsets followed by sizes.
int label_args(const int32_t* nz, const int32_t* nx, int32_t n_models, const double* vel, double threshold,
               const int32_t* labels, const int32_t* counts, std::vector<int64_t>& offs, std::vector<int>& dims) {
  if (!nz || !nx || n_models < 1 || !vel || !labels || !counts) return fail(SEIS_EINVAL, "NULL argument or no grids");
  if (!(threshold >= 0.0)) return fail(SEIS_EINVAL, "threshold must be >= 0");
  dims.assign(2 * (size_t)n_models, 0);
  offs.assign(n_models, 0);
  int64_t off = 0;
  for (int m = 0; m < n_models; ++m) {
    if (nz[m] < 1 || nx[m] < 1 || (int64_t)nz[m] * nx[m] > kLabelMax)
      return fail(SEIS_EINVAL, "grid " + std::to_string(m) + ": need 1 <= nz*nx <= " + std::to_string(kLabelMax));
    dims[2 * m] = nz[m];
    dims[2 * m + 1] = nx[m];
    offs[m] = off;
    off += (int64_t)nz[m] * nx[m];
  }
  return SEIS_OK;
}

// Stage the grid table at device address d (8-byte aligned, room for
// 12 bytes per grid) on stream s and launch one labelling workgroup per grid.
int launch_label(char* d, const std::vector<int64_t>& offs, const std::vector<int>& dims, const double* vel,
                 double threshold, int32_t* labels, int32_t* counts, hipStream_t s) {
  const size_t bo = sizeof(int64_t) * offs.size(), bd = sizeof(int) * dims.size();
  HIP_TRY(hipMemcpyAsync(d, offs.data(), bo, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d + bo, dims.data(), bd, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_label, dim3((unsigned)offs.size()), dim3(kLabelThreads), 0, s,
                     reinterpret_cast<const int*>(d + bo), reinterpret_cast<const int64_t*>(d), vel, threshold,
                     labels, counts);
  HIP_TRY(hipGetLastError());
  return SEIS_OK;
}

int static_scratch(char** out);

}  // namespace

// ---------------------------------------------------------------- allocation-free labelling
// Device staging of seis_label_flat_static: the grid table of one call (12
// bytes per grid).  Each GPU holds its own copy of this array until the
// process ends, and every call on that GPU stages into it.  Host code never
// uses the array's name as a device address: static_scratch looks the
// current device's copy up once (hipGetSymbolAddress) and hands out that.
static __device__ __align__(256) char scratch[SEIS_LABEL_STATIC_MAX_GRIDS * (sizeof(int64_t) + 2 * sizeof(int))];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

int static_scratch(char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SEIS_EINVAL, "device index " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<char*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return SEIS_OK;
}

}  // namespace

extern "C" {

const char* seis_last_error(void) { return g_last_error.c_str(); }

int64_t seis_plan_bytes(const seis_model_desc* models, int32_t n_models, const seis_time_desc* time,
                        int32_t ckpt_every, int32_t caps) {
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
  p->hist_ok.assign(n_models, 0);
  for (int k = 0; k < SEIS_KIND_COUNT; ++k) p->cfg[k] = kDefaultCfg[k];
  p->cfg[kTail] = kTailCfg;
  const bool tan = caps & SEIS_CAP_TANGENT, adj = caps & SEIS_CAP_ADJOINT;
  auto bail = [&](int code) { plan_free(p); return code; };
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&p->num_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    fail(SEIS_EHIP, "cannot query the device");
    return bail(SEIS_EHIP);
  }
  if ((rc = dev_alloc(&p->d_models, n_models)) || (rc = dev_alloc(&p->d_waves, p->n_waves)) ||
      (rc = dev_alloc(&p->d_params, n_models)) || (rc = dev_alloc(&p->d_wparam, p->n_waves)) ||
      (rc = dev_alloc(&p->d_fold, adj ? L.fold : 0)) ||
      (rc = dev_alloc(&p->d_part, (int64_t)kRedBlocks * n_models)) ||
      (rc = dev_alloc(&p->d_alpha, L.cells)) || (rc = dev_alloc(&p->d_dalpha, tan ? L.cells : 0)) ||
      (rc = dev_alloc(&p->d_wavelet, p->nt)) || (rc = dev_alloc(&p->d_arena, L.arena)) ||
      (rc = dev_alloc(&p->d_seis, L.seis)) || (rc = dev_alloc(&p->d_dseis, tan ? L.seis : 0)) ||
      (rc = dev_alloc(&p->d_dd, L.seis)))
    return bail(rc);
  std::vector<float> wl(p->nt);
  for (int i = 0; i < p->nt; ++i) wl[i] = (float)time->wavelet[i];
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMemcpy(p->d_models, L.models.data(), sizeof(ModelDev) * n_models, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_waves, L.waves.data(), sizeof(WaveDev) * p->n_waves, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_wavelet, wl.data(), sizeof(float) * p->nt, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(p->d_wparam, 0, sizeof(WaveParam) * p->n_waves);
  if (e == hipSuccess) e = hipMemset(p->d_params, 0, sizeof(ParamDev) * n_models);
  // full history: slot 0 is q[0], the zero initial state, which no pass writes
  if (L.fwd_hist)
    for (const WaveDev& w : L.waves)
      if (e == hipSuccess) e = hipMemset(p->d_arena + w.off[R_HIST], 0, sizeof(float) * (size_t)w.NZ * w.NX);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    fail(SEIS_EHIP, std::string("plan setup: ") + hipGetErrorString(e));
    return bail(SEIS_EHIP);
  }
  for (int k = 0; k < kLists; ++k)
    if ((rc = setup_kind(p, k))) return bail(rc);
  *plan_out = p;
  return SEIS_OK;
}

int seis_plan_destroy(seis_plan_t* plan) {
  if (!plan) return SEIS_OK;
  (void)hipDeviceSynchronize();
  plan_free(plan);
  return SEIS_OK;
}

int seis_lbfgs_direction(const seis_lbfgs_req* reqs, int32_t n_req, seis_stream_t stream) {
  if (n_req < 0 || (n_req > 0 && !reqs)) return fail(SEIS_EINVAL, "bad request list");
  for (int i = 0; i < n_req; ++i) {
    const seis_lbfgs_req& r = reqs[i];
    const std::string who = "request " + std::to_string(i) + ": ";
    if (r.n < 1 || r.n > kLbfgsThreads * kLbfgsMaxPer)
      return fail(SEIS_EINVAL, who + "n must be in 1.." + std::to_string(kLbfgsThreads * kLbfgsMaxPer));
    if (r.cap < 1) return fail(SEIS_EINVAL, who + "cap must be >= 1");
    if (!r.g || !r.d || !r.scal || !r.ring || !r.hdiag)
      return fail(SEIS_EINVAL, who + "NULL argument");
    if (!r.first && (!r.S || !r.Y || !r.rho || !r.al || !r.g_prev || !r.d_prev))
      return fail(SEIS_EINVAL, who + "NULL history argument");
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (int c0 = 0; c0 < n_req; c0 += kLbfgsChunk) {
    const int nc = std::min(kLbfgsChunk, n_req - c0);
    LbfgsBatch b{};
    int nmax = 1;
    for (int i = 0; i < nc; ++i) {
      b.r[i] = reqs[c0 + i];
      nmax = std::max(nmax, (int)reqs[c0 + i].n);
    }
    const int per = (nmax + kLbfgsThreads - 1) / kLbfgsThreads;
    const dim3 grid(nc), block(kLbfgsThreads);
    if (per <= 1) hipLaunchKernelGGL(k_lbfgs<1>, grid, block, 0, s, b);
    else if (per <= 2) hipLaunchKernelGGL(k_lbfgs<2>, grid, block, 0, s, b);
    else if (per <= 4) hipLaunchKernelGGL(k_lbfgs<4>, grid, block, 0, s, b);
    else if (per <= 8) hipLaunchKernelGGL(k_lbfgs<8>, grid, block, 0, s, b);
    else if (per <= 16) hipLaunchKernelGGL(k_lbfgs<16>, grid, block, 0, s, b);
    else hipLaunchKernelGGL(k_lbfgs<32>, grid, block, 0, s, b);
    HIP_TRY(hipGetLastError());
  }
  return SEIS_OK;
}

int seis_label_flat(const int32_t* nz, const int32_t* nx, int32_t n_models, const double* vel, double threshold,
                    int32_t* labels, int32_t* counts, seis_stream_t stream) {
  std::vector<int64_t> offs;
  std::vector<int> dims;
  TRY(label_args(nz, nx, n_models, vel, threshold, labels, counts, offs, dims));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const size_t bo = sizeof(int64_t) * offs.size(), bd = sizeof(int) * dims.size();
  char* d = nullptr;
  HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d), bo + bd, s));
  TRY(launch_label(d, offs, dims, vel, threshold, labels, counts, s));
  HIP_TRY(hipFreeAsync(d, s));
  return SEIS_OK;
}

int seis_label_flat_static(const int32_t* nz, const int32_t* nx, int32_t n_models, const double* vel,
                           double threshold, int32_t* labels, int32_t* counts, seis_stream_t stream) {
  if (n_models > SEIS_LABEL_STATIC_MAX_GRIDS)
    return fail(SEIS_EINVAL, "seis_label_flat_static takes at most " + std::to_string(SEIS_LABEL_STATIC_MAX_GRIDS) +
                                 " grids per call, got " + std::to_string(n_models));
  std::vector<int64_t> offs;
  std::vector<int> dims;
  TRY(label_args(nz, nx, n_models, vel, threshold, labels, counts, offs, dims));
  char* d = nullptr;
  TRY(static_scratch(&d));
  return launch_label(d, offs, dims, vel, threshold, labels, counts, reinterpret_cast<hipStream_t>(stream));
}

int64_t seis_plan_device_bytes(const seis_plan_t* plan) { return plan ? plan->L.bytes : 0; }

int seis_plan_layout(const seis_plan_t* plan, int64_t* vec_off, int64_t* seis_off) {
  if (!plan) return fail(SEIS_EINVAL, "plan is NULL");
  if (vec_off) std::memcpy(vec_off, plan->L.vec_off.data(), sizeof(int64_t) * (plan->n_models + 1));
  if (seis_off) std::memcpy(seis_off, plan->L.seis_off.data(), sizeof(int64_t) * (plan->n_models + 1));
  return SEIS_OK;
}

int seis_plan_set_config(seis_plan_t* p, int32_t kind, int32_t rx, int32_t rz, int32_t k, int32_t strips) {
  if (!p || kind < 0 || kind >= SEIS_KIND_COUNT) return fail(SEIS_EINVAL, "bad plan or kind");
  const KCfg c{rx, rz, k, strips};
  if (!cfg_ok(kind, c))
    return fail(SEIS_EINVAL, "unsupported config: window " + std::to_string(rx) + "x" + std::to_string(rz) +
                                 ", k=" + std::to_string(k) + ", strips=" + std::to_string(strips));
  if (c.strips == 2) {
    for (const WaveDev& w : p->L.waves)
      if (w.NZ < 4 * c.k) return fail(SEIS_EINVAL, "streamed kernels need padded grids of at least 4k rows");
    // the streamed kernels address a wavefield's state with 32-bit byte offsets
    if (p->L.max_wave_bytes >= (int64_t)UINT32_MAX)
      return fail(SEIS_EINVAL, "streamed kernels need each wavefield's state region below 4 GiB");
  }
  HIP_TRY(hipDeviceSynchronize());
  drop_graphs(p);
  p->cfg[kind] = c;
  return setup_kind(p, kind);
}

int seis_plan_get_config(const seis_plan_t* p, int32_t kind, int32_t* rx, int32_t* rz, int32_t* k,
                         int32_t* strips) {
  if (!p || kind < 0 || kind >= SEIS_KIND_COUNT || !rx || !rz || !k || !strips)
    return fail(SEIS_EINVAL, "bad argument");
  *rx = p->cfg[kind].rx;
  *rz = p->cfg[kind].rz;
  *k = p->cfg[kind].k;
  *strips = p->cfg[kind].strips;
  return SEIS_OK;
}

int seis_plan_set_stream_threshold(seis_plan_t* p, int32_t min_active_tiles) {
  if (!p) return fail(SEIS_EINVAL, "plan is NULL");
  p->stream_min = min_active_tiles < 0 ? -1 : min_active_tiles;
  return SEIS_OK;
}

int seis_set_active(seis_plan_t* p, const int32_t* active) {
  if (!p || !active) return fail(SEIS_EINVAL, "NULL argument");
  HIP_TRY(hipDeviceSynchronize());
  for (int i = 0; i < p->n_models; ++i) p->L.models[i].active = active[i] ? 1 : 0;
  HIP_TRY(hipMemcpy(p->d_models, p->L.models.data(), sizeof(ModelDev) * p->n_models,
                    hipMemcpyHostToDevice));
  for (int k = 0; k < kLists; ++k) TRY(upload_tiles(p, k));
  count_active(p);
  return SEIS_OK;
}

int seis_set_active_async(seis_plan_t* p, const int32_t* active, seis_stream_t stream) {
  if (!p || !active) return fail(SEIS_EINVAL, "NULL argument");
  if (p->n_models > kMaskWords * 32) {   // mask too large for kernel arguments
    HIP_TRY(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
    return seis_set_active(p, active);
  }
  CompactArgs a{};
  for (int i = 0; i < p->n_models; ++i) {
    p->L.models[i].active = active[i] ? 1 : 0;
    if (active[i]) a.bits[i >> 5] |= 1u << (i & 31);
  }
  count_active(p);
  for (int k = 0; k < kLists; ++k) {
    const TileList& t = p->tl[k];
    a.all[k] = t.d_all;
    a.tmp[k] = t.d_tmp;
    a.out[k] = t.d;
    a.count[k] = t.d_count;
    a.n_all[k] = t.n_all;
  }
  hipLaunchKernelGGL(k_compact, dim3(kLists), dim3(1024), 0, reinterpret_cast<hipStream_t>(stream), a,
                     p->d_waves, p->d_models, p->n_models);
  HIP_TRY(hipGetLastError());
  return SEIS_OK;
}

int seis_forward(seis_plan_t* p, const float* vec, float* seis, seis_stream_t stream) {
  if (!p || !vec) return fail(SEIS_EINVAL, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, nullptr));
  TRY(launch_pass_forward(p, s, false, false));
  return copy_out(p, s, p->d_seis, seis);
}

int seis_forward_misfit(seis_plan_t* p, const float* vec, const float* dd, float* seis, double* misfit,
                        seis_stream_t stream) {
  if (!p || !vec || !dd || !misfit) return fail(SEIS_EINVAL, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, nullptr));
  HIP_TRY(hipMemcpyAsync(p->d_dd, dd, sizeof(float) * p->L.seis, hipMemcpyDeviceToDevice, s));
  TRY(launch_pass_forward(p, s, false, false));
  launch_misfit(p, s, misfit);
  HIP_TRY(hipGetLastError());
  return copy_out(p, s, p->d_seis, seis);
}

int seis_forward_save(seis_plan_t* p, const float* vec, float* seis, seis_stream_t stream) {
  if (!p || !vec) return fail(SEIS_EINVAL, "NULL argument");
  if (!(p->caps & SEIS_CAP_ADJOINT)) return fail(SEIS_ESTATE, "plan lacks SEIS_CAP_ADJOINT");
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, nullptr));
  TRY(launch_pass_forward(p, s, false, true));
  mark_saved(p);
  return copy_out(p, s, p->d_seis, seis);
}

int seis_adjoint(seis_plan_t* p, const float* vec, const float* dd, int32_t on_residual, double* misfit,
                 double* grad, seis_stream_t stream) {
  if (!p || !vec || !dd || !grad) return fail(SEIS_EINVAL, "NULL argument");
  if (!(p->caps & SEIS_CAP_ADJOINT)) return fail(SEIS_ESTATE, "plan lacks SEIS_CAP_ADJOINT");
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, nullptr));   // same alpha as the saving forward; clears the source correlation
  HIP_TRY(hipMemcpyAsync(p->d_dd, dd, sizeof(float) * p->L.seis, hipMemcpyDeviceToDevice, s));
  if (on_residual && misfit) launch_misfit(p, s, misfit);
  const int kind = on_residual ? 0 : 1;
  TRY(launch_pass_adjoint(p, s, kind));
  return finish_grad(p, s, vec, grad);
}

int seis_tangent(seis_plan_t* p, const float* vec, const float* dvec, float* seis, float* dseis,
                 seis_stream_t stream) {
  if (!p || !vec || !dvec) return fail(SEIS_EINVAL, "NULL argument");
  if (!(p->caps & SEIS_CAP_TANGENT)) return fail(SEIS_ESTATE, "plan lacks SEIS_CAP_TANGENT");
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, dvec));
  TRY(launch_pass_forward(p, s, true, false));
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
  TRY(launch_pass_forward(p, s, false, true));
  mark_saved(p);
  if (on_residual && misfit) launch_misfit(p, s, misfit);
  const int kind = on_residual ? 0 : 1;
  TRY(launch_pass_adjoint(p, s, kind));
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
  TRY(launch_pass_forward(p, s, true, true));
  mark_saved(p);
  TRY(launch_pass_adjoint(p, s, 2));
  TRY(finish_grad(p, s, vec, grad));
  return copy_out(p, s, p->d_dseis, dseis);
}

int seis_tangent_saved(seis_plan_t* p, const float* vec, const float* dvec, float* dseis,
                       seis_stream_t stream) {
  if (!p || !vec || !dvec) return fail(SEIS_EINVAL, "NULL argument");
  TRY(check_saved(p));
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, dvec));
  TRY(launch_pass_born(p, s));
  return copy_out(p, s, p->d_dseis, dseis);
}

int seis_gn_product_saved(seis_plan_t* p, const float* vec, const float* dvec, float* dseis, double* grad,
                          seis_stream_t stream) {
  if (!p || !vec || !dvec || !grad) return fail(SEIS_EINVAL, "NULL argument");
  TRY(check_saved(p));
  hipStream_t s = (hipStream_t)stream;
  TRY(run_prep(p, s, vec, dvec));
  TRY(launch_pass_born(p, s));
  TRY(launch_pass_adjoint(p, s, 2));
  TRY(finish_grad(p, s, vec, grad));
  return copy_out(p, s, p->d_dseis, dseis);
}

}  // extern "C"
