#pragma once
// Device code of the gfx950 half-sarcomere engine: the persistent per-instance
// step kernel template and what it calls.  Included by every translation unit
// that instantiates kernel variants (mf_variants_*.hip) and by the host code
// (mf_engine.hip).
//
// One workgroup owns one instance for its whole run.  Node vectors of the
// Newton-CG solve live in registers, chain-major: wave w owns filament chains
// w, w+NW, ...; lane l owns contiguous nodes [l*c, l*c+c) of each, so the
// preconditioner (a pair of prefix/suffix sums along each chain, the exact
// inverse of the backbone tridiagonal) runs as wave scans without LDS traffic.
// Only the CG search direction, positions, per-XB force/stiffness and the
// binding bookkeeping are shared through LDS.
#include "mf_common.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <type_traits>

using namespace mf;

// Library-internal launch flag (above the public MF_LAUNCH_* bits): the
// instance's XB sampling and absorbing-chain rows are copied into LDS after
// its other arrays.
constexpr int LAUNCH_TABLES_IN_LDS = 1 << 16;
// Internal launch flag: float positions of every (thin, face) candidate list
// in LDS for the nearest-site search (set when they fit).
constexpr int LAUNCH_FACE_POS = 1 << 17;

// Pointers into the workgroup's LDS carry the local address space so every
// access compiles to ds_* instructions (a generic pointer would go through
// flat addressing).
#define LDS __attribute__((address_space(3)))

namespace {

// Launch shape helper: table sets per kinetics.
__host__ __device__ inline int table_sets(const mf_launch& L, int flag) {
  return (L.pop_mode == MF_POP_EXPLICIT && (L.pop_flags & flag)) ? L.n_pops : 1;
}

// ---------------------------------------------------------------------------
// state layout
// ---------------------------------------------------------------------------
struct StateLayout {
  int64_t pos, ls, xb_bound, xb_key, xb_near, xb_xdist, thin_bound, xb_state, tm, bytes;
};
__host__ __device__ inline int64_t al8(int64_t v) { return (v + 7) & ~int64_t(7); }
__host__ __device__ inline StateLayout state_layout(int Tk, int nc, int Tn, int ns, int nx) {
  const int64_t N = (int64_t)Tk * nc + (int64_t)Tn * ns, X = (int64_t)Tk * nc * nx,
                S = (int64_t)Tn * ns;
  StateLayout L;
  L.pos = 0;
  L.ls = L.pos + 8 * N;
  L.xb_bound = L.ls + 8;
  L.xb_key = al8(L.xb_bound + 4 * X);
  L.xb_near = al8(L.xb_key + 4 * X);
  L.xb_xdist = al8(L.xb_near + 4 * X);
  L.thin_bound = al8(L.xb_xdist + 4 * X);
  L.xb_state = al8(L.thin_bound + 4 * S);
  L.tm = L.xb_state + X;
  L.bytes = al8(L.tm + S);
  return L;
}

// ---------------------------------------------------------------------------
// transition tables: one matrix per thread
// ---------------------------------------------------------------------------
// The populations of one instance (core/subpopulation.py).  Without
// subpopulations K = 1 and row 0 is the instance's parameter row.
struct Pops {
  const float* prm;   // [K][MF_N_PARAMS], row 0 = wild type
  const float* frac;  // [K] mean-field weights
  int K;
  bool blend_tm, blend_xb;  // mean field: generator = sum_k f_k Q_k
};

__device__ __forceinline__ Pops pops_of(const mf_launch& L, int b) {
  Pops p;
  p.K = L.pop_mode != MF_POP_NONE ? L.n_pops : 1;
  p.prm = p.K > 1 ? L.pop_params + (int64_t)b * p.K * MF_N_PARAMS : L.params + (int64_t)b * MF_N_PARAMS;
  p.frac = p.K > 1 ? L.pop_fractions + (int64_t)b * p.K : nullptr;
  const bool mf = L.pop_mode == MF_POP_MEAN_FIELD && p.K > 1;
  p.blend_tm = mf && (L.pop_flags & MF_POP_TM);
  p.blend_xb = mf && (L.pop_flags & MF_POP_XB);
  return p;
}

// TM row block `cfg` (27 neighbour configurations) of table set `set`: FP64
// expm of Q dt, stored as FP32 cumulative rows (transitions.py:434-513,
// 99-190, 580-602).
__device__ __forceinline__ void tm_table_entry(const Pops& P, int set, double pCa, double dt, int cfg,
                                               float* tmrows) {
  double Q[4][4];
  if (P.blend_tm) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) Q[i][j] = 0.0;
    for (int k = 0; k < P.K; ++k)
      tm_generator<double, true>(P.prm + k * MF_N_PARAMS, pCa, cfg, Q, (double)P.frac[k]);
  } else {
    tm_generator<double>(P.prm + set * MF_N_PARAMS, pCa, cfg, Q);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) Q[i][j] *= dt;
  expm_ref<4, double>(Q);
  float* row = tmrows + (set * 27 + cfg) * 16;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double c = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) { c += Q[i][j]; row[i * 4 + j] = (float)c; }
  }
}

// The XB rates' offset- and spacing-free parts (XbPre) of parameter set k
// at pCa: from the cache (five doubles per set) when there is one, else
// computed (three powers and two exponentials).
__device__ __forceinline__ XbPre<double> xb_pre_of(const Pops& P, const double* cache, double pCa, int k) {
  if (cache) {
    const double* c = cache + 5 * k;
    XbPre<double> p;
    p.kt = c[0]; p.ikt = c[1]; p.r50 = c[2]; p.e10 = c[3]; p.e32 = c[4];
    return p;
  }
  return xb_pre<double>(P.prm + k * MF_N_PARAMS, pCa);
}
// Fill the cache for every parameter set at pCa (threads 0..K-1; the caller
// puts a barrier between this and the readers).
__device__ __forceinline__ void fill_xb_pre(const Pops& P, double pCa, LDS double* cache, int tid) {
  if (tid < P.K) {
    const XbPre<double> p = xb_pre<double>(P.prm + tid * MF_N_PARAMS, pCa);
    LDS double* c = cache + 5 * tid;
    c[0] = p.kt; c[1] = p.ikt; c[2] = p.r50; c[3] = p.e10; c[4] = p.e32;
  }
}

// The absorbing-chain entry of XB key (set * 2 nb + perm * nb + bin) at
// spacing d: P[3][4], P[2][0], P[3][0] of the chain with states 0 and 4
// absorbing (transitions.py:1029-1078), from its 3x3 transient block
// (absorbing_probs); mean-field populations blend the block.  ``pre``: the
// XbPre cache at pCa, or null.
__device__ __forceinline__ void absorbing_entry(const Pops& P, const double* pre, double pCa, double d, double dt,
                                                double center, int key, int nb, float* absrows) {
  const int set = key / (2 * nb), k2 = key % (2 * nb);
  const double perm = (double)(k2 / nb);
  double a[3][3], e10 = 0.0, e34 = 0.0;
  if (P.blend_xb) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) a[i][j] = 0.0;
    for (int k = 0; k < P.K; ++k) {
      const XbRates<double> r =
          xb_rates<double>(P.prm + k * MF_N_PARAMS, xb_pre_of(P, pre, pCa, k), d, center, perm);
      absorbing_block<double, true>(r, dt, (double)P.frac[k], a, e10, e34);
    }
  } else {
    const XbRates<double> r =
        xb_rates<double>(P.prm + set * MF_N_PARAMS, xb_pre_of(P, pre, pCa, set), d, center, perm);
    absorbing_block<double, false>(r, dt, 1.0, a, e10, e34);
  }
  double o[3];
  absorbing_probs<double>(a, e10, e34, o);
  float* row = absrows + (int64_t)key * 4;
  row[0] = (float)o[0];
  row[1] = (float)o[1];
  row[2] = (float)o[2];
  row[3] = 0.0f;
}

// XB key (set * 2 nb + perm * nb + bin) at spacing d: the 6 cumulative rows
// of the sampling table (transitions.py:632-922), or, for the absorbing
// chain (rows 0 and 4 zeroed, transitions.py:1029-1078), the three entries
// the metrics use.  Mean-field populations blend their generators
// (transitions.py:925-975).
__device__ __forceinline__ void xb_table_entry(const Pops& P, double pCa, double d, double dt, double center,
                                               int key, int nb, bool absorbing, float* xbrows, float* absrows) {
  const int set = key / (2 * nb), k2 = key % (2 * nb);
  const double perm = (double)(k2 / nb);
  if (absorbing) {
    absorbing_entry(P, nullptr, pCa, d, dt, center, key, nb, absrows);
    return;
  }
  double Q[6][6];
  if (P.blend_xb) {
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) Q[i][j] = 0.0;
    for (int k = 0; k < P.K; ++k)
      xb_generator<double, true>(P.prm + k * MF_N_PARAMS, pCa, d, center, perm, absorbing, Q, (double)P.frac[k]);
  } else {
    xb_generator<double>(P.prm + set * MF_N_PARAMS, pCa, d, center, perm, absorbing, Q);
  }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) Q[i][j] *= dt;
  expm_ref<6, double>(Q);
  if (!absorbing) {
    float* row = xbrows + (int64_t)key * 48;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double c = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) { c += Q[i][j]; row[i * 8 + j] = (float)c; }
      row[i * 8 + 6] = 0.0f;
      row[i * 8 + 7] = 0.0f;
    }
  } else {
    float* row = absrows + (int64_t)key * 4;
    row[0] = (float)Q[3][4];
    row[1] = (float)Q[2][0];
    row[2] = (float)Q[3][0];
    row[3] = 0.0f;
  }
}

// The rows of one XB key's entry from the quad's generator rows A (this
// lane's two rows, rp = min(l, 2)): scaled by dt, expm_quad6
// (xb_table_entry's result, bit for bit up to the sign of a zero), stored;
// l = lane & 3.
__device__ __forceinline__ void xb_quad_finish(double A[2][6], double dt, int key, bool absorbing, float* xbrows,
                                               float* absrows, int l) {
  const int rp = l < 2 ? l : 2;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) A[i][j] *= dt;
  expm_quad6(A, l);
  if (l == 3) return;
  if (!absorbing) {
    float* row = xbrows + (int64_t)key * 48;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      float* r = row + (2 * rp + i) * 8;
      double c = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) { c += A[i][j]; r[j] = (float)c; }
      r[6] = 0.0f;
      r[7] = 0.0f;
    }
  } else if (rp == 1) {  // rows 2 and 3: P[3][4], P[2][0], P[3][0]
    float* row = absrows + (int64_t)key * 4;
    row[0] = (float)A[1][4];
    row[1] = (float)A[0][0];
    row[2] = (float)A[1][0];
    row[3] = 0.0f;
  }
}

// The same for a pair's row triples (h = lane & 1): expm_pair6, rows
// 3h..3h+2 stored.  Sampling rows only.
__device__ __forceinline__ void xb_pair_finish(double A[3][6], double dt, int key, float* xbrows, int h) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) A[i][j] *= dt;
  expm_pair6(A, h);
  float* row = xbrows + (int64_t)key * 48;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float* r = row + (3 * h + i) * 8;
    double c = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) { c += A[i][j]; r[j] = (float)c; }
    r[6] = 0.0f;
    r[7] = 0.0f;
  }
}

// One mean-field-blended XB key's entry computed by the four lanes of a
// quad: the populations' generators accumulated in order, as
// xb_generator<double, true> does; l = lane & 3.
__device__ __forceinline__ void xb_blend_entry_quad(const Pops& P, const double* pre, double pCa, double d,
                                                    double dt, double center, int key, int nb, bool absorbing,
                                                    float* xbrows, float* absrows, int l) {
  const double perm = (double)((key % (2 * nb)) / nb);
  const int rp = l < 2 ? l : 2;
  double A[2][6];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) A[i][j] = 0.0;
  for (int k = 0; k < P.K; ++k) {
    const XbRates<double> r =
        xb_rates<double>(P.prm + k * MF_N_PARAMS, xb_pre_of(P, pre, pCa, k), d, center, perm);
    const double w = (double)P.frac[k];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      double g[6];
      xb_gen_row<double>(r, 2 * rp + i, absorbing, g);
#pragma unroll
      for (int j = 0; j < 6; ++j) A[i][j] += w * g[j];
    }
  }
  xb_quad_finish(A, dt, key, absorbing, xbrows, absrows, l);
}

// The kxb * 2 nb XB (or absorbing) entries of one instance at spacing d.
// With ``list`` only the listed keys (count of them) are rebuilt.  Sampling
// rows, in three passes over the listed keys:
//  1. one lane per key evaluates the rates (the transcendental part, which
//     the four lanes of a quad would otherwise repeat) and parks them in the
//     key's own row slot, with the key's scaling power s: twice the largest
//     exit rate times dt, the infinity norm of Q dt that expm_quad6 scales
//     by (a third of the keys have a rate at the cap and need 16 squarings
//     at dt = 1 ms, the others 1-11);
//  2. (with ``order``) a counting sort of the keys by s, largest first,
//     through the 20 counters of ``hist``;
//  3. groups of LPK lanes (pairs, expm_pair6, or quads, expm_quad6) take
//     the keys 64 / LPK at a time (a sub-round of one wave, whose groups
//     square in lockstep, so it lasts as long as its largest s) and write
//     the rows over the parked rates.  Sub-round j of the sorted list goes
//     to wave j mod nw, in reverse order on every other pass (a snake), so
//     the waves' sums of sub-round costs even out.  Pairs need 36 VGPRs per
//     matrix against the quads' 24: they are for blocks of up to 512
//     threads (256 VGPRs per lane).
// Each key's result depends only on its own rates, not on the order.  All
// the workgroup's threads call it; ``hist`` must be zero on entry and is
// zero again on return.
template <int LPK>
__device__ __forceinline__ void xb_table_rebuild_body(const Pops& P, float pCa, double d, float dt,
                                                      const float* centers, int nb, int kxb, bool absorbing,
                                                      float* xbrows, float* absrows, int tid, int nthreads,
                                                      const unsigned short* list, int count,
                                                      unsigned short* order, uint32_t* hist, const double* pre) {
  const int l = tid & 3;
  const int n = list ? count : kxb * 2 * nb;
  if (absorbing) {
    // one lane per key: a 3x3 exponential (absorbing_entry)
    for (int i = tid; i < n; i += nthreads) {
      const int key = list ? list[i] : i;
      absorbing_entry(P, pre, (double)pCa, d, (double)dt, (double)centers[key % nb], key, nb, absrows);
    }
    return;
  }
  if (P.blend_xb) {
    for (int i = tid >> 2; i < n; i += nthreads >> 2) {
      const int key = list ? list[i] : i;
      xb_blend_entry_quad(P, pre, (double)pCa, d, (double)dt, (double)centers[key % nb], key, nb, absorbing,
                          xbrows, absrows, l);
    }
    return;
  }
  const int lane = tid & 63, wave = tid >> 6, nw = nthreads >> 6, rp = l < 2 ? l : 2;
  // 1. rates and s per key, parked in the key's slot (11 doubles, then s)
  for (int i = tid; i < n; i += nthreads) {
    const int key = list ? list[i] : i;
    const int set = key / (2 * nb), k2 = key % (2 * nb);
    {
      const XbRates<double> r = xb_rates<double>(P.prm + set * MF_N_PARAMS, xb_pre_of(P, pre, (double)pCa, set), d,
                                                 (double)centers[key % nb], (double)(k2 / nb));
      double* slot = (double*)(xbrows + (int64_t)key * 48);
      slot[0] = r.r01; slot[1] = r.r04; slot[2] = r.r05; slot[3] = r.r10; slot[4] = r.r12; slot[5] = r.r21;
      slot[6] = r.r23; slot[7] = r.r32; slot[8] = r.r34; slot[9] = r.r40; slot[10] = r.r50;
      if (order) {
        const double ex = fmax(fmax(fmax(r.r10 + r.r12, r.r21 + r.r23), fmax(r.r32 + r.r34, r.r50)),
                               fmax(r.r01 + r.r04 + r.r05, r.r40));
        const int sv = (int)ceilf(log2f((float)(2.0 * ex * (double)dt) / 0.5f + 1e-10f));
        const int sb = sv < 0 ? 0 : (sv > 19 ? 19 : sv);
        ((int*)slot)[22] = sb;
        atomicAdd(hist + sb, 1u);
      }
    }
  }
  __syncthreads();
  // 2. counting sort, largest s first
  if (order) {
    if (tid < 20) {
      uint32_t off = 0;
      for (int b = 19; b > tid; --b) off += hist[b];
      hist[20 + tid] = off;
    }
    __syncthreads();
    for (int i = tid; i < n; i += nthreads) {
      const int key = list ? list[i] : i;
      const int sb = ((const int*)(xbrows + (int64_t)key * 48))[22];
      order[atomicAdd(hist + 20 + sb, 1u)] = (unsigned short)key;
    }
    if (tid < 20) hist[tid] = 0u;
    __syncthreads();
  }
  // 3. the lane groups, sub-round by sub-round in snake order
  constexpr int KPW = 64 / LPK;
  const int rounds = (n + KPW - 1) / KPW;
  for (int pass = 0;; ++pass) {
    const int j = pass * nw + ((pass & 1) ? nw - 1 - wave : wave);
    if (j >= rounds) break;
    const int pos = j * KPW + lane / LPK;
    if (pos < n) {
      const int kk = order ? order[pos] : (list ? list[pos] : pos);
      const double* slot = (const double*)(xbrows + (int64_t)kk * 48);
      XbRates<double> rk;
      rk.r01 = slot[0]; rk.r04 = slot[1]; rk.r05 = slot[2]; rk.r10 = slot[3]; rk.r12 = slot[4]; rk.r21 = slot[5];
      rk.r23 = slot[6]; rk.r32 = slot[7]; rk.r34 = slot[8]; rk.r40 = slot[9]; rk.r50 = slot[10];
      if constexpr (LPK == 2) {
        const int h = lane & 1;
        double A[3][6];
#pragma unroll
        for (int i2 = 0; i2 < 3; ++i2) xb_gen_row<double>(rk, 3 * h + i2, false, A[i2]);
        xb_pair_finish(A, (double)dt, kk, xbrows, h);
      } else {
        double A[2][6];
#pragma unroll
        for (int i2 = 0; i2 < 2; ++i2) xb_gen_row<double>(rk, 2 * rp + i2, false, A[i2]);
        xb_quad_finish(A, (double)dt, kk, false, xbrows, absrows, l);
      }
    }
  }
}

// The rebuild out of line, so the step kernel around it keeps its own
// register allocation (templated on the block size, so each copy gets the
// register budget of the kernels that call it).
template <int BS>
__device__ __attribute__((noinline)) void xb_table_rebuild(const Pops P, float pCa, double d, float dt,
                                                           const float* centers, int nb, int kxb,
                                                           bool absorbing, float* xbrows, float* absrows,
                                                           int tid, int nthreads, const unsigned short* list,
                                                           int count, unsigned short* order, uint32_t* hist,
                                                           const double* pre) {
  xb_table_rebuild_body<(BS <= 512 ? 2 : 4)>(P, pCa, d, dt, centers, nb, kxb, absorbing, xbrows, absrows, tid,
                                             nthreads, list, count, order, hist, pre);
}

// The absorbing rows alone, one lane per key, out of line in a function of
// their own: its few registers are all the call saves and restores (the
// sampling-row rebuild above uses the whole register file).
template <int BS>
__device__ __attribute__((noinline)) void absorbing_rebuild(const Pops P, float pCa, double d, float dt,
                                                            const float* centers, int nb, int kxb, float* absrows,
                                                            int tid, const unsigned short* list, int count,
                                                            const double* pre) {
  const int n = list ? count : kxb * 2 * nb;
  for (int i = tid; i < n; i += BS) {
    const int key = list ? list[i] : i;
    absorbing_entry(P, pre, (double)pCa, d, (double)dt, (double)centers[key % nb], key, nb, absrows);
  }
}

// The step kernel's rebuild: a call for the LDS-resident variants, inline
// for the global-workspace ones.  Values live across a call sit in the
// callee-saved registers or are spilled, and at 12 node slots per lane the
// compiler then reloads the CG state from scratch on every CG pass
// (~180 scratch operations per pass at 6x6, against ~90 inline): a 6x6
// dynamic-spacing step takes 1.26 ms inline and 1.6 ms with the call.  At
// 2-8 slots per lane the call is as fast or faster (4x4: 331 against
// 368 us per dynamic step).
template <int BS, int GW>
__device__ __forceinline__ void rebuild_rows(const Pops& P, float pCa, double d, float dt, const float* centers,
                                             int nb, int kxb, bool absorbing, float* xbrows, float* absrows,
                                             int tid, const unsigned short* list, int count,
                                             unsigned short* order, uint32_t* hist, const double* pre) {
  if constexpr (GW == 0) {
    if (absorbing && !P.blend_xb) {
      absorbing_rebuild<BS>(P, pCa, d, dt, centers, nb, kxb, absrows, tid, list, count, pre);
      return;
    }
  }
  if constexpr (GW > 0)
    xb_table_rebuild_body<(BS <= 512 ? 2 : 4)>(P, pCa, d, dt, centers, nb, kxb, absorbing, xbrows, absrows, tid,
                                               BS, list, count, order, hist, pre);
  else
    xb_table_rebuild<BS>(P, pCa, d, dt, centers, nb, kxb, absorbing, xbrows, absrows, tid, BS, list, count, order,
                         hist, pre);
}

// The keys marked in a 64-word bitmap, in increasing order, into list[]; run
// by one whole wave (lane w takes word w); returns the count to every lane.
__device__ __forceinline__ int compact_keys(const LDS uint32_t* bits, unsigned short* list, int lane) {
  uint32_t w = bits[lane];
  const float c = (float)__popc(w);
  const float incl = wave_prefix(c, lane);
  int at = (int)wave_prefix_excl(incl, lane);
  while (w) {
    const int b = __ffs(w) - 1;
    list[at++] = (unsigned short)(lane * 32 + b);
    w &= w - 1;
  }
  return (int)read_lane(incl, 63);
}

// ---------------------------------------------------------------------------
// persistent step kernel
// ---------------------------------------------------------------------------
struct Dims {
  int Tk, nc, Tn, ns, nx, W, nf, nb;
  int nk, N, X, S, nch;
  int ktm, kxb;  // TM / XB table sets (explicit subpopulations), else 1
};

// Per-lane node slots (chain-major).
template <int MS>
struct Slots {
  int node[MS];   // global DOF index, -1 if none
  int li[MS];     // index along the chain
  int chk[MS];    // ordinal of the chain among this wave's chains
  int fil[MS];    // filament index within its type
  bool thick[MS];
  int nch;        // chains owned by this wave
  int per0;       // slots of the wave's first chain (slots 0..per0-1)
  bool thick0;    // the first chain is a thick filament
};

// XB partners of a lane's slot nodes, fixed during a solve: for a crown the
// sites its bound heads hold, for a site the crowns of the heads bound to it
// (list order).  Packed (partner DOF << 16 | XB index); -1 = none.  A site
// with more than PM bound heads walks the LDS list instead (overflow).
// Lanes that own at most two nodes keep PM = 4 partners per node in
// registers; lanes with more nodes (large lattices) keep none (PM = 0) and
// read a crown's heads and a site's list from LDS, since 4 partners x
// their node count would not fit the register file.
constexpr int MAXP = 4;
template <int MS>
struct Partners {
  static constexpr int PM = MS <= 2 ? MAXP : 0;
  static constexpr int P = PM > 0 ? PM : 1;
  int pk[MS][P];
  float k[MS][P];
  bool overflow[MS];
};
__device__ __forceinline__ int p_node(int pk) { return pk >> 16; }
__device__ __forceinline__ int p_xb(int pk) { return pk & 0xffff; }

// Integer metric counters (LDS, one row per wave, rewritten every step).
enum Counter { C_XB0 = 0, C_TM0 = 6, C_OV = 10, C_OV2, C_OV3, C_ATPC, C_NEWLY, NC = 16 };
// Per-thick-filament metric values written by the owning lane; the first
// NF_SUM are summed over filaments, F_X0 / F_XT hold the pre-solve positions
// of the M-line and tip crowns.
enum FilVal { F_FORCE = 0, F_E1, F_DE1, F_TIT, F_DTIT, NF_SUM, F_X0 = NF_SUM, F_XT, F_EO, NF };
// Per-wave partial sums: pass 1 (node sums / maxima, XB expectations), pass 2,
// and the pre-solve sum of thick positions (work = sum x_new - sum x_pre).
// W_RG / W_RGD carry the radial sums of a dynamic-spacing residual.
enum WavePart { W_SUMK = 0, W_SUMN, W_WORK, W_MAXK, W_NMINK, W_MAXN, W_NMINN, W_ATPE, W_TEAR, W_SQK, W_SQN,
                W_XPRE, W_RG, W_RGD, NWP };

// All per-instance arrays of one workgroup.  Positions, the arrays updated
// by atomics (site lists, thin.bound_to) and the small tables always sit in
// LDS.  The other per-node / per-XB / per-site arrays ("Big") sit in LDS
// too, except in the global-workspace (GW) variants for lattices whose
// arrays exceed the 160 KB of LDS: there they live in the instance's slice
// of a device workspace.  Only the waves of this workgroup touch that slice,
// with plain loads and stores, and they share one CU and its write-through
// L1, so a workgroup barrier orders those accesses as it orders LDS ones
// (atomics, which execute in L2, are kept to LDS arrays).
template <bool REF64, int GW = 0>
struct Shared {
  template <class T> using Big = typename std::conditional<(GW > 0), T*, LDS T*>::type;
  // level 2 (HG: lattices above 6x6) also keeps the positions, the per-head
  // site and state and the TM states in the workspace
  template <class T> using Huge = typename std::conditional<(GW > 1), T*, LDS T*>::type;
  Huge<float> x32;       // [N]   positions (reference mode)
  Huge<double> x64;      // [N]   positions (refined mode)
  Big<float> pv;         // [N]   CG direction
  Big<float> fx;         // [X]   XB axial force (reference mode)
  Big<double> fx64;      // [X]   XB axial force (refined mode)
  Big<float> kx;         // [X]   XB axial stiffness df/dx
  Huge<short> xsite;     // [X]   DOF index of the bound site, -1 unbound
  Big<short> xnext;      // [X]   next XB in the site's bound list
  Big<short> xkey;       // [X]   table row of the last kinetics step (< 2 * MF_MAX_BINS)
  Big<short> xslot;      // [X]   (thin, face) slot | target-valid << 15
  Big<float> org;        // [N]   displacement origin: crown offset / cumulative thin rest
  Big<float> cbn;        // [N]   GW + dynamic spacing: the border column per node, which
                         //       other variants keep in registers
  Huge<unsigned char> xst;  // [X] XB state
  LDS int* shead;        // [S]   first XB bound to the site, -1 none
  LDS int* tbound;       // [S]   thin.bound_to
  Huge<unsigned char> tm;   // [2][S] TM state (double buffered)
  Big<short> fsites;     // [Tn*nf*W] candidate sites per (thin, face)
  Big<short> fcount;     // [Tn*nf]
  Big<signed char> tprev;  // [S] same-chain neighbours as offsets along the filament
  Big<signed char> tnext;  // [S] (0: none, the site references itself)
  LDS float* tmcum;      // [ktm*27*16]
  LDS float* edges;      // [nb+1]
  LDS float* prm;        // [MF_N_PARAMS]
  LDS float* red;        // [8][NW] scalar reduction ring
  LDS double* wpart;     // [NW][NWP] metric wave partials
  LDS double* fil;       // [Tk][NF] metric per-filament values
  LDS float* ktip;       // [Tk] titin tip stiffness of each thick filament (last residual)
  LDS int* cnt;          // [NW][NC] metric counters, one row per wave
  LDS uint32_t* need;    // [2][64] bitmaps of the XB keys a step reads (sampling, absorbing),
                         // then the two compacted key counts, then the rebuild's
                         // [20] counters and [20] cursors of its sort by squaring count
  LDS double* xbpre;     // [MF_MAX_POPS][5] XbPre of each parameter set at the rebuild's pCa
};

// LDS / workspace carve-up.  One routine computes the byte offsets (host:
// the dynamic LDS size and the workspace slice; device: the typed
// pointers), so the two cannot disagree.  Offsets of "Big" arrays are into
// the workspace slice when GW is set.
struct LdsOffsets {
  int64_t x, pv, fx, kx, xsite, xnext, xkey, xslot, org, cbn, xst, shead, tbound, tm, fsites, fcount, tprev, tnext, tmcum, need,
      edges, prm, red, wpart, fil, ktip, cnt, xbpre, total, ws_total;
};
template <int BS, bool REF64, int GW = 0>
__host__ __device__ inline LdsOffsets lds_offsets(const Dims& d) {
  constexpr int NW = BS / WAVE;
  LdsOffsets L;
  int64_t o = 0, w = 0;
  auto take = [&](int64_t bytes) { const int64_t at = o; o = (o + bytes + 15) & ~int64_t(15); return at; };
  auto big = [&](int64_t bytes) {
    if (!GW) return take(bytes);
    const int64_t at = w;
    w = (w + bytes + 255) & ~int64_t(255);
    return at;
  };
  auto huge = [&](int64_t bytes) { return GW > 1 ? big(bytes) : take(bytes); };
  L.x = huge((REF64 ? 8LL : 4LL) * d.N);
  L.pv = big(4LL * d.N);
  L.fx = big((REF64 ? 8LL : 4LL) * d.X);
  L.kx = big(4LL * d.X);
  L.xsite = huge(2LL * d.X);
  L.xnext = big(2LL * d.X);
  L.xkey = big(2LL * d.X);
  L.xslot = big(2LL * d.X);
  L.org = big(4LL * d.N);
  L.cbn = GW ? big(4LL * d.N) : 0;
  L.xst = huge(d.X);
  L.shead = take(4LL * d.S);
  L.tbound = take(4LL * d.S);
  L.tm = huge(2LL * d.S);
  L.fsites = big(2LL * d.Tn * d.nf * d.W);
  L.fcount = big(2LL * d.Tn * d.nf);
  L.tprev = big(d.S);
  L.tnext = big(d.S);
  L.tmcum = take(4 * 27 * 16 * d.ktm);
  L.edges = take(4LL * (d.nb + 1));
  L.prm = take(4 * MF_N_PARAMS);
  L.red = take(4 * 8 * NW);
  L.wpart = take(8LL * NW * NWP);
  L.fil = take(8LL * d.Tk * NF);
  L.ktip = take(4LL * d.Tk);
  L.cnt = take(4 * NC * NW);
  L.need = take(4 * (2 * 64 + 2 + 40));
  L.xbpre = take(8 * 5 * MF_MAX_POPS);
  L.total = o;
  L.ws_total = w;
  return L;
}
template <int BS, bool REF64, int GW>
int64_t lds_bytes(const Dims& d) { return lds_offsets<BS, REF64, GW>(d).total; }
template <int BS, bool REF64, int GW>
int64_t ws_bytes(const Dims& d) { return lds_offsets<BS, REF64, GW>(d).ws_total; }

template <int BS, bool REF64, int GW>
__device__ __forceinline__ Shared<REF64, GW> carve(LDS char* base, char* ws, const Dims& d) {
  const LdsOffsets o = lds_offsets<BS, REF64, GW>(d);
  using S = Shared<REF64, GW>;
  // a Big array's base: the workspace slice in GW variants, LDS otherwise
  auto bb = [&](int64_t off) {
    if constexpr (GW) return (char*)(ws + off); else return (LDS char*)(base + off);
  };
  S t;
  auto hb = [&](int64_t off) {
    if constexpr (GW > 1) return (char*)(ws + off); else return (LDS char*)(base + off);
  };
  t.x32 = (typename S::template Huge<float>)hb(o.x);
  t.x64 = (typename S::template Huge<double>)hb(o.x);
  t.pv = (typename S::template Big<float>)bb(o.pv);
  t.fx = (typename S::template Big<float>)bb(o.fx);
  t.fx64 = (typename S::template Big<double>)bb(o.fx);
  t.kx = (typename S::template Big<float>)bb(o.kx);
  t.xsite = (typename S::template Huge<short>)hb(o.xsite);
  t.xnext = (typename S::template Big<short>)bb(o.xnext);
  t.xkey = (typename S::template Big<short>)bb(o.xkey);
  t.xslot = (typename S::template Big<short>)bb(o.xslot);
  t.org = (typename S::template Big<float>)bb(o.org);
  t.cbn = (typename S::template Big<float>)bb(o.cbn);
  t.xst = (typename S::template Huge<unsigned char>)hb(o.xst);
  t.shead = (LDS int*)(base + o.shead);
  t.tbound = (LDS int*)(base + o.tbound);
  t.tm = (typename S::template Huge<unsigned char>)hb(o.tm);
  t.fsites = (typename S::template Big<short>)bb(o.fsites);
  t.fcount = (typename S::template Big<short>)bb(o.fcount);
  t.tprev = (typename S::template Big<signed char>)bb(o.tprev);
  t.tnext = (typename S::template Big<signed char>)bb(o.tnext);
  t.tmcum = (LDS float*)(base + o.tmcum);
  t.edges = (LDS float*)(base + o.edges);
  t.prm = (LDS float*)(base + o.prm);
  t.red = (LDS float*)(base + o.red);
  t.wpart = (LDS double*)(base + o.wpart);
  t.fil = (LDS double*)(base + o.fil);
  t.ktip = (LDS float*)(base + o.ktip);
  t.cnt = (LDS int*)(base + o.cnt);
  t.need = (LDS uint32_t*)(base + o.need);
  t.xbpre = (LDS double*)(base + o.xbpre);
  return t;
}

// Workgroup scalar reductions through a ring of LDS rows (one barrier each).
template <int BS>
struct Red {
  static constexpr int NW = BS / WAVE;
  LDS float* buf;  // [8][NW]
  int ring;
  __device__ LDS float* row() { LDS float* p = buf + (ring & 7) * NW; ++ring; return p; }
};
// Cross-wave step of these reductions, after the barrier.  TREE = false:
// every wave adds the NW partials in wave order.  TREE = true: lane w of row
// q (w < NW <= 16) reads wave w's partial of quantity q and a 16-lane DPP
// tree per row folds them (fewer instructions, a longer dependent chain:
// it pays in the many-slot variants, not at two slots per lane).  Either
// way every wave runs the same operations on the same partials and gets the
// same bits.
template <int BS, bool TREE = false>
__device__ __forceinline__ float wg_sum(Red<BS>& rd, float v, int lane, int wave) {
  static_assert(Red<BS>::NW <= 16, "one 16-lane row per quantity");
  LDS float* r = rd.row();
  const float s = wave_sum(v, lane);
  if (lane == 0) r[wave] = s;
  __syncthreads();
  if constexpr (TREE) {
    const float x = lane < Red<BS>::NW ? r[lane] : 0.0f;
    return read_lane(row_reduce16(x), 15);
  } else {
    float t = 0.0f;
#pragma unroll
    for (int w = 0; w < Red<BS>::NW; ++w) t += r[w];
    return t;
  }
}
// Two sums with one barrier (consumes two ring rows; a row is rewritten only
// after another reduction's barrier, by which time every wave has read it).
template <int BS, bool TREE = false>
__device__ __forceinline__ void wg_sum2(Red<BS>& rd, float& a, float& b, int lane, int wave) {
  static_assert(Red<BS>::NW <= 16, "one 16-lane row per quantity");
  LDS float* ra = rd.row();
  LDS float* rb = rd.row();
  const float sa = wave_sum(a, lane);
  const float sb = wave_sum(b, lane);
  if (lane == 0) { ra[wave] = sa; rb[wave] = sb; }
  __syncthreads();
  if constexpr (TREE) {
    const int w = lane & 15;
    const float x = w < Red<BS>::NW && lane < 32 ? (lane < 16 ? ra[w] : rb[w]) : 0.0f;
    const float t = row_reduce16(x);
    a = read_lane(t, 15);
    b = read_lane(t, 31);
  } else {
    float ta = 0.0f, tb = 0.0f;
#pragma unroll
    for (int w = 0; w < Red<BS>::NW; ++w) { ta += ra[w]; tb += rb[w]; }
    a = ta;
    b = tb;
  }
}
template <int BS, bool TREE = false>
__device__ __forceinline__ void wg_sum3(Red<BS>& rd, float& a, float& b, float& c, int lane, int wave) {
  static_assert(Red<BS>::NW <= 16, "one 16-lane row per quantity");
  LDS float* ra = rd.row();
  LDS float* rb = rd.row();
  LDS float* rc = rd.row();
  const float sa = wave_sum(a, lane), sb = wave_sum(b, lane), sc = wave_sum(c, lane);
  if (lane == 0) { ra[wave] = sa; rb[wave] = sb; rc[wave] = sc; }
  __syncthreads();
  if constexpr (TREE) {
    const int w = lane & 15;
    const float x = w < Red<BS>::NW && lane < 48 ? (lane < 16 ? ra[w] : lane < 32 ? rb[w] : rc[w]) : 0.0f;
    const float t = row_reduce16(x);
    a = read_lane(t, 15);
    b = read_lane(t, 31);
    c = read_lane(t, 47);
  } else {
    float ta = 0.0f, tb = 0.0f, tc = 0.0f;
#pragma unroll
    for (int w = 0; w < Red<BS>::NW; ++w) { ta += ra[w]; tb += rb[w]; tc += rc[w]; }
    a = ta;
    b = tb;
    c = tc;
  }
}
template <int BS, bool TREE = false>
__device__ __forceinline__ float wg_max(Red<BS>& rd, float v, int lane, int wave) {
  static_assert(Red<BS>::NW <= 16, "one 16-lane row per quantity");
  LDS float* r = rd.row();
  const float s = wave_max(v, lane);
  if (lane == 0) r[wave] = s;
  __syncthreads();
  if constexpr (TREE) {
    const float x = lane < Red<BS>::NW ? r[lane] : -INFINITY;
    return read_lane(row_reduce16(x, OpMax()), 15);
  } else {
    float t = r[0];
#pragma unroll
    for (int w = 1; w < Red<BS>::NW; ++w) t = fmaxf(t, r[w]);
    return t;
  }
}

// Scan of slot values along every chain of this wave.  dir 0: inclusive
// prefix in node order, 1: inclusive suffix.  Direction per chain type.
template <int MS>
__device__ __forceinline__ void chain_scan(float (&v)[MS], const Slots<MS>& sl, int lane, int thick_dir,
                                           int thin_dir) {
  // (two-slot kernels only: in the many-slot ones the second copy of the
  // scan costs the CG loop registers, 6x6 steps 25 % slower)
  if (MS <= 2 && sl.nch == 1) {
    // one chain in the wave, in slots 0..per0-1: the same operations as
    // the general case below without its per-slot chain tests
    const int dir = sl.thick0 ? thick_dir : thin_dir;
    float acc = 0.0f;
    if (dir == 0) {
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < sl.per0) { acc += v[s]; v[s] = acc; }
      const float ex = wave_prefix_excl(wave_prefix(acc, lane), lane);
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < sl.per0) v[s] += ex;
    } else {
#pragma unroll
      for (int s = MS - 1; s >= 0; --s)
        if (s < sl.per0) { acc += v[s]; v[s] = acc; }
      const float incl = wave_prefix(acc, lane);
      const float ex = read_lane(incl, 63) - incl;  // sum over the lanes after this one
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < sl.per0) v[s] += ex;
    }
    return;
  }
  for (int k = 0; k < sl.nch; ++k) {
    bool th = false;
#pragma unroll
    for (int s = 0; s < MS; ++s) th = (sl.chk[s] == k) ? sl.thick[s] : th;
    const int dir = th ? thick_dir : thin_dir;
    float acc = 0.0f;
    if (dir == 0) {
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (sl.chk[s] == k) { acc += v[s]; v[s] = acc; }
      const float ex = wave_prefix_excl(wave_prefix(acc, lane), lane);
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (sl.chk[s] == k) v[s] += ex;
    } else {
#pragma unroll
      for (int s = MS - 1; s >= 0; --s)
        if (sl.chk[s] == k) { acc += v[s]; v[s] = acc; }
      const float incl = wave_prefix(acc, lane);
      const float ex = read_lane(incl, 63) - incl;  // sum over the lanes after this one
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (sl.chk[s] == k) v[s] += ex;
    }
  }
}

// z = M^-1 r with M the negated backbone tridiagonal (state.py:269-316):
// thick (M-line anchored, tip free): z_i = sum_j (min(i,j)+1) r_j / k
//   = prefix over i of (suffix sums of r) / k;
// thin (free end at 0, Z-line anchored): z_i = sum_j (n - max(i,j)) r_j / k
//   = suffix over i of (prefix sums of r) / k.
template <int MS>
__device__ __forceinline__ void precond(const float (&r)[MS], float (&z)[MS], const Slots<MS>& sl, int lane,
                                        float inv_kk, float inv_kn) {
#pragma unroll
  for (int s = 0; s < MS; ++s) z[s] = sl.node[s] >= 0 ? r[s] : 0.0f;
  chain_scan<MS>(z, sl, lane, 1, 0);
  // A chain's padding slots (past its last node: lanes 43-63 of an 86-site
  // thin chain) leave the first scan holding its running total; zeroed, so
  // the second scan sums the chain's nodes only.  Left in, they add
  // (padding count) x (sum of r) / k to every node of the chain: still SPD,
  // but a far worse preconditioner (about 1.5 more CG passes per solve).
#pragma unroll
  for (int s = 0; s < MS; ++s) z[s] = sl.node[s] >= 0 ? z[s] : 0.0f;
  chain_scan<MS>(z, sl, lane, 0, 1);
#pragma unroll
  for (int s = 0; s < MS; ++s) z[s] *= sl.thick[s] ? inv_kk : inv_kn;
}

template <bool REF64, int GW>
__device__ __forceinline__ double posd(const Shared<REF64, GW>& sh, int i) {
  if constexpr (REF64) return sh.x64[i]; else return (double)sh.x32[i];
}
template <bool REF64, int GW>
__device__ __forceinline__ float posf(const Shared<REF64, GW>& sh, int i) {
  if constexpr (REF64) return (float)sh.x64[i]; else return sh.x32[i];
}

// e^x in the residual's type.  OWN (kernels with three or more node slots
// per lane, whose register pressure makes the compiler hold and spill the
// library exp's constants): exp_f64 in double; otherwise the library's exp
// (the two-slot kernels keep its constants in registers at no cost: the
// 2x2 fixed-spacing step is 7 % slower with exp_f64).
template <typename T, bool OWN>
__device__ __forceinline__ T exp_t(T x) {
  if constexpr (OWN && std::is_same<T, double>::value) return exp_f64(x); else return exp(x);
}

// XB spring constants of a state (forces.py:378-394).
struct Springs { float gk, gr, ck, cr; };
__device__ __forceinline__ Springs springs(const LDS float* prm, int st) {
  const bool strong = (st == 2) || (st == 3);
  Springs s;
  s.gk = strong ? prm[P_G_K_S] : prm[P_G_K_W];
  s.gr = strong ? prm[P_G_REST_S] : prm[P_G_REST_W];
  s.ck = strong ? prm[P_C_K_S] : prm[P_C_K_W];
  s.cr = strong ? prm[P_C_REST_S] : prm[P_C_REST_W];
  return s;
}

// Angle of a bound head, theta = atan2(d, xd).  The fixed-spacing refined
// residual takes it from a float atan2 (|error| ~1e-7 rad, ~2e-7 pN per
// head, far below the stopping tolerance at any one node).  The radial
// residual of dynamic spacing sums over every bound head of the lattice,
// where that rounding would add up to ~1e-5 pN, so it uses a double atan2:
// atan2_pos_f64 (the spacing is positive), whose coefficients, unlike the
// library function's, the compiler does not hold in registers across the
// time loop.
template <typename T, bool EXACT_ANGLE>
__device__ __forceinline__ T xb_angle(T d, T xd) {
  if constexpr (EXACT_ANGLE && std::is_same<T, double>::value) return atan2_pos_f64(d, xd);
  else if constexpr (EXACT_ANGLE) return atan2(d, xd);
  else return T(atan2f((float)d, (float)xd));
}

// Axial force of a bound XB at offset xd and angle th (forces.py:366-399)
// and its derivative d f / d xd (what jax.jvp produces at solver.py:413).
template <typename T>
__device__ __forceinline__ void xb_force(const Springs& sp, T xd, T d, T th, T& f, float& kf) {
  const T r = sqrt(xd * xd + d * d);
  const T rs = r > T(1e-10) ? r : T(1e-10);
  if constexpr (std::is_same<T, double>::value) {
    // refined mode: one double division for the force; the stiffness only
    // steers CG, so it is formed in float from the float reciprocal
    const double inv = 1.0 / rs;
    f = (double(sp.gk) * (r - double(sp.gr)) * xd - inv * double(sp.ck) * (th - double(sp.cr)) * d) * inv;
    const float dfl = (float)d, xf = (float)xd, dth = (float)th - sp.cr;
    const float i1 = __builtin_amdgcn_rcpf((float)rs), i2 = i1 * i1;
    kf = sp.gk * (1.0f - sp.gr * dfl * dfl * (i2 * i1)) + sp.ck * dfl * (dfl + 2.0f * xf * dth) * (i2 * i2);
  } else {
    // reference mode: the reference's float expression (forces.py:398-399)
    const T cs = xd / rs, sn = d / rs;
    f = T(sp.gk) * (r - T(sp.gr)) * cs - (T(1) / rs) * T(sp.ck) * (th - T(sp.cr)) * sn;
    const T r2 = rs * rs;
    kf = (float)(T(sp.gk) * (T(1) - T(sp.gr) * d * d / (r2 * rs)) +
                 T(sp.ck) * d * (d + T(2) * xd * (th - T(sp.cr))) / (r2 * r2));
  }
}

// A bound XB's axial force and stiffness (as xb_force) together with its
// spacing derivatives, for the bordered system of dynamic lattice spacing
// (solver.py:552-656, forces.py:649-749): f_d = d f / d d, the radial force
// g (forces.py:708-709) and g_d = d g / d d.  With E the head's energy,
// f = dE/dxd and g = dE/dd, so d g / d xd = f_d.  The forces f and g enter
// the residuals and are formed in T from one reciprocal of r; the three
// derivatives only steer the Newton and CG iterations and are formed in
// float, like the stiffness.
template <typename T>
__device__ __forceinline__ void xb_force_radial(const Springs& sp, T xd, T d, T th, T& f, float& kf, float& f_d,
                                                T& g, float& g_d) {
  if constexpr (!std::is_same<T, double>::value) {
    // reference mode: the reference's float expressions
    xb_force<T>(sp, xd, d, th, f, kf);
    const T r = sqrt(xd * xd + d * d);
    const T rs = r > T(1e-10) ? r : T(1e-10);
    const T r3 = rs * rs * rs, r4 = r3 * rs;
    const T dth = th - T(sp.cr);
    f_d = (float)(T(sp.gk) * T(sp.gr) * xd * d / r3 - T(sp.ck) * (xd * d + dth * (xd * xd - d * d)) / r4);
    g = T(sp.gk) * (rs - T(sp.gr)) * (d / rs) + (T(1) / rs) * T(sp.ck) * dth * (xd / rs);
    g_d = (float)(T(sp.gk) * (T(1) - T(sp.gr) * xd * xd / r3) + T(sp.ck) * xd * (xd - T(2) * d * dth) / r4);
    return;
  }
  const T r = sqrt(xd * xd + d * d);
  const T rs = r > T(1e-10) ? r : T(1e-10);
  const T inv = T(1) / rs;
  const T dth = th - T(sp.cr);
  f = (T(sp.gk) * (r - T(sp.gr)) * xd - inv * T(sp.ck) * dth * d) * inv;
  g = T(sp.gk) * (rs - T(sp.gr)) * (d * inv) + inv * T(sp.ck) * dth * (xd * inv);
  const float dfl = (float)d, xf = (float)xd, dthf = (float)th - sp.cr;
  const float i1 = __builtin_amdgcn_rcpf((float)rs), i2 = i1 * i1, i3 = i2 * i1, i4 = i2 * i2;
  kf = sp.gk * (1.0f - sp.gr * dfl * dfl * i3) + sp.ck * dfl * (dfl + 2.0f * xf * dthf) * i4;
  f_d = sp.gk * sp.gr * xf * dfl * i3 - sp.ck * (xf * dfl + dthf * (xf * xf - dfl * dfl)) * i4;
  g_d = sp.gk * (1.0f - sp.gr * xf * xf * i3) + sp.ck * xf * (xf - 2.0f * dfl * dthf) * i4;
}

// The same spacing derivatives on their own, for the workspace variants
// (those keep this split form: the merged one, though it does less work,
// moves their spills into the CG loop and costs them about 5 %).
template <typename T>
__device__ __forceinline__ void xb_radial(const Springs& sp, T xd, T d, T th, float& f_d, T& g, float& g_d) {
  const T r = sqrt(xd * xd + d * d);
  const T rs = r > T(1e-10) ? r : T(1e-10);
  const T r3 = rs * rs * rs, r4 = r3 * rs;
  const T dth = th - T(sp.cr);
  f_d = (float)(T(sp.gk) * T(sp.gr) * xd * d / r3 - T(sp.ck) * (xd * d + dth * (xd * xd - d * d)) / r4);
  g = T(sp.gk) * (rs - T(sp.gr)) * (d / rs) + (T(1) / rs) * T(sp.ck) * dth * (xd / rs);
  g_d = (float)(T(sp.gk) * (T(1) - T(sp.gr) * xd * xd / r3) + T(sp.ck) * xd * (xd - T(2) * d * dth) / r4);
}

// Titin axial force on the tip crown and its derivative with respect to the
// tip position (forces.py:136-157).  RAD: also, for the bordered system, the
// radial force tr (N_TITIN F d / L), its spacing derivative tr_d, and
// tt_d = d(axial titin force)/d d, which is also d(radial residual)/d x_tip.
template <typename T, bool RAD = false, bool OWN = false>
__device__ __forceinline__ void titin(const LDS float* prm, T tip, T z, T d, T& tt, float& ktt, T& energy,
                                      T* tr = nullptr, float* tr_d = nullptr, float* tt_d = nullptr) {
  const T ad = z - tip;
  const T Lq = sqrt(ad * ad + d * d);
  const T arg = T(prm[P_TITIN_B]) * (Lq - T(prm[P_TITIN_REST]));
  const T argc = arg < T(-100) ? T(-100) : (arg > T(100) ? T(100) : arg);
  const T E = exp_t<T, OWN>(argc);
  // stored energy a/b (exp(b (L - rest)) - 1) (metrics_fn.py), from the
  // unclamped exponent
  energy = T(prm[P_TITIN_A]) / T(prm[P_TITIN_B]) * ((arg == argc ? E : exp_t<T, OWN>(arg)) - T(1));
  T F = T(prm[P_TITIN_A]) * E;
  F = F > T(0) ? F : T(0);
  const T cs = Lq > T(0) ? ad / Lq : T(0);
  tt = T(N_TITIN) * F * cs;
  const T Fp = (arg > T(-100) && arg < T(100)) ? T(prm[P_TITIN_B]) * F : T(0);
  ktt = (float)(-T(N_TITIN) * (Fp * ad * ad / (Lq * Lq) + F * d * d / (Lq * Lq * Lq)));
  if constexpr (RAD) {
    const T Ls = Lq > T(1e-10) ? Lq : T(1e-10);
    const T il = T(1) / Ls, il2 = il * il, il3 = il2 * il;
    *tr = T(N_TITIN) * F * (d * il);
    *tr_d = (float)(T(N_TITIN) * (Fp * d * d * il2 + F * ad * ad * il3));
    *tt_d = (float)(T(N_TITIN) * ad * d * (Fp * il2 - F * il3));
  }
}

// The titin radial terms on their own (as titin<T, true> forms them), for
// the workspace variants.
template <typename T, bool OWN = false>
__device__ __forceinline__ void titin_radial(const LDS float* prm, T tip, T z, T d, T& tr, float& tr_d,
                                             float& tt_d) {
  const T ad = z - tip;
  const T Lq = sqrt(ad * ad + d * d);
  const T Ls = Lq > T(1e-10) ? Lq : T(1e-10);
  const T arg = T(prm[P_TITIN_B]) * (Lq - T(prm[P_TITIN_REST]));
  const T argc = arg < T(-100) ? T(-100) : (arg > T(100) ? T(100) : arg);
  T F = T(prm[P_TITIN_A]) * exp_t<T, OWN>(argc);
  F = F > T(0) ? F : T(0);
  const T bF = (arg > T(-100) && arg < T(100)) ? T(prm[P_TITIN_B]) * F : T(0);
  tr = T(N_TITIN) * F * (d / Ls);
  const T L2 = Ls * Ls, L3 = L2 * Ls;
  tr_d = (float)(T(N_TITIN) * (bF * d * d / L2 + F * ad * ad / L3));
  tt_d = (float)(T(N_TITIN) * ad * d * (bF / L2 - F / L3));
}

// A site's list of bound heads, relinked into increasing head order.  The
// heads insert themselves with atomicExch in the order their waves happen
// to arrive, which varies from run to run, and the solve sums a site's
// partner forces and stiffness terms in list order: unsorted, the solved
// positions would vary in their last bits between identical runs.  Run by
// the lane that owns the site node, the only lane that walks the list
// during the solve; lists hold a few heads.
template <bool REF64, int GW>
__device__ __forceinline__ void sort_site_list(const Shared<REF64, GW>& sh, int site) {
  int j = sh.shead[site];
  if (j < 0 || sh.xnext[j] < 0) return;
  int first = -1;
  while (j >= 0) {
    const int nx = sh.xnext[j];
    if (first < 0 || j < first) {
      sh.xnext[j] = (short)first;
      first = j;
    } else {
      int q = first;
      for (int r = sh.xnext[q]; r >= 0 && r < j; r = sh.xnext[q]) q = r;
      sh.xnext[j] = sh.xnext[q];
      sh.xnext[q] = (short)j;
    }
    j = nx;
  }
  sh.shead[site] = first;
}

// Partners of the slot nodes for the current bindings (start of a solve).
template <int MS, bool REF64, int GW>
__device__ __forceinline__ void build_partners(const Shared<REF64, GW>& sh, const Dims& dm, const Slots<MS>& sl,
                                               Partners<MS>& pt) {
  constexpr int PM = Partners<MS>::PM;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    pt.overflow[s] = false;
#pragma unroll
    for (int h = 0; h < Partners<MS>::P; ++h) { pt.pk[s][h] = -1; pt.k[s][h] = 0.0f; }
    const int n = sl.node[s];
    if (n < 0) continue;
    if (!sl.thick[s]) sort_site_list(sh, n - dm.nk);
    if (PM == 0) {
      pt.overflow[s] = !sl.thick[s];  // sites walk their whole list
      continue;
    }
    if (sl.thick[s]) {
      // the crown's heads: all loads first (h past n_xb_per_crown reads
      // head 0 and is discarded), and the node index kept from being
      // hoisted out of the time loop (and spilled)
      int n_ = n;
      asm volatile("" : "+v"(n_));
      constexpr int PQ = Partners<MS>::P;
      int site[PQ], st[PQ];
#pragma unroll
      for (int h = 0; h < PQ; ++h) {
        const int j = n_ * dm.nx + (h < dm.nx ? h : 0);
        site[h] = sh.xsite[j];
        st[h] = sh.xst[j];
      }
#pragma unroll
      for (int h = 0; h < PM; ++h)
        if (h < dm.nx && site[h] >= 0 && st[h] >= 1 && st[h] <= 3) pt.pk[s][h] = (site[h] << 16) | (n_ * dm.nx + h);
    } else {
      int j = sh.shead[n - dm.nk];
#pragma unroll
      for (int h = 0; h < PM; ++h) {
        if (j < 0) break;
        pt.pk[s][h] = ((j / dm.nx) << 16) | j;
        j = sh.xnext[j];
      }
      pt.overflow[s] = j >= 0;
    }
  }
}

// Dynamic lattice spacing: the spacing d is one more unknown, and the
// Jacobian gains a border (solver.py:552-656).  All members are uniform
// across the workgroup (every lane computes them from the same reductions).
struct Radial {
  bool on;      // this instance solves for d
  float K;      // K_lat, already multiplied by n_thick (simulation.py:1002-1003)
  float dref;   // d_ref = ls (z0 / z)^nu (simulation.py:694-696)
  float frad;   // radial residual at the last residual evaluation
  float jdd;    // d F_rad / d d at the last residual evaluation
  float dinv;   // d-block preconditioner -1 / J_dd, taken at the solve's start
};

// Residual at the current positions: per-XB forces/stiffnesses (XB phase),
// then per-node passive + XB sums (node phase), which also refreshes the
// partners' stiffnesses.  Returns max|F| over the workgroup; f[] receives the
// lane's slot residuals, sh.ktip each thick filament's titin tip stiffness.
// With dynamic spacing the XB phase also forms d f / d d per head (kept in
// the CG-direction buffer, idle here) and the radial sums; the node phase
// gathers the border column cb = d F_axial / d d = d F_rad / d x, and the
// radial residual joins the max.
template <int BS, int MS, bool REF64, bool DYN, typename T, int GW>
__device__ __forceinline__ float residual(const Shared<REF64, GW>& sh, const Dims& dm, const Slots<MS>& sl,
                                          const float (&rA)[MS], const float (&rB)[MS], T zl, T d,
                                          float (&f)[MS], float (&cb)[MS], Partners<MS>& pt,
                                          Radial& rad, Red<BS>& rd, int tid, int lane, int wave,
                                          bool first) {
  const LDS float* prm = sh.prm;
  const bool dyn = DYN && rad.on;
  const auto fdx = sh.pv;  // [X] d f / d d per head, dynamic spacing only
  T sg = T(0), sgd = T(0);
  for (int j = tid; j < dm.X; j += BS) {
    const int site = sh.xsite[j];
    T fj = T(0);
    float kj = 0.0f, fd = 0.0f;
    const int st = sh.xst[j];
    if (site >= 0 && st >= 1 && st <= 3) {
      const int crown = j / dm.nx;
      const T xd = T(posd<REF64>(sh, site)) - T(posd<REF64>(sh, crown));
      const Springs sp = springs(prm, st);
      if constexpr (GW == 0) {
        if (dyn) {
          T g;
          float gd;
          xb_force_radial<T>(sp, xd, d, xb_angle<T, true>(d, xd), fj, kj, fd, g, gd);
          sg += g;
          sgd += T(gd);
        } else {
          xb_force<T>(sp, xd, d, xb_angle<T, false>(d, xd), fj, kj);
        }
      } else {
        const T th = dyn ? xb_angle<T, true>(d, xd) : xb_angle<T, false>(d, xd);
        xb_force<T>(sp, xd, d, th, fj, kj);
        if (dyn) {
          T g;
          float gd;
          xb_radial<T>(sp, xd, d, th, fd, g, gd);
          sg += g;
          sgd += T(gd);
        }
      }
    }
    if constexpr (REF64) sh.fx64[j] = (double)fj; else sh.fx[j] = (float)fj;
    sh.kx[j] = kj;
    if (dyn) fdx[j] = fd;
  }
  __syncthreads();
  const T kk = T(prm[P_THICK_K]), kn = T(prm[P_THIN_K]);
  float mx = 0.0f;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    f[s] = 0.0f;
    cb[s] = 0.0f;
    const int n = sl.node[s];
    if (n < 0) continue;
    const int i = sl.li[s];
    const T x0 = T(posd<REF64>(sh, n));
    T xs = T(0);
    float cs = 0.0f;
    constexpr int PM = Partners<MS>::PM;
#pragma unroll
    for (int h = 0; h < PM; ++h) {
      const int pk = pt.pk[s][h];
      if (pk < 0) continue;
      const int j = p_xb(pk);
      if constexpr (REF64) xs += T(sh.fx64[j]); else xs += T(sh.fx[j]);
      pt.k[s][h] = sh.kx[j];
      if (dyn) cs += fdx[j];
    }
    if (PM == 0 && sl.thick[s]) {
      // the crown's own heads (force and d f / d d are 0 for unbound ones)
      for (int h = 0; h < dm.nx; ++h) {
        const int j = n * dm.nx + h;
        if constexpr (REF64) xs += T(sh.fx64[j]); else xs += T(sh.fx[j]);
        if (dyn) cs += fdx[j];
      }
    }
    T net;
    if (sl.thick[s]) {
      const T xm = i > 0 ? T(posd<REF64>(sh, n - 1)) : T(0);
      const T s_lo = (x0 - xm - T(rA[s])) * kk;
      T s_hi;
      if (i < dm.nc - 1) {
        s_hi = (T(posd<REF64>(sh, n + 1)) - x0 - T(rB[s])) * kk;
      } else {
        float kt;
        T en;
        T tr = T(0);
        float trd = 0.0f, ttd = 0.0f;
        if constexpr (GW == 0) titin<T, DYN, (MS >= 3)>(prm, x0, zl, d, s_hi, kt, en, &tr, &trd, &ttd);
        else titin<T, false, (MS >= 3)>(prm, x0, zl, d, s_hi, kt, en);
        sh.ktip[sl.fil[s]] = kt;
        if constexpr (REF64) {
          // the tip's titin energy for the metrics: the step's first
          // evaluation is at the pre-solve positions, its last at the solved
          // ones (same lane, read back by it after the solve)
          LDS double* fv = sh.fil + sl.fil[s] * NF;
          fv[F_TIT] = (double)en;
          if (first) fv[F_EO] = (double)en;
        }
        if (dyn) {
          if constexpr (GW > 0) titin_radial<T, (MS >= 3)>(prm, x0, zl, d, tr, trd, ttd);
          sg += tr;
          sgd += T(trd);
          cs += ttd;
        }
      }
      net = s_hi - s_lo;
      net += xs;
      cb[s] = cs;
      if (GW && dyn) sh.cbn[n] = cs;
    } else {
      const T xp = i < dm.ns - 1 ? T(posd<REF64>(sh, n + 1)) : zl;
      const T s_hi = (xp - x0 - T(rA[s])) * kn;
      const T s_lo = i > 0 ? (x0 - T(posd<REF64>(sh, n - 1)) - T(rB[s])) * kn : T(0);
      net = s_hi - s_lo;
      if (pt.overflow[s]) {
        int j = sh.shead[n - dm.nk];
        for (int h = 0; h < PM; ++h) j = sh.xnext[j];
        for (; j >= 0; j = sh.xnext[j]) {
          if constexpr (REF64) xs += T(sh.fx64[j]); else xs += T(sh.fx[j]);
          if (dyn) cs += fdx[j];
        }
      }
      net -= xs;
      cb[s] = -cs;
      if (GW && dyn) sh.cbn[n] = -cs;
    }
    f[s] = (float)net;  // the CG right-hand side is FP32
    // max|F| is taken in float: it is compared with the stopping tolerance
    // and reported as a float32 metric.
    mx = fmaxf(mx, fabsf((float)net));
  }
  if (dyn) {
    // F_rad = -K (d - d_ref) - sum g - sum tr (solver.py:502-528);
    // J_dd = -K - sum g_d - sum tr_d.  Summed in the residual's type, in a
    // fixed wave order, so every lane gets the same bits.
    const T a = wave_sum(sg, lane), c = wave_sum(sgd, lane);
    if (lane == 0) { sh.wpart[wave * NWP + W_RG] = (double)a; sh.wpart[wave * NWP + W_RGD] = (double)c; }
    __syncthreads();
    T ga = T(0), gc = T(0);
    for (int w = 0; w < BS / WAVE; ++w) { ga += T(sh.wpart[w * NWP + W_RG]); gc += T(sh.wpart[w * NWP + W_RGD]); }
    const T fr = -T(rad.K) * (d - T(rad.dref)) - ga;
    rad.frad = (float)fr;
    rad.jdd = (float)(-T(rad.K) - gc);
    mx = fmaxf(mx, fabsf(rad.frad));
  }
  return wg_max<BS, (MS > 2)>(rd, mx, lane, wave);
}

// -J p for the slot nodes, p read from the shared vector buffer pv.
// Branch-free per lane: a missing partner reads the node's own entry with
// stiffness 0, neighbour indices are clamped into range and the chain ends
// selected, so the wave runs one straight path (the filament type is
// wave-uniform; only the rare over-full site list takes a branch).
template <int MS, bool REF64, int GW, typename V>
__device__ __forceinline__ void neg_jvp(const Shared<REF64, GW>& sh, const Dims& dm, const Slots<MS>& sl,
                                        const Partners<MS>& pt, float kk, float kn,
                                        V pv, float (&Ap)[MS]) {
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    const int n0 = sl.node[s];
    const int n = n0 >= 0 ? n0 : 0;
    const int i = sl.li[s];
    const float p0 = pv[n];
    float q = 0.0f;  // sum over partners of k (p_partner - p0)
    constexpr int PM = Partners<MS>::PM;
    // the partners' entries are gathered first, all loads in flight
    // together; the empty asm consumes the first and last, so the compiler
    // cannot interleave the loads with the sums (it would, to save
    // registers, serialising four LDS round trips)
    float pp[Partners<MS>::P];
#pragma unroll
    for (int h = 0; h < PM; ++h) {
      const int pk = pt.pk[s][h];
      pp[h] = pv[pk >= 0 ? p_node(pk) : n];
    }
    if constexpr (PM > 0) asm volatile("" ::"v"(pp[0]), "v"(pp[PM - 1]));
#pragma unroll
    for (int h = 0; h < PM; ++h) q += pt.k[s][h] * (pp[h] - p0);
    if (PM == 0 && sl.thick[s]) {
      // the crown's own heads (stiffness 0 for unbound ones)
      for (int h = 0; h < dm.nx; ++h) {
        const int j = n * dm.nx + h;
        const int site = sh.xsite[j];
        q += sh.kx[j] * (pv[site >= 0 ? site : n] - p0);
      }
    }
    const float pprev = pv[n > 0 ? n - 1 : 0];
    float net;
    if (sl.thick[s]) {
      const bool last = i >= dm.nc - 1;
      const float pnext = pv[last ? n : n + 1];
      const float lo = kk * (p0 - (i > 0 ? pprev : 0.0f));
      const float hi = last ? sh.ktip[sl.fil[s]] * p0 : kk * (pnext - p0);
      net = hi - lo + q;
    } else {
      const bool last = i >= dm.ns - 1;
      const float pnext = pv[last ? n : n + 1];
      const float hi = kn * ((last ? 0.0f : pnext) - p0);
      const float lo = i > 0 ? kn * (p0 - pprev) : 0.0f;
      if (pt.overflow[s]) {
        int j = sh.shead[n - dm.nk];
        for (int h = 0; h < PM; ++h) j = sh.xnext[j];
        for (; j >= 0; j = sh.xnext[j]) q += sh.kx[j] * (pv[j / dm.nx] - p0);
      }
      net = hi - lo + q;
    }
    Ap[s] = n0 >= 0 ? -net : 0.0f;
  }
}

// Preconditioned CG on A = -J from x0 = 0, n_cg iterations (solver.py:329-366).
//
// Reference mode runs the textbook recurrence the reference unrolls: two
// reductions (rz, pAp) and three barriers per iteration.  Refined mode runs
// the Chronopoulos-Gear form of the same method (identical iterates in exact
// arithmetic): u = M r and w = A u are formed first and (r.u, w.u) reduced
// together, so each iteration costs one fused reduction and two barriers;
// p and s = A p follow from u, w by the beta recurrence.  Refined mode only
// needs the correction to the accuracy the Newton step can use: it stops
// once r.u / r0.u0 <= eta2 (an inexact-Newton forcing term set by the
// caller), or as soon as the ratio of the last two r.u predicts that the
// update just made reaches it, and always once r.u reaches FP32 rounding.
// (The pipelined variant that shares one barrier between the reduction and
// the matvec was tried: its recurred residual drifts from the true one once
// converged, and passive instances, which converge in one or two passes,
// then lose accuracy.)
//
// With dynamic spacing the system is bordered: the unknown d (uniform, held
// by every lane) couples to the nodes through the column cb, so
// A [v; v_d] = [-J v - cb v_d; -(cb . v) - J_dd v_d], and the d block of the
// preconditioner is rad.dinv (solver.py:531-549).  cb . v joins the
// iteration's existing reduction.
template <int BS, int MS, bool REF64, bool DYN, int GW>
__device__ __forceinline__ void pcg(const Shared<REF64, GW>& sh, const Dims& dm, const Slots<MS>& sl,
                                    const Partners<MS>& pt, const float (&b)[MS], const float (&cb)[MS], const Radial& rad, float b_d, int n_cg,
                                    float eta2, float (&x)[MS], float& x_d, Red<BS>& rd, int lane,
                                    int wave) {
  const float kk = sh.prm[P_THICK_K], kn = sh.prm[P_THIN_K];
  const float ikk = 1.0f / kk, ikn = 1.0f / kn;
  const bool dyn = DYN && rad.on;
  x_d = 0.0f;
  if constexpr (REF64) {
    float r[MS], u[MS], w[MS], p[MS], q[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) { r[s] = b[s]; x[s] = 0.0f; p[s] = 0.0f; q[s] = 0.0f; }
    float r_d = dyn ? b_d : 0.0f, p_d = 0.0f, q_d = 0.0f;
    float gamma = 0.0f, alpha = 1.0f, gamma0 = 0.0f;
    Slots<MS> sl_ = sl;
    for (int it = 0;; ++it) {
      if constexpr (GW || MS >= 5) {
        // Hide the node indices from loop-invariant hoisting: with 5-12 node
        // slots per lane the compiler would otherwise keep per-slot
        // addresses, masks and scale factors live across iterations and
        // spill them, reloading each one every pass; recomputing them from
        // the index is a few VALU operations.
#pragma unroll
        for (int s = 0; s < MS; ++s) asm volatile("" : "+v"(sl_.node[s]), "+v"(sl_.li[s]));
      }
      const Slots<MS>& sl = sl_;
      precond<MS>(r, u, sl, lane, ikk, ikn);
      const float u_d = rad.dinv * r_d;
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (sl.node[s] >= 0) sh.pv[sl.node[s]] = u[s];
      __syncthreads();
      neg_jvp<MS, REF64>(sh, dm, sl, pt, kk, kn, sh.pv, w);
      float g = 0.0f, dl = 0.0f, cu = 0.0f;
      if (dyn) {
#pragma unroll
        for (int s = 0; s < MS; ++s) {
          const float c = GW ? (sl.node[s] >= 0 ? sh.cbn[sl.node[s]] : 0.0f) : cb[s];
          w[s] -= c * u_d;
          cu += c * u[s];
        }
      }
#pragma unroll
      for (int s = 0; s < MS; ++s) { g += r[s] * u[s]; dl += w[s] * u[s]; }
      float w_d = 0.0f;
      if (dyn) {
        wg_sum3<BS, (MS > 2)>(rd, g, dl, cu, lane, wave);
        w_d = -cu - rad.jdd * u_d;
        g += r_d * u_d;
        dl += w_d * u_d;
      } else {
        wg_sum2<BS, (MS > 2)>(rd, g, dl, lane, wave);
      }
      // converged to FP32 rounding (r.u ~ |r|^2 has fallen 14 orders) or
      // broken down: further passes only add rounding noise (every lane sees
      // the same reduced g, so the exit is uniform)
      if (it == 0) gamma0 = g;
      if (!(g > eta2 * gamma0)) break;
      // r.u shrinks by a near-constant factor per pass: predict the next one
      const bool last = eta2 > 1e-14f && it > 0 && g * g <= eta2 * gamma0 * gamma;
      // the recurrence coefficients take the hardware reciprocal (1 ulp):
      // CG only has to produce a correction, the FP64 residual checks it
      const float beta = it == 0 ? 0.0f : g * __builtin_amdgcn_rcpf(gamma + 1e-30f);
      const float den = it == 0 ? dl : dl - beta * g * __builtin_amdgcn_rcpf(alpha);
      alpha = g * __builtin_amdgcn_rcpf(den + 1e-30f);
      gamma = g;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        p[s] = u[s] + beta * p[s];
        q[s] = w[s] + beta * q[s];
        x[s] += alpha * p[s];
        r[s] -= alpha * q[s];
      }
      if (dyn) {
        p_d = u_d + beta * p_d;
        q_d = w_d + beta * q_d;
        x_d += alpha * p_d;
        r_d -= alpha * q_d;
      }
      if (last || it >= n_cg - 1) break;
    }
    return;
  }
  float r[MS], z[MS], p[MS], Ap[MS];
#pragma unroll
  for (int s = 0; s < MS; ++s) { r[s] = b[s]; x[s] = 0.0f; p[s] = 0.0f; }
  float r_d = dyn ? b_d : 0.0f, p_d = 0.0f;
  float rz = 0.0f;
  for (int it = 0;; ++it) {
    precond<MS>(r, z, sl, lane, ikk, ikn);
    const float z_d = rad.dinv * r_d;
    float loc = 0.0f;
#pragma unroll
    for (int s = 0; s < MS; ++s) loc += r[s] * z[s];
    float rz_new = wg_sum<BS, (MS > 2)>(rd, loc, lane, wave);
    if (dyn) rz_new += r_d * z_d;
    const float beta = it == 0 ? 0.0f : rz_new / (rz + 1e-30f);
    rz = rz_new;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      p[s] = it == 0 ? z[s] : z[s] + beta * p[s];
      if (sl.node[s] >= 0) sh.pv[sl.node[s]] = p[s];
    }
    p_d = it == 0 ? z_d : z_d + beta * p_d;
    __syncthreads();
    neg_jvp<MS, REF64>(sh, dm, sl, pt, kk, kn, sh.pv, Ap);
    loc = 0.0f;
    float pAp, Ap_d = 0.0f;
    if (dyn) {
      float cp = 0.0f;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        const float c = GW ? (sl.node[s] >= 0 ? sh.cbn[sl.node[s]] : 0.0f) : cb[s];
        Ap[s] -= c * p_d;
        cp += c * p[s];
        loc += p[s] * Ap[s];
      }
      wg_sum2<BS, (MS > 2)>(rd, loc, cp, lane, wave);
      Ap_d = -cp - rad.jdd * p_d;
      pAp = loc + p_d * Ap_d;
    } else {
#pragma unroll
      for (int s = 0; s < MS; ++s) loc += p[s] * Ap[s];
      pAp = wg_sum<BS, (MS > 2)>(rd, loc, lane, wave);
    }
    const float alpha = rz / (pAp + 1e-30f);
#pragma unroll
    for (int s = 0; s < MS; ++s) { x[s] += alpha * p[s]; r[s] -= alpha * Ap[s]; }
    if (dyn) {
      x_d += alpha * p_d;
      r_d -= alpha * Ap_d;
    }
    if (it >= n_cg - 1) break;
  }
}

// Equilibrium solve at Z-line zl (solver.py:663-771).  Fixed spacing: d is
// the given spacing.  Dynamic spacing (rad.on): d is the initial guess and
// returns the solved spacing, clamped at >= 1 nm after every Newton step
// (solver.py:599-656).
// Reference mode: FP32 residual, stop at max|F| <= tol or n_newton.
// Refined mode: FP64 residual, FP32 CG correction, stop at refine_tol or
// n_refine iterations; past n_refine it continues (up to 4 n_refine) while
// each iteration still halves max|F|, so a large Z-line jump (the first step
// after the initial equilibration at the parameter row's Z-line) converges
// and a residual stuck at a clamp does not run on.
template <int BS, int MS, bool REF64, bool DYN, int GW>
__device__ __forceinline__ void solve(const Shared<REF64, GW>& sh, const Dims& dm, const Slots<MS>& sl,
                                      const float (&rA)[MS], const float (&rB)[MS], float zl,
                                      typename std::conditional<REF64, double, float>::type& d,
                                      const mf_solver_cfg& cfg, Radial& rad, Red<BS>& rd, int tid, int lane,
                                      int wave, int& iters, float& resid) {
  using T = typename std::conditional<REF64, double, float>::type;
  float dx[MS], fb[MS], cb[MS];
  float tol;
  int cap;
  if constexpr (REF64) {
    tol = cfg.refine_tol;
    cap = cfg.n_refine;
  } else {
    const float t0 = isnan(cfg.tol) ? sh.prm[P_SOLVER_TOL] : cfg.tol;
    tol = fmaxf(t0, fmaxf(sh.prm[P_THICK_K] * 1e-4f, 0.25f));
    cap = cfg.n_newton;
  }
  Partners<MS> pt;
  build_partners<MS, REF64>(sh, dm, sl, pt);
  int it = 0;
  float mx, prev = INFINITY;
  for (;;) {
    mx = residual<BS, MS, REF64, DYN, T>(sh, dm, sl, rA, rB, T(zl), d, fb, cb, pt, rad, rd, tid, lane,
                                         wave, it == 0);
    // the d-block preconditioner uses J_dd at the initial state (solver.py:633-638)
    if (DYN && rad.on && it == 0) rad.dinv = -1.0f / rad.jdd;
    if (!(mx > tol)) break;
    if (it >= cap && !(REF64 && it < 4 * cap && mx < 0.5f * prev)) break;
    prev = mx;
    float dd;
    // forcing term: the step's linear error (max-norm ~ 10 eta |F|) well
    // below the tolerance, but no tighter than what the nonlinearity leaves
    // after one Newton step anyway (quadratic: ~1e-5 |F| relative, |F| in
    // pN); FP32 rounding bounds it from below
    float eta2 = 1e-14f;
    if (REF64 && cfg.cg_forcing > 0.0f) {
      const float eta = fmaxf(cfg.cg_forcing * tol / mx, 1e-5f * mx);
      eta2 = fmaxf(eta * eta, 1e-14f);
    }
    pcg<BS, MS, REF64, DYN>(sh, dm, sl, pt, fb, cb, rad, rad.frad, cfg.n_cg, eta2, dx, dd, rd, lane,
                            wave);
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      const int n = sl.node[s];
      if (n < 0) continue;
      const float v = isfinite(dx[s]) ? dx[s] : 0.0f;
      if constexpr (REF64) sh.x64[n] += (double)v; else sh.x32[n] += v;
    }
    if (DYN && rad.on) {
      d += T(isfinite(dd) ? dd : 0.0f);
      d = d > T(1) ? d : T(1);
    }
    __syncthreads();
    ++it;
  }
  iters = it;
  resid = mx;
}

// Nearest reachable site (geometry.py:60-175): search from crown+13, mask
// padding slots and sites at x <= 0, first minimum wins; returns the site
// index along the thin filament and the float32 axial offset from the crown.
// Candidates are gathered in chunks of NEAR_CHUNK so a chunk's LDS loads are
// in flight together without holding all W positions in registers.
constexpr int NEAR_CHUNK = 8;
template <bool REF64, int GW>
__device__ __forceinline__ void nearest_site(const Shared<REF64, GW>& sh, const Dims& dm, int slot, int thin,
                                             int crown, int& nearest, float& xdist, const LDS float* fp,
                                             int Wp) {
  const float base = posf<REF64>(sh, crown);
  const float head = base + XB_REACH;
  const int cnt = sh.fcount[slot];
  const auto cs = sh.fsites + slot * dm.W;
  const int tb = dm.nk + thin * dm.ns;
  float best = INFINITY;
  int bj = 0;
  if (fp) {
    // the face's candidate positions, masked entries already +inf
    typedef float f4 __attribute__((ext_vector_type(4)));
    const LDS f4* row = (const LDS f4*)(fp + slot * Wp);
#pragma unroll 2
    for (int q = 0; q < Wp / 4; ++q) {
      const f4 v = row[q];
      const float d0 = fabsf(v.x - head), d1 = fabsf(v.y - head), d2 = fabsf(v.z - head), d3 = fabsf(v.w - head);
      if (d0 < best) { best = d0; bj = 4 * q; }
      if (d1 < best) { best = d1; bj = 4 * q + 1; }
      if (d2 < best) { best = d2; bj = 4 * q + 2; }
      if (d3 < best) { best = d3; bj = 4 * q + 3; }
    }
    nearest = cs[bj];
    xdist = posf<REF64>(sh, tb + nearest) - base;
    return;
  }
  for (int j0 = 0; j0 < dm.W; j0 += NEAR_CHUNK) {
    float sp[NEAR_CHUNK];
#pragma unroll
    for (int c = 0; c < NEAR_CHUNK; ++c) sp[c] = j0 + c < dm.W ? posf<REF64>(sh, tb + cs[j0 + c]) : 0.0f;
#pragma unroll
    for (int c = 0; c < NEAR_CHUNK; ++c) {
      const float dd = fabsf(sp[c] - head);
      const bool ok = (j0 + c < cnt) && (sp[c] > 0.0f);
      if (ok && dd < best) { best = dd; bj = j0 + c; }
    }
  }
  nearest = cs[bj];
  xdist = posf<REF64>(sh, tb + nearest) - base;
}

// Float positions of each (thin, face) candidate list, row stride Wp (W
// rounded up to 4): the sites nearest_site compares, with padding slots and
// sites at x <= 0 (geometry.py:113-116) stored as +inf so the search needs
// no masks.
template <bool REF64, int GW>
__device__ __forceinline__ void face_positions(const Shared<REF64, GW>& sh, const Dims& dm, LDS float* fp, int Wp,
                                               int tid, int nthreads) {
  const int n = dm.Tn * dm.nf * Wp;
  for (int e = tid; e < n; e += nthreads) {
    const int slot = e / Wp, j = e - slot * Wp;
    float v = INFINITY;
    if (j < sh.fcount[slot]) {
      const float x = posf<REF64>(sh, dm.nk + (slot / dm.nf) * dm.ns + sh.fsites[slot * dm.W + j]);
      v = x > 0.0f ? x : INFINITY;
    }
    fp[e] = v;
  }
}

// clip(digitize(x, edges) - 1, 0, nb - 1) (transitions.py:915-917): the
// number of edges <= x, minus one; NaN sorts after every edge.  The edges
// are a float32 linspace, so the affine estimate is off by at most one bin;
// two edge loads settle it.
__device__ __forceinline__ int bin_of(const LDS float* edges, int nb, float inv_w, float x) {
  if (isnan(x)) return nb - 1;
  const float t = fminf(fmaxf((x - edges[0]) * inv_w, 0.0f), (float)(nb - 1));
  int i = (int)t;
  const float a = edges[i], c = edges[i + 1];
  if (x < a) --i;
  else if (x >= c) ++i;
  return i < 0 ? 0 : (i > nb - 1 ? nb - 1 : i);
}

// Add popcount(pred over the wave) to an LDS counter (one atomic per wave).
// Lanes of the wave for which pred holds (wave-uniform result).
__device__ __forceinline__ int wave_count(bool pred) { return __popcll(__ballot(pred)); }

template <int BS, int MS, int MX, bool REF64, bool DYN, int GW>
__global__ __launch_bounds__(BS) void run_kernel(const mf_launch L) {
  extern __shared__ __align__(16) char smem[];
  constexpr int NW = BS / WAVE;
  // the wave index is uniform: readfirstlane keeps it, and everything derived
  // from it (chain ownership), in scalar registers
  const int tid = threadIdx.x, lane = tid & (WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(tid / WAVE);
  const int b = blockIdx.x;
  // descriptors are read in place (scalar loads); a by-value copy of these
  // structs would live in scratch
  const mf_instance& I = L.instances[b];
  if (I.n_steps <= 0 && !(I.flags & MF_INST_INIT)) return;
  const mf_lattice& G = L.lattices[I.lattice];
  Dims dm;
  dm.Tk = G.n_thick; dm.nc = G.n_crowns; dm.Tn = G.n_thin; dm.ns = G.n_sites; dm.nx = G.n_xb_per_crown;
  dm.W = G.width; dm.nf = G.n_faces; dm.nb = G.n_bins;
  dm.nk = dm.Tk * dm.nc; dm.N = dm.nk + dm.Tn * dm.ns; dm.X = dm.nk * dm.nx; dm.S = dm.Tn * dm.ns;
  dm.nch = dm.Tk + dm.Tn;
  dm.ktm = table_sets(L, MF_POP_TM);
  dm.kxb = table_sets(L, MF_POP_XB);
  // GW variants: this instance's slice of the device workspace
  char* ws = GW ? (char*)L.workspace + (int64_t)b * L.workspace_stride : nullptr;
  Shared<REF64, GW> sh = carve<BS, REF64, GW>((LDS char*)smem, ws, dm);
  Red<BS> rd{sh.red, 0};

  const float* lf = L.lat_f32;
  const int32_t* li = L.lat_i32;
  const float* prm_g = L.params + (int64_t)b * MF_N_PARAMS;
  const float* drv = L.drivers + (int64_t)b * L.t_stride * 3;
  float* tab = L.tables + I.table_off;
  const StateLayout SL = state_layout(dm.Tk, dm.nc, dm.Tn, dm.ns, dm.nx);
  uint8_t* stb = L.state + I.state_off;
  int32_t* st_near = (int32_t*)(stb + SL.xb_near);
  float* st_xdist = (float*)(stb + SL.xb_xdist);
  const int S = dm.S, X = dm.X, nb = dm.nb;

  // ---- static per-workgroup data ------------------------------------------
  for (int i = tid; i < MF_N_PARAMS; i += BS) sh.prm[i] = prm_g[i];
  for (int i = tid; i <= nb; i += BS) sh.edges[i] = lf[G.bin_edges + i];
  for (int i = tid; i < dm.ktm * 27 * 16; i += BS) sh.tmcum[i] = tab[TAB_TM + i];
  for (int i = tid; i < dm.Tn * dm.nf * dm.W; i += BS) sh.fsites[i] = (short)li[G.face_sites + i];
  for (int i = tid; i < dm.Tn * dm.nf; i += BS) sh.fcount[i] = (short)li[G.face_count + i];
  for (int i = tid; i < S; i += BS) {
    const int k = i % dm.ns;  // host checks |neighbour - k| <= 127
    sh.tprev[i] = (signed char)(li[G.tm_prev + i] - k);
    sh.tnext[i] = (signed char)(li[G.tm_next + i] - k);
  }
  for (int i = tid; i < NC * NW; i += BS) sh.cnt[i] = 0;
  for (int i = tid; i < 2 * 64 + 2 + 40; i += BS) sh.need[i] = 0;

  // per-XB (thin, face) slot with the target-valid flag; displacement origins
  for (int j = tid; j < X; j += BS)
    sh.xslot[j] = (short)(li[G.xb_slot + j] | (li[G.xb_valid + j] != 0 ? 0x8000 : 0));
  for (int i = tid; i < dm.N; i += BS)
    sh.org[i] = i < dm.nk ? lf[G.crown_offsets + i] : lf[G.thin_rest_cum + (i - dm.nk)];

  // chain-major slot map
  Slots<MS> sl;
  {
    int base = 0, k = 0;
#pragma unroll
    for (int s = 0; s < MS; ++s) { sl.node[s] = -1; sl.li[s] = 0; sl.chk[s] = -1; sl.fil[s] = 0; sl.thick[s] = false; }
    for (int c = wave; c < dm.nch; c += NW, ++k) {
      const bool th = c < dm.Tk;
      const int len = th ? dm.nc : dm.ns;
      const int per = (len + WAVE - 1) / WAVE;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        const int j = s - base;
        if (j >= 0 && j < per) {
          const int loc = lane * per + j;
          sl.chk[s] = k;
          sl.thick[s] = th;
          sl.fil[s] = th ? c : c - dm.Tk;
          sl.li[s] = loc;
          sl.node[s] = loc < len ? (th ? c * dm.nc + loc : dm.nk + (c - dm.Tk) * dm.ns + loc) : -1;
        }
      }
      base += per;
    }
    sl.nch = k;
    // the first chain's shape, uniform per wave (readfirstlane: scalar
    // registers, so the single-chain scan below branches on them for free)
    {
      const int c0 = wave < dm.nch ? wave : 0;
      const bool th0 = c0 < dm.Tk;
      sl.per0 = __builtin_amdgcn_readfirstlane(((th0 ? dm.nc : dm.ns) + WAVE - 1) / WAVE);
      sl.thick0 = __builtin_amdgcn_readfirstlane(th0 ? 1 : 0) != 0;
    }
  }
  // per-slot rests: thick rA = rest_i, rB = rest_{i+1}; thin rA = rest_i, rB = rest_{i-1}
  float rA[MS], rB[MS];
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    rA[s] = rB[s] = 0.0f;
    const int n = sl.node[s];
    if (n < 0) continue;
    const int i = sl.li[s];
    if (sl.thick[s]) {
      const float* cr = lf + G.crown_rests + (int64_t)sl.fil[s] * dm.nc;
      rA[s] = cr[i];
      rB[s] = i < dm.nc - 1 ? cr[i + 1] : 0.0f;
    } else {
      const float* br = lf + G.binding_rests + (int64_t)sl.fil[s] * dm.ns;
      rA[s] = br[i];
      rB[s] = i > 0 ? br[i - 1] : 0.0f;
    }
  }

  // ---- load or realize the state -----------------------------------------
  const bool init = (I.flags & MF_INST_INIT) != 0;
  double ls_carry;
  if (!init) {
    const double* pos = (const double*)(stb + SL.pos);
    for (int i = tid; i < dm.N; i += BS) {
      if constexpr (REF64) sh.x64[i] = pos[i]; else sh.x32[i] = (float)pos[i];
    }
    const int32_t* xbb = (const int32_t*)(stb + SL.xb_bound);
    const int32_t* xbk = (const int32_t*)(stb + SL.xb_key);
    const uint8_t* xbs = stb + SL.xb_state;
    for (int j = tid; j < X; j += BS) {
      const int bt = xbb[j];
      sh.xsite[j] = (short)(bt >= 0 ? dm.nk + li[G.xb_thin + j] * dm.ns + bt : -1);
      sh.xkey[j] = (short)xbk[j];
      sh.xst[j] = xbs[j];
    }
    const int32_t* tbd = (const int32_t*)(stb + SL.thin_bound);
    const uint8_t* tms = stb + SL.tm;
    for (int s = tid; s < S; s += BS) { sh.tbound[s] = tbd[s]; sh.tm[s] = tms[s]; }
    ls_carry = *(const double*)(stb + SL.ls);
  } else {
    // realize_state (state.py:191-262) at z0 = drivers[t0]
    const float z0 = drv[(I.t0 + L.t_base) * 3 + 0];
    for (int i = tid; i < dm.N; i += BS) {
      const float v = i < dm.nk ? lf[G.crown_offsets + i] : z0 - lf[G.binding_offsets + (i - dm.nk)];
      if constexpr (REF64) sh.x64[i] = (double)v; else sh.x32[i] = v;
    }
    for (int j = tid; j < X; j += BS) { sh.xsite[j] = -1; sh.xst[j] = 0; }
    for (int s = tid; s < S; s += BS) { sh.tbound[s] = -1; sh.tm[s] = 0; }
    ls_carry = (double)drv[(I.t0 + L.t_base) * 3 + 2];
  }
  for (int s = tid; s < S; s += BS) sh.shead[s] = -1;
  __syncthreads();

  // dynamic lattice spacing (DYN variants, MF_INST_DYNAMIC instances)
  Radial rad;
  rad.on = DYN && (I.flags & MF_INST_DYNAMIC) != 0;
  rad.K = I.K_lat;
  rad.dref = 0.0f;
  rad.frad = 0.0f;
  rad.jdd = -1.0f;
  rad.dinv = 0.0f;
  const float z_trace0 = drv[0];  // l0 of the Poisson reference spacing (simulation.py:676)
  const float* centers = lf + G.bin_centers;

  const uint64_t seed = I.seed;
  const int n_steps = I.n_steps, t_first = I.t0 + L.t_base;
  const float ov_lo = G.overlap_lo, ov_hi = G.overlap_hi;
  int tmc = 0;  // which TM buffer is current
  const float inv_w = (float)nb / (lf[G.bin_edges + nb] - lf[G.bin_edges]);
  const double inv_nk = 1.0 / (double)dm.nk, inv_S = 1.0 / (double)S;
  // XB sampling and absorbing rows: in LDS when the launch's LDS budget
  // holds them (the host sets LAUNCH_TABLES_IN_LDS and sizes the dynamic
  // LDS), else read from the instance's table in device memory.  Read
  // through generic pointers, which address either.
  float* xbtab = tab + tab_xb(dm.ktm);
  float* abstab = tab + tab_abs(nb, dm.ktm, dm.kxb);
  // explicit subpopulations: per-head and per-site population labels
  const uint8_t* xlab = dm.kxb > 1 ? L.pop_labels + I.label_off : nullptr;
  const uint8_t* tlab = dm.ktm > 1 ? L.pop_labels + I.label_off + X : nullptr;
  const Pops pops = pops_of(L, b);
  int64_t lds_extra = lds_offsets<BS, REF64, GW>(dm).total;
  if (L.launch_flags & LAUNCH_TABLES_IN_LDS) {
    LDS float* tl = (LDS float*)(smem + lds_extra);
    const int nt = dm.kxb * (2 * nb * 48 + 2 * nb * 4);
    // the XB rows follow the instance's ktm TM row sets
    const float* src = tab + tab_xb(dm.ktm);
    for (int i = tid; i < nt; i += BS) tl[i] = src[i];
    xbtab = (float*)tl;
    abstab = (float*)(tl + dm.kxb * 2 * nb * 48);
    lds_extra += 4LL * nt;
    __syncthreads();
  }
  // candidate-site positions for the nearest-site search (when they fit)
  const int Wp = (dm.W + 3) & ~3;
  LDS float* fpos =
      (L.launch_flags & LAUNCH_FACE_POS) ? (LDS float*)(smem + ((lds_extra + 15) & ~int64_t(15))) : nullptr;
  const float* U = L.uniforms ? L.uniforms + (int64_t)b * L.t_stride * L.u_stride : nullptr;

  // ---- time loop; step -1 is create_and_equilibrate (simulation.py:633-642)
  // drivers are read one step ahead, so that the load has landed by the
  // time the step needs them; z_prev is the previous step's z_line
  float z_prev = 0.0f, d_z = 0.0f, d_p = 0.0f, d_l = 0.0f;
  if (init || n_steps > 0) { d_z = drv[t_first * 3 + 0]; d_p = drv[t_first * 3 + 1]; d_l = drv[t_first * 3 + 2]; }
  if (n_steps > 0 && t_first > 0) z_prev = drv[(t_first - 1) * 3];
  const int tid_k = tid;
  bool fpos_ready = false;  // fpos already holds this step's candidate positions
  // the split halves of a step (public launch flags, uniform per launch)
  const bool kin_only = (L.launch_flags & MF_LAUNCH_KINETICS_ONLY) != 0;
  // in-kernel table rebuild on driver changes; the tables hold the drivers
  // of row t_first when the launch starts
  // (only the dynamic-spacing variants carry the out-of-line rebuild: a call
  // inside the time loop would change the register allocation of the
  // fixed-spacing kernels; the host runs traces on these variants)
  const bool trace = DYN && (L.launch_flags & MF_LAUNCH_REBUILD) != 0;
  float tab_pca = d_p, tab_ls = d_l;
  // pCa of the XbPre cache (NaN: not filled yet)
  float pre_pca = __builtin_nanf("");
  const bool solve_only = (L.launch_flags & MF_LAUNCH_SOLVE_ONLY) != 0;
  for (int step = init ? -1 : 0; step < n_steps; ++step) {
    // The thread index is made opaque once per step: otherwise the compiler
    // hoists every tid-derived LDS address of the strided per-head and
    // per-site loops out of the time loop and, short of registers, spills
    // them, paying a scratch round trip per use; recomputing an address is
    // one VALU operation.
    int tid = tid_k;
    asm volatile("" : "+v"(tid));
    const bool kin = step >= 0;
    const int t = t_first + (kin ? step : 0);
    const float zl = d_z, pCa = d_p, lsd = d_l;
    if (kin && step + 1 < n_steps) {
      d_z = drv[(t + 1) * 3 + 0]; d_p = drv[(t + 1) * 3 + 1]; d_l = drv[(t + 1) * 3 + 2];
    }
    float z_solve, d_solve;
    // old key and state of each head, for the dynamic-spacing metrics (their
    // absorbing chain is evaluated at the solved spacing)
    int q_kold[MX], q_sold[MX];
    bool chg = false;  // this step rebuilds a fixed-spacing instance's tables
    if (!kin) {
      // Nearest update, then a solve against the parameter row's z_line and
      // lattice_spacing (the reference's run() leaves those at their defaults).
      if (fpos) {
        face_positions(sh, dm, fpos, Wp, tid, BS);
        __syncthreads();
      }
      for (int j = tid; j < X; j += BS) {
        int nr; float xd;
        const int slot = sh.xslot[j] & 0x7fff;
        nearest_site<REF64>(sh, dm, slot, slot / dm.nf, j / dm.nx, nr, xd, fpos, Wp);
        // tm == 0 everywhere: permissiveness 0
        sh.xkey[j] = (short)((xlab ? xlab[j] * 2 * nb : 0) + bin_of(sh.edges, nb, inv_w, xd));
        if (n_steps == 0) { st_near[j] = nr; st_xdist[j] = xd; }
      }
      __syncthreads();
      z_solve = sh.prm[P_ZLINE];
      d_solve = sh.prm[P_LS];
    } else if (solve_only) {
      // solve_equilibrium alone (solver.py:663-771) on the loaded state
      z_prev = zl;
      z_solve = zl;
      d_solve = lsd;
      if (DYN && rad.on) rad.dref = lsd * powf(z_trace0 / zl, I.nu);
      // the site -> bound-head lists the solve walks, which K2 builds in a
      // full step: cleared, then every head with a site inserted
      for (int s2 = tid; s2 < S; s2 += BS) sh.shead[s2] = -1;
      __syncthreads();
      for (int j = tid; j < X; j += BS) {
        const int site = sh.xsite[j];
        if (site >= 0) sh.xnext[j] = (short)atomicExch((int*)(sh.shead + (site - dm.nk)), j);
      }
      __syncthreads();
    } else {
      const float dz = t > 0 ? zl - z_prev : 0.0f;
      z_prev = zl;
      z_solve = zl;
      d_solve = lsd;
      // d_ref = ls (z0 / z)^nu (simulation.py:694-696)
      if (DYN && rad.on) rad.dref = lsd * powf(z_trace0 / zl, I.nu);
      // A fixed-spacing instance whose pCa or spacing differs from the one
      // its tables hold rebuilds them here (the reference forms its
      // generators from each step's drivers, transitions.py:889-912): the TM
      // rows now, the XB rows this step reads after the nearest-site search.
      // (dynamic spacing rebuilds its XB rows every step; its TM rows follow pCa)
      const bool new_pca = trace && pCa != tab_pca;
      chg = trace && !(DYN && rad.on) && (new_pca || lsd != tab_ls);
      if (DYN && (chg || (rad.on && new_pca))) {
        for (int e = tid; e < dm.ktm * 27; e += BS)
          tm_table_entry(pops, e / 27, (double)pCa, (double)L.cfg.dt, e % 27, (float*)sh.tmcum);
        tab_pca = pCa;
        tab_ls = lsd;
        __syncthreads();
      }
      const bool rebuild = DYN && (rad.on || chg);
      // thin.axial += dz (simulation.py:683-687)
      if (dz != 0.0f) {
        for (int i = dm.nk + tid; i < dm.N; i += BS) {
          if constexpr (REF64) sh.x64[i] += (double)dz; else sh.x32[i] += dz;
        }
        __syncthreads();
      }
      if (fpos && !fpos_ready) {
        face_positions(sh, dm, fpos, Wp, tid, BS);
        __syncthreads();
      }
      fpos_ready = false;

      // ---- K1: TM transitions (transitions.py:516-625) + nearest sites -------
      const auto tmo = sh.tm + tmc * S;
      const auto tmn = sh.tm + (tmc ^ 1) * S;
      // device draws: when the sites and the heads both fit the MX strided
      // slots of a thread, the site and head uniforms of slot k come from
      // one Philox block per lane (every thread is active here)
      const bool paired = U == nullptr && S <= MX * BS && X <= MX * BS;
      float u_tm[MX], u_xb[MX];
#pragma unroll
      for (int k = 0; k < MX; ++k) {
        u_tm[k] = u_xb[k] = 0.0f;
        if (paired) philox_uniform_pair(seed, t, tid + k * BS, u_tm[k], u_xb[k]);
      }
      auto tm_site = [&](int s, float u) {
        const int pv = sh.tprev[s], nv = sh.tnext[s];
        const int sp_ = tmo[s + pv], sn_ = tmo[s + nv];
        const bool pr = pv != 0, nr = nv != 0;
        const int n2 = (pr && sp_ == 2) + (nr && sn_ == 2);
        const int n3 = (pr && sp_ == 3) + (nr && sn_ == 3);
        const int ncl = (pr && sp_ <= 1) + (nr && sn_ <= 1);
        const int cfg = 9 * n2 + 3 * n3 + ncl;
        const int st = tmo[s];
        int ns;
        if (st == 3 && sh.tbound[s] >= 0) {
          ns = 3;  // locked open under a bound head: [0, 0, 0, 1]
        } else {
          const LDS float* c = sh.tmcum + ((tlab ? tlab[s] * 27 : 0) + cfg) * 16 + st * 4;
          ns = u < c[0] ? 0 : u < c[1] ? 1 : u < c[2] ? 2 : u < c[3] ? 3 : 0;
        }
        tmn[s] = (unsigned char)ns;
        sh.shead[s] = -1;
      };
      if (paired) {
#pragma unroll
        for (int k = 0; k < MX; ++k) {
          const int s = tid + k * BS;
          if (s < S) tm_site(s, u_tm[k]);
        }
      } else {
        for (int s = tid; s < S; s += BS)
          tm_site(s, U ? U[(int64_t)t * L.u_stride + s] : philox_uniform(seed, t, 0, s));
      }
      // nearest sites, and the candidate table rows of both permissiveness
      // levels fetched ahead of the barrier (K2 only picks one)
      int q_near[MX], q_bin[MX], q_old[MX];
      bool q_occ[MX], q_clr[MX];
      float4 q_r0[MX], q_r1[MX], q_ab[MX];
      float2 q_s0[MX], q_s1[MX];
#pragma unroll
      for (int q = 0; q < MX; ++q) {
        const int j = tid + q * BS;
        q_near[q] = 0; q_bin[q] = 0; q_old[q] = 0; q_occ[q] = false; q_clr[q] = false;
        q_r0[q] = q_r1[q] = q_ab[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        q_s0[q] = q_s1[q] = make_float2(0.f, 0.f);
        if (j < X) {
          const int old = sh.xst[j];
          const int keyold = sh.xkey[j];
          float xd;
          const int slot = sh.xslot[j] & 0x7fff;
          const int thin = slot / dm.nf;
          nearest_site<REF64>(sh, dm, slot, thin, j / dm.nx, q_near[q], xd, fpos, Wp);
          // the launch's last nearest-site update goes to the state record
          // (xb_nearest_bs, xb_distances[..., 0] of the reference's State)
          if (step == n_steps - 1) { st_near[j] = q_near[q]; st_xdist[j] = xd; }
          q_near[q] += thin * dm.ns;  // thin-global site index from here on
          const int bn = bin_of(sh.edges, nb, inv_w, xd);
          const int lb = xlab ? xlab[j] * 2 * nb : 0;
          if (rebuild) {
            // the absorbing row of the head's previous key; its sampling
            // row is marked after this barrier, once its site's new TM
            // state (the permissiveness) is known
            if (old == 2 || old == 3)
              atomicOr((uint32_t*)(sh.need + 64 + (keyold >> 5)), 1u << (keyold & 31));
          } else {
            const float* r0 = xbtab + (int64_t)(lb + bn) * 48 + old * 8;
            const float* r1 = xbtab + (int64_t)(lb + nb + bn) * 48 + old * 8;
            q_r0[q] = *(const float4*)r0; q_s0[q] = *(const float2*)(r0 + 4);
            q_r1[q] = *(const float4*)r1; q_s1[q] = *(const float2*)(r1 + 4);
            if (!(DYN && rad.on) && (old == 2 || old == 3)) q_ab[q] = *(const float4*)(abstab + (int64_t)keyold * 4);
          }
          q_kold[q] = keyold;
          q_sold[q] = old;
          q_bin[q] = lb + bn;
          q_occ[q] = sh.tbound[q_near[q]] >= 0;
          // The owner of this head's own bound site before any K2 write.
          // The reference clears every unbinding site before it sets the
          // new bindings (transitions.py:1203-1219).  A head binds a site
          // only if the site's owner read -1 here, and an unbinding head's
          // clear then finds -1 already: so a clear is stored only when
          // this value is >= 0, no bind can target that site this step, and
          // a store and an atomicMin never meet on one site.
          const int own = sh.xsite[j];
          q_clr[q] = own >= 0 && sh.tbound[own - dm.nk] >= 0;
          q_old[q] = old;
        }
      }
      __syncthreads();
      tmc ^= 1;
      if (DYN && rebuild) {
        // the marked XB rows at the kinetics spacing (dynamic: the carried
        // one, simulation.py:691-693), and for a changed fixed-spacing
        // instance the marked absorbing rows too, then the prefetch
        const double d_kin = (DYN && rad.on) ? ls_carry : (double)lsd;
        if (pCa != pre_pca) {  // (barriers follow before the rebuild reads it)
          fill_xb_pre(pops, (double)pCa, sh.xbpre, tid);
          pre_pca = pCa;
        }
        // the marked keys as compact lists (in the CG buffer, idle during
        // the kinetics), so the busy quads are contiguous
        // (one list per kind, 2 bytes per key; the CG buffer holds 2 N; a
        // second list that does not fit rebuilds every absorbing row)
        unsigned short* ks = (unsigned short*)sh.pv;
        const bool two = 2 * dm.kxb * 2 * nb <= 2 * dm.N;
        unsigned short* ka = two ? ks + dm.kxb * 2 * nb : nullptr;
        // the one sampling row each head reads: its bin at its nearest
        // site's permissiveness after this step's TM transitions (K2's rule)
        {
          const auto tnew = sh.tm + tmc * S;
#pragma unroll
          for (int q = 0; q < MX; ++q) {
            const int j = tid + q * BS;
            if (j < X) {
              const bool perm = (tnew[q_near[q]] == 3) && (sh.xslot[j] < 0);
              const int key = (perm ? nb : 0) + q_bin[q];
              atomicOr((uint32_t*)(sh.need + (key >> 5)), 1u << (key & 31));
            }
          }
        }
        __syncthreads();
        if (wave == 0) {
          const int ns_ = compact_keys(sh.need, ks, lane);
          const int na_ = (chg && two) ? compact_keys(sh.need + 64, ka, lane) : 0;
          if (lane == 0) { sh.need[128] = (uint32_t)ns_; sh.need[129] = (uint32_t)na_; }
        }
        __syncthreads();
        const int n_s = (int)sh.need[128], n_a = (int)sh.need[129];
        // the keys sorted by their squaring count go after the two lists,
        // when the buffer holds a third
        unsigned short* ko = 3 * dm.kxb * 2 * nb <= 2 * dm.N ? ks + 2 * dm.kxb * 2 * nb : nullptr;
        rebuild_rows<BS, GW>(pops, pCa, d_kin, L.cfg.dt, centers, nb, dm.kxb, false, xbtab, abstab, tid,
                             ks, n_s, ko, (uint32_t*)(sh.need + 130), (const double*)sh.xbpre);
        if (chg)
          rebuild_rows<BS, GW>(pops, pCa, d_kin, L.cfg.dt, centers, nb, dm.kxb, true, xbtab, abstab, tid,
                               ka, n_a, nullptr, nullptr, (const double*)sh.xbpre);
        __syncthreads();
        for (int i = tid; i < (chg ? 128 : 64); i += BS) sh.need[i] = 0;
#pragma unroll
        for (int q = 0; q < MX; ++q) {
          const int j = tid + q * BS;
          if (j < X) {
            const int old = q_old[q];
            const float* r0 = xbtab + (int64_t)q_bin[q] * 48 + old * 8;
            const float* r1 = xbtab + (int64_t)(q_bin[q] + nb) * 48 + old * 8;
            q_r0[q] = *(const float4*)r0; q_s0[q] = *(const float2*)(r0 + 4);
            q_r1[q] = *(const float4*)r1; q_s1[q] = *(const float2*)(r1 + 4);
            if (chg && (old == 2 || old == 3)) q_ab[q] = *(const float4*)(abstab + (int64_t)q_kold[q] * 4);
          }
        }
      }
      const auto tmc_p = sh.tm + tmc * S;
      // pre-solve thick positions the metrics need (work, energy changes).
      // Written after K1's barrier: wave 0 may still be reading the previous
      // step's partials until then.
      {
        double wxp = 0.0;
#pragma unroll
        for (int s = 0; s < MS; ++s) {
          const int n = sl.node[s];
          if (n < 0 || !sl.thick[s]) continue;
          const double xp = posd<REF64>(sh, n);
          wxp += xp;
          if (sl.li[s] == 0) sh.fil[sl.fil[s] * NF + F_X0] = xp;
          if (sl.li[s] == dm.nc - 1) sh.fil[sl.fil[s] * NF + F_XT] = xp;
        }
        wxp = wave_sum(wxp, lane);
        if (lane == 0) sh.wpart[wave * NWP + W_XPRE] = wxp;
      }

      // ---- K2: XB transitions + binding bookkeeping (transitions.py:1081-1245)
      float m_atpe = 0.0f, m_tear = 0.0f;
      int c_xb[6] = {0, 0, 0, 0, 0, 0}, c_atp = 0, c_new = 0;  // wave-uniform counts
#pragma unroll
      for (int q = 0; q < MX; ++q) {
        const int j = tid + q * BS;
        const bool on = j < X;
        int nw = 0, old = 0;
        if (on) {
          const int gsite = q_near[q];
          const bool perm = (tmc_p[gsite] == 3) && (sh.xslot[j] < 0);
          const int key = (perm ? nb : 0) + q_bin[q];
          old = q_old[q];
          const float u = U ? U[(int64_t)t * L.u_stride + S + j] : paired ? u_xb[q] : philox_uniform(seed, t, 1, j);
          const float4 c0 = perm ? q_r1[q] : q_r0[q];
          const float2 c1 = perm ? q_s1[q] : q_s0[q];
          nw = u < c0.x ? 0 : u < c0.y ? 1 : u < c0.z ? 2 : u < c0.w ? 3 : u < c1.x ? 4 : u < c1.y ? 5 : 0;
          const bool ob = old >= 1 && old <= 3, nbd = nw >= 1 && nw <= 3;
          const bool binding = !ob && nbd, unbinding = ob && !nbd;
          const bool can_bind = binding && perm && !q_occ[q];
          int site = sh.xsite[j];
          if (unbinding) {
            if (site >= 0 && q_clr[q]) sh.tbound[site - dm.nk] = -1;
            site = -1;
          }
          if (can_bind) {
            site = dm.nk + gsite;
            atomicMin((unsigned int*)(sh.tbound + (site - dm.nk)), (unsigned int)j);
          }
          if (binding && !can_bind) nw = 0;
          sh.xsite[j] = (short)site;
          sh.xst[j] = (unsigned char)nw;
          sh.xkey[j] = (short)key;
          if (site >= 0) sh.xnext[j] = (short)atomicExch((int*)(sh.shead + (site - dm.nk)), j);
          // expected ATP / tear from the previous step's key (metrics_fn.py:297-335)
          if (!(DYN && rad.on) && (old == 2 || old == 3)) {
            const float4 a = q_ab[q];
            if (old == 3) m_atpe += a.x;
            m_tear += old == 3 ? a.z : a.y;
          }
        }
        // state counts of the new XB states and per-step transition counts
#pragma unroll
        for (int k = 0; k < 6; ++k) c_xb[k] += wave_count(on && nw == k);
        c_atp += wave_count(on && old == 3 && (nw == 4 || nw == 0));
        c_new += wave_count(on && old == 0 && nw == 1);
      }
      if (lane == 0) {
        LDS int* c = sh.cnt + wave * NC;
#pragma unroll
        for (int k = 0; k < 6; ++k) c[C_XB0 + k] = c_xb[k];
        c[C_ATPC] = c_atp;
        c[C_NEWLY] = c_new;
      }
      {
        const float ae = wave_sum(m_atpe, lane), te = wave_sum(m_tear, lane);
        if (lane == 0) { sh.wpart[wave * NWP + W_ATPE] = ae; sh.wpart[wave * NWP + W_TEAR] = te; }
      }
      __syncthreads();
    }  // kinetics
    // kinetics_step alone (timestep.py:65-129): no solve, no metrics row
    if (kin && kin_only) continue;

    // ---- equilibrium solve -------------------------------------------------
    int iters;
    float resid;
    using TS = typename std::conditional<REF64, double, float>::type;
    const bool rad_on = rad.on;
    rad.on = rad_on && kin;  // the initial equilibration runs at fixed spacing
    TS d_sol = rad.on ? TS(ls_carry) : TS(d_solve);
    solve<BS, MS, REF64, DYN>(sh, dm, sl, rA, rB, z_solve, d_sol, L.cfg, rad, rd, tid, lane, wave, iters,
                              resid);
    rad.on = rad_on;
    if (!kin) continue;
    const double ls_new = (DYN && rad.on) ? (double)d_sol : (double)d_solve;
    ls_carry = ls_new;
    if (DYN && rad.on && !solve_only) {
      // the metrics' absorbing chain at the solved spacing (simulation.py:711-719),
      // evaluated on the pre-kinetics keys and states (metrics_fn.py:297-335)
      unsigned short* ka = (unsigned short*)sh.pv;   // the CG buffer is idle after the solve
      if (pCa != pre_pca) {
        fill_xb_pre(pops, (double)pCa, sh.xbpre, tid);
        pre_pca = pCa;
      }
      if (wave == 0) {
        const int na_ = compact_keys(sh.need + 64, ka, lane);
        if (lane == 0) sh.need[129] = (uint32_t)na_;
      }
      __syncthreads();
      rebuild_rows<BS, GW>(pops, pCa, ls_new, L.cfg.dt, centers, nb, dm.kxb, true, xbtab, abstab, tid,
                           ka, (int)sh.need[129], nullptr, nullptr, (const double*)sh.xbpre);
      __syncthreads();
      for (int i = 64 + tid; i < 128; i += BS) sh.need[i] = 0;
      float m_atpe = 0.0f, m_tear = 0.0f;
#pragma unroll
      for (int q = 0; q < MX; ++q) {
        const int j = tid + q * BS;
        if (j < X && (q_sold[q] == 2 || q_sold[q] == 3)) {
          const float4 a = *(const float4*)(abstab + (int64_t)q_kold[q] * 4);
          if (q_sold[q] == 3) m_atpe += a.x;
          m_tear += q_sold[q] == 3 ? a.z : a.y;
        }
      }
      const float ae = wave_sum(m_atpe, lane), te = wave_sum(m_tear, lane);
      if (lane == 0) { sh.wpart[wave * NWP + W_ATPE] = ae; sh.wpart[wave * NWP + W_TEAR] = te; }
    }

    // ---- metrics (metrics_fn.py:154-417) -----------------------------------
    if (L.metrics != nullptr || L.force != nullptr) {
      const auto tmc_p = sh.tm + tmc * S;
      const double kk = (double)sh.prm[P_THICK_K];
      const double ta = sh.prm[P_TITIN_A], tbp = sh.prm[P_TITIN_B], trest = sh.prm[P_TITIN_REST];
      // one pass: node sums, sums of squares and extrema (wave partials),
      // per-filament values
      double sk = 0, sn = 0, wk = 0, mxk = -INFINITY, nmk = -INFINITY, mxn = -INFINITY, nmn = -INFINITY;
      double qk = 0, qn = 0;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        const int n = sl.node[s];
        if (n < 0) continue;
        const double x = posd<REF64>(sh, n);
        const double o = (double)sh.org[n];
        const double dd = x - o;
        if (sl.thick[s]) {
          sk += dd; qk += dd * dd; mxk = fmax(mxk, dd); nmk = fmax(nmk, -dd);
          wk += x;
          LDS double* fv = sh.fil + sl.fil[s] * NF;
          if (sl.li[s] == 0) {
            const double e1 = 0.5 * kk * dd * dd;
            const double d0 = fv[F_X0] - o;
            fv[F_FORCE] = dd * kk;
            fv[F_E1] = e1;
            fv[F_DE1] = e1 - 0.5 * kk * d0 * d0;
          }
          if (sl.li[s] == dm.nc - 1) {
            if (REF64 && !(DYN && rad.on)) {
              // the solve evaluated both at this Z-line and spacing: F_TIT
              // holds the energy at the solved tip, F_EO at the pre-solve one
              fv[F_DTIT] = fv[F_TIT] - fv[F_EO];
            } else {
              const double an = (double)zl - x, ao = (double)zl - fv[F_XT];
              const double Ln = sqrt(an * an + ls_new * ls_new), Lo = sqrt(ao * ao + ls_new * ls_new);
              const double en = (ta / tbp) * (exp_t<double, (MS >= 3)>(tbp * (Ln - trest)) - 1.0);
              const double eo = (ta / tbp) * (exp_t<double, (MS >= 3)>(tbp * (Lo - trest)) - 1.0);
              fv[F_TIT] = en;
              fv[F_DTIT] = en - eo;
            }
          }
        } else {
          sn += dd; qn += dd * dd; mxn = fmax(mxn, dd); nmn = fmax(nmn, -dd);
        }
      }
      // A wave that owns no thick (or no thin) node skips that kind's
      // reductions: every lane already holds the identity (0 or -inf).
      bool any_k = false, any_n = false;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        any_k = any_k || (sl.node[s] >= 0 && sl.thick[s]);
        any_n = any_n || (sl.node[s] >= 0 && !sl.thick[s]);
      }
      if (__ballot(any_k)) {
        sk = wave_sum(sk, lane); wk = wave_sum(wk, lane); qk = wave_sum(qk, lane);
        mxk = wave_max(mxk, lane); nmk = wave_max(nmk, lane);
      }
      if (__ballot(any_n)) {
        sn = wave_sum(sn, lane); qn = wave_sum(qn, lane);
        mxn = wave_max(mxn, lane); nmn = wave_max(nmn, lane);
      }
      if (lane == 0) {
        LDS double* wp = sh.wpart + wave * NWP;
        wp[W_SUMK] = sk; wp[W_SUMN] = sn; wp[W_WORK] = wk;
        wp[W_MAXK] = mxk; wp[W_NMINK] = nmk; wp[W_MAXN] = mxn; wp[W_NMINN] = nmn;
        wp[W_SQK] = qk; wp[W_SQN] = qn;
      }
      // TM state and overlap counts on the post-solve positions
      int c_tm[4] = {0, 0, 0, 0}, c_ov = 0, c_ov2 = 0, c_ov3 = 0;  // wave-uniform
      for (int s2 = tid; s2 < S + (BS - S % BS) % BS; s2 += BS) {
        const bool on = s2 < S;
        const int st = on ? tmc_p[s2] : -1;
        bool ov = false;
        if (on) {
          const float x = posf<REF64>(sh, dm.nk + s2);
          ov = x >= ov_lo && x <= ov_hi && x > 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) c_tm[k] += wave_count(st == k);
        c_ov += wave_count(ov);
        c_ov2 += wave_count(ov && st == 2);
        c_ov3 += wave_count(ov && st == 3);
      }
      if (lane == 0) {
        LDS int* c = sh.cnt + wave * NC;
#pragma unroll
        for (int k = 0; k < 4; ++k) c[C_TM0 + k] = c_tm[k];
        c[C_OV] = c_ov;
        c[C_OV2] = c_ov2;
        c[C_OV3] = c_ov3;
      }
      // The next step's candidate positions, when its Z-line equals this
      // one: then nothing moves the thin filaments before its search, and
      // the positions are ready behind this barrier instead of one of its own.
      fpos_ready = fpos != nullptr && step + 1 < n_steps && d_z == zl;
      if (fpos_ready) face_positions(sh, dm, fpos, Wp, tid, BS);
      __syncthreads();
      // The cross-wave fold is split over waves 0-2, which run it
      // concurrently: lane k folds one column (across the waves' partials or
      // the filaments) and lane 0 writes the wave's part of the metrics row.
      float* mrow = L.metrics ? L.metrics + ((int64_t)b * L.t_stride + t) * MF_N_METRICS : nullptr;
      if (wave == 0) {
        // node moments and the XB expectations; lanes past NWP fold column 0
        // again so that every lane's loads are unconditional
        const int col = lane < NWP ? lane : 0;
        const bool ismax = col == W_MAXK || col == W_NMINK || col == W_MAXN || col == W_NMINN;
        double v = ismax ? -INFINITY : 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          const double x = sh.wpart[w * NWP + col];
          v = ismax ? fmax(v, x) : v + x;
        }
        // population variances (jnp.std) from the sums of squares, in double
        // (displacements are O(100) nm at most: the cancellation costs a few
        // of double's 16 digits, far from the float32 result)
        const double mean_k = read_lane(v, W_SUMK) * inv_nk, mean_n = read_lane(v, W_SUMN) * inv_S;
        const double var_k = fmax(read_lane(v, W_SQK) * inv_nk - mean_k * mean_k, 0.0);
        const double var_n = fmax(read_lane(v, W_SQN) * inv_S - mean_n * mean_n, 0.0);
        // every lane reads the folded values (cross-lane reads stay outside
        // the lane-0 branch, where the other lanes' registers are not live)
        const double mxk = read_lane(v, W_MAXK), nmk = read_lane(v, W_NMINK);
        const double mxn = read_lane(v, W_MAXN), nmn = read_lane(v, W_NMINN);
        const double atpe = read_lane(v, W_ATPE), tear = read_lane(v, W_TEAR);
        if (lane == 0 && mrow) {
          float* m = mrow;
          m[34] = (float)mean_k; m[35] = (float)mxk; m[36] = (float)(-nmk);
          m[37] = (float)sqrt(var_k);
          m[38] = (float)mean_n; m[39] = (float)mxn; m[40] = (float)(-nmn);
          m[41] = (float)sqrt(var_n);
          m[48] = (float)atpe; m[49] = (float)tear;
        }
      } else if (wave == 1) {
        // force and energies (lanes < NF_SUM: per-filament columns), work
        // and expected ATP (lanes 8-10: wave-partial columns)
        double v = 0.0;
        if (lane < NF_SUM) {
          for (int f = 0; f < dm.Tk; ++f) v += sh.fil[f * NF + lane];
        } else if (lane >= 8 && lane < 11) {
          const int col = lane == 8 ? W_WORK : lane == 9 ? W_XPRE : W_ATPE;
#pragma unroll
          for (int w = 0; w < NW; ++w) v += sh.wpart[w * NWP + col];
        }
        const double force = read_lane(v, F_FORCE);
        const double work = force * ((read_lane(v, 8) - read_lane(v, 9)) * inv_nk);
        const double atpe = read_lane(v, 10);
        const double e1 = read_lane(v, F_E1), de1 = read_lane(v, F_DE1);
        const double tit = read_lane(v, F_TIT), dtit = read_lane(v, F_DTIT);
        if (lane == 0) {
          if (mrow) {
            float* m = mrow;
            const float invTk = 1.0f / (float)dm.Tk;
            m[0] = (float)force; m[1] = resid; m[2] = zl; m[3] = pCa; m[4] = (float)ls_new;
            m[42] = (float)e1 * invTk; m[43] = (float)de1 * invTk;
            m[44] = (float)tit * invTk; m[45] = (float)dtit * invTk;
            m[46] = (float)work; m[47] = (float)(work * inv_nk);
            m[50] = (float)(atpe > 0.01 ? work / atpe : 0.0);
            m[51] = (float)iters;
          }
          if (L.force) L.force[(int64_t)b * L.t_stride + t] = force;
        }
      } else if (wave == 2 && mrow) {
        // state and overlap counts (lane k < NC: counter k)
        const int col = lane < NC ? lane : 0;
        int cv = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) cv += sh.cnt[w * NC + col];
        float c[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) c[k] = (float)__builtin_amdgcn_readlane(cv, k);
        if (lane == 0) {
          const float nbound = c[C_XB0 + 1] + c[C_XB0 + 2] + c[C_XB0 + 3];
          const float ov = c[C_OV];
          const float invX = 1.0f / (float)X, invS = 1.0f / (float)S;
          float* m = mrow;
          m[5] = nbound;
#pragma unroll
          for (int k = 0; k < 6; ++k) m[6 + k] = c[C_XB0 + k];
          m[12] = nbound * invX;
#pragma unroll
          for (int k = 0; k < 6; ++k) m[13 + k] = c[C_XB0 + k] * invX;
#pragma unroll
          for (int k = 0; k < 4; ++k) m[19 + k] = c[C_TM0 + k];
#pragma unroll
          for (int k = 0; k < 4; ++k) m[23 + k] = c[C_TM0 + k] * invS;
          m[27] = c[C_TM0 + 3] * invS;
          m[28] = c[C_OV2] / ov;
          m[29] = c[C_OV3] / ov;
          m[30] = (c[C_OV2] + c[C_OV3]) / ov;
          m[31] = ov;
          m[32] = c[C_ATPC]; m[33] = c[C_NEWLY];
        }
      }
    }
  }

  // ---- write the state back ------------------------------------------------
  __syncthreads();
  {
    double* pos = (double*)(stb + SL.pos);
    for (int i = tid; i < dm.N; i += BS) pos[i] = posd<REF64>(sh, i);
    int32_t* xbb = (int32_t*)(stb + SL.xb_bound);
    int32_t* xbk = (int32_t*)(stb + SL.xb_key);
    uint8_t* xbs = stb + SL.xb_state;
    for (int j = tid; j < X; j += BS) {
      const int site = sh.xsite[j];
      xbb[j] = site >= 0 ? site - dm.nk - li[G.xb_thin + j] * dm.ns : -1;
      xbk[j] = sh.xkey[j];
      xbs[j] = sh.xst[j];
    }
    int32_t* tbd = (int32_t*)(stb + SL.thin_bound);
    uint8_t* tms = stb + SL.tm;
    const auto tcur = sh.tm + tmc * S;
    for (int s = tid; s < S; s += BS) { tbd[s] = sh.tbound[s]; tms[s] = tcur[s]; }
    if (tid == 0) *(double*)(stb + SL.ls) = ls_carry;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// kernel variants: each translation unit mf_variants_*.hip instantiates a
// group of them; the host code concatenates the groups in a fixed order.
// ---------------------------------------------------------------------------
namespace mf {
struct Variant {
  int bs, ms, mx;
  bool ref64, dyn;
  int gw;  // placement level: 0 LDS, 1 per-head arrays in the workspace, 2 positions too
  const void* fn;
  int64_t (*lds)(const Dims&);
  int64_t (*ws)(const Dims&);
};

template <int BS, int MS, int MX, bool R, bool D = false, int GW = 0>
inline Variant make_variant() {
  return Variant{BS, MS, MX, R, D, GW, (const void*)&run_kernel<BS, MS, MX, R, D, GW>, &lds_bytes<BS, R, GW>,
                 &ws_bytes<BS, R, GW>};
}

}  // namespace mf
