// Core distances through a certified MFMA Gram filter (float32 windows).
//
// The core distance of row i is the ms-th smallest exact distance of the row,
// self included (reference hdbscan_gpu/_cuda.py:114-118 partitions the
// materialised matrix).  Exact means the reference's own arithmetic:
// s = sum over k ascending of (x_ik - x_jk)^2 in float32, no contraction,
// then the correctly rounded square root (reference hdbscan_gpu/_kernels.py:
// 45-73).  sqrt is monotone, so the ms-th smallest distance is
// sqrt(ms-th smallest s).
//
// Filter.  Matrix cores (v_mfma_f32_16x16x32_bf16) evaluate, for every pair,
//   g = |x_i|^2 (1 +- c) + |x_j|^2 (1 +- c) +- eabs - 2 x_i . x_j
// with x split into bf16 pieces, x = hi + lo + rest (|rest| <= 2^-16 |x|), and
// the three products hi.hi + hi.lo + lo.hi accumulated in float32 on top of
// the two norm terms (the accumulator's initial value).  With "+" the result
// is an upper bound ub >= s, with "-" a lower bound lb <= s.  gram_prep_kernel
// writes the pieces of x and of -2 x and the norm terms once per point.
//   Pass A: U_i = ms-th smallest ub of row i, taken over the per-lane
//           minima (each of the 16 lanes of a row keeps its M smallest ub),
//           a subset of the row, so U_i >= the ms-th smallest ub >= t_i,
//           the exact statistic (ms columns have s <= ub).
//   Pass B: every column with lb <= U_i is written to the row's candidate
//           list.  Each column with s <= t_i has lb <= s <= t_i <= U_i, so
//           the ms smallest exact values of the row are all candidates.
// core_exact_kernel then recomputes the candidates exactly (one lane per
// pair, coordinate order) and selects the ms-th smallest.  A row with more
// candidates than list slots (ties, or the filter off) is recomputed
// exactly over all columns.
//
// Error bound (u = 2^-24, D = padded dimension, S = |x_i|^2 + |x_j|^2,
// r = real squared distance <= 2 S):
//   reference sum          |s - r| <= gamma_{D+2} r
//   float32 squared norms  |n - |x|^2| <= gamma_D |x|^2
//   bf16 split residual    <= 3.02 2^-16 |x_k| |y_k| per coordinate product
//   float32 accumulation   <= 2u per multiply-add on magnitudes <= 2.1 S
//                          (allowing any internal order and rounding),
//                          one more rounding for the initial value
// so c = (16 D + 48) u + 7 2^-16 bounds all of them relative to S.  Any flush
// of subnormal bf16 operands or products stays below 2^-110 (1 + S), covered
// by eabs = 2^-109 and 2^-110 added to c.  The filter is used only when every
// squared norm of the window is <= 2^100, where no pair can overflow;
// otherwise every row takes the exact path (which also checks overflow).
#pragma once

#include "hdb_common.h"

namespace hdb {

typedef float gram_f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 gram_bf16x8 __attribute__((ext_vector_type(8)));
typedef short gram_s16x8 __attribute__((ext_vector_type(8)));

// bf16 bits of v rounded to nearest even (finite v).
__device__ __forceinline__ uint32_t bf16_bits(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_value(uint32_t bits) { return __uint_as_float(bits << 16); }

// Minimum over the 16 lanes of a DPP row (row_ror 1, 2, 4, 8).
__device__ __forceinline__ float row16_min(float v) {
  v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x121, 0xF, 0xF, false)));
  v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x122, 0xF, 0xF, false)));
  v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xF, 0xF, false)));
  v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xF, 0xF, false)));
  return v;
}

// Ascending register list of the K smallest values offered.
template <int K>
__device__ __forceinline__ void list_offer(float* lst, float v) {
  if (v < lst[K - 1]) {
    // Ascending list, v not NaN (it compared below lst[K-1]): the new q-th
    // smallest is the median of (lst[q-1], lst[q], v).
#pragma unroll
    for (int q = K - 1; q > 0; --q) lst[q] = __builtin_amdgcn_fmed3f(lst[q - 1], lst[q], v);
    lst[0] = v < lst[0] ? v : lst[0];
  }
}

// The kth smallest (1-based) of the union of the sorted lists of the 16
// lanes of each DPP row; every lane of the row returns it.  Equal heads are
// removed together, so the value at which the running count first reaches
// kth is the kth smallest.  Uniform trip count: all lanes active.
template <int K>
__device__ __forceinline__ float group16_kth(float* lst, int kth, int lane) {
  int cnt = 0;
  float ans = INFINITY;
  bool done = false;
  for (int it = 0; it < kth; ++it) {
    const float h = lst[0];
    const float m = row16_min(h);
    const bool eq = h == m;
    const uint64_t bal = __ballot(eq);
    const int c = __popcll((bal >> (lane & 48)) & 0xFFFFull);
    if (!done) {
      cnt += c;
      if (cnt >= kth) {
        ans = m;
        done = true;
      }
    }
    if (eq) {
#pragma unroll
      for (int q = 0; q < K - 1; ++q) lst[q] = lst[q + 1];
      lst[K - 1] = INFINITY;
    }
  }
  return ans;
}

// Launch shape of the Gram candidate kernel for padded dimension DK and
// RSET 16-row sets per wave (each B operand read from LDS feeds RSET MFMAs).
template <int DK, int RSET> struct GramShape {
  static constexpr int KP = (3 * DK + 31) / 32 * 32;    // three bf16 segments, MFMA K = 32
  static constexpr int NQ = KP / 32;                    // MFMAs per 16x16 tile
  static constexpr int KS = KP + 8;                     // LDS column stride (bf16)
  static constexpr int TJ = DK >= 128 ? 32 : DK <= 32 ? 128 : 64;  // columns per chunk
  static constexpr int RS = RSET;
  static constexpr int WR = 16 * RS;                    // rows per wave
  static constexpr int MAXT = DK >= 64 ? 256 : 512;     // threads per workgroup
  static constexpr int VC = TJ * KP / 8;                // 16-byte vectors per chunk
  static constexpr int PF = (VC + MAXT - 1) / MAXT;     // prefetched vectors per thread
};

// Candidate slots per row (columns beyond it send the row to the exact path).
constexpr int kGramCap = 32;
// +inf norm terms after each window's columns: a 16-column step starting
// inside the window reads at most 15 past it.
constexpr int kGramPad = 16;

// Bound constant c and half the absolute term for padded dimension dk.
__host__ __device__ constexpr float gram_c(int dk) {
  return float(16 * dk + 48) * 5.9604644775390625e-08f + 7.0f * 1.52587890625e-05f + 7.7037197775489434e-34f;
}
constexpr float kGramEh = 7.7037197775489434e-34f;  // 2^-110

// Per point p of the batch: pa[p] = bf16 (hi(x) | lo(x)), pb[p] = (hi(-2x) |
// lo(-2x)), each DK wide (zero padded); nb[p] = |x|^2 (1 + c) + eh,
// nb[P + p] = |x|^2 (1 - c) - eh, nb[2P + p] = |x|^2.
// Non-finite coordinates OR HDB_STATUS_NONFINITE (2) into *status.
template <int DK>
__global__ __launch_bounds__(256) void gram_prep_kernel(const float* x, int64_t points, int d, uint16_t* pa,
                                                        uint16_t* pb, float* nb, int32_t* status) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int bad = 0;
  if (p < points) {
  const float* xp = x + (size_t)p * d;
  uint16_t* ap = pa + (size_t)p * 2 * DK;
  uint16_t* bp = pb + (size_t)p * 2 * DK;
  const float c = gram_c(DK);
  float sq = 0.f;
  for (int k0 = 0; k0 < DK; k0 += 8) {
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = k0 + e < d ? xp[k0 + e] : 0.f;
#pragma unroll
    for (int e = 0; e < 8; ++e) bad |= !isfinite(v[e]);
    gram_s16x8 ah, al, bh, bl;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      sq = fmaf(v[e], v[e], sq);
      const uint32_t h = bf16_bits(v[e]);
      ah[e] = (short)h;
      al[e] = (short)bf16_bits(v[e] - bf16_value(h));
      const float w = -2.f * v[e];
      const uint32_t g = bf16_bits(w);
      bh[e] = (short)g;
      bl[e] = (short)bf16_bits(w - bf16_value(g));
    }
    *reinterpret_cast<gram_s16x8*>(ap + k0) = ah;
    *reinterpret_cast<gram_s16x8*>(ap + DK + k0) = al;
    *reinterpret_cast<gram_s16x8*>(bp + k0) = bh;
    *reinterpret_cast<gram_s16x8*>(bp + DK + k0) = bl;
  }
  nb[p] = fmaf(sq, c, sq) + kGramEh;
  nb[points + p] = fmaf(sq, -c, sq) - kGramEh;
  nb[2 * points + p] = sq;
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 2);
}

// LDS bytes of core_gram_kernel<.., DK, RSET> for a window of n points.
template <int DK, int RSET>
HDB_HD uint32_t gram_lds_bytes(int n) {
  using G = GramShape<DK, RSET>;
  return (uint32_t)G::TJ * G::KS * 2u + 2u * (uint32_t)(n + kGramPad) * 4u;
}

template <int DK> constexpr int gram_max_threads() { return GramShape<DK, 1>::MAXT; }

// Source offset (in bf16) inside a point's (hi | lo) pair of the 8 operand
// entries starting at K index k, for the segment order [s0 | s1 | s2];
// -1: zero padding.
template <int DK>
__device__ __forceinline__ int gram_src(int k, int lo_segment) {
  const int seg = k / DK, m = k - seg * DK;
  return seg >= 3 ? -1 : (seg == lo_segment ? DK : 0) + m;
}

// Pass A and B of the filter.  One workgroup: WR rows per wave (RS sets of
// 16, each with its MFMA A operand in registers); columns streamed through
// LDS in chunks of TJ points, twice; the next chunk's loads are in flight
// while the current chunk is consumed.  Output: cand[b][i][kGramCap] column
// indices and ncand[b][i] (> kGramCap: recompute the row exactly).
// Grid: (ceil(n / rows), batch).  M: per-lane minima kept in pass A (1, 2).
template <int M, int DK, int RSET>
__global__ __launch_bounds__(gram_max_threads<DK>()) void core_gram_kernel(Windows win, int ms, const uint16_t* pa,
                                                                          const uint16_t* pb, const float* nb,
                                                                          int64_t points, uint16_t* cand,
                                                                          int32_t* ncand) {
  using G = GramShape<DK, RSET>;
  constexpr int NQ = G::NQ, KS = G::KS, TJ = G::TJ, PF = G::PF, VC = G::VC;
  constexpr int RS = G::RS, WR = G::WR, CAP = kGramCap;
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ uint32_t sh_nmax;
  const int tid = threadIdx.x, T = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, W = T >> 6;
  const int row0 = blockIdx.x * W * WR;
  int64_t wbase;
  int n;
  int wid;
  win.get(blockIdx.y, wid, wbase, n);
  if (row0 >= n) return;
  uint16_t* bcol = reinterpret_cast<uint16_t*>(lds);          // [TJ][KS] bf16
  // [2][n + kGramPad]: upper then lower norm terms, +inf past the window.
  float* nbound = reinterpret_cast<float*>(bcol + TJ * KS);
  const int np = n + kGramPad;
  const size_t rbase = (size_t)wbase;

  if (tid == 0) sh_nmax = 0;
  // Zero padding of the column operands (never overwritten).
  for (int idx = tid; idx < TJ * KS; idx += T) {
    const int k = idx % KS;
    if (k >= 3 * DK) bcol[idx] = 0;
  }
  __syncthreads();  // sh_nmax initialised before the atomics below
  uint32_t nmax = 0;
  for (int j = tid; j < n; j += T) {
    nbound[j] = nb[rbase + j];
    nbound[np + j] = nb[points + rbase + j];
    const float sq = nb[2 * points + rbase + j];
    // NaN compares false below, so it disables the filter like a huge norm.
    nmax = max(nmax, sq <= 1.2676506002282294e30f ? __float_as_uint(sq) : 0xFFFFFFFFu);
  }
  if (tid < 2 * kGramPad) nbound[(tid / kGramPad) * np + n + tid % kGramPad] = INFINITY;
  atomicMax(&sh_nmax, nmax);
  __syncthreads();
  if (sh_nmax > __float_as_uint(1.2676506002282294e30f)) {
    // Filter off for this window: every row goes to the exact path.
    for (int r = tid; r < W * WR; r += T)
      if (row0 + r < n) ncand[rbase + row0 + r] = CAP + 1;
    return;
  }

  // Accumulator layout: register r holds row 4 (lane >> 4) + r of the set,
  // column lane & 15.
  const int grp = lane >> 4;
  const int nchunks = (n + TJ - 1) / TJ;
  const uint32_t below = (1u << (lane & 15)) - 1u;

  // Operand A of set g: row g*16 + (lane & 15) of the wave's strip as
  // [hi(x) | hi(x) | lo(x)].
  gram_bf16x8 av[RS][NQ];
#pragma unroll
  for (int g = 0; g < RS; ++g) {
    const int ai = min(row0 + wave * WR + g * 16 + (lane & 15), n - 1);
    const uint16_t* ap = pa + (rbase + ai) * 2 * DK;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int src = gram_src<DK>(32 * q + 8 * grp, 2);
      av[g][q] = src < 0 ? gram_bf16x8{} : *reinterpret_cast<const gram_bf16x8*>(ap + src);
    }
  }
  float U[RS][4];
  int cnt[RS][4];
  float mn[RS][4][M];
#pragma unroll
  for (int g = 0; g < RS; ++g)
#pragma unroll
    for (int r = 0; r < 4; ++r) cnt[g][r] = 0;
  for (int pass = 0; pass < 2; ++pass) {
    float nrow[RS][4];
#pragma unroll
    for (int g = 0; g < RS; ++g)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = min(row0 + wave * WR + g * 16 + 4 * grp + r, n - 1);
        nrow[g][r] = nbound[pass * np + i];
#pragma unroll
        for (int q = 0; q < M; ++q) mn[g][r][q] = INFINITY;
      }
    gram_s16x8 pre[PF];
    // Column chunk: operand vectors [hi(-2x) | lo(-2x) | hi(-2x)], copied.
    auto fetch = [&](int c0) {
#pragma unroll
      for (int e = 0; e < PF; ++e) {
        const int v = tid + e * T;
        const int c = v / (G::KP / 8), k = (v - c * (G::KP / 8)) * 8, j = c0 + c;
        const int src = gram_src<DK>(k, 1);
        pre[e] = (v < VC && j < n && src >= 0)
                     ? *reinterpret_cast<const gram_s16x8*>(pb + (rbase + j) * 2 * DK + src)
                     : gram_s16x8{};
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int e = 0; e < PF; ++e) {
        const int v = tid + e * T;
        const int c = v / (G::KP / 8), k = (v - c * (G::KP / 8)) * 8;
        if (v < VC) *reinterpret_cast<gram_s16x8*>(bcol + c * KS + k) = pre[e];
      }
    };
    // Threads beyond the prefetch width (small workgroups) load directly.
    auto direct = [&](int c0) {
      for (int v = tid + PF * T; v < VC; v += T) {
        const int c = v / (G::KP / 8), k = (v - c * (G::KP / 8)) * 8, j = c0 + c;
        const int src = gram_src<DK>(k, 1);
        *reinterpret_cast<gram_s16x8*>(bcol + c * KS + k) =
            (j < n && src >= 0) ? *reinterpret_cast<const gram_s16x8*>(pb + (rbase + j) * 2 * DK + src)
                                : gram_s16x8{};
      }
    };
    __syncthreads();
    fetch(0);
    store();
    direct(0);
    __syncthreads();
    for (int ch = 0; ch < nchunks; ++ch) {
      const int c0 = ch * TJ;
      const bool more = ch + 1 < nchunks;
      if (more) fetch(c0 + TJ);
      const int cn = min(TJ, n - c0);
      for (int s0 = 0; s0 < cn; s0 += 16) {
        const int cl = s0 + (lane & 15);
        const int j = c0 + cl;
        const float ncj = nbound[pass * np + j];  // j < n + 15: +inf past the window
        gram_f32x4 acc[RS];
#pragma unroll
        for (int g = 0; g < RS; ++g)
          acc[g] = gram_f32x4{nrow[g][0] + ncj, nrow[g][1] + ncj, nrow[g][2] + ncj, nrow[g][3] + ncj};
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const gram_bf16x8 bv = *reinterpret_cast<const gram_bf16x8*>(bcol + cl * KS + 32 * q + 8 * grp);
#pragma unroll
          for (int g = 0; g < RS; ++g) acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[g][q], bv, acc[g], 0, 0, 0);
        }
        if (pass == 0) {
#pragma unroll
          for (int g = 0; g < RS; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              // With mn[0] <= mn[1], the second smallest of {mn[0], mn[1], v}
              // is their median (one instruction); the smallest is an unsigned
              // minimum of the bit patterns (one instruction, where fminf adds
              // NaN canonicalisation): the upper bounds are positive (>= the
              // exact sum plus a positive absolute term) or +inf (columns
              // past the window), and positive floats order like their bits.
              // No NaN can occur: the filter runs only on finite norms.
              const float v = acc[g][r];
              if constexpr (M == 2) mn[g][r][1] = __builtin_amdgcn_fmed3f(mn[g][r][0], mn[g][r][1], v);
              mn[g][r][0] = __uint_as_float(min(__float_as_uint(v), __float_as_uint(mn[g][r][0])));
            }
        } else {
          // Candidates of row (set g, 4 grp + r): the 16 lanes of a group
          // hold its 16 columns; slots follow the lane order.
#pragma unroll
          for (int g = 0; g < RS; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const bool c = !(acc[g][r] > U[g][r]);
              const uint64_t mask = __ballot(c);
              if (mask) {
                const uint32_t m16 = (uint32_t)(mask >> (16 * grp)) & 0xFFFFu;
                const int pos = cnt[g][r] + __popc(m16 & below);
                const int i = row0 + wave * WR + g * 16 + 4 * grp + r;
                if (c && pos < CAP && i < n) cand[(rbase + i) * CAP + pos] = (uint16_t)j;
                cnt[g][r] += __popc(m16);
              }
            }
        }
      }
      __syncthreads();
      if (more) {
        store();
        direct(c0 + TJ);
        __syncthreads();
      }
    }
    if (pass == 0) {
#pragma unroll
      for (int g = 0; g < RS; ++g)
#pragma unroll
        for (int r = 0; r < 4; ++r) U[g][r] = group16_kth<M>(mn[g][r], ms, lane);
    }
  }
  if ((lane & 15) == 0) {
#pragma unroll
    for (int g = 0; g < RS; ++g)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wave * WR + g * 16 + 4 * grp + r;
        if (i < n) ncand[rbase + i] = cnt[g][r];
      }
  }
}

// Per-window form of both passes for windows of at most kWinMaxN points and
// at most 16 dimensions (padded to 16).  One workgroup owns one window and
// computes the bf16 pieces itself (no prep kernel).  Every column's B
// operand is written into LDS once, in MFMA fragment order: fragment q of
// 16-column block cb is 64 lanes x 16 bytes, lane l's eight K entries at
// byte 16 l, so each operand read is one contiguous, conflict-free 16-byte
// read per lane.  Each wave then runs both passes over all columns for its
// 64 rows (four 16-row sets) with no barrier between column blocks (the
// chunked kernel re-reads every column chunk per row block, behind two
// barriers per chunk).
//
// The norm terms ride in K slots 48..53 of the second MFMA, which the
// three 16-wide segments leave free: A holds the row term as three bf16
// pieces in slots 48..50 and 1 in 51..53, B holds 1 in 48..50 and the
// column term's pieces in 51..53.  The accumulator starts at zero, so no
// VALU addition per tile forms the initial value.  Pieces of v: p1 =
// bf16(v), p2 = bf16(v - p1), p3 = bf16(v - p1 - p2); each difference is
// exact in float32 (it is v's low-order bits) and |v - p1 - p2 - p3| <=
// 2^-24 |v|.  Bound: against the chunked form the initial-value rounding
// disappears and six multiply-adds of magnitude <= 2.1 S (2u each) plus the
// piece residuals (u S per term) appear, about 27 u S; the constant
// gram_c(16 + 4) adds 64 u S.  Columns past the window carry a +inf term
// (pieces +inf, 0, 0; the A side multiplies them by 1), so their bounds are
// +inf: never among the smallest, never a candidate.
constexpr int kWinDK = 16;
constexpr int kWinMaxN = 1024;
// Norm-term constant of the per-window form.
constexpr float kWinC = gram_c(kWinDK + 4);

// LDS bytes of core_window_kernel for windows of up to n points of d
// dimensions: two counters per row, then either the column fragments (two
// 1 KiB fragments per 16-column block) or, after the passes, the window's
// points and the rows' candidate lists.
HDB_HD inline uint32_t gram_window_head(int n) { return (8u * (uint32_t)n + 15u) & ~15u; }
HDB_HD inline uint32_t gram_window_lds_bytes(int n, int d) {
  const uint32_t frags = (uint32_t)((n + 15) / 16) * 2048u;
  const uint32_t exact = ((4u * (uint32_t)n * (uint32_t)d + 15u) & ~15u) + 2u * kGramCap * (uint32_t)n;
  return gram_window_head(n) + (frags > exact ? frags : exact);
}

// The three bf16 pieces of v (finite) as bit patterns.
__device__ __forceinline__ void bf16_pieces(float v, uint32_t& p1, uint32_t& p2, uint32_t& p3) {
  p1 = bf16_bits(v);
  const float r1 = v - bf16_value(p1);
  p2 = bf16_bits(r1);
  p3 = bf16_bits(r1 - bf16_value(p2));
}

// Eight K entries (slots 48..55) carrying a norm term: the term's pieces
// first and three ones after them (row side), or the ones first (column
// side).  inf: a column past the window.
__device__ __forceinline__ gram_s16x8 norm_slots(float v, bool row_side, bool inf) {
  uint32_t p1, p2, p3;
  if (inf) {
    p1 = 0x7F80u;  // +inf
    p2 = p3 = 0;
  } else {
    bf16_pieces(v, p1, p2, p3);
  }
  constexpr short one = 0x3F80;  // bf16 1.0
  gram_s16x8 s = {};
  if (row_side) {
    s[0] = (short)p1; s[1] = (short)p2; s[2] = (short)p3;
    s[3] = one; s[4] = one; s[5] = one;
  } else {
    s[0] = one; s[1] = one; s[2] = one;
    s[3] = (short)p1; s[4] = (short)p2; s[5] = (short)p3;
  }
  return s;
}

// Loads a point's DK padded coordinates (zeros past d).  VEC (d == DK and
// every row 16-byte aligned): four 16-byte loads.  One path per
// instantiation: with both in one function the compiler keeps the array in
// scratch memory.
template <bool VEC>
__device__ __forceinline__ void window_load(const float* xp, int d, float (&v)[kWinDK]) {
  if constexpr (VEC) {
#pragma unroll
    for (int q = 0; q < kWinDK / 4; ++q) {
      const float4 t = reinterpret_cast<const float4*>(xp)[q];
      v[4 * q] = t.x;
      v[4 * q + 1] = t.y;
      v[4 * q + 2] = t.z;
      v[4 * q + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kWinDK; ++k) v[k] = k < d ? xp[k] : 0.f;
  }
}

// Squared norm of a loaded point (FMA chain in coordinate order, as
// gram_prep_kernel); ORs 1 into bad for a non-finite coordinate.
__device__ __forceinline__ float window_norm(const float (&v)[kWinDK], int& bad) {
  float sq = 0.f;
#pragma unroll
  for (int k = 0; k < kWinDK; ++k) {
    bad |= !isfinite(v[k]);
    sq = fmaf(v[k], v[k], sq);
  }
  return sq;
}

// The same outputs for DK <= 16 with a block's 256 points staged through
// LDS: the rows are read and the bf16 pieces written as contiguous 16-byte
// vectors across lanes (one point per lane would touch 64 cache lines per
// instruction); the per-point arithmetic is that of gram_prep_kernel.
template <int DK>
__global__ __launch_bounds__(256) void gram_prep_lds_kernel(const float* x, int64_t points, int d, uint16_t* pa,
                                                            uint16_t* pb, float* nb, int32_t* status) {
  static_assert(DK <= 16, "LDS staging sized for DK <= 16");
  __shared__ __align__(16) float xin[256 * DK];
  __shared__ __align__(16) uint16_t oa[256 * 2 * DK];
  __shared__ __align__(16) uint16_t ob[256 * 2 * DK];
  const int t = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * 256;
  const int cnt = (int)min((int64_t)256, points - p0);
  const float* xb = x + (size_t)p0 * d;
  const int total = cnt * d;  // d <= DK
  if (((((uintptr_t)xb) & 15) == 0) && (total & 3) == 0) {
    for (int k = t; k < total / 4; k += 256) reinterpret_cast<float4*>(xin)[k] = reinterpret_cast<const float4*>(xb)[k];
  } else {
    for (int k = t; k < total; k += 256) xin[k] = xb[k];
  }
  __syncthreads();
  int bad = 0;
  if (t < cnt) {
    const float c = gram_c(DK);
    float sq = 0.f;
    uint16_t* ap = oa + t * 2 * DK;
    uint16_t* bp = ob + t * 2 * DK;
#pragma unroll
    for (int k0 = 0; k0 < DK; k0 += 8) {
      float v[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = k0 + e < d ? xin[t * d + k0 + e] : 0.f;
#pragma unroll
      for (int e = 0; e < 8; ++e) bad |= !isfinite(v[e]);
      gram_s16x8 ah, al, bh, bl;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        sq = fmaf(v[e], v[e], sq);
        const uint32_t h = bf16_bits(v[e]);
        ah[e] = (short)h;
        al[e] = (short)bf16_bits(v[e] - bf16_value(h));
        const float w = -2.f * v[e];
        const uint32_t g = bf16_bits(w);
        bh[e] = (short)g;
        bl[e] = (short)bf16_bits(w - bf16_value(g));
      }
      *reinterpret_cast<gram_s16x8*>(ap + k0) = ah;
      *reinterpret_cast<gram_s16x8*>(ap + DK + k0) = al;
      *reinterpret_cast<gram_s16x8*>(bp + k0) = bh;
      *reinterpret_cast<gram_s16x8*>(bp + DK + k0) = bl;
    }
    const int64_t p = p0 + t;
    nb[p] = fmaf(sq, c, sq) + kGramEh;
    nb[points + p] = fmaf(sq, -c, sq) - kGramEh;
    nb[2 * points + p] = sq;
  }
  __syncthreads();
  // cnt points of 2 DK bf16 values per array: cnt * DK / 4 vectors of 16 B
  // (the block's first point is 256 * 4 DK bytes into each array).
  const int vecs = cnt * DK / 4;
  uint4* ga = reinterpret_cast<uint4*>(pa + (size_t)p0 * 2 * DK);
  uint4* gb = reinterpret_cast<uint4*>(pb + (size_t)p0 * 2 * DK);
  for (int k = t; k < vecs; k += 256) {
    ga[k] = reinterpret_cast<const uint4*>(oa)[k];
    gb[k] = reinterpret_cast<const uint4*>(ob)[k];
  }
  if (__any(bad) && (t & 63) == 0) atomicOr(status, 2);
}

// Exact coordinate-order squared distance of rows xi and xj (d floats), 16
// loads in flight per operand; 16-byte loads when both rows are aligned.
__device__ __forceinline__ float exact_sq(const float* xi, const float* xj, int d) {
  float t = 0.f;
  const bool vec = ((d & 3) == 0) && ((((uintptr_t)xi) | ((uintptr_t)xj)) & 15) == 0;
  for (int k0 = 0; k0 < d; k0 += 16) {
    float a[16], c[16];
    if (vec && k0 + 16 <= d) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 va = reinterpret_cast<const float4*>(xi + k0)[q];
        const float4 vc = reinterpret_cast<const float4*>(xj + k0)[q];
        a[4 * q] = va.x; a[4 * q + 1] = va.y; a[4 * q + 2] = va.z; a[4 * q + 3] = va.w;
        c[4 * q] = vc.x; c[4 * q + 1] = vc.y; c[4 * q + 2] = vc.z; c[4 * q + 3] = vc.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        a[k] = k0 + k < d ? xi[k0 + k] : 0.f;
        c[k] = k0 + k < d ? xj[k0 + k] : 0.f;
      }
    }
    // Zero padding adds (0 - 0)^2 = +0, which leaves the sum unchanged.
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const float dl = a[k] - c[k];
      t = t + dl * dl;
    }
  }
  return t;
}

// Core distance of each row from its candidate list: one thread per row
// recomputes its candidates exactly (or reads them from the exact distance
// matrix when one is materialised) and keeps the KMAX smallest; a row
// marked for the exact path streams every column instead.
// Grid: (ceil(n / 256), windows).
// Exact selection for one row: the ms-th smallest coordinate-order squared
// distance over its candidates (at most kGramCap), or over the whole window
// when the row has more (ties, or the filter off for the window).  xb: the
// window's points (device memory or LDS), di: the row of a materialised
// matrix (sqrt values) or null.  Sets bad for a non-finite full-row value.
template <int KMAX>
__device__ __forceinline__ float exact_row(const float* xb, const float* di, int n, int d, int ms, int i,
                                           const uint16_t* cr, int m, int& bad) {
  constexpr int CAP = kGramCap;
  const float* xi = xb + (size_t)i * d;
  float lst[KMAX];
#pragma unroll
  for (int q = 0; q < KMAX; ++q) lst[q] = INFINITY;
  if (m <= CAP) {
    // Four candidates at a time, their reads in flight together (the K
    // smallest values do not depend on the order they are offered in).  The
    // list's slots are 64-byte aligned, so four entries are one 8-byte load.
    int p = 0;
    for (; p + 4 <= m; p += 4) {
      const uint2 jj = *reinterpret_cast<const uint2*>(cr + p);
      const int j[4] = {(int)(jj.x & 0xFFFFu), (int)(jj.x >> 16), (int)(jj.y & 0xFFFFu), (int)(jj.y >> 16)};
      float t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] = di ? di[j[u]] : exact_sq(xi, xb + (size_t)j[u] * d, d);
#pragma unroll
      for (int u = 0; u < 4; ++u) list_offer<KMAX>(lst, t[u]);
    }
    for (; p < m; ++p) {
      const int j = cr[p];
      // The matrix holds sqrt(s); squares of its values order like s.
      const float t = di ? di[j] : exact_sq(xi, xb + (size_t)j * d, d);
      list_offer<KMAX>(lst, t);
    }
  } else {
    int j = 0;
    for (; j + 4 <= n; j += 4) {
      float t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] = di ? di[j + u] : exact_sq(xi, xb + (size_t)(j + u) * d, d);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        bad |= !isfinite(t[u]);
        list_offer<KMAX>(lst, t[u]);
      }
    }
    for (; j < n; ++j) {
      const float t = di ? di[j] : exact_sq(xi, xb + (size_t)j * d, d);
      bad |= !isfinite(t);
      list_offer<KMAX>(lst, t);
    }
  }
  float t = lst[0];
#pragma unroll
  for (int q = 1; q < KMAX; ++q)
    if (q == ms - 1) t = lst[q];
  return di ? t : sqrt(t);
}

template <int KMAX>
__global__ __launch_bounds__(256) void core_exact_kernel(const float* x, const float* dist, Windows win, int d,
                                                         int ms, const uint16_t* cand, const int32_t* ncand,
                                                         float* core, int32_t* status) {
  int64_t base;
  int n;
  int wid;
  win.get(blockIdx.y, wid, base, n);
  if ((int)(blockIdx.x * 256) >= n) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  int bad = 0;
  if (i < n) {
    const int64_t row = base + i;
    const float* di = dist ? dist + win.matrix(wid, base) + (size_t)i * n : nullptr;
    core[row] = exact_row<KMAX>(x + (size_t)base * d, di, n, d, ms, i, cand + row * kGramCap, ncand[row], bad);
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 1);
}

// LDS bytes above which core_exact_lds_kernel is not used.
constexpr uint32_t kExactLdsBytes = 65536;

// The same selection with the window's points staged in LDS first (one
// workgroup per window, n * d floats <= kExactLdsBytes): every candidate is
// read from LDS instead of gathered from device memory.
template <int KMAX>
__global__ __launch_bounds__(1024) void core_exact_lds_kernel(const float* x, Windows win, int d, int ms,
                                                             const uint16_t* cand, const int32_t* ncand,
                                                             float* core, int32_t* status) {
  extern __shared__ __align__(16) float xs[];
  int64_t base;
  int n;
  int wid;
  win.get(blockIdx.x, wid, base, n);
  const float* xb = x + (size_t)base * d;
  const int total = n * d;
  if (((((uintptr_t)xb) & 15) == 0) && (total & 3) == 0) {
    const float4* src = reinterpret_cast<const float4*>(xb);
    float4* dst = reinterpret_cast<float4*>(xs);
    for (int k = threadIdx.x; k < total / 4; k += blockDim.x) dst[k] = src[k];
  } else {
    for (int k = threadIdx.x; k < total; k += blockDim.x) xs[k] = xb[k];
  }
  __syncthreads();
  int bad = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int64_t row = base + i;
    core[row] = exact_row<KMAX>(xs, nullptr, n, d, ms, i, cand + row * kGramCap, ncand[row], bad);
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 1);
}

// Core distances core[point] of windows of at most kWinMaxN points and 16
// dimensions, both filter passes and the exact selection in one workgroup
// per window.  Grid: one workgroup per window of the launch, blockDim.x >=
// the largest window rounded up to 64, LDS gram_window_lds_bytes(largest
// window, d).  pairs: per window 16 n 32-bit entries of scratch (the
// candidate-list buffer of the chunked form, 64 bytes per point).
//
// The column fragment is the MFMA's first operand and the row fragment the
// second, so accumulator register r of lane l holds column 4 (l >> 4) + r of
// the block and row l & 15 of the set: each lane owns one row per set and
// sees four of its columns per block (the row's 16 columns of a block are
// spread over the lanes l & 15 + 16 k, k = 0..3).
//   Pass A: each lane keeps the M smallest upper bounds of its columns; U =
//           the ms-th smallest of the 4 M values of the row's four lanes (a
//           subset of the row, so U >= the ms-th smallest upper bound >= the
//           exact statistic).  M = 2 for ms <= 8, 4 for ms <= 16: 4 M >= ms.
//   Pass B: a column is a candidate when its lower bound is <= U.  Each
//           lane ORs the test results of eight blocks into one mask word
//           (bit 4 c + r: block c of the eight, register r) and only set
//           bits are expanded: the row's LDS counter counts the candidate,
//           and while the count is below kGramCap the pair (row << 16 |
//           column) is appended to the window's pair array (one LDS counter
//           per window): consecutive 4-byte stores, not scattered list
//           slots.  A pair past the array's capacity marks its row for the
//           exact path (bit 31 of its counter).
//   Exact:  the fragments are dead; the window's points are staged in LDS,
//           the pairs are sorted into per-row lists in LDS, and each thread
//           selects its row's core distance from its list, or from the
//           whole window when the row had more than kGramCap candidates
//           (exact_row, as core_exact_lds_kernel).
// M: values kept per lane in pass A (2, 4); KMAX: the exact selection's
// list length (8 for ms <= 8, 16 for ms <= 16); VEC: 16-dimensional rows
// read as 16-byte vectors (the input 16-byte aligned).
template <int M, int KMAX, bool VEC>
__global__ __launch_bounds__(1024) void core_window_kernel(const float* x, Windows win, int d, int ms,
                                                           uint32_t* pairs, float* core, int32_t* status) {
  constexpr int DK = kWinDK, CAP = kGramCap, RS = 4;
  constexpr uint32_t kFull = 0x80000000u;  // row counter bit: exact path
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ uint32_t sh_off, sh_pairs;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int grp = lane >> 4;
  int64_t base;
  int n;
  int wid;
  win.get(blockIdx.x, wid, base, n);
  const float* xb = x + (size_t)base * d;
  const int ncb = (n + 15) >> 4;
  uint32_t* rowcnt = reinterpret_cast<uint32_t*>(lds);         // [n] candidates per row
  uint32_t* fillcnt = rowcnt + n;                               // [n] list fill (exact stage)
  unsigned char* region = lds + gram_window_head(n);
  gram_s16x8* frag = reinterpret_cast<gram_s16x8*>(region);    // [cb][q][64 lanes]
  uint32_t* wpairs = pairs + (size_t)base * 16;
  const uint32_t cap = 16u * (uint32_t)n;
  if (tid == 0) sh_off = sh_pairs = 0;
  if (tid < n) rowcnt[tid] = fillcnt[tid] = 0;
  __syncthreads();

  // Column operands: thread j writes column j's eight 16-byte entries
  // [hi(-2x) | lo(-2x) | hi(-2x) | norm slots | zeros].  One column per
  // thread: blockDim.x >= 64 ceil(n / 64) >= 16 ncb.
  int bad = 0;
  float sqj = 0.f;
  const int j = tid;
  const bool colw = j < 16 * ncb;
  if (colw) {
    gram_s16x8 bh[2] = {}, bl[2] = {};
    if (j < n) {
      float v[DK];
      window_load<VEC>(xb + (size_t)j * d, d, v);
      sqj = window_norm(v, bad);
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float w = -2.f * v[8 * h + e];
          const uint32_t g = bf16_bits(w);
          bh[h][e] = (short)g;
          bl[h][e] = (short)bf16_bits(w - bf16_value(g));
        }
      // NaN compares false: a non-finite norm turns the filter off too.
      if (!(sqj <= 1.2676506002282294e30f)) atomicOr(&sh_off, 1u);
    }
    const int cl = j & 15;
    gram_s16x8* f0 = frag + (size_t)(j >> 4) * 128 + cl;
    f0[0] = bh[0];
    f0[16] = bh[1];
    f0[32] = bl[0];
    f0[48] = bl[1];
    f0[64] = bh[0];
    f0[80] = bh[1];
    f0[96] = norm_slots(fmaf(sqj, kWinC, sqj) + kGramEh, false, j >= n);
    f0[112] = gram_s16x8{};
  }
  if (__any(bad) && lane == 0) atomicOr(status, 2);
  __syncthreads();
  const int row0 = wave * 16 * RS;
  const bool active = row0 < n;
  if (sh_off) {
    // Filter off for this window: every row goes to the exact path.
    if (tid < n) rowcnt[tid] = kFull;
  } else {

  // Row operands of set g: row row0 + 16 g + (lane & 15) as
  // [hi(x) | hi(x) | lo(x) | norm slots | zeros] (lane holds K entries
  // 8 grp .. 8 grp + 7 of each 32-wide MFMA step).
  gram_bf16x8 av[RS][2];
  float sqr[RS];
  if (active) {
    // All four rows' loads are issued before any arithmetic.
    float vr[RS][DK];
#pragma unroll
    for (int g = 0; g < RS; ++g) window_load<VEC>(xb + (size_t)min(row0 + 16 * g + (lane & 15), n - 1) * d, d, vr[g]);
#pragma unroll
    for (int g = 0; g < RS; ++g) {
      const float (&v)[DK] = vr[g];
      int ignore = 0;
      sqr[g] = window_norm(v, ignore);
      // Both halves are split and the lane's half selected as a vector: an
      // index that depends on the lane (v[8 (grp & 1) + e]) would put the
      // coordinates in scratch memory.
      gram_s16x8 hi0, hi1, lo0, lo1;  // named values, not an array the select could index
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const uint32_t b0 = bf16_bits(v[e]), b1 = bf16_bits(v[8 + e]);
        hi0[e] = (short)b0;
        lo0[e] = (short)bf16_bits(v[e] - bf16_value(b0));
        hi1[e] = (short)b1;
        lo1[e] = (short)bf16_bits(v[8 + e] - bf16_value(b1));
      }
      const bool h1 = (grp & 1) != 0;
      av[g][0] = __builtin_bit_cast(gram_bf16x8, h1 ? hi1 : hi0);
      const gram_s16x8 lo_h = h1 ? lo1 : lo0;
      const gram_s16x8 q1 = grp < 2    ? lo_h
                            : grp == 2 ? norm_slots(fmaf(sqr[g], kWinC, sqr[g]) + kGramEh, true, false)
                                       : gram_s16x8{};
      av[g][1] = __builtin_bit_cast(gram_bf16x8, q1);
    }
  }

  // Bounds of one column block for the four sets: all eight MFMAs issue
  // before any result is read.
  auto block = [&](const gram_bf16x8& b0, const gram_bf16x8& b1, gram_f32x4 (&acc)[RS]) {
#pragma unroll
    for (int g = 0; g < RS; ++g) acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b0, av[g][0], gram_f32x4{}, 0, 0, 0);
#pragma unroll
    for (int g = 0; g < RS; ++g) acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b1, av[g][1], acc[g], 0, 0, 0);
  };
  auto load = [&](int cb, gram_bf16x8& b0, gram_bf16x8& b1) {
    b0 = __builtin_bit_cast(gram_bf16x8, frag[cb * 128 + lane]);
    b1 = __builtin_bit_cast(gram_bf16x8, frag[cb * 128 + 64 + lane]);
  };

  float U[RS];
  if (active) {
    // Pass A.
    float mn[RS][M];
#pragma unroll
    for (int g = 0; g < RS; ++g)
#pragma unroll
      for (int q = 0; q < M; ++q) mn[g][q] = INFINITY;
    gram_bf16x8 b0, b1;
    load(0, b0, b1);
    for (int cb = 0; cb < ncb; ++cb) {
      // The next block's fragments are read before this block's MFMAs (the
      // scheduling barrier keeps the compiler from sinking the reads to
      // their use), so the LDS latency overlaps this block's work.
      gram_bf16x8 n0, n1;
      load(min(cb + 1, ncb - 1), n0, n1);
      __builtin_amdgcn_sched_barrier(0);
      gram_f32x4 acc[RS];
      block(b0, b1, acc);
      b0 = n0;
      b1 = n1;
#pragma unroll
      for (int g = 0; g < RS; ++g)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // Ascending list of the M smallest: the new q-th smallest is the
          // median of (lst[q-1], lst[q], v); the smallest an unsigned
          // minimum of the bit patterns (positive bounds or +inf order like
          // their bits; no NaN: the filter runs on finite norms only).
          const float v = acc[g][r];
#pragma unroll
          for (int q = M - 1; q > 0; --q) mn[g][q] = __builtin_amdgcn_fmed3f(mn[g][q - 1], mn[g][q], v);
          mn[g][0] = __uint_as_float(min(__float_as_uint(v), __float_as_uint(mn[g][0])));
        }
    }
    // U: the ms-th smallest of the row's 4 M kept values (lanes l & 15 +
    // 16 k), by ms extractions of the minimum, one instance at a time.
#pragma unroll
    for (int g = 0; g < RS; ++g) {
      float all[4 * M];
#pragma unroll
      for (int q = 0; q < M; ++q) {
        all[q] = mn[g][q];
        all[M + q] = __shfl_xor(mn[g][q], 16);
        all[2 * M + q] = __shfl_xor(mn[g][q], 32);
        all[3 * M + q] = __shfl_xor(mn[g][q], 48);
      }
      float u = INFINITY;
      for (int it = 0; it < ms; ++it) {
        float m = all[0];
#pragma unroll
        for (int q = 1; q < 4 * M; ++q) m = fminf(m, all[q]);
        bool done = false;
#pragma unroll
        for (int q = 0; q < 4 * M; ++q) {
          const bool hit = !done && all[q] == m;
          all[q] = hit ? INFINITY : all[q];
          done = done || hit;
        }
        u = m;
      }
      U[g] = u;
    }
  }
  // Pass B operands: the lower-bound norm terms replace the upper ones.
  __syncthreads();  // every wave has finished reading pass A's fragments
  if (colw) frag[(size_t)(j >> 4) * 128 + 96 + (j & 15)] = norm_slots(fmaf(sqj, -kWinC, sqj) - kGramEh, false, j >= n);
  if (active && grp == 2) {
#pragma unroll
    for (int g = 0; g < RS; ++g)
      av[g][1] = __builtin_bit_cast(gram_bf16x8, norm_slots(fmaf(sqr[g], -kWinC, sqr[g]) - kGramEh, true, false));
  }
  __syncthreads();
  if (active) {
  // Pass B.
  gram_bf16x8 b0, b1;
  load(0, b0, b1);
  for (int c8 = 0; c8 < ncb; c8 += 8) {
    uint32_t word[RS] = {};
#pragma unroll
    for (int cc = 0; cc < 8; ++cc) {
      const int cb = c8 + cc;
      if (cb < ncb) {
        gram_bf16x8 n0, n1;
        load(min(cb + 1, ncb - 1), n0, n1);
        __builtin_amdgcn_sched_barrier(0);
        gram_f32x4 acc[RS];
        block(b0, b1, acc);
        b0 = n0;
        b1 = n1;
#pragma unroll
        for (int g = 0; g < RS; ++g)
#pragma unroll
          for (int r = 0; r < 4; ++r) word[g] |= (acc[g][r] > U[g] ? 0u : 1u) << (4 * cc + r);
      }
    }
    // Expand the (rare) set bits: column 16 (c8 + b / 4) + 4 grp + b % 4 of
    // row row0 + 16 g + (lane & 15); rows past the window are never listed.
#pragma unroll
    for (int g = 0; g < RS; ++g) {
      const int i = row0 + 16 * g + (lane & 15);
      uint32_t m = i < n ? word[g] : 0u;
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1u;
        const int jc = 16 * (c8 + (b >> 2)) + 4 * grp + (b & 3);
        if (jc < n) {
          const uint32_t pos = atomicAdd(&rowcnt[i], 1u);
          if (pos < (uint32_t)CAP) {
            const uint32_t k = atomicAdd(&sh_pairs, 1u);
            if (k < cap)
              wpairs[k] = ((uint32_t)i << 16) | (uint32_t)jc;
            else
              atomicOr(&rowcnt[i], kFull);
          }
        }
      }
    }
  }
  }  // active
  }  // filter on
  __syncthreads();  // every pair is appended; the fragments are dead

  // Exact selection.  The pairs were written by this workgroup before the
  // barrier, so its reads see them.
  float* xs = reinterpret_cast<float*>(region);  // [n][d]
  uint16_t* lists = reinterpret_cast<uint16_t*>(region + ((4u * (uint32_t)n * (uint32_t)d + 15u) & ~15u));
  const int total = n * d;
  if (((((uintptr_t)xb) & 15) == 0) && (total & 3) == 0) {
    for (int q = tid; q < total / 4; q += blockDim.x)
      reinterpret_cast<float4*>(xs)[q] = reinterpret_cast<const float4*>(xb)[q];
  } else {
    for (int q = tid; q < total; q += blockDim.x) xs[q] = xb[q];
  }
  const uint32_t np = min(sh_pairs, cap);
  for (uint32_t q = tid; q < np; q += blockDim.x) {
    const uint32_t pr = wpairs[q];
    const uint32_t i = pr >> 16;
    const uint32_t slot = atomicAdd(&fillcnt[i], 1u);  // < kGramCap: appended only below it
    lists[i * CAP + slot] = (uint16_t)(pr & 0xFFFFu);
  }
  __syncthreads();
  int badx = 0;
  for (int i = tid; i < n; i += blockDim.x) {
    const uint32_t rc = rowcnt[i];
    const int m = (rc & kFull) ? CAP + 1 : (int)rc;
    core[base + i] = exact_row<KMAX>(xs, nullptr, n, d, ms, i, lists + i * CAP, m, badx);
  }
  if (__any(badx) && lane == 0) atomicOr(status, 1);
}

}  // namespace hdb
