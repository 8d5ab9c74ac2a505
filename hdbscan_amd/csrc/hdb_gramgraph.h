// The MFMA Gram form of the graph stage (float32 windows on the
// materialised-matrix path).
//
// Instead of the exact N x N distance matrix (3 VALU operations per
// coordinate per pair), matrix cores evaluate for every pair
//   a_ij = |x_i|^2 + |x_j|^2 - 2 x_i . x_j
// from the split bf16 pieces of gram_prep_kernel (products hi.hi + hi.lo +
// lo.hi accumulated in float32 on top of the two norms).  The error analysis
// of hdb_gram.h bounds |a_ij - s_ij| <= c S + 2^-109, S = |x_i|^2 + |x_j|^2,
// where s_ij is the reference's coordinate-order float sum (reference
// hdbscan_gpu/_kernels.py:45-73) whose correctly rounded square root is the
// exact distance.  Consumers use
//   E_ij = c2 (n_i + n_j) + 2^-107,  c2 = 1.001 c + 4 u,
// which also covers the float32 rounding of n_i + n_j, of the product and of
// a_ij -+ E_ij, so lb = fl(a - E) <= s_ij <= fl(a + E) = ub.  Pairs with a
// squared norm above 2^100 (or not finite) are not bounded this way: tiles
// holding such a point store the exact sum s_ij itself, and E_ij = 0 there.
//
// From the bounds, fl(sqrt(lb)) <= d_ij <= fl(sqrt(ub)) for the exact
// distance d_ij (sqrt is correctly rounded and monotone), so every weight
// w = max(core_i, core_j, d_ij) is known to lie in an interval.  The
// consumers decide only what the intervals decide and compute the exact
// coordinate-order distance for what they leave open:
//   core_from_gram_kernel  the ms-th smallest upper bound U_i of a row, then
//                          the exact values of every column with lb <= U_i
//                          (they include the ms nearest);
//   prim_gram_kernel       Prim with interval keys; ties inside a bound are
//                          resolved exactly when a step's choice depends on
//                          them, otherwise deferred;
//   edge_fix_kernel        the exact weight (and, for deferred ties, the
//                          source) of every edge whose weight is an interval.
// The result is the reference's Prim edge sequence bit for bit.
#pragma once

#include "hdb_common.h"
#include "hdb_gram.h"
#include "hdb_prim.h"

namespace hdb {

constexpr float kGramNormLimit = 1.2676506002282294e30f;  // 2^100
constexpr float kGramEabs = 6.1629758220391547e-33f;      // 2^-107
constexpr int kGramTile = 128;                            // G tile side (one workgroup)
constexpr int kGramStageStride = kGramTile + 4;           // LDS row stride of the output stage (floats)
constexpr uint32_t kGramMatrixLds = (uint32_t)kGramTile * kGramStageStride * 4u;

__host__ __device__ constexpr float gram_c2(int dk) { return gram_c(dk) * 1.001f + 4.0f * 5.9604644775390625e-08f; }

// E_ij (0 when either norm is outside the bounded range).
__device__ __forceinline__ float gram_err(float c2, float ni, float nj) {
  return (ni <= kGramNormLimit && nj <= kGramNormLimit) ? c2 * (ni + nj) + kGramEabs : 0.f;
}

// Squared norms of the batch's points (an FMA chain; any order stays within
// the bound's gamma_D term) and the input check (HDB_STATUS_NONFINITE).
__global__ __launch_bounds__(256) void gram_norm_kernel(const float* x, int64_t points, int d, float* norms,
                                                        int32_t* status) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int bad = 0;
  if (p < points) {
    const float* xp = x + (size_t)p * d;
    float sq = 0.f;
    for (int k = 0; k < d; ++k) {
      const float v = xp[k];
      bad |= !isfinite(v);
      sq = fmaf(v, v, sq);
    }
    norms[p] = sq;
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 2);
}

// MFMA operand arrays of the Gram matrix, in the order the matrix cores take
// them: for window w, 16-point block k (points past the window are zero) and
// K-step q (32 bf16), entry (((w * nblk + k) * NQ + q) * 64 + lane) holds the
// 8 bf16 operand values of lane `lane`: point 16 k + (lane & 15), K indices
// 32 q + 8 (lane >> 4) ..., of A = [hi(x) | hi(x) | lo(x)] and of
// B = [hi(-2x) | lo(-2x) | hi(-2x)] (zero past 3 DK).  A wave then loads an
// operand block as one contiguous kilobyte.
template <int DK>
__global__ __launch_bounds__(256) void gram_operand_kernel(const float* x, int n, int d, int64_t entries, int nblk,
                                                           gram_s16x8* aop, gram_s16x8* bop) {
  constexpr int NQ = GramShape<DK, 1>::NQ;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= entries) return;
  const int lane = (int)(e & 63);
  const int64_t kq = e >> 6;
  const int q = (int)(kq % NQ);
  const int64_t kb = kq / NQ;
  const int64_t w = kb / nblk;
  const int i = (int)(kb % nblk) * 16 + (lane & 15);
  const int kk = 32 * q + 8 * (lane >> 4);
  const int seg = kk / DK, m = kk - seg * DK;
  gram_s16x8 av = {}, bv = {};
  if (seg < 3 && i < n) {
    const float* xp = x + ((size_t)w * n + i) * d;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const float v = m + t < d ? xp[m + t] : 0.f;
      const uint32_t h = bf16_bits(v);
      const uint32_t l = bf16_bits(v - bf16_value(h));
      const float v2 = -2.f * v;
      const uint32_t h2 = bf16_bits(v2);
      const uint32_t l2 = bf16_bits(v2 - bf16_value(h2));
      av[t] = (short)(seg == 2 ? l : h);
      bv[t] = (short)(seg == 1 ? l2 : h2);
    }
  }
  aop[e] = av;
  bop[e] = bv;
}

// One workgroup computes a kGramTile x kGramTile block of G[b] (256 threads,
// each wave a 64 x 64 quarter as 4 x 4 MFMA blocks of 16 x 16).  The
// diagonal is stored as +0 (s_ii = 0).  Grid: (tiles per side ^ 2, batch).
template <int DK>
__global__ __launch_bounds__(256) void gram_matrix_kernel(const float* x, int n, int d, const gram_s16x8* aop,
                                                          const gram_s16x8* bop, const float* norms, float* G,
                                                          int32_t* status) {
  using Sh = GramShape<DK, 1>;
  constexpr int NQ = Sh::NQ;
  extern __shared__ __align__(16) float stage[];  // [kGramTile][kGramStageStride]
  const int b = blockIdx.y;
  const int tps = (n + kGramTile - 1) / kGramTile;
  const int r0 = (blockIdx.x / tps) * kGramTile, c0 = (blockIdx.x % tps) * kGramTile;
  const int64_t base = (int64_t)b * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  float* gb = G + (size_t)base * n;

  // Points of the tile with a norm outside the bounded range send the whole
  // tile to the exact path.
  const int p = tid < kGramTile ? r0 + tid : c0 + tid - kGramTile;
  const float np = norms[base + min(p, n - 1)];
  if (__syncthreads_or(!(np <= kGramNormLimit))) {
    int bad = 0;
    for (int idx = tid; idx < kGramTile * kGramTile; idx += 256) {
      const int i = r0 + idx / kGramTile, j = c0 + idx % kGramTile;
      if (i >= n || j >= n) continue;
      const float s = i == j ? 0.f : exact_sq(x + (size_t)(base + i) * d, x + (size_t)(base + j) * d, d);
      bad |= !isfinite(s);
      gb[(size_t)i * n + j] = s;
    }
    if (__any(bad) && lane == 0) atomicOr(status, 1);
    return;
  }

  const int grp = lane >> 4, l16 = lane & 15;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  gram_f32x4 acc[4][4];
  {
    float nc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) nc[cb] = norms[base + min(c0 + wc + 16 * cb + l16, n - 1)];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      float nr[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) nr[r] = norms[base + min(r0 + wr + 16 * rb + 4 * grp + r, n - 1)];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = gram_f32x4{nr[0] + nc[cb], nr[1] + nc[cb], nr[2] + nc[cb], nr[3] + nc[cb]};
    }
  }
  // Operand blocks of the wave's 4 row blocks and 4 column blocks (blocks
  // past the window are clamped to its last one; their results are not
  // stored).
  const int nblk = (n + 15) / 16;
  const gram_s16x8* arow[4];
  const gram_s16x8* bcol[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ka = min((r0 + wr) / 16 + q, nblk - 1), kb = min((c0 + wc) / 16 + q, nblk - 1);
    arow[q] = aop + ((size_t)b * nblk + ka) * NQ * 64 + lane;
    bcol[q] = bop + ((size_t)b * nblk + kb) * NQ * 64 + lane;
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    gram_bf16x8 av[4], bv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      av[k] = __builtin_bit_cast(gram_bf16x8, arow[k][q * 64]);
      bv[k] = __builtin_bit_cast(gram_bf16x8, bcol[k][q * 64]);
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[rb], bv[cb], acc[rb][cb], 0, 0, 0);
  }
  // Through LDS so that each row of the tile is written as 16-byte vectors.
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wr + 16 * rb + 4 * grp + r, col = wc + 16 * cb + l16;
        stage[row * kGramStageStride + col] = (r0 + row == c0 + col) ? 0.f : acc[rb][cb][r];
      }
  __syncthreads();
  const int rows = min(kGramTile, n - r0), cols = min(kGramTile, n - c0);
  if (cols == kGramTile && (n & 3) == 0) {
    for (int idx = tid; idx < rows * (kGramTile / 4); idx += 256) {
      const int row = idx / (kGramTile / 4), c4 = idx % (kGramTile / 4);
      *reinterpret_cast<float4*>(gb + (size_t)(r0 + row) * n + c0 + 4 * c4) =
          *reinterpret_cast<const float4*>(stage + row * kGramStageStride + 4 * c4);
    }
  } else {
    for (int idx = tid; idx < rows * kGramTile; idx += 256) {
      const int row = idx / kGramTile, col = idx % kGramTile;
      if (col < cols) gb[(size_t)(r0 + row) * n + c0 + col] = stage[row * kGramStageStride + col];
    }
  }
}

// Core distances from G: one wave per row.  Pass A keeps each lane's KMAX
// smallest upper bounds and takes the ms-th smallest of their union, U (an
// upper bound of the exact statistic); pass B lists every column whose lower
// bound is <= U (all ms nearest are among them); their exact sums are then
// recomputed in coordinate order and the ms-th smallest gives the core
// distance (sqrt is monotone).  More than kCoreCands columns: the whole row
// is recomputed exactly.  Grid: batch * n / 4 workgroups of 256.
constexpr int kCoreCands = 128;

// The kth smallest (1-based) of the union of the 64 lanes' ascending lists
// of non-negative values (equal heads pop together; the minimum is an
// unsigned minimum of the bit patterns over DPP levels).  Uniform trip
// count: all lanes active.
template <int K>
__device__ __forceinline__ float wave_kth(float* lst, int kth) {
  int cnt = 0;
  float ans = INFINITY;
  for (int it = 0; it < kth && cnt < kth; ++it) {
    const float m = __uint_as_float(wave_umin32(__float_as_uint(lst[0])));
    const bool eq = lst[0] == m;
    cnt += __popcll(__ballot(eq));
    ans = m;
    if (eq) {
#pragma unroll
      for (int q = 0; q < K - 1; ++q) lst[q] = lst[q + 1];
      lst[K - 1] = INFINITY;
    }
  }
  return ans;
}

// Exact coordinate-order squared distances from row xj to up to three rows
// xa[0..P-1] (d <= 256 floats), evaluated by the whole wave: lane l loads
// coordinates l, l + 64, ... of every row (all loads issued together), squares
// the differences, and each pair's terms are then added in ascending
// coordinate order from the lanes' registers, exactly as the reference's loop
// adds them (t = 0; t = t + dl*dl); the P sums advance together.  Terms past
// d are +0 and leave a sum unchanged.  Arguments uniform; all lanes active.
template <int P>
__device__ __forceinline__ void wave_exact_sq(const float* xj, const float* const* xa, int d, float* out) {
  const int lane = threadIdx.x & 63;
  float vj[4], va[P][4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int k = lane + 64 * m;
    const bool in = k < d;
    vj[m] = in ? xj[k] : 0.f;
#pragma unroll
    for (int p = 0; p < P; ++p) va[p][m] = in ? xa[p][k] : 0.f;
  }
  float sq[P][4], t[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    t[p] = 0.f;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const float dl = va[p][m] - vj[m];
      sq[p][m] = dl * dl;
    }
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    if (64 * m >= d) break;
#pragma unroll
    for (int l = 0; l < 64; ++l)
#pragma unroll
      for (int p = 0; p < P; ++p) t[p] = t[p] + __uint_as_float(read_lane(__float_as_uint(sq[p][m]), l));
  }
#pragma unroll
  for (int p = 0; p < P; ++p) out[p] = t[p];
}

__device__ __forceinline__ float wave_exact_sq1(const float* xa, const float* xb, int d) {
  float o[1];
  const float* rows[1] = {xa};
  wave_exact_sq<1>(xb, rows, d, o);
  return o[0];
}

template <int KMAX>
__global__ __launch_bounds__(256) void core_from_gram_kernel(const float* G, const float* x, const float* norms,
                                                             int64_t rows, int n, int d, int ms, float c2,
                                                             float* core, int32_t* status) {
  __shared__ uint16_t cand[4][kCoreCands];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t r = (int64_t)blockIdx.x * 4 + wave;
  if (r >= rows) return;  // uniform per wave
  const int64_t base = r - r % n;
  const int i = (int)(r - base);
  const float* gi = G + (size_t)r * n;
  const float ni = norms[r];
  const float* xi = x + (size_t)r * d;
  // Pass A: U = the ms-th smallest upper bound of the row (>= the exact
  // statistic t), and Emax = the row's largest bound term.  The ms columns of
  // the exact ms smallest values have ub - s <= 2 E_j plus the rounding of
  // g +- E (below 2u |g| <= 0.01 E_j for every c2 used here), so
  // t >= U - 2.01 Emax =: L, taken a few ulps lower still (its own rounding).
  // Eight loads in flight.
  float lub[KMAX];
#pragma unroll
  for (int q = 0; q < KMAX; ++q) lub[q] = INFINITY;
  float emax = 0.f;
  for (int j0 = 0; j0 < n; j0 += 8 * 64) {
    float g[8], nj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = j0 + u * 64 + lane;
      g[u] = j < n ? gi[j] : INFINITY;
      nj[u] = j < n ? norms[base + j] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = j0 + u * 64 + lane;
      const float e = gram_err(c2, ni, nj[u]);
      emax = fmaxf(emax, e);
      const float ub = j == i ? 0.f : fmaxf(g[u] + e, 0.f);  // +inf past the row
#pragma unroll
      for (int q = KMAX - 1; q > 0; --q) lub[q] = __builtin_amdgcn_fmed3f(lub[q - 1], lub[q], ub);
      lub[0] = fminf(lub[0], ub);
    }
  }
  const float U = wave_kth<KMAX>(lub, ms);
  emax = __uint_as_float(wave_umax32(__float_as_uint(emax)));
  const float lc = U - 2.01f * emax;
  const float L = lc > 0.f ? __uint_as_float(__float_as_uint(lc) - 4u) : 0.f;
  // Pass B: columns with ub < L are certainly below t (count them); t is the
  // (ms - below)-th smallest exact value among the columns whose bounds meet
  // [L, U]; all others lie above U >= t.
  int below = 0, cnt = 0;
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    bool c = false, lo = false;
    if (j < n) {
      const float g = gi[j];
      const float e = gram_err(c2, ni, norms[base + j]);
      const float ub = j == i ? 0.f : fmaxf(g + e, 0.f);
      const float lb = j == i ? 0.f : fmaxf(g - e, 0.f);
      lo = ub < L;
      c = !lo && lb <= U;
    }
    below += __popcll(__ballot(lo));
    const uint64_t m = __ballot(c);
    const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
    if (c && pos < kCoreCands) cand[wave][pos] = (uint16_t)j;
    cnt += __popcll(m);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const int k = max(1, ms - below);
  int bad = 0;
  float t;
  if (cnt <= 64) {
    // One exact value per lane, each evaluated by the whole wave.
    float v = INFINITY;
    for (int q = 0; q < cnt; ++q) {
      const int j = cand[wave][q];
      const float s = j == i ? 0.f : wave_exact_sq1(xi, x + (size_t)(base + j) * d, d);
      if (lane == q) v = s;
    }
    float one[1] = {v};
    t = wave_kth<1>(one, k);
  } else {
    // Many values inside the bounds (ties): every column exactly.
    float lst[KMAX];
#pragma unroll
    for (int q = 0; q < KMAX; ++q) lst[q] = INFINITY;
    for (int j = lane; j < n; j += 64) {
      const float s = j == i ? 0.f : exact_sq(xi, x + (size_t)(base + j) * d, d);
      bad |= !isfinite(s);
      list_offer<KMAX>(lst, s);
    }
    t = wave_kth<KMAX>(lst, ms);
  }
  if (lane == 0) core[r] = sqrtf(t);
  if (__any(bad) && lane == 0) atomicOr(status, 1);
}

struct GramPrimArgs {
  const float* g;       // G [B][n][n]
  const float* x;       // points [B][n][d]
  const float* core;    // per point, or null when unit_core
  const float* norms;   // squared norms per point
  int n, d, unit_core;
  float c2;
  int32_t* edge_u;      // per point: edge i of a window at its first point + i
  int32_t* edge_v;
  float* edge_w;
  int32_t* edge_alt;    // -2: weight exact; -1: weight to recompute; >= 0: deferred tie, this source is newer
  int32_t* wflag;
  int32_t* status;
};

// Exact weights max(core_a, core_j, d_aj) of vertex j and the sources a[0..2]
// (coordinate-order distances, correctly rounded sqrt), evaluated by the
// whole wave (uniform arguments, all lanes active); +inf for a < 0.
__device__ __forceinline__ void wave_exact_w3(const GramPrimArgs& A, const float* xw, const float* cs, const int* a,
                                              int j, float* w) {
  const float* rows[3];
#pragma unroll
  for (int p = 0; p < 3; ++p) rows[p] = xw + (size_t)(a[p] < 0 ? j : a[p]) * A.d;
  float sq[3];
  wave_exact_sq<3>(xw + (size_t)j * A.d, rows, A.d, sq);
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const float dd = sqrtf(sq[p]);
    w[p] = a[p] < 0 ? INFINITY : A.unit_core ? dd : fmaxf(fmaxf(cs[a[p]], cs[j]), dd);
  }
}

// Interval state of one unvisited vertex: best weight in [lo, hi], its
// source sa, and sb >= 0 when a newer relaxation sb tied inside the bounds
// (the best is then sb's weight if strictly below sa's, else sa's); pc >= 0:
// a third relaxation by pc left open, to be resolved exactly.
struct GState {
  float lo, hi;
  int sa, sb, pc;
};

__device__ __forceinline__ bool gstate_exact(const GState& st) { return st.lo == st.hi && st.sb < 0; }

// Relaxation of vertex j by vertex c with the Gram value g (the reference's
// update: a strictly smaller weight replaces the best; reference
// hdbscan_gpu/_kernels.py:215-320).  Decided from the bounds when they
// decide it; a first undecided relaxation is deferred (sb), a second one is
// left pending (pc) for the wave's exact pass.
__device__ __forceinline__ void gstate_relax(GState& st, int c, float g, float mc, float E) {
  if (!(mc < st.hi)) return;  // w >= mc >= hi >= best
  const float ls = g - E;
  if (ls >= square_bound(st.hi)) return;  // d >= hi
  const float dlo = sqrtf(fmaxf(ls, 0.f)), dhi = sqrtf(g + E);
  const float wlo = fmaxf(mc, dlo), whi = fmaxf(mc, dhi);
  if (whi < st.lo) {
    st.lo = wlo;
    st.hi = whi;
    st.sa = c;
    st.sb = -1;
    return;
  }
  if (wlo >= st.hi) return;
  if (st.sb < 0) {
    st.sb = c;
    st.lo = fminf(st.lo, wlo);
    st.hi = fminf(st.hi, whi);
    return;
  }
  st.pc = c;
}

// Exact passes of a wave over its slots: every slot whose flag is set is
// replaced by its exact value and source (sa, then sb and pc in the order
// they relaxed it, each replacing the best only when strictly smaller).
// Called with all lanes active; `want` is evaluated per lane.
template <int VPT, typename Want>
__device__ __forceinline__ void gstate_exact_pass(GState* S, const GramPrimArgs& A, const float* xw, const float* cs,
                                                  int T, Want want) {
  const int lane = threadIdx.x & 63, w0 = threadIdx.x - lane;
#pragma unroll
  for (int s = 0; s < VPT; ++s) {
    uint64_t m = __builtin_amdgcn_ballot_w64(want(s, S[s]));
    while (m) {
      const int l = (int)__builtin_ctzll(m);
      m &= m - 1;
      const int j = w0 + l + s * T;
      const int src3[3] = {(int)read_lane((uint32_t)S[s].sa, l), (int)read_lane((uint32_t)S[s].sb, l),
                           (int)read_lane((uint32_t)S[s].pc, l)};
      float w3[3];
      wave_exact_w3(A, xw, cs, src3, j, w3);
      float w = w3[0];
      int src = src3[0];
#pragma unroll
      for (int p = 1; p < 3; ++p)
        if (w3[p] < w) {
          w = w3[p];
          src = src3[p];
        }
      if (lane == l) {
        S[s].lo = S[s].hi = w;
        S[s].sa = src;
        S[s].sb = -1;
        S[s].pc = -1;
      }
    }
  }
}

// (lo, vertex) key and the upper bound of the keyed state.
struct KeyHi {
  Cand<float> k;
  float h;
};

__device__ __forceinline__ KeyHi keyhi_pick(bool c, const KeyHi& x, const KeyHi& y) {
  KeyHi r;
  r.k = cand_pick(c, x.k, y.k);
  r.h = c ? x.h : y.h;
  return r;
}

// Two smallest keys (first < second), with the first's upper bound.
__device__ __forceinline__ void keyhi_top2(KeyHi& first, Cand<float>& second, const KeyHi& x) {
  const bool lt1 = x.k.less(first.k);
  const bool lt2 = x.k.less(second);
  second = cand_pick(lt1, first.k, cand_pick(lt2, x.k, second));
  first = keyhi_pick(lt1, x, first);
}

// Wave minimum of (key, hi) and the wave's second smallest key.
__device__ __forceinline__ void wave_top2(const KeyHi& a, const Cand<float>& b, KeyHi& m, Cand<float>& s) {
  m.k = wave_min(a.k);
  const uint64_t at = __builtin_amdgcn_ballot_w64(a.k.k == m.k.k);
  const int l = at ? (int)__builtin_ctzll(at) : 0;
  m.h = __uint_as_float(read_lane(__float_as_uint(a.h), l));
  const bool mine = a.k.k == m.k.k;
  s = wave_min(cand_pick(mine, b, a.k));
}

// The same over lanes 0..15 (one entry per wave of the workgroup).
__device__ __forceinline__ void row0_top2(const KeyHi& a, const Cand<float>& b, KeyHi& m, Cand<float>& s) {
  m.k = row0_min(a.k);
  const uint64_t at = __builtin_amdgcn_ballot_w64(a.k.k == m.k.k) & 0xFFFFull;
  const int l = at ? (int)__builtin_ctzll(at) : 0;
  m.h = __uint_as_float(read_lane(__float_as_uint(a.h), l));
  const bool mine = a.k.k == m.k.k;
  s = row0_min(cand_pick(mine, b, a.k));
}

// A choice is certain when the winner's upper bound (with its index) is
// below the runner-up's lower bound (with its index): no other vertex can
// hold a smaller (weight, index) key.
__device__ __forceinline__ bool certain(const KeyHi& m, const Cand<float>& s) {
  const uint64_t hk = ((uint64_t)__float_as_uint(m.h) << 32) | (uint32_t)m.k.k;
  return hk < s.k;
}

// Speculative Prim with interval keys (the round structure of
// prim_spec_kernel: the rows of the current vertex c and of the predicted
// next vertex p are read together; K' = keys after c, K'' = K' after p).
// Each choice is taken only when the interval keys make it certain;
// otherwise the open candidates are resolved exactly and the reduction is
// repeated (once suffices: afterwards the candidates are exact).  A choice
// in K' that needs resolution drops the round's speculation.
template <int VPT>
__global__ __launch_bounds__(1024) void prim_gram_kernel(GramPrimArgs A) {
  using C = Cand<float>;
  const int b = blockIdx.x;
  const int n = A.n;
  const int64_t base = (int64_t)b * n;
  const int tid = threadIdx.x, T = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, W = T >> 6;
  const bool unit = A.unit_core != 0;
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ C red_k[2][4][16];  // [parity][min K', second K', min K'', second K''][wave]
  __shared__ float red_h[2][2][16];
  __shared__ int sh_err;
  float* cs = reinterpret_cast<float*>(lds);  // core distances (0 with unit core)
  float* ns = cs + n;                          // squared norms
  const float* gg = A.g + (size_t)base * n;
  const float* xw = A.x + (size_t)base * A.d;
  if (tid == 0) sh_err = 0;
  for (int i = tid; i < n; i += T) {
    cs[i] = unit ? 0.f : A.core[base + i];
    ns[i] = A.norms[base + i];
  }
  __syncthreads();
  GState st[VPT];
  float cr[VPT], nr[VPT];
  uint32_t vis = 0;
#pragma unroll
  for (int s = 0; s < VPT; ++s) {
    const int j = tid + s * T;
    st[s].lo = st[s].hi = INFINITY;
    st[s].sa = st[s].sb = st[s].pc = -1;
    cr[s] = j < n ? cs[j] : 0.f;
    nr[s] = j < n ? ns[j] : 0.f;
    if (j >= n) vis |= 1u << s;
  }
  if (tid == 0) vis |= 1u;  // vertex 0 starts the tree
  const float rmax = RealTraits<float>::max_value();
  const C none = C::make(rmax, 0);
  const float c2 = A.c2;
  int cur = 0, pr = -1, par = 0, bad = 0;
  for (int step = 0; step < n - 1;) {
    bool spec = pr >= 0 && n - 1 - step >= 2;
    const float ccur = cs[cur], ncur = ns[cur];
    const float cpr = spec ? cs[pr] : 0.f, npr = spec ? ns[pr] : 0.f;
    float g1[VPT], g2[VPT];
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      const int j = tid + s * T;
      const bool un = !((vis >> s) & 1u);
      const bool l1 = un && fmaxf(ccur, cr[s]) < st[s].hi;
      const bool l2 = spec && un && j != pr && fmaxf(cpr, cr[s]) < st[s].hi;
      g1[s] = l1 ? gg[(size_t)cur * n + j] : 0.f;
      g2[s] = l2 ? gg[(size_t)pr * n + j] : 0.f;
    }
    GState s1[VPT], s2[VPT];
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      s1[s] = st[s];
      if (!((vis >> s) & 1u)) gstate_relax(s1[s], cur, g1[s], fmaxf(ccur, cr[s]), gram_err(c2, ncur, nr[s]));
    }
    gstate_exact_pass<VPT>(s1, A, xw, cs, T, [](int, const GState& e) { return e.pc >= 0; });
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      const int j = tid + s * T;
      s2[s] = s1[s];
      if (spec && j != pr && !((vis >> s) & 1u))
        gstate_relax(s2[s], pr, g2[s], fmaxf(cpr, cr[s]), gram_err(c2, npr, nr[s]));
    }
    if (spec) gstate_exact_pass<VPT>(s2, A, xw, cs, T, [](int, const GState& e) { return e.pc >= 0; });
    KeyHi m1, m2;
    C sec1, sec2;
    // Reductions of K' (and of K'' when speculating), at most twice each.
    for (int pass = 0; pass < 4; ++pass) {
      KeyHi a1{none, INFINITY}, a2{none, INFINITY};
      C b1 = none, b2 = none;
#pragma unroll
      for (int s = 0; s < VPT; ++s) {
        const int j = tid + s * T;
        if (!((vis >> s) & 1u)) {
          keyhi_top2(a1, b1, KeyHi{C::make(s1[s].lo, j), s1[s].hi});
          if (spec && j != pr) keyhi_top2(a2, b2, KeyHi{C::make(s2[s].lo, j), s2[s].hi});
        }
      }
      KeyHi w1, w2;
      C v1s, v2s;
      wave_top2(a1, b1, w1, v1s);
      if (spec) wave_top2(a2, b2, w2, v2s);
      if (lane == 0) {
        red_k[par][0][wave] = w1.k;
        red_k[par][1][wave] = v1s;
        red_h[par][0][wave] = w1.h;
        if (spec) {
          red_k[par][2][wave] = w2.k;
          red_k[par][3][wave] = v2s;
          red_h[par][1][wave] = w2.h;
        }
      }
      __syncthreads();
      KeyHi q1{none, INFINITY}, q2{none, INFINITY};
      C t1 = none, t2 = none;
      if (lane < W) {
        q1 = KeyHi{red_k[par][0][lane], red_h[par][0][lane]};
        t1 = red_k[par][1][lane];
        if (spec) {
          q2 = KeyHi{red_k[par][2][lane], red_h[par][1][lane]};
          t2 = red_k[par][3][lane];
        }
      }
      par ^= 1;
      row0_top2(q1, t1, m1, sec1);
      if (spec) row0_top2(q2, t2, m2, sec2);
      const bool ok1 = certain(m1, sec1);
      const bool hit = spec && ok1 && m1.k.index() == pr;
      const bool ok2 = !hit || certain(m2, sec2);
      if (ok1 && ok2) break;
      if (pass >= 2) {  // still open after resolution: only non-finite weights do this
        bad = 1;
        break;
      }
      // Resolve the open candidates exactly: in K' (the round then keeps
      // only K'), or in K'' after a certain hit.
      const bool in1 = !ok1;
      const float lim = in1 ? m1.h : m2.h;
      if (in1) spec = false;
      const int prv = pr;
      auto open = [&](int s, const GState& e) {
        return !((vis >> s) & 1u) && e.lo <= lim && !gstate_exact(e) && (in1 || tid + s * T != prv);
      };
      if (in1)
        gstate_exact_pass<VPT>(s1, A, xw, cs, T, open);
      else
        gstate_exact_pass<VPT>(s2, A, xw, cs, T, open);
    }
    const bool hit = spec && m1.k.index() == pr;
    const int v1 = m1.k.index();
    const int v2 = hit ? m2.k.index() : -1;
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      const int j = tid + s * T;
      if (j == v1 || j == v2) {
        const GState& e = j == v1 ? s1[s] : s2[s];
        // A visited winner means no vertex had a weight below REAL_MAX: the
        // graph is degenerate (only possible with non-finite weights).
        if ((vis >> s) & 1u) bad = 1;
        vis |= 1u << s;
        const size_t o = (size_t)base + step + (j == v1 ? 0 : 1);
        const bool exact = gstate_exact(e);
        A.edge_u[o] = exact && !(e.lo < rmax) ? 0 : e.sa;
        A.edge_v[o] = j;
        A.edge_w[o] = e.lo;
        A.edge_alt[o] = exact ? -2 : e.sb;
      }
      st[s] = hit ? s2[s] : s1[s];
    }
    step += hit ? 2 : 1;
    cur = hit ? v2 : v1;
    const C nxt = hit ? sec2 : sec1;
    pr = nxt.less(none) ? nxt.index() : -1;
  }
  if (bad) atomicOr(&sh_err, 1);
  __syncthreads();
  if (tid == 0) {
    A.wflag[b] = sh_err;
    if (sh_err) atomicOr(A.status, sh_err);
  }
}

// Exact weights (and deferred sources) of the edges whose weight Prim kept
// as an interval.  Grid: (ceil((n - 1) / 256), batch).
__global__ __launch_bounds__(256) void edge_fix_kernel(GramPrimArgs A) {
  const int b = blockIdx.y;
  const int e = blockIdx.x * 256 + threadIdx.x;
  const int n = A.n;
  int bad = 0;
  if (e < n - 1) {
    const int64_t base = (int64_t)b * n;
    const size_t o = (size_t)base + e;
    const int alt = A.edge_alt[o];
    if (alt != -2) {
      const float* xw = A.x + (size_t)base * A.d;
      const float* cs = A.unit_core ? nullptr : A.core + base;
      const int v = A.edge_v[o];
      int src = A.edge_u[o];
      auto wt = [&](int a) {
        if (a < 0) return (float)INFINITY;
        const float dd = sqrtf(exact_sq(xw + (size_t)a * A.d, xw + (size_t)v * A.d, A.d));
        return A.unit_core ? dd : fmaxf(fmaxf(cs[a], cs[v]), dd);
      };
      float w = wt(src);
      if (alt >= 0) {
        const float wb = wt(alt);
        if (wb < w) {
          w = wb;
          src = alt;
        }
      }
      const float rmax = RealTraits<float>::max_value();
      bad = !isfinite(w);
      A.edge_w[o] = w;
      A.edge_u[o] = w < rmax ? src : 0;
    }
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) {
    atomicOr(A.status, 1);
    atomicOr(A.wflag + b, 1);
  }
}

}  // namespace hdb
