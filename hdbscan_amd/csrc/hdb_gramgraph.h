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

// One workgroup computes a kGramTile x kGramTile block of G[b] (256 threads,
// each wave a 64 x 64 quarter as 4 x 4 MFMA blocks of 16 x 16).  The
// diagonal is stored as +0 (s_ii = 0).  Grid: (tiles per side ^ 2, batch).
template <int DK>
__global__ __launch_bounds__(256) void gram_matrix_kernel(const float* x, int n, int d, const uint16_t* pa,
                                                          const uint16_t* pb, const float* norms, float* G,
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
  const uint16_t* arow[4];
  const uint16_t* bcol[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    arow[q] = pa + (size_t)(base + min(r0 + wr + 16 * q + l16, n - 1)) * 2 * DK;
    bcol[q] = pb + (size_t)(base + min(c0 + wc + 16 * q + l16, n - 1)) * 2 * DK;
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    // Operand A of a row block: [hi(x) | hi(x) | lo(x)]; B of a column
    // block: [hi(-2x) | lo(-2x) | hi(-2x)] (zero past 3 DK).
    const int sa = gram_src<DK>(32 * q + 8 * grp, 2);
    const int sb = gram_src<DK>(32 * q + 8 * grp, 1);
    gram_bf16x8 av[4], bv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      av[k] = sa < 0 ? gram_bf16x8{} : *reinterpret_cast<const gram_bf16x8*>(arow[k] + sa);
      bv[k] = sb < 0 ? gram_bf16x8{} : *reinterpret_cast<const gram_bf16x8*>(bcol[k] + sb);
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
// (equal heads pop together); uniform trip count.
template <int K>
__device__ __forceinline__ float wave_kth(float* lst, int kth) {
  int cnt = 0;
  float ans = INFINITY;
  for (int it = 0; it < kth && cnt < kth; ++it) {
    float m = lst[0];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) m = fminf(m, __shfl_xor(m, o));
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
  float lst[KMAX];
#pragma unroll
  for (int q = 0; q < KMAX; ++q) lst[q] = INFINITY;
  for (int j = lane; j < n; j += 64) {
    const float g = gi[j];
    const float ub = j == i ? 0.f : g + gram_err(c2, ni, norms[base + j]);
    if (ub < lst[KMAX - 1]) {
#pragma unroll
      for (int q = KMAX - 1; q > 0; --q) lst[q] = __builtin_amdgcn_fmed3f(lst[q - 1], lst[q], ub);
      lst[0] = ub < lst[0] ? ub : lst[0];
    }
  }
  const float U = wave_kth<KMAX>(lst, ms);
  int cnt = 0;
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    bool c = false;
    if (j < n) {
      const float g = gi[j];
      const float lb = j == i ? 0.f : g - gram_err(c2, ni, norms[base + j]);
      c = lb <= U;
    }
    const uint64_t m = __ballot(c);
    const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
    if (c && pos < kCoreCands) cand[wave][pos] = (uint16_t)j;
    cnt += __popcll(m);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int q = 0; q < KMAX; ++q) lst[q] = INFINITY;
  int bad = 0;
  if (cnt <= kCoreCands) {
    for (int k = lane; k < cnt; k += 64) {
      const int j = cand[wave][k];
      const float s = j == i ? 0.f : exact_sq(xi, x + (size_t)(base + j) * d, d);
      list_offer<KMAX>(lst, s);
    }
  } else {
    for (int j = lane; j < n; j += 64) {
      const float s = j == i ? 0.f : exact_sq(xi, x + (size_t)(base + j) * d, d);
      bad |= !isfinite(s);
      list_offer<KMAX>(lst, s);
    }
  }
  const float t = wave_kth<KMAX>(lst, ms);
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

// Exact weight of the pair (a, j): max(core_a, core_j, d_aj) from the
// points (coordinate order, correctly rounded sqrt); +inf for a < 0 (never
// relaxed).
__device__ __forceinline__ float gram_exact_w(const GramPrimArgs& A, const float* xw, const float* cs, int a,
                                              int j) {
  if (a < 0) return INFINITY;
  const float dd = sqrtf(exact_sq(xw + (size_t)a * A.d, xw + (size_t)j * A.d, A.d));
  return A.unit_core ? dd : fmaxf(fmaxf(cs[a], cs[j]), dd);
}

// Interval state of one unvisited vertex: best weight in [lo, hi], its
// source sa, and sb >= 0 when a newer relaxation sb tied inside the bounds
// (the best is then sb's weight if strictly below sa's, else sa's).
struct GState {
  float lo, hi;
  int sa, sb;
};

// Exact value and source of a state (a deferred tie resolved).
__device__ __forceinline__ void gstate_resolve(GState& st, const GramPrimArgs& A, const float* xw, const float* cs,
                                               int j) {
  if (st.lo == st.hi && st.sb < 0) return;
  float w = gram_exact_w(A, xw, cs, st.sa, j);
  int src = st.sa;
  if (st.sb >= 0) {
    const float wb = gram_exact_w(A, xw, cs, st.sb, j);
    if (wb < w) {
      w = wb;
      src = st.sb;
    }
  }
  st.lo = st.hi = w;
  st.sa = src;
  st.sb = -1;
}

// Relaxation of vertex j by vertex c with the Gram value g (the reference's
// update: a strictly smaller weight replaces the best; reference
// hdbscan_gpu/_kernels.py:215-320).  Decided from the bounds when they
// decide it; a first undecided relaxation is deferred (sb), a second one is
// resolved exactly at once.
__device__ __forceinline__ void gstate_relax(GState& st, const GramPrimArgs& A, const float* xw, const float* cs,
                                             int j, int c, float g, float mc, float E) {
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
  gstate_resolve(st, A, xw, cs, j);
  const float wc = gram_exact_w(A, xw, cs, c, j);
  if (wc < st.lo) {
    st.lo = st.hi = wc;
    st.sa = c;
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
    st[s].sa = st[s].sb = -1;
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
      const int j = tid + s * T;
      s1[s] = st[s];
      if (!((vis >> s) & 1u)) {
        gstate_relax(s1[s], A, xw, cs, j, cur, g1[s], fmaxf(ccur, cr[s]), gram_err(c2, ncur, nr[s]));
        s2[s] = s1[s];
        if (spec && j != pr)
          gstate_relax(s2[s], A, xw, cs, j, pr, g2[s], fmaxf(cpr, cr[s]), gram_err(c2, npr, nr[s]));
      } else {
        s2[s] = s1[s];
      }
    }
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
#pragma unroll
      for (int s = 0; s < VPT; ++s) {
        const int j = tid + s * T;
        if ((vis >> s) & 1u) continue;
        if (in1) {
          if (s1[s].lo <= lim) gstate_resolve(s1[s], A, xw, cs, j);
        } else if (j != pr && s2[s].lo <= lim) {
          gstate_resolve(s2[s], A, xw, cs, j);
        }
      }
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
        const bool exact = e.lo == e.hi && e.sb < 0;
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
