// Exact distance matrix through a certified MFMA filter (float32 windows on
// the materialised-matrix path).
//
// The plain matrix path evaluates every pair's exact distance on the VALU
// (3 operations per coordinate) and Prim then reads rows of that matrix.
// Prim, however, needs the exact value of a pair only where it can decide a
// step: almost every pair is far from both of its points' neighbourhoods and
// enters Prim only as a key that some nearer pair beats.  Here matrix cores
// bound every pair and only the pairs the bounds cannot dismiss are
// evaluated exactly:
//
//   gx_prep_kernel    squared norms, the input check, and the bf16 pieces
//                     hi(x) | lo(x) in the order the matrix cores take them;
//   gx_bound_kernel   per row i, U2_i: a certified upper bound of the R-th
//                     smallest squared distance of the row (R = max(ms, 8),
//                     self included), from the hi.hi product over a sample of
//                     columns (a subset's order statistic bounds the row's);
//   gx_matrix_kernel  per pair of the upper triangle, the three-product lower
//                     bound lb <= s of hdb_gram.h.  Pairs with lb <= U2_i or
//                     lb <= U2_j are "near" and listed; every other pair
//                     stores minus a lower bound of its distance (the sign
//                     bit marks it) in both halves of the matrix;
//   gx_near_kernel    the exact coordinate-order distance of every near pair,
//                     stored in both halves (so row r holds the same value
//                     for (r, c) as row c for (c, r)); near pairs with
//                     s <= U2_i are appended to row i's core list (they
//                     include the ms nearest);
//   gx_core_kernel    the core distance: the ms-th smallest of the row
//                     (self included), from the core list;
//   prim_gx_kernel    the speculative two-row Prim of hdb_prim.h over these
//                     rows.  Every unvisited vertex keeps the smallest weight
//                     seen in the rows read so far -- exact, or a lower bound
//                     when the last entry that lowered it was a far one (a
//                     flag bit).  If a step's argmin is flagged, its exact
//                     key is computed first (resolution, below) and the
//                     argmin is taken again, in a round that reads no row
//                     and speculates on the resolved vertex with the row
//                     the resolution read.
//
// Resolution of vertex r: its row holds, for each visited c, the exact
// distance or a lower bound lbd; for the latter an upper bound follows from
// lbd and the pair's bound width, so T = the smallest upper weight over
// visited c bounds r's key from above, and only far pairs whose lower weight
// is <= T are evaluated exactly.  r's key is the smallest weight over visited
// c; its source the earliest visited c with that weight -- what the
// reference's strict-less update (hdbscan_gpu/_kernels.py:215-320) keeps.
//
// Why the steps equal the reference's: an unvisited vertex's key b_j is the
// smallest of the (exact or lower-bound) weights of the entries read for it,
// each <= its true weight, so the true key K_j >= b_j.  If b_j is not
// flagged it is an exact weight and every lower bound read was >= it, so
// K_j = b_j, and its source is the first visited vertex that reached it (an
// equal weight read later does not replace it, as in the reference).  An
// unflagged winner of the argmin of (b, vertex) therefore has the smallest
// true key, lowest index among equals, and the reference's source.
//
// Error bounds.  The matrix kernel's lb is hdb_gram.h's lower form: the
// three split-bf16 products accumulated on the per-point terms
// t_i = n_i (1 - c) - 2^-110, so lb <= s for the reference's coordinate-order
// sum s.  Its upper form (1 + c, + 2^-110) would bound s from above, and the
// two differ by 2 c S + 2^-108 plus at most c S of rounding each, so
// s <= lb + 4 E with E = c2 S + 2^-107 (c2 = 1.001 c + 4 u, S = n_i + n_j).
// Pass A uses the hi.hi product alone, which differs from the three-product
// value by 2 |sum hi lo + lo hi| <= 2^-7 (1 + 2^-6) S: its upper bound adds
// E1 = (c2 + 0.008) S + 2^-107 to n_i + n_j - 2 hi . hi.  Tiles holding a
// point whose squared norm exceeds 2^100 store exact distances only (and pass
// A gives such rows U2 = +inf).  Lower bounds of distances are sqrt(lb) from
// the hardware square root scaled by 1 - 2^-18 (the instruction's error is a
// few ulp); resolution's upper bounds square the stored lower bound back with
// margins 2^-14 and 2^-18, which exceed the roundings involved.
#pragma once

#include "hdb_common.h"
#include "hdb_gram.h"
#include "hdb_gramgraph.h"
#include "hdb_prim.h"

namespace hdb {

constexpr int kGxTile = 128;       // matrix tile side
constexpr int kGxCoreCap = 64;     // core-list entries per point before the row is scanned
constexpr int kGxNearPerPoint = 32;  // near-pair list capacity per window: 32 entries per point
constexpr float kGxPassA = 0.008f;                    // hi.hi-only extra relative to S
constexpr float kGxDown = 0.99999618530273438f;       // 1 - 2^-18
constexpr float kGxUp = 1.0000038146972656f;          // 1 + 2^-18
constexpr float kGxSq = 1.00006103515625f;            // 1 + 2^-14

// Operand entry of window w, 16-point block k, K-step q (of 2 * DK / 32: the
// hi pieces, then the lo pieces) and lane: 8 bf16 of point 16 k + (lane & 15),
// coordinates 32 q' + 8 (lane >> 4) ... (q' = q mod DK / 32).  A wave loads a
// block's K-step as one contiguous kilobyte; the same layout serves as the
// MFMA's A and B operand.
__host__ __device__ constexpr size_t gx_operand_entries(int dk, int64_t batch, int64_t n) {
  return (size_t)batch * (size_t)((n + 15) / 16) * (size_t)(2 * dk / 32) * 64;
}

typedef float gx_f32x4 __attribute__((ext_vector_type(4)));

// Matrix entries are written once and read by the Prim much later: streaming
// (non-temporal) stores keep them from evicting the MFMA operands from L2.
__device__ __forceinline__ void gx_store(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void gx_store4(float* p, float a, float b, float c, float d) {
  __builtin_nontemporal_store(gx_f32x4{a, b, c, d}, reinterpret_cast<gx_f32x4*>(p));
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup L to XCD
// L mod 8), each XCD with its own L2.  The kernels that reread a window's
// operands map the workgroups of window b to XCD b mod 8, in window order:
// L = 8 (per * (b / 8) + item) + b mod 8 for the item-th of `per`
// workgroups of window b.  Grid: 8 * ceil(batch / 8) * per workgroups; the
// ones past the batch exit.
constexpr int kXcds = 8;
__device__ __forceinline__ bool xcd_item(int64_t L, int64_t batch, int per, int& b, int& item) {
  const int64_t k = L / kXcds;
  b = (int)((k / per) * kXcds + L % kXcds);
  item = (int)(k % per);
  return b < batch;
}
__host__ inline unsigned xcd_grid(int64_t batch, int per) {
  return (unsigned)(((batch + kXcds - 1) / kXcds) * kXcds * per);
}

// One workgroup per 16-point block: the block's coordinates are read as
// contiguous vectors into LDS, one thread per point forms its squared norm
// (an FMA chain; any order stays within the bound's gamma_D term), and each
// thread writes whole 16-byte operand entries; the window's near-pair count
// is zeroed.  Grid: (nblk, batch).
template <int DK>
__global__ __launch_bounds__(256) void gx_prep_kernel(const float* x, int n, int d, int nblk, gram_s16x8* op,
                                                      float* norms, int32_t* lcnt, int32_t* ncnt, int32_t* status) {
  constexpr int NQ1 = DK / 32;
  __shared__ float sx[16][DK + 1];
  const int w = blockIdx.y, k = blockIdx.x, tid = threadIdx.x;
  if (k == 0 && tid == 0) ncnt[w] = 0;
  const int p0 = 16 * k, pts = min(16, n - p0);
  const float* xb = x + ((size_t)w * n + p0) * d;
  int bad = 0;
  for (int e = tid; e < 16 * DK; e += 256) {
    const int r = e / DK, c = e - r * DK;
    const float v = r < pts && c < d ? xb[(size_t)r * d + c] : 0.f;
    bad |= !isfinite(v);
    sx[r][c] = v;
  }
  __syncthreads();
  if (tid < pts) {
    float sq = 0.f;
    for (int c = 0; c < d; ++c) sq = fmaf(sx[tid][c], sx[tid][c], sq);
    norms[(size_t)w * n + p0 + tid] = sq;
    lcnt[(size_t)w * n + p0 + tid] = 0;
  }
  gram_s16x8* ob = op + ((size_t)w * nblk + k) * (2 * NQ1) * 64;
  for (int e = tid; e < NQ1 * 64; e += 256) {
    const int q = e >> 6, lane = e & 63;
    const int r = lane & 15, c0 = 32 * q + 8 * (lane >> 4);
    gram_s16x8 hv, lv;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const float v = sx[r][c0 + t];
      const uint32_t h = bf16_bits(v);
      hv[t] = (short)h;
      lv[t] = (short)bf16_bits(v - bf16_value(h));
    }
    ob[(size_t)q * 64 + lane] = hv;
    ob[(size_t)(NQ1 + q) * 64 + lane] = lv;
  }
  if (__any(bad) && (tid & 63) == 0) atomicOr(status, 2);
}

// Pass A.  Each wave takes RB 16-row blocks of one window and sweeps the
// columns [0, cols); every lane keeps, per row it holds, its two smallest
// upper bounds (one column residue class mod 16).  The R-th smallest of the
// 16 lanes' lists of a row is >= the R-th smallest over all sampled columns,
// which is >= the row's R-th smallest exact squared distance (the pair
// (i, i) enters with an upper bound of its zero; rows outside the sample get
// their zero offered once).  Norms outside the bounded range give +inf
// bounds.  Grid: xcd_grid(batch, ceil(n / (64 RB))).
template <int DK, int RB>
__global__ __launch_bounds__(256) void gx_bound_kernel(const gram_s16x8* op, const float* norms, int64_t batch,
                                                       int n, int nblk, int cols, int R, float e1c, float* U2) {
  constexpr int NQ1 = DK / 32, NQ2 = 2 * NQ1;
  int b, item;
  if (!xcd_item(blockIdx.x, batch, (n + 64 * RB - 1) / (64 * RB), b, item)) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int row0 = (item * 4 + wave) * 16 * RB;
  if (row0 >= n) return;  // uniform per wave; no barrier below
  const int g = lane >> 4, l16 = lane & 15;
  const float* nb = norms + (size_t)b * n;
  gram_bf16x8 av[RB][NQ1];
  float hi[RB][4], ei[RB][4], lst[RB][4][2];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    const int k = min(row0 / 16 + rb, nblk - 1);
#pragma unroll
    for (int q = 0; q < NQ1; ++q)
      av[rb][q] = __builtin_bit_cast(gram_bf16x8, op[((size_t)(b * nblk + k) * NQ2 + q) * 64 + lane]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int i = row0 + 16 * rb + 4 * g + t;
      const float ni = nb[min(i, n - 1)];
      const bool ok = ni <= kGramNormLimit;
      hi[rb][t] = ok ? -0.5f * ni : -INFINITY;
      ei[rb][t] = ok ? e1c * ni : INFINITY;
      lst[rb][t][0] = (i >= cols && l16 == 0) ? 0.f : INFINITY;
      lst[rb][t][1] = INFINITY;
    }
  }
  const int cblocks = (cols + 15) / 16;
  for (int cb = 0; cb < cblocks; ++cb) {
    const int j = 16 * cb + l16;
    const float nj = nb[min(j, n - 1)];
    const bool ok = nj <= kGramNormLimit && j < cols;
    const float hj = ok ? -0.5f * nj : -INFINITY;
    const float ej = ok ? e1c * nj + kGramEabs : INFINITY;
    gram_bf16x8 bv[NQ1];
#pragma unroll
    for (int q = 0; q < NQ1; ++q)
      bv[q] = __builtin_bit_cast(gram_bf16x8, op[((size_t)(b * nblk + cb) * NQ2 + q) * 64 + lane]);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      gram_f32x4 acc;
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = hi[rb][t] + hj;
#pragma unroll
      for (int q = 0; q < NQ1; ++q) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[rb][q], bv[q], acc, 0, 0, 0);
#pragma unroll
      for (int t = 0; t < 4; ++t) list_offer<2>(lst[rb][t], fmaf(-2.f, acc[t], ei[rb][t] + ej));
    }
  }
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float u = group16_kth<2>(lst[rb][t], R, lane);
      const int i = row0 + 16 * rb + 4 * g + t;
      if (l16 == 0 && i < n) U2[(size_t)b * n + i] = u;
    }
}

__device__ __forceinline__ void gx_append(int32_t* lcnt, float* lval, int64_t p, float v) {
  const int k = atomicAdd(lcnt + p, 1);
  if (k < kGxCoreCap) lval[(size_t)p * kGxCoreCap + k] = v;
}

// 128 x 128 tiles of the upper triangle of each window's matrix (tile t of
// the triangle in row order), each quarter of 64 x 64 as 4 x 4 MFMA blocks of
// 16 x 16: the three split-bf16 products accumulated on the lower norm terms
// give lb (a lower bound of the reference's squared distance).  A pair is near
// when lb <= max(U2_i, U2_j); far entries store minus a lower bound of the
// distance in both halves, near pairs go to the window's list for
// gx_near_kernel (pairs past the list's capacity are evaluated here).
// Off-diagonal tiles inside the window store every entry, a near one with its
// lower bound, which the exact value replaces later; other tiles store far
// entries only.  Tiles holding a point whose squared norm exceeds 2^100 store
// exact distances.
//
// Wave-specialised (512 threads, one workgroup per CU): waves 0-3 (MFMA
// waves) compute tile k's quarters while waves 4-7 (store waves) classify
// and store tile k - 1; the accumulators pass through LDS
// (lane-major, one 64 x 64 quarter per wave pair) between two barriers per
// tile.  A wave's memory counter is in order and counts stores too, so a
// wave that both stores and loads waits for its stores at its next load;
// here the MFMA waves never store to device memory and the store waves load
// from it only on the rare exact paths (exact tiles, pairs past the list's
// capacity); their near-list reservation is made before the tile's stores,
// so it waits only for stores a tile older.  Per iteration k:
//   A: MFMA waves: products of tile k;  store waves: tile k - 1 from LDS;
//   barrier;
//   B: MFMA waves: accumulators of tile k to LDS, norm terms and U2 of
//      tile k + 1 (parity (k + 1) & 1), its exact-tile flag (k + 1) % 3;
//   barrier.
// Tile ids: workgroup g takes g, g + G, ... (G a multiple of 8 keeps g on
// its XCD).
template <int DK>
__global__ __launch_bounds__(512) void gx_matrix_kernel(const float* x, int64_t batch, int n, int d, int nblk,
                                                           const gram_s16x8* op, const float* norms, const float* U2,
                                                           float* M, int32_t* lcnt, float* lval, int32_t* ncnt,
                                                           uint32_t* nlist, int cap, int32_t* status) {
  constexpr int NQ1 = DK / 32, NQ2 = 2 * NQ1, NQ3 = 3 * NQ1;
  constexpr float kLower = gram_c(DK);
  __shared__ float s_acc[4][64 * 64];
  __shared__ float s_n[2][2][kGxTile];
  __shared__ float s_u[2][2][kGxTile];
  __shared__ int s_exact[3];
  __shared__ uint16_t s_pos[4][kGxTile * kGxTile / 4];
  const int nt = (n + kGxTile - 1) / kGxTile;
  const int per = nt * (nt + 1) / 2;
  const int64_t ids = ((batch + kXcds - 1) / kXcds) * kXcds * per;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool mw = wave < 4;
  const int q = wave & 3;
  const int wr = (q >> 1) * 64, wc = (q & 1) * 64;
  const int g = lane >> 4, l16 = lane & 15;
  const int qt = tid & 255;  // thread index within its four waves
  int bad = 0;
  // Tile k of this workgroup: false past the last tile id or for a window
  // index past the batch (uniform).
  auto tile = [&](int64_t k, int& b, int& r0, int& c0, bool& diag) -> bool {
    const int64_t L = blockIdx.x + k * gridDim.x;
    int t;
    if (L >= ids || !xcd_item(L, batch, per, b, t)) return false;
    int ti = 0;
    while (t >= nt - ti) {
      t -= nt - ti;
      ++ti;
    }
    r0 = ti * kGxTile;
    c0 = (ti + t) * kGxTile;
    diag = t == 0;
    return true;
  };
  // MFMA waves: the lower norm terms n (1 - c) - 2^-110 and U2 of tile k's
  // rows and columns; a point outside the bounded range marks the tile exact.
  auto load_terms = [&](int64_t k) {
    int b, r0, c0;
    bool diag;
    if (!tile(k, b, r0, c0, diag)) return;
    const int side = qt >> 7, kk = qt & (kGxTile - 1);
    const int p = (side ? c0 : r0) + kk;
    const bool in = p < n;
    const int64_t base = (int64_t)b * n;
    const float npt = in ? norms[base + p] : 0.f;
    s_n[k & 1][side][kk] = fmaf(npt, -kLower, npt) - kGramEh;
    s_u[k & 1][side][kk] = in ? U2[base + p] : -INFINITY;
    if (!(npt <= kGramNormLimit)) atomicOr(&s_exact[k % 3], 1);
  };
  if (tid == 0) s_exact[0] = s_exact[1] = s_exact[2] = 0;
  __syncthreads();
  if (mw) load_terms(0);
  __syncthreads();
  gram_f32x4 acc[4][4];
  for (int64_t k = 0;; ++k) {
    const bool have_k = blockIdx.x + k * gridDim.x < ids;
    if (!have_k && (k == 0 || blockIdx.x + (k - 1) * gridDim.x >= ids)) break;  // uniform
    bool prod = false;  // MFMA waves: this quarter of tile k was computed
    if (mw) {
      if (tid == 0) s_exact[(k + 1) % 3] = 0;
      int b, r0, c0;
      bool diag;
      if (have_k && tile(k, b, r0, c0, diag) && !s_exact[k % 3] && !(diag && wr > wc)) {
        prod = true;
        const float (*sn)[kGxTile] = s_n[k & 1];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              acc[rb][cb][r] = -0.5f * (sn[0][wr + 16 * rb + 4 * g + r] + sn[1][wc + 16 * cb + l16]);
        const gram_s16x8* opl = op + lane;
        size_t aoff[4], boff[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          aoff[kk] = (size_t)(b * nblk + min((r0 + wr) / 16 + kk, nblk - 1)) * NQ2 * 64;
          boff[kk] = (size_t)(b * nblk + min((c0 + wc) / 16 + kk, nblk - 1)) * NQ2 * 64;
        }
        // K-steps: A = hi | hi | lo, B = hi | lo | hi (products hi.hi,
        // hi.lo, lo.hi); blocks past the window are clamped to its last one
        // (their results are not stored).
#pragma unroll 2
        for (int qq3 = 0; qq3 < NQ3; ++qq3) {
          const int seg = qq3 / NQ1, qq = qq3 % NQ1;
          const int qa = seg == 2 ? NQ1 + qq : qq, qb = seg == 1 ? NQ1 + qq : qq;
          gram_bf16x8 av[4], bv[4];
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            av[kk] = __builtin_bit_cast(gram_bf16x8, opl[aoff[kk] + (size_t)qa * 64]);
            bv[kk] = __builtin_bit_cast(gram_bf16x8, opl[boff[kk] + (size_t)qb * 64]);
          }
#pragma unroll
          for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
              acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[rb], bv[cb], acc[rb][cb], 0, 0, 0);
        }
      }
    } else if (k > 0) {
      int b, r0, c0;
      bool diag;
      if (tile(k - 1, b, r0, c0, diag)) {
        const int64_t base = (int64_t)b * n;
        float* mb = M + (size_t)base * n;
        const float* xb = x + (size_t)base * d;
        const float (*su)[kGxTile] = s_u[(k - 1) & 1];
        if (s_exact[(k - 1) % 3]) {
          // Every pair of the tile exactly (a point beyond the bounded range).
          for (int idx = qt; idx < kGxTile * kGxTile; idx += 256) {
            const int li = idx >> 7, lj = idx & 127;
            const int i = r0 + li, j = c0 + lj;
            if (i >= n || j >= n || (diag && i > j)) continue;
            if (i == j) {
              gx_store(mb + (size_t)i * n + j, 0.f);
              continue;
            }
            const float sq = exact_sq(xb + (size_t)i * d, xb + (size_t)j * d, d);
            bad |= !isfinite(sq);
            const float dv = sqrtf(sq);
            gx_store(mb + (size_t)i * n + j, dv);
            gx_store(mb + (size_t)j * n + i, dv);
            if (sq <= su[0][li]) gx_append(lcnt, lval, base + i, dv);
            if (sq <= su[1][lj]) gx_append(lcnt, lval, base + j, dv);
          }
        } else if (!(diag && wr > wc)) {
          // The lower-left quarter of a diagonal tile holds no pair i < j.
          const float* sa = s_acc[q];
#pragma unroll
          for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
              for (int r = 0; r < 4; ++r) acc[rb][cb][r] = sa[((rb * 4 + cb) * 4 + r) * 64 + lane];
          const bool vec = !diag && (n & 3) == 0;
          const bool interior = vec && r0 + kGxTile <= n && c0 + kGxTile <= n;
          // Classification: acc becomes the stored value (minus the lower
          // bound), near entries go to the wave's list.
          uint64_t nrm = 0;
          int cnt = 0;
#pragma unroll
          for (int rb = 0; rb < 4; ++rb) {
            const int li0 = wr + 16 * rb + 4 * g, i0 = r0 + li0;
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
              const int lj = wc + 16 * cb + l16, j = c0 + lj;
              const float uc = su[1][lj];
              bool nr[4];
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int i = i0 + r;
                const bool valid = interior || ((i < n) & (j < n) & (!diag | (i < j)));
                const float lbs = -2.f * acc[rb][cb][r];
                nr[r] = valid & (lbs <= fmaxf(su[0][li0 + r], uc));
                // sqrt(+0) = +0 for lbs <= 0: the bound -0.
                acc[rb][cb][r] = __builtin_amdgcn_sqrtf(fmaxf(lbs, 0.f)) * -kGxDown;
                nrm |= (uint64_t)nr[r] << ((rb * 4 + cb) * 4 + r);
              }
              if (__ballot(nr[0] | nr[1] | nr[2] | nr[3])) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  const uint64_t bal = __ballot(nr[r]);
                  const int slot = cnt + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                  if (nr[r]) s_pos[q][slot] = (uint16_t)(((li0 + r) << 7) | lj);
                  cnt += __popcll(bal);
                }
              }
            }
          }
          // The wave's near pairs get their place in the window's list
          // before this tile's stores are issued.
          int first = 0;
          if (lane == 0 && cnt) first = atomicAdd(ncnt + b, cnt);
          first = __shfl(first, 0);
          if (interior) {
            // Every entry, near ones with their lower bound (replaced later).
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
              const int i0 = r0 + wr + 16 * rb + 4 * g;
#pragma unroll
              for (int cb = 0; cb < 4; ++cb) {
                const int j = c0 + wc + 16 * cb + l16;
                const gram_f32x4 v = acc[rb][cb];
#pragma unroll
                for (int r = 0; r < 4; ++r) gx_store(mb + (size_t)(i0 + r) * n + j, v[r]);
                gx_store4(mb + (size_t)j * n + i0, v[0], v[1], v[2], v[3]);
              }
            }
          } else {
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
              const int i0 = r0 + wr + 16 * rb + 4 * g;
#pragma unroll
              for (int cb = 0; cb < 4; ++cb) {
                const int j = c0 + wc + 16 * cb + l16;
                const gram_f32x4 v = acc[rb][cb];
                const uint32_t m4 = (uint32_t)(nrm >> ((rb * 4 + cb) * 4)) & 15u;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  const int i = i0 + r;
                  if (i < n && j < n && (!diag || i <= j) && !((m4 >> r) & 1u))
                    gx_store(mb + (size_t)i * n + j, i == j ? 0.f : v[r]);
                }
                // Mirror: row j, columns i0 .. i0 + 3.
                if (j < n) {
                  if (vec && i0 + 3 < n && m4 == 0) {
                    gx_store4(mb + (size_t)j * n + i0, v[0], v[1], v[2], v[3]);
                  } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                      const int i = i0 + r;
                      if (i < n && (!diag || i < j) && !((m4 >> r) & 1u)) gx_store(mb + (size_t)j * n + i, v[r]);
                    }
                  }
                }
              }
            }
          }
          // Near pairs to the list for gx_near_kernel; pairs past its
          // capacity exactly here (after this wave's stores of the tile).
          const int room = max(0, min(cnt, cap - first));
          for (int e = lane; e < cnt; e += 64) {
            const int pos = s_pos[q][e];
            const int li = pos >> 7, lj = pos & 127;
            const int i = r0 + li, j = c0 + lj;
            if (e < room) {
              nlist[(size_t)b * cap + first + e] = ((uint32_t)i << 12) | (uint32_t)j;
              continue;
            }
            const float sq = exact_sq(xb + (size_t)i * d, xb + (size_t)j * d, d);
            bad |= !isfinite(sq);
            const float dv = sqrtf(sq);
            gx_store(mb + (size_t)i * n + j, dv);
            gx_store(mb + (size_t)j * n + i, dv);
            if (sq <= su[0][li]) gx_append(lcnt, lval, base + i, dv);
            if (sq <= su[1][lj]) gx_append(lcnt, lval, base + j, dv);
          }
        }
      }
    }
    __syncthreads();
    if (mw) {
      if (prod) {
        float* sa = s_acc[q];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) sa[((rb * 4 + cb) * 4 + r) * 64 + lane] = acc[rb][cb][r];
      }
      load_terms(k + 1);
    }
    __syncthreads();
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 1);
}

// Exact coordinate-order squared distance of rows xi and xj for d <= DK:
// the loads of 64 coordinates at a time are all issued before the first add
// (the row pairs of one wave are scattered, so each phase costs one memory
// round trip); the adds run in coordinate order, t = t + dl * dl.
template <int DK>
__device__ __forceinline__ float gx_exact_sq(const float* xi, const float* xj, int d) {
  constexpr int PH = DK < 64 ? DK : 64;
  const bool vec = ((d & 3) == 0) && ((((uintptr_t)xi) | ((uintptr_t)xj)) & 15) == 0;
  float t = 0.f;
#pragma unroll
  for (int k0 = 0; k0 < DK; k0 += PH) {
    if (k0 >= d) break;  // d is uniform
    float a[PH], c[PH];
    if (vec && k0 + PH <= d) {
#pragma unroll
      for (int q = 0; q < PH / 4; ++q) {
        const float4 va = reinterpret_cast<const float4*>(xi + k0)[q];
        const float4 vc = reinterpret_cast<const float4*>(xj + k0)[q];
        a[4 * q] = va.x; a[4 * q + 1] = va.y; a[4 * q + 2] = va.z; a[4 * q + 3] = va.w;
        c[4 * q] = vc.x; c[4 * q + 1] = vc.y; c[4 * q + 2] = vc.z; c[4 * q + 3] = vc.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < PH; ++k) {
        a[k] = k0 + k < d ? xi[k0 + k] : 0.f;
        c[k] = k0 + k < d ? xj[k0 + k] : 0.f;
      }
    }
    // Padding adds (0 - 0)^2 = +0, which leaves the sum unchanged.
#pragma unroll
    for (int k = 0; k < PH; ++k) {
      const float dl = a[k] - c[k];
      t = t + dl * dl;
    }
  }
  return t;
}

template <int DK>
__global__ __launch_bounds__(256) void gx_near_kernel(const float* x, int64_t batch, int n, int d, int cap,
                                                      const int32_t* ncnt, const uint32_t* nlist, const float* U2,
                                                      float* M, int32_t* lcnt, float* lval, int32_t* status) {
  int b, chunk;
  if (!xcd_item(blockIdx.x, batch, (cap + 255) / 256, b, chunk)) return;
  const int cnt = min(ncnt[b], cap);
  if (chunk * 256 >= cnt) return;  // uniform per workgroup
  const int e = chunk * 256 + (int)threadIdx.x;
  int bad = 0;
  if (e < cnt) {
    const uint32_t pe = nlist[(size_t)b * cap + e];
    const int i = (int)(pe >> 12), j = (int)(pe & 4095u);
    const int64_t base = (int64_t)b * n;
    const float sq = gx_exact_sq<DK>(x + (size_t)(base + i) * d, x + (size_t)(base + j) * d, d);
    bad = !isfinite(sq);
    const float dv = sqrtf(sq);
    float* mb = M + (size_t)base * n;
    gx_store(mb + (size_t)i * n + j, dv);
    gx_store(mb + (size_t)j * n + i, dv);
    if (sq <= U2[base + i]) gx_append(lcnt, lval, base + i, dv);
    if (sq <= U2[base + j]) gx_append(lcnt, lval, base + j, dv);
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 1);
}

// Core distance of every point: the ms-th smallest distance of its row, self
// (+0) included -- the (ms - 1)-th smallest of its core list, which holds
// every other point at squared distance <= U2 (a superset of the ms - 1
// nearest).  A list that overflowed: the row's exact entries (sign bit
// clear) are scanned instead; they include every near pair.  One wave per
// point; ms in 3 .. K.
template <int K>
__global__ __launch_bounds__(256) void gx_core_kernel(const float* M, int64_t points, int n, int ms,
                                                      const int32_t* lcnt, const float* lval, float* core) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t p = (int64_t)blockIdx.x * 4 + wave;
  if (p >= points) return;  // uniform per wave
  const int cnt = lcnt[p];
  float res;
  if (cnt <= kGxCoreCap) {
    float lst[1] = {lane < cnt ? lval[(size_t)p * kGxCoreCap + lane] : INFINITY};
    res = wave_kth<1>(lst, ms - 1);
  } else {
    const float* row = M + (size_t)p * n;
    float lst[K];
#pragma unroll
    for (int q = 0; q < K; ++q) lst[q] = INFINITY;
    for (int j = lane; j < n; j += 64) {
      const float v = row[j];
      if (!(__float_as_uint(v) >> 31)) list_offer<K>(lst, v);
    }
    res = wave_kth<K>(lst, ms);
  }
  if (lane == 0) core[p] = res;
}

struct GxPrimArgs {
  const float* m;      // [B][n][n]: exact distance, or -(lower bound) with the sign bit set
  const float* x;      // points [B][n][d]
  const float* core;   // per point, null when unit_core
  const float* norms;  // squared norms per point
  int n, d, unit_core;
  float c2;
  int32_t* edge_u;
  int32_t* edge_v;
  float* edge_w;
  int32_t* wflag;
  int32_t* status;
};

// Key (weight bits, vertex) with the lower-bound flag (0 or 1) in the lowest
// bit of the vertex word: keys order by weight, then vertex.
__device__ __forceinline__ Cand<float> gx_key(uint32_t wb, int j, uint32_t flag) {
  Cand<float> c;
  c.k = ((uint64_t)wb << 32) | ((uint32_t)j << 1) | flag;
  return c;
}

// Workgroup-wide unsigned 64-bit minimum (all threads, uniform result).
__device__ __forceinline__ uint64_t gx_block_umin64(uint64_t v, uint64_t* sh, int W) {
  v = wave_umin64(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  uint64_t r = ~0ull;
  for (int w = 0; w < W; ++w) r = min(r, sh[w]);
  return r;
}

template <int VPT>
__global__ __launch_bounds__(1024) void prim_gx_kernel(GxPrimArgs a) {
  using C = Cand<float>;
  const int b = blockIdx.x;
  const int n = a.n;
  const int64_t base = (int64_t)b * n;
  const int tid = threadIdx.x, T = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, W = T >> 6;
  const bool unit = a.unit_core != 0;
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ C red[2][3][16];
  __shared__ uint64_t rsh[2][16];
  __shared__ int sh_err, rc_n;
  __shared__ int rc_c[64];
  __shared__ float rc_v[64];
  float* cs = reinterpret_cast<float*>(lds);
  const float* mg = a.m + (size_t)base * n;
  if (tid == 0) sh_err = 0;
  if (!unit)
    for (int i = tid; i < n; i += T) cs[i] = a.core[base + i];
  // Keys, cores and row entries are handled as bit patterns: every weight is
  // a non-negative float (or +inf), whose order is that of its bits, so max
  // and < are single unsigned instructions.  A slot's flag is 0 or 1.
  uint32_t bb[VPT], crb[VPT], fg[VPT];
  float cr[VPT];
  int src[VPT], vst[VPT], jc[VPT];
  uint32_t vis = 0;
#pragma unroll
  for (int s = 0; s < VPT; ++s) {
    const int j = tid + s * T;
    bb[s] = 0x7F800000u;  // +inf
    fg[s] = 0;
    src[s] = 1;
    vst[s] = 0;
    jc[s] = min(j, n - 1);
    cr[s] = j < n && !unit ? a.core[base + j] : 0.f;
    crb[s] = __float_as_uint(cr[s]);
    if (j >= n) vis |= 1u << s;
  }
  if (tid == 0) vis |= 1u;  // vertex 0 starts the tree (visit step 0)
  __syncthreads();

  const float rmax = RealTraits<float>::max_value();
  const C none = C::make(rmax, 0);
  int cur = 0, pr = -1, par = 0, bad = 0;
  bool rowless = false;
  // Row of the vertex resolved last, kept for the round after its
  // resolution, which speculates on that vertex without reading a row.
  uint32_t rrow[VPT];
  bool keep = false;

  // Exact key and source of unvisited vertex r over the visited vertices.
  // Far pairs whose lower weight is <= T are listed in LDS and evaluated one
  // pair per wave (all coordinates loaded at once; entries past the list's
  // 64 slots by their owner lane).
  auto resolve = [&](int r) {
    const float cor = unit ? 0.f : cs[r];
    const float nrm = a.norms[base + r];
    const float* mr = mg + (size_t)r * n;
    if (tid == 0) rc_n = 0;
    float tmin = INFINITY, wx[VPT], kl[VPT], cm[VPT];
    uint32_t farm = 0, okm = 0;
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      const int c = tid + s * T;
      wx[s] = INFINITY;
      kl[s] = INFINITY;
      cm[s] = unit ? 0.f : fmaxf(cr[s], cor);
      rrow[s] = __float_as_uint(mr[jc[s]]);
      if (((vis >> s) & 1u) && c < n) {
        okm |= 1u << s;
        const float v = __uint_as_float(rrow[s]);
        if (__float_as_int(v) < 0) {
          const float lbd = -v;
          kl[s] = fmaxf(cm[s], lbd);
          const float E = a.c2 * (a.norms[base + c] + nrm) + kGramEabs;
          const float ubs = (lbd * lbd * kGxSq + 4.f * E) * kGxUp;
          tmin = fminf(tmin, fmaxf(cm[s], __builtin_amdgcn_sqrtf(ubs) * kGxUp));
          farm |= 1u << s;
        } else {
          wx[s] = fmaxf(cm[s], v);
          tmin = fminf(tmin, wx[s]);
        }
      }
    }
    const uint64_t tk = gx_block_umin64((uint64_t)__float_as_uint(tmin) << 32, rsh[0], W);
    const float tb = __uint_as_float((uint32_t)(tk >> 32));
    int slot[VPT];
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      slot[s] = -1;
      if (((farm >> s) & 1u) && kl[s] <= tb) {
        slot[s] = atomicAdd(&rc_n, 1);
        if (slot[s] < 64) rc_c[slot[s]] = tid + s * T;
      }
    }
    __syncthreads();
    const int nc = min(rc_n, 64);
    for (int e = wave; e < nc; e += W) {
      const float sq = wave_exact_sq1(a.x + (size_t)(base + rc_c[e]) * a.d, a.x + (size_t)(base + r) * a.d, a.d);
      if (lane == 0) rc_v[e] = sq;
    }
    __syncthreads();
    uint64_t key = ~0ull;
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      if (!((okm >> s) & 1u)) continue;
      const int c = tid + s * T;
      float w = wx[s];
      if ((farm >> s) & 1u) {
        if (slot[s] < 0) continue;
        float sq;
        if (slot[s] < 64) {
          sq = rc_v[slot[s]];
        } else {
          const float* xc = a.x + (size_t)(base + c) * a.d;
          const float* xr = a.x + (size_t)(base + r) * a.d;
          sq = 0.f;
#pragma unroll 1
          for (int k = 0; k < a.d; ++k) {
            const float dl = xc[k] - xr[k];
            sq = sq + dl * dl;
          }
        }
        bad |= !isfinite(sq);
        w = fmaxf(cm[s], sqrtf(sq));
      }
      key = min(key, ((uint64_t)__float_as_uint(w) << 32) | ((uint32_t)vst[s] << 12) | (uint32_t)c);
    }
    key = gx_block_umin64(key, rsh[1], W);
#pragma unroll
    for (int s = 0; s < VPT; ++s)
      if (tid + s * T == r) {
        bb[s] = (uint32_t)(key >> 32);
        src[s] = (int)(key & 0xFFFu);
        fg[s] = 0;
      }
  };

  for (int step = 0; step < n - 1;) {
    cur = __builtin_amdgcn_readfirstlane(cur);  // uniform: row addresses in scalar registers
    pr = __builtin_amdgcn_readfirstlane(pr);
    // After a resolution: no row for c, and the resolved vertex's row (kept
    // in registers) as p's.
    const bool spec = (keep || (!rowless && pr >= 0)) && n - 1 - step >= 2;
    const uint32_t ccur = rowless || unit ? 0u : __float_as_uint(cs[cur]);
    const uint32_t cpr = spec && !unit ? __float_as_uint(cs[pr]) : 0u;
    // Both rows are read whole (clamped index past the window), all loads
    // issued together; entries are used only where they can lower a key.
    uint32_t d1[VPT], d2[VPT];
    const float* row1 = mg + (size_t)cur * n;
    const float* row2 = mg + (size_t)(spec ? pr : cur) * n;
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      d1[s] = rowless ? 0u : __float_as_uint(row1[jc[s]]);
      d2[s] = spec ? (keep ? rrow[s] : __float_as_uint(row2[jc[s]])) : 0u;
    }
    uint32_t b1[VPT], b2[VPT], f1[VPT], f2[VPT];
    int s1[VPT], s2[VPT];
    C k1a = none, k1b = none, k2a = none, k2b = none;
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      const int j = tid + s * T;
      const bool un = !((vis >> s) & 1u);
      // Relaxation by c's row: w = max(core_c, core_j, |entry|); a weight
      // strictly below the key replaces it, the entry's sign bit its flag.
      const uint32_t w1 = max(max(ccur, crb[s]), d1[s] & 0x7FFFFFFFu);
      const bool lt1 = !rowless && un && w1 < bb[s];
      b1[s] = lt1 ? w1 : bb[s];
      s1[s] = lt1 ? cur : src[s];
      f1[s] = lt1 ? d1[s] >> 31 : fg[s];
      const C o1 = un ? gx_key(b1[s], j, f1[s]) : C::none();
      if (spec) {
        k1a = cand_pick(o1.less(k1a), o1, k1a);
        // ... and then by p's row, for p's runner-up step.
        const bool e2 = un && j != pr;
        const uint32_t w2 = max(max(cpr, crb[s]), d2[s] & 0x7FFFFFFFu);
        const bool lt2 = e2 && w2 < b1[s];
        b2[s] = lt2 ? w2 : b1[s];
        s2[s] = lt2 ? pr : s1[s];
        f2[s] = lt2 ? d2[s] >> 31 : f1[s];
        cand_top2(k2a, k2b, e2 ? gx_key(b2[s], j, f2[s]) : C::none());
      } else {
        b2[s] = b1[s];
        s2[s] = s1[s];
        f2[s] = f1[s];
        cand_top2(k1a, k1b, o1);
      }
    }
    const C pa = cand_pick(spec, k2a, k1a), pb = cand_pick(spec, k2b, k1b);
    const C w1 = wave_min(pa);
    const C wc = spec ? wave_min(k1a) : w1;
    const C w2 = wave_min(cand_pick(cand_eq(pa, w1), pb, pa));
    if (lane == 0) {
      red[par][0][wave] = wc;
      red[par][1][wave] = w1;
      red[par][2][wave] = w2;
    }
    __syncthreads();
    C qc = none, q1 = none, q2 = none;
    if (lane < W) {
      qc = red[par][0][lane];
      q1 = red[par][1][lane];
      q2 = red[par][2][lane];
    }
    par ^= 1;
    const C g1 = row0_min(q1);
    const C gc = spec ? row0_min(qc) : g1;
    const C g2 = row0_min(cand_pick(cand_eq(q1, g1), q2, q1));
    const C first = cand_pick(spec, gc, g1);
    const int v1 = first.index() >> 1;
    const bool fl1 = first.index() & 1;
    // A winner whose key is a lower bound is not added: this step's keys
    // are kept (c's row; p's row is dropped) and it is resolved below.
    const bool hit = !fl1 && spec && v1 == pr;
    const bool fl2 = hit && (g1.index() & 1);
    const int v2 = hit && !fl2 ? g1.index() >> 1 : -1;
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      const int j = tid + s * T;
      if (j == v1 && !fl1) {
        if ((vis >> s) & 1u) bad |= 1;
        vis |= 1u << s;
        vst[s] = step + 1;
        const float gw = first.weight();
        const size_t o = (size_t)base + step;
        a.edge_u[o] = gw < rmax ? s1[s] : 0;
        a.edge_v[o] = v1;
        a.edge_w[o] = gw;
      }
      if (j == v2) {
        if ((vis >> s) & 1u) bad |= 1;
        vis |= 1u << s;
        vst[s] = step + 2;
        const float gw = g1.weight();
        const size_t o = (size_t)base + step + 1;
        a.edge_u[o] = gw < rmax ? s2[s] : 0;
        a.edge_v[o] = v2;
        a.edge_w[o] = gw;
      }
      bb[s] = hit ? b2[s] : b1[s];
      src[s] = hit ? s2[s] : s1[s];
      fg[s] = hit ? f2[s] : f1[s];
    }
    if (fl1 || fl2) {
      // The winner (or, after p was added, the runner-up) needs its exact
      // key before it can be added; the next round reads no row and
      // speculates on the resolved vertex with the row just read.
      step += fl2 ? 1 : 0;
      const int r = fl1 ? v1 : g1.index() >> 1;
      resolve(r);
      rowless = true;
      pr = r;
      keep = true;
      continue;
    }
    step += hit ? 2 : 1;
    cur = hit ? v2 : v1;
    rowless = false;
    keep = false;
    pr = (!spec || hit) && g2.less(none) ? g2.index() >> 1 : -1;
  }
  if (bad) atomicOr(&sh_err, 1);
  __syncthreads();
  if (tid == 0) {
    a.wflag[b] = sh_err;
    if (sh_err) atomicOr(a.status, sh_err);
  }
}

}  // namespace hdb
