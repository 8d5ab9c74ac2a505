// Distance, validation, core-distance and tree kernels.
#pragma once

#include "hdb_common.h"
#include "hdb_tree.h"

namespace hdb {

// ORs HDB_STATUS_NONFINITE (2) into *status if any input value is not finite.
template <typename Real>
__global__ __launch_bounds__(256) void check_finite_kernel(const Real* x, size_t count, int32_t* status) {
  int bad = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    bad |= !isfinite(x[i]);
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 2);
}

// Dense distances out[b][i][j], 64x64 tiles (4x4 per thread).  Only tiles
// on or above the diagonal are computed; each is written as is and, through
// an LDS transpose, to its mirror position.  Coordinates are staged through
// LDS in chunks of 16 dimensions; each output sums (x_ik - x_jk)^2 for k
// ascending, then takes the correctly rounded square root.  x_i - x_j is the
// exact negation of x_j - x_i, so the mirror holds the same bits the
// reference's triangle-and-mirror kernel writes (reference
// hdbscan_gpu/_kernels.py:45-73).  Grid: (nt (nt + 1) / 2, items) for nt
// tiles per side of the launch's largest window; block t is the t-th tile of
// the upper triangle in row order (blocks past a smaller window's triangle
// exit).  Window w's matrix starts at element win.matrix(w, base).
// Input check: the diagonal tiles stage every coordinate of the window
// exactly once, so they OR HDB_STATUS_NONFINITE (2) into *status for a
// non-finite coordinate (the reference's input check, hdbscan_gpu/_cuda.py:
// 96-99, which outranks the overflow bit 1 set for a non-finite distance).
template <typename Real>
__global__ __launch_bounds__(256) void distance_kernel(const Real* x, Windows win, int d, Real* out,
                                                       int32_t* status) {
  constexpr int TILE = 64, KC = 16;
  __shared__ Real si[KC][TILE];
  __shared__ Real sj[KC][TILE];
  __shared__ Real st[TILE][TILE + 1];
  int64_t base;
  int n, w;
  win.get(blockIdx.y, w, base, n);
  const int nt = (n + TILE - 1) / TILE;
  if ((int)blockIdx.x >= nt * (nt + 1) / 2) return;
  int t = blockIdx.x, rt = 0;
  while (t >= nt - rt) {
    t -= nt - rt;
    ++rt;
  }
  const int ti = rt * TILE, tj = (rt + t) * TILE;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const Real* xb = x + (size_t)base * d;
  const bool diag = ti == tj;
  int nonfinite = 0;
  Real tot[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) tot[r][q] = 0;
  for (int k0 = 0; k0 < d; k0 += KC) {
    __syncthreads();
    for (int idx = tid; idx < TILE * KC; idx += 256) {
      const int row = idx / KC, kk = idx - row * KC, k = k0 + kk;
      const int gi = ti + row, gj = tj + row;
      const Real vi = (gi < n && k < d) ? xb[(size_t)gi * d + k] : Real(0);
      nonfinite |= diag && !isfinite(vi);
      si[kk][row] = vi;
      sj[kk][row] = (gj < n && k < d) ? xb[(size_t)gj * d + k] : Real(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) {
      Real av[4], cv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) av[r] = si[kk][ty * 4 + r];
#pragma unroll
      for (int q = 0; q < 4; ++q) cv[q] = sj[kk][tx * 4 + q];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const Real dl = av[r] - cv[q];
          tot[r][q] = tot[r][q] + dl * dl;
        }
    }
  }
  int bad = 0;
  Real* ob = out + win.matrix(w, base);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gi = ti + ty * 4 + r;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gj = tj + tx * 4 + q;
      const Real v = sqrt(tot[r][q]);
      st[ty * 4 + r][tx * 4 + q] = v;
      if (gi < n && gj < n) {
        bad |= !isfinite(v);
        ob[(size_t)gi * n + gj] = v;
      }
    }
  }
  if (tj != ti) {
    __syncthreads();
    // Mirror tile: row tj + c, column ti + r holds st[r][c]; consecutive
    // threads write consecutive columns.
    for (int idx = tid; idx < TILE * TILE; idx += 256) {
      const int c = idx / TILE, r = idx - c * TILE;
      const int gi = tj + c, gj = ti + r;
      if (gi < n && gj < n) ob[(size_t)gi * n + gj] = st[r][c];
    }
  }
  if (__any(bad) && (tid & 63) == 0) atomicOr(status, 1);
  if (__any(nonfinite) && (tid & 63) == 0) atomicOr(status, 2);
}

// Launch lists of a ragged batch.  Window w of size n_w belongs to bucket k
// (sizes up to top[k], above top[k-1]) and goes to lists[start[k] + rank],
// rank = its position among the bucket's windows by size, largest first
// (equal sizes by index): the claim order of the bucket's launches, longest
// first.  Windows of buckets flagged in matrix_mask get the element offset of
// their n x n distance matrix (those windows in index order).  The launches'
// claim counters are zeroed here (no memset node in a captured graph).
struct RaggedLists {
  int32_t start[8];
  int32_t top[8];
  uint32_t matrix_mask;
  int32_t queues;
};

// One block per 64 windows (one per lane); the batch's window sizes pass
// through LDS in tiles of 256 (size << 3 | bucket: sizes are at most 4096)
// and the four waves split each tile, their partial ranks and offsets summed
// at the end.  Grid: max(ceil(windows / 64), 1) blocks of 256 threads.
__global__ __launch_bounds__(256) void ragged_lists_kernel(const int64_t* off, int windows, RaggedLists rl,
                                                           int32_t* lists, int64_t* moff, int32_t* queue) {
  __shared__ uint32_t tile[256];
  __shared__ int part_rank[4][64];
  __shared__ int64_t part_mo[4][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (blockIdx.x == 0 && tid < rl.queues) queue[tid] = 0;
  auto bucket = [&](int64_t nv) {
    int k = 0;
    while (k < 7 && nv > rl.top[k]) ++k;
    return k;
  };
  const int g = blockIdx.x * 64 + lane;
  const int n = g < windows ? (int)(off[g + 1] - off[g]) : 0;
  const int k = g < windows ? bucket(n) : 8;
  int rank = 0;
  int64_t mo = 0;
  for (int c0 = 0; c0 < windows; c0 += 256) {
    const int v = c0 + tid;
    if (v < windows) {
      const int nv = (int)(off[v + 1] - off[v]);
      tile[tid] = ((uint32_t)nv << 3) | (uint32_t)bucket(nv);
    }
    __syncthreads();
    const int cnt = min(256, windows - c0);
#pragma unroll 4
    for (int i = wave; i < cnt; i += 4) {
      const uint32_t e = tile[i];
      const int nv = (int)(e >> 3), kv = (int)(e & 7u), vv = c0 + i;
      rank += kv == k && (nv > n || (nv == n && vv < g));
      if (vv < g && ((rl.matrix_mask >> kv) & 1u)) mo += (int64_t)nv * nv;
    }
    __syncthreads();
  }
  part_rank[wave][lane] = rank;
  part_mo[wave][lane] = mo;
  __syncthreads();
  if (wave == 0 && g < windows) {
    rank = part_rank[0][lane] + part_rank[1][lane] + part_rank[2][lane] + part_rank[3][lane];
    mo = part_mo[0][lane] + part_mo[1][lane] + part_mo[2][lane] + part_mo[3][lane];
    lists[rl.start[k] + rank] = g;
    moff[g] = mo;
  }
}

// Issue-rate probe for the exact distance's instruction mix: per coordinate
// one subtraction, one multiplication and one addition, no FMA (built with
// -ffp-contract=off) and no packed fp32.  Each thread runs 8 independent
// chains t = t + (a - t) * (a - t) for `iters` rounds: 24 operations per
// round; the result is stored so nothing is removed.  Values may become
// infinite or NaN: the issue rate does not depend on them.
__global__ __launch_bounds__(256) void valu_probe_kernel(int iters, float seed, float* out) {
  float a[8], t[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[i] = seed + (float)threadIdx.x * 1e-3f + (float)i;
    t[i] = (float)i * 1e-2f;
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const float dl = a[i] - t[i];
      t[i] = t[i] + dl * dl;
    }
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) s = s + t[i];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Upper bound of b*b (see square_bound in hdb_prim.h).
__device__ __forceinline__ float sq_bound(float b) {
  const float p = b * b;
  return isfinite(p) ? __uint_as_float(__float_as_uint(p) + 1u) : p;
}
__device__ __forceinline__ double sq_bound(double b) {
  const double p = b * b;
  return isfinite(p) ? __longlong_as_double(__double_as_longlong(p) + 1ll) : p;
}

// Offers the squared distance tot to the ascending register list lst of the
// KMAX smallest distances.  tot >= lim (an upper bound of lst[KMAX-1]^2)
// means sqrt(tot) >= lst[KMAX-1]: the list cannot change and the correctly
// rounded square root is skipped.
template <typename Real, int KMAX>
__device__ __forceinline__ void topk_offer(Real* lst, Real& lim, Real tot) {
  if (tot < lim) {
    const Real v = sqrt(tot);
    if (v < lst[KMAX - 1]) {
      if constexpr (sizeof(Real) == 4) {
        // Ascending list: the new q-th smallest is the median of
        // (lst[q-1], lst[q], v); v is finite (tot < lim).
#pragma unroll
        for (int q = KMAX - 1; q > 0; --q) lst[q] = __builtin_amdgcn_fmed3f(lst[q - 1], lst[q], v);
        lst[0] = v < lst[0] ? v : lst[0];
      } else {
#pragma unroll
        for (int q = KMAX - 1; q > 0; --q) lst[q] = fmin(lst[q], fmax(lst[q - 1], v));
        lst[0] = fmin(lst[0], v);
      }
      lim = sq_bound(lst[KMAX - 1]);
    }
  }
}

// Core distance of row i: the (ms-1)-th smallest of d(i, j) over all j
// (self included), exactly as an order statistic of the reference's row
// (reference hdbscan_gpu/_cuda.py:114-118).  One thread per row keeps the
// KMAX smallest distances in a register insertion network while streaming
// 64-point tiles of the window through LDS.  DK > 0: dimensions padded to DK
// (zero padding adds +0 exactly); DK == 0: any d, 32-dimension chunks.
template <typename Real, int KMAX, int DK>
__global__ __launch_bounds__(256) void core_topk_kernel(const Real* x, Windows win, int d, int ms, Real* core,
                                                        int32_t* status) {
  constexpr int TJ = DK > 0 ? 64 : 32;
  constexpr int KC = DK > 0 ? DK : 32;
  __shared__ Real tile[TJ * KC];
  int64_t base;
  int n;
  int wid;
  win.get(blockIdx.y, wid, base, n);
  if ((int)(blockIdx.x * blockDim.x) >= n) return;
  const int tid = threadIdx.x;
  const int i = blockIdx.x * blockDim.x + tid;
  const Real* xb = x + (size_t)base * d;
  Real lst[KMAX];
#pragma unroll
  for (int q = 0; q < KMAX; ++q) lst[q] = Real(INFINITY);
  Real lim = Real(INFINITY);
  int bad = 0;
  Real xi[KC];
  if (DK > 0) {
#pragma unroll
    for (int k = 0; k < KC; ++k) xi[k] = (i < n && k < d) ? xb[(size_t)i * d + k] : Real(0);
  }
  for (int j0 = 0; j0 < n; j0 += TJ) {
    const int jn = min(TJ, n - j0);
    if constexpr (DK > 0) {
      __syncthreads();
      for (int idx = tid; idx < TJ * KC; idx += blockDim.x) {
        const int jj = idx / KC, k = idx - jj * KC, j = j0 + jj;
        tile[idx] = (j < n && k < d) ? xb[(size_t)j * d + k] : Real(0);
      }
      __syncthreads();
      for (int jj = 0; jj < jn; ++jj) {
        Real t = 0;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          const Real dl = xi[k] - tile[jj * KC + k];
          t = t + dl * dl;
        }
        bad |= !isfinite(t);
        topk_offer<Real, KMAX>(lst, lim, t);
      }
    } else {
      Real tot[TJ];
#pragma unroll
      for (int jj = 0; jj < TJ; ++jj) tot[jj] = 0;
      for (int k0 = 0; k0 < d; k0 += KC) {
        __syncthreads();
        for (int idx = tid; idx < TJ * KC; idx += blockDim.x) {
          const int jj = idx / KC, kk = idx - jj * KC, j = j0 + jj, k = k0 + kk;
          tile[idx] = (j < n && k < d) ? xb[(size_t)j * d + k] : Real(0);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KC; ++k) xi[k] = (i < n && k0 + k < d) ? xb[(size_t)i * d + k0 + k] : Real(0);
#pragma unroll
        for (int jj = 0; jj < TJ; ++jj) {
          Real t = tot[jj];
#pragma unroll
          for (int k = 0; k < KC; ++k) {
            const Real dl = xi[k] - tile[jj * KC + k];
            t = t + dl * dl;
          }
          tot[jj] = t;
        }
      }
      for (int jj = 0; jj < jn; ++jj) {
        bad |= !isfinite(tot[jj]);
        topk_offer<Real, KMAX>(lst, lim, tot[jj]);
      }
    }
  }
  if (i < n) {
    Real v = lst[0];
#pragma unroll
    for (int q = 1; q < KMAX; ++q)
      if (q == ms - 1) v = lst[q];
    core[(size_t)base + i] = v;
  }
  // Every ordered pair passes through here, so this is the overflow check of
  // the whole distance matrix (Prim may skip pairs that cannot matter).
  bad = bad && i < n;
  if (__any(bad) && (tid & 63) == 0) atomicOr(status, 1);
}

template <typename Real> struct BitsOf;
template <> struct BitsOf<float> {
  using U = uint32_t;
  __device__ static U get(float v) { return __float_as_uint(v); }
  __device__ static float put(U u) { return __uint_as_float(u); }
};
template <> struct BitsOf<double> {
  using U = uint64_t;
  __device__ static U get(double v) { return (U)__double_as_longlong(v); }
  __device__ static double put(U u) { return __longlong_as_double((long long)u); }
};

// Core distances for any min_samples from a materialised matrix: one wave per
// row; the row is staged in LDS and the (ms-1)-th smallest value is built
// bit by bit from the top (non-negative IEEE values order like their bits).
template <typename Real>
__global__ __launch_bounds__(64) void core_select_kernel(const Real* dist, Windows win, int ms, Real* core) {
  using U = typename BitsOf<Real>::U;
  extern __shared__ __align__(16) unsigned char lds[];
  Real* row = reinterpret_cast<Real*>(lds);
  int64_t base;
  int n, w;
  win.get(blockIdx.y, w, base, n);
  if ((int)blockIdx.x >= n) return;
  const size_t r = (size_t)base + blockIdx.x;
  const Real* src = dist + win.matrix(w, base) + (size_t)blockIdx.x * n;
  const int lane = threadIdx.x;
  for (int j = lane; j < n; j += 64) row[j] = src[j];
  __syncthreads();
  const int k = ms - 1;
  U ans = 0;
  for (int bit = (int)sizeof(U) * 8 - 2; bit >= 0; --bit) {
    const U cand = ans | ((U)1 << bit);
    int cnt = 0;
    for (int j = lane; j < n; j += 64) cnt += BitsOf<Real>::get(row[j]) < cand;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) cnt += __shfl_xor(cnt, m);
    if (cnt <= k) ans = cand;
  }
  if (lane == 0) core[r] = BitsOf<Real>::put(ans);
}

// Core distances for min_samples <= KMAX from a materialised matrix: one
// wave per row.  Lane l keeps the KMAX smallest of row[l], row[l + 64], ...
// (coalesced reads straight from memory); the ms-th smallest of the union
// of the 64 sorted lists is then found by popping wave minima (equal heads
// pop together, so the value at which the count first reaches ms is the
// order statistic).  Grid: (ceil(n / 4), items), one wave per row.
template <typename Real, int KMAX>
__global__ __launch_bounds__(256) void core_select_list_kernel(const Real* dist, Windows win, int ms, Real* core) {
  const int lane = threadIdx.x & 63;
  int64_t base;
  int n, w;
  win.get(blockIdx.y, w, base, n);
  const int i = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (i >= n) return;  // uniform per wave
  const int64_t r = base + i;
  const Real* src = dist + win.matrix(w, base) + (size_t)i * n;
  Real lst[KMAX];
#pragma unroll
  for (int q = 0; q < KMAX; ++q) lst[q] = Real(INFINITY);
  if constexpr (sizeof(Real) == 4) {
    // Eight coalesced loads in flight per lane, then a branch-free insertion:
    // with lst ascending, the new q-th smallest of lst and v is the median of
    // (lst[q-1], lst[q], v) (one instruction), the smallest a compare-select.
    // Distances are never NaN here (non-finite input is rejected); columns
    // past the row read as +inf and change nothing.
    for (int j0 = 0; j0 < n; j0 += 8 * 64) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = j0 + u * 64 + lane;
        v[u] = j < n ? src[j] : INFINITY;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
#pragma unroll
        for (int q = KMAX - 1; q > 0; --q) lst[q] = __builtin_amdgcn_fmed3f(lst[q - 1], lst[q], v[u]);
        lst[0] = v[u] < lst[0] ? v[u] : lst[0];
      }
    }
  } else {
    // fp64 (no median-of-three instruction): eight loads in flight, then the
    // guarded insertion.
    for (int j0 = 0; j0 < n; j0 += 8 * 64) {
      Real v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = j0 + u * 64 + lane;
        v[u] = j < n ? src[j] : Real(INFINITY);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (v[u] < lst[KMAX - 1]) {
#pragma unroll
          for (int q = KMAX - 1; q > 0; --q) lst[q] = fmin(lst[q], fmax(lst[q - 1], v[u]));
          lst[0] = fmin(lst[0], v[u]);
        }
      }
    }
  }
  int cnt = 0;
  Real ans = Real(INFINITY);
  for (int it = 0; it < ms && cnt < ms; ++it) {
    Real m = lst[0];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) m = fmin(m, __shfl_xor(m, o));
    const bool eq = lst[0] == m;
    cnt += __popcll(__ballot(eq));
    ans = m;
    if (eq) {
#pragma unroll
      for (int q = 0; q < KMAX - 1; ++q) lst[q] = lst[q + 1];
      lst[KMAX - 1] = Real(INFINITY);
    }
  }
  if (lane == 0) core[r] = ans;
}

struct TreeArgs {
  const int32_t* edge_u;
  const int32_t* edge_v;
  const void* edge_w;
  const int32_t* wflag;
  Windows win;
  long long mcs;
  double epsilon;
  int sort_depth;
  int serial_sort;       // one-lane introsort instead of the workgroup-parallel one
  int serial_link;       // linkage by one wave in sort order instead of the parallel run form
  int32_t* labels;
  unsigned char* arena;  // global arena (when not in LDS)
  size_t arena_stride;   // dense: bytes per window; 0: ragged addressing
  unsigned char* slots;  // cluster records of the condensation walks
  size_t slot_stride;    // dense: bytes per window; 0: per point
  int32_t* status;
  int32_t* export_ints;  // [B][10][n] or null
  void* export_reals;    // [B][4][n] or null
  uint32_t* sorted_uv;   // split tree stage: sorted endpoints per point (u | v << 16)
  void* sorted_w;        // split tree stage: sorted weights per point (Real)
};

// Tree stages of one window (item of the launch) by one workgroup: the
// window's arena, its Prim edges as sort keys, then the NumPy-order sort.
template <typename Real, bool LDSA>
__device__ __forceinline__ unsigned char* tree_arena(const TreeArgs& a, int b, int64_t p0) {
  extern __shared__ __align__(16) unsigned char lds[];
  return LDSA ? lds
              : a.arena + (a.arena_stride ? (size_t)b * a.arena_stride
                                          : (size_t)p0 * kArenaPerPoint + (size_t)b * kArenaPerWindow);
}

template <typename Real>
__device__ __forceinline__ void tree_sort(const TreeArgs& a, const TreeView<Real>& t, int n, int* counters) {
  if (a.serial_sort) {
    if (threadIdx.x == 0) introsort(t.keys, n - 1, a.sort_depth, t.stack);
    __syncthreads();
  } else {
    // The lists live in birth..stack, which is free until condensation and
    // at least sort_lists_bytes(n) long (tree_layout guarantees it).
    SortLists lists;
    lists.cap = sort_task_cap(n);
    lists.cap_small = sort_small_cap(n);
    lists.cur = reinterpret_cast<int16_t*>(t.birth);
    lists.next = lists.cur + 3 * lists.cap;
    lists.heap = lists.next + 3 * lists.cap;
    lists.small = lists.heap + 2 * lists.cap;
    parallel_introsort(t.keys, n - 1, a.sort_depth, t.uf, t.comp, lists, counters);
  }
}

// Linkage, condensation, selection and labels of one window whose sorted
// keys (weights; idx = the edge's index in edge_u / edge_v) are in the
// arena.  eu / ev / ew: the window's Prim edges (for the exports only).
template <typename Real>
__device__ __forceinline__ void tree_finish(const TreeArgs& a, const TreeView<Real>& t, int n, int b, int64_t p0,
                                            const int32_t* eu, const int32_t* ev, const Real* ew, int* sh_counters,
                                            long long* sh_clk, long long clk0) {
  const int tid = threadIdx.x, T = blockDim.x;
  int32_t* lab = a.labels + (size_t)p0;
  long long clk[6];
  clk[0] = clk0;
  {
    clk[1] = wall_clock64();
    if (a.serial_link) {
      // parent / point_parent / birth are free until condensation.
      int16_t* su = t.parent;
      int16_t* sv = t.pparent;
      uint32_t* rec = reinterpret_cast<uint32_t*>(t.birth);
      linkage_gather(t, n, su, sv, rec);
      __syncthreads();
      clk[2] = wall_clock64();
      if (tid < 64) linkage_wave(t, n, su, sv, rec);
    } else {
      clk[2] = wall_clock64();
      linkage_parallel(t, n);
    }
    clk[3] = wall_clock64();
  }
  __syncthreads();
  int32_t* xi = a.export_ints ? a.export_ints + (size_t)b * 10 * n : nullptr;
  Real* xr = a.export_reals ? static_cast<Real*>(a.export_reals) + (size_t)b * 4 * n : nullptr;
  if (xi) {
    for (int i = tid; i < n; i += T) {
      const bool e = i < n - 1;
      xi[0 * n + i] = e ? eu[i] : -1;
      xi[1 * n + i] = e ? ev[i] : -1;
      xi[2 * n + i] = e ? t.keys[i].idx : -1;
      xi[3 * n + i] = e ? t.left[i] : -1;
      xi[4 * n + i] = e ? t.right[i] : -1;
      xi[5 * n + i] = e ? t.sizes[i] : -1;
      xr[0 * n + i] = e ? ew[i] : Real(0);
      xr[1 * n + i] = e ? t.keys[i].w : Real(0);
    }
    __syncthreads();
  }
  // Condensation scratch (arrays dead after linkage): edge_u holds the slot
  // of each visited node, parent the pruned point counts (until the final
  // cluster arrays are written), uf|comp the split keys (n int32) and then
  // the parents of all 2n-1 nodes.
  int16_t* node_slot = t.edge_u;
  condense_prepare(t, n, a.mcs, node_slot, t.parent);
  __syncthreads();
  ClusterSlot<Real>* slots = reinterpret_cast<ClusterSlot<Real>*>(
      a.slots + (a.slot_stride ? (size_t)b * a.slot_stride : (size_t)p0 * sizeof(ClusterSlot<Real>)));
  const int nl = condense_clusters(t, n, a.mcs, slots, node_slot, t.parent, sh_counters);
  if (tid == 0) clk[4] = wall_clock64();
  for (int i = tid; i < n - 1; i += T) {
    t.uf[t.left[i]] = (int16_t)(n + i);
    t.uf[t.right[i]] = (int16_t)(n + i);
  }
  if (tid == 0) t.uf[2 * n - 2] = (int16_t)(2 * n - 2);  // the root is its own parent
  __syncthreads();
  condense_points(t, n, t.uf, node_slot);
  if (xr) {
    for (int i = tid; i < n; i += T) xr[2 * n + i] = i < nl - n ? t.birth[i] : Real(0);
  }
  __syncthreads();
  if (tid < 64) select_clusters_wave(t, n, nl, (Real)a.epsilon);
  if (tid == 0) {
    clk[5] = wall_clock64();
    for (int q = 0; q < 6; ++q) sh_clk[q] = clk[q];
  }
  __syncthreads();
  if (xi) {
    for (int i = tid; i < n; i += T) {
      const bool c = i < nl - n;
      xi[6 * n + i] = c ? t.parent[i] : -1;
      xi[7 * n + i] = c ? t.selected[i] : 0;
      xi[8 * n + i] = t.pparent[i];
      xi[9 * n + i] = i == 0 ? nl : (i < 6 ? (int32_t)(sh_clk[i] - sh_clk[i - 1]) : 0);
      xr[3 * n + i] = c ? t.stab[i] : Real(0);
    }
  }
  const int16_t* resolved = t.resolved();
  for (int i = tid; i < n; i += T) {
    const int pp = t.pparent[i];
    lab[i] = pp >= n ? resolved[pp - n] : -1;
  }
}


template <typename Real, bool LDSA>
__device__ __forceinline__ void tree_window(const TreeArgs& a, int item) {
  __shared__ int sh_counters[4];
  __shared__ long long sh_clk[6];
  const int tid = threadIdx.x, T = blockDim.x;
  int64_t p0;
  int n, b;
  a.win.get(item, b, p0, n);
  int32_t* lab = a.labels + (size_t)p0;
  if (a.wflag[b]) {
    for (int i = tid; i < n; i += T) lab[i] = -1;
    return;
  }
  const TreeView<Real> t = tree_view<Real>(tree_arena<Real, LDSA>(a, b, p0), n);
  const int32_t* eu = a.edge_u + (size_t)p0;
  const int32_t* ev = a.edge_v + (size_t)p0;
  const Real* ew = static_cast<const Real*>(a.edge_w) + (size_t)p0;
  for (int i = tid; i < n - 1; i += T) {
    EdgeKey<Real> k;
    k.w = ew[i];
    k.idx = i;
    t.keys[i] = k;
    t.edge_u[i] = (int16_t)eu[i];
    t.edge_v[i] = (int16_t)ev[i];
  }
  __syncthreads();
  // Phase timestamps (constant 100 MHz clock) for the diagnostic export.
  const long long clk0 = wall_clock64();
  tree_sort<Real>(a, t, n, sh_counters);
  tree_finish<Real>(a, t, n, b, p0, eu, ev, ew, sh_counters, sh_clk, clk0);
}

// Split tree stage, first half (one workgroup per window): the NumPy-order
// sort of the window's Prim edges, written as the sorted-edge transport --
// endpoints as two int16 in one int32 (u | v << 16) and the weight, in sorted
// order, at the window's first point.
template <typename Real, bool LDSA>
__device__ __forceinline__ void sort_window(const TreeArgs& a, int item) {
  __shared__ int sh_counters[4];
  const int tid = threadIdx.x, T = blockDim.x;
  int64_t p0;
  int n, b;
  a.win.get(item, b, p0, n);
  if (a.wflag[b]) return;
  const TreeView<Real> t = tree_view<Real>(tree_arena<Real, LDSA>(a, b, p0), n);
  const int32_t* eu = a.edge_u + (size_t)p0;
  const int32_t* ev = a.edge_v + (size_t)p0;
  const Real* ew = static_cast<const Real*>(a.edge_w) + (size_t)p0;
  for (int i = tid; i < n - 1; i += T) {
    EdgeKey<Real> k;
    k.w = ew[i];
    k.idx = i;
    t.keys[i] = k;
  }
  __syncthreads();
  tree_sort<Real>(a, t, n, sh_counters);
  for (int i = tid; i < n - 1; i += T) {
    const EdgeKey<Real> k = t.keys[i];
    a.sorted_uv[p0 + i] = (uint32_t)(uint16_t)eu[k.idx] | ((uint32_t)(uint16_t)ev[k.idx] << 16);
    static_cast<Real*>(a.sorted_w)[p0 + i] = k.w;
  }
}

// Split tree stage, second half: linkage, condensation, selection and labels
// from the sorted-edge transport.  A workgroup needs only max(64, (n-1)/4)
// threads here, so small windows run several to a compute unit.
template <typename Real, bool LDSA>
__device__ __forceinline__ void hierarchy_window(const TreeArgs& a, int item) {
  __shared__ int sh_counters[4];
  __shared__ long long sh_clk[6];
  const int tid = threadIdx.x, T = blockDim.x;
  int64_t p0;
  int n, b;
  a.win.get(item, b, p0, n);
  int32_t* lab = a.labels + (size_t)p0;
  if (a.wflag[b]) {
    for (int i = tid; i < n; i += T) lab[i] = -1;
    return;
  }
  const TreeView<Real> t = tree_view<Real>(tree_arena<Real, LDSA>(a, b, p0), n);
  // The linkage works on the Prim order: the step of each sorted edge is the
  // position of its added vertex in the window's Prim output, minus one.
  const Real* sw = static_cast<const Real*>(a.sorted_w) + (size_t)p0;
  int16_t* step_of = t.uf;
  for (int s = tid; s < n - 1; s += T) step_of[a.edge_v[p0 + s]] = (int16_t)s;
  __syncthreads();
  for (int i = tid; i < n - 1; i += T) {
    const uint32_t uv = a.sorted_uv[p0 + i];
    const int v = (int)(uv >> 16);
    const int s = step_of[v];
    EdgeKey<Real> k;
    k.w = sw[i];
    k.idx = s;
    t.keys[i] = k;
    t.edge_u[s] = (int16_t)(uv & 0xFFFFu);
    t.edge_v[s] = (int16_t)v;
  }
  __syncthreads();
  tree_finish<Real>(a, t, n, b, p0, nullptr, nullptr, nullptr, sh_counters, sh_clk, wall_clock64());
}

// Tree stages, one window per workgroup or, with a queue, windows claimed
// one after another.  Lane 0 runs the sequential stages; the workgroup
// stages inputs and writes labels and exports.
template <typename Real, bool LDSA>
__global__ __launch_bounds__(1024) void tree_kernel(TreeArgs a) {
  for_each_item(a.win, [&](int item) { tree_window<Real, LDSA>(a, item); });
}

template <typename Real, bool LDSA>
__global__ __launch_bounds__(1024) void sort_edges_kernel(TreeArgs a) {
  for_each_item(a.win, [&](int item) { sort_window<Real, LDSA>(a, item); });
}

template <typename Real, bool LDSA>
__global__ __launch_bounds__(1024) void hierarchy_kernel(TreeArgs a) {
  for_each_item(a.win, [&](int item) { hierarchy_window<Real, LDSA>(a, item); });
}

}  // namespace hdb
