// Tree stages of one window: NumPy-order introsort of the Prim edges,
// union-find linkage, breadth-first condensation, excess-of-mass selection,
// epsilon merge and label resolution.
//
// One workgroup per window runs them out of the window's arena (LDS when it
// fits), each stage in the parallel form that gives the reference's exact
// result: partitions by one wave each, linkage broadcast through one wave,
// condensation by pointer jumping with sequential per-cluster sums,
// selection level by level on one wave.  Semantics follow the reference
// stage by stage:
//   sort      reference hdbscan_gpu/_kernels.py:358-516
//   linkage   reference hdbscan_gpu/_kernels.py:518-556
//   condense  reference hdbscan_gpu/_kernels.py:558-643
//   eom       reference hdbscan_gpu/_kernels.py:671-709
//   labels    reference hdbscan_gpu/_kernels.py:712-760
// heap_sort / introsort follow NumPy's scalar aheapsort / aquicksort step for
// step, and the tree semantics are scikit-learn's _tree.pyx (BSD-3; notices
// in THIRD_PARTY_NOTICES.md at the repository root).
#pragma once

#include "hdb_common.h"

namespace hdb {

// NumPy's scalar quicksort hands spans with hi - lo <= 15 to insertion sort.
constexpr int kSmallSort = 15;

template <typename Real>
__device__ __forceinline__ void key_swap(EdgeKey<Real>* k, int i, int j) {
  EdgeKey<Real> t = k[i];
  k[i] = k[j];
  k[j] = t;
}

// Heapsort of keys[lo..hi] (one-based heap, NumPy's aheapsort steps).
template <typename Real>
__device__ void heap_sort(EdgeKey<Real>* keys, int lo, int hi) {
  EdgeKey<Real>* a = keys + lo - 1;
  int n = hi - lo + 1;
  for (int l = n >> 1; l > 0; --l) {
    EdgeKey<Real> tmp = a[l];
    int i = l, j = l << 1;
    while (j <= n) {
      if (j < n && a[j].w < a[j + 1].w) j += 1;
      if (tmp.w < a[j].w) {
        a[i] = a[j];
        i = j;
        j += j;
      } else {
        break;
      }
    }
    a[i] = tmp;
  }
  while (n > 1) {
    EdgeKey<Real> tmp = a[n];
    a[n] = a[1];
    n -= 1;
    int i = 1, j = 2;
    while (j <= n) {
      if (j < n && a[j].w < a[j + 1].w) j++;
      if (tmp.w < a[j].w) {
        a[i] = a[j];
        i = j;
        j += j;
      } else {
        break;
      }
    }
    a[i] = tmp;
  }
}

// Introsort of keys[0..m-1] by weight in NumPy's scalar aquicksort order:
// depth 2*floor(log2 m), median of three with the pivot parked at hi-1,
// larger side deferred with the decremented depth, smaller side continued in
// place without a depth check, insertion sort for spans hi-lo <= 15.
template <typename Real>
__device__ void introsort(EdgeKey<Real>* keys, int m, int depth_override, int* stack) {
  if (m < 2) return;
  int* st_lo = stack;
  int* st_hi = stack + kSortStack;
  int* st_depth = stack + 2 * kSortStack;
  int top = 0, lo = 0, hi = m - 1, depth = 0;
  for (int c = m; c > 1; c >>= 1) depth += 2;
  if (depth_override > -1000) depth = depth_override;
  for (;;) {
    if (depth < 0) {
      heap_sort(keys, lo, hi);
    } else {
      while (hi - lo > kSmallSort) {
        if (top == kSortStack) {
          // Unreachable: the deferred side is the larger one, so at most
          // log2(m) spans are pending.  Kept so the loop always terminates.
          heap_sort(keys, lo, hi);
          break;
        }
        int mid = lo + ((hi - lo) >> 1);
        if (keys[mid].w < keys[lo].w) key_swap(keys, mid, lo);
        if (keys[hi].w < keys[mid].w) key_swap(keys, hi, mid);
        if (keys[mid].w < keys[lo].w) key_swap(keys, mid, lo);
        const Real pivot = keys[mid].w;
        int i = lo, j = hi - 1;
        key_swap(keys, mid, j);
        for (;;) {
          do { ++i; } while (keys[i].w < pivot);
          do { --j; } while (pivot < keys[j].w);
          if (i >= j) break;
          key_swap(keys, i, j);
        }
        key_swap(keys, i, hi - 1);
        --depth;
        if (i - lo < hi - i) {
          st_lo[top] = i + 1; st_hi[top] = hi; hi = i - 1;
        } else {
          st_lo[top] = lo; st_hi[top] = i - 1; lo = i + 1;
        }
        st_depth[top++] = depth;
      }
      for (int i = lo + 1; i <= hi; ++i) {
        EdgeKey<Real> v = keys[i];
        int j = i;
        while (j > lo && v.w < keys[j - 1].w) {
          keys[j] = keys[j - 1];
          --j;
        }
        keys[j] = v;
      }
    }
    if (top == 0) break;
    --top;
    lo = st_lo[top]; hi = st_hi[top]; depth = st_depth[top];
  }
}

// ---------------------------------------------------------------------------
// Workgroup-parallel introsort with NumPy's exact result.
//
// NumPy's partition of keys[lo..hi] (after median of three and parking the
// pivot at hi-1) scans i upward stopping at weights >= pivot and j downward
// stopping at weights <= pivot, swapping until i >= j.  Because the scans
// only ever cross unmodified positions, the k-th swap exchanges L_k (k-th
// position >= pivot counted from lo+1 upward) with R_k (k-th position
// <= pivot counted from hi-2 downward) for every k with L_k < R_k, and the
// final i is min(L_{K+1}, R_K) (L_1 when no swap happens).  One wave computes
// both position lists with ballots and prefix counts, performs the K swaps in
// parallel and parks the pivot: the same permutation as the serial scan.
// Spans are independent, so waves process different spans concurrently in
// rounds, one partition per span and round (breadth first: the longest
// chain of dependent scans is about 2m keys instead of 4m when a wave keeps
// partitioning one side); spans of <= 16 keys are insertion-sorted one lane
// each at the end, spans whose depth ran out are heap-sorted one lane each.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Orders one wave's LDS accesses: those after it see the wave's earlier
// writes (a wave's LDS operations complete in order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Partition of keys[lo..hi] by one full wave; returns the pivot's final index.
__device__ __forceinline__ float lane0_value(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
}
__device__ __forceinline__ double lane0_value(double v) {
  const long long b = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 0);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 0);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Field by field (a copy of the padded fp64 key as a whole goes through
// scratch memory).
template <typename Real>
__device__ __forceinline__ void key_exchange(EdgeKey<Real>& x, EdgeKey<Real>& y) {
  const Real w = x.w;
  const auto i = x.idx;
  x.w = y.w;
  x.idx = y.idx;
  y.w = w;
  y.idx = i;
}

// Spans handed to a partition have hi - lo > 15, so lo, mid, hi - 1 and hi
// are four distinct positions.
template <typename Real>
__device__ int wave_partition(EdgeKey<Real>* keys, int lo, int hi, int16_t* lpos, int16_t* rpos, int lane) {
  // NumPy's median of three (swap mid/lo, hi/mid, mid/lo when strictly
  // smaller, then park the median at hi - 1), on registers: the four keys
  // are read at once and written back once.
  Real piv = 0;
  if (lane == 0) {
    const int mid = lo + ((hi - lo) >> 1);
    EdgeKey<Real> a = keys[lo], m = keys[mid], h = keys[hi];
    const Real park_w = keys[hi - 1].w;
    const auto park_idx = keys[hi - 1].idx;
    if (m.w < a.w) key_exchange(m, a);
    if (h.w < m.w) key_exchange(h, m);
    if (m.w < a.w) key_exchange(m, a);
    keys[lo] = a;
    keys[hi] = h;
    keys[mid].w = park_w;
    keys[mid].idx = park_idx;
    keys[hi - 1] = m;
    piv = m.w;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const Real pivot = lane0_value(piv);
  // One ascending pass over [lo, hi-1]: positions >= pivot in [lo+1, hi-1]
  // (L_1 < L_2 < ...) and positions <= pivot in [lo, hi-2], also ascending
  // (R_k, the k-th from the top, is entry n_le - k).
  int n_ge = 0, n_le = 0;
  for (int base = lo; base <= hi - 1; base += 64) {
    const int p = base + lane;
    const bool in = p <= hi - 1;
    const Real w = in ? keys[p].w : Real(0);
    const bool ge = in && p >= lo + 1 && !(w < pivot);
    const bool le = in && p <= hi - 2 && !(pivot < w);
    const uint64_t mg = __ballot(ge), ml = __ballot(le);
    if (ge) lpos[lo + n_ge + (int)lanes_below(mg)] = (int16_t)p;
    if (le) rpos[lo + n_le + (int)lanes_below(ml)] = (int16_t)p;
    n_ge += __popcll(mg);
    n_le += __popcll(ml);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int16_t* rtop = rpos + lo + n_le - 1;  // rtop[-(k-1)] = R_k
  // K = number of leading k with L_k < R_k (a prefix property).
  const int mlen = min(n_ge, n_le);
  int swaps = 0;
  for (int base = 0; base < mlen; base += 64) {
    const int k = base + lane;
    const bool valid = k < mlen;
    const bool c = valid && lpos[lo + k] < rtop[-k];
    swaps += __popcll(__ballot(c));
    if (__ballot(valid && !c) != 0) break;  // the first false ends the prefix
  }
  for (int base = 0; base < swaps; base += 64) {
    const int k = base + lane;
    if (k < swaps) {
      const int a = lpos[lo + k], b = rtop[-k];
      const EdgeKey<Real> ka = keys[a], kb = keys[b];
      keys[a] = kb;
      keys[b] = ka;
    }
  }
  const int i = swaps == 0 ? lpos[lo] : min((int)lpos[lo + swaps], (int)rtop[-(swaps - 1)]);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) key_swap(keys, i, hi - 1);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return i;
}

// NumPy's insertion sort of keys[lo..hi] (hi - lo <= 15) in registers.
// Insertion sort with a strict comparison is stable, so its result is the
// stable order of the span; adjacent compare-exchanges that swap only on a
// strict inequality (odd-even transposition, as many rounds as slots) give
// exactly that order.  Slots past the span never move.
template <typename Real>
__device__ void small_sort(EdgeKey<Real>* keys, int lo, int hi) {
  constexpr int S = kSmallSort + 1;
  const int len = hi - lo + 1;
  Real w[S];
  int32_t id[S];
#pragma unroll
  for (int q = 0; q < S; ++q) {
    w[q] = q < len ? keys[lo + q].w : Real(0);
    id[q] = q < len ? keys[lo + q].idx : 0;
  }
#pragma unroll
  for (int round = 0; round < S; ++round) {
#pragma unroll
    for (int q = round & 1; q + 1 < S; q += 2) {
      const bool sw = q + 1 < len && w[q + 1] < w[q];
      const Real a = w[q], b = w[q + 1];
      const int32_t ia = id[q], ib = id[q + 1];
      w[q] = sw ? b : a;
      w[q + 1] = sw ? a : b;
      id[q] = sw ? ib : ia;
      id[q + 1] = sw ? ia : ib;
    }
  }
#pragma unroll
  for (int q = 0; q < S; ++q) {
    if (q < len) {
      keys[lo + q].w = w[q];
      keys[lo + q].idx = id[q];
    }
  }
}

// Task lists of the parallel introsort (int16 entries).  Spans queued for a
// round are disjoint with >= 8 keys each (the deferred side of a partition of
// >= 17 keys, or a kept side of >= 18), so a round holds at most m/8 + 1
// tasks; spans left for insertion sort are disjoint with >= 2 keys (at most
// m/2); spans handed to heapsort are popped tasks (at most m/8 + 1 in
// total).
struct SortLists {
  int16_t* cur;    // [3][cap]: lo, hi, 2 depth + (1 for a kept side)
  int16_t* next;   // [3][cap]
  int16_t* heap;   // [2][cap]: lo, hi
  int16_t* small;  // [2][cap_small]: lo, hi
  int cap, cap_small;
};

HDB_HD int sort_task_cap(int n) { return n / 4 + 8; }
HDB_HD int sort_small_cap(int n) { return n / 2 + 8; }
// Bytes of the task lists for a window of n points.
HDB_HD uint32_t sort_lists_bytes(int n) { return 2u * (8u * sort_task_cap(n) + 2u * sort_small_cap(n)); }

// Called by every thread of the workgroup (blockDim.x a multiple of 64).
template <typename Real>
__device__ void parallel_introsort(EdgeKey<Real>* keys, int m, int depth_override, int16_t* lpos,
                                   int16_t* rpos, SortLists L, int* counters) {
  // counters: [0] cur count, [1] next count, [2] small count, [3] heap count
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, waves = blockDim.x >> 6;
  if (tid == 0) {
    int depth = 0;
    for (int c = m; c > 1; c >>= 1) depth += 2;
    if (depth_override > -1000) depth = depth_override;
    L.cur[0] = 0;
    L.cur[L.cap] = (int16_t)(m - 1);
    L.cur[2 * L.cap] = (int16_t)(depth * 2);
    counters[0] = m >= 2 ? 1 : 0;
    counters[1] = 0;
    counters[2] = 0;
    counters[3] = 0;
  }
  __syncthreads();
  int16_t* cur = L.cur;
  int16_t* next = L.next;
  for (;;) {
    const int ncur = counters[0];
    if (ncur == 0) break;
    for (int t = wave; t < ncur; t += waves) {
      const int lo = cur[t], hi = cur[L.cap + t], field = cur[2 * L.cap + t];
      int depth = field >> 1;
      // NumPy tests the depth budget only on spans popped from its stack,
      // never on the side it keeps partitioning (bit 0: a kept side).
      if (!(field & 1) && depth < 0) {
        if (lane == 0) {
          const int h = atomicAdd(&counters[3], 1);
          L.heap[h] = (int16_t)lo;
          L.heap[L.cap + h] = (int16_t)hi;
        }
        continue;
      }
      if (hi - lo > kSmallSort) {
        // One partition per span and round: both sides go to the next round
        // (breadth first), NumPy's pushed side and its kept side alike.
        const int i = wave_partition(keys, lo, hi, lpos, rpos, lane);
        --depth;
        int plo, phi, klo, khi;
        if (i - lo < hi - i) {
          plo = i + 1; phi = hi; klo = lo; khi = i - 1;
        } else {
          plo = lo; phi = i - 1; klo = i + 1; khi = hi;
        }
        if (lane == 0) {
          const int q = atomicAdd(&counters[1], 1);
          next[q] = (int16_t)plo;
          next[L.cap + q] = (int16_t)phi;
          next[2 * L.cap + q] = (int16_t)(depth * 2);
          if (khi - klo > kSmallSort) {
            const int r = atomicAdd(&counters[1], 1);
            next[r] = (int16_t)klo;
            next[L.cap + r] = (int16_t)khi;
            next[2 * L.cap + r] = (int16_t)(depth * 2 + 1);
          } else if (khi > klo) {
            const int r = atomicAdd(&counters[2], 1);
            L.small[r] = (int16_t)klo;
            L.small[L.cap_small + r] = (int16_t)khi;
          }
        }
      } else if (hi > lo && lane == 0) {
        const int q = atomicAdd(&counters[2], 1);
        L.small[q] = (int16_t)lo;
        L.small[L.cap_small + q] = (int16_t)hi;
      }
    }
    __syncthreads();
    if (tid == 0) {
      counters[0] = counters[1];
      counters[1] = 0;
    }
    int16_t* tmp = cur;
    cur = next;
    next = tmp;
    __syncthreads();
  }
  const int nsmall = counters[2], nheap = counters[3];
  for (int s = tid; s < nsmall; s += blockDim.x) small_sort(keys, L.small[s], L.small[L.cap_small + s]);
  for (int h = tid; h < nheap; h += blockDim.x) heap_sort(keys, (int)L.heap[h], (int)L.heap[L.cap + h]);
  __syncthreads();
}

template <typename Real> struct TreeView {
  EdgeKey<Real>* keys;
  int16_t* edge_u;
  int16_t* edge_v;
  int16_t* uf;
  int16_t* comp;
  int16_t* left;
  int16_t* right;
  int16_t* sizes;
  int16_t* parent;
  Real* birth;
  Real* stab;
  uint8_t* selected;
  int16_t* pparent;
  int* stack;
  // Aliases (see TreeLayout lifetimes).
  __device__ int16_t* first_child() const { return left; }
  __device__ int16_t* next_sibling() const { return right; }
  __device__ int16_t* blocked() const { return sizes; }
  __device__ int16_t* stop() const { return uf; }
  __device__ int16_t* mark() const { return comp; }
  __device__ int16_t* label_map() const { return edge_u; }
  __device__ int16_t* resolved() const { return edge_v; }
};

template <typename Real>
__device__ __forceinline__ TreeView<Real> tree_view(unsigned char* base, int n) {
  const TreeLayout L = tree_layout(n, (int)sizeof(Real));
  TreeView<Real> v;
  v.keys = reinterpret_cast<EdgeKey<Real>*>(base + L.keys);
  v.edge_u = reinterpret_cast<int16_t*>(base + L.edges);
  v.edge_v = v.edge_u + n;
  v.uf = reinterpret_cast<int16_t*>(base + L.uf);
  v.comp = v.uf + n;  // contiguous: uf|comp also serves as one 2n array
  v.left = reinterpret_cast<int16_t*>(base + L.left);
  v.right = reinterpret_cast<int16_t*>(base + L.right);
  v.sizes = reinterpret_cast<int16_t*>(base + L.sizes);
  v.parent = reinterpret_cast<int16_t*>(base + L.parent);
  v.birth = reinterpret_cast<Real*>(base + L.birth);
  v.stab = reinterpret_cast<Real*>(base + L.stab);
  v.selected = reinterpret_cast<uint8_t*>(base + L.selected);
  v.pparent = reinterpret_cast<int16_t*>(base + L.pparent);
  v.stack = reinterpret_cast<int*>(base + L.stack);
  return v;
}

// Kruskal over the sorted keys: node n+i joins the newest nodes containing
// edge_u and edge_v of the i-th lightest edge.  Union by size;
// representatives stay leaves, and rec[] holds each representative's newest
// node id | component size << 16 (one 32-bit LDS word: ids < 2^13, sizes <=
// 4096).  left/right/sizes are indexed by node - n.  su/sv hold the
// endpoints already gathered in sorted order (linkage_gather).
template <typename Real>
__device__ void linkage_gather(const TreeView<Real>& t, int n, int16_t* su, int16_t* sv, uint32_t* rec) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (i < n - 1) {
      const int e = t.keys[i].idx;
      su[i] = t.edge_u[e];
      sv[i] = t.edge_v[e];
    }
    t.uf[i] = (int16_t)i;
    rec[i] = (uint32_t)i | (1u << 16);
  }
}

// The linkage run by one wave, 64 edges at a time.  Each lane finds the
// roots of its edge's endpoints in the union-find state at the start of the
// batch (with their records); the unions are then applied in edge order.
// For edge k every lane forms the union of its own edge from its current
// roots (winner, loser, merged record), lane k stores its union and the
// three values are broadcast from lane k, so that later lanes replace a root
// that just lost by the winner and take the winner's new record.  Every edge
// therefore sees exactly the roots, sizes and node ids the sequential
// algorithm sees, and the unions (by size, ties to the first endpoint's
// root) are the same.  Depth stays <= log2(n) without path compression
// because unions are by size.  Per edge: five vector operations for the
// union (no scalar unpacking of lane reads), one lane's two stores, three
// lane reads, two compares and two selects per root.
template <typename Real>
__device__ void linkage_wave(const TreeView<Real>& t, int n, const int16_t* su, const int16_t* sv,
                             uint32_t* rec) {
  const int lane = threadIdx.x & 63;
  int16_t* uf = t.uf;
  for (int i0 = 0; i0 < n - 1; i0 += 64) {
    const int i = i0 + lane;
    const bool valid = i < n - 1;
    int ra = 0, rc = 0;
    uint32_t ka = 0, kc = 0;  // records of the two roots
    if (valid) {
      int a = su[i], c = sv[i];
      for (int p = uf[a]; p != a; p = uf[a]) a = p;
      for (int p = uf[c]; p != c; p = uf[c]) c = p;
      ra = a;
      rc = c;
      ka = rec[a];
      kc = rec[c];
    }
    uint32_t oa = 0, oc = 0;  // this lane's records at its own edge
    // Wave-uniform edge count: the loop test is a scalar compare.
    const int cnt = __builtin_amdgcn_readfirstlane(min(64, n - 1 - i0));
    // Wait for the root and record loads here: a wait inside the loop below
    // would also wait, every edge, for the previous edge's LDS stores.
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    for (int k = 0; k < cnt; ++k) {
      // Every lane forms the union of its own edge from its current roots;
      // lane k's is edge k's (the earlier edges of the batch are applied).
      const bool cw = (ka >> 16) < (kc >> 16);
      const int win = cw ? rc : ra, los = cw ? ra : rc;
      const uint32_t merged = (uint32_t)(n + i0 + lane) | (((ka >> 16) + (kc >> 16)) << 16);
      if (lane == k) {
        uf[los] = (int16_t)win;
        rec[win] = merged;
      }
      const int swin = __builtin_amdgcn_readlane(win, k);
      const int slos = __builtin_amdgcn_readlane(los, k);
      const uint32_t smerged = (uint32_t)__builtin_amdgcn_readlane((int)merged, k);
      // Lane k's own records hold the edge's children and sizes.
      const bool mine = lane == k;
      oa = mine ? ka : oa;
      oc = mine ? kc : oc;
      const bool ma = ra == slos || ra == swin;
      const bool mc = rc == slos || rc == swin;
      ra = ma ? swin : ra;
      ka = ma ? smerged : ka;
      rc = mc ? swin : rc;
      kc = mc ? smerged : kc;
    }
    if (valid) {
      t.left[i] = (int16_t)(oa & 0xFFFFu);
      t.right[i] = (int16_t)(oc & 0xFFFFu);
      t.sizes[i] = (int16_t)((oa >> 16) + (oc >> 16));
    }
  }
}

// Linkage in parallel, from the Prim order.  Prim visits every connected
// component of {MST edges of weight < w} (and of weight <= w) as one run of
// consecutive steps: while such a component has an unvisited vertex, some
// unvisited vertex of it has a key below (or at) w, and no vertex outside it
// has.  Position 0 holds vertex 0 and position p >= 1 the vertex added at step
// p - 1; gap p (between positions p - 1 and p) is that step's edge.  Before the
// edge of sort rank i (weight w) is applied, the component of each endpoint is
// therefore a run of positions bounded by gaps of weight >= w -- gaps whose
// rank is >= r0, the first rank of weight w -- and the newest node of a run of
// more than one position is n + the largest rank inside it, so:
//  * an edge whose weight no other edge has (gap g) joins the run ending at
//    position g - 1 and the run starting at g;
//  * the edges of one weight are applied in rank order over the runs of that
//    threshold with a small union-find (one lane per weight): the run of the
//    added vertex starts at its gap g, the run of the source vertex is found
//    from the source's position.  A run that starts at no gap of the group is
//    the leftmost run of its component at weight <= w and is followed by a gap
//    of the group; that gap's rank names its union-find slot.
// The run queries go through a max tree of gap ranks (leaves: gap p at P + p;
// gap 0 and the padding are sentinels above every rank).  Results are the
// sequential Kruskal's left/right/sizes exactly (reference
// hdbscan_gpu/_kernels.py:518-556).
constexpr int16_t kRankSentinel = 32767;

// Nearest gap p <= q whose rank is >= r (gap 0 is a sentinel: always found).
__device__ __forceinline__ int rank_left(const int16_t* tr, int P, int q, int r) {
  int node = P + q;
  if (tr[node] >= r) return q;
  while (node > 1) {
    if ((node & 1) && tr[node - 1] >= r) {
      node -= 1;
      while (node < P) node = tr[2 * node + 1] >= r ? 2 * node + 1 : 2 * node;
      return node - P;
    }
    node >>= 1;
  }
  return 0;
}

// Nearest gap p > q whose rank is >= r, or n.
__device__ __forceinline__ int rank_right(const int16_t* tr, int P, int q, int r, int n) {
  int node = P + q;
  while (node > 1) {
    if (!(node & 1) && tr[node + 1] >= r) {
      node += 1;
      while (node < P) node = tr[2 * node] >= r ? 2 * node : 2 * node + 1;
      return min(node - P, n);
    }
    node >>= 1;
  }
  return n;
}

// Largest rank over gaps lo..hi (lo <= hi).
__device__ __forceinline__ int rank_max(const int16_t* tr, int P, int lo, int hi) {
  int l = lo + P, r = hi + P + 1, res = -1;
  while (l < r) {
    if (l & 1) res = max(res, (int)tr[l++]);
    if (r & 1) res = max(res, (int)tr[--r]);
    l >>= 1;
    r >>= 1;
  }
  return res;
}

template <typename Real>
__device__ void linkage_parallel(const TreeView<Real>& t, int n) {
  const int tid = threadIdx.x, T = blockDim.x, m = n - 1;
  if (m <= 0) return;
  int P = 1;
  while (P < n) P <<= 1;
  // Arrays free until condensation: uf (positions, then union-find parents of
  // 2n slots over uf|comp), birth..stab (the rank tree, 2P int16 <= 8n bytes,
  // then the slot records, 2n uint32), parent / pparent (tied edges: the run
  // of the added vertex), left / right / sizes (tied edges: the run of the
  // source and its slot, until the union-find pass writes the results).
  int16_t* pos = t.uf;
  int16_t* tr = reinterpret_cast<int16_t*>(t.birth);
  if (tid == 0) pos[0] = 0;
  for (int i = tid; i < P; i += T) {
    if (i < m) {
      pos[t.edge_v[i]] = (int16_t)(i + 1);
      tr[P + t.keys[i].idx + 1] = (int16_t)i;
    }
    if (i == 0 || i >= n) tr[P + i] = kRankSentinel;
  }
  __syncthreads();
  for (int half = P >> 1; half >= 1; half >>= 1) {
    for (int i = half + tid; i < 2 * half; i += T) tr[i] = max(tr[2 * i], tr[2 * i + 1]);
    __syncthreads();
  }
  // The newest node of the run of positions s..e.
  auto run_node = [&](int s, int e) -> int {
    return s == e ? (s == 0 ? 0 : (int)t.edge_v[s - 1]) : n + rank_max(tr, P, s + 1, e);
  };
  for (int i = tid; i < m; i += T) {
    const Real w = t.keys[i].w;
    const bool tie_lo = i > 0 && t.keys[i - 1].w == w;
    const bool tie_hi = i + 1 < m && t.keys[i + 1].w == w;
    int r0 = i, r1 = i + 1;
    if (tie_lo) {
      int lo = 0, hi = i;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (t.keys[mid].w < w) lo = mid + 1; else hi = mid;
      }
      r0 = lo;
    }
    if (tie_hi) {
      int lo = i + 1, hi = m;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (t.keys[mid].w > w) hi = mid; else lo = mid + 1;
      }
      r1 = lo;
    }
    const int g = t.keys[i].idx + 1;
    const int ec = rank_right(tr, P, g, r0, n) - 1;  // the added vertex's run: g..ec
    if (r1 - r0 == 1) {
      const int sa = rank_left(tr, P, g - 1, r0);    // the source's run: sa..g-1
      t.left[i] = (int16_t)run_node(sa, g - 1);
      t.right[i] = (int16_t)run_node(g, ec);
      t.sizes[i] = (int16_t)(ec + 1 - sa);
    } else {
      const int q = pos[t.edge_u[g - 1]];
      const int sa = rank_left(tr, P, q, r0);
      const int ea = rank_right(tr, P, q, r0, n) - 1;
      t.parent[i] = (int16_t)run_node(g, ec);
      t.pparent[i] = (int16_t)(ec - g + 1);
      t.left[i] = (int16_t)run_node(sa, ea);
      t.right[i] = (int16_t)(ea - sa + 1);
      const int rs = tr[P + sa];
      t.sizes[i] = (int16_t)(sa >= 1 && rs >= r0 && rs < r1 ? rs : n + tr[P + ea + 1]);
    }
  }
  __syncthreads();
  // One lane per weight shared by several edges: slots 0..m-1 are the runs
  // starting at the group's gaps (by rank), n + k the leftmost run before the
  // gap of rank k; records hold newest node | size << 16.
  int16_t* par = t.uf;
  uint32_t* rec = reinterpret_cast<uint32_t*>(t.birth);
  for (int i = tid; i < m; i += T) {
    const Real w = t.keys[i].w;
    if ((i > 0 && t.keys[i - 1].w == w) || !(i + 1 < m && t.keys[i + 1].w == w)) continue;
    int r1 = i + 2;
    while (r1 < m && t.keys[r1].w == w) ++r1;
    for (int k = i; k < r1; ++k) {
      par[k] = (int16_t)k;
      rec[k] = (uint32_t)(uint16_t)t.parent[k] | ((uint32_t)(uint16_t)t.pparent[k] << 16);
      const int sl = t.sizes[k];
      if (sl >= n) {
        par[sl] = (int16_t)sl;
        rec[sl] = (uint32_t)(uint16_t)t.left[k] | ((uint32_t)(uint16_t)t.right[k] << 16);
      }
    }
    for (int k = i; k < r1; ++k) {
      int fa = t.sizes[k], fc = k;
      while (par[fa] != fa) {
        par[fa] = par[par[fa]];
        fa = par[fa];
      }
      while (par[fc] != fc) {
        par[fc] = par[par[fc]];
        fc = par[fc];
      }
      const uint32_t ka = rec[fa], kc = rec[fc];
      const uint32_t size = (ka >> 16) + (kc >> 16);
      t.left[k] = (int16_t)(ka & 0xFFFFu);
      t.right[k] = (int16_t)(kc & 0xFFFFu);
      t.sizes[k] = (int16_t)size;
      const uint32_t merged = (uint32_t)(n + k) | (size << 16);
      if ((ka >> 16) < (kc >> 16)) {
        par[fa] = (int16_t)fc;
        rec[fc] = merged;
      } else {
        par[fc] = (int16_t)fa;
        rec[fa] = merged;
      }
    }
  }
}

// Condensation, prepared in parallel.  After linkage the sort index of each
// key is dead; for node n+i:
//   keys[i].w   = lambda = 1/distance (+inf for distance 0)
//   keys[i].idx = next | pre_offset << 13 | split << 27, where next is the
//                 child that keeps the node's cluster (its only child with
//                 >= mcs points; 0 if none), pre_offset that child's preorder
//                 offset (1 for the left child, 1 + nodes(left) = 2*size(left)
//                 for the right) and split marks two such children
//   pruned[i]   = points in children with < mcs points (pruned at this node)
//   node_slot[i] = -1 (set to the cluster slot when a walk visits the node)
template <typename Real>
__device__ void condense_prepare(const TreeView<Real>& t, int n, long long mcs, int16_t* node_slot,
                                 int16_t* pruned) {
  for (int i = threadIdx.x; i < n - 1; i += blockDim.x) {
    const int a = t.left[i], c = t.right[i];
    const int sa = a < n ? 1 : t.sizes[a - n];
    const int sc = c < n ? 1 : t.sizes[c - n];
    const bool la = sa >= mcs, lc = sc >= mcs;
    const Real dist = t.keys[i].w;
    const int next = la && !lc ? a : (lc && !la ? c : 0);
    const int off = la && !lc ? 1 : 2 * sa;
    EdgeKey<Real> k;
    k.w = dist > Real(0) ? Real(1) / dist : Real(INFINITY);
    k.idx = next | (off << 13) | ((la && lc) ? (1 << 27) : 0);
    t.keys[i] = k;
    pruned[i] = (int16_t)((la ? 0 : sa) + (lc ? 0 : sc));
    node_slot[i] = -1;
  }
}

// Per-cluster record of the condensation (device memory, one per cluster,
// indexed by cluster id - n).
template <typename Real> struct ClusterSlot {
  Real birth;        // lambda at which the cluster appears (0 for the root)
  int32_t parent;    // parent cluster id (-1 for the root)
};

// Exclusive prefix sum of v[0..count-1] in place (count <= 4 * blockDim.x),
// v[count] = total.  Called by every thread of the workgroup.
__device__ void block_exclusive_scan(int16_t* v, int count, int* wave_sums) {
  const int tid = threadIdx.x, T = blockDim.x, lane = tid & 63, wave = tid >> 6;
  const int per = (count + T - 1) / T;
  const int lo = min(count, tid * per), hi = min(count, lo + per);
  int sum = 0;
  for (int i = lo; i < hi; ++i) sum += v[i];
  int inc = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wave_sums[wave] = inc;
  __syncthreads();
  int run = inc - sum;
  for (int w = 0; w < wave; ++w) run += wave_sums[w];
  for (int i = lo; i < hi; ++i) {
    const int x = v[i];
    v[i] = (int16_t)run;
    run += x;
  }
  if (hi == count && lo < hi) v[count] = (int16_t)run;
  if (count == 0 && tid == 0) v[0] = 0;
  __syncthreads();
}

// Condensation (after condense_prepare), in parallel over the internal nodes
// n+i of the single-linkage tree.  The nodes kept by the reference's
// breadth-first condensation are the root and every node of >= mcs points;
// per cluster they form one downward path: from the cluster's first node
// (the root, or a child of a split: a node whose two children both have
// >= mcs points) each node's only child of >= mcs points inherits the
// cluster, and the path ends at a split or at a node whose children are
// both smaller.  The reference numbers new clusters in breadth-first order
// of the split nodes, i.e. by (depth, preorder) of those nodes: split number
// r gives its children n+1+2r (left) and n+2+2r (right).  It adds to a
// cluster's stability only along that path, in path order: each smaller
// child adds lambda - birth once per point (all its points leave at the same
// lambda), a split adds (lambda - birth) * size for the left then the right
// child -- so each cluster's sum is formed by one lane, in that order.
//   1. parent of every internal node (scatter from left/right);
//   2. depth and preorder of every kept node by pointer jumping (a child's
//      preorder offset is 1 for the left child, 1 + nodes(left) = 2*size(left)
//      for the right);
//   3. the first node of each kept node's cluster by pointer jumping;
//   4. split ranks by counting over the compacted (depth, preorder) keys,
//      cluster ids, node_slot[i] = cluster id - n of each kept node n+i;
//   5. every path's additions laid out in path order (a scan of their
//      counts places each node's), then each cluster's sequential sum.
// Scratch: uf, comp, edge_v, point_parent, birth and stability (all dead or
// not yet written).  Leaves birth, stability, parent and selected per cluster
// (indexed by cluster id - n).  Needs n - 1 <= 4 * blockDim.x.  Called by
// every thread of the workgroup; returns next_label (uniform).
template <typename Real>
__device__ int condense_clusters(const TreeView<Real>& t, int n, long long mcs, ClusterSlot<Real>* slots,
                                 int16_t* node_slot, const int16_t* pruned, int* counters) {
  __shared__ int wave_sums[16];
  const int tid = threadIdx.x, T = blockDim.x;
  const int m = n - 1;  // internal nodes n+i, i < m; the root is i = m-1
  const int root = m - 1;
  int16_t* par = t.uf;
  int16_t* aux = t.comp;   // jump pointers, then split ranks, then path offsets
  int16_t* dep = t.edge_v;
  int16_t* pre = t.pparent;  // preorder, then pruned counts in path order
  int16_t* hd = reinterpret_cast<int16_t*>(t.birth);
  int16_t* lab = hd + n;
  int16_t* pc = t.pparent;  // additions per path position, then their offsets
  constexpr int PER = 4;
  for (int i = tid; i < m; i += T) {
    const int a = t.left[i], c = t.right[i];
    if (a >= n) par[a - n] = (int16_t)i;
    if (c >= n) par[c - n] = (int16_t)i;
  }
  if (tid == 0) par[root] = -1;
  __syncthreads();
  // 2. depth and preorder.
  bool kept[PER];
  int vd[PER], vp[PER], va[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = tid + r * T;
    kept[r] = i < m && (i == root || t.sizes[i] >= mcs);
    vd[r] = 0;
    vp[r] = 0;
    va[r] = -1;
    if (kept[r] && i != root) {
      const int p = par[i];
      const int a = t.left[p];
      vd[r] = 1;
      vp[r] = a == n + i ? 1 : 2 * (a < n ? 1 : t.sizes[a - n]);
      va[r] = p;
    }
    if (kept[r]) {
      dep[i] = (int16_t)vd[r];
      pre[i] = (int16_t)vp[r];
      aux[i] = (int16_t)va[r];
    }
  }
  __syncthreads();
  for (;;) {
    int nd[PER], np[PER], na[PER];
    bool more = false;
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      nd[r] = vd[r];
      np[r] = vp[r];
      na[r] = va[r];
      if (kept[r] && va[r] >= 0) {
        const int a = va[r];
        nd[r] += dep[a];
        np[r] += pre[a];
        na[r] = aux[a];
        more |= na[r] >= 0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int i = tid + r * T;
      if (kept[r] && va[r] >= 0) {
        vd[r] = nd[r];
        vp[r] = np[r];
        va[r] = na[r];
        dep[i] = (int16_t)nd[r];
        pre[i] = (int16_t)np[r];
        aux[i] = (int16_t)na[r];
      }
    }
    if (!__syncthreads_or(more)) break;
  }
  // 3. first node of each kept node's cluster.
  bool first[PER], split[PER];
  int vh[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = tid + r * T;
    first[r] = false;
    split[r] = false;
    vh[r] = -1;
    if (kept[r]) {
      const int p = par[i];
      first[r] = p < 0 || ((t.keys[p].idx >> 27) & 1);
      split[r] = (t.keys[i].idx >> 27) & 1;
      vh[r] = first[r] ? i : p;
      hd[i] = (int16_t)vh[r];
    }
  }
  __syncthreads();
  for (;;) {
    int nh[PER];
    bool more = false;
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      nh[r] = vh[r];
      if (kept[r] && !first[r]) {
        nh[r] = hd[vh[r]];
        more |= nh[r] != vh[r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int i = tid + r * T;
      if (kept[r] && !first[r] && nh[r] != vh[r]) {
        vh[r] = nh[r];
        hd[i] = (int16_t)nh[r];
      }
    }
    if (!__syncthreads_or(more)) break;
  }
  // 4. split ranks by (depth, preorder), cluster ids.  The split keys are
  // compacted into the path-lambda storage (free until step 5).
  int32_t* keys32 = reinterpret_cast<int32_t*>(t.stab);
  if (tid == 0) counters[3] = 0;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; ++r)
    if (split[r]) keys32[atomicAdd(&counters[3], 1)] = (vd[r] << 13) | vp[r];
  __syncthreads();
  const int splits = counters[3];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    if (split[r]) {
      const int k = (vd[r] << 13) | vp[r];
      int rank = 0;
      for (int q = 0; q < splits; ++q) rank += keys32[q] < k ? 1 : 0;
      aux[tid + r * T] = (int16_t)rank;
    }
  }
  __syncthreads();
  int vl[PER];  // cluster id - n of each first node
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = tid + r * T;
    vl[r] = -1;
    if (first[r]) {
      const int p = par[i];
      vl[r] = p < 0 ? 0 : 1 + 2 * aux[p] + (t.left[p] == n + i ? 0 : 1);
      lab[i] = (int16_t)vl[r];
    }
  }
  __syncthreads();
  const int count = 1 + 2 * splits;  // clusters, root included
  // Path lengths: the last node of each path (a split, or no child kept).
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = tid + r * T;
    if (kept[r]) {
      const int L = lab[vh[r]];
      node_slot[i] = (int16_t)L;
      if (split[r] || (t.keys[i].idx & 0x1fff) == 0) aux[L] = (int16_t)(vd[r] - dep[vh[r]] + 1);
    }
  }
  // Per cluster: birth and parent cluster from its first node.
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    if (first[r]) {
      const int i = tid + r * T;
      const int p = par[i];
      ClusterSlot<Real>& cs = slots[vl[r]];
      cs.birth = p < 0 ? Real(0) : t.keys[p].w;
      cs.parent = p < 0 ? -1 : n + lab[hd[p]];
    }
  }
  __syncthreads();
  block_exclusive_scan(aux, count, wave_sums);
  // 5. the additions of every path, in path order: per position the
  // number of additions (the pruned points, or the two children of a
  // split), an exclusive scan for each position's first one, then every
  // kept node writes its own; one lane per cluster adds its range in order.
  const int npos = aux[count];
  int qpos[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = tid + r * T;
    qpos[r] = 0;
    if (kept[r]) {
      qpos[r] = aux[node_slot[i]] + vd[r] - dep[vh[r]];
      pc[qpos[r]] = split[r] ? (int16_t)2 : pruned[i];
    }
  }
  __syncthreads();
  block_exclusive_scan(pc, npos, wave_sums);
  Real* ev = t.birth;  // birth and stability regions, contiguous: >= 2n Reals
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = tid + r * T;
    if (kept[r]) {
      const Real step = t.keys[i].w - slots[node_slot[i]].birth;
      const int e0 = pc[qpos[r]];
      if (split[r]) {
        const int a = t.left[i], c = t.right[i];
        ev[e0] = step * Real(a < n ? 1 : t.sizes[a - n]);
        ev[e0 + 1] = step * Real(c < n ? 1 : t.sizes[c - n]);
      } else {
        for (int k = 0; k < pruned[i]; ++k) ev[e0 + k] = step;
      }
    }
  }
  __syncthreads();
  Real cb[PER], cstab[PER];
  int cpar[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int L = tid + r * T;
    cb[r] = 0;
    cstab[r] = 0;
    cpar[r] = -1;
    if (L < count) {
      const ClusterSlot<Real> cs = slots[L];
      const int e1 = pc[aux[L + 1]];
      int e = pc[aux[L]];
      Real stab = 0;
      for (; e + 8 <= e1; e += 8) {
        Real v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = ev[e + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) stab += v[u];
      }
      for (; e < e1; ++e) stab += ev[e];
      cb[r] = cs.birth;
      cstab[r] = stab;
      cpar[r] = cs.parent;
    }
  }
  __syncthreads();
  // Final cluster arrays, indexed by cluster id - n (the scratch is dead).
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int L = tid + r * T;
    if (L < count) {
      t.birth[L] = cb[r];
      t.stab[L] = cstab[r];
      t.parent[L] = (int16_t)cpar[r];
      t.selected[L] = (uint8_t)(L == 0 ? 0 : 1);
    }
  }
  __syncthreads();
  return n + count;
}

// Point parents: each point climbs to its first kept ancestor (the node that
// pruned it) and takes that node's cluster.
template <typename Real>
__device__ void condense_points(const TreeView<Real>& t, int n, const int16_t* parent_of,
                                const int16_t* node_slot) {
  for (int leaf = threadIdx.x; leaf < n; leaf += blockDim.x) {
    int x = parent_of[leaf], guard = 0;
    while (x >= n && node_slot[x - n] < 0 && guard++ < n) x = parent_of[x];
    const int s = x >= n ? node_slot[x - n] : -1;
    t.pparent[leaf] = (int16_t)(s >= 0 ? n + s : -1);
  }
}

// Excess-of-mass selection, epsilon merge and label resolution, run by one
// wave level by level over the cluster tree.  Leaves resolved()[c - n] =
// label of cluster c (or -1).  Cluster ids are relative (c = id - n, root 0);
// the children of a split are consecutive, the first at an odd id, so every
// cluster has no child or two (c1, c1 + 1).  Sibling clusters never depend
// on each other, so each pass handles one depth level at a time, in the
// order the serial reference passes imply:
//   EOM (reference hdbscan_gpu/_kernels.py:685-699): deepest level first,
//     sub = (0 + stab(c1)) + stab(c2) (0 for a leaf); sub > stab deselects
//     and replaces stab;
//   suppression under a selected ancestor (:700-709), the epsilon stop
//     (nearest ancestor with 1/birth > eps, else the root's child on the
//     path) and the chosen marks (:716-739): root side first;
//   label numbering in ascending id order (:740-758), then each cluster's
//     label or its parent's, root side first.
// A level reads what the previous level wrote (wave_sync between levels).
template <typename Real>
__device__ void select_clusters_wave(const TreeView<Real>& t, int n, int next_label, Real epsilon) {
  const int lane = threadIdx.x & 63;
  const int count = next_label - n;
  int16_t* child = t.first_child();  // first child, or -1
  int16_t* order = t.next_sibling(); // clusters by depth, root first
  int16_t* blocked = t.blocked();
  int16_t* stop = t.stop();
  int16_t* mark = t.mark();
  int16_t* label_map = t.label_map();
  int16_t* resolved = t.resolved();
  // The sort pairs are dead after condensation: they hold the level ends.
  int16_t* level_end = reinterpret_cast<int16_t*>(t.keys);
  for (int c = lane; c < count; c += 64) {
    child[c] = -1;
    mark[c] = 0;
  }
  wave_sync();
  for (int c = 1 + 2 * lane; c < count; c += 128) child[t.parent[c] - n] = (int16_t)c;
  if (lane == 0) {
    order[0] = 0;
    level_end[0] = 1;
  }
  wave_sync();
  // Breadth-first levels: the children of level k-1 (in its order) form k.
  int levels = 1, lo = 0, hi = 1;
  while (lo < hi) {
    int end = hi;
    for (int i0 = lo; i0 < hi; i0 += 64) {
      const int i = i0 + lane;
      const int ch = i < hi ? child[order[i]] : -1;
      const uint64_t m = __ballot(ch >= 0);
      if (ch >= 0) {
        const int q = end + 2 * (int)lanes_below(m);
        order[q] = (int16_t)ch;
        order[q + 1] = (int16_t)(ch + 1);
      }
      end += 2 * (int)__popcll(m);
    }
    wave_sync();
    if (end > hi) {
      if (lane == 0) level_end[levels] = (int16_t)end;
      ++levels;
    }
    lo = hi;
    hi = end;
    wave_sync();
  }
  // Excess of mass, deepest level first (the root is never selected).
  for (int k = levels - 1; k >= 1; --k) {
    const int b0 = level_end[k - 1], b1 = level_end[k];
    for (int i = b0 + lane; i < b1; i += 64) {
      const int c = order[i];
      const int ch = child[c];
      const Real sub = ch < 0 ? Real(0) : (Real(0) + t.stab[ch]) + t.stab[ch + 1];
      if (sub > t.stab[c]) {
        t.selected[c] = 0;
        t.stab[c] = sub;
      }
    }
    wave_sync();
  }
  // Suppression, epsilon stop and marks, root side first.
  if (lane == 0) blocked[0] = 0;
  wave_sync();
  for (int k = 1; k < levels; ++k) {
    const int b0 = level_end[k - 1], b1 = level_end[k];
    for (int i = b0 + lane; i < b1; i += 64) {
      const int c = order[i];
      const int p = t.parent[c] - n;
      const bool inherited = blocked[p] != 0;
      const bool sel = !inherited && t.selected[c];
      blocked[c] = (int16_t)(inherited || t.selected[c]);
      if (inherited) t.selected[c] = 0;
      if (epsilon > Real(0)) {
        // Ancestor test is strict ">", start test strict "<": they differ at
        // equality on purpose.
        const int st = p == 0 ? c + n : (Real(1) / t.birth[p] > epsilon ? p + n : stop[p]);
        stop[c] = (int16_t)st;
        if (sel) {
          const int chosen = Real(1) / t.birth[c] < epsilon ? st : c + n;
          mark[chosen - n] = 1;
        }
      }
    }
    wave_sync();
  }
  if (epsilon > Real(0)) {
    for (int c = 1 + lane; c < count; c += 64) t.selected[c] = (uint8_t)(mark[c] == 1);
    wave_sync();
  }
  // Labels in ascending id order.
  int label = 0;
  for (int c0 = 1; c0 < count; c0 += 64) {
    const int c = c0 + lane;
    const bool sel = c < count && t.selected[c];
    const uint64_t m = __ballot(sel);
    if (c < count) label_map[c] = sel ? (int16_t)(label + (int)lanes_below(m)) : (int16_t)-1;
    label += (int)__popcll(m);
  }
  if (lane == 0) resolved[0] = -1;
  wave_sync();
  for (int k = 1; k < levels; ++k) {
    const int b0 = level_end[k - 1], b1 = level_end[k];
    for (int i = b0 + lane; i < b1; i += 64) {
      const int c = order[i];
      resolved[c] = t.selected[c] ? label_map[c] : resolved[t.parent[c] - n];
    }
    wave_sync();
  }
}

}  // namespace hdb
