// Tree stages of one window: NumPy-order introsort of the Prim edges,
// union-find linkage, breadth-first condensation, excess-of-mass selection,
// epsilon merge and label resolution.
//
// These are sequential by definition (each step depends on the previous
// one); one lane runs them out of the window's arena (LDS when it fits).
// Semantics follow the reference stage by stage:
//   sort      reference hdbscan_gpu/_kernels.py:358-516
//   linkage   reference hdbscan_gpu/_kernels.py:518-556
//   condense  reference hdbscan_gpu/_kernels.py:558-643
//   eom       reference hdbscan_gpu/_kernels.py:671-709
//   labels    reference hdbscan_gpu/_kernels.py:712-760
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
// rounds; spans of <= 16 keys are insertion-sorted one lane each at the end,
// spans whose depth ran out are heap-sorted one lane each.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Partition of keys[lo..hi] by one full wave; returns the pivot's final index.
template <typename Real>
__device__ int wave_partition(EdgeKey<Real>* keys, int lo, int hi, int16_t* lpos, int16_t* rpos, int lane) {
  if (lane == 0) {
    const int mid = lo + ((hi - lo) >> 1);
    if (keys[mid].w < keys[lo].w) key_swap(keys, mid, lo);
    if (keys[hi].w < keys[mid].w) key_swap(keys, hi, mid);
    if (keys[mid].w < keys[lo].w) key_swap(keys, mid, lo);
    key_swap(keys, mid, hi - 1);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const Real pivot = keys[hi - 1].w;
  // Positions >= pivot in [lo+1, hi-1], ascending.
  int n_ge = 0;
  for (int base = lo + 1; base <= hi - 1; base += 64) {
    const int p = base + lane;
    const bool ge = p <= hi - 1 && !(keys[p].w < pivot);
    const uint64_t m = __ballot(ge);
    if (ge) lpos[lo + n_ge + (int)lanes_below(m)] = (int16_t)p;
    n_ge += __popcll(m);
  }
  // Positions <= pivot in [lo, hi-2], descending.
  int n_le = 0;
  for (int top = hi - 2; top >= lo; top -= 64) {
    const int p = top - lane;
    const bool le = p >= lo && !(pivot < keys[p].w);
    const uint64_t m = __ballot(le);
    if (le) rpos[lo + n_le + (int)lanes_below(m)] = (int16_t)p;
    n_le += __popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // K = number of leading k with L_k < R_k (a prefix property).
  const int mlen = min(n_ge, n_le);
  int swaps = 0;
  for (int base = 0; base < mlen; base += 64) {
    const int k = base + lane;
    const bool valid = k < mlen;
    const bool c = valid && lpos[lo + k] < rpos[lo + k];
    swaps += __popcll(__ballot(c));
    if (__ballot(valid && !c) != 0) break;  // the first false ends the prefix
  }
  for (int base = 0; base < swaps; base += 64) {
    const int k = base + lane;
    if (k < swaps) {
      const int a = lpos[lo + k], b = rpos[lo + k];
      const EdgeKey<Real> ka = keys[a], kb = keys[b];
      keys[a] = kb;
      keys[b] = ka;
    }
  }
  const int i = swaps == 0 ? lpos[lo] : min((int)lpos[lo + swaps], (int)rpos[lo + swaps - 1]);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) key_swap(keys, i, hi - 1);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return i;
}

// Task lists of the parallel introsort (int16 entries).  Spans pushed in one
// round are disjoint with >= 8 keys each (the deferred side of a partition of
// >= 17 keys), so a round holds at most m/8 + 1 tasks; spans left for
// insertion sort are disjoint with >= 2 keys (at most m/2); spans handed to
// heapsort are popped tasks (at most m/8 + 1 in total).
struct SortLists {
  int16_t* cur;    // [3][cap]: lo, hi, depth
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
    L.cur[2 * L.cap] = (int16_t)depth;
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
      int lo = cur[t], hi = cur[L.cap + t], depth = cur[2 * L.cap + t];
      if (depth < 0) {
        if (lane == 0) {
          const int h = atomicAdd(&counters[3], 1);
          L.heap[h] = (int16_t)lo;
          L.heap[L.cap + h] = (int16_t)hi;
        }
        continue;
      }
      while (hi - lo > kSmallSort) {
        const int i = wave_partition(keys, lo, hi, lpos, rpos, lane);
        --depth;
        int plo, phi;
        if (i - lo < hi - i) {
          plo = i + 1; phi = hi; hi = i - 1;
        } else {
          plo = lo; phi = i - 1; lo = i + 1;
        }
        if (lane == 0) {
          const int q = atomicAdd(&counters[1], 1);
          next[q] = (int16_t)plo;
          next[L.cap + q] = (int16_t)phi;
          next[2 * L.cap + q] = (int16_t)depth;
        }
      }
      if (hi > lo && lane == 0) {
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
  for (int s = tid; s < nsmall; s += blockDim.x) {
    const int lo = L.small[s], hi = L.small[L.cap_small + s];
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
// edge_u and edge_v of the i-th lightest edge.  Union by size with path
// halving; representatives stay leaves so comp[] maps them to their newest
// node and rsize[] holds their component size.  left/right/sizes are indexed
// by node - n.  su/sv hold the endpoints already gathered in sorted order
// (linkage_gather), so the only dependent loads per edge are the two finds,
// which are advanced together.
template <typename Real>
__device__ void linkage_gather(const TreeView<Real>& t, int n, int16_t* su, int16_t* sv, int16_t* rsize) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (i < n - 1) {
      const int e = t.keys[i].idx;
      su[i] = t.edge_u[e];
      sv[i] = t.edge_v[e];
    }
    t.uf[i] = (int16_t)i;
    t.comp[i] = (int16_t)i;
    rsize[i] = 1;
  }
}

template <typename Real>
__device__ void linkage(const TreeView<Real>& t, int n, const int16_t* su, const int16_t* sv, int16_t* rsize) {
  int16_t* uf = t.uf;
  for (int i = 0; i < n - 1; ++i) {
    int a = su[i], c = sv[i];
    for (;;) {
      const int pa = uf[a], pc = uf[c];
      if (pa == a && pc == c) break;
      if (pa != a) {
        const int g = uf[pa];
        uf[a] = (int16_t)g;
        a = g;
      }
      if (pc != c) {
        const int g = uf[pc];
        uf[c] = (int16_t)g;
        c = g;
      }
    }
    const int ca = t.comp[a], cc = t.comp[c];
    const int sa = rsize[a], sc = rsize[c];
    t.left[i] = (int16_t)ca;
    t.right[i] = (int16_t)cc;
    t.sizes[i] = (int16_t)(sa + sc);
    if (sa < sc) {
      uf[a] = (int16_t)c;
      t.comp[c] = (int16_t)(n + i);
      rsize[c] = (int16_t)(sa + sc);
    } else {
      uf[c] = (int16_t)a;
      t.comp[a] = (int16_t)(n + i);
      rsize[a] = (int16_t)(sa + sc);
    }
  }
}

// Condensation, prepared in parallel.  After linkage the sort index of each
// key is dead: keys[i] becomes {lambda of node n+i, child sizes packed
// size_left | size_right << 16}.
template <typename Real>
__device__ void condense_prepare(const TreeView<Real>& t, int n, int16_t* prune_cluster) {
  for (int i = threadIdx.x; i < n - 1; i += blockDim.x) {
    const int a = t.left[i], c = t.right[i];
    const int sa = a < n ? 1 : t.sizes[a - n];
    const int sc = c < n ? 1 : t.sizes[c - n];
    const Real dist = t.keys[i].w;
    EdgeKey<Real> k;
    k.w = dist > Real(0) ? Real(1) / dist : Real(INFINITY);
    k.idx = sa | (sc << 16);
    t.keys[i] = k;
  }
  for (int x = threadIdx.x; x < 2 * n; x += blockDim.x) prune_cluster[x] = -1;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    t.parent[i] = -1;
    t.selected[i] = 0;
    t.birth[i] = 0;
    t.stab[i] = 0;
  }
}

// Breadth-first condensation from root 2n-2 (cluster n), visiting only the
// nodes that stay in the condensed tree (the root and nodes of >= mcs
// points).  Children of a split with both sides >= mcs get consecutive new
// cluster IDs; a smaller child is pruned: each of its `size` points adds
// lambda - birth to the cluster's stability, one addition at a time in the
// reference's order (all its points share the same lambda), and the subtree
// root remembers its cluster for condense_points.  The queue carries
// (node, cluster) pairs.  Returns next_label, or -1 on a queue overflow
// (malformed tree).
template <typename Real>
__device__ int condense(const TreeView<Real>& t, int n, long long mcs, int16_t* qnode, int16_t* qclus,
                        int16_t* prune_cluster) {
  const int root = 2 * n - 2;
  int head = 0, tail = 1, next_label = n + 1;
  qnode[0] = (int16_t)root;
  qclus[0] = (int16_t)n;
  while (head < tail) {
    const int node = qnode[head], cluster = qclus[head];
    ++head;
    const int i = node - n;
    const EdgeKey<Real> info = t.keys[i];
    const int a = t.left[i], c = t.right[i];
    const Real cluster_birth = t.birth[cluster - n];
    Real stab = t.stab[cluster - n];
    const Real lambda = info.w;
    const int size_a = info.idx & 0xffff, size_c = (info.idx >> 16) & 0xffff;
    const bool large_a = size_a >= mcs, large_c = size_c >= mcs;
    if (large_a && large_c) {
      if (tail + 2 > n || next_label + 2 > 2 * n) return -1;
      for (int k = 0; k < 2; ++k) {
        const int child = k == 0 ? a : c;
        const int label = next_label++;
        t.parent[label - n] = (int16_t)cluster;
        t.birth[label - n] = lambda;
        t.selected[label - n] = 1;
        stab += (lambda - cluster_birth) * Real(k == 0 ? size_a : size_c);
        qnode[tail] = (int16_t)child;
        qclus[tail] = (int16_t)label;
        ++tail;
      }
    } else {
      const Real step = lambda - cluster_birth;
      for (int k = 0; k < 2; ++k) {
        const int child = k == 0 ? a : c;
        if (k == 0 ? large_a : large_c) {
          if (tail + 1 > n) return -1;
          qnode[tail] = (int16_t)child;
          qclus[tail] = (int16_t)cluster;
          ++tail;
        } else {
          prune_cluster[child] = (int16_t)cluster;
          const int size = k == 0 ? size_a : size_c;
          for (int q = 0; q < size; ++q) stab += step;
        }
      }
    }
    t.stab[cluster - n] = stab;
  }
  return next_label;
}

// Point parents: each point climbs to the root of the pruned subtree that
// contains it and takes the cluster recorded there.
template <typename Real>
__device__ void condense_points(const TreeView<Real>& t, int n, const int16_t* parent_of,
                                const int16_t* prune_cluster) {
  for (int leaf = threadIdx.x; leaf < n; leaf += blockDim.x) {
    int x = leaf, guard = 0;
    while (prune_cluster[x] < 0 && guard++ < 2 * n) x = parent_of[x];
    t.pparent[leaf] = prune_cluster[x];
  }
}

// Excess-of-mass selection, epsilon merge and label resolution.  Leaves
// resolved()[c - n] = label of cluster c (or -1).
template <typename Real>
__device__ void select_clusters(const TreeView<Real>& t, int n, int next_label, Real epsilon) {
  int16_t* first_child = t.first_child();
  int16_t* next_sibling = t.next_sibling();
  int16_t* blocked = t.blocked();
  const int count = next_label - n;
  for (int c = 0; c < count; ++c) {
    first_child[c] = -1;
    next_sibling[c] = -1;
  }
  // Descending insertion keeps each child list in ascending ID order.
  for (int child = next_label - 1; child > n; --child) {
    const int p = t.parent[child - n];
    next_sibling[child - n] = first_child[p - n];
    first_child[p - n] = (int16_t)child;
  }
  for (int node = next_label - 1; node > n; --node) {
    Real sub = 0;
    for (int ch = first_child[node - n]; ch != -1; ch = next_sibling[ch - n]) sub += t.stab[ch - n];
    if (sub > t.stab[node - n]) {
      t.selected[node - n] = 0;
      t.stab[node - n] = sub;
    }
  }
  // Parents carry smaller IDs: one forward pass suppresses every descendant
  // of a selected cluster.
  blocked[0] = 0;
  for (int node = n + 1; node < next_label; ++node) {
    const int inherited = blocked[t.parent[node - n] - n];
    blocked[node - n] = (int16_t)(inherited || t.selected[node - n]);
    if (inherited) t.selected[node - n] = 0;
  }
  if (epsilon > Real(0)) {
    int16_t* stop = t.stop();
    int16_t* mark = t.mark();
    for (int c = 0; c < count; ++c) mark[c] = 0;
    for (int node = n + 1; node < next_label; ++node) {
      const int anc = t.parent[node - n];
      // Ancestor test is strict ">", start test strict "<": they differ at
      // equality on purpose.
      stop[node - n] = (int16_t)(anc == n ? node
                                          : (Real(1) / t.birth[anc - n] > epsilon ? anc : stop[anc - n]));
      if (!t.selected[node - n]) continue;
      const int chosen = Real(1) / t.birth[node - n] < epsilon ? stop[node - n] : node;
      mark[chosen - n] = 1;
    }
    for (int node = n + 1; node < next_label; ++node) t.selected[node - n] = (uint8_t)(mark[node - n] == 1);
  }
  int16_t* label_map = t.label_map();
  int16_t* resolved = t.resolved();
  int label = 0;
  for (int node = n + 1; node < next_label; ++node)
    label_map[node - n] = t.selected[node - n] ? (int16_t)label++ : (int16_t)-1;
  resolved[0] = -1;
  for (int node = n + 1; node < next_label; ++node)
    resolved[node - n] = t.selected[node - n] ? label_map[node - n] : resolved[t.parent[node - n] - n];
}

}  // namespace hdb
