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
  __device__ int16_t* relabel() const { return uf; }
  __device__ int16_t* queue() const { return comp; }
  __device__ int16_t* prune() const { return edge_u; }  // 2n entries: edge_u | edge_v
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
  v.comp = reinterpret_cast<int16_t*>(base + L.comp);
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
// node.  left/right/sizes are indexed by node - n.
template <typename Real>
__device__ void linkage(const TreeView<Real>& t, int n) {
  for (int i = 0; i < n; ++i) {
    t.uf[i] = (int16_t)i;
    t.comp[i] = (int16_t)i;
  }
  for (int i = 0; i < n - 1; ++i) {
    const int e = t.keys[i].idx;
    int a = t.edge_u[e], c = t.edge_v[e];
    for (int p = t.uf[a]; p != a; p = t.uf[a]) {
      const int g = t.uf[p];
      t.uf[a] = (int16_t)g;
      a = g;
    }
    for (int p = t.uf[c]; p != c; p = t.uf[c]) {
      const int g = t.uf[p];
      t.uf[c] = (int16_t)g;
      c = g;
    }
    const int ca = t.comp[a], cc = t.comp[c];
    const int sa = ca < n ? 1 : t.sizes[ca - n];
    const int sc = cc < n ? 1 : t.sizes[cc - n];
    t.left[i] = (int16_t)ca;
    t.right[i] = (int16_t)cc;
    t.sizes[i] = (int16_t)(sa + sc);
    if (sa < sc) {
      t.uf[a] = (int16_t)c;
      t.comp[c] = (int16_t)(n + i);
    } else {
      t.uf[c] = (int16_t)a;
      t.comp[a] = (int16_t)(n + i);
    }
  }
}

// Breadth-first condensation from root 2n-2 (cluster n).  Returns next_label
// or -1 if a queue bound would be exceeded (malformed tree).
template <typename Real>
__device__ int condense(const TreeView<Real>& t, int n, long long mcs) {
  int16_t* relabel = t.relabel();
  int16_t* queue = t.queue();
  int16_t* prune = t.prune();
  for (int i = 0; i < n; ++i) {
    relabel[i] = -1;
    t.parent[i] = -1;
    t.selected[i] = 0;
    t.birth[i] = 0;
    t.stab[i] = 0;
    t.pparent[i] = -1;
  }
  const int root = 2 * n - 2;
  int head = 0, tail = 1, next_label = n + 1;
  queue[0] = (int16_t)root;
  relabel[root - n] = (int16_t)n;
  while (head < tail) {
    const int node = queue[head++];
    const int a = t.left[node - n], c = t.right[node - n];
    const int cluster = relabel[node - n];
    const Real cluster_birth = t.birth[cluster - n];
    Real cluster_stab = t.stab[cluster - n];
    const Real dist = t.keys[node - n].w;
    const Real lambda = dist > Real(0) ? Real(1) / dist : Real(INFINITY);
    const long long size_a = a < n ? 1 : t.sizes[a - n];
    const long long size_c = c < n ? 1 : t.sizes[c - n];
    const bool large_a = size_a >= mcs, large_c = size_c >= mcs;
    if (large_a && large_c) {
      if (tail + 2 > n - 1 || next_label + 2 > 2 * n) return -1;
      relabel[a - n] = (int16_t)next_label;
      t.parent[next_label - n] = (int16_t)cluster;
      t.birth[next_label - n] = lambda;
      t.selected[next_label - n] = 1;
      ++next_label;
      cluster_stab += (lambda - cluster_birth) * Real(size_a);
      queue[tail++] = (int16_t)a;
      relabel[c - n] = (int16_t)next_label;
      t.parent[next_label - n] = (int16_t)cluster;
      t.birth[next_label - n] = lambda;
      t.selected[next_label - n] = 1;
      ++next_label;
      cluster_stab += (lambda - cluster_birth) * Real(size_c);
      queue[tail++] = (int16_t)c;
    } else {
      for (int k = 0; k < 2; ++k) {
        const int child = k == 0 ? a : c;
        if (k == 0 ? large_a : large_c) {
          if (tail + 1 > n - 1) return -1;
          relabel[child - n] = (int16_t)cluster;
          queue[tail++] = (int16_t)child;
        } else {
          int ph = 0, pt = 0, leaf = child;
          for (;;) {
            if (leaf < n) {
              t.pparent[leaf] = (int16_t)cluster;
              cluster_stab += lambda - cluster_birth;
            } else {
              if (pt + 2 > 2 * n) return -1;
              prune[pt++] = t.left[leaf - n];
              prune[pt++] = t.right[leaf - n];
            }
            if (ph == pt) break;
            leaf = prune[ph++];
          }
        }
      }
    }
    t.stab[cluster - n] = cluster_stab;
  }
  return next_label;
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
