// Shared device/host definitions of the gfx950 HDBSCAN kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hdb {

#define HDB_HD __host__ __device__ __forceinline__

// Largest LDS allocation of one workgroup on gfx950 (160 KiB per CU).
constexpr uint32_t kLdsLimit = 160u * 1024u;

template <typename Real> struct RealTraits;
template <> struct RealTraits<float> {
  HDB_HD static float max_value() { return 3.4028234663852886e+38F; }
};
template <> struct RealTraits<double> {
  HDB_HD static double max_value() { return 1.7976931348623157e+308; }
};

HDB_HD uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }
HDB_HD size_t align256(size_t v) { return (v + 255u) & ~size_t(255u); }

// Sort element: the Prim weight and the Prim step that produced it.  The
// introsort compares weights only; the step index travels with it, so
// sorting pairs performs exactly the swaps of NumPy's indirect sort.
template <typename Real> struct alignas(2 * sizeof(Real) > 8 ? 16 : 8) EdgeKey {
  Real w;
  int32_t idx;
};

// Per-window tree arena.  Byte offsets from the window's arena base.
// Lifetimes (arrays sharing storage are never live together):
//   fixed   keys[n] (sort pairs; sorted weights feed condensation)
//           edges[2n] int16: edge_u | edge_v, later the per-node prune
//           clusters (2n entries), later label_map | resolved
//   region  uf|comp (2n entries: linkage) -> condensation queue (node, cluster)
//           -> parents of all nodes -> stop, mark (labels)
//           left, right, sizes (linkage, condense) -> first_child,
//           next_sibling, blocked (selection)
//           parent, birth, stability, selected, point_parent
//           sort stack
struct TreeLayout {
  uint32_t keys, edges, region;
  uint32_t uf, comp, left, right, sizes, parent, birth, stab, selected, pparent, stack;
  uint32_t total;
};

constexpr int kSortStack = 64;

HDB_HD TreeLayout tree_layout(int n, int real_size) {
  TreeLayout L;
  uint32_t o = 0;
  const uint32_t un = (uint32_t)n;
  L.keys = o;
  o = align16(o + (real_size == 4 ? 8u : 16u) * un);
  L.edges = o;
  o = align16(o + 4u * un);
  L.region = o;
  L.uf = o; o = align16(o + 4u * un);
  L.comp = L.uf + 2u * un;
  L.left = o; o = align16(o + 2u * un);
  L.right = o; o = align16(o + 2u * un);
  L.sizes = o; o = align16(o + 2u * un);
  L.parent = o; o = align16(o + 2u * un);
  L.birth = o; o = align16(o + (uint32_t)real_size * un);
  L.stab = o; o = align16(o + (uint32_t)real_size * un);
  L.selected = o; o = align16(o + un);
  L.pparent = o; o = align16(o + 2u * un);
  L.stack = o; o = align16(o + 3u * 4u * kSortStack);
  // The parallel introsort's task lists reuse birth..end (free until
  // condensation); make sure that span is long enough for tiny windows.
  const uint32_t lists = 2u * (8u * (un / 4u + 8u) + 2u * (un / 2u + 8u));
  if (o < L.birth + lists) o = align16(L.birth + lists);
  L.total = o;
  return L;
}

// The windows of one launch.  Dense: window w holds points w*n .. w*n+n-1.
// Ragged: window w holds points off[w] .. off[w+1]-1 and a launch covers the
// windows whose size lies in (lo, n]; a workgroup of another window exits
// before any barrier.
struct Windows {
  const int64_t* off;  // [count + 1] point offsets, or null (dense)
  int n;               // dense size, or the largest size of the launch
  int lo;              // ragged: sizes <= lo belong to another launch
  int count;           // number of windows (ragged)
  __device__ __forceinline__ bool get(int w, int64_t& base, int& nw) const {
    if (!off) {
      base = (int64_t)w * n;
      nw = n;
      return true;
    }
    base = off[w];
    nw = (int)(off[w + 1] - base);
    return nw > lo && nw <= n;
  }
  // Window of point p (ragged: binary search over the offsets).
  __device__ __forceinline__ int window_of(int64_t p) const {
    if (!off) return (int)(p / n);
    int a = 0, b = count;
    while (b - a > 1) {
      const int m = (a + b) >> 1;
      if (off[m] <= p) a = m; else b = m;
    }
    return a;
  }
};

// Global tree arena and cluster slots of ragged launches are addressed per
// point: window w with first point base uses bytes
// [base * kArenaPerPoint + w * kArenaPerWindow, ...), which covers
// tree_layout(n).total <= 51 n + 1100 for every n.
constexpr size_t kArenaPerPoint = 64, kArenaPerWindow = 4096;

}  // namespace hdb
