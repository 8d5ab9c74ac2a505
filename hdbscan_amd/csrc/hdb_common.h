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

// The windows of one launch.  Dense: item w is window w, points w*n ..
// w*n+n-1.  Ragged: item i is window list[i] (the launch's windows, largest
// first), points off[w] .. off[w+1]-1.  Persistent launches (queue != null)
// let each workgroup claim items from the device counter *queue until all
// `count` items are taken.
struct Windows {
  const int64_t* off;   // [windows + 1] point offsets, or null (dense)
  const int32_t* list;  // ragged: the launch's windows in claim order
  const int64_t* moff;  // ragged matrix path: element offset of each window's n x n matrix
  int n;                // dense size, or the largest size of the launch
  int count;            // items of the launch
  int* queue;           // claim counter of a persistent launch, or null
  __device__ __forceinline__ void get(int i, int& w, int64_t& base, int& nw) const {
    if (!off) {
      w = i;
      base = (int64_t)i * n;
      nw = n;
      return;
    }
    w = list[i];
    base = off[w];
    nw = (int)(off[w + 1] - base);
  }
  // Element offset of window w's distance matrix (dense: base * n).
  __device__ __forceinline__ int64_t matrix(int w, int64_t base) const { return moff ? moff[w] : base * n; }
};

// Runs body(item) for every item of a launch: one item per workgroup
// (blockIdx.x), or, with a queue, items claimed one at a time until none is
// left (every thread of the workgroup takes part in each body call).
template <typename Body>
__device__ __forceinline__ void for_each_item(const Windows& win, Body body) {
  if (!win.queue) {
    body((int)blockIdx.x);
    return;
  }
  __shared__ int sh_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) sh_item = atomicAdd(win.queue, 1);
    __syncthreads();
    const int item = sh_item;
    if (item >= win.count) break;
    body(item);
  }
}

// Global tree arena and cluster slots of ragged launches are addressed per
// point: window w with first point base uses bytes
// [base * kArenaPerPoint + w * kArenaPerWindow, ...), which covers
// tree_layout(n).total <= 51 n + 1100 for every n.
constexpr size_t kArenaPerPoint = 64, kArenaPerWindow = 4096;

}  // namespace hdb
