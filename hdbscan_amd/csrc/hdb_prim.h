// Cooperative Prim over the mutual-reachability graph of one window.
//
// One workgroup owns one window.  Thread t owns vertices t, t+T, t+2T, ...
// (VPT slots) and keeps, per slot, the vertex's coordinates (points path),
// core distance, best weight, best source and a visited bit in registers.
// Each step:
//   1. every thread forms the exact distance from the current vertex to its
//      unvisited vertices (coordinate order, no FMA, correctly rounded sqrt;
//      the current vertex is broadcast from LDS) -- or reads the row of a
//      materialised distance matrix (matrix path),
//   2. w = max(core_cur, core_j, d); strictly smaller w replaces best and
//      source (equal weights keep the older source),
//   3. the block-wide argmin of (best, vertex) -- lowest vertex on ties -- is
//      one wave64 butterfly per wave plus one LDS exchange and ONE barrier.
// Reference semantics: hdbscan_gpu/_kernels.py:164-320 (and sklearn
// _linkage.pyx:111-224).  Output edges are (source, next, weight) in step
// order, with source 0 when the weight is not below REAL_MAX
// (reference hdbscan_gpu/_kernels.py:307-317).  The update and tie rules are
// scikit-learn's (BSD-3; notices in THIRD_PARTY_NOTICES.md at the repository
// root).
#pragma once

#include "hdb_common.h"

namespace hdb {

// Candidate (weight, vertex) ordered lexicographically.  Weights are
// non-negative or +inf, so their IEEE bit patterns order like the values.
template <typename Real> struct Cand;

template <> struct Cand<float> {
  uint64_t k;
  __device__ __forceinline__ static Cand make(float w, int j) {
    Cand c;
    c.k = ((uint64_t)__float_as_uint(w) << 32) | (uint32_t)j;
    return c;
  }
  __device__ __forceinline__ bool less(const Cand& o) const { return k < o.k; }
  __device__ __forceinline__ Cand xor_shuffle(int m) const {
    Cand c;
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)k, m);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(k >> 32), m);
    c.k = ((uint64_t)hi << 32) | lo;
    return c;
  }
  __device__ __forceinline__ int index() const { return (int)(uint32_t)k; }
  __device__ __forceinline__ float weight() const { return __uint_as_float((uint32_t)(k >> 32)); }
  // Above every (weight, vertex) key, the REAL_MAX start value included.
  __device__ __forceinline__ static Cand none() {
    Cand c;
    c.k = ~0ull;
    return c;
  }
};

template <> struct Cand<double> {
  uint64_t w;
  uint32_t j;
  __device__ __forceinline__ static Cand make(double v, int idx) {
    Cand c;
    c.w = (uint64_t)__double_as_longlong(v);
    c.j = (uint32_t)idx;
    return c;
  }
  __device__ __forceinline__ bool less(const Cand& o) const { return w < o.w || (w == o.w && j < o.j); }
  __device__ __forceinline__ Cand xor_shuffle(int m) const {
    Cand c;
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)w, m);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(w >> 32), m);
    c.w = ((uint64_t)hi << 32) | lo;
    c.j = (uint32_t)__shfl_xor((int)j, m);
    return c;
  }
  __device__ __forceinline__ int index() const { return (int)j; }
  __device__ __forceinline__ double weight() const { return __longlong_as_double((long long)w); }
  __device__ __forceinline__ static Cand none() {
    Cand c;
    c.w = ~0ull;
    c.j = ~0u;
    return c;
  }
};

// c ? x : y, field by field (a select of whole structs goes through
// scratch memory for the padded fp64 candidate).
__device__ __forceinline__ Cand<float> cand_pick(bool c, const Cand<float>& x, const Cand<float>& y) {
  Cand<float> r;
  r.k = c ? x.k : y.k;
  return r;
}
__device__ __forceinline__ Cand<double> cand_pick(bool c, const Cand<double>& x, const Cand<double>& y) {
  Cand<double> r;
  r.w = c ? x.w : y.w;
  r.j = c ? x.j : y.j;
  return r;
}

// DPP lane moves within 16-lane rows (quad_perm xor 1, xor 2; row_ror 4, 8).
// Every lane is written (all rows and banks, sources inside the row), so no
// old value is needed: mov_dpp avoids copying v into the destination first.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_move(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
}
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppRor4 = 0x124, kDppRor8 = 0x128;

__device__ __forceinline__ uint32_t read_lane(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

template <int CTRL>
__device__ __forceinline__ Cand<float> dpp_cand(const Cand<float>& c) {
  Cand<float> o;
  o.k = ((uint64_t)dpp_move<CTRL>((uint32_t)(c.k >> 32)) << 32) | dpp_move<CTRL>((uint32_t)c.k);
  return o;
}
template <int CTRL>
__device__ __forceinline__ Cand<double> dpp_cand(const Cand<double>& c) {
  Cand<double> o;
  o.w = ((uint64_t)dpp_move<CTRL>((uint32_t)(c.w >> 32)) << 32) | dpp_move<CTRL>((uint32_t)c.w);
  o.j = dpp_move<CTRL>(c.j);
  return o;
}
__device__ __forceinline__ Cand<float> lane_cand(const Cand<float>& c, int lane) {
  Cand<float> o;
  o.k = ((uint64_t)read_lane((uint32_t)(c.k >> 32), lane) << 32) | read_lane((uint32_t)c.k, lane);
  return o;
}
__device__ __forceinline__ Cand<double> lane_cand(const Cand<double>& c, int lane) {
  Cand<double> o;
  o.w = ((uint64_t)read_lane((uint32_t)(c.w >> 32), lane) << 32) | read_lane((uint32_t)c.w, lane);
  o.j = read_lane(c.j, lane);
  return o;
}

// Minimum within each 16-lane row, left in every lane of the row.
template <typename Real>
__device__ __forceinline__ Cand<Real> row_min(Cand<Real> c) {
  Cand<Real> o = dpp_cand<kDppXor1>(c);
  if (o.less(c)) c = o;
  o = dpp_cand<kDppXor2>(c);
  if (o.less(c)) c = o;
  o = dpp_cand<kDppRor4>(c);
  if (o.less(c)) c = o;
  o = dpp_cand<kDppRor8>(c);
  if (o.less(c)) c = o;
  return c;
}

// DPP move that writes only the rows in ROWS; the other rows hold undefined
// values (wave_min reads only lanes that every step wrote).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_rows(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, ROWS, 0xF, false);
}
constexpr int kDppBcast15 = 0x142, kDppBcast31 = 0x143;
template <int CTRL, int ROWS>
__device__ __forceinline__ Cand<float> dpp_cand_rows(const Cand<float>& c) {
  Cand<float> o;
  o.k = ((uint64_t)dpp_rows<CTRL, ROWS>((uint32_t)(c.k >> 32)) << 32) | dpp_rows<CTRL, ROWS>((uint32_t)c.k);
  return o;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ Cand<double> dpp_cand_rows(const Cand<double>& c) {
  Cand<double> o;
  o.w = ((uint64_t)dpp_rows<CTRL, ROWS>((uint32_t)(c.w >> 32)) << 32) | dpp_rows<CTRL, ROWS>((uint32_t)c.w);
  o.j = dpp_rows<CTRL, ROWS>(c.j);
  return o;
}

// Wave-wide minimum, returned wave-uniform (scalar registers).  All 64 lanes
// must be active.  Row minima, then row broadcasts: lane 15 into row 1 and
// lane 47 into row 3, then lane 31 into rows 2 and 3; lane 63 holds the
// minimum of all four rows.  Rows 0 and 2 after the first broadcast, and
// rows 0 and 1 after the second, are never read: row 3 takes its own row
// minimum, row 2's (lane 47) and row 1's combined value (lane 31).
template <typename Real>
__device__ __forceinline__ Cand<Real> wave_min(Cand<Real> c) {
  c = row_min(c);
  Cand<Real> o = dpp_cand_rows<kDppBcast15, 0xA>(c);
  if (o.less(c)) c = o;
  o = dpp_cand_rows<kDppBcast31, 0xC>(c);
  if (o.less(c)) c = o;
  return lane_cand(c, 63);
}

// DPP move whose unwritten lanes hold all ones, the identity of an unsigned
// minimum: with that immediate as the old value the move folds into the min
// that consumes it (one v_min_u32 with a DPP source per level).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_umin_src(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)v, CTRL, ROWS, 0xF, false);
}

// Wave-wide unsigned minimum of a 32-bit value, the same lane pattern as
// wave_min.
__device__ __forceinline__ uint32_t wave_umin32(uint32_t v) {
  v = min(v, dpp_umin_src<kDppXor1, 0xF>(v));
  v = min(v, dpp_umin_src<kDppXor2, 0xF>(v));
  v = min(v, dpp_umin_src<kDppRor4, 0xF>(v));
  v = min(v, dpp_umin_src<kDppRor8, 0xF>(v));
  v = min(v, dpp_umin_src<kDppBcast15, 0xA>(v));
  v = min(v, dpp_umin_src<kDppBcast31, 0xC>(v));
  return read_lane(v, 63);
}

// Wave-wide unsigned maximum of a 32-bit value (the same lane pattern; the
// unwritten lanes of each move read 0, the identity of the maximum).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_umax_src(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ uint32_t wave_umax32(uint32_t v) {
  v = max(v, dpp_umax_src<kDppXor1, 0xF>(v));
  v = max(v, dpp_umax_src<kDppXor2, 0xF>(v));
  v = max(v, dpp_umax_src<kDppRor4, 0xF>(v));
  v = max(v, dpp_umax_src<kDppRor8, 0xF>(v));
  v = max(v, dpp_umax_src<kDppBcast15, 0xA>(v));
  v = max(v, dpp_umax_src<kDppBcast31, 0xC>(v));
  return read_lane(v, 63);
}

// The fp32 candidate key in two 32-bit steps: the smallest weight pattern
// (six single-instruction levels), then the smallest vertex among the lanes
// holding it -- the lexicographic minimum of (weight, vertex), i.e. the
// 64-bit key minimum.  One lane holds that weight almost always (a wave's
// minimum is tied in about 1 step in 300 at N=1024, D=16): its vertex is
// read from it directly; ties take a second six-level pass over vertices.
__device__ __forceinline__ Cand<float> wave_min(Cand<float> c) {
  const uint32_t w = (uint32_t)(c.k >> 32);
  const uint32_t wm = wave_umin32(w);
  const uint64_t tied = __builtin_amdgcn_ballot_w64(w == wm);
  uint32_t jm;
  if (__builtin_popcountll(tied) == 1)
    jm = read_lane((uint32_t)c.k, (int)__builtin_ctzll(tied));
  else
    jm = wave_umin32(w == wm ? (uint32_t)c.k : 0xFFFFFFFFu);
  Cand<float> r;
  r.k = ((uint64_t)wm << 32) | jm;
  return r;
}

// Wave-wide unsigned 64-bit minimum, the same lane pattern as wave_min.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  return ((uint64_t)dpp_rows<CTRL, ROWS>((uint32_t)(v >> 32)) << 32) | dpp_rows<CTRL, ROWS>((uint32_t)v);
}
__device__ __forceinline__ uint64_t read_lane64(uint64_t v, int lane) {
  return ((uint64_t)read_lane((uint32_t)(v >> 32), lane) << 32) | read_lane((uint32_t)v, lane);
}
__device__ __forceinline__ uint64_t wave_umin64(uint64_t v) {
  v = min(v, dpp_u64<kDppXor1, 0xF>(v));
  v = min(v, dpp_u64<kDppXor2, 0xF>(v));
  v = min(v, dpp_u64<kDppRor4, 0xF>(v));
  v = min(v, dpp_u64<kDppRor8, 0xF>(v));
  v = min(v, dpp_u64<kDppBcast15, 0xA>(v));
  v = min(v, dpp_u64<kDppBcast31, 0xC>(v));
  return read_lane64(v, 63);
}

// The fp64 candidate key in the same two steps as the fp32 one: the
// smallest weight pattern (64-bit levels), then the vertex from the only
// lane holding it, or a second pass over vertices when it is tied.
__device__ __forceinline__ Cand<double> wave_min(Cand<double> c) {
  const uint64_t wm = wave_umin64(c.w);
  const uint64_t tied = __builtin_amdgcn_ballot_w64(c.w == wm);
  Cand<double> r;
  r.w = wm;
  if (__builtin_popcountll(tied) == 1)
    r.j = read_lane(c.j, (int)__builtin_ctzll(tied));
  else
    r.j = wave_umin32(c.w == wm ? c.j : 0xFFFFFFFFu);
  return r;
}

// Minimum over lanes 0..15, returned wave-uniform (the other rows are not
// read).  All 64 lanes must be active.
template <typename Real>
__device__ __forceinline__ Cand<Real> row0_min(Cand<Real> c) {
  return lane_cand(row_min(c), 0);
}

// Unsigned minimum over each 16-lane row, read from lane 0.
__device__ __forceinline__ uint32_t row0_umin32(uint32_t v) {
  v = min(v, dpp_umin_src<kDppXor1, 0xF>(v));
  v = min(v, dpp_umin_src<kDppXor2, 0xF>(v));
  v = min(v, dpp_umin_src<kDppRor4, 0xF>(v));
  v = min(v, dpp_umin_src<kDppRor8, 0xF>(v));
  return read_lane(v, 0);
}

// fp64 keys over lanes 0..15, as the fp64 wave_min.
__device__ __forceinline__ Cand<double> row0_min(Cand<double> c) {
  uint64_t v = c.w;
  v = min(v, dpp_u64<kDppXor1, 0xF>(v));
  v = min(v, dpp_u64<kDppXor2, 0xF>(v));
  v = min(v, dpp_u64<kDppRor4, 0xF>(v));
  v = min(v, dpp_u64<kDppRor8, 0xF>(v));
  const uint64_t wm = read_lane64(v, 0);
  const uint64_t tied = __builtin_amdgcn_ballot_w64(c.w == wm) & 0xFFFFull;
  Cand<double> r;
  r.w = wm;
  if (__builtin_popcountll(tied) == 1)
    r.j = read_lane(c.j, (int)__builtin_ctzll(tied));
  else
    r.j = row0_umin32(c.w == wm ? c.j : 0xFFFFFFFFu);
  return r;
}

// fp32 keys over lanes 0..15 in two 32-bit passes, as wave_min.
__device__ __forceinline__ Cand<float> row0_min(Cand<float> c) {
  const uint32_t w = (uint32_t)(c.k >> 32);
  const uint32_t wm = row0_umin32(w);
  const uint32_t jm = row0_umin32(w == wm ? (uint32_t)c.k : 0xFFFFFFFFu);
  Cand<float> r;
  r.k = ((uint64_t)wm << 32) | jm;
  return r;
}

// Unit roundoff of Real.
template <typename Real> struct Unit;
template <> struct Unit<float> { static constexpr float u = 5.9604644775390625e-08f; };  // 2^-24
template <> struct Unit<double> { static constexpr double u = 1.1102230246251565e-16; };  // 2^-53

template <typename Real> struct BitsOfReal;
template <> struct BitsOfReal<float> {
  __device__ static uint64_t get(float v) { return __float_as_uint(v); }
  __device__ static float put(uint64_t u) { return __uint_as_float((uint32_t)u); }
};
template <> struct BitsOfReal<double> {
  __device__ static uint64_t get(double v) { return (uint64_t)__double_as_longlong(v); }
  __device__ static double put(uint64_t u) { return __longlong_as_double((long long)u); }
};
// Bit pattern of a non-negative weight (or +inf): unsigned order of the bits
// is the order of the values, so max and < are single integer operations.
template <typename Real> struct WBits;
template <> struct WBits<float> {
  using U = uint32_t;
  __device__ static U of(float v) { return __float_as_uint(v); }
  __device__ static bool finite(U u) { return (u & 0x7FFFFFFFu) < 0x7F800000u; }
  __device__ static Cand<float> key(U w, int j) {
    Cand<float> c;
    c.k = ((uint64_t)w << 32) | (uint32_t)j;
    return c;
  }
};
template <> struct WBits<double> {
  using U = uint64_t;
  __device__ static U of(double v) { return (uint64_t)__double_as_longlong(v); }
  __device__ static bool finite(U u) { return (u & 0x7FFFFFFFFFFFFFFFull) < 0x7FF0000000000000ull; }
  __device__ static Cand<double> key(U w, int j) {
    Cand<double> c;
    c.w = w;
    c.j = (uint32_t)j;
    return c;
  }
};

// Largest squared norm for which no squared distance of the window can
// overflow (sum of squared differences <= 2 (|x|^2 + |y|^2) <= 4 * limit).
template <typename Real> struct GramLimit;
template <> struct GramLimit<float> { static constexpr float value = 1.2676506002282294e30f; };  // 2^100
template <> struct GramLimit<double> { static constexpr double value = 8.6e270; };

// An upper bound of b*b (exact real square): the product rounded, then one
// ulp up.  For tot >= square_bound(b), sqrt(tot) >= b exactly, hence the
// correctly rounded sqrt(tot) >= b.
__device__ __forceinline__ float square_bound(float b) {
  const float p = b * b;
  return isfinite(p) ? __uint_as_float(__float_as_uint(p) + 1u) : p;
}
__device__ __forceinline__ double square_bound(double b) {
  const double p = b * b;
  return isfinite(p) ? __longlong_as_double(__double_as_longlong(p) + 1ll) : p;
}

// Vertex state moved by the periodic repacking of the points path.
template <typename Real> struct PackRec {
  Real best;
  Real lim;
  int32_t id;
  int32_t src;
};

struct PrimArgs {
  const void* x;      // points [B][n][d] (points path)
  const void* dist;   // distances [B][n][n] (matrix path)
  const void* core;   // per point, or null when unit_core
  Windows win;        // window sizes and offsets (edges and core per point)
  int d;
  int unit_core;
  int prefilter;      // certified Gram prefilter before exact distances
  int repack;         // points path: repack live vertices every `repack` steps (0: never)
  void* pack;         // per point PackRec<Real> scratch for the repacking
  int32_t* edge_u;    // per point: edge i of a window at its first point + i
  int32_t* edge_v;
  void* edge_w;       // Real
  int32_t* wflag;     // [B] per-window error bits
  int32_t* status;    // OR of all error bits
};

// MODE 0: points, coordinates of owned vertices in registers (DK padded dims,
// zero padding adds +0 exactly).  The current vertex comes from the window's
// points staged in LDS (XLDS) or, when they do not fit, from the wave that
// found it: each wave's argmin lane writes its vertex's coordinates to the
// wave's LDS slot before the step's barrier, and after the barrier every
// thread reads the slot of the winner's wave (two slot sets by step parity:
// a set is rewritten one step after its last read, behind a barrier).
// Repacking (points path): every `repack` steps the unvisited vertices are
// renumbered over the workgroup's (slot, thread) positions, vertices that
// can still improve (best > own core) first and the others (best == own
// core: no weight can be below it) from the top, so that waves whose
// vertices can no longer improve skip the distance work as a whole.  Keys
// carry the original vertex ids, so the result does not depend on where a
// vertex sits.
// MODE 1: matrix rows from device memory.  MODE 2: the points path with the
// certified FMA Gram prefilter (a separate instantiation keeps the plain
// path free of its branches).
template <typename Real, int VPT, int DK, int MODE, bool XLDS>
__device__ __forceinline__ void prim_window(const PrimArgs& a, int item) {
  using C = Cand<Real>;
  constexpr bool PTS = MODE != 1;  // points path (MODE 0, or 2 with the Gram prefilter)
  int64_t base;
  int n, b;
  a.win.get(item, b, base, n);
  const int tid = threadIdx.x, T = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, W = T >> 6;
  const int d = a.d;
  const bool unit = a.unit_core != 0;
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ C red[2][16];
  __shared__ unsigned long long kmin[3];
  constexpr bool WSLOT = PTS && !XLDS;
  __shared__ Real wslot[WSLOT ? 2 * 16 * (DK > 0 ? DK : 1) : 1];
  __shared__ unsigned long long sh_nmax;
  __shared__ int sh_err;
  __shared__ int sh_pack[2][2];  // [parity] live count, count from the top
  constexpr int DKS = DK > 0 ? DK : 1;
  Real* xs = reinterpret_cast<Real*>(lds);
  Real* cs = reinterpret_cast<Real*>(lds + (XLDS ? (size_t)n * DK * sizeof(Real) : 0));
  const Real* xg = static_cast<const Real*>(a.x) + (size_t)base * (PTS ? d : 0);
  const Real* dg = static_cast<const Real*>(a.dist) + (MODE == 1 ? a.win.matrix(b, base) : 0);
  const Real* cg = static_cast<const Real*>(a.core) + (size_t)base;
  if (tid == 0) {
    sh_err = 0;
    sh_nmax = 0;
    kmin[0] = kmin[1] = kmin[2] = ~0ull;
    sh_pack[0][0] = sh_pack[0][1] = sh_pack[1][0] = sh_pack[1][1] = 0;
  }
  __syncthreads();
  int err = 0;
  if constexpr (PTS && XLDS) {
    for (int idx = tid; idx < n * DKS; idx += T) {
      const int i = idx / DKS, k = idx - i * DKS;
      const Real v = k < d ? xg[(size_t)i * d + k] : Real(0);
      if (!isfinite(v)) err |= 2;
      xs[idx] = v;
    }
  }
  Real* xn = cs + (unit ? 0 : n);  // squared norms of all points (points path)
  // Points path without LDS points: the wave holding each vertex (the wave
  // that publishes it when it wins a step).
  uint8_t* owner = reinterpret_cast<uint8_t*>(xn + n);
  if (WSLOT)
    for (int j = tid; j < n; j += T) owner[j] = (uint8_t)((j % T) >> 6);
  if (!unit)
    for (int i = tid; i < n; i += T) cs[i] = cg[i];

  Real xr[VPT][DKS];
  Real best[VPT], cr[VPT], lim[VPT], nr[VPT];
  int src[VPT];
  int vid[VPT];  // vertex held by each slot (n: none)
  C key[VPT];    // (best, vertex) of an unvisited slot, or C::none()
  // Liveness floor: the core distance (-inf with unit core distances) of an
  // unvisited slot, +inf once visited or empty.
  Real crx[VPT];
  uint32_t vis = 0;
#pragma unroll
  for (int s = 0; s < VPT; ++s) {
    const int j = tid + s * T;
    vid[s] = j < n ? j : n;
    best[s] = Real(INFINITY);
    lim[s] = Real(INFINITY);
    src[s] = 1;
    cr[s] = 0;
#pragma unroll
    for (int k = 0; k < DKS; ++k) xr[s][k] = 0;
    if (j < n) {
      if (PTS) {
#pragma unroll
        for (int k = 0; k < DKS; ++k) {
          const Real v = k < d ? xg[(size_t)j * d + k] : Real(0);
          if (!XLDS && !isfinite(v)) err |= 2;
          xr[s][k] = v;
        }
      }
      if (!unit) cr[s] = cg[j];
    } else {
      vis |= 1u << s;
    }
    key[s] = j < n && j != 0 ? C::make(Real(INFINITY), j) : C::none();
    crx[s] = j < n && j != 0 ? (unit ? Real(-INFINITY) : cr[s]) : Real(INFINITY);
    Real nn = 0;
#pragma unroll
    for (int k = 0; k < DKS; ++k) nn = fma(xr[s][k], xr[s][k], nn);
    nr[s] = nn;
  }
  if (tid == 0) vis |= 1u;  // vertex 0 starts the tree
  if (err) atomicOr(&sh_err, err);
  // Gram prefilter: norms of every point (for the current vertex) and the
  // window-wide largest norm.  With every squared norm <= 2^100 (fp32) no
  // pair can overflow, so skipping pairs never hides an overflow.
  Real nmax = 0;
  if (PTS) {
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      const int j = tid + s * T;
      if (j < n) {
        xn[j] = nr[s];
        nmax = fmax(nmax, nr[s]);
      }
    }
  }
  // Non-negative IEEE values order like their bit patterns.
  if (PTS) atomicMax(&sh_nmax, (unsigned long long)BitsOfReal<Real>::get(nmax));
  __syncthreads();
  const bool gram = MODE == 2 && BitsOfReal<Real>::put(sh_nmax) <= GramLimit<Real>::value;
  // Bound on |a - r| for a = fl(|x_j|^2 + |x_c|^2 - 2 x_j.x_c) (FMA chains of
  // DK terms) and r the exact real squared distance: 2.03 (DK + 3) u
  // (n_j + n_c) is enough; E uses 2 (DK + 4) u and the test below doubles it.
  const Real gram_k = Real(2 * (DKS + 4)) * Unit<Real>::u;

  const Real rmax = RealTraits<Real>::max_value();
  int cur = 0, par = 0, bad = 0, pp = 0;
  // Countdown to the next repacking, at steps repack, 2 repack, ... (0: never).
  int to_repack = a.repack ? a.repack + 1 : 0;
  uint64_t bad_lanes = 0;  // lanes that formed a non-finite weight (wave-uniform)
  PackRec<Real>* pk = static_cast<PackRec<Real>*>(a.pack) + base;
  for (int step = 0; step < n - 1; ++step) {
    Real xc[DKS];
    if (PTS) {
      if constexpr (XLDS) {
#pragma unroll
        for (int k = 0; k < DKS; ++k) xc[k] = xs[cur * DK + k];
      } else if (step == 0) {
#pragma unroll
        for (int k = 0; k < DKS; ++k) xc[k] = k < d ? xg[k] : Real(0);
      } else {
        const Real* sl = wslot + ((step - 1) & 1) * 16 * DKS + owner[cur] * DKS;
#pragma unroll
        for (int k = 0; k < DKS; ++k) xc[k] = sl[k];
      }
    }
    if (PTS && to_repack && --to_repack == 0) {
      to_repack = a.repack;
#pragma unroll
      for (int s = 0; s < VPT; ++s) {
        if (!((vis >> s) & 1u)) {
          const bool can = best[s] > cr[s];
          const int pos = can ? atomicAdd(&sh_pack[pp][0], 1) : n - 1 - atomicAdd(&sh_pack[pp][1], 1);
          PackRec<Real> r;
          r.best = best[s];
          r.lim = lim[s];
          r.id = vid[s];
          r.src = src[s];
          pk[pos] = r;
        }
      }
      __syncthreads();
      const int live_n = sh_pack[pp][0], top_n = sh_pack[pp][1];
      if (tid == 0) sh_pack[pp ^ 1][0] = sh_pack[pp ^ 1][1] = 0;
      pp ^= 1;
      vis = 0;
#pragma unroll
      for (int s = 0; s < VPT; ++s) {
        const int p = tid + s * T;
        if (p < live_n || (p >= n - top_n && p < n)) {
          const PackRec<Real> r = pk[p];
          const int j = r.id;
          vid[s] = j;
          best[s] = r.best;
          lim[s] = r.lim;
          src[s] = r.src;
          cr[s] = unit ? Real(0) : cs[j];
#pragma unroll
          for (int k = 0; k < DKS; ++k)
            xr[s][k] = XLDS ? xs[j * DK + k] : (k < d ? xg[(size_t)j * d + k] : Real(0));
          nr[s] = gram ? xn[j] : Real(0);
          if (WSLOT) owner[j] = (uint8_t)wave;
          key[s] = C::make(r.best, j);
          crx[s] = unit ? Real(-INFINITY) : cr[s];
        } else {
          vid[s] = n;
          vis |= 1u << s;
          key[s] = C::none();
          crx[s] = Real(INFINITY);
        }
      }
    }
    const Real ccur = unit ? Real(0) : cs[cur];
    const Real ccx = unit ? Real(-INFINITY) : ccur;
    const Real ncur = gram ? xn[cur] : Real(0);
    C c = C::none();
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      // Only a weight strictly below best can change this vertex.  With core
      // distances, max(core_cur, core_j) >= best rules it out before any
      // distance is formed (w = max(core_cur, core_j, d) >= best).
      // As one comparison against the slot's floor (+inf once visited, -inf
      // with unit core distances, which keeps every unvisited vertex live).
      // Core distances are never NaN here (non-finite input is rejected).
      const Real lv = ccx > crx[s] ? ccx : crx[s];
      const bool live = lv < best[s];
      const uint64_t live_lanes = __builtin_amdgcn_ballot_w64(lv < best[s]);
      bool need = live;
      if (gram && __ballot(live)) {
        // Certified skip: a - 2E >= lim (1 + 2^-16) implies the exact
        // coordinate-order sum of squares is >= lim, hence no update.
        Real g = 0;
#pragma unroll
        for (int k = 0; k < DKS; ++k) g = fma(xr[s][k], xc[k], g);
        const Real sn = nr[s] + ncur;
        const Real approx = sn - Real(2) * g;
        const Real err = gram_k * sn;
        need = live && !(approx - Real(2) * err >= lim[s] * Real(1.0000152587890625));
      }
      if (MODE == 2 ? __ballot(need) != 0 : live_lanes != 0) {
        const int j = vid[s];
        if (PTS) {
          Real tot = 0;
#pragma unroll
          for (int k = 0; k < DKS; ++k) {
            const Real dl = xr[s][k] - xc[k];
            tot = tot + dl * dl;
          }
          bad_lanes |= __builtin_amdgcn_ballot_w64(!isfinite(tot)) & live_lanes;
          // tot >= lim[s] (an upper bound of best^2) implies sqrt(tot) >= best:
          // no update, so the correctly rounded square root is skipped.
          const bool near = tot < (live ? lim[s] : Real(-INFINITY));
          if (__ballot(near)) {
            const Real dd = sqrt(tot);
            const Real w = unit ? dd : fmax(fmax(ccur, cr[s]), dd);
            if (near && w < best[s]) {
              best[s] = w;
              src[s] = cur;
              lim[s] = square_bound(w);
              key[s] = C::make(w, j);
            }
          }
        } else {
          const Real dd = live ? dg[(size_t)cur * n + j] : Real(0);
          bad_lanes |= __builtin_amdgcn_ballot_w64(!isfinite(dd)) & live_lanes;
          const Real w = unit ? dd : fmax(fmax(ccur, cr[s]), dd);
          if (live && w < best[s]) {
            best[s] = w;
            src[s] = cur;
            key[s] = C::make(w, j);
          }
        }
      }
      // Visited slots hold C::none(), above every candidate.
      c = s == 0 ? key[0] : cand_pick(key[s].less(c), key[s], c);
    }
    c = wave_min(c);
    {
      // The start value (REAL_MAX, 0) of the reference's argmin, once per
      // wave on its minimum instead of once per lane.
      const C start = C::make(rmax, 0);
      c = cand_pick(c.less(start), c, start);
    }
    if constexpr (WSLOT) {
      // The wave's argmin vertex, if this lane owns it, publishes its
      // coordinates (a wave with every vertex visited publishes nothing).
      const int v = c.index();
      Real* sl = wslot + (step & 1) * 16 * DKS + wave * DKS;
#pragma unroll
      for (int s = 0; s < VPT; ++s)
        if (vid[s] == v && !((vis >> s) & 1u)) {
#pragma unroll
          for (int k = 0; k < DKS; ++k) sl[k] = xr[s][k];
        }
    }
    C g;
    if constexpr (sizeof(Real) == 4) {
      // One 64-bit LDS atomic min per wave; slot step%3 is reset two steps
      // ahead (its last reads happened before the previous barrier, its
      // next atomics come after the next one), so one barrier per step.
      if (lane == 0) atomicMin(&kmin[par], c.k);
      __syncthreads();
      g.k = kmin[par];
      if (tid == 0) kmin[par == 0 ? 2 : par - 1] = ~0ull;
      par = par == 2 ? 0 : par + 1;
    } else {
      if (lane == 0) red[par & 1][wave] = c;
      __syncthreads();
      // Cross-wave minimum: lane w holds wave w's key, one more reduction.
      g = lane < W ? red[par & 1][lane] : C::make(rmax, 0x7fffffff);
      g = wave_min(g);
      par ^= 1;
    }
    const int nxt = g.index();
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      if (nxt == vid[s]) {
        // A visited winner means no vertex had a weight below REAL_MAX: the
        // graph is degenerate (only possible with non-finite weights).
        if ((vis >> s) & 1u) bad |= 1;
        vis |= 1u << s;
        key[s] = C::none();
        crx[s] = Real(INFINITY);
        const Real gw = g.weight();
        const size_t o = (size_t)base + step;
        a.edge_u[o] = gw < rmax ? src[s] : 0;
        a.edge_v[o] = nxt;
        static_cast<Real*>(a.edge_w)[o] = gw;
      }
    }
    cur = nxt;
  }
  if (bad || (bad_lanes && lane == 0)) atomicOr(&sh_err, 1);
  __syncthreads();
  if (tid == 0) {
    a.wflag[b] = sh_err;
    if (sh_err) atomicOr(a.status, sh_err);
  }
}

// One window per workgroup, or windows claimed from the launch's queue.
template <typename Real, int VPT, int DK, int MODE, bool XLDS>
__global__ __launch_bounds__(VPT == 8 ? 512 : 1024) void prim_kernel(PrimArgs a) {
  for_each_item(a.win, [&](int item) { prim_window<Real, VPT, DK, MODE, XLDS>(a, item); });
}

template <typename Real>
__device__ __forceinline__ bool cand_eq(const Cand<Real>& x, const Cand<Real>& y) {
  return !x.less(y) && !y.less(x);
}

// Two smallest candidates seen so far (first < second).  Value selects only:
// a branchy swap becomes a store through a selected address, which the
// compiler places in scratch memory.
template <typename Real>
__device__ __forceinline__ void cand_top2(Cand<Real>& first, Cand<Real>& second, const Cand<Real>& x) {
  const bool lt1 = x.less(first);
  const bool lt2 = x.less(second);
  second = cand_pick(lt1, first, cand_pick(lt2, x, second));
  first = cand_pick(lt1, x, first);
}

// Speculative Prim over matrix rows: up to two vertices per round.
//
// A step of the matrix path waits on one row read from device memory.  The
// vertex that wins the step after the current one is usually the runner-up
// of the previous step (the current vertex rarely lowers a key below it), so
// each round reads the row of the current vertex c AND the row of the
// predicted next vertex p together:
//   K'  = keys after c's row (the true keys of this step),
//   K'' = K' after p's row, over unvisited vertices other than p.
// One barrier exchanges per wave min(K'), and the two smallest of K''.  If
// the exact argmin of K' is p, the prediction held: p is added with its K'
// key, K'' are then the true keys, and argmin K'' is added as well -- the
// same two steps, edges and tie-breaks as one vertex per step.  Otherwise the
// argmin of K' is added and K' kept.  The second smallest key of the round's
// final keys predicts the next round (after a miss the next round reads one
// row and yields the prediction).
template <typename Real, int VPT>
__device__ __forceinline__ void prim_spec_window(const PrimArgs& a, int item) {
  using C = Cand<Real>;
  int64_t base;
  int n, b;
  a.win.get(item, b, base, n);
  const int tid = threadIdx.x, T = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, W = T >> 6;
  const bool unit = a.unit_core != 0;
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ C red[2][3][16];  // [round parity][min K', min K'' (or K'), second][wave]
  __shared__ int sh_err;
  Real* cs = reinterpret_cast<Real*>(lds);
  const Real* dg = static_cast<const Real*>(a.dist) + a.win.matrix(b, base);
  const Real* cg = static_cast<const Real*>(a.core) + (size_t)base;
  if (tid == 0) sh_err = 0;
  if (!unit)
    for (int i = tid; i < n; i += T) cs[i] = cg[i];
  // Keys, cores and weights as bit patterns (WBits): no divergent branch in
  // the round, both rows read whole with clamped indices.
  using WB = WBits<Real>;
  using U = typename WB::U;
  U bb[VPT], crb[VPT];
  int src[VPT], jc[VPT];
  uint32_t vis = 0;
#pragma unroll
  for (int s = 0; s < VPT; ++s) {
    const int j = tid + s * T;
    bb[s] = WB::of(Real(INFINITY));
    src[s] = 1;
    jc[s] = min(j, n - 1);
    crb[s] = WB::of(j < n && !unit ? cg[j] : Real(0));
    if (j >= n) vis |= 1u << s;
  }
  if (tid == 0) vis |= 1u;  // vertex 0 starts the tree
  __syncthreads();

  const Real rmax = RealTraits<Real>::max_value();
  const C none = C::make(rmax, 0);
  int cur = 0, pr = -1, par = 0, bad = 0;
  for (int step = 0; step < n - 1;) {
    cur = __builtin_amdgcn_readfirstlane(cur);  // uniform: row addresses in scalar registers
    pr = __builtin_amdgcn_readfirstlane(pr);
    // Speculate only with two or more unvisited vertices.
    const bool spec = pr >= 0 && n - 1 - step >= 2;
    const U ccur = unit ? U(0) : WB::of(cs[cur]);
    const U cpr = spec && !unit ? WB::of(cs[pr]) : U(0);
    const Real* row1 = dg + (size_t)cur * n;
    const Real* row2 = dg + (size_t)(spec ? pr : cur) * n;
    U d1[VPT], d2[VPT];
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      d1[s] = WB::of(row1[jc[s]]);
      d2[s] = spec ? WB::of(row2[jc[s]]) : U(0);
    }
    U b1[VPT], b2[VPT];
    int s1[VPT], s2[VPT];
    C k1a = none, k1b = none, k2a = none, k2b = none;
    int bad2 = 0;
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      const int j = tid + s * T;
      const bool un = !((vis >> s) & 1u);
      // Only a weight strictly below the key can change it:
      // max(core_c, core_j) < key is needed before the entry matters.
      const U m1 = max(ccur, crb[s]);
      const bool l1 = un && m1 < bb[s];
      bad |= l1 && !WB::finite(d1[s]);
      const U w1 = max(m1, d1[s]);
      const bool lt1 = l1 && w1 < bb[s];
      b1[s] = lt1 ? w1 : bb[s];
      s1[s] = lt1 ? cur : src[s];
      // K': the two smallest predict the next round of a plain round; a
      // speculative round needs only the smallest.
      const C o1 = un ? WB::key(b1[s], j) : C::none();
      if (spec) {
        k1a = cand_pick(o1.less(k1a), o1, k1a);
        const bool e2 = un && j != pr;
        const U m2 = max(cpr, crb[s]);
        const bool l2 = e2 && m2 < b1[s];
        bad2 |= l2 && !WB::finite(d2[s]);
        const U w2 = max(m2, d2[s]);
        const bool lt2 = l2 && w2 < b1[s];
        b2[s] = lt2 ? w2 : b1[s];
        s2[s] = lt2 ? pr : s1[s];
        cand_top2(k2a, k2b, e2 ? WB::key(b2[s], j) : C::none());
      } else {
        b2[s] = b1[s];
        s2[s] = s1[s];
        cand_top2(k1a, k1b, o1);
      }
    }
    // Per wave: the list that predicts (K'' when speculating, else K') to
    // its two smallest, and min K'.
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
    // Across waves (a set is rewritten two rounds later, behind a barrier).
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
    const bool hit = spec && gc.index() == pr;
    // First vertex of the round: p on a hit, else argmin K'.
    const C first = cand_pick(spec, gc, g1);
    const int v1 = first.index();
    const int v2 = hit ? g1.index() : -1;
#pragma unroll
    for (int s = 0; s < VPT; ++s) {
      const int j = tid + s * T;
      if (j == v1) {
        // A visited winner means no vertex had a weight below REAL_MAX: the
        // graph is degenerate (only possible with non-finite weights).
        if ((vis >> s) & 1u) bad |= 1;
        vis |= 1u << s;
        const Real gw = first.weight();
        const size_t o = (size_t)base + step;
        a.edge_u[o] = gw < rmax ? s1[s] : 0;
        a.edge_v[o] = v1;
        static_cast<Real*>(a.edge_w)[o] = gw;
      }
      if (j == v2) {
        if ((vis >> s) & 1u) bad |= 1;
        vis |= 1u << s;
        const Real gw = g1.weight();
        const size_t o = (size_t)base + step + 1;
        a.edge_u[o] = gw < rmax ? s2[s] : 0;
        a.edge_v[o] = v2;
        static_cast<Real*>(a.edge_w)[o] = gw;
      }
      bb[s] = hit ? b2[s] : b1[s];
      src[s] = hit ? s2[s] : s1[s];
    }
    if (hit) bad |= bad2;
    step += hit ? 2 : 1;
    cur = hit ? v2 : v1;
    // The next round's prediction: the runner-up of the keys just kept.
    pr = !spec && g2.less(none) ? g2.index() : hit && g2.less(none) ? g2.index() : -1;
  }
  if (bad) atomicOr(&sh_err, 1);
  __syncthreads();
  if (tid == 0) {
    a.wflag[b] = sh_err;
    if (sh_err) atomicOr(a.status, sh_err);
  }
}

template <typename Real, int VPT>
__global__ __launch_bounds__(VPT == 8 ? 512 : 1024) void prim_spec_kernel(PrimArgs a) {
  for_each_item(a.win, [&](int item) { prim_spec_window<Real, VPT>(a, item); });
}

}  // namespace hdb
