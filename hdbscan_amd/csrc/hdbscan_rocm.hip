// C ABI of the gfx950 batched HDBSCAN hot path (see include/hdbscan_rocm.h).
//
// Pipeline per call, all on the caller's stream:
//   [core]  gram_prep_kernel + core_gram_kernel + core_exact_kernel (the
//           certified MFMA Gram filter: fp32, min_samples 3..16, points
//           path), core_topk_kernel (exact top-k from points), or
//           core_select_list_kernel / core_select_kernel (selection from a
//           materialised matrix)
//   [prim]  prim_kernel: points path (distances formed on the fly from
//           register/LDS-resident coordinates), or distance_kernel and then
//           prim_spec_kernel reading matrix rows (two per round)
//   [tree]  tree_kernel: sort, linkage, condensation, selection, labels
// Workspace (caller-owned device memory, hdb_workspace_bytes):
//   edge_u[B][n] i32 | edge_v[B][n] i32 | edge_w[B][n] Real | wflag[B] i32 |
//   core[B][n] Real | tree arena[B][stride] | cluster records[B][n] |
//   distances[B][n][n] Real (matrix path) | Gram candidates, operands and
//   norm terms (filter) | Prim repacking records (points path)
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_set>

#include "../../include/hdbscan_rocm.h"
#include "hdb_common.h"
#include "hdb_gram.h"
#include "hdb_gramexact.h"
#include "hdb_gramgraph.h"
#include "hdb_kernels.h"
#include "hdb_prim.h"

using namespace hdb;

namespace {

constexpr int kTreeThreads = 1024;

thread_local std::string g_error;

// Optional per-stage HIP event timers of the calling thread (hdb_profile).
struct StageTimers {
  bool on = false;
  int count = 0;
  hipEvent_t events[16] = {};
  const char* names[16] = {};
};
thread_local StageTimers g_timers;

void stage_mark(hipStream_t s, const char* name) {
  StageTimers& t = g_timers;
  if (!t.on || t.count >= 16) return;
  if (!t.events[t.count]) hipEventCreate(&t.events[t.count]);
  hipEventRecord(t.events[t.count], s);
  t.names[t.count] = name;
  ++t.count;
}

int fail(const std::string& msg) {
  g_error = msg;
  return -1;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

// LDS bytes one workgroup may allocate on the current device (160 KiB on
// gfx950); falls back to that figure when no device is visible (sizing only).
uint32_t device_lds_limit() {
  static std::mutex mu;
  static int cached_device = -1;
  static uint32_t cached = kLdsLimit;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return kLdsLimit;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (dev != cached_device) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0) {
      cached = (uint32_t)v < kLdsLimit ? (uint32_t)v : kLdsLimit;
    } else {
      (void)hipGetLastError();
      cached = 64u * 1024u;
    }
    cached_device = dev;
  }
  return cached;
}

// Raises a kernel's dynamic LDS limit to the bytes a launch needs.  A
// rejected request is cleared (the launch itself then reports any error).
int allow_lds(const void* fn, uint32_t bytes) {
  static std::mutex mu;
  static std::unordered_set<uint64_t> done;
  if (bytes <= 64u * 1024u) return 0;
  const uint64_t key = ((uint64_t)(uintptr_t)fn << 20) ^ bytes;
  std::lock_guard<std::mutex> lock(mu);
  if (done.count(key)) return 0;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
    (void)hipGetLastError();
  done.insert(key);
  return 0;
}

struct Plan {
  int rs = 4;             // sizeof(Real)
  bool matrix = false;    // Prim reads a materialised matrix
  int vpt = 1, dk = 0, threads = 64;
  bool xlds = false;      // points path: all points in LDS
  uint32_t prim_lds = 0;
  bool core = false;      // core distances needed (min_samples > 2)
  bool core_topk = false; // core from points with a register list
  int kmax = 0, core_dk = 0;
  bool core_gram = false; // core through the certified MFMA Gram filter
  bool gram_window = false;  // ... in its per-window form (core_window_kernel)
  int gram_dk = 0, gram_rs = 0, gram_threads = 0;
  uint32_t gram_lds = 0;
  bool tree_lds = true;
  uint32_t tree_lds_bytes = 0;
  size_t arena_stride = 0;
  size_t slot_stride = 0;
  bool need_matrix = false;
  int flags = 0;
  size_t off_eu = 0, off_ev = 0, off_ew = 0, off_flag = 0, off_core = 0, off_arena = 0, off_slots = 0, off_matrix = 0;
  size_t off_cand = 0, off_ncand = 0, off_pa = 0, off_pb = 0, off_nb = 0, off_pack = 0, total = 0;
  int repack = 0;  // Prim repacking interval (points path)
  bool gram_graph = false;  // matrix path through the MFMA Gram matrix (hdb_gramgraph.h)
  bool gram_exact = false;  // matrix path with MFMA-bounded near/far entries (hdb_gramexact.h)
  int gx_dk = 0, gx_cols = 0, gx_rank = 0;
  size_t off_gop = 0, off_gnorm = 0, off_gu2 = 0, off_gcnt = 0, off_gval = 0, off_gncnt = 0, off_gnlist = 0;
  size_t off_alt = 0;
  bool split_tree = false;  // tree stage as sort + hierarchy kernels (sorted-edge transport)
  int sort_threads = 0, hier_threads = 0;
  size_t off_suv = 0, off_sw = 0;
};

// Workgroup widths of the split tree stage for windows of up to n points:
// the sort takes about one lane per two edges, the hierarchy the least the
// condensation allows (n - 1 <= 4 T), so that small windows share a CU.
int split_sort_threads(int64_t n) { return (int)std::min<int64_t>(1024, std::max<int64_t>(64, (n / 2 + 63) / 64 * 64)); }
int split_hier_threads(int64_t n) {
  return (int)std::min<int64_t>(1024, std::max<int64_t>(64, ((n - 1 + 3) / 4 + 63) / 64 * 64));
}

// One launch of the pipeline: the kernels' window descriptor, the grid's
// window count, the points covered and the largest window.
struct Launch {
  Windows win;
  int64_t count = 0;   // items (windows) of this launch
  int64_t points = 0;  // points of the whole batch (per-point buffers)
  int64_t nmax = 0;    // largest window of this launch
  int* prim_queue = nullptr;  // claim counters of persistent Prim / tree launches (ragged)
  int* tree_queue = nullptr;
  int64_t persist = 0;        // workgroups of a persistent launch
};

// Compute units of the current device (256 on MI355X when none is visible).
int device_cus() {
  static std::mutex mu;
  static int cached_device = -1, cached = 256;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 256;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (dev != cached_device) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cached = v;
    else (void)hipGetLastError();
    cached_device = dev;
  }
  return cached;
}

Launch dense_launch(int64_t batch, int64_t n) {
  Launch L;
  L.win.off = nullptr;
  L.win.list = nullptr;
  L.win.moff = nullptr;
  L.win.queue = nullptr;
  L.win.n = (int)n;
  L.win.count = (int)batch;
  L.count = batch;
  L.points = batch * n;
  L.nmax = n;
  return L;
}

int pad_dims(int64_t d) {
  if (d <= 8) return 8;
  if (d <= 16) return 16;
  if (d <= 32) return 32;
  if (d <= 64) return 64;
  return 0;
}

// Row sets per wave of the Gram core kernel: 2 or 4 up to 32 dimensions,
// 1 or 2 up to 128, 1 beyond (register budget).  Flag bits 12..13 pick one
// (1 -> 1, 2 -> 2, 3 -> 4); 0 chooses by window size (below).
int gram_rsets(int dk, int64_t n, int flags) {
  const int want = (flags >> 12) & 3;
  // Default: two 16-row sets per wave; four for windows of 1025..2048 points
  // (each column chunk is read by half as many row blocks; measured on
  // ragged batches: 12.66 -> 11.89 ms per call, while four sets for the
  // windows above 2048 points, the batch's critical path, cost 5 %).
  const int rs = want == 1 ? 1 : want == 3 ? 4 : want == 2 ? 2 : (n > 1024 && n <= 2048) ? 4 : 2;
  if (dk <= 32) return rs == 1 ? 2 : rs;
  if (dk <= 128) return rs == 4 ? 2 : rs;
  return 1;
}

template <int DK>
uint32_t gram_bytes_rs(int rs, int64_t n) {
  if (rs == 4) return gram_lds_bytes<DK, 4>((int)n);
  if (rs == 2) return gram_lds_bytes<DK, 2>((int)n);
  return gram_lds_bytes<DK, 1>((int)n);
}

uint32_t gram_bytes(int dk, int rs, int64_t n) {
  switch (dk) {
    case 8: return gram_bytes_rs<8>(rs, n);
    case 16: return gram_bytes_rs<16>(rs, n);
    case 32: return gram_bytes_rs<32>(rs, n);
    case 64: return gram_bytes_rs<64>(rs, n);
    case 128: return gram_bytes_rs<128>(rs, n);
    default: return gram_bytes_rs<256>(rs, n);
  }
}

// Gram-filtered core distances: float32, 3 <= min_samples <= 16, at most
// 256 dimensions.  Waves per workgroup: as many as the padding allows, fewer
// for small windows.
void set_gram(Plan& P, int64_t n, int64_t d, int32_t ms, uint32_t lds_limit) {
  P.core_gram = false;
  if (!P.core || P.rs != 4 || ms < 3 || ms > 16 || d > 256 || (P.flags & HDB_FLAG_PLAIN)) return;
  // With a materialised exact matrix the core distances are one selection
  // away; the filter then only pays when asked for.
  if (P.matrix && !(P.flags & HDB_FLAG_FILTER)) return;
  const int dk = d <= 8 ? 8 : d <= 16 ? 16 : d <= 32 ? 32 : d <= 64 ? 64 : d <= 128 ? 128 : 256;
  const int rs = gram_rsets(dk, n, P.flags);
  int waves = (dk >= 64 ? 256 : 512) / 64;
  const int wr = 16 * rs;  // rows per wave
  while (waves > 1 && (int64_t)(waves / 2) * wr >= n) waves /= 2;
  if (gram_bytes(dk, rs, n) > lds_limit - 1024u) return;
  P.core_gram = true;
  // Windows of up to 1024 points and 16 dimensions: one workgroup per window
  // with the column operands in LDS once, no prep kernel.
  P.gram_window = dk <= kWinDK && n <= kWinMaxN && !P.matrix && !(P.flags & HDB_FLAG_GRAM_CHUNKED) &&
                  gram_window_lds_bytes((int)n, (int)d) <= lds_limit;
  P.gram_dk = dk;
  P.gram_rs = rs;
  P.gram_threads = waves * 64;
  P.gram_lds = gram_bytes(dk, rs, n);
}

// Bytes of one MFMA operand array of the Gram matrix (gram_operand_kernel).
size_t gram_operand_bytes(int dk, int64_t batch, int64_t n) {
  const int nq = (3 * dk + 31) / 32;
  return (size_t)batch * ((n + 15) / 16) * nq * 64 * 16;
}


int make_plan(int64_t batch, int64_t n, int64_t d, int dtype, int32_t ms, int flags, Plan* p) {
  if (dtype != HDB_F32 && dtype != HDB_F64) return fail("dtype must be HDB_F32 or HDB_F64");
  if (batch < 1 || d < 1 || n < 2) return fail("X requires nonempty batches/features and at least 2 samples");
  if (n > HDB_MAX_SAMPLES) return fail("at most 4096 samples per window are supported");
  // hdb_cluster_matrix sizes with min_samples 3 to reserve core storage.
  if (ms < 1 || (ms > n && !(flags & HDB_FLAG_MATRIX_INPUT))) return fail("min_samples must be in [1, n]");
  if (d > (1 << 24) || batch > (1 << 24)) return fail("dimension or batch too large");
  Plan P;
  const uint32_t lds_limit = device_lds_limit();
  P.flags = flags;
  P.rs = dtype == HDB_F32 ? 4 : 8;
  const int rs = P.rs;
  P.core = ms > 2;
  P.core_topk = P.core && ms <= 64;
  P.kmax = ms <= 8 ? 8 : ms <= 16 ? 16 : ms <= 32 ? 32 : 64;
  P.core_dk = pad_dims(d);
  // Points path: owned coordinates must fit 64 VGPRs per thread.
  // Vertices per Prim thread: about 512 threads per window measured best at
  // N=1024 without core distances (fewer waves to synchronise per step, two
  // slots of ILP), 1024 threads with core distances and at most 16 fp32
  // coordinates per point (1.5 % faster at N=1024, D=16); larger windows
  // use 1024 threads (four slots per thread past 2048 points).
  const int dk = pad_dims(d);
  P.dk = dk;
  const int want_vpt = (flags >> 8) & 15;
  if (want_vpt != 0 && want_vpt != 1 && want_vpt != 2 && want_vpt != 4 && want_vpt != 8)
    return fail("vertices per thread must be 1, 2, 4 or 8");
  int vpt = want_vpt == 8 ? 0 : want_vpt;
  if (!vpt) {
    vpt = n <= 512 ? 1 : n <= 1024 ? 2 : 4;
    // Points path up to 1024 with at most 64 bytes of coordinates: one slot
    // per thread and 1024 threads beat two slots and 512 threads (1.5 % at
    // N=1024, D=16 with core distances; 2.4 % at D=8 without).
    if (vpt == 2 && dk > 0 && (int64_t)dk * rs <= 64) vpt = 1;
    // Points path up to 2048: two slots per thread and 1024 threads beat
    // four slots and 512 threads (20 % at N=2048, D=16).  The matrix path
    // too (speculative Prim 3.12 -> 2.71 ms at N=2048, D=128).
    if (vpt == 4 && n <= 2048 && (dk == 0 || (int64_t)2 * dk * rs <= 64 * 4)) vpt = 2;
    while (vpt > 1 && (int64_t)vpt * dk * rs > 64 * 4 && (int64_t)(vpt / 2) * 1024 >= n) vpt /= 2;
  }
  while ((int64_t)vpt * 1024 < n) vpt *= 2;
  if (vpt > 4) return fail("window too large for the vertex layout");
  // Points path above 2048 points, at most 16 float32 coordinates: eight
  // slots on 512 threads.  Four slots' coordinates and state on 1024 threads
  // exceed the 128 registers such a workgroup allows, and the compiler
  // spills them to scratch inside the step loop; 512 threads allow 256.
  if (vpt == 4 && want_vpt != 4 && rs == 4 && dk > 0 && dk <= 16 && !(flags & HDB_FLAG_MATRIX)) vpt = 8;
  P.vpt = vpt;
  P.threads = (int)(((n + vpt - 1) / vpt + 63) / 64 * 64);
  P.matrix = (flags & HDB_FLAG_MATRIX) || dk == 0 || (int64_t)vpt * dk * rs > (vpt == 8 ? 64 * 8 : 64 * 4);
  // The speculative matrix-row Prim repeats its reductions in every wave:
  // 8 vertices per thread halves the waves of a window (no coordinates are
  // held in registers on this path).
  if (want_vpt == 8) {
    if (!P.matrix || (flags & HDB_FLAG_ONE_STEP))
      return fail("8 vertices per thread need the speculative matrix path");
    P.vpt = vpt = 8;
    P.threads = (int)(((n + vpt - 1) / vpt + 63) / 64 * 64);
  }
  // The MFMA Gram form of the graph stage (float32, matrix path) when asked
  // for: measured slower than the exact matrix at N=2048, D=128 (the
  // interval-key Prim's rounds cost more than the exact matrix saves).
  P.gram_graph = P.matrix && rs == 4 && d <= 256 && ms <= 64 && (flags & HDB_FLAG_GRAM_GRAPH) &&
                 !(flags & (HDB_FLAG_PLAIN | HDB_FLAG_MATRIX_INPUT | HDB_FLAG_ONE_STEP)) && want_vpt != 8;
  // The MFMA-bounded matrix (hdb_gramexact.h): float32 matrix path, core
  // distances for min_samples <= 16 (or none), up to 256 dimensions.
  P.gram_exact = P.matrix && rs == 4 && !P.gram_graph && d <= 256 && ms <= 16 && want_vpt != 8 &&
                 !(flags & (HDB_FLAG_PLAIN | HDB_FLAG_MATRIX_INPUT | HDB_FLAG_ONE_STEP | HDB_FLAG_NO_GRAM_EXACT));
  if (P.gram_graph) {
    P.gram_dk = d <= 8 ? 8 : d <= 16 ? 16 : d <= 32 ? 32 : d <= 64 ? 64 : d <= 128 ? 128 : 256;
    if (P.vpt > 4) return fail("window too large for the vertex layout");
  } else if (P.gram_exact) {
    P.gx_dk = d <= 32 ? 32 : d <= 64 ? 64 : d <= 128 ? 128 : 256;
    P.gx_cols = (int)n;
    P.gx_rank = std::max<int>(ms, 8);
  } else {
    set_gram(P, n, d, ms, lds_limit);
  }
  const uint32_t core_bytes = P.core ? (uint32_t)(n * rs) : 0u;
  if (!P.matrix) {
    // points (optional) | core distances (when needed) | squared norms
    const uint64_t xs = (uint64_t)n * dk * rs;
    const uint64_t norms = (uint64_t)n * rs;
    P.xlds = !(flags & HDB_FLAG_NO_POINT_LDS) && xs + core_bytes + norms + 3072 <= lds_limit;
    // Without LDS points, one byte per vertex records the wave holding it.
    P.prim_lds = (uint32_t)((P.xlds ? xs : 0) + core_bytes + norms + (P.xlds ? 0 : align16((uint32_t)n)));
  } else {
    P.prim_lds = P.gram_graph ? (uint32_t)(2 * n * 4) : core_bytes;
  }
  const TreeLayout L = tree_layout((int)n, rs);
  P.tree_lds = !(flags & HDB_FLAG_GLOBAL_ARENA) && L.total + 1024 <= lds_limit;
  P.tree_lds_bytes = P.tree_lds ? L.total : 0u;
  P.arena_stride = P.tree_lds ? 0 : align256(L.total);
  P.need_matrix = !(flags & HDB_FLAG_MATRIX_INPUT) && (P.matrix || (P.core && !P.core_topk));
  const bool pieces_needed = (P.core_gram && !P.gram_window) || P.gram_graph;
  size_t o = 0;
  const size_t bn = (size_t)batch * n;
  P.off_eu = o; o = align256(o + bn * 4);
  P.off_ev = o; o = align256(o + bn * 4);
  P.off_ew = o; o = align256(o + bn * rs);
  P.off_flag = o; o = align256(o + (size_t)batch * 4);
  P.off_core = o; o = align256(o + (P.core ? bn * rs : 0));
  P.off_arena = o; o = align256(o + (size_t)batch * P.arena_stride);
  P.slot_stride = align256((size_t)n * (rs == 4 ? sizeof(ClusterSlot<float>) : sizeof(ClusterSlot<double>)));
  P.off_slots = o; o = align256(o + (size_t)batch * P.slot_stride);
  P.off_matrix = o; o = align256(o + (P.need_matrix ? bn * (size_t)n * rs : 0));
  P.off_cand = o; o = align256(o + (P.core_gram ? bn * kGramCap * 2 : 0));
  P.off_ncand = o; o = align256(o + (P.core_gram ? bn * 4 : 0));
  const size_t pieces = P.gram_graph    ? gram_operand_bytes(P.gram_dk, batch, n)
                        : pieces_needed ? bn * 2 * (size_t)P.gram_dk * 2
                                        : 0;
  P.off_pa = o; o = align256(o + pieces);
  P.off_pb = o; o = align256(o + pieces);
  P.off_nb = o; o = align256(o + (pieces_needed ? 3 * bn * 4 : 0));
  P.off_alt = o; o = align256(o + (P.gram_graph ? bn * 4 : 0));
  P.off_gop = o; o = align256(o + (P.gram_exact ? gx_operand_entries(P.gx_dk, batch, n) * 16 : 0));
  P.off_gnorm = o; o = align256(o + (P.gram_exact ? bn * 4 : 0));
  P.off_gu2 = o; o = align256(o + (P.gram_exact ? bn * 4 : 0));
  P.off_gcnt = o; o = align256(o + (P.gram_exact ? bn * 4 : 0));
  P.off_gval = o; o = align256(o + (P.gram_exact ? bn * kGxCoreCap * 4 : 0));
  P.off_gncnt = o; o = align256(o + (P.gram_exact ? (size_t)batch * 4 : 0));
  P.off_gnlist = o; o = align256(o + (P.gram_exact ? bn * kGxNearPerPoint * 4 : 0));
  // Split tree stage when windows outnumber the compute units at least two
  // to one: its narrow hierarchy workgroups then share CUs (B=1024 windows of
  // 1024 points: tree 0.73 -> 0.27 ms; at B=256 the fused kernel is 10 %
  // faster).  HDB_FLAG_SPLIT_TREE / HDB_FLAG_FUSED_TREE force either; export
  // calls keep the fused kernel, whose exports include the sort permutation.
  P.split_tree = !(flags & HDB_FLAG_FUSED_TREE) &&
                 ((flags & HDB_FLAG_SPLIT_TREE) || batch >= 2 * (int64_t)device_cus());
  P.sort_threads = split_sort_threads(n);
  P.hier_threads = split_hier_threads(n);
  P.off_suv = o; o = align256(o + (P.split_tree ? bn * 4 : 0));
  P.off_sw = o; o = align256(o + (P.split_tree ? bn * rs : 0));
  // Prim repacking of the points path: flag bits 14..15 (0: every 128
  // steps, 1: never, 2: every 32, 3: every 64).  128 measured best: equal
  // or faster than 64 at N=1024 and 4 % faster on ragged batches, whose
  // largest windows reload coordinates from device memory at each repacking.
  const int rp = (flags >> 14) & 3;
  P.repack = P.matrix ? 0 : rp == 1 ? 0 : rp == 2 ? 32 : rp == 3 ? 64 : 128;
  P.off_pack = o; o = align256(o + (P.repack ? bn * (rs == 4 ? sizeof(PackRec<float>) : sizeof(PackRec<double>)) : 0));
  P.total = o;
  *p = P;
  return 0;
}

template <typename Real>
int launch_distances(hipStream_t s, const Real* x, const Launch& L, int64_t d, Real* out, int32_t* status) {
  const unsigned nt = (unsigned)((L.nmax + 63) / 64);
  const dim3 grid(nt * (nt + 1) / 2, (unsigned)L.count);
  distance_kernel<Real><<<grid, 256, 0, s>>>(x, L.win, (int)d, out, status);
  return check_launch("distance_kernel");
}

template <typename Real, int KMAX>
int launch_topk_k(hipStream_t s, const Real* x, const Launch& L, int64_t d, int ms, Real* core, int dk,
                  int32_t* status) {
  const dim3 grid((unsigned)((L.nmax + 255) / 256), (unsigned)L.count);
  const Windows w = L.win;
  switch (dk) {
    case 8: core_topk_kernel<Real, KMAX, 8><<<grid, 256, 0, s>>>(x, w, (int)d, ms, core, status); break;
    case 16: core_topk_kernel<Real, KMAX, 16><<<grid, 256, 0, s>>>(x, w, (int)d, ms, core, status); break;
    case 32: core_topk_kernel<Real, KMAX, 32><<<grid, 256, 0, s>>>(x, w, (int)d, ms, core, status); break;
    case 64: core_topk_kernel<Real, KMAX, 64><<<grid, 256, 0, s>>>(x, w, (int)d, ms, core, status); break;
    default: core_topk_kernel<Real, KMAX, 0><<<grid, 256, 0, s>>>(x, w, (int)d, ms, core, status); break;
  }
  return check_launch("core_topk_kernel");
}

struct GramBuffers {
  uint16_t* pa;
  uint16_t* pb;
  float* nb;
  uint16_t* cand;
  int32_t* ncand;
};

template <int M, int DK, int RSET>
int launch_gram_k(hipStream_t s, const Plan& P, const Launch& L, int ms, const GramBuffers& g) {
  auto fn = core_gram_kernel<M, DK, RSET>;
  if (int r = allow_lds((const void*)fn, P.gram_lds)) return r;
  const int rows = P.gram_threads / 64 * GramShape<DK, RSET>::WR;
  const dim3 grid((unsigned)((L.nmax + rows - 1) / rows), (unsigned)L.count);
  fn<<<grid, P.gram_threads, P.gram_lds, s>>>(L.win, ms, g.pa, g.pb, g.nb, L.points, g.cand, g.ncand);
  return check_launch("core_gram_kernel");
}

template <int M, int DK, int RA, int RB>
int launch_gram_r(hipStream_t s, const Plan& P, const Launch& L, int ms, const GramBuffers& g) {
  if (P.gram_rs == RB) return launch_gram_k<M, DK, RB>(s, P, L, ms, g);
  return launch_gram_k<M, DK, RA>(s, P, L, ms, g);
}

template <int M>
int launch_gram_d(hipStream_t s, const Plan& P, const Launch& L, int ms, const GramBuffers& g) {
  switch (P.gram_dk) {
    case 8: return launch_gram_r<M, 8, 2, 4>(s, P, L, ms, g);
    case 16: return launch_gram_r<M, 16, 2, 4>(s, P, L, ms, g);
    case 32: return launch_gram_r<M, 32, 2, 4>(s, P, L, ms, g);
    case 64: return launch_gram_r<M, 64, 1, 2>(s, P, L, ms, g);
    case 128: return launch_gram_r<M, 128, 1, 2>(s, P, L, ms, g);
    case 256: return launch_gram_r<M, 256, 1, 1>(s, P, L, ms, g);
  }
  return fail("internal: unsupported Gram padding");
}

GramBuffers gram_buffers(const Plan& P, unsigned char* ws) {
  GramBuffers g;
  g.pa = reinterpret_cast<uint16_t*>(ws + P.off_pa);
  g.pb = reinterpret_cast<uint16_t*>(ws + P.off_pb);
  g.nb = reinterpret_cast<float*>(ws + P.off_nb);
  g.cand = reinterpret_cast<uint16_t*>(ws + P.off_cand);
  g.ncand = reinterpret_cast<int32_t*>(ws + P.off_ncand);
  return g;
}

// bf16 pieces and norm terms of every point of the batch.
int launch_gram_prep(hipStream_t s, const Plan& P, const float* x, int64_t points, int64_t d, unsigned char* ws,
                     int32_t* status) {
  const GramBuffers g = gram_buffers(P, ws);
  const unsigned blocks = (unsigned)((points + 255) / 256);
  switch (P.gram_dk) {
    case 8: gram_prep_lds_kernel<8><<<blocks, 256, 0, s>>>(x, points, (int)d, g.pa, g.pb, g.nb, status); break;
    case 16: gram_prep_lds_kernel<16><<<blocks, 256, 0, s>>>(x, points, (int)d, g.pa, g.pb, g.nb, status); break;
    case 32: gram_prep_kernel<32><<<blocks, 256, 0, s>>>(x, points, (int)d, g.pa, g.pb, g.nb, status); break;
    case 64: gram_prep_kernel<64><<<blocks, 256, 0, s>>>(x, points, (int)d, g.pa, g.pb, g.nb, status); break;
    case 128: gram_prep_kernel<128><<<blocks, 256, 0, s>>>(x, points, (int)d, g.pa, g.pb, g.nb, status); break;
    default: gram_prep_kernel<256><<<blocks, 256, 0, s>>>(x, points, (int)d, g.pa, g.pb, g.nb, status); break;
  }
  return check_launch("gram_prep_kernel");
}

// The MFMA Gram form of the graph stage (hdb_gramgraph.h): G, core
// distances from G, interval-key Prim, exact weights of the open edges.
template <int DK>
int launch_gram_matrix(hipStream_t s, const float* x, int64_t batch, int64_t n, int64_t d, const GramBuffers& g,
                       int64_t points, float* G, int32_t* status) {
  constexpr int NQ = GramShape<DK, 1>::NQ;
  gram_s16x8* aop = reinterpret_cast<gram_s16x8*>(g.pa);
  gram_s16x8* bop = reinterpret_cast<gram_s16x8*>(g.pb);
  const int nblk = (int)((n + 15) / 16);
  const int64_t entries = batch * nblk * NQ * 64;
  gram_operand_kernel<DK><<<(unsigned)((entries + 255) / 256), 256, 0, s>>>(x, (int)n, (int)d, entries, nblk, aop, bop);
  if (int r = check_launch("gram_operand_kernel")) return r;
  stage_mark(s, "gram_operand_kernel");
  auto fn = gram_matrix_kernel<DK>;
  if (int r = allow_lds((const void*)fn, kGramMatrixLds)) return r;
  const unsigned tps = (unsigned)((n + kGramTile - 1) / kGramTile);
  fn<<<dim3(tps * tps, (unsigned)batch), 256, kGramMatrixLds, s>>>(x, (int)n, (int)d, aop, bop, g.nb + 2 * points,
                                                                     G, status);
  return check_launch("gram_matrix_kernel");
}

template <int VPT>
int launch_prim_gram(hipStream_t s, const Plan& P, int64_t batch, const GramPrimArgs& a) {
  auto fn = prim_gram_kernel<VPT>;
  if (int r = allow_lds((const void*)fn, P.prim_lds)) return r;
  fn<<<(unsigned)batch, P.threads, P.prim_lds, s>>>(a);
  return check_launch("prim_gram_kernel");
}

int run_gram_graph(hipStream_t s, const Plan& P, const float* x, const Launch& L, int64_t d, int ms, bool unit,
                   unsigned char* ws, int32_t* status) {
  const int64_t batch = L.count, n = L.nmax, points = L.points;
  const GramBuffers g = gram_buffers(P, ws);
  float* G = reinterpret_cast<float*>(ws + P.off_matrix);
  float* core = reinterpret_cast<float*>(ws + P.off_core);
  gram_norm_kernel<<<(unsigned)((points + 255) / 256), 256, 0, s>>>(x, points, (int)d, g.nb + 2 * points, status);
  if (int e = check_launch("gram_norm_kernel")) return e;
  stage_mark(s, "gram_norm_kernel");
  int r = 0;
  switch (P.gram_dk) {
    case 8: r = launch_gram_matrix<8>(s, x, batch, n, d, g, points, G, status); break;
    case 16: r = launch_gram_matrix<16>(s, x, batch, n, d, g, points, G, status); break;
    case 32: r = launch_gram_matrix<32>(s, x, batch, n, d, g, points, G, status); break;
    case 64: r = launch_gram_matrix<64>(s, x, batch, n, d, g, points, G, status); break;
    case 128: r = launch_gram_matrix<128>(s, x, batch, n, d, g, points, G, status); break;
    default: r = launch_gram_matrix<256>(s, x, batch, n, d, g, points, G, status); break;
  }
  if (r) return r;
  stage_mark(s, "gram_matrix_kernel");
  const float c2 = gram_c2(P.gram_dk);
  const float* norms = g.nb + 2 * points;
  if (!unit) {
    const unsigned blocks = (unsigned)((points + 3) / 4);
    if (ms <= 8)
      core_from_gram_kernel<8><<<blocks, 256, 0, s>>>(G, x, norms, points, (int)n, (int)d, ms, c2, core, status);
    else if (ms <= 16)
      core_from_gram_kernel<16><<<blocks, 256, 0, s>>>(G, x, norms, points, (int)n, (int)d, ms, c2, core, status);
    else if (ms <= 32)
      core_from_gram_kernel<32><<<blocks, 256, 0, s>>>(G, x, norms, points, (int)n, (int)d, ms, c2, core, status);
    else
      core_from_gram_kernel<64><<<blocks, 256, 0, s>>>(G, x, norms, points, (int)n, (int)d, ms, c2, core, status);
    if (int e = check_launch("core_from_gram_kernel")) return e;
    stage_mark(s, "core_from_gram_kernel");
  }
  GramPrimArgs a;
  a.g = G;
  a.x = x;
  a.core = unit ? nullptr : core;
  a.norms = norms;
  a.n = (int)n;
  a.d = (int)d;
  a.unit_core = unit ? 1 : 0;
  a.c2 = c2;
  a.edge_u = reinterpret_cast<int32_t*>(ws + P.off_eu);
  a.edge_v = reinterpret_cast<int32_t*>(ws + P.off_ev);
  a.edge_w = reinterpret_cast<float*>(ws + P.off_ew);
  a.edge_alt = reinterpret_cast<int32_t*>(ws + P.off_alt);
  a.wflag = reinterpret_cast<int32_t*>(ws + P.off_flag);
  a.status = status;
  switch (P.vpt) {
    case 1: r = launch_prim_gram<1>(s, P, batch, a); break;
    case 2: r = launch_prim_gram<2>(s, P, batch, a); break;
    default: r = launch_prim_gram<4>(s, P, batch, a); break;
  }
  if (r) return r;
  stage_mark(s, "prim_kernel");
  edge_fix_kernel<<<dim3((unsigned)((n - 1 + 255) / 256), (unsigned)batch), 256, 0, s>>>(a);
  if (int e = check_launch("edge_fix_kernel")) return e;
  stage_mark(s, "edge_fix_kernel");
  return 0;
}

// Candidates (MFMA bounds), then the exact selection (after the prep, which
// a ragged batch runs once for all its launches).  Per-lane minima kept in
// pass A: one for min_samples <= 3, else two (the row bound stays tight
// unless three of the nearest share a lane).
int launch_gram(hipStream_t s, const Plan& P, const float* x, const float* dist, const Launch& L, int64_t d,
                int ms, float* core, unsigned char* ws, int32_t* status, bool prep) {
  const GramBuffers g = gram_buffers(P, ws);
  if (P.gram_window && !dist) {
    // Both passes and the exact selection in one workgroup per window; the
    // candidate-list buffer holds each window's (row, column) pairs.
    const int threads = (int)std::min<int64_t>(1024, (L.nmax + 63) / 64 * 64);
    const uint32_t lds = gram_window_lds_bytes((int)L.nmax, (int)d);
    const bool vec = d == kWinDK && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
    auto fn = ms <= 8 ? (vec ? core_window_kernel<2, 8, true> : core_window_kernel<2, 8, false>)
                      : (vec ? core_window_kernel<4, 16, true> : core_window_kernel<4, 16, false>);
    if (int r = allow_lds((const void*)fn, lds)) return r;
    fn<<<(unsigned)L.count, threads, lds, s>>>(x, L.win, (int)d, ms, reinterpret_cast<uint32_t*>(g.cand), core,
                                              status);
    return check_launch("core_window_kernel");
  } else {
    if (prep) {
      if (int r = launch_gram_prep(s, P, x, L.points, d, ws, status)) return r;
      stage_mark(s, "gram_prep_kernel");
    }
    const int r = ms <= 3 ? launch_gram_d<1>(s, P, L, ms, g) : launch_gram_d<2>(s, P, L, ms, g);
    if (r) return r;
    stage_mark(s, "core_gram_kernel");
  }
  const uint64_t xbytes = (uint64_t)L.nmax * d * sizeof(float);
  if (!dist && xbytes <= kExactLdsBytes) {
    // The window's points fit in LDS: one workgroup per window reads them
    // once and every candidate from LDS.
    const int threads = (int)std::min<int64_t>(1024, (L.nmax + 63) / 64 * 64);
    auto fn = ms <= 8 ? core_exact_lds_kernel<8> : core_exact_lds_kernel<16>;
    if (int r = allow_lds((const void*)fn, (uint32_t)xbytes)) return r;
    fn<<<(unsigned)L.count, threads, (uint32_t)xbytes, s>>>(x, L.win, (int)d, ms, g.cand, g.ncand, core, status);
    return check_launch("core_exact_kernel");
  }
  const dim3 grid((unsigned)((L.nmax + 255) / 256), (unsigned)L.count);
  if (ms <= 8)
    core_exact_kernel<8><<<grid, 256, 0, s>>>(x, dist, L.win, (int)d, ms, g.cand, g.ncand, core, status);
  else
    core_exact_kernel<16><<<grid, 256, 0, s>>>(x, dist, L.win, (int)d, ms, g.cand, g.ncand, core, status);
  return check_launch("core_exact_kernel");
}

template <typename Real>
int launch_core(hipStream_t s, const Plan& P, const Real* x, const Launch& L, int64_t d, int ms, Real* core,
                Real* matrix, unsigned char* ws, int32_t* status, bool matrix_ready, bool gram_prep) {
  if constexpr (sizeof(Real) == 4) {
    if (P.core_gram)
      return launch_gram(s, P, x, matrix_ready ? matrix : nullptr, L, d, ms, core, ws, status, gram_prep);
  }
  if (P.core_topk && !matrix_ready) {
    switch (P.kmax) {
      case 8: return launch_topk_k<Real, 8>(s, x, L, d, ms, core, P.core_dk, status);
      case 16: return launch_topk_k<Real, 16>(s, x, L, d, ms, core, P.core_dk, status);
      case 32: return launch_topk_k<Real, 32>(s, x, L, d, ms, core, P.core_dk, status);
      default: return launch_topk_k<Real, 64>(s, x, L, d, ms, core, P.core_dk, status);
    }
  }
  const int64_t n = L.nmax;
  if (!matrix_ready) {
    if (int r = launch_distances<Real>(s, x, L, d, matrix, status)) return r;
  }
  const dim3 grid((unsigned)((n + 3) / 4), (unsigned)L.count);
  if (ms <= 8) {
    core_select_list_kernel<Real, 8><<<grid, 256, 0, s>>>(matrix, L.win, ms, core);
  } else if (ms <= 16) {
    core_select_list_kernel<Real, 16><<<grid, 256, 0, s>>>(matrix, L.win, ms, core);
  } else if (ms <= 32) {
    core_select_list_kernel<Real, 32><<<grid, 256, 0, s>>>(matrix, L.win, ms, core);
  } else if (ms <= 64) {
    core_select_list_kernel<Real, 64><<<grid, 256, 0, s>>>(matrix, L.win, ms, core);
  } else {
    const uint32_t bytes = (uint32_t)(n * sizeof(Real));
    core_select_kernel<Real><<<dim3((unsigned)n, (unsigned)L.count), 64, bytes, s>>>(matrix, L.win, ms, core);
  }
  return check_launch("core_select_kernel");
}

template <typename Real, int VPT, int DK, int MODE, bool XLDS>
int launch_prim_one(hipStream_t s, const Plan& P, int64_t batch, const PrimArgs& a) {
  auto fn = prim_kernel<Real, VPT, DK, MODE, XLDS>;
  if (int r = allow_lds((const void*)fn, P.prim_lds)) return r;
  fn<<<(unsigned)batch, P.threads, P.prim_lds, s>>>(a);  // batch: workgroups along the window axis
  return check_launch("prim_kernel");
}

template <typename Real, int VPT>
int launch_prim_spec(hipStream_t s, const Plan& P, int64_t batch, const PrimArgs& a) {
  auto fn = prim_spec_kernel<Real, VPT>;
  if (int r = allow_lds((const void*)fn, P.prim_lds)) return r;
  fn<<<(unsigned)batch, P.threads, P.prim_lds, s>>>(a);  // batch: workgroups along the window axis
  return check_launch("prim_spec_kernel");
}

template <typename Real, int VPT, int DK>
int launch_prim_x(hipStream_t s, const Plan& P, int64_t batch, const PrimArgs& a) {
  if constexpr (VPT * DK * sizeof(Real) <= (VPT == 8 ? 64 * 8 : 64 * 4)) {
    if (a.prefilter) {
      if (P.xlds) return launch_prim_one<Real, VPT, DK, 2, true>(s, P, batch, a);
      return launch_prim_one<Real, VPT, DK, 2, false>(s, P, batch, a);
    }
    if (P.xlds) return launch_prim_one<Real, VPT, DK, 0, true>(s, P, batch, a);
    return launch_prim_one<Real, VPT, DK, 0, false>(s, P, batch, a);
  }
  return fail("internal: unsupported points layout");
}

template <typename Real, int VPT>
int launch_prim_v(hipStream_t s, const Plan& P, int64_t batch, const PrimArgs& a) {
  if (P.matrix) {
    if (P.flags & HDB_FLAG_ONE_STEP) return launch_prim_one<Real, VPT, 0, 1, false>(s, P, batch, a);
    return launch_prim_spec<Real, VPT>(s, P, batch, a);
  }
  switch (P.dk) {
    case 8: return launch_prim_x<Real, VPT, 8>(s, P, batch, a);
    case 16: return launch_prim_x<Real, VPT, 16>(s, P, batch, a);
    case 32: return launch_prim_x<Real, VPT, 32>(s, P, batch, a);
    case 64: return launch_prim_x<Real, VPT, 64>(s, P, batch, a);
  }
  return fail("internal: unsupported dimension padding");
}

template <typename Real>
int launch_prim(hipStream_t s, const Plan& P, int64_t batch, const PrimArgs& a) {
  switch (P.vpt) {
    case 1: return launch_prim_v<Real, 1>(s, P, batch, a);
    case 2: return launch_prim_v<Real, 2>(s, P, batch, a);
    case 4: return launch_prim_v<Real, 4>(s, P, batch, a);
    case 8:
      if (P.matrix) return launch_prim_spec<Real, 8>(s, P, batch, a);
      if constexpr (sizeof(Real) == 4) {
        if (P.dk == 8) return launch_prim_x<Real, 8, 8>(s, P, batch, a);
        if (P.dk == 16) return launch_prim_x<Real, 8, 16>(s, P, batch, a);
      }
      return fail("internal: unsupported points layout");
  }
  return fail("internal: unsupported vertices per thread");
}

template <typename Real>
int launch_tree(hipStream_t s, const Plan& P, int64_t batch, const TreeArgs& a) {
  if (P.split_tree && !a.export_ints) {
    if (P.tree_lds) {
      auto fs = sort_edges_kernel<Real, true>;
      auto fh = hierarchy_kernel<Real, true>;
      if (int r = allow_lds((const void*)fs, P.tree_lds_bytes)) return r;
      if (int r = allow_lds((const void*)fh, P.tree_lds_bytes)) return r;
      fs<<<(unsigned)batch, P.sort_threads, P.tree_lds_bytes, s>>>(a);
      if (int r = check_launch("sort_edges_kernel")) return r;
      stage_mark(s, "sort_edges_kernel");
      fh<<<(unsigned)batch, P.hier_threads, P.tree_lds_bytes, s>>>(a);
    } else {
      sort_edges_kernel<Real, false><<<(unsigned)batch, P.sort_threads, 0, s>>>(a);
      if (int r = check_launch("sort_edges_kernel")) return r;
      stage_mark(s, "sort_edges_kernel");
      hierarchy_kernel<Real, false><<<(unsigned)batch, P.hier_threads, 0, s>>>(a);
    }
    return check_launch("hierarchy_kernel");
  }
  if (P.tree_lds) {
    auto fn = tree_kernel<Real, true>;
    if (int r = allow_lds((const void*)fn, P.tree_lds_bytes)) return r;
    fn<<<(unsigned)batch, kTreeThreads, P.tree_lds_bytes, s>>>(a);
  } else {
    tree_kernel<Real, false><<<(unsigned)batch, kTreeThreads, 0, s>>>(a);
  }
  return check_launch("tree_kernel");
}

// The MFMA-bounded matrix, core distances from its lists, and Prim over it
// (hdb_gramexact.h).  Dense batches only.
template <int DK>
int launch_gx_matrix(hipStream_t s, const Plan& P, const float* x, int64_t batch, int64_t n, int64_t d,
                     unsigned char* ws, float* M, int32_t* status) {
  const int nblk = (int)((n + 15) / 16);
  const gram_s16x8* op = reinterpret_cast<const gram_s16x8*>(ws + P.off_gop);
  const float* norms = reinterpret_cast<const float*>(ws + P.off_gnorm);
  float* u2 = reinterpret_cast<float*>(ws + P.off_gu2);
  constexpr int RB = DK >= 128 ? 2 : 4;
  const float c2 = gram_c2(DK);
  gx_bound_kernel<DK, RB><<<xcd_grid(batch, (int)((n + 64 * RB - 1) / (64 * RB))), 256, 0, s>>>(
      op, norms, batch, (int)n, nblk, P.gx_cols, P.gx_rank, c2 + kGxPassA, u2);
  if (int r = check_launch("gx_bound_kernel")) return r;
  stage_mark(s, "gx_bound_kernel");
  const unsigned nt = (unsigned)((n + kGxTile - 1) / kGxTile);
  int32_t* lcnt = reinterpret_cast<int32_t*>(ws + P.off_gcnt);
  float* lval = reinterpret_cast<float*>(ws + P.off_gval);
  int32_t* ncnt = reinterpret_cast<int32_t*>(ws + P.off_gncnt);
  uint32_t* nlist = reinterpret_cast<uint32_t*>(ws + P.off_gnlist);
  const int cap = kGxNearPerPoint * (int)n;
  // Persistent: one wave-specialised workgroup per CU (a multiple of 8 keeps
  // each on its XCD).
  const unsigned tiles = xcd_grid(batch, (int)(nt * (nt + 1) / 2));
  gx_matrix_kernel<DK><<<std::min<unsigned>(tiles, (unsigned)((device_cus() + 7) / 8 * 8)), 512, 0, s>>>(
      x, batch, (int)n, (int)d, nblk, op, norms, u2, M, lcnt, lval, ncnt, nlist, cap, status);
  if (int r = check_launch("gx_matrix_kernel")) return r;
  stage_mark(s, "gx_matrix_kernel");
  gx_near_kernel<DK><<<xcd_grid(batch, (cap + 255) / 256), 256, 0, s>>>(x, batch, (int)n, (int)d, cap, ncnt, nlist, u2,
                                                                   M, lcnt, lval, status);
  return check_launch("gx_near_kernel");
}

template <int DK>
int launch_gx_prep(hipStream_t s, const Plan& P, const float* x, int64_t batch, int64_t n, int64_t d,
                   unsigned char* ws, int32_t* status) {
  const int nblk = (int)((n + 15) / 16);
  gx_prep_kernel<DK><<<dim3((unsigned)nblk, (unsigned)batch), 256, 0, s>>>(
      x, (int)n, (int)d, nblk, reinterpret_cast<gram_s16x8*>(ws + P.off_gop),
      reinterpret_cast<float*>(ws + P.off_gnorm), reinterpret_cast<int32_t*>(ws + P.off_gcnt),
      reinterpret_cast<int32_t*>(ws + P.off_gncnt), status);
  return check_launch("gx_prep_kernel");
}

template <int VPT>
int launch_prim_gx(hipStream_t s, const Plan& P, int64_t batch, const GxPrimArgs& a) {
  auto fn = prim_gx_kernel<VPT>;
  if (int r = allow_lds((const void*)fn, P.prim_lds)) return r;
  fn<<<(unsigned)batch, P.threads, P.prim_lds, s>>>(a);
  return check_launch("prim_gx_kernel");
}

// Prep, pass A, the bounded matrix M and the exact near pairs.
int run_gram_matrix_only(hipStream_t s, const Plan& P, const float* x, int64_t batch, int64_t n, int64_t d,
                         unsigned char* ws, float* M, int32_t* status) {
  int r = 0;
  switch (P.gx_dk) {
    case 32: r = launch_gx_prep<32>(s, P, x, batch, n, d, ws, status); break;
    case 64: r = launch_gx_prep<64>(s, P, x, batch, n, d, ws, status); break;
    case 128: r = launch_gx_prep<128>(s, P, x, batch, n, d, ws, status); break;
    default: r = launch_gx_prep<256>(s, P, x, batch, n, d, ws, status); break;
  }
  if (r) return r;
  stage_mark(s, "gx_prep_kernel");
  switch (P.gx_dk) {
    case 32: r = launch_gx_matrix<32>(s, P, x, batch, n, d, ws, M, status); break;
    case 64: r = launch_gx_matrix<64>(s, P, x, batch, n, d, ws, M, status); break;
    case 128: r = launch_gx_matrix<128>(s, P, x, batch, n, d, ws, M, status); break;
    default: r = launch_gx_matrix<256>(s, P, x, batch, n, d, ws, M, status); break;
  }
  if (r) return r;
  stage_mark(s, "gx_near_kernel");
  return 0;
}

int run_gram_exact(hipStream_t s, const Plan& P, const float* x, const Launch& L, int64_t d, int ms, bool unit,
                   unsigned char* ws, int32_t* status) {
  const int64_t batch = L.count, n = L.nmax, points = L.points;
  if (points != batch * n) return fail("internal: ragged batch on the MFMA-bounded matrix path");
  int r = run_gram_matrix_only(s, P, x, batch, n, d, ws, reinterpret_cast<float*>(ws + P.off_matrix), status);
  if (r) return r;
  const float* M = reinterpret_cast<const float*>(ws + P.off_matrix);
  float* core = reinterpret_cast<float*>(ws + P.off_core);
  if (!unit) {
    gx_core_kernel<16><<<(unsigned)((points + 3) / 4), 256, 0, s>>>(
        M, points, (int)n, ms, reinterpret_cast<const int32_t*>(ws + P.off_gcnt),
        reinterpret_cast<const float*>(ws + P.off_gval), core);
    if (int e = check_launch("gx_core_kernel")) return e;
    stage_mark(s, "gx_core_kernel");
  }
  GxPrimArgs a;
  a.m = M;
  a.x = x;
  a.core = unit ? nullptr : core;
  a.norms = reinterpret_cast<const float*>(ws + P.off_gnorm);
  a.n = (int)n;
  a.d = (int)d;
  a.unit_core = unit ? 1 : 0;
  a.c2 = gram_c2(P.gx_dk);
  a.edge_u = reinterpret_cast<int32_t*>(ws + P.off_eu);
  a.edge_v = reinterpret_cast<int32_t*>(ws + P.off_ev);
  a.edge_w = reinterpret_cast<float*>(ws + P.off_ew);
  a.wflag = reinterpret_cast<int32_t*>(ws + P.off_flag);
  a.status = status;
  switch (P.vpt) {
    case 1: r = launch_prim_gx<1>(s, P, batch, a); break;
    case 2: r = launch_prim_gx<2>(s, P, batch, a); break;
    default: r = launch_prim_gx<4>(s, P, batch, a); break;
  }
  if (r) return r;
  stage_mark(s, "prim_kernel");
  return 0;
}

template <typename Real>
int run_cluster(hipStream_t s, const Plan& P, const Real* x, const Real* given_dist, const Real* given_core,
                const Launch& L, int64_t d, long long mcs, int ms, bool unit, double eps, int sort_depth,
                int32_t* labels, unsigned char* ws, int32_t* status, int32_t* exp_i, void* exp_r,
                bool gram_prep = true, hipEvent_t core_done = nullptr) {
  int32_t* eu = reinterpret_cast<int32_t*>(ws + P.off_eu);
  int32_t* ev = reinterpret_cast<int32_t*>(ws + P.off_ev);
  Real* ew = reinterpret_cast<Real*>(ws + P.off_ew);
  int32_t* wflag = reinterpret_cast<int32_t*>(ws + P.off_flag);
  Real* core = reinterpret_cast<Real*>(ws + P.off_core);
  Real* matrix = reinterpret_cast<Real*>(ws + P.off_matrix);
  const Real* dist = given_dist;
  const Real* core_in = given_core;
  stage_mark(s, "start");
  bool prim_done = false;
  if constexpr (sizeof(Real) == 4) {
    if (P.gram_graph && !given_dist) {
      if (L.win.off) return fail("internal: ragged batch on the matrix path");
      if (int r = run_gram_graph(s, P, x, L, d, ms, unit, ws, status)) return r;
      prim_done = true;
    } else if (P.gram_exact && !given_dist) {
      if (L.win.off) return fail("internal: ragged batch on the MFMA-bounded matrix path");
      if (int r = run_gram_exact(s, P, x, L, d, ms, unit, ws, status)) return r;
      prim_done = true;
    }
  }
  if (!given_dist && !prim_done) {
    bool matrix_ready = false;
    if (P.matrix) {
      if (int r = launch_distances<Real>(s, x, L, d, matrix, status)) return r;
      matrix_ready = true;
      dist = matrix;
      stage_mark(s, "distance_kernel");
    }
    if (!unit) {
      if (int r = launch_core<Real>(s, P, x, L, d, ms, core, matrix, ws, status, matrix_ready, gram_prep)) return r;
      core_in = core;
      stage_mark(s, P.core_gram ? (P.gram_window && !matrix_ready ? "core_window_kernel" : "core_exact_kernel")
                                : matrix_ready ? "core_select_kernel" : "core_topk_kernel");
    }
  }
  if (core_done && hipEventRecord(core_done, s) != hipSuccess) return fail("event record failed");
  PrimArgs pa;
  pa.x = x;
  pa.dist = dist;
  pa.core = unit ? nullptr : core_in;
  pa.win = L.win;
  pa.d = (int)d;
  pa.unit_core = unit ? 1 : 0;
  pa.prefilter = (P.flags & HDB_FLAG_PREFILTER) ? 1 : 0;
  pa.repack = P.repack;
  pa.pack = ws + P.off_pack;
  pa.edge_u = eu;
  pa.edge_v = ev;
  pa.edge_w = ew;
  pa.wflag = wflag;
  pa.status = status;
  if (!prim_done) {
    pa.win.queue = L.prim_queue;
    const int64_t grid = L.prim_queue ? std::min(L.count, L.persist) : L.count;
    if (int r = launch_prim<Real>(s, P, grid, pa)) return r;
    stage_mark(s, "prim_kernel");
  }
  TreeArgs ta;
  ta.edge_u = eu;
  ta.edge_v = ev;
  ta.edge_w = ew;
  ta.wflag = wflag;
  ta.win = L.win;
  ta.mcs = mcs;
  ta.epsilon = eps;
  ta.sort_depth = sort_depth;
  ta.serial_sort = (P.flags & HDB_FLAG_SERIAL_SORT) ? 1 : 0;
  ta.serial_link = (P.flags & HDB_FLAG_SERIAL_LINKAGE) ? 1 : 0;
  ta.labels = labels;
  ta.arena = ws + P.off_arena;
  ta.arena_stride = P.arena_stride;
  ta.slots = ws + P.off_slots;
  ta.slot_stride = P.slot_stride;
  ta.status = status;
  ta.export_ints = exp_i;
  ta.export_reals = exp_r;
  ta.sorted_uv = reinterpret_cast<uint32_t*>(ws + P.off_suv);
  ta.sorted_w = ws + P.off_sw;
  // The split tree stage runs two kernels over the windows: one item per
  // workgroup (a claim counter serves one launch).
  ta.win.queue = P.split_tree ? nullptr : L.tree_queue;
  if (int r = launch_tree<Real>(s, P, ta.win.queue ? std::min(L.count, L.persist) : L.count, ta)) return r;
  stage_mark(s, "tree_kernel");
  return 0;
}

// ---------------------------------------------------------------------------
// Ragged batches: windows of uneven size, packed point after point.
//
// Windows are grouped by size into buckets (64, 128], (128, 256], ...,
// (2048, 4096] (the first bucket also takes every size <= 64).  Each bucket
// is one pass of the pipeline planned for its largest window; its kernels
// run over all windows and workgroups of other buckets exit at once.  The
// buckets are independent, so they run concurrently on internal streams
// forked from and joined back into the caller's stream.  Per-point buffers
// are shared by all buckets; the split bf16 operands of the Gram filter are
// written once for the whole batch before the fork.

constexpr int kBuckets = 7;
constexpr int kBucketTop[kBuckets] = {64, 128, 256, 512, 1024, 2048, 4096};
constexpr int kSideStreams = 4;

struct RaggedPlan {
  int64_t windows = 0, points = 0;
  int64_t count[kBuckets] = {}, nmax[kBuckets] = {}, start[kBuckets + 1] = {};
  Plan plan[kBuckets];
  bool used[kBuckets] = {};
  bool gram = false, global_arena = false;  // gram: some bucket reads the prep kernel's pieces
  bool gram_cand = false;                   // some bucket writes Gram candidate lists
  int gram_bucket = -1;
  int top = -1;       // bucket of the largest windows
  int64_t head = 0;   // its leading windows run as a separate pass (0: one pass)
  uint32_t matrix_mask = 0;
  int64_t matrix_elems = 0;
  size_t off_eu = 0, off_ev = 0, off_ew = 0, off_flag = 0, off_core = 0, off_arena = 0, off_slots = 0;
  size_t off_cand = 0, off_ncand = 0, off_pa = 0, off_pb = 0, off_nb = 0, off_pack = 0;
  size_t off_list = 0, off_moff = 0, off_queue = 0, off_matrix = 0, off_suv = 0, off_sw = 0, total = 0;
};

int bucket_of(int64_t n) {
  int k = 0;
  while (n > kBucketTop[k]) ++k;
  return k;
}

// Per-bucket tuning flags (vertices per thread, Gram row sets, repacking,
// arena): bits of these masks from a tuning table replace the call's.
constexpr int kTuneMask = 0xFF00 | HDB_FLAG_GLOBAL_ARENA | HDB_FLAG_NO_POINT_LDS;
std::mutex g_tune_mutex;
int g_tune[kBuckets] = {};
bool g_tune_set = false;

int make_ragged_plan(const int64_t* off, int64_t windows, int64_t d, int dtype, int32_t ms, int flags,
                     RaggedPlan* out) {
  if (!off) return fail("null pointer argument");
  if (windows < 1) return fail("X requires nonempty batches/features and at least 2 samples");
  if (windows > (1 << 20)) return fail("at most 2^20 windows per ragged call");
  if (off[0] != 0) return fail("window offsets must start at 0");
  RaggedPlan R;
  R.windows = windows;
  for (int64_t w = 0; w < windows; ++w) {
    const int64_t n = off[w + 1] - off[w];
    if (n < 2) return fail("X requires nonempty batches/features and at least 2 samples");
    if (n > HDB_MAX_SAMPLES) return fail("at most 4096 samples per window are supported");
    if (ms > n) return fail("min_samples must be in [1, n] for every window");
    const int k = bucket_of(n);
    R.used[k] = true;
    R.count[k] += 1;
    R.nmax[k] = std::max<int64_t>(R.nmax[k], n);
  }
  for (int k = 0; k < kBuckets; ++k) R.start[k + 1] = R.start[k] + R.count[k];
  R.points = off[windows];
  // The largest windows' Prim is the critical path: the windows of the top
  // bucket within 85 % of its largest size get their core distances and Prim
  // first, the rest of the bucket follows (the lists order windows largest
  // first, so they are the list's leading entries).
  for (int k = 0; k < kBuckets; ++k)
    if (R.used[k]) R.top = k;
  {
    const int64_t thr = R.nmax[R.top] * 17 / 20;
    for (int64_t w = 0; w < windows; ++w) {
      const int64_t n = off[w + 1] - off[w];
      R.head += bucket_of(n) == R.top && n >= thr;
    }
    if (R.head >= R.count[R.top] || R.nmax[R.top] <= 512) R.head = 0;
  }
  const int rs = dtype == HDB_F32 ? 4 : 8;
  int tune[kBuckets] = {};
  bool tuned = false;
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    tuned = g_tune_set;
    for (int k = 0; k < kBuckets; ++k) tune[k] = g_tune[k];
  }
  // The Gram graph stage and the MFMA-bounded matrix lay matrices out
  // densely; ragged buckets on the matrix path use the exact matrix.
  const int base_flags = (flags & ~HDB_FLAG_GRAM_GRAPH) | HDB_FLAG_NO_GRAM_EXACT;
  for (int k = 0; k < kBuckets; ++k) {
    if (!R.used[k]) continue;
    const int fk = tuned ? (base_flags & ~kTuneMask) | (tune[k] & kTuneMask) : base_flags;
    if (make_plan(R.count[k], R.nmax[k], d, dtype, ms, fk, &R.plan[k])) return -1;
    const Plan& P = R.plan[k];
    if (P.matrix) R.matrix_mask |= 1u << k;
    if (P.core_gram) R.gram_cand = true;
    if (P.core_gram && !P.gram_window) {  // the chunked form reads the prep kernel's pieces
      R.gram = true;
      R.gram_bucket = k;
    }
    if (!P.tree_lds) R.global_arena = true;
  }
  for (int k = 0; k < kBuckets; ++k)
    if (R.used[k] && R.plan[k].core_gram && !R.plan[k].gram_window &&
        R.plan[k].gram_dk != R.plan[R.gram_bucket].gram_dk)
      return fail("internal: Gram paddings differ between buckets");
  for (int64_t w = 0; w < windows; ++w) {
    const int64_t n = off[w + 1] - off[w];
    if ((R.matrix_mask >> bucket_of(n)) & 1u) R.matrix_elems += n * n;
  }
  const size_t pts = (size_t)R.points;
  const bool core = ms > 2;
  const int gram_dk = R.gram ? R.plan[R.gram_bucket].gram_dk : 0;
  size_t o = 0;
  R.off_eu = o; o = align256(o + pts * 4);
  R.off_ev = o; o = align256(o + pts * 4);
  R.off_ew = o; o = align256(o + pts * rs);
  R.off_flag = o; o = align256(o + (size_t)windows * 4);
  R.off_core = o; o = align256(o + (core ? pts * rs : 0));
  R.off_arena = o; o = align256(o + (R.global_arena ? pts * kArenaPerPoint + (size_t)windows * kArenaPerWindow : 0));
  R.off_slots = o;
  o = align256(o + pts * (rs == 4 ? sizeof(ClusterSlot<float>) : sizeof(ClusterSlot<double>)));
  R.off_cand = o; o = align256(o + (R.gram_cand ? pts * kGramCap * 2 : 0));
  R.off_ncand = o; o = align256(o + (R.gram_cand ? pts * 4 : 0));
  const size_t pieces = R.gram ? pts * 2 * (size_t)gram_dk * 2 : 0;
  R.off_pa = o; o = align256(o + pieces);
  R.off_pb = o; o = align256(o + pieces);
  R.off_nb = o; o = align256(o + (R.gram ? 3 * pts * 4 : 0));
  R.off_pack = o; o = align256(o + pts * (rs == 4 ? sizeof(PackRec<float>) : sizeof(PackRec<double>)));
  bool split = false;
  for (int k = 0; k < kBuckets; ++k) split |= R.used[k] && R.plan[k].split_tree;
  R.off_suv = o; o = align256(o + (split ? pts * 4 : 0));
  R.off_sw = o; o = align256(o + (split ? pts * rs : 0));
  R.off_list = o; o = align256(o + (size_t)windows * 4);
  R.off_moff = o; o = align256(o + (size_t)windows * 8);
  R.off_queue = o; o = align256(o + (2 * kBuckets + 2) * 4);
  R.off_matrix = o; o = align256(o + (size_t)R.matrix_elems * rs);
  R.total = o;
  // Point the bucket plans at the shared buffers (per-point addressing).
  for (int k = 0; k < kBuckets; ++k) {
    Plan& P = R.plan[k];
    P.off_eu = R.off_eu; P.off_ev = R.off_ev; P.off_ew = R.off_ew; P.off_flag = R.off_flag;
    P.off_core = R.off_core; P.off_arena = R.off_arena; P.off_slots = R.off_slots;
    P.off_cand = R.off_cand; P.off_ncand = R.off_ncand; P.off_pa = R.off_pa; P.off_pb = R.off_pb;
    P.off_nb = R.off_nb; P.off_matrix = R.off_matrix; P.total = R.total; P.off_pack = R.off_pack;
    P.off_suv = R.off_suv; P.off_sw = R.off_sw;
    P.arena_stride = 0;
    P.slot_stride = 0;
  }
  *out = R;
  return 0;
}

// Internal streams (default priority) and fork/join events, one set per
// (device, caller stream): ragged calls on different caller streams get
// different internal streams and may overlap; calls on one caller stream
// serialise on its set, as they would on the stream itself.
struct SideStreams {
  std::mutex mu;  // held while a call enqueues on this set
  hipStream_t s[kSideStreams];
  hipEvent_t fork, head, tail, join[kSideStreams];
};

int side_streams(hipStream_t caller, SideStreams** out) {
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, SideStreams*> sets;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail("no current HIP device");
  std::lock_guard<std::mutex> lock(mu);
  SideStreams*& S = sets[{dev, caller}];
  if (!S) {
    SideStreams* fresh = new SideStreams;
    for (int i = 0; i < kSideStreams; ++i)
      if (hipStreamCreateWithFlags(&fresh->s[i], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&fresh->join[i], hipEventDisableTiming) != hipSuccess)
        return fail("cannot create internal streams");
    if (hipEventCreateWithFlags(&fresh->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&fresh->head, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&fresh->tail, hipEventDisableTiming) != hipSuccess)
      return fail("cannot create events");
    S = fresh;
  }
  *out = S;
  return 0;
}

// One ragged call: the launch lists (device, largest windows first), the
// Gram operands of the whole batch, then every bucket's pipeline on the
// internal streams.  Prim and tree launches are persistent: a few
// workgroups per compute unit claim the bucket's windows from a device
// counter in list order.
template <typename Real>
int run_ragged(hipStream_t s, const RaggedPlan& R, const Real* x, const int64_t* off_dev, int64_t d,
               long long mcs, int ms, double eps, int32_t* labels, unsigned char* ws, int32_t* status) {
  SideStreams* S = nullptr;
  if (int r = side_streams(s, &S)) return r;
  std::lock_guard<std::mutex> lock(S->mu);
  const bool timers = g_timers.on;
  g_timers.on = false;  // stage events assume one stream
  int err = 0;
  int32_t* lists = reinterpret_cast<int32_t*>(ws + R.off_list);
  int64_t* moff = reinterpret_cast<int64_t*>(ws + R.off_moff);
  int32_t* queues = reinterpret_cast<int32_t*>(ws + R.off_queue);
  {
    RaggedLists rl;
    for (int k = 0; k < 8; ++k) {
      rl.start[k] = k <= kBuckets ? (int32_t)R.start[std::min(k, kBuckets)] : 0;
      rl.top[k] = k < kBuckets ? kBucketTop[k] : HDB_MAX_SAMPLES;
    }
    rl.matrix_mask = R.matrix_mask;
    rl.queues = 2 * kBuckets + 2;
    const unsigned blocks = (unsigned)std::max<int64_t>((R.windows + 63) / 64, 1);
    ragged_lists_kernel<<<blocks, 256, 0, s>>>(off_dev, (int)R.windows, rl, lists, moff, queues);
    err = check_launch("ragged_lists_kernel");
  }
  if constexpr (sizeof(Real) == 4) {
    if (!err && R.gram) err = launch_gram_prep(s, R.plan[R.gram_bucket], x, R.points, d, ws, status);
  }
  if (!err && hipEventRecord(S->fork, s) != hipSuccess) err = fail("event record failed");
  const int cus = device_cus();
  // The bucket of the largest windows is the critical path (its Prim runs
  // longest): it gets the first internal stream to itself, and the other
  // buckets, spread over the remaining streams, start once its core
  // distances are done, so that its first stages have the whole device.
  // With a head (the top bucket's largest windows), the head runs first,
  // the rest of the top bucket starts after the head's core distances and
  // the other buckets after the rest's.
  int used = 0;
  bool waited[kSideStreams] = {false, false, false, false};
  struct Pass {
    int k;
    int64_t first, count;
    int32_t* queue;
  };
  Pass passes[kBuckets + 1];
  int npass = 0;
  for (int k = kBuckets - 1; k >= 0; --k) {
    if (!R.used[k]) continue;
    if (k == R.top && R.head > 0) {
      passes[npass++] = {k, 0, R.head, queues + 2 * k};
      passes[npass++] = {k, R.head, R.count[k] - R.head, queues + 2 * kBuckets};
    } else {
      passes[npass++] = {k, 0, R.count[k], queues + 2 * k};
    }
  }
  const bool split_top = R.head > 0;
  for (int q = 0; q < npass && !err; ++q) {
    const Pass& pq = passes[q];
    const int k = pq.k;
    // Streams: 0 the (head of the) top bucket; with a head, 1 the rest of
    // the top bucket and 2, 3 the other buckets; else 1..3 the others.
    const int si = used == 0 ? 0
                   : split_top ? (used == 1 ? 1 : 2 + (used - 2) % (kSideStreams - 2))
                               : 1 + (used - 1) % (kSideStreams - 1);
    hipStream_t t = S->s[si];
    if (!waited[si]) {
      hipEvent_t after = si == 0 ? S->fork : (split_top && used >= 2 ? S->tail : S->head);
      if (hipStreamWaitEvent(t, after, 0) != hipSuccess) {
        err = fail("stream wait failed");
        break;
      }
      waited[si] = true;
    }
    Launch L;
    L.win.off = off_dev;
    L.win.list = lists + R.start[k] + pq.first;
    L.win.moff = moff;
    L.win.queue = nullptr;
    L.win.n = (int)R.nmax[k];
    L.win.count = (int)pq.count;
    L.count = pq.count;
    L.points = R.points;
    L.nmax = R.nmax[k];
    L.prim_queue = pq.queue;
    L.tree_queue = pq.queue + 1;
    L.persist = (int64_t)cus * std::max(1, 1024 / std::max(64, R.plan[k].threads));
    hipEvent_t core_done = used == 0 ? S->head : (split_top && used == 1 ? S->tail : nullptr);
    err = run_cluster<Real>(t, R.plan[k], x, nullptr, nullptr, L, d, mcs, ms, ms <= 2, eps, -100000, labels, ws,
                            status, nullptr, nullptr, false, core_done);
    ++used;
  }
  for (int i = 0; i < kSideStreams; ++i) {
    if (!waited[i]) continue;
    if (hipEventRecord(S->join[i], S->s[i]) != hipSuccess || hipStreamWaitEvent(s, S->join[i], 0) != hipSuccess)
      err = err ? err : fail("stream join failed");
  }
  g_timers.on = timers;
  return err;
}

// Ragged calls are captured into a HIP graph on an internal stream the first
// time a call signature (buffers, sizes, parameters, window offsets) is seen
// and the graph is replayed onto the caller's stream afterwards: the host
// then enqueues one graph launch instead of some thirty kernel launches and
// stream waits.  The graphs hold kernel launches and event dependencies
// only.
struct RaggedGraph {
  uint64_t key = 0;
  int dev = -1;
  hipGraphExec_t exec = nullptr;
};
constexpr int kGraphCache = 64;

uint64_t mix64(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h * 0xFF51AFD7ED558CCDull;
}

template <typename Real>
int run_ragged_graph(hipStream_t s, const RaggedPlan& R, const Real* x, const int64_t* off, const int64_t* off_dev,
                     int64_t d, long long mcs, int ms, double eps, int32_t* labels, unsigned char* ws,
                     int32_t* status, int flags) {
  static std::mutex mu;
  static RaggedGraph cache[kGraphCache];
  static int filled = 0;
  static hipStream_t capture[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return fail("no current HIP device");
  uint64_t key = mix64(0, (uint64_t)dev);
  for (uint64_t v : {(uint64_t)(uintptr_t)x, (uint64_t)(uintptr_t)off_dev, (uint64_t)(uintptr_t)labels,
                     (uint64_t)(uintptr_t)ws, (uint64_t)(uintptr_t)status, (uint64_t)R.windows, (uint64_t)d,
                     (uint64_t)sizeof(Real), (uint64_t)mcs, (uint64_t)ms, (uint64_t)flags})
    key = mix64(key, v);
  uint64_t eb;
  std::memcpy(&eb, &eps, sizeof(eb));
  key = mix64(key, eb);
  for (int64_t w = 0; w <= R.windows; ++w) key = mix64(key, (uint64_t)off[w]);
  std::lock_guard<std::mutex> lock(mu);
  for (int i = 0; i < filled; ++i)
    if (cache[i].key == key && cache[i].dev == dev) {
      if (hipGraphLaunch(cache[i].exec, s) != hipSuccess) return fail("hipGraphLaunch failed");
      return 0;
    }
  if (filled == kGraphCache) return run_ragged<Real>(s, R, x, off_dev, d, mcs, ms, eps, labels, ws, status);
  if (!capture[dev] && hipStreamCreateWithFlags(&capture[dev], hipStreamNonBlocking) != hipSuccess)
    return fail("cannot create the capture stream");
  hipStream_t c = capture[dev];
  if (hipStreamBeginCapture(c, hipStreamCaptureModeRelaxed) != hipSuccess) {
    (void)hipGetLastError();
    return run_ragged<Real>(s, R, x, off_dev, d, mcs, ms, eps, labels, ws, status);
  }
  const int err = run_ragged<Real>(c, R, x, off_dev, d, mcs, ms, eps, labels, ws, status);
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(c, &graph);
  if (err || ce != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (err) return err;
    return run_ragged<Real>(s, R, x, off_dev, d, mcs, ms, eps, labels, ws, status);
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess || !exec) {
    (void)hipGetLastError();
    return run_ragged<Real>(s, R, x, off_dev, d, mcs, ms, eps, labels, ws, status);
  }
  cache[filled].key = key;
  cache[filled].dev = dev;
  cache[filled].exec = exec;
  ++filled;
  if (hipGraphLaunch(exec, s) != hipSuccess) return fail("hipGraphLaunch failed");
  return 0;
}

// Workspace of hdb_cluster_noalloc: one array per GPU in the library's own
// device memory, held until the process ends.  In host code `scratch` names
// a host-side stand-in; the device address comes from hipGetSymbolAddress
// (scratch_base) and is the only form host code passes on.
constexpr size_t kScratchBytes = (size_t)HDB_NOALLOC_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device, resolved once per device.
int scratch_base(unsigned char** out) {
  static std::mutex mu;
  static void* base[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return fail("no current HIP device");
  std::lock_guard<std::mutex> lock(mu);
  if (!base[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    base[dev] = p;
  }
  *out = static_cast<unsigned char*>(base[dev]);
  return 0;
}

}  // namespace

extern "C" {

int hdb_version(void) { return 100; }

int hdb_profile(int enable) {
  g_timers.on = enable != 0;
  g_timers.count = 0;
  return 0;
}

int hdb_stage_count(void) { return g_timers.count > 0 ? g_timers.count - 1 : 0; }

const char* hdb_stage_name(int i) {
  if (i < 0 || i + 1 >= g_timers.count) return "";
  return g_timers.names[i + 1];
}

float hdb_stage_ms(int i) {
  if (i < 0 || i + 1 >= g_timers.count) return -1.0f;
  float ms = -1.0f;
  if (hipEventElapsedTime(&ms, g_timers.events[i], g_timers.events[i + 1]) != hipSuccess) return -1.0f;
  return ms;
}

const char* hdb_last_error(void) { return g_error.c_str(); }

int hdb_probe_valu(void* stream, int blocks, int iters, float* out, float* ms) {
  g_error.clear();
  if (!out || !ms || blocks < 1 || iters < 1) return fail("invalid probe arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return fail("cannot create events");
  valu_probe_kernel<<<(unsigned)blocks, 256, 0, s>>>(iters, 1.0f, out);  // warm-up
  (void)hipEventRecord(e0, s);
  valu_probe_kernel<<<(unsigned)blocks, 256, 0, s>>>(iters, 1.0f, out);
  (void)hipEventRecord(e1, s);
  int rc = check_launch("valu_probe_kernel");
  if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = fail("probe synchronisation failed");
  if (!rc && hipEventElapsedTime(ms, e0, e1) != hipSuccess) rc = fail("probe timing failed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

size_t hdb_workspace_bytes(int64_t batch, int64_t n, int64_t d, int dtype, int32_t min_samples, int flags) {
  Plan P;
  if (make_plan(batch, n, d, dtype, min_samples, flags, &P)) return 0;
  return P.total;
}

int hdb_cluster(void* stream, const void* x, int64_t batch, int64_t n, int64_t d, int dtype,
                int64_t min_cluster_size, int32_t min_samples, double epsilon, int32_t* labels,
                void* workspace, size_t workspace_bytes, int32_t* status, int flags) {
  g_error.clear();
  if (!x || !labels || !status) return fail("null pointer argument");
  if (min_cluster_size < 2) return fail("min_cluster_size must be >= 2");
  Plan P;
  if (make_plan(batch, n, d, dtype, min_samples, flags, &P)) return -1;
  if (workspace_bytes < P.total || (P.total && !workspace)) return fail("workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned char* ws = static_cast<unsigned char*>(workspace);
  const bool unit = min_samples <= 2;
  const Launch L = dense_launch(batch, n);
  g_timers.count = 0;
  if (dtype == HDB_F32)
    return run_cluster<float>(s, P, static_cast<const float*>(x), nullptr, nullptr, L, d, min_cluster_size,
                              min_samples, unit, epsilon, -100000, labels, ws, status, nullptr, nullptr);
  return run_cluster<double>(s, P, static_cast<const double*>(x), nullptr, nullptr, L, d, min_cluster_size,
                             min_samples, unit, epsilon, -100000, labels, ws, status, nullptr, nullptr);
}

int hdb_cluster_noalloc(void* stream, const void* x, int64_t batch, int64_t n, int64_t d, int dtype,
                        int64_t min_cluster_size, int32_t min_samples, double epsilon, int32_t* labels,
                        int32_t* status, int flags) {
  g_error.clear();
  if (!x || !labels || !status) return fail("null pointer argument");
  if (min_cluster_size < 2) return fail("min_cluster_size must be >= 2");
  Plan P;
  if (make_plan(batch, n, d, dtype, min_samples, flags, &P)) return -1;
  if (P.total > kScratchBytes)
    return fail("hdb_cluster_noalloc: this call needs " + std::to_string(P.total) +
                " bytes of workspace, above its fixed " + std::to_string(kScratchBytes) +
                "-byte scratch; split the batch or call hdb_cluster with a workspace");
  unsigned char* ws = nullptr;
  if (int r = scratch_base(&ws)) return r;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool unit = min_samples <= 2;
  const Launch L = dense_launch(batch, n);
  g_timers.count = 0;
  if (dtype == HDB_F32)
    return run_cluster<float>(s, P, static_cast<const float*>(x), nullptr, nullptr, L, d, min_cluster_size,
                              min_samples, unit, epsilon, -100000, labels, ws, status, nullptr, nullptr);
  return run_cluster<double>(s, P, static_cast<const double*>(x), nullptr, nullptr, L, d, min_cluster_size,
                             min_samples, unit, epsilon, -100000, labels, ws, status, nullptr, nullptr);
}

int hdb_cluster_export(void* stream, const void* x, int64_t batch, int64_t n, int64_t d, int dtype,
                       int64_t min_cluster_size, int32_t min_samples, double epsilon, int sort_depth,
                       int32_t* labels, void* workspace, size_t workspace_bytes, int32_t* status, int flags,
                       int32_t* export_ints, void* export_reals) {
  g_error.clear();
  if (!x || !labels || !status) return fail("null pointer argument");
  if (min_cluster_size < 2) return fail("min_cluster_size must be >= 2");
  if ((export_ints == nullptr) != (export_reals == nullptr)) return fail("export buffers must be given together");
  Plan P;
  if (make_plan(batch, n, d, dtype, min_samples, flags, &P)) return -1;
  if (workspace_bytes < P.total || (P.total && !workspace)) return fail("workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned char* ws = static_cast<unsigned char*>(workspace);
  const bool unit = min_samples <= 2;
  const Launch L = dense_launch(batch, n);
  g_timers.count = 0;
  if (dtype == HDB_F32)
    return run_cluster<float>(s, P, static_cast<const float*>(x), nullptr, nullptr, L, d, min_cluster_size,
                              min_samples, unit, epsilon, sort_depth, labels, ws, status, export_ints, export_reals);
  return run_cluster<double>(s, P, static_cast<const double*>(x), nullptr, nullptr, L, d, min_cluster_size,
                             min_samples, unit, epsilon, sort_depth, labels, ws, status, export_ints, export_reals);
}

int hdb_distances(void* stream, const void* x, int64_t batch, int64_t n, int64_t d, int dtype, void* out,
                  int32_t* status) {
  g_error.clear();
  if (!x || !out || !status) return fail("null pointer argument");
  Plan P;
  if (make_plan(batch, n, d, dtype, 1, 0, &P)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // distance_kernel checks every input coordinate (bit 2) on its diagonal tiles.
  if (dtype == HDB_F32)
    return launch_distances<float>(s, static_cast<const float*>(x), dense_launch(batch, n), d,
                                   static_cast<float*>(out), status);
  return launch_distances<double>(s, static_cast<const double*>(x), dense_launch(batch, n), d,
                                  static_cast<double*>(out), status);
}

int hdb_bounded_distances(void* stream, const void* x, int64_t batch, int64_t n, int64_t d, int dtype,
                          int32_t min_samples, void* out, void* workspace, size_t workspace_bytes,
                          int32_t* status) {
  g_error.clear();
  if (!x || !out || !status) return fail("null pointer argument");
  if (dtype != HDB_F32) return fail("the bounded matrix is a float32 form");
  Plan P;
  if (make_plan(batch, n, d, dtype, min_samples, HDB_FLAG_MATRIX, &P)) return -1;
  if (!P.gram_exact) return fail("these arguments do not take the bounded matrix (min_samples <= 16, d <= 256)");
  if (workspace_bytes < P.total || (P.total && !workspace)) return fail("workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned char* ws = static_cast<unsigned char*>(workspace);
  return run_gram_matrix_only(s, P, static_cast<const float*>(x), batch, n, d, ws, static_cast<float*>(out),
                              status);
}

int hdb_core_distances(void* stream, const void* x, int64_t batch, int64_t n, int64_t d, int dtype,
                       int32_t min_samples, void* core, void* workspace, size_t workspace_bytes,
                       int32_t* status, int flags) {
  g_error.clear();
  if (!x || !core || !status) return fail("null pointer argument");
  // The cluster plan with a materialised matrix lays out the matrix (for
  // min_samples > 64) and the Gram candidate lists (3 <= min_samples <= 16).
  Plan P;
  // HDB_FLAG_NO_GRAM_EXACT: the MFMA-bounded matrix selects core distances
  // inside its own stage; this entry runs the core stage alone.
  const int filter = (flags & HDB_FLAG_PLAIN) ? 0 : HDB_FLAG_FILTER;
  if (make_plan(batch, n, d, dtype, min_samples, flags | HDB_FLAG_MATRIX | HDB_FLAG_NO_GRAM_EXACT | filter, &P))
    return -1;
  P.core = true;
  P.core_topk = min_samples <= 64;
  P.kmax = min_samples <= 8 ? 8 : min_samples <= 16 ? 16 : min_samples <= 32 ? 32 : 64;
  // No matrix is materialised before the core stage here, so the Gram filter
  // takes its per-window form where the clustering path would.
  P.gram_window = P.core_gram && P.gram_dk <= kWinDK && n <= kWinMaxN && !(flags & HDB_FLAG_GRAM_CHUNKED) &&
                  gram_window_lds_bytes((int)n, (int)d) <= device_lds_limit();
  const size_t bn = (size_t)batch * n;
  const size_t need = P.core_gram ? P.total : P.core_topk ? 0 : P.off_matrix + bn * n * P.rs;
  if (workspace_bytes < need || (need && !workspace)) return fail("workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned char* ws = static_cast<unsigned char*>(workspace);
  const Launch L = dense_launch(batch, n);
  // The core kernels check distances (bit 1); the input check (bit 2) is a
  // separate pass here, as no Prim kernel follows.
  const size_t count = (size_t)batch * n * d;
  const unsigned blocks = (unsigned)std::min<size_t>((count + 255) / 256, 4096);
  if (dtype == HDB_F32)
    check_finite_kernel<float><<<blocks, 256, 0, s>>>(static_cast<const float*>(x), count, status);
  else
    check_finite_kernel<double><<<blocks, 256, 0, s>>>(static_cast<const double*>(x), count, status);
  if (int r = check_launch("check_finite_kernel")) return r;
  if (dtype == HDB_F32)
    return launch_core<float>(s, P, static_cast<const float*>(x), L, d, min_samples, static_cast<float*>(core),
                              reinterpret_cast<float*>(ws + P.off_matrix), ws, status, false, true);
  return launch_core<double>(s, P, static_cast<const double*>(x), L, d, min_samples, static_cast<double*>(core),
                             reinterpret_cast<double*>(ws + P.off_matrix), ws, status, false, true);
}

int hdb_ragged_set_tuning(const int32_t* bucket_flags, int count) {
  g_error.clear();
  if (count < 0 || count > kBuckets || (count && !bucket_flags)) return fail("tuning table: 0..7 bucket flags");
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  for (int k = 0; k < kBuckets; ++k) g_tune[k] = k < count ? (bucket_flags[k] & kTuneMask) : 0;
  g_tune_set = count > 0;
  return 0;
}

size_t hdb_ragged_workspace_bytes(const int64_t* offsets, int64_t windows, int64_t d, int dtype,
                                  int32_t min_samples, int flags) {
  g_error.clear();
  RaggedPlan R;
  if (make_ragged_plan(offsets, windows, d, dtype, min_samples, flags, &R)) return 0;
  return R.total;
}

int hdb_cluster_ragged(void* stream, const void* x, const int64_t* offsets, const int64_t* offsets_dev,
                       int64_t windows, int64_t d, int dtype, int64_t min_cluster_size, int32_t min_samples,
                       double epsilon, int32_t* labels, void* workspace, size_t workspace_bytes, int32_t* status,
                       int flags) {
  g_error.clear();
  if (!x || !labels || !status || !offsets_dev) return fail("null pointer argument");
  if (min_cluster_size < 2) return fail("min_cluster_size must be >= 2");
  RaggedPlan R;
  if (make_ragged_plan(offsets, windows, d, dtype, min_samples, flags, &R)) return -1;
  if (workspace_bytes < R.total || !workspace) return fail("workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned char* ws = static_cast<unsigned char*>(workspace);
  const bool graph = !(flags & HDB_FLAG_NO_GRAPH);
  if (dtype == HDB_F32) {
    const float* xf = static_cast<const float*>(x);
    return graph ? run_ragged_graph<float>(s, R, xf, offsets, offsets_dev, d, min_cluster_size, min_samples,
                                           epsilon, labels, ws, status, flags)
                 : run_ragged<float>(s, R, xf, offsets_dev, d, min_cluster_size, min_samples, epsilon, labels, ws,
                                     status);
  }
  const double* xd = static_cast<const double*>(x);
  return graph ? run_ragged_graph<double>(s, R, xd, offsets, offsets_dev, d, min_cluster_size, min_samples,
                                          epsilon, labels, ws, status, flags)
               : run_ragged<double>(s, R, xd, offsets_dev, d, min_cluster_size, min_samples, epsilon, labels, ws,
                                    status);
}

int hdb_cluster_matrix(void* stream, const void* distances, const void* core, int64_t batch, int64_t n, int dtype,
                       int64_t min_cluster_size, int unit_core, double epsilon, int sort_depth, int32_t* labels,
                       void* workspace, size_t workspace_bytes, int32_t* status, int flags, int32_t* export_ints,
                       void* export_reals) {
  g_error.clear();
  if (!distances || !labels || !status) return fail("null pointer argument");
  if (!unit_core && !core) return fail("core distances required unless unit_core");
  if (min_cluster_size < 2) return fail("min_cluster_size must be >= 2");
  if ((export_ints == nullptr) != (export_reals == nullptr)) return fail("export buffers must be given together");
  Plan P;
  if (make_plan(batch, n, 1, dtype, unit_core ? 1 : 3, flags | HDB_FLAG_MATRIX | HDB_FLAG_MATRIX_INPUT, &P)) return -1;
  if (workspace_bytes < P.total || !workspace) return fail("workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned char* ws = static_cast<unsigned char*>(workspace);
  const Launch L = dense_launch(batch, n);
  g_timers.count = 0;
  if (dtype == HDB_F32)
    return run_cluster<float>(s, P, nullptr, static_cast<const float*>(distances), static_cast<const float*>(core),
                              L, 1, min_cluster_size, unit_core ? 1 : 3, unit_core != 0, epsilon, sort_depth,
                              labels, ws, status, export_ints, export_reals);
  return run_cluster<double>(s, P, nullptr, static_cast<const double*>(distances), static_cast<const double*>(core),
                             L, 1, min_cluster_size, unit_core ? 1 : 3, unit_core != 0, epsilon, sort_depth,
                             labels, ws, status, export_ints, export_reals);
}

}  // extern "C"
