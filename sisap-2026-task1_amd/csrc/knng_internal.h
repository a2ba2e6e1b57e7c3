// knng_internal.h -- host-side context of libknng.so shared by the build
// (knng.hip) and serving (knng_serve.hip) translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

#include "knng.h"
#include "knng_device.h"

using knng_dev::Node;


struct knng_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  knng_params p{};
  int64_t n = 0;
  int32_t d = 0, words = 0;
  int num_cu = 256;
  std::string err;

  float* x = nullptr;          // owned fp32 rows (keep_fp32)
  const float* x_ext = nullptr;
  int8_t* q = nullptr;
  uint64_t* sk = nullptr;
  int64_t ingested = 0;

  // forest
  int32_t *perm = nullptr, *perm2 = nullptr, *nodepos = nullptr, *nodepos2 = nullptr;
  int32_t *flags = nullptr, *scan = nullptr, *blocksums = nullptr;
  Node *nodes = nullptr, *nodes2 = nullptr;
  int4* leaves = nullptr;
  Node* hier = nullptr;        // every level's nodes, global child links
  int64_t hier_count = 0;
  // cells: spherical k-means partitions of the int8 rows that seed the trees
  int32_t ncells = 0, nparts = 0;   // ncells: the largest partition's cell count (the stride)
  std::vector<int32_t> cell_count;  // cells of each partition
  int8_t* cz = nullptr;        // [nparts][ncells][d] int8 centroid directions
  float* cw = nullptr;         // [nparts][ncells] 1 / |centroid|
  int32_t* cassign = nullptr;  // [nparts][n] cell of every row
  int32_t* cell_root = nullptr;   // [T][ncells]: subtree root (node >= 0, leaf -(l+1), empty INT32_MIN)
  int32_t* dcount = nullptr;   // [0] next nodes, [1] leaves, [2] overflow, [3] scan total
  int64_t node_cap = 0, leaf_cap = 0, nleaves = 0;
  bool have_forest = false;

  // graph
  int32_t *gid[2] = {nullptr, nullptr}, *gsc[2] = {nullptr, nullptr};
  uint64_t* fresh[2] = {nullptr, nullptr};   // per row: bit k = entry k entered last iteration
  bool fresh_all = false;                    // fresh[cur] is all ones (no join iteration yet)
  int32_t* ovf = nullptr;                    // rows passed on to k_join [n]
  int32_t* jcbuf = nullptr;                  // phased join: candidates per work position [n][kJCap]
  int32_t* jccount = nullptr;                // their counts (-1: the row went to k_join) [n]
  unsigned long long* ovf_n = nullptr;       // their count
  int cur = 0;
  bool have_graph = false;
  // leaf of every row in every tree ([n][leaf_tp], first perm position of the
  // leaf), valid while the graph descends from this forest's leaf init
  int32_t* leafkey = nullptr;
  int leaf_tp = 4;
  bool leaf_valid = false;
  uint8_t* gmask[2] = {nullptr, nullptr};    // per list entry: bit t = row and entry share tree t's leaf
  // reverse lists (reverse_width R > 0): packed (score, id) top-R slots [n][R]
  // (all zero between iterations), and the lists built from them
  unsigned long long* rv_slots = nullptr;
  int32_t* rv_id = nullptr;      // [n][R], -1 past the last entry
  uint8_t* rv_mask = nullptr;    // [n][R] mask bytes of the edges u -> v
  unsigned long long *counters = nullptr, *csum = nullptr;

  // ingest staging
  void* pin[2] = {nullptr, nullptr};
  float* dstage[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  size_t stage_bytes = 64ull << 20;

  // export scratch
  int32_t* ex_knns = nullptr;
  float* ex_dists = nullptr;
  int64_t ex_cap = 0;
  float* rr_part = nullptr;   // sliced FP32 rerank: running sums [n][R], R <= k_build
  unsigned int* rr_q = nullptr;   // FP32 rerank: per-pass, per-XCD row-queue counters

  // independent builds packed into one context (knng_build_batch): segment s
  // is rows [seg_host[s], seg_host[s+1]); trees, leaves and edges never
  // cross a segment, and each segment's lists are exactly those of a build
  // of its rows alone
  int32_t nseg = 0;
  std::vector<int64_t> seg_host;
  int64_t* seg_base = nullptr;   // device copy of seg_host
  int64_t* tune_n = nullptr;     // device scalar: rows of an autotuning launch
  int join_variant = -1;         // k_join shape in use (-1: not chosen yet)
  int leaf_variant = -1;         // leaf-tile shape for leaves <= 64 rows (-1: not chosen yet)

  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  knng_stats st{};
};

namespace knng_host {

inline int fail(knng_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                 \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(ctx, KNNG_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

template <typename T>
inline int dalloc(knng_ctx* c, T** ptr, size_t count, const char* what) {
  if (*ptr) return 0;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(T) + 256);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(c, KNNG_ENOMEM, std::string("hipMalloc failed for ") + what + " (" +
                                     std::to_string(count * sizeof(T)) + " B): " +
                                     hipGetErrorString(e));
  }
  return 0;
}

template <typename T>
void dfree(T** ptr) {
  if (*ptr) (void)hipFree(reinterpret_cast<void*>(const_cast<typename std::remove_const<T>::type*>(*ptr)));
  *ptr = nullptr;
}

// A device buffer owned by one scope: freed (dfree) when the scope ends.
template <typename T>
struct ScopedDev {
  T* p = nullptr;
  ScopedDev() = default;
  ScopedDev(const ScopedDev&) = delete;
  ScopedDev& operator=(const ScopedDev&) = delete;
  ~ScopedDev() { dfree(&p); }
};

// Stage tracing: with KNNG_TRACE=1 in the environment, trace_mark(c, label)
// waits for the context's stream and prints the wall time since the previous
// mark to stderr.  Off (and free) otherwise.
struct TraceClock {
  bool on = false;
  std::mutex mu;   // marks may come from several host threads (the k-means partitions)
  std::chrono::steady_clock::time_point last;
};

inline TraceClock& trace_clock() {
  static TraceClock tc;
  static const bool init = [] {
    const char* e = std::getenv("KNNG_TRACE");
    tc.on = e && e[0] == '1';
    tc.last = std::chrono::steady_clock::now();
    return true;
  }();
  (void)init;
  return tc;
}

inline void trace_mark(knng_ctx* c, const char* label) {
  TraceClock& tc = trace_clock();
  if (!tc.on) return;
  (void)hipStreamSynchronize(c->stream);
  std::lock_guard<std::mutex> lk(tc.mu);
  const auto now = std::chrono::steady_clock::now();
  std::fprintf(stderr, "[knng] %-28s %9.2f ms\n", label,
               std::chrono::duration<double, std::milli>(now - tc.last).count());
  tc.last = now;
}

inline int grid_for(const knng_ctx* c, int64_t work, int per_block, int blocks_per_cu) {
  int64_t g = (work + per_block - 1) / per_block;
  const int64_t cap = int64_t(c->num_cu) * blocks_per_cu;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return int(g);
}

// k_rerank_slice's grid: four rows per block, a multiple of 8 blocks (the
// kernel gives each XCD its own eighth of the rows)
inline int rerank_grid(const knng_ctx* c, int64_t n) { return (grid_for(c, n, 4, 16) + 7) & ~7; }

// Cells of the k-means partition for n rows (0: plain RP trees).
inline int32_t cells_for(const knng_params& p, int64_t n) {
  if (p.cells >= 0) return p.cells;
  const int64_t c = n / 1536;
  return c >= 2 ? int32_t(c) : 0;
}

// Packed independent builds (knng.hip).  set_segments: the segment bases of
// a fresh context (seg[0] = 0, seg[nseg] = n, every segment >= 2 rows, no
// cells).  build_packed: forest + init + join for every segment at once, each
// with its own stop rule; iters[s] = the segment's join iterations, cur[s] =
// the graph buffer with its final lists.  use_graph: each join iteration is
// one hipGraph launch (captured once per buffer direction).
// export_segment: the best k of segment s into device buffers, ids relative
// to the segment.
int set_segments(knng_ctx* c, const std::vector<int64_t>& seg);
int build_packed(knng_ctx* c, std::vector<int32_t>* iters, std::vector<int32_t>* cur, bool use_graph,
                 bool* graph_used);
int export_segment(knng_ctx* c, int s, int cur, int k, int rerank, int32_t* dk, float* dd);

// K1 on an arbitrary stream (defined in knng.hip); xcopy (optional) also
// receives the float32 rows
int launch_quantize(hipStream_t stream, const float* dx, int64_t nrows, int d, int8_t* q,
                    uint64_t* sk, int num_cu, float* xcopy = nullptr);
// route int8 rows to cells: out[j] = argmax_c float(<q_r, Z_c>) * w[c] for the rows
// r = row_off + j * row_stride (knng.hip); exclude_self skips c == j, score gets the best value.
// With partial arrays pbest/pidx of part_cap entries each, a launch with few rows splits the
// cells over several blocks per row tile (same result).
int launch_assign(hipStream_t stream, const int8_t* q, int64_t nrows, int d, const int8_t* Z,
                  const float* w, int C, int32_t* out, int num_cu, int64_t row_off = 0,
                  int64_t row_stride = 1, bool exclude_self = false, float* score = nullptr,
                  float* pbest = nullptr, int32_t* pidx = nullptr, int64_t part_cap = 0);
// partial-array entries a split launch_assign over nrows rows and C cells can use
inline int64_t assign_part_entries(int64_t nrows, int num_cu) {
  return nrows * std::max<int64_t>(1, (2 * int64_t(num_cu) * 128 + nrows - 1) / nrows);
}

}  // namespace knng_host
