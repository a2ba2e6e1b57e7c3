// knng_device.h -- device-side building blocks shared by the build kernels
// (knng.hip) and the search kernels (knng_serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

namespace knng_dev {

constexpr int kMaxDim = 1024;
constexpr int kMaxDepth = 64;           // forest depth at which a node is cut into leaves
constexpr int kMaxK = 64;
constexpr int kMaxTrees = 64;
constexpr int kLeafSlots = 64;          // candidate slots per (point, tree)
constexpr int kMaxLeaf = 2048;          // largest leaf (leaves above kLeafSlots keep 64 mates)
constexpr int kNumCounters = 1024;      // spread of the per-block atomic counters
// fields per counter: changed slots, unique candidates, re-scored candidates,
// inserted ids, candidates whose sketches were gathered, 2-hop paths dropped
// by the leaf-mate rule
constexpr int kCounterFields = 6;
constexpr int kMaskTrees = 8;           // trees whose leaves a list entry's mask byte records
constexpr int kRowQueueWords = 8 * 32;  // k_join row-queue counters, one 256 B line per XCD
constexpr float kFltMin = 1.17549435e-38f;

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m));
  return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m));
  return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

__device__ __forceinline__ int dot4_i8(int a, int b, int c) {
  return __builtin_amdgcn_sdot4(a, b, c, false);
}

__device__ __forceinline__ int dot16_i8(const v4i& a, const v4i& b, int c) {
  c = dot4_i8(a.x, b.x, c);
  c = dot4_i8(a.y, b.y, c);
  c = dot4_i8(a.z, b.z, c);
  return dot4_i8(a.w, b.w, c);
}

// Mask byte of a list entry: bit t (trees 0..7) = the row and the entry share
// tree t's leaf.  own[t] = the row's leaf keys, lk = the entry's leaf-key row.
__device__ __forceinline__ int32_t edge_mask(const int32_t* own, const int32_t* __restrict__ lk, int T) {
  int32_t m = 0;
  const int tm = T < kMaskTrees ? T : kMaskTrees;
  for (int t = 0; t < tm; ++t) m |= int32_t(lk[t] == own[t]) << t;
  return m;
}

// (score desc, id asc) as a descending u64 key; 0 is "no candidate".
__device__ __host__ __forceinline__ uint64_t pack_key(int32_t score, int32_t id) {
  return (uint64_t(uint32_t(score) ^ 0x80000000u) << 32) | uint64_t(0xFFFFFFFFu - uint32_t(id));
}
__device__ __host__ __forceinline__ int32_t key_score(uint64_t k) {
  return int32_t(uint32_t(k >> 32) ^ 0x80000000u);
}
__device__ __host__ __forceinline__ int32_t key_id(uint64_t k) {
  return int32_t(0xFFFFFFFFu - uint32_t(k));
}

// xorshift32 step of the reference's init padding (src/main.cpp:273-275).
__device__ __host__ __forceinline__ uint32_t xorshift32(uint32_t s) {
  s ^= s << 13;
  s ^= s >> 17;
  s ^= s << 5;
  return s;
}

// splitmix64 finaliser (Steele, Lea, Flood, OOPSLA 2014): node pivot hash.
__device__ __host__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Pivot positions of node (tree, depth, start, size): p1 in [0,size),
// p2 in [0,size)\{p1}.
__device__ __host__ __forceinline__ void node_pivots(uint64_t seed, int tree, int depth,
                                                     int64_t start, int64_t size, int32_t* p1,
                                                     int32_t* p2) {
  const uint64_t key =
      (uint64_t(uint32_t(tree)) << 48) ^ (uint64_t(uint32_t(depth)) << 40) ^ uint64_t(start);
  const uint64_t h1 = mix64(seed ^ mix64(key));
  const uint64_t h2 = mix64(h1);
  const int64_t a = int64_t(h1 % uint64_t(size));
  int64_t b = int64_t(h2 % uint64_t(size - 1));
  if (b >= a) ++b;
  *p1 = int32_t(a);
  *p2 = int32_t(b);
}

// ---------------------------------------------------------------------------
// quantiser (src/main.cpp:179-192)
// ---------------------------------------------------------------------------


// src/main.cpp:184-186: multiply, min(127), max(-128); x86 DAZ (the reference
// is built with -ffast-math) makes a denormal input read as zero.
__device__ __forceinline__ float q_clamp(float v) {
  v = (fabsf(v) < kFltMin) ? 0.0f : v;
  const float s = v * 127.0f;
  const float hi = (s < 127.0f) ? s : 127.0f;
  return (hi > -128.0f) ? hi : -128.0f;
}


// forest node of one level: a contiguous segment of one tree's perm
// Children: >= 0 is a node of the next level (or, once copied into the
// persistent hierarchy, a global node index); < 0 is leaf -(child+1).
struct Node {
  int32_t tree, start, size, p1, p2;   // p1/p2: pivot positions inside the segment
  int32_t left, right, nleft;
  int32_t p1id, p2id;                  // pivot row ids
  int32_t base;   // first row of the node's build when independent builds share one forest
                  // (pivots are drawn from start - base, as in that build alone); else 0
};

// Build (segment) of row r when independent builds are packed into one
// context: the largest s with seg_base[s] <= r (seg_base ascending).
__device__ __host__ __forceinline__ int seg_of(const int64_t* seg_base, int nseg, int64_t r) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg_base[mid] <= r)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  int off = 0;
  const int nw = blockDim.x >> 6;
  for (int j = 0; j < w; ++j) off += lds[j];
  int tot = 0;
  for (int j = 0; j < nw; ++j) tot += lds[j];
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// open-addressing set of int32 keys in LDS (-1 = empty); true if newly inserted
__device__ __forceinline__ bool hs_insert(int32_t* hs, int bits, int32_t key) {
  const uint32_t mask = (1u << bits) - 1u;
  uint32_t h = (uint32_t(key) * 2654435761u) >> (32 - bits);
  while (true) {
    const int32_t prev = atomicCAS(&hs[h], -1, key);
    if (prev == -1) return true;
    if (prev == key) return false;
    h = (h + 1u) & mask;
  }
}

// Wave-aggregated append: lanes with pred get consecutive slots from *counter
// (one LDS atomic per wave).  Every lane of the wave must call it.
__device__ __forceinline__ int wave_append(bool pred, int* counter) {
  const unsigned long long m = __ballot(pred);
  if (m == 0ull) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll(m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader);
  return pred ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// a before b in (score desc, id asc, slot asc)
__device__ __forceinline__ bool before(int32_t sa, int32_t ia, int32_t la, int32_t sb, int32_t ib,
                                       int32_t lb) {
  if (sa != sb) return sa > sb;
  if (ia != ib) return ia < ib;
  return la < lb;
}

}  // namespace knng_dev
