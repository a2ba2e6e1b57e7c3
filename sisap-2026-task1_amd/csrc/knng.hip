// knng.hip -- MI355X (gfx950) kNN-graph builder: HIP kernels + the C-ABI of
// include/knng.h.
//
// Pipeline (one context = one GPU, one stream), replacing src/main.cpp:143-457
// of the reference engine:
//   K1 k_quantize        fused float32 -> int8 (truncating) + 1-bit sign sketch
//   K2 k_forest_*        level-synchronous two-pivot RP forest, all trees at once
//   K3 k_leaf_gram       leaf brute force as int8 MFMA tiles (exact int32)
//   K4 k_init_merge      per-point top-K of all leaf-mates, (score desc, id asc)
//   K5-7 k_join          pull-style 2-hop join: LDS hash dedup, sketch popcount
//                        prefilter, gathered int8 dots, rank merge, change count
//   K8 k_export_*        plain export or FP32 rerank of each row's list
//
// Data layout in HBM: q int8 [n][d] (row = d bytes, d % 64 == 0), sketch u64
// [n][d/64] (bit b of word w <=> dim 64w+b), graph as SoA int32 ids[n][K] and
// scores[n][K], double-buffered (Jacobi iterations).  Forest state is
// perm[T][n] (point ids, each tree's leaves contiguous) with a per-position
// node index.

#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>
#include <limits>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "knng.h"
#include "knng_device.h"
#include "knng_internal.h"

using namespace knng_dev;


namespace {

// ---------------------------------------------------------------------------
// K1: quantise + sketch.  One wave per row; lane l owns dims 256g+4l..+3, so a
// wave reads 1 KB of float32 per step and the four ballots of one step hold
// sketch words 4g..4g+3, interleaved by 4 (bit 4i+j <- lane i, element j).
// ---------------------------------------------------------------------------

// xcopy (null: none) also receives the float32 rows: a server's small launch
// quantises straight from pinned host staging and keeps a device copy of the
// rows for the search's FP32 rerank, in place of a separate upload
__global__ __launch_bounds__(256) void k_quantize(const float* __restrict__ x, int64_t nrows,
                                                  int d, int8_t* __restrict__ q,
                                                  uint64_t* __restrict__ sk, float* __restrict__ xcopy) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  const int words = d >> 6;
  for (int64_t r = wave; r < nrows; r += nwaves) {
    const float* xr = x + r * d;
    int8_t* qr = q + r * d;
    uint64_t mine = 0;
    // the row's four 1 KB quarters (d <= 1024) loaded before any is used
    float4 vq[kMaxDim / 256];
#pragma unroll
    for (int g = 0; g < kMaxDim / 256; ++g) {
      const int col = g * 256 + lane * 4;
      vq[g] = col < d ? *reinterpret_cast<const float4*>(xr + col) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (xcopy) {
#pragma unroll
      for (int g = 0; g < kMaxDim / 256; ++g) {
        const int col = g * 256 + lane * 4;
        if (col < d) *reinterpret_cast<float4*>(xcopy + r * d + col) = vq[g];
      }
    }
#pragma unroll
    for (int g = 0; g < kMaxDim / 256; ++g) {
      if (g * 256 >= d) break;
      const int col = g * 256 + lane * 4;
      const bool ok = col < d;
      const float4 v = vq[g];
      const float c0 = q_clamp(v.x), c1 = q_clamp(v.y), c2 = q_clamp(v.z), c3 = q_clamp(v.w);
      const uint32_t packed = uint32_t(uint8_t(int8_t(int(c0)))) |
                              (uint32_t(uint8_t(int8_t(int(c1)))) << 8) |
                              (uint32_t(uint8_t(int8_t(int(c2)))) << 16) |
                              (uint32_t(uint8_t(int8_t(int(c3)))) << 24);
      if (ok) *reinterpret_cast<uint32_t*>(qr + col) = packed;
      // sketch word 4g + qd covers dims 256g + 64qd .. + 63, held by lanes
      // 16qd .. 16qd + 15: each lane's four sign bits at bit 4 (lane % 16),
      // OR-reduced over its 16-lane group on the vector ALU (the scalar unit
      // a CU's four SIMDs share would otherwise bound the kernel)
      const uint32_t nib = ok ? uint32_t(c0 > 0.0f) | (uint32_t(c1 > 0.0f) << 1) | (uint32_t(c2 > 0.0f) << 2) |
                                    (uint32_t(c3 > 0.0f) << 3)
                              : 0u;
      const int sh = 4 * (lane & 15);
      uint32_t lo = sh < 32 ? nib << sh : 0u, hi = sh >= 32 ? nib << (sh - 32) : 0u;
#pragma unroll
      for (int m = 8; m >= 1; m >>= 1) {
        lo |= __shfl_xor(lo, m, 16);
        hi |= __shfl_xor(hi, m, 16);
      }
      // lane 4g + qd takes group qd's word
      const uint32_t wlo = __shfl(lo, 16 * (lane & 3)), whi = __shfl(hi, 16 * (lane & 3));
      if ((lane >> 2) == g) mine = (uint64_t(whi) << 32) | wlo;
    }
    if (lane < words) sk[r * words + lane] = mine;
  }
}

// ---------------------------------------------------------------------------
// K2: RP forest, one level of every tree per pass.
// ---------------------------------------------------------------------------


__global__ void k_forest_seed(int64_t n, int T, int leaf, uint64_t seed, int32_t* perm,
                              int32_t* nodepos, Node* nodes) {
  const int64_t total = int64_t(T) * n;
  for (int64_t gp = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; gp < total;
       gp += int64_t(gridDim.x) * blockDim.x) {
    const int t = int(gp / n);
    perm[gp] = int32_t(gp - int64_t(t) * n);
    nodepos[gp] = t;
    if (gp - int64_t(t) * n == 0) {
      Node nd;
      nd.tree = t;
      nd.start = 0;
      nd.size = int32_t(n);
      node_pivots(seed, t, 0, 0, n, &nd.p1, &nd.p2);
      nd.left = nd.right = -1;
      nd.nleft = 0;
      nd.base = 0;
      nodes[t] = nd;
    }
  }
}

// Independent builds packed into one forest (segments [seg_base[s],
// seg_base[s+1]) of the rows): perm = identity and every position points at
// its segment's root node in its tree, segnode[t * nseg + s] (-1 where the
// segment is a single leaf).
__global__ void k_forest_seed_segs(int64_t n, int T, const int64_t* __restrict__ seg_base, int nseg,
                                   const int32_t* __restrict__ segnode, int32_t* __restrict__ perm,
                                   int32_t* __restrict__ nodepos) {
  const int64_t total = int64_t(T) * n;
  for (int64_t gp = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; gp < total;
       gp += int64_t(gridDim.x) * blockDim.x) {
    const int t = int(gp / n);
    const int64_t r = gp - int64_t(t) * n;
    perm[gp] = int32_t(r);
    nodepos[gp] = segnode[int64_t(t) * nseg + seg_of(seg_base, nseg, r)];
  }
}

// side flag of every position in an active node: 1 = left.  16 lanes per
// position (a 1 KB row each, four rows per wave in flight); pivot rows stay
// in registers while consecutive positions share a node.
template <int CPL>   // CPL 16-byte chunks per lane of a 64-lane row (d <= 1024 CPL)
__global__ __launch_bounds__(256) void k_forest_flags(const int8_t* __restrict__ q, int d,
                                                      int64_t n, int64_t total,
                                                      const int32_t* __restrict__ perm,
                                                      const int32_t* __restrict__ nodepos,
                                                      const Node* __restrict__ nodes,
                                                      int32_t* __restrict__ flags, int run) {
  // A wave takes `run` consecutive positions; its four 16-lane groups take
  // every fourth of them, so four rows load at once, each group keeping its
  // node's two pivot rows while consecutive positions stay in that node.
  constexpr int CH = 4 * CPL;   // 16-byte chunks per lane of a 16-lane group
  const int lane = threadIdx.x & 63, g = lane >> 4, s16 = lane & 15;
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  const int nch = d >> 4;
  v4i pa[CH], pb[CH];
  for (int64_t base = wave * run; base < total; base += nwaves * run) {
    int cur = -1;
    int64_t seg0 = 0;
    int32_t p1 = 0, p2 = 0;
    const int64_t end = base + run < total ? base + run : total;
    for (int64_t g0 = base; g0 < end; g0 += 4) {
      const int64_t gp = g0 + g;
      const int nd = gp < end ? nodepos[gp] : -1;
      if (nd >= 0 && nd != cur) {
        const Node node = nodes[nd];
        seg0 = int64_t(node.tree) * n + node.start;
        p1 = node.p1;
        p2 = node.p2;
        const int8_t* ra = q + int64_t(perm[seg0 + p1]) * d;
        const int8_t* rb = q + int64_t(perm[seg0 + p2]) * d;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int ch = s16 + 16 * c;
          pa[c] = ch < nch ? *reinterpret_cast<const v4i*>(ra + 16 * ch) : v4i{0, 0, 0, 0};
          pb[c] = ch < nch ? *reinterpret_cast<const v4i*>(rb + 16 * ch) : v4i{0, 0, 0, 0};
        }
        cur = nd;
      }
      const int64_t rel = gp - seg0;
      const bool dot = nd >= 0 && rel != p1 && rel != p2;
      int s = 0;
      if (dot) {
        const int8_t* rx = q + int64_t(perm[gp]) * d;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int ch = s16 + 16 * c;
          if (ch < nch) {
            const v4i v = *reinterpret_cast<const v4i*>(rx + 16 * ch);
            s = dot16_i8(v, pa[c], s);
            s -= dot16_i8(v, pb[c], 0);
          }
        }
      }
#pragma unroll
      for (int m = 8; m >= 1; m >>= 1) s += __shfl_xor(s, m, 16);
      // <x,p1> > <x,p2> goes left (1), ties go right; the pivots themselves
      // go left (p1) and right (p2); positions outside active nodes get 0
      if (s16 == 0 && gp < end) flags[gp] = nd < 0 ? 0 : (rel == p1 ? 1 : (rel == p2 ? 0 : (s > 0 ? 1 : 0)));
    }
  }
}

// exclusive scan over int32 (tile = 256 threads x 16 elements)
constexpr int kScanTile = 4096;


__global__ __launch_bounds__(256) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                     int32_t* __restrict__ sums) {
  __shared__ int lds[8];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + threadIdx.x * 16;
  int s = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (base + j < n) s += in[base + j];
  int tot;
  block_exclusive_scan(s, lds, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_top(int32_t* sums, int64_t nb, int32_t* out_total) {
  __shared__ int lds[16];
  int carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += blockDim.x) {
    const int64_t j = b0 + threadIdx.x;
    const int v = j < nb ? sums[j] : 0;
    int tot;
    const int ex = block_exclusive_scan(v, lds, &tot);
    if (j < nb) sums[j] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(256) void k_scan_down(const int32_t* __restrict__ in, int64_t n,
                                                   const int32_t* __restrict__ sums,
                                                   int32_t* __restrict__ out) {
  __shared__ int lds[8];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + threadIdx.x * 16;
  int v[16];
  int s = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    v[j] = base + j < n ? in[base + j] : 0;
    s += v[j];
  }
  int tot;
  int run = block_exclusive_scan(s, lds, &tot) + sums[blockIdx.x];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
}

// Children of every active node; children larger than a leaf become the next
// level's nodes, the rest are appended to the leaf table.
__global__ void k_forest_nodes(int64_t n, int leaf, uint64_t seed, int depth, Node* nodes,
                               int nnodes, const int32_t* __restrict__ S,
                               const int32_t* __restrict__ perm, Node* next,
                               int32_t* next_count, int4* leaves, int32_t* leaf_count,
                               int64_t leaf_cap, int32_t* overflow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  Node nd = nodes[i];
  const int64_t gs = int64_t(nd.tree) * n + nd.start;
  nodes[i].p1id = perm[gs + nd.p1];
  nodes[i].p2id = perm[gs + nd.p2];
  const int32_t nl = S[gs + nd.size] - S[gs];
  const int32_t cs[2] = {nd.start, nd.start + nl};
  const int32_t cz[2] = {nl, nd.size - nl};
  int32_t child[2];
  for (int c = 0; c < 2; ++c) {
    if (cz[c] > leaf && depth + 1 >= kMaxDepth) {
      // degenerate input (e.g. many identical rows, whose ties always go
      // right): cut the child into consecutive leaves of <= leaf rows
      const int nch = (cz[c] + leaf - 1) / leaf;
      const int id = atomicAdd(leaf_count, nch);
      for (int k = 0; k < nch; ++k) {
        const int32_t s0 = cs[c] + k * leaf;
        const int32_t z = (cz[c] - k * leaf) < leaf ? (cz[c] - k * leaf) : leaf;
        if (id + k < leaf_cap)
          leaves[id + k] = make_int4(nd.tree, s0, z, 0);
        else
          *overflow = 1;
      }
      child[c] = -(id + 1);
    } else if (cz[c] > leaf) {
      const int id = atomicAdd(next_count, 1);
      Node ch;
      ch.tree = nd.tree;
      ch.start = cs[c];
      ch.size = cz[c];
      node_pivots(seed, nd.tree, depth + 1, cs[c] - nd.base, cz[c], &ch.p1, &ch.p2);
      ch.left = ch.right = -1;
      ch.nleft = 0;
      ch.base = nd.base;
      next[id] = ch;
      child[c] = id;
    } else {
      const int id = atomicAdd(leaf_count, 1);
      if (id < leaf_cap)
        leaves[id] = make_int4(nd.tree, cs[c], cz[c], 0);
      else
        *overflow = 1;
      child[c] = -(id + 1);
    }
  }
  nodes[i].left = child[0];
  nodes[i].right = child[1];
  nodes[i].nleft = nl;
}

// stable partition of each active node by its flags
__global__ __launch_bounds__(256) void k_forest_partition(
    int64_t n, int64_t total, const int32_t* __restrict__ perm, const int32_t* __restrict__ nodepos,
    const Node* __restrict__ nodes, const int32_t* __restrict__ flags,
    const int32_t* __restrict__ S, int32_t* __restrict__ perm2, int32_t* __restrict__ nodepos2) {
  for (int64_t gp = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; gp < total;
       gp += int64_t(gridDim.x) * blockDim.x) {
    const int nd = nodepos[gp];
    if (nd < 0) {
      perm2[gp] = perm[gp];
      nodepos2[gp] = -1;
      continue;
    }
    const Node node = nodes[nd];
    const int64_t gs = int64_t(node.tree) * n + node.start;
    const int32_t lr = S[gp] - S[gs];
    int64_t np;
    int32_t child;
    if (flags[gp]) {
      np = gs + lr;
      child = node.left;
    } else {
      np = gs + node.nleft + (gp - gs - lr);
      child = node.right;
    }
    perm2[np] = perm[gp];
    nodepos2[np] = child;
  }
}

// leafkey[r * TP + t] = first perm position of row r's leaf in tree t, one
// wave per leaf (every row lies in exactly one leaf of every tree)
__global__ __launch_bounds__(256) void k_leaf_keys(const int4* __restrict__ leaves, int64_t nleaves,
                                                   const int32_t* __restrict__ perm, int64_t n, int TP,
                                                   int32_t* __restrict__ leafkey) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  for (int64_t l = wave; l < nleaves; l += nwaves) {
    const int4 lf = leaves[l];
    const int32_t* mem = perm + int64_t(lf.x) * n + lf.y;
    for (int j = lane; j < lf.z; j += 64) leafkey[int64_t(mem[j]) * TP + lf.x] = lf.y;
  }
}

// the mask byte of every list entry of the graph gid (see edge_mask): one
// wave per row, lane k for entry k (K <= 64), rows in `order` (tree 0's leaf
// order; null: index order) with XCD x = block % 8 taking the contiguous
// eighth x of it, so the rows in flight on an XCD are cluster-mates whose
// neighbours' leaf keys share its L2
__global__ __launch_bounds__(256) void k_edge_masks(const int32_t* __restrict__ gid, int64_t n, int K,
                                                    const int32_t* __restrict__ leafkey, int T, int TP,
                                                    uint8_t* __restrict__ gmask,
                                                    const int32_t* __restrict__ order) {
  const int lane = threadIdx.x & 63;
  const int xcd = blockIdx.x & 7;
  const int64_t hi = n * (xcd + 1) / 8, step = int64_t(gridDim.x >> 3) * 4;
  for (int64_t j = n * xcd / 8 + int64_t(blockIdx.x >> 3) * 4 + (threadIdx.x >> 6); j < hi; j += step) {
    const int64_t r = order ? int64_t(order[j]) : j;
    if (lane < K)
      gmask[r * K + lane] =
          uint8_t(edge_mask(leafkey + r * TP, leafkey + int64_t(gid[r * K + lane]) * TP, T));
  }
}

// Append one level's nodes to the persistent hierarchy, turning next-level
// child indices into global ones (the next level is appended right after).
__global__ void k_hier_append(const Node* __restrict__ nodes, int nnodes, Node* __restrict__ hier,
                              int64_t next_base) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  Node nd = nodes[i];
  if (nd.left >= 0) nd.left = int32_t(next_base + nd.left);
  if (nd.right >= 0) nd.right = int32_t(next_base + nd.right);
  hier[i] = nd;
}

// ---------------------------------------------------------------------------
// Cells (spherical k-means on the int8 rows), used to seed the forest.
// k_assign: a[j] = argmax_c float(<q_r, Z_c>) * w[c] for the rows
// r = row_off + j * row_stride, ties to the smaller c; EXCL skips c == j
// (rows and centroids are the same set); score[j] (if given) = the best value.
// One block = 128 rows, wave v owns rows 32v..32v+31 and keeps their A
// fragments for the whole pass in registers.  32-cell centroid tiles are
// double-buffered in LDS, filled by direct global->LDS loads (16 bytes per
// lane, 64 lanes = 1 KB contiguous) while the MFMAs run on the other one.
// At d = 1024 the k-step count is a template constant, so the 32 MFMAs and
// their LDS fragment reads form one straight-line block the scheduler can
// interleave (a runtime bound puts each MFMA behind its own branch).  Row r of a tile holds its 16-byte chunk c at slot
// (c - r) mod (d/16), so the 32 rows a B fragment read touches fall on
// different banks.  Even and odd k-steps accumulate into two chains
// (exact int32 sums, added at the end) to keep two MFMAs in flight.
// With gridDim.y = S > 1 the cell tiles are split into S consecutive slices,
// block (x, y) scans slice y only and writes its per-row best to the partial
// arrays pbest/pidx [S][nrows]; k_assign_reduce then takes the best over the
// slices (larger value, then smaller cell), which is the single-pass result.
// Few rows (queries, centroid merging) thus still fill the GPU.
// ---------------------------------------------------------------------------

constexpr int kAssignRows = 128;
constexpr int kAssignTile = 32;

// issue the direct loads of centroid tile [c0, c0 + nc) into `buf` (this wave's share of nw waves)
__device__ __forceinline__ void assign_issue_tile(const int8_t* __restrict__ Z, int d, int cpr, int c0,
                                                  int nc, int8_t* buf, int wv, int lane, int nw = 4) {
  const int chunks = nc * cpr;
  for (int j = wv; j * 64 < chunks; j += nw) {
    const int f = j * 64 + lane;
    const int r = f / cpr, p = f - r * cpr;
    if (f < chunks) {
      int c = p + r % cpr;
      if (c >= cpr) c -= cpr;
      __builtin_amdgcn_global_load_lds(Z + int64_t(c0 + r) * d + 16 * c, buf + 1024 * j, 16, 0, 0);
    }
  }
}

// Ordinary zero stores over an LDS array that direct global->LDS loads
// (__builtin_amdgcn_global_load_lds) then fill, by the whole block, at the
// kernel's start.  Without an ordinary store the optimiser may treat the array
// as never written, drop it and fold its reads to undefined values; the build
// checks every such kernel's LDS size against its declared arrays
// (kernel-resource-usage remarks) so a dropped array fails the build.
__device__ __forceinline__ void lds_claim(void* p, int bytes) {
  v4i* v = static_cast<v4i*>(p);
  for (int j = int(threadIdx.x); j < (bytes >> 4); j += int(blockDim.x)) v[j] = v4i{0, 0, 0, 0};
}

// s_waitcnt vmcnt(n), other counters unconstrained (gfx9 encoding:
// vmcnt[3:0] | expcnt 7 << 4 | lgkmcnt 15 << 8 | vmcnt[5:4] << 14)
#define KNNG_VMCNT(n) (((n) & 15) | (7 << 4) | (15 << 8) | (((n) >> 4) << 14))
template <int N>
__device__ __forceinline__ void wait_vmcnt_le(int n) {
  if constexpr (N == 0) {
    __builtin_amdgcn_s_waitcnt(KNNG_VMCNT(0));
  } else {
    if (n >= N)
      __builtin_amdgcn_s_waitcnt(KNNG_VMCNT(N));
    else
      wait_vmcnt_le<N - 1>(n);
  }
}

// this wave's direct loads for centroid tile u (16 B chunks, plus the tile's
// 32 weights on wave 0): what assign_issue_tile issues
__device__ __forceinline__ int assign_tile_loads(int C, int cpr, int u, int wv) {
  const int c0 = u * kAssignTile;
  const int nc = (C - c0) < kAssignTile ? (C - c0) : kAssignTile;
  const int jmax = (nc * cpr + 63) / 64;
  return (jmax > wv ? (jmax - wv + 3) / 4 : 0) + (wv == 0 ? 1 : 0);
}

constexpr int kAssignStages = 2;   // centroid tiles in LDS per block (2 x 33 KB)
constexpr int kAssignStageW = kAssignTile * kMaxDim;             // weights' offset in a stage
constexpr int kAssignStageBytes = kAssignStageW + 1024;          // rows + weights, 1 KB aligned

template <bool EXCL, int KS>
__global__ __launch_bounds__(256) void k_assign(const int8_t* __restrict__ q, int64_t nrows,
                                                int64_t row_off, int64_t row_stride, int d,
                                                const int8_t* __restrict__ Z,
                                                const float* __restrict__ w, int C,
                                                int32_t* __restrict__ out,
                                                float* __restrict__ score,
                                                float* __restrict__ pbest,
                                                int32_t* __restrict__ pidx) {
  // Centroid tiles stream through an LDS ring of kAssignStages buffers, each
  // with the tile's 32 weights, so the epilogue issues no global load (one
  // would make the next wait cover the tile being prefetched as well): the
  // MFMAs of tile t overlap the loads of the following tiles, and each wave
  // waits (exact vmcnt) only for its own share of tile t.  A deeper ring (4)
  // measured slower than 2 at 6.35M x 1024.
  // stage = the tile's rows, then its 32 weights at kAssignStageW (one array:
  // the optimizer sees LDS that only direct loads write as unwritten and
  // folds reads of a separate weight array away)
  __shared__ __attribute__((aligned(1024))) int8_t ztile[kAssignStages][kAssignStageBytes];
  lds_claim(ztile, int(sizeof(ztile)));   // ordered before the first fill by the row loop's barrier
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // KS > 0: d = 32 * KS known at compile time (straight-line MFMA chain the
  // scheduler can feed from LDS ahead of use); KS = 0: any d
  const int ksteps = KS > 0 ? KS : d >> 5, cpr = KS > 0 ? 2 * KS : d >> 4;
  const int ntiles_all = (C + kAssignTile - 1) / kAssignTile;
  const int per_slice = (ntiles_all + int(gridDim.y) - 1) / int(gridDim.y);
  const int t_lo = int(blockIdx.y) * per_slice;
  const int ntiles = min(ntiles_all, t_lo + per_slice);   // this block scans tiles [t_lo, ntiles)
  const int cl = lane & 31, hi = lane >> 5;
  const int rrot = cl % cpr;   // rotation of the B row this lane reads
  auto issue = [&](int u) {
    const int c0 = u * kAssignTile;
    const int nc = (C - c0) < kAssignTile ? (C - c0) : kAssignTile;
    const int buf = u % kAssignStages;
    assign_issue_tile(Z, d, cpr, c0, nc, ztile[buf], wv, lane);
    if (wv == 0 && lane < nc)
      __builtin_amdgcn_global_load_lds(w + c0 + lane, ztile[buf] + kAssignStageW, 4, 0, 0);
  };
  for (int64_t r0 = int64_t(blockIdx.x) * kAssignRows; r0 < nrows;
       r0 += int64_t(gridDim.x) * kAssignRows) {
    // A fragments: row r0 + 32*wv + (lane&31), bytes 32*ks + 16*(lane>>5)
    const int64_t arow = r0 + 32 * wv + cl;
    const bool arow_ok = arow < nrows;
    const int8_t* abase = q + (row_off + (arow_ok ? arow : 0) * row_stride) * d + 16 * hi;
    v4i a[32];
#pragma unroll
    for (int ks = 0; ks < 32; ++ks)
      a[ks] = (arow_ok && ks < ksteps) ? *reinterpret_cast<const v4i*>(abase + 32 * ks)
                                       : v4i{0, 0, 0, 0};
    float best[16];
    int32_t bidx[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      best[j] = -INFINITY;
      bidx[j] = 0x7FFFFFFF;
    }
    __syncthreads();   // every ring buffer free (previous row tile done)
    for (int u = t_lo; u < ntiles && u < t_lo + kAssignStages - 1; ++u) issue(u);
    for (int t = t_lo; t < ntiles; ++t) {
      // loads issued after tile t's: tiles t+1 .. t+stages-2 (those that exist)
      int later = 0;
      for (int u = t + 1; u < ntiles && u <= t + kAssignStages - 2; ++u) later += assign_tile_loads(C, cpr, u, wv);
      if constexpr (kAssignStages == 2)
        __builtin_amdgcn_s_waitcnt(KNNG_VMCNT(0));   // nothing issued after tile t yet
      else
        wait_vmcnt_le<31>(later);
      __syncthreads();   // tile t visible to all waves; tile t-1's buffer free
      if (t + kAssignStages - 1 < ntiles) issue(t + kAssignStages - 1);
      const int c0 = t * kAssignTile;
      const int nc = (C - c0) < kAssignTile ? (C - c0) : kAssignTile;
      const int8_t* rb = ztile[t % kAssignStages] + cl * d;
      v16i acc0, acc1;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        acc0[j] = 0;
        acc1[j] = 0;
      }
#pragma unroll
      for (int ks = 0; ks < 32; ks += 2) {
        if (ks < ksteps) {
          int p0 = 2 * ks + hi - rrot;
          if (p0 < 0) p0 += cpr;
          int p1 = 2 * ks + 2 + hi - rrot;
          if (p1 < 0) p1 += cpr;
          acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[ks], *reinterpret_cast<const v4i*>(rb + 16 * p0),
                                                       acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[ks + 1], *reinterpret_cast<const v4i*>(rb + 16 * p1),
                                                       acc1, 0, 0, 0);
        }
      }
      const int cell = c0 + cl;
      const bool ok = cl < nc;
      const float* wt = reinterpret_cast<const float*>(ztile[t % kAssignStages] + kAssignStageW);
      const float wc = ok ? wt[cl] : 0.0f;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const float sc = float(acc0[j] + acc1[j]) * wc;
        // acc element j of lane l holds row 8*(j/4) + 4*(l/32) + (j%4) of the wave's 32
        const bool self = EXCL && (r0 + 32 * wv + 8 * (j >> 2) + 4 * hi + (j & 3)) == cell;
        if (ok && !self && sc > best[j]) {   // cells arrive in increasing order: ties keep the smaller
          best[j] = sc;
          bidx[j] = cell;
        }
      }
    }
    // reduce over the 32 lanes that share the accumulator rows
#pragma unroll
    for (int j = 0; j < 16; ++j) {
#pragma unroll
      for (int m = 16; m >= 1; m >>= 1) {
        const float ov = __shfl_xor(best[j], m);
        const int32_t oi = __shfl_xor(bidx[j], m);
        if (ov > best[j] || (ov == best[j] && oi < bidx[j])) {
          best[j] = ov;
          bidx[j] = oi;
        }
      }
      const int64_t row = r0 + 32 * wv + 8 * (j >> 2) + 4 * hi + (j & 3);
      if (cl == 0 && row < nrows) {
        if (gridDim.y > 1) {
          pbest[int64_t(blockIdx.y) * nrows + row] = best[j];
          pidx[int64_t(blockIdx.y) * nrows + row] = bidx[j];
        } else {
          out[row] = bidx[j];
          if (score) score[row] = best[j];
        }
      }
    }
  }
}

// Centroid rows in LDS with a 1040-byte pitch: one direct 1 KB load per row,
// and the 16 B column reads of 8 consecutive rows cover all 32 banks, so
// every lane's fragment address is one base register plus an immediate.
constexpr int kAssignPitch = kMaxDim + 16;
constexpr int kAssign2StageW = kAssignTile * kAssignPitch;        // weights' offset in a stage
constexpr int kAssign2StageBytes = kAssign2StageW + 256;

// k_assign at d = 1024 with two blocks per CU: 4 waves per block, each with
// the A fragments of 32 rows over all of k (128 VGPRs), centroid tiles
// double-buffered in LDS in the padded layout (67 KB per block).  The two
// blocks on a CU wait at independent barriers, so on each SIMD one block's
// epilogue overlaps the other's MFMAs.
template <bool EXCL>
__global__ __launch_bounds__(256, 2) void k_assign_wide(const int8_t* __restrict__ q, int64_t nrows,
                                                        int64_t row_off, int64_t row_stride,
                                                        const int8_t* __restrict__ Z,
                                                        const float* __restrict__ w, int C,
                                                        int32_t* __restrict__ out,
                                                        float* __restrict__ score,
                                                        float* __restrict__ pbest,
                                                        int32_t* __restrict__ pidx) {
  constexpr int d = kMaxDim, KS = kMaxDim / 32;
  __shared__ __attribute__((aligned(1024))) int8_t ztile[2][kAssign2StageBytes];
  lds_claim(ztile, int(sizeof(ztile)));   // ordered before the first fill by the row loop's barrier
  const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ntiles_all = (C + kAssignTile - 1) / kAssignTile;
  const int per_slice = (ntiles_all + int(gridDim.y) - 1) / int(gridDim.y);
  const int t_lo = int(blockIdx.y) * per_slice;
  const int ntiles = min(ntiles_all, t_lo + per_slice);
  const int cl = lane & 31, hi = lane >> 5;
  auto issue = [&](int u) {
    const int c0 = u * kAssignTile;
    const int nc = (C - c0) < kAssignTile ? (C - c0) : kAssignTile;
    int8_t* buf = ztile[u & 1];
    for (int r = wv; r < nc; r += 4)   // one 1 KB row per load
      __builtin_amdgcn_global_load_lds(Z + int64_t(c0 + r) * d + 16 * lane, buf + r * kAssignPitch, 16, 0, 0);
    if (wv == 0 && lane < nc)
      __builtin_amdgcn_global_load_lds(w + c0 + lane, buf + kAssign2StageW, 4, 0, 0);
  };
  for (int64_t r0 = int64_t(blockIdx.x) * kAssignRows; r0 < nrows;
       r0 += int64_t(gridDim.x) * kAssignRows) {
    const int64_t arow = r0 + 32 * wv + cl;
    const bool arow_ok = arow < nrows;
    const int8_t* abase = q + (row_off + (arow_ok ? arow : 0) * row_stride) * d + 16 * hi;
    v4i a[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      a[ks] = arow_ok ? *reinterpret_cast<const v4i*>(abase + 32 * ks) : v4i{0, 0, 0, 0};
    float best[16];
    int32_t bidx[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      best[j] = -INFINITY;
      bidx[j] = 0x7FFFFFFF;
    }
    // fold tile t's scores into best/bidx with selects; on a partial tile
    // (FULL false) a column past its cells scores -inf, which never wins
    auto fold = [&](int t, const v16i& acc, auto full) {
      constexpr bool FULL = decltype(full)::value;
      const int c0 = t * kAssignTile;
      const int cell = c0 + cl;
      const bool ok = FULL || cl < C - c0;
      const float* wt = reinterpret_cast<const float*>(ztile[t & 1] + kAssign2StageW);
      const float wc = ok ? wt[cl] : 0.f;
      // self exclusion: this lane's column is the row of element j exactly
      // when its offset from the lane's first row equals 8 (j / 4) + j % 4
      int self_at = -1;
      if constexpr (EXCL) {
        const int64_t off = int64_t(cell) - (r0 + 32 * wv + 4 * hi);
        self_at = (off >= 0 && off < 32) ? int(off) : -1;
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        float sc = float(acc[j]) * wc;
        if constexpr (EXCL) {
          const bool self = self_at == 8 * (j >> 2) + (j & 3);
          sc = (ok && !self) ? sc : -INFINITY;
        } else if constexpr (!FULL) {
          sc = ok ? sc : -INFINITY;
        }
        const bool take = sc > best[j];   // cells arrive in increasing order: ties keep the smaller
        best[j] = take ? sc : best[j];
        bidx[j] = take ? cell : bidx[j];
      }
    };
    __syncthreads();   // both stages free (previous row tile done)
    if (t_lo < ntiles) issue(t_lo);
    for (int t = t_lo; t < ntiles; ++t) {
      __builtin_amdgcn_s_waitcnt(KNNG_VMCNT(0));   // this wave's loads of tile t have landed
      __syncthreads();   // tile t visible; stage (t+1) & 1 no longer read
      if (t + 1 < ntiles) issue(t + 1);
      const int8_t* rb = ztile[t & 1] + cl * kAssignPitch + 16 * hi;
      v4i b[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) b[ks] = *reinterpret_cast<const v4i*>(rb + 32 * ks);
      v16i acc;
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] = 0;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[ks], b[ks], acc, 0, 0, 0);
      // keep kAhead fragment reads in flight ahead of the MFMAs
      constexpr int kAhead = 4;
      __builtin_amdgcn_sched_group_barrier(0x100, kAhead, 0);   // DS reads
#pragma unroll
      for (int i = 0; i < KS - kAhead; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);      // MFMA
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);      // DS read
      }
      __builtin_amdgcn_sched_group_barrier(0x008, kAhead, 0);
      if (C - t * kAssignTile >= kAssignTile)
        fold(t, acc, std::true_type{});
      else
        fold(t, acc, std::false_type{});
    }
    // reduce over the 32 lanes that share the accumulator rows
#pragma unroll
    for (int j = 0; j < 16; ++j) {
#pragma unroll
      for (int m = 16; m >= 1; m >>= 1) {
        const float ov = __shfl_xor(best[j], m);
        const int32_t oi = __shfl_xor(bidx[j], m);
        if (ov > best[j] || (ov == best[j] && oi < bidx[j])) {
          best[j] = ov;
          bidx[j] = oi;
        }
      }
      const int64_t row = r0 + 32 * wv + 8 * (j >> 2) + 4 * hi + (j & 3);
      if (cl == 0 && row < nrows) {
        if (gridDim.y > 1) {
          pbest[int64_t(blockIdx.y) * nrows + row] = best[j];
          pidx[int64_t(blockIdx.y) * nrows + row] = bidx[j];
        } else {
          out[row] = bidx[j];
          if (score) score[row] = best[j];
        }
      }
    }
  }
}

// best over the S cell slices of a split k_assign: larger value, then smaller cell
__global__ __launch_bounds__(256) void k_assign_reduce(const float* __restrict__ pbest,
                                                       const int32_t* __restrict__ pidx, int S,
                                                       int64_t nrows, int32_t* __restrict__ out,
                                                       float* __restrict__ score) {
  for (int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x; r < nrows; r += int64_t(gridDim.x) * 256) {
    float b = pbest[r];
    int32_t bi = pidx[r];
#pragma unroll 8
    for (int y = 1; y < S; ++y) {
      const float v = pbest[int64_t(y) * nrows + r];
      const int32_t vi = pidx[int64_t(y) * nrows + r];
      if (v > b || (v == b && vi < bi)) {
        b = v;
        bi = vi;
      }
    }
    out[r] = bi;
    if (score) score[r] = b;
  }
}

// out[ids[j]] = rows[j] (d bytes each)
__global__ __launch_bounds__(256) void k_scatter_rows(const int8_t* __restrict__ rows, int d,
                                                      const int64_t* __restrict__ ids, int64_t m,
                                                      int8_t* __restrict__ out) {
  const int per = d >> 4;
  for (int64_t u = int64_t(blockIdx.x) * 256 + threadIdx.x; u < m * per;
       u += int64_t(gridDim.x) * 256) {
    const int64_t j = u / per;
    const int c = int(u - j * per);
    *reinterpret_cast<v4i*>(out + ids[j] * d + 16 * c) =
        *reinterpret_cast<const v4i*>(rows + j * d + 16 * c);
  }
}

// Centroid of cell c from its exact sum S_c (int8 rows or int64 sums):
// z = trunc(S * 127 / max |S|), *w = float(1 / sqrt(sum z^2)) for one cell's
// sum row S (d <= 1024) by the whole block; an all-zero sum gives z = 0,
// w = 0.  In place is allowed (every S value is read before any z value is
// written).
template <typename TS>
__device__ __forceinline__ void quant_cell(const TS* S, int d, int8_t* z, float* w, long long* red) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int64_t v[4];
  long long mx = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = tid + 256 * u;
    v[u] = k < d ? int64_t(S[k]) : 0;
    const long long av = v[u] < 0 ? -v[u] : v[u];
    mx = av > mx ? av : mx;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const long long ot = __shfl_xor(mx, o);
    mx = ot > mx ? ot : mx;
  }
  if (lane == 0) red[wv] = mx;
  __syncthreads();
  long long M = red[0];
  for (int u = 1; u < 4; ++u) M = red[u] > M ? red[u] : M;
  __syncthreads();
  long long ss = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = tid + 256 * u;
    if (k < d) {
      const int8_t zk = M ? int8_t((v[u] * 127) / M) : int8_t(0);
      z[k] = zk;
      ss += int(zk) * int(zk);
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o);
  if (lane == 0) red[wv] = ss;
  __syncthreads();
  if (tid == 0) {
    const long long t = red[0] + red[1] + red[2] + red[3];
    *w = t > 0 ? float(1.0 / sqrt(double(t))) : 0.0f;
  }
  __syncthreads();
}

// z_k, w_k from S_k (quant_cell) for the listed cells (all m if cells is
// null); one block per cell, 256 threads
template <typename TS>
__global__ __launch_bounds__(256) void k_quant_centroids(const TS* S, int d, const int32_t* __restrict__ cells,
                                                         int m, int8_t* z, float* __restrict__ w) {
  __shared__ long long red[4];
  for (int i = blockIdx.x; i < m; i += gridDim.x) {
    const int64_t cc = cells ? cells[i] : i;
    quant_cell(S + cc * d, d, z + cc * d, w + cc, red);
  }
}

// A k-means merging round's decisions from the round's nearest cells: live
// cell j (weight not NaN) absorbs k = best[j] when k > j, best[k] = j and the
// cosine bval[j] * w[j] exceeds `merge`; partner[j] = k, else -1.  Mutually
// nearest pairs are disjoint.  *npairs counts the pairs.
__global__ __launch_bounds__(256) void k_merge_decide(const int32_t* __restrict__ best,
                                                      const float* __restrict__ bval,
                                                      const float* __restrict__ w, int C, float merge,
                                                      int32_t* __restrict__ partner,
                                                      unsigned long long* __restrict__ npairs) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= C) return;
  int32_t p = -1;
  const float wj = w[j];
  if (wj == wj) {
    const int32_t k = best[j];
    if (k > j && k < C && best[k] == j && bval[j] * wj > merge) p = k;
  }
  partner[j] = p;
  if (p >= 0) atomicAdd(npairs, 1ull);
}

// The round's merges: S[j] += S[k], cnt[j] += cnt[k], z_j / w_j requantised,
// w[k] = NaN (dead: never wins an argmax again) for every pair (j, k).
__global__ __launch_bounds__(256) void k_merge_apply(int64_t* __restrict__ S, int64_t* __restrict__ cnt,
                                                     const int32_t* __restrict__ partner, int C, int d,
                                                     int8_t* z, float* __restrict__ w) {
  __shared__ long long red[4];
  for (int j = blockIdx.x; j < C; j += gridDim.x) {
    const int32_t k = partner[j];
    if (k < 0) continue;
    int64_t* sj = S + int64_t(j) * d;
    const int64_t* sk = S + int64_t(k) * d;
    for (int t = threadIdx.x; t < d; t += 256) sj[t] += sk[t];
    if (threadIdx.x == 0) {
      cnt[j] += cnt[k];
      w[k] = __builtin_nanf("");
    }
    __syncthreads();
    quant_cell(sj, d, z + int64_t(j) * d, w + j, red);
  }
}

// Cell member lists of an assignment on the device (k_cell_count, then
// these two): offs[c] = sum of cnt[c'] for c' < c (one block), then row
// off + j * stride of list position j at members[offs[a[j]] + k].  The order
// within a cell is arbitrary: the cell sums over it are exact integer sums.
__global__ __launch_bounds__(1024) void k_count_offsets(const int32_t* __restrict__ cnt, int C,
                                                        int64_t* __restrict__ offs) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x, per = (C + 1023) / 1024;
  const int b = t * per, e = min(C, b + per);
  int64_t sum = 0;
  for (int k = b; k < e; ++k) sum += cnt[k];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int64_t v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = part[t] - sum;
  if (t == 0) offs[0] = 0;
  for (int k = b; k < e; ++k) {
    run += cnt[k];
    offs[k + 1] = run;
  }
}

__global__ __launch_bounds__(256) void k_cell_scatter(const int32_t* __restrict__ a, int64_t nrows, int C,
                                                      int64_t off, int64_t stride,
                                                      const int64_t* __restrict__ offs,
                                                      int32_t* __restrict__ cur, int32_t* __restrict__ members) {
  for (int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x; j < nrows; j += int64_t(gridDim.x) * 256) {
    const int32_t c = a[j];
    if (c < 0 || c >= C) continue;
    members[offs[c] + atomicAdd(&cur[c], 1)] = int32_t(off + j * stride);
  }
}

// z_c, w_c from S_c (quant_cell) for every cell with cnt[c] > 0; an empty
// cell keeps its centroid
__global__ __launch_bounds__(256) void k_quant_nonempty(const int64_t* S, int d, const int32_t* __restrict__ cnt,
                                                        int C, int8_t* z, float* __restrict__ w) {
  __shared__ long long red[4];
  for (int c = blockIdx.x; c < C; c += gridDim.x) {
    if (cnt[c] == 0) continue;
    quant_cell(S + int64_t(c) * d, d, z + int64_t(c) * d, w + c, red);
  }
}

// w[cells[i]] = NaN: a dead cell never wins an argmax
__global__ void k_kill_cells(float* __restrict__ w, const int32_t* __restrict__ cells, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) w[cells[i]] = __builtin_nanf("");
}

// out[j] = row ids[j] (d bytes each)
__global__ __launch_bounds__(256) void k_gather_rows(const int8_t* __restrict__ q, int d,
                                                     const int64_t* __restrict__ ids, int64_t m,
                                                     int8_t* __restrict__ out) {
  const int per = d >> 4;
  for (int64_t u = int64_t(blockIdx.x) * 256 + threadIdx.x; u < m * per;
       u += int64_t(gridDim.x) * 256) {
    const int64_t j = u / per;
    const int c = int(u - j * per);
    *reinterpret_cast<v4i*>(out + j * d + 16 * c) =
        *reinterpret_cast<const v4i*>(q + ids[j] * d + 16 * c);
  }
}

// Per-cell exact sums of member rows: block per cell, members listed in
// `members` (cells contiguous, offsets in `offs`).  int64 accumulators: a
// cell may hold more than 2^24 rows of +-128.
__global__ __launch_bounds__(256) void k_cell_sum(const int8_t* __restrict__ q, int d,
                                                  const int32_t* __restrict__ members,
                                                  const int64_t* __restrict__ offs, int C,
                                                  int64_t* __restrict__ sums) {
  for (int c = blockIdx.x; c < C; c += gridDim.x) {
    const int64_t b = offs[c], e = offs[c + 1];
    for (int k0 = threadIdx.x * 4; k0 < d; k0 += 1024) {
      int64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
      int64_t m = b;
      for (; m + 4 <= e; m += 4) {
        uint32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          v[u] = *reinterpret_cast<const uint32_t*>(q + int64_t(members[m + u]) * d + k0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          s0 += int8_t(v[u] & 0xFF);
          s1 += int8_t((v[u] >> 8) & 0xFF);
          s2 += int8_t((v[u] >> 16) & 0xFF);
          s3 += int8_t(v[u] >> 24);
        }
      }
      for (; m < e; ++m) {
        const uint32_t v = *reinterpret_cast<const uint32_t*>(q + int64_t(members[m]) * d + k0);
        s0 += int8_t(v & 0xFF);
        s1 += int8_t((v >> 8) & 0xFF);
        s2 += int8_t((v >> 16) & 0xFF);
        s3 += int8_t(v >> 24);
      }
      int64_t* o = sums + int64_t(c) * d + k0;
      o[0] = s0;
      o[1] = s1;
      o[2] = s2;
      o[3] = s3;
    }
  }
}

// ---------------------------------------------------------------------------
// K3: leaf Gram tiles on int8 MFMA.  One block per leaf (<= 64 members), four
// waves, wave w computes the 32x32 quadrant (w>>1, w&1) with
// v_mfma_i32_32x32x32_i8 over K-chunks staged in LDS.  A and B fragments are
// read with the same lane->k mapping, so the k order inside a fragment does
// not change the exact int32 sums.  Every member's 64 slots for this tree are
// written: key(score, mate id), or 0 for self / beyond the leaf.
// ---------------------------------------------------------------------------

constexpr int kGramKC = 512;

__global__ __launch_bounds__(256) void k_leaf_gram(const int8_t* __restrict__ q, int d, int64_t n,
                                                   int T, const int32_t* __restrict__ perm,
                                                   const int4* __restrict__ leaves, int64_t nleaves,
                                                   uint64_t* __restrict__ cand) {
  __shared__ int32_t mid[64];
  __shared__ __attribute__((aligned(16))) int8_t tile[64 * (kGramKC + 16)];
  constexpr int kStride = kGramKC + 16;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int mq = w >> 1, nq = w & 1;
  for (int64_t L = blockIdx.x; L < nleaves; L += gridDim.x) {
    const int4 lf = leaves[L];
    const int t = lf.x, start = lf.y, size = lf.z;
    if (size > kLeafSlots) continue;   // big leaves: k_leaf_big (uniform per block)
    if (tid < 64) mid[tid] = tid < size ? perm[int64_t(t) * n + start + tid] : -1;
    __syncthreads();
    const bool rows_ok = mq * 32 < size, cols_ok = nq * 32 < size;
    v16i acc;
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0;
    for (int k0 = 0; k0 < d; k0 += kGramKC) {
      const int kc = (d - k0) < kGramKC ? (d - k0) : kGramKC;
      const int upr = kc >> 4;   // 16-byte units per row
      for (int u = tid; u < size * upr; u += 256) {
        const int r = u / upr, c = u - r * upr;
        *reinterpret_cast<v4i*>(tile + r * kStride + 16 * c) =
            *reinterpret_cast<const v4i*>(q + int64_t(mid[r]) * d + k0 + 16 * c);
      }
      __syncthreads();
      if (rows_ok && cols_ok) {
        const int8_t* ra = tile + (mq * 32 + (lane & 31)) * kStride + 16 * (lane >> 5);
        const int8_t* rb = tile + (nq * 32 + (lane & 31)) * kStride + 16 * (lane >> 5);
        for (int ks = 0; ks < kc; ks += 32) {
          const v4i a = *reinterpret_cast<const v4i*>(ra + ks);
          const v4i b = *reinterpret_cast<const v4i*>(rb + ks);
          acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc, 0, 0, 0);
        }
      }
      __syncthreads();
    }
    if (rows_ok) {
      const int ncol = nq * 32 + (lane & 31);
      const int32_t mate = (cols_ok && ncol < size) ? mid[ncol] : -1;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int m = mq * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        if (m < size) {
          const uint64_t key = (mate >= 0 && ncol != m) ? pack_key(acc[j], mate) : 0ull;
          cand[int64_t(mid[m]) * (T * kLeafSlots) + t * kLeafSlots + ncol] = key;
        }
      }
    }
    __syncthreads();
  }
}

// The second leaf-tile shape: one wave per leaf of at most 64 rows (blocks of
// four waves take four leaves), no LDS and no block barriers.  Lane (c, h)
// loads the 16-byte fragments of rows c and 32 + c at k-offset 16 h of every
// 32-dim step straight from HBM; the same fragments serve as A (rows) and B
// (columns) of v_mfma_i32_32x32x32_i8, and the Gram matrix is symmetric, so a
// leaf of 33..64 rows takes three 32 x 32 tiles (top-left, bottom-left,
// bottom-right; the bottom-left one also written transposed as the
// top-right) and a leaf of at most 32 rows one.  The keys equal
// k_leaf_gram's, slot for slot.
__global__ __launch_bounds__(256) void k_leaf_gram_wave(const int8_t* __restrict__ q, int d, int64_t n, int T,
                                                        const int32_t* __restrict__ perm,
                                                        const int4* __restrict__ leaves, int64_t nleaves,
                                                        uint64_t* __restrict__ cand) {
  const int lane = threadIdx.x & 63, cl = lane & 31, hi = lane >> 5;
  const int64_t w0 = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = (int64_t(gridDim.x) * blockDim.x) >> 6;
  const int ks_n = d >> 5;
  for (int64_t L = w0; L < nleaves; L += nw) {
    const int4 lf = leaves[L];
    const int t = lf.x, start = lf.y, size = lf.z;
    if (size > kLeafSlots) continue;   // big leaves: the big-leaf kernels
    const int32_t* mp = perm + int64_t(t) * n + start;
    const int32_t mtop = cl < size ? mp[cl] : -1;
    const int32_t mbot = 32 + cl < size ? mp[32 + cl] : -1;
    const bool two = size > 32;   // wave-uniform
    const int8_t* pt = q + int64_t(mtop < 0 ? 0 : mtop) * d + 16 * hi;
    const int8_t* pb = q + int64_t(mbot < 0 ? 0 : mbot) * d + 16 * hi;
    v16i a00, a10, a11;
#pragma unroll
    for (int j = 0; j < 16; ++j) a00[j] = a10[j] = a11[j] = 0;
    for (int k0 = 0; k0 < ks_n; k0 += 8) {
      v4i ft[8], fb[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const bool in = k0 + u < ks_n;
        ft[u] = (in && mtop >= 0) ? *reinterpret_cast<const v4i*>(pt + 32 * (k0 + u)) : v4i{0, 0, 0, 0};
        fb[u] = (in && two && mbot >= 0) ? *reinterpret_cast<const v4i*>(pb + 32 * (k0 + u)) : v4i{0, 0, 0, 0};
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        a00 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ft[u], ft[u], a00, 0, 0, 0);
        if (two) {
          a10 = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb[u], ft[u], a10, 0, 0, 0);
          a11 = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb[u], fb[u], a11, 0, 0, 0);
        }
      }
    }
    // element j of lane (cl, hi): row r = 8 (j / 4) + 4 hi + j % 4 of the
    // tile's row half, column cl of its column half.  Member m's slot s gets
    // (score, id of member s), 0 for s = m or s >= size, as k_leaf_gram.
    auto put = [&](int32_t member, int slot, int32_t mate, bool self, int32_t score) {
      const uint64_t key = (mate >= 0 && !self) ? pack_key(score, mate) : 0ull;
      cand[int64_t(member) * (T * kLeafSlots) + t * kLeafSlots + slot] = key;
    };
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int r = 8 * (j >> 2) + 4 * hi + (j & 3);
      const int32_t top_r = __shfl(mtop, r), bot_r = __shfl(mbot, r);   // members r and 32 + r
      if (r < size) put(top_r, cl, mtop, r == cl, a00[j]);                 // top-left
      if (cl < size) put(mtop, 32 + r, bot_r, false, two ? a10[j] : 0);    // top-right (transposed)
      if (two && 32 + r < size) {
        put(bot_r, cl, mtop, false, a10[j]);                               // bottom-left
        put(bot_r, 32 + cl, mbot, r == cl, a11[j]);                        // bottom-right
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K3b: leaves of 65..kMaxLeaf members.  A work item is (leaf, 128 members);
// wave v owns members r0+32v .. r0+32v+31 and holds their A fragments in
// registers; the leaf's members stream through LDS as 32-row B tiles shared
// by the four waves, double-buffered with direct global->LDS loads (rows in
// the rotated chunk layout of k_assign); v_mfma_i32_32x32x32_i8 scores go
// to the block's scratch rows.  Then each wave keeps, per member, the 64
// best mates by (score desc, id asc), self excluded: a bisection finds the
// 64th score, and a second bisection over ids settles ties at that score;
// counts are ballot popcounts.  The keys fill the member's 64 slots for this
// tree (order irrelevant: K4 sorts them).
// ---------------------------------------------------------------------------

constexpr int kBigRows = 128;

__device__ __forceinline__ void big_issue_tile(const int8_t* __restrict__ q, int d, int cpr,
                                               const int32_t* mid, int c0, int nc, int8_t* buf,
                                               int wv, int lane, int nw = 4) {
  const int chunks = nc * cpr;
  for (int j = wv; j * 64 < chunks; j += nw) {
    const int f = j * 64 + lane;
    const int r = f / cpr, p = f - r * cpr;
    if (f < chunks) {
      int c = p + r % cpr;
      if (c >= cpr) c -= cpr;
      __builtin_amdgcn_global_load_lds(q + int64_t(mid[c0 + r]) * d + 16 * c, buf + 1024 * j, 16, 0, 0);
    }
  }
}

__device__ __forceinline__ int ballot_count(bool p) { return __popcll(__ballot(p)); }

// Scores of item it = item0 + blockIdx.x (128 members of a big leaf against
// all its members) into scratch[blockIdx.x][128][mpad], int32 exact.  A
// fragments stay in registers (one wave per SIMD); member tiles are
// double-buffered in LDS by direct loads.
template <int KS>   // KS > 0: d = 32 * KS at compile time (straight-line MFMA chain); 0: any d
__global__ __launch_bounds__(256) void k_leaf_scores(const int8_t* __restrict__ q, int d, int64_t n,
                                                     const int32_t* __restrict__ perm,
                                                     const int4* __restrict__ items, int64_t item0,
                                                     int64_t nitems, int32_t* __restrict__ scratch,
                                                     int mpad) {
  __shared__ int32_t mid[kMaxLeaf];
  __shared__ __attribute__((aligned(1024))) int8_t tile[2][32 * kMaxDim];
  lds_claim(tile, int(sizeof(tile)));   // ordered before the first fill by the item loop's barrier
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int ksteps = KS > 0 ? KS : d >> 5, cpr = KS > 0 ? 2 * KS : d >> 4;
  const int cl = lane & 31, hi = lane >> 5;
  const int rrot = cl % cpr;
  for (int64_t bi = blockIdx.x; bi < nitems; bi += gridDim.x) {
    const int4 item = items[item0 + bi];   // {tree, start, size, first member}
    const int t = item.x, start = item.y, size = item.z, r0 = item.w;
    int32_t* scr = scratch + bi * kBigRows * mpad;
    __syncthreads();   // the previous item is done with mid[] and tile[]
    for (int j = tid; j < size; j += 256) mid[j] = perm[int64_t(t) * n + start + j];
    __syncthreads();
    const bool wave_on = r0 + 32 * wv < size;
    v4i a[32];
    {
      const int arow = r0 + 32 * wv + cl;
      const bool ok = arow < size;
      const int8_t* abase = q + int64_t(ok ? mid[arow] : 0) * d + 16 * hi;
#pragma unroll
      for (int ks = 0; ks < 32; ++ks)
        a[ks] = (ok && ks < ksteps) ? *reinterpret_cast<const v4i*>(abase + 32 * ks) : v4i{0, 0, 0, 0};
    }
    const int ntiles = (size + 31) >> 5;
    big_issue_tile(q, d, cpr, mid, 0, size < 32 ? size : 32, tile[0], wv, lane);
    for (int tt = 0; tt < ntiles; ++tt) {
      const int c0 = tt * 32;
      const int nc = (size - c0) < 32 ? (size - c0) : 32;
      __builtin_amdgcn_s_waitcnt(0);   // this wave's loads of tile tt have landed
      __syncthreads();                 // tile tt visible; tile tt-1's buffer free
      if (tt + 1 < ntiles) {
        const int c1 = c0 + 32;
        big_issue_tile(q, d, cpr, mid, c1, (size - c1) < 32 ? (size - c1) : 32, tile[(tt + 1) & 1], wv,
                       lane);
      }
      if (wave_on) {
        const int8_t* rb = tile[tt & 1] + cl * d;
        v16i acc0, acc1;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          acc0[j] = 0;
          acc1[j] = 0;
        }
#pragma unroll
        for (int ks = 0; ks < 32; ks += 2)
          if (ks < ksteps) {
            int p0 = 2 * ks + hi - rrot;
            if (p0 < 0) p0 += cpr;
            int p1 = 2 * ks + 2 + hi - rrot;
            if (p1 < 0) p1 += cpr;
            acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[ks], *reinterpret_cast<const v4i*>(rb + 16 * p0),
                                                         acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[ks + 1],
                                                         *reinterpret_cast<const v4i*>(rb + 16 * p1), acc1,
                                                         0, 0, 0);
          }
        if (cl < nc) {
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            // acc element j of lane l: row 8*(j/4) + 4*(l/32) + (j%4) of the wave's 32, column l%32
            const int rl = 32 * wv + 8 * (j >> 2) + 4 * hi + (j & 3);
            scr[int64_t(rl) * mpad + c0 + cl] = acc0[j] + acc1[j];
          }
        }
      }
    }
  }
}

// k_leaf_scores at d = 1024 with two blocks per CU, as k_assign_wide: 4
// waves, each with the A fragments of 32 member rows over all of k, member
// tiles double-buffered in LDS in the padded layout (one 1 KB direct load per
// member row), fragment reads kept ahead of the MFMAs; each wave writes its
// 32 x 32 exact int32 scores per tile.
__global__ __launch_bounds__(256, 2) void k_leaf_scores_wide(const int8_t* __restrict__ q, int64_t n,
                                                             const int32_t* __restrict__ perm,
                                                             const int4* __restrict__ items, int64_t item0,
                                                             int64_t nitems, int32_t* __restrict__ scratch,
                                                             int mpad) {
  constexpr int d = kMaxDim, KS = kMaxDim / 32;
  __shared__ int32_t mid[kMaxLeaf];
  __shared__ __attribute__((aligned(1024))) int8_t tile[2][kAssign2StageW];
  lds_claim(tile, int(sizeof(tile)));   // ordered before the first fill by the item loop's barrier
  const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int cl = lane & 31, hi = lane >> 5;
  auto issue = [&](int c0, int nc, int8_t* buf) {
    for (int r = wv; r < nc; r += 4)   // member row c0 + r: one 1 KB load
      __builtin_amdgcn_global_load_lds(q + int64_t(mid[c0 + r]) * d + 16 * lane, buf + r * kAssignPitch, 16, 0,
                                       0);
  };
  for (int64_t bi = blockIdx.x; bi < nitems; bi += gridDim.x) {
    const int4 item = items[item0 + bi];   // {tree, start, size, first member}
    const int t = item.x, start = item.y, size = item.z, r0 = item.w;
    int32_t* scr = scratch + bi * kBigRows * mpad;
    __syncthreads();   // the previous item is done with mid[] and tile[]
    for (int j = tid; j < size; j += 256) mid[j] = perm[int64_t(t) * n + start + j];
    __syncthreads();
    const bool wave_on = r0 + 32 * wv < size;
    v4i a[KS];
    {
      const int arow = r0 + 32 * wv + cl;
      const bool ok = arow < size;
      const int8_t* abase = q + int64_t(ok ? mid[arow] : 0) * d + 16 * hi;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
        a[ks] = ok ? *reinterpret_cast<const v4i*>(abase + 32 * ks) : v4i{0, 0, 0, 0};
    }
    const int ntiles = (size + 31) >> 5;
    issue(0, size < 32 ? size : 32, tile[0]);
    for (int tt = 0; tt < ntiles; ++tt) {
      __builtin_amdgcn_s_waitcnt(KNNG_VMCNT(0));   // this wave's loads of tile tt have landed
      __syncthreads();                              // tile tt visible; tile tt-1's buffer free
      if (tt + 1 < ntiles) {
        const int c1 = (tt + 1) * 32;
        issue(c1, (size - c1) < 32 ? (size - c1) : 32, tile[(tt + 1) & 1]);
      }
      if (!wave_on) continue;
      const int8_t* rb = tile[tt & 1] + cl * kAssignPitch + 16 * hi;
      v4i b[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) b[ks] = *reinterpret_cast<const v4i*>(rb + 32 * ks);
      v16i acc;
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] = 0;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[ks], b[ks], acc, 0, 0, 0);
      constexpr int kAhead = 4;   // fragment reads in flight ahead of the MFMAs
      __builtin_amdgcn_sched_group_barrier(0x100, kAhead, 0);
#pragma unroll
      for (int i = 0; i < KS - kAhead; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, kAhead, 0);
      const int c0 = tt * 32;
      if (cl < size - c0) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          // acc element j of lane l: row 8*(j/4) + 4*(l/32) + (j%4) of the wave's 32, column l%32
          const int rl = 32 * wv + 8 * (j >> 2) + 4 * hi + (j & 3);
          scr[int64_t(rl) * mpad + c0 + cl] = acc[j];
        }
      }
    }
  }
}

// The R entries of largest (score, then smaller id) among v[] (N per lane,
// INT32_MIN = none) whose scores lie in [lo, hi], lo low enough that at least
// R entries reach it; written to dst[0..R) in the order of index lane + 64 e.
template <int N>
__device__ __forceinline__ void leaf_top_r(const int32_t (&v)[N], const int32_t (&ids)[N], int lo, int hi,
                                           int R, int64_t n, uint64_t* dst, int lane) {
  // largest theta with |{v >= theta}| >= R
  while (lo < hi) {
    const int md = lo + (hi - lo + 1) / 2;
    int cnt = 0;
#pragma unroll
    for (int e = 0; e < N; ++e) cnt += ballot_count(v[e] >= md);
    if (cnt >= R)
      lo = md;
    else
      hi = md - 1;
  }
  const int theta = lo;
  int cgt = 0, ceq = 0;
#pragma unroll
  for (int e = 0; e < N; ++e) {
    cgt += ballot_count(v[e] > theta);
    ceq += ballot_count(v[e] == theta);
  }
  const int need = R - cgt;
  int idcut = INT32_MAX;
  if (ceq > need) {
    // smallest idcut with |{v == theta, id <= idcut}| >= need
    int ilo = 0, ihi = int(n - 1);
    while (ilo < ihi) {
      const int md = ilo + (ihi - ilo) / 2;
      int cnt = 0;
#pragma unroll
      for (int e = 0; e < N; ++e) cnt += ballot_count(v[e] == theta && ids[e] <= md);
      if (cnt >= need)
        ihi = md;
      else
        ilo = md + 1;
    }
    idcut = ilo;
  }
  int base = 0;
#pragma unroll
  for (int e = 0; e < N; ++e) {
    const bool sel = v[e] > theta || (v[e] == theta && ids[e] <= idcut);
    const uint64_t bal = __ballot(sel);
    const int pos = base + __popcll(bal & ((1ull << lane) - 1ull));
    if (sel) dst[pos] = pack_key(v[e], ids[e]);
    base += __popcll(bal);
  }
  if (lane >= R) dst[lane] = 0ull;
}

constexpr int kSelCap = 512;   // compacted candidates per wave in k_leaf_select

// Top-R mates (R = min(size - 1, 64)) of one member row of a big leaf by
// (score desc, id asc), self excluded, from its scores row; one wave per row,
// rows (bi, rr) of the chunk's items taken in turn.  The member's 64 slots of
// this tree get the keys (0 beyond R; the merge sorts them).  Each lane loads
// four consecutive columns per 16-byte load.  The entries at or above the
// smallest lane maximum (at least 64 of them when every lane holds one) are
// compacted into LDS (typically a few hundred of the 2048), and the selection
// runs on those, 8 per lane instead of 32.  When more than kSelCap entries
// reach it, the cut rises to the largest score that at least R entries reach
// (the R-th largest score: every selected entry reaches it); a row that still
// holds more than kSelCap entries at that cut (ties) is appended to `slow` for
// k_leaf_select_slow.  Keeping that case out of this kernel keeps its
// registers at 8 waves per SIMD.
__global__ __launch_bounds__(256, 5) void k_leaf_select(int64_t n, int T, const int32_t* __restrict__ perm,
                                                     const int4* __restrict__ items, int64_t item0,
                                                     int64_t nitems, const int32_t* __restrict__ scratch,
                                                     int mpad, uint64_t* __restrict__ cand,
                                                     int32_t* __restrict__ slow, unsigned int* __restrict__ nslow) {
  constexpr int kPer = kMaxLeaf / 64, kC = kSelCap / 64;
  __shared__ int2 cbuf[4][kSelCap];   // (score, column) of this wave's compacted entries
  const int lane = threadIdx.x & 63;
  int2* cb = cbuf[threadIdx.x >> 6];
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  for (int64_t wr = wave; wr < nitems * kBigRows; wr += nwaves) {
    const int64_t bi = wr / kBigRows;
    const int rr = int(wr - bi * kBigRows);
    const int4 item = items[item0 + bi];
    const int t = item.x, start = item.y, size = item.z;
    const int m = item.w + rr;
    if (m >= size) continue;
    const int32_t* mid = perm + int64_t(t) * n + start;
    const int R = (size - 1) < kLeafSlots ? (size - 1) : kLeafSlots;
    const int32_t* row = scratch + (bi * kBigRows + rr) * mpad;
    // v[4 e + k] holds column 4 lane + 256 e + k (mpad is a multiple of 64).
    // lane4 is opaque to the compiler, so the 32 per-element column values
    // are not hoisted out of the row loop into registers of their own.
    int lane4 = 4 * lane;
    asm volatile("" : "+v"(lane4));
    const int lim = size - lane4, self = m - lane4;   // column c: c - lane4 < lim, != self
    int32_t v[kPer];
    int lmax = INT32_MIN;
#pragma unroll
    for (int e = 0; e < kPer / 4; ++e) {
      // chunks wholly past the leaf are skipped (a wave-uniform test); columns
      // past mpad in the last chunk read the next row (or the kMaxLeaf pad
      // behind the last one) and are masked below
      const int4 x4 = 256 * e < size ? *reinterpret_cast<const int4*>(row + lane4 + 256 * e)
                                     : make_int4(INT32_MIN, INT32_MIN, INT32_MIN, INT32_MIN);
      const int32_t xs[4] = {x4.x, x4.y, x4.z, x4.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int o = 256 * e + k;
        v[4 * e + k] = (o < lim && o != self) ? xs[k] : INT32_MIN;   // scores are > -2^25
        lmax = max(lmax, v[4 * e + k]);
      }
    }
    // When every lane holds a valid entry, the 64 lane maxima are distinct
    // entries, so their minimum has at least 64 >= R entries at or above it;
    // otherwise every entry is a candidate.
    int lo = wave_min_i(lmax);
    const int hi = wave_max_i(lmax);
    if (ballot_count(lmax != INT32_MIN) < 64) lo = -(1 << 25);
    int c0 = 0;
#pragma unroll
    for (int e = 0; e < kPer; ++e) c0 += ballot_count(v[e] >= lo);
    if (c0 > kSelCap) {
      // largest cut with at least R entries at or above it
      int a = lo, b = hi;
      while (a < b) {
        const int md = a + (b - a + 1) / 2;
        int cnt = 0;
#pragma unroll
        for (int e = 0; e < kPer; ++e) cnt += ballot_count(v[e] >= md);
        if (cnt >= R)
          a = md;
        else
          b = md - 1;
      }
      lo = a;
      c0 = 0;
#pragma unroll
      for (int e = 0; e < kPer; ++e) c0 += ballot_count(v[e] >= lo);
      if (c0 > kSelCap) {
        if (lane == 0) slow[atomicAdd(nslow, 1u)] = int32_t(wr);
        continue;
      }
    }
    int base = 0;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const bool sel = v[e] >= lo;
      const uint64_t bal = __ballot(sel);
      if (sel) cb[base + __popcll(bal & ((1ull << lane) - 1ull))] = make_int2(v[e], lane4 + 256 * (e >> 2) + (e & 3));
      base += __popcll(bal);
    }
    __builtin_amdgcn_wave_barrier();
    int32_t cv[kC], ids[kC];
#pragma unroll
    for (int k = 0; k < kC; ++k) {
      const int i = lane + 64 * k;
      const int2 p = i < c0 ? cb[i] : make_int2(INT32_MIN, 0);
      cv[k] = p.x;
      ids[k] = i < c0 ? mid[p.y] : 0;
    }
    __builtin_amdgcn_wave_barrier();   // cbuf is free for the next row once read
    uint64_t* dst = cand + int64_t(mid[m]) * (T * kLeafSlots) + t * kLeafSlots;
    leaf_top_r<kC>(cv, ids, lo, hi, R, n, dst, lane);
  }
}

// The rows k_leaf_select passed on (more than kSelCap entries tie at the
// R-th largest score): the same selection over all of the row's entries.
// One wave per listed row.
__global__ __launch_bounds__(256) void k_leaf_select_slow(int64_t n, int T, const int32_t* __restrict__ perm,
                                                          const int4* __restrict__ items, int64_t item0,
                                                          const int32_t* __restrict__ scratch, int mpad,
                                                          uint64_t* __restrict__ cand,
                                                          const int32_t* __restrict__ slow,
                                                          const unsigned int* __restrict__ nslow) {
  constexpr int kPer = kMaxLeaf / 64;
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  const int64_t cnt = *nslow;
  for (int64_t s = wave; s < cnt; s += nwaves) {
    const int64_t wr = slow[s];
    const int64_t bi = wr / kBigRows;
    const int rr = int(wr - bi * kBigRows);
    const int4 item = items[item0 + bi];
    const int t = item.x, start = item.y, size = item.z;
    const int m = item.w + rr;
    const int32_t* mid = perm + int64_t(t) * n + start;
    const int R = (size - 1) < kLeafSlots ? (size - 1) : kLeafSlots;
    const int32_t* row = scratch + (bi * kBigRows + rr) * mpad;
    int32_t v[kPer], ids[kPer];
    int lo = INT32_MAX, hi = INT32_MIN;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const int col = lane + 64 * e;
      v[e] = (col < size && col != m) ? row[col] : INT32_MIN;
      ids[e] = col < size ? mid[col] : 0;
      if (v[e] != INT32_MIN) {
        lo = min(lo, v[e]);
        hi = max(hi, v[e]);
      }
    }
    lo = wave_min_i(lo);
    hi = wave_max_i(hi);
    uint64_t* dst = cand + int64_t(mid[m]) * (T * kLeafSlots) + t * kLeafSlots;
    leaf_top_r<kPer>(v, ids, lo, hi, R, n, dst, lane);
  }
}

// ---------------------------------------------------------------------------
// K4: per-point top-K of the leaf candidates.  Bitonic sort (descending) of
// the point's T*64 keys in LDS; duplicates (same mate from several trees)
// carry identical keys and become adjacent; the first K distinct keys are the
// list.  Fewer than K distinct candidates: the reference's xorshift32 padding
// (src/main.cpp:271-312), done by one thread.
// ---------------------------------------------------------------------------

// K4 with one wave per row (S2 <= 1024 slots): the row's S2 keys sit in
// registers, E = S2 / 64 per lane at positions lane * E + e, and are bitonic
// sorted descending there (within-lane steps as register swaps, the others
// with lane shuffles), so no block barrier is needed.  The sorted order of a
// multiset is unique, so the list equals k_init_merge's; so does the padding
// path (fewer than K distinct candidates), run by lane 0 as there by thread 0.
template <int S2>
__global__ __launch_bounds__(256) void k_init_merge_wave(const int8_t* __restrict__ q, int d, int64_t n, int T,
                                                         int K, const uint64_t* __restrict__ cand,
                                                         int32_t* __restrict__ gid, int32_t* __restrict__ gsc,
                                                         const int64_t* __restrict__ seg_base, int nseg) {
  constexpr int E = S2 / 64;
  static_assert(E >= 1 && E <= 16, "S2 must be in [64, 1024]");
  __shared__ uint64_t pool_w[4][2 * kMaxK];
  __shared__ int32_t pad_id_w[4][kMaxK];
  __shared__ int32_t pad_sc_w[4][kMaxK];
  __shared__ int np_w[4];
  __shared__ uint32_t state_w[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t* pool = pool_w[w];
  int32_t* pad_id = pad_id_w[w];
  int32_t* pad_sc = pad_sc_w[w];
  const int slots = T * kLeafSlots;
  for (int64_t i = int64_t(blockIdx.x) * 4 + w; i < n; i += int64_t(gridDim.x) * 4) {
    const uint64_t* src = cand + i * slots;
    uint64_t v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int x = lane * E + e;
      v[e] = x < slots ? src[x] : 0ull;
    }
#pragma unroll
    for (int k = 2; k <= S2; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        if (j < E) {   // partner in this lane's registers
#pragma unroll
          for (int e = 0; e < E; ++e)
            if ((e & j) == 0) {
              const bool desc = ((lane * E + e) & k) == 0;
              const uint64_t a = v[e], b = v[e + j];
              const bool sw = (a < b) == desc;
              v[e] = sw ? b : a;
              v[e + j] = sw ? a : b;
            }
        } else {       // partner lane lane ^ (j / E), same register
          const int lj = j / E;
          const bool lower = (lane & lj) == 0;
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const uint64_t o = __shfl_xor(v[e], lj);
            const int lo = (lower ? lane : (lane ^ lj)) * E + e;   // the pair's lower position
            const bool desc = (lo & k) == 0;
            const uint64_t mx = v[e] > o ? v[e] : o, mn = v[e] > o ? o : v[e];
            v[e] = (lower == desc) ? mx : mn;
          }
        }
      }
    }
    // distinct valid keys in order: the first K go to the list
    const uint64_t prev_last = __shfl_up(v[E - 1], 1);
    int flag[E];
    int cnt = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint64_t pv = e > 0 ? v[e - 1] : prev_last;
      flag[e] = (v[e] != 0ull && (lane * E + e == 0 || pv != v[e])) ? 1 : 0;
      cnt += flag[e];
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o);
      if (lane >= o) incl += t;
    }
    const int U0 = __shfl(incl, 63);
    int pos = incl - cnt;
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (flag[e]) {
        if (pos < K) {
          gid[i * K + pos] = key_id(v[e]);
          gsc[i * K + pos] = key_score(v[e]);
          pool[pos] = v[e];
        }
        ++pos;
      }
    if (U0 < K) {   // uniform across the wave
      __builtin_amdgcn_wave_barrier();
      int64_t sb = 0, sn = n;
      if (nseg > 0) {
        const int sg = seg_of(seg_base, nseg, i);
        sb = seg_base[sg];
        sn = seg_base[sg + 1] - sb;
      }
      // padding ids: src/main.cpp:271-286, in the build's own row numbers
      if (lane == 0) {
        uint32_t st = uint32_t(1337 + (i - sb));
        int np = 0;
        while (U0 + np < K) {
          st = xorshift32(st);
          const int32_t r = int32_t(sb + int64_t(st % uint64_t(sn)));
          if (r == i) continue;
          pad_id[np++] = r;
        }
        np_w[w] = np;
        state_w[w] = st;
      }
      __builtin_amdgcn_wave_barrier();
      const int np = np_w[w];
      for (int p = 0; p < np; ++p) {
        const int8_t* a = q + i * d;
        const int8_t* b = q + int64_t(pad_id[p]) * d;
        int sdot = 0;
        for (int c = lane; c < (d >> 4); c += 64)
          sdot = dot16_i8(*reinterpret_cast<const v4i*>(a + 16 * c), *reinterpret_cast<const v4i*>(b + 16 * c),
                          sdot);
        sdot = wave_sum_i(sdot);
        if (lane == 0) pad_sc[p] = sdot;
      }
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) {
        int m = U0;
        for (int p = 0; p < np; ++p) pool[m++] = pack_key(pad_sc[p], pad_id[p]);
        for (int a = 1; a < m; ++a) {   // insertion sort, descending
          const uint64_t val = pool[a];
          int b = a;
          while (b > 0 && pool[b - 1] < val) {
            pool[b] = pool[b - 1];
            --b;
          }
          pool[b] = val;
        }
        int kept = 0;
        int32_t prev = -1;
        for (int a = 0; a < m && kept < K; ++a) {
          const int32_t id = key_id(pool[a]);
          if (id != prev) {
            gid[i * K + kept] = id;
            gsc[i * K + kept] = key_score(pool[a]);
            prev = id;
            ++kept;
          }
        }
        uint32_t st = state_w[w];
        while (kept < K) {   // src/main.cpp:303-312
          st = xorshift32(st);
          const int32_t r = int32_t(sb + int64_t(st % uint64_t(sn)));
          if (r != i) {
            gid[i * K + kept] = r;
            gsc[i * K + kept] = INT32_MIN;
            ++kept;
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

template <int S2>
__global__ __launch_bounds__(256) void k_init_merge(const int8_t* __restrict__ q, int d, int64_t n,
                                                    int T, int K, const uint64_t* __restrict__ cand,
                                                    int32_t* __restrict__ gid,
                                                    int32_t* __restrict__ gsc,
                                                    const int64_t* __restrict__ seg_base, int nseg) {
  __shared__ uint64_t keys[S2];
  __shared__ int lds[8];
  __shared__ int s_unique;
  __shared__ uint64_t pool[2 * kMaxK];
  __shared__ int32_t pad_id[kMaxK];
  __shared__ int32_t pad_sc[kMaxK];
  __shared__ int s_npad;
  const int tid = threadIdx.x;
  const int slots = T * kLeafSlots;
  constexpr int PER = S2 / 256;   // elements per thread in the compaction step
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint64_t* src = cand + i * slots;
    for (int j = tid; j < S2; j += 256) keys[j] = j < slots ? src[j] : 0ull;
    __syncthreads();
    for (int k = 2; k <= S2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int p = tid; p < S2 / 2; p += 256) {
          const int a = 2 * p - (p & (j - 1));
          const int b = a + j;
          const uint64_t x = keys[a], y = keys[b];
          const bool desc = (a & k) == 0;
          if ((x < y) == desc) {
            keys[a] = y;
            keys[b] = x;
          }
        }
        __syncthreads();
      }
    }
    // distinct valid keys, in order
    int flag[PER];
    int cnt = 0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int j = tid * PER + e;
      const uint64_t kj = keys[j];
      flag[e] = (kj != 0ull && (j == 0 || keys[j - 1] != kj)) ? 1 : 0;
      cnt += flag[e];
    }
    int tot;
    int pos = block_exclusive_scan(cnt, lds, &tot);
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      if (flag[e]) {
        if (pos < K) {
          const uint64_t kj = keys[tid * PER + e];
          gid[i * K + pos] = key_id(kj);
          gsc[i * K + pos] = key_score(kj);
          pool[pos] = kj;
        }
        ++pos;
      }
    }
    if (tid == 0) s_unique = tot;
    __syncthreads();
    const int U0 = s_unique;
    if (U0 < K) {   // uniform across the block
      // the row's own build: rows [sb, sb + sn) (packed builds), else all n
      int64_t sb = 0, sn = n;
      if (nseg > 0) {
        const int sg = seg_of(seg_base, nseg, i);
        sb = seg_base[sg];
        sn = seg_base[sg + 1] - sb;
      }
      // padding ids: src/main.cpp:271-286, in the build's own row numbers
      if (tid == 0) {
        uint32_t s = uint32_t(1337 + (i - sb));
        int np = 0;
        while (U0 + np < K) {
          s = xorshift32(s);
          const int32_t r = int32_t(sb + int64_t(s % uint64_t(sn)));
          if (r == i) continue;
          pad_id[np++] = r;
        }
        pad_sc[0] = int32_t(s);   // carry the xorshift state
        s_npad = np;
      }
      __syncthreads();
      const int np = s_npad;
      const uint32_t state = uint32_t(pad_sc[0]);
      __syncthreads();
      // dots of the padded ids: one wave each
      const int lane = tid & 63, w = tid >> 6;
      for (int p = w; p < np; p += 4) {
        const int8_t* a = q + i * d;
        const int8_t* b = q + int64_t(pad_id[p]) * d;
        int s = 0;
        for (int c = lane; c < (d >> 4); c += 64)
          s = dot16_i8(*reinterpret_cast<const v4i*>(a + 16 * c),
                       *reinterpret_cast<const v4i*>(b + 16 * c), s);
        s = wave_sum_i(s);
        if (lane == 0) pad_sc[p] = s;
      }
      __syncthreads();
      if (tid == 0) {
        int m = U0;
        for (int p = 0; p < np; ++p) pool[m++] = pack_key(pad_sc[p], pad_id[p]);
        for (int a = 1; a < m; ++a) {   // insertion sort, descending
          const uint64_t v = pool[a];
          int b = a;
          while (b > 0 && pool[b - 1] < v) {
            pool[b] = pool[b - 1];
            --b;
          }
          pool[b] = v;
        }
        int kept = 0;
        int32_t prev = -1;
        for (int a = 0; a < m && kept < K; ++a) {
          const int32_t id = key_id(pool[a]);
          if (id != prev) {
            gid[i * K + kept] = id;
            gsc[i * K + kept] = key_score(pool[a]);
            prev = id;
            ++kept;
          }
        }
        uint32_t s = state;
        while (kept < K) {   // src/main.cpp:303-312
          s = xorshift32(s);
          const int32_t r = int32_t(sb + int64_t(s % uint64_t(sn)));
          if (r != i) {
            gid[i * K + kept] = r;
            gsc[i * K + kept] = INT32_MIN;
            ++kept;
          }
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// K5-K7: one Jacobi join iteration, one block (4 waves) per point.
//  1. own list + self into an LDS hash set (excluded from candidates)
//  2. 2-hop ids ids_in[own[k]][k2] -> hash set; first insert appends to cid[]
//  3. sketch similarity, 8 lanes per candidate (128 B gather), LDS histogram
//  4. threshold T = max t with |{h >= t}| >= M (ties at T kept)
//  5. int8 dots of the selected, 16 lanes per candidate (1 KB gather)
//  6. rank merge of own list and new candidates by (score desc, id asc, slot)
//  7. write the new list, count changed slots (src/main.cpp:421-427)
// ---------------------------------------------------------------------------

constexpr int kRevMax = 8;   // reverse_width limit

struct JoinScratch {
  uint64_t own_key[kMaxK];    // pack_key(own score, own id)
  int32_t own_id[kMaxK];
  int32_t own_sc[kMaxK];
  int32_t new_id[kMaxK];
  int32_t new_sc[kMaxK];
  int32_t new_fresh[kMaxK];   // 1: the entry came from this iteration's candidates
  uint8_t own_bit[kMaxK];     // 1: own entry k entered the list last iteration (one byte per entry:
                              // 64-bit masks per neighbour would be hoisted out of the row loop and
                              // spilled)
  uint8_t own_mask[kMaxK];    // own entry k's leaf mask byte
  uint8_t new_mask[kMaxK];
  uint64_t own_sk[kMaxDim / 64];
  int32_t own_leaf[kMaxTrees];   // the row's leaf (first perm position) in every tree
  int32_t rv_id[kRevMax];     // the row's reverse entries (-1: none) and their mask bytes
  int32_t rv_mk[kRevMax];
  int32_t hist[kMaxDim + 1];
  int32_t cnt[6];   // 0 unique candidates, 1 selected, 2 survivors, 3 threshold, 4 new ids,
                    // 5 candidates left after dropping leaf-mates
  int lds[8];
};



template <int HSBITS, int CAP, int OCC, int DOTPER, int SKPER, int DIM, bool SMALL = false, bool REV = false>
__global__ __launch_bounds__(256, OCC) void k_join(
    const int8_t* __restrict__ q, const uint64_t* __restrict__ sk, int d, int64_t n, int K, int M,
    const int32_t* __restrict__ ids_in, const int32_t* __restrict__ sc_in,
    int32_t* __restrict__ ids_out, int32_t* __restrict__ sc_out,
    unsigned long long* __restrict__ counters, unsigned long long* __restrict__ rowq,
    const int32_t* __restrict__ order,
    const uint64_t* __restrict__ fresh_in, uint64_t* __restrict__ fresh_out,
    const int64_t* __restrict__ nwork_dev, const int64_t* __restrict__ seg_base, int nseg,
    unsigned long long* __restrict__ segc, int32_t* __restrict__ ovf,
    unsigned long long* __restrict__ ovf_n, const uint8_t* __restrict__ masks_in,
    uint8_t* __restrict__ masks_out, const int32_t* __restrict__ leafkey, int T, int TP,
    const int32_t* __restrict__ rv_id, const uint8_t* __restrict__ rv_mask, int R) {
  // order[0 .. nwork) lists the rows to process (nwork = *nwork_dev if given,
  // else n).  With segments (independent builds packed together, nseg > 0),
  // each row's changed-slot and new-id counts are also added to its build's
  // pair segc[2s], segc[2s+1]; a block accumulates while consecutive rows
  // stay in one build and flushes on a change.
  // fresh_in != nullptr: incremental join -- the pair (i -> n1 -> c) is a
  // candidate only if n1 is fresh in i's list or c is fresh in n1's list
  // (otherwise the previous iteration already considered it)
  // SMALL: the variant for rows with few candidates (iterations under the
  // new/old rule after the first): a small hash and candidate array, no
  // sketches (every row it finishes has at most M candidates, so all are
  // re-scored, as the full kernel would), more blocks per CU.  A row whose
  // unique candidates pass lim = min(M, CAP - 256) stops inserting (each of
  // the 4 waves checks before inserting at most 64 more, so at most CAP are
  // ever held) and is appended to ovf[*ovf_n]; the full kernel then runs
  // those rows.  It writes nothing for such a row.
  // masks_in != nullptr (the graph came from this forest's leaf init): list
  // entry e of row r carries a mask byte, bit t = r and e share tree t's leaf
  // (trees 0..7).  The path i -> n1 -> c is dropped when both of its edges
  // share a leaf of the same tree: then c shares that leaf with i, the init
  // scored it exactly and kept it in i's 64 slots of that tree unless 64 mates
  // beat it, so it is either in i's list already or below the init list's
  // K-th entry (K <= 64), and lists only improve: it could never enter.  New
  // entries get their mask byte from the leaf keys (leafkey[r * TP + t] = the
  // first perm position of row r's leaf in tree t).
  // REV (with R > 0): reverse-neighbour candidates from the R-entry reverse
  // lists rv_id (k_rev_final), as k_jcand takes them (CAP must hold
  // K * K + R + 2 K R); without REV the stage is not compiled in.
  constexpr int HS = 1 << HSBITS;
  static_assert(HS * 4 >= CAP * 2 + CAP * 4, "selection arrays must fit in the hash area");
  __shared__ __attribute__((aligned(16))) int32_t hs[HS];
  __shared__ __attribute__((aligned(16))) int32_t cid[CAP];
  __shared__ uint16_t ch[CAP];
  __shared__ __attribute__((aligned(16))) int8_t own_row[kMaxDim];
  __shared__ JoinScratch S;
  uint16_t* sel_idx = reinterpret_cast<uint16_t*>(hs);            // [CAP]
  int32_t* sel_sc = hs + CAP / 2;                                 // [CAP]
  uint16_t* surv = ch;                                            // [CAP] after selection

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // DIM > 0: the row length is a compile-time constant (d = DIM), which turns
  // the per-chunk bounds of the sketch and row gathers into straight-line code
  if (DIM > 0) d = DIM;
  const int words = DIM > 0 ? DIM / 64 : d >> 6, nch = DIM > 0 ? DIM / 16 : d >> 4;
  // the block's counters (changed slots, candidates, selected, new ids,
  // sketched candidates), thread 0 only; LDS keeps them out of the register
  // budget, where they would stay live through every phase of the row loop
  __shared__ unsigned long long my_ctr[kCounterFields];

  // Schedule: blocks are dealt to the 8 XCDs round-robin, so XCD x = block % 8
  // takes the contiguous slice [x*n/8, (x+1)*n/8) of the processing order and
  // its blocks walk it together, each taking its next row from the slice's
  // queue counter rowq[32x] (one atomic per row, issued a row ahead).  With
  // `order` = tree 0's leaf order, concurrently processed rows share most
  // neighbours, which then hit that XCD's L2.  Blocks that start late (the
  // grid may exceed what is resident) or draw cheap rows simply take more or
  // fewer rows, so the slices drain together.  The result does not depend on
  // the order (Jacobi iteration).
  __shared__ int64_t s_row;
  // the next row's queue slot, parked in LDS as soon as the atomic returns: a
  // register held through the whole row would be spilled (a scratch store per
  // thread and row) by the small-row kernel
  __shared__ unsigned long long s_next;
  const int xcd = blockIdx.x & 7;
  const int64_t nwork = nwork_dev ? *nwork_dev : n;
  const int64_t lo = nwork * xcd / 8, hi = nwork * (xcd + 1) / 8;
  // per-build counts of the current run of rows (thread 0 only; LDS keeps
  // them out of the register budget)
  __shared__ int seg_cur;
  __shared__ unsigned long long seg_upd, seg_ins;
  if (tid == 0) {
    seg_cur = -1;
    seg_upd = seg_ins = 0;
  }
  if (tid < kCounterFields) my_ctr[tid] = 0;
  unsigned long long* const qctr = rowq + 32 * xcd + (SMALL ? 1 : 0);
  const int lim = SMALL ? min(M, CAP - 256) : CAP;
  if (tid == 0) s_row = lo + int64_t(atomicAdd(qctr, 1ull));
  __syncthreads();
  for (int64_t j = s_row; j < hi; j = s_row) {
    unsigned long long next_row = 0;
    if (tid == 0) next_row = atomicAdd(qctr, 1ull);
    const int64_t i = order ? int64_t(order[j]) : j;
    // ---- 0/1: reset, own list
    for (int j = tid; j < HS; j += 256) hs[j] = -1;
    if constexpr (!SMALL)
      for (int j = tid; j <= d; j += 256) S.hist[j] = 0;
    if (tid < 6) S.cnt[tid] = 0;
    if (masks_out && tid < T && tid < kMaskTrees) S.own_leaf[tid] = leafkey[i * TP + tid];
    if (tid < K) S.own_mask[tid] = masks_in ? masks_in[i * K + tid] : uint8_t(0);
    if (REV && tid < R) {
      S.rv_id[tid] = rv_id[i * R + tid];
      S.rv_mk[tid] = rv_mask ? int32_t(rv_mask[i * R + tid]) : 0;
    }
    for (int c = tid; c < nch; c += 256)
      *reinterpret_cast<v4i*>(own_row + 16 * c) = *reinterpret_cast<const v4i*>(q + i * d + 16 * c);
    if (tid < words) S.own_sk[tid] = sk[i * words + tid];
    if (tid < K) {
      const int32_t oid = ids_in[i * K + tid], osc = sc_in[i * K + tid];
      S.own_id[tid] = oid;
      S.own_sc[tid] = osc;
      S.own_key[tid] = pack_key(osc, oid);
    }
    if (tid < K) S.own_bit[tid] = fresh_in ? uint8_t((fresh_in[i] >> tid) & 1ull) : uint8_t(1);
    // wave 0 waits for its own-row loads here anyway, so the atomic's latency is hidden
    if (tid == 0) s_next = next_row;
    __syncthreads();
    if (tid < K) hs_insert(hs, HSBITS, S.own_id[tid]);
    if (tid == K) hs_insert(hs, HSBITS, int32_t(i));
    __syncthreads();
    // ---- 2: 2-hop candidates; each wave loads all its neighbour rows first
    {
      constexpr int kRowsPerWave = kMaxK / 4;
      int32_t c[kRowsPerWave];
#pragma unroll
      for (int r = 0; r < kRowsPerWave; ++r) {
        const int k = w + 4 * r;
        c[r] = -1;
        if (k < K && lane < K) {
          const int32_t n1 = S.own_id[k];
          const bool take = S.own_bit[k] != 0 || !fresh_in || ((fresh_in[n1] >> lane) & 1ull);
          if (take) {
            c[r] = ids_in[int64_t(n1) * K + lane];
            if (masks_in && (masks_in[int64_t(n1) * K + lane] & S.own_mask[k])) c[r] = -2;   // leaf-mate path
          }
        }
      }
      if (masks_in) {
        int dropped = 0;
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) dropped += __popcll(__ballot(c[r] == -2));
        if (lane == 0 && dropped) atomicAdd(&S.cnt[5], dropped);
      }
#pragma unroll
      for (int r = 0; r < kRowsPerWave; ++r) {
        // the row overflows (a wave-uniform read): insert nothing more (a guard,
        // not a break, so the loop unrolls and c[] stays in fixed registers)
        if (SMALL && S.cnt[0] > lim) c[r] = -1;
        const bool fresh_id = c[r] >= 0 && hs_insert(hs, HSBITS, c[r]);
        const int slot = wave_append(fresh_id, &S.cnt[0]);
        if (fresh_id) cid[slot] = c[r];
      }
    }
    // ---- 2b: reverse-neighbour candidates: the row's reverse entries u (path
    //          i <- u), its neighbours' reverse entries w (i -> n1 <- w) and the
    //          list entries w of its reverse entries (i <- u -> w); every reverse
    //          edge is new, and a path is dropped when its edges' mask bytes meet
    //          (i <- u: when the edge's own byte is not 0)
    if (REV && R > 0) {
      const int nrev = R + 2 * K * R;
      int dropped = 0;
      for (int e0 = 0; e0 < nrev; e0 += 256) {
        const int e = e0 + tid;
        int32_t id = -1;
        if (e < R) {
          id = S.rv_id[e];
          if (id >= 0 && S.rv_mk[e]) id = -2;
        } else if (e < R + K * R) {
          const int e2 = e - R, k = e2 / R;
          const int64_t at = int64_t(S.own_id[k]) * R + (e2 - k * R);
          id = rv_id[at];
          if (id >= 0 && rv_mask && (int32_t(rv_mask[at]) & int32_t(S.own_mask[k]))) id = -2;
        } else if (e < nrev) {
          const int e3 = e - R - K * R, jr = e3 / K;
          const int32_t u = S.rv_id[jr];
          if (u >= 0) {
            const int64_t at = int64_t(u) * K + (e3 - jr * K);
            id = ids_in[at];
            if (masks_in && (int32_t(masks_in[at]) & S.rv_mk[jr])) id = -2;
          }
        }
        dropped += __popcll(__ballot(id == -2));
        if (SMALL && S.cnt[0] > lim) id = -1;
        const bool fresh_id = id >= 0 && hs_insert(hs, HSBITS, id);
        const int slot = wave_append(fresh_id, &S.cnt[0]);
        if (fresh_id) cid[slot] = id;
      }
      if (masks_in && lane == 0 && dropped) atomicAdd(&S.cnt[5], dropped);
    }
    __syncthreads();
    int U = S.cnt[0];
    if (SMALL && U > lim) {   // uniform: the full kernel runs this row
      if (tid == 0) ovf[atomicAdd(ovf_n, 1ull)] = int32_t(i);
    } else {
      if constexpr (SMALL) {
        // at most M candidates: the selection is the identity
        for (int u = tid; u < U; u += 256) sel_idx[u] = uint16_t(u);
        if (tid == 0) S.cnt[1] = U;
      } else {
        // ---- 3: sketch similarity, 8 lanes per candidate, 8 candidates per group
        //         in flight (16 B per lane each).  With at most M candidates the
        //         threshold is 0 and every candidate is re-scored whatever its
        //         sketch, so the sketches are not gathered.
        if (U > M) {
          constexpr int kSkPer = SKPER;
          const int sub = tid & 7, grp = tid >> 3;
          uint64_t o0 = 0, o1 = 0;
          if (2 * sub < words) o0 = S.own_sk[2 * sub];
          if (2 * sub + 1 < words) o1 = S.own_sk[2 * sub + 1];
          for (int u0 = 0; u0 < U; u0 += 32 * kSkPer) {
            uint64_t a[kSkPer], b[kSkPer];
    #pragma unroll
            for (int j = 0; j < kSkPer; ++j) {
              const int u = u0 + grp + 32 * j;
              a[j] = b[j] = 0;
              if (u < U) {
                const uint64_t* cs = sk + int64_t(cid[u]) * words;
                if (2 * sub < words) a[j] = cs[2 * sub] ^ o0;
                if (2 * sub + 1 < words) b[j] = cs[2 * sub + 1] ^ o1;
              }
            }
    #pragma unroll
            for (int j = 0; j < kSkPer; ++j) {
              const int u = u0 + grp + 32 * j;
              int pc = __popcll(a[j]) + __popcll(b[j]);
              pc += __shfl_xor(pc, 1);
              pc += __shfl_xor(pc, 2);
              pc += __shfl_xor(pc, 4);
              if (u < U && sub == 0) {
                const int h = d - pc;
                ch[u] = uint16_t(h);
                atomicAdd(&S.hist[h], 1);
              }
            }
          }
        }
        __syncthreads();
        // ---- 4: threshold (wave 0): largest t with count(h >= t) >= M
        if (w == 0) {
          int thr = 0;
          if (U > M) {
            const int per = (d + 1 + 63) / 64;
            const int hi = d - lane * per;            // this lane's bins: hi, hi-1, ..., hi-per+1
            int local = 0;
            for (int b = 0; b < per; ++b)
              if (hi - b >= 0) local += S.hist[hi - b];
            int inc = local;
            for (int o = 1; o < 64; o <<= 1) {
              const int t = __shfl_up(inc, o);
              if (lane >= o) inc += t;
            }
            const int excl = inc - local;
            const unsigned long long hit = __ballot(inc >= M && excl < M);
            const int src = __ffsll(hit) - 1;
            int t_lane = 0;
            if (lane == src) {
              int acc = excl;
              for (int b = 0; b < per; ++b) {
                const int bin = hi - b;
                if (bin < 0) break;
                acc += S.hist[bin];
                if (acc >= M) {
                  t_lane = bin;
                  break;
                }
              }
            }
            thr = __shfl(t_lane, src);
          }
          if (lane == 0) S.cnt[3] = thr;
        }
        __syncthreads();
        // ---- selection
        {
          const int thr = S.cnt[3];
          for (int u0 = 0; u0 < U; u0 += 256) {
            const int u = u0 + tid;
            const bool take = u < U && (U <= M || ch[u] >= thr);   // U <= M: every candidate
            const int s = wave_append(take, &S.cnt[1]);
            if (take) sel_idx[s] = uint16_t(u);
          }
        }
      }
      __syncthreads();
      const int NS = S.cnt[1];
      // ---- 5: int8 dots, 16 lanes per candidate, 4 candidates per wave step
      {
        const int g = lane >> 4, s16 = lane & 15;
        // this lane's four 16-byte chunks of the own row: held in registers by
        // the full kernel; the small-row kernel (64 VGPRs) re-reads them from
        // LDS at each step instead
        auto own_chunk = [&](int c) -> v4i {
          const int chn = s16 + 16 * c;
          return chn < nch ? *reinterpret_cast<const v4i*>(own_row + 16 * chn) : v4i{0, 0, 0, 0};
        };
        v4i own[4];
        if constexpr (!SMALL) {
  #pragma unroll
          for (int c = 0; c < 4; ++c) own[c] = own_chunk(c);
        }
        // DOTPER candidates per 16-lane group in flight (64 B per lane each);
        // at 2 the kernel fits 128 VGPRs, i.e. four blocks per CU
        constexpr int kDotPer = DOTPER;
        for (int s0 = 0; s0 < NS; s0 += 16 * kDotPer) {
          v4i v[kDotPer][4];
  #pragma unroll
          for (int j = 0; j < kDotPer; ++j) {
            const int sj = s0 + w * 4 + g + 16 * j;
            const int8_t* rj = sj < NS ? q + int64_t(cid[sel_idx[sj]]) * d : nullptr;
  #pragma unroll
            for (int c = 0; c < 4; ++c) {
              const int chn = s16 + 16 * c;
              v[j][c] = (rj && chn < nch) ? *reinterpret_cast<const v4i*>(rj + 16 * chn) : v4i{0, 0, 0, 0};
            }
          }
  #pragma unroll
          for (int j = 0; j < kDotPer; ++j) {
            int dj = 0;
  #pragma unroll
            for (int c = 0; c < 4; ++c) dj = dot16_i8(v[j][c], SMALL ? own_chunk(c) : own[c], dj);
  #pragma unroll
            for (int m = 8; m >= 1; m >>= 1) dj += __shfl_xor(dj, m);
            const int sj = s0 + w * 4 + g + 16 * j;
            if (s16 == 0 && sj < NS) sel_sc[sj] = dj;
          }
        }
      }
      __syncthreads();
      // ---- 6a: survivors (fewer than K own entries precede them).  Keys order
      // (score desc, id asc); a candidate's id is never an own id, and own ids
      // are distinct, so key comparisons decide every pair the slot rule would.
      for (int s0 = 0; s0 < NS; s0 += 256) {
        const int s = s0 + tid;
        bool keep = false;
        if (s < NS) {
          const uint64_t key = pack_key(sel_sc[s], cid[sel_idx[s]]);
          int c = 0;
          for (int k = 0; k < K; ++k) c += S.own_key[k] > key ? 1 : 0;
          keep = c < K;
        }
        const int p = wave_append(keep, &S.cnt[2]);
        if (keep) surv[p] = uint16_t(s);
      }
      __syncthreads();
      const int P = S.cnt[2];
      // survivors' keys packed over cid[] (no longer needed once read) when they fit
      constexpr int kPackCap = CAP / 2, kPackPer = (kPackCap + 255) / 256;
      uint64_t* pkey = reinterpret_cast<uint64_t*>(cid);
      const bool packed = P <= kPackCap;
      if (packed) {
        uint64_t kk[kPackPer];
  #pragma unroll
        for (int j = 0; j < kPackPer; ++j) {
          const int p = tid + 256 * j;
          kk[j] = 0;
          if (p < P) {
            const int sv = surv[p];
            kk[j] = pack_key(sel_sc[sv], cid[sel_idx[sv]]);
          }
        }
        __syncthreads();
  #pragma unroll
        for (int j = 0; j < kPackPer; ++j) {
          const int p = tid + 256 * j;
          if (p < P) pkey[p] = kk[j];
        }
        __syncthreads();
      }
      // ---- 6b: ranks
      for (int e = tid; e < K + P; e += 256) {
        int32_t sc, id, sl;
        if (e < K) {
          sc = S.own_sc[e];
          id = S.own_id[e];
          sl = e;
        } else if (packed) {
          const uint64_t key = pkey[e - K];
          sc = key_score(key);
          id = key_id(key);
          sl = K;
        } else {
          const int s = surv[e - K];
          sc = sel_sc[s];
          id = cid[sel_idx[s]];
          sl = K;
        }
        int r = 0;
        for (int k = 0; k < K; ++k) r += before(S.own_sc[k], S.own_id[k], k, sc, id, sl) ? 1 : 0;
        if (packed) {
          const uint64_t key = pack_key(sc, id);
          for (int p = 0; p < P && r < K; ++p) r += pkey[p] > key ? 1 : 0;
        } else {
          for (int p = 0; p < P && r < K; ++p) {
            const int s = surv[p];
            r += before(sel_sc[s], cid[sel_idx[s]], K, sc, id, sl) ? 1 : 0;
          }
        }
        if (r < K) {
          S.new_id[r] = id;
          S.new_sc[r] = sc;
          S.new_fresh[r] = e >= K ? 1 : 0;
          if (masks_out)
            S.new_mask[r] = uint8_t(e < K ? S.own_mask[e] : edge_mask(S.own_leaf, leafkey + int64_t(id) * TP, T));
          if (e >= K) atomicAdd(&S.cnt[4], 1);   // a new id entered the list
        }
      }
      __syncthreads();
      // ---- 7: write back, count changes
      if (w == 0) {
        bool changed = false, fresh = false;
        if (lane < K) {
          ids_out[i * K + lane] = S.new_id[lane];
          sc_out[i * K + lane] = S.new_sc[lane];
          if (masks_out) masks_out[i * K + lane] = S.new_mask[lane];
          changed = S.new_id[lane] != S.own_id[lane];
          fresh = S.new_fresh[lane] != 0;
        }
        const unsigned long long m = __ballot(changed);
        const unsigned long long fm = __ballot(fresh);
        if (fresh_out && lane == 0) fresh_out[i] = fm;
        if (lane == 0) {
          my_ctr[0] += __popcll(m);
          my_ctr[1] += U;
          my_ctr[2] += NS;
          my_ctr[3] += S.cnt[4];
          my_ctr[5] += S.cnt[5];
          if (U > M) my_ctr[4] += U;
          if (segc && tid == 0) {
            const int sg = seg_of(seg_base, nseg, i);
            if (sg != seg_cur) {
              if (seg_cur >= 0 && (seg_upd | seg_ins)) {
                atomicAdd(&segc[2 * seg_cur], seg_upd);
                atomicAdd(&segc[2 * seg_cur + 1], seg_ins);
              }
              seg_cur = sg;
              seg_upd = seg_ins = 0;
            }
            seg_upd += __popcll(m);
            seg_ins += S.cnt[4];
          }
        }
      }
    }
    if (tid == 0) s_row = lo + int64_t(s_next);
    __syncthreads();
  }
  if (tid == 0 && segc && seg_cur >= 0 && (seg_upd | seg_ins)) {
    atomicAdd(&segc[2 * seg_cur], seg_upd);
    atomicAdd(&segc[2 * seg_cur + 1], seg_ins);
  }
  if (tid == 0) {
    unsigned long long* c = counters + kCounterFields * (blockIdx.x % kNumCounters);
    for (int f = 0; f < kCounterFields; ++f) atomicAdd(c + f, my_ctr[f]);
  }
}

// ---------------------------------------------------------------------------
// Reverse-neighbour lists (NN-Descent's reverse candidates).  Row v's reverse
// list holds the R best rows u, by (score desc, u asc), whose lists hold v as
// an entry the coming iteration treats as new (every entry in an iteration
// without the new/old rule, else the entries that came in during the previous
// one).  k_rev_insert offers every such edge u -> v to v's R slots of packed
// 64-bit (score, id) keys with a lock-free top-R insertion: find the slot
// holding the smallest key; stop if the offered key is not larger; else swap
// it in with a 64-bit compare-and-swap, and start over if another insertion
// changed that slot first.  A key leaves the slots only for a larger one, so
// once every offer has finished the slots hold exactly the R largest keys
// offered, whatever the order of the offers (keys of one v are distinct: a
// list holds v at most once).  Slots are read with agent-scope atomic loads,
// so an offer never works from a stale copy in another XCD's L2.
// k_rev_final sorts each row's slots, writes the ids (-1 past the last) and
// the mask byte of each edge u -> v (trees whose leaf u and v share, from the
// leaf keys; 0 without them), and zeroes the slots for the next iteration.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_rev_insert(const int32_t* __restrict__ gid,
                                                    const int32_t* __restrict__ gsc,
                                                    const uint64_t* __restrict__ fresh, int64_t n, int K,
                                                    int R, unsigned long long* __restrict__ slots) {
  const int64_t total = n * K;
  for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
    const int64_t u = e / K;
    const int k = int(e - u * K);
    if (fresh && !((fresh[u] >> k) & 1ull)) continue;
    const int32_t v = gid[e];
    if (v < 0 || v >= n) continue;
    const unsigned long long key = pack_key(gsc[e], int32_t(u));
    unsigned long long* sl = slots + int64_t(v) * R;
    while (true) {
      unsigned long long mn = ~0ull;
      int at = 0;
      for (int j = 0; j < R; ++j) {
        const unsigned long long x = __hip_atomic_load(sl + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (x < mn) {
          mn = x;
          at = j;
        }
      }
      if (key <= mn) break;                                  // R larger keys are already in
      if (atomicCAS(sl + at, mn, key) == mn) break;          // swapped in
    }
  }
}

__global__ __launch_bounds__(256) void k_rev_final(unsigned long long* __restrict__ slots, int64_t n, int R,
                                                   const int32_t* __restrict__ leafkey, int T, int TP,
                                                   int32_t* __restrict__ rv_id, uint8_t* __restrict__ rv_mask) {
  for (int64_t v = int64_t(blockIdx.x) * 256 + threadIdx.x; v < n; v += int64_t(gridDim.x) * 256) {
    unsigned long long key[kRevMax];
#pragma unroll
    for (int j = 0; j < kRevMax; ++j) key[j] = j < R ? slots[v * R + j] : 0ull;
    // descending insertion sort of at most kRevMax keys
#pragma unroll
    for (int j = 1; j < kRevMax; ++j)
#pragma unroll
      for (int t = j; t > 0; --t)
        if (key[t] > key[t - 1]) {
          const unsigned long long x = key[t];
          key[t] = key[t - 1];
          key[t - 1] = x;
        }
    for (int j = 0; j < R; ++j) {
      const int32_t u = key[j] ? int32_t(0xFFFFFFFFu - uint32_t(key[j] & 0xFFFFFFFFull)) : -1;
      rv_id[v * R + j] = u;
      rv_mask[v * R + j] = (u >= 0 && leafkey) ? uint8_t(edge_mask(leafkey + v * TP, leafkey + int64_t(u) * TP, T))
                                               : uint8_t(0);
      slots[v * R + j] = 0ull;
    }
  }
}

// ---------------------------------------------------------------------------
// Phased join (single builds): the iteration of k_join split into two
// wave-per-row kernels that hold little LDS, so each CU keeps 24-32 rows in
// flight instead of four, and the rows no wave can hold go to k_join.
//   k_jcand       own list -> 2-hop ids (new/old rule) -> leaf-mate rule ->
//                 dedup in a per-wave LDS hash -> sketch prefilter when more
//                 than M remain -> the row's candidates in cbuf[j][0..S),
//                 ccount[j] = S; a row with more than kJCap candidates gets
//                 ccount[j] = -1 and goes to ovf[] for k_join.
//   k_jdot_merge  int8 dots of the candidates (16 lanes per 1 KB row), rank
//                 merge with the own list by (score desc, id asc, slot),
//                 write-back, change / new-id counts.
// The lists are exactly k_join's: the candidate set, the prefilter threshold
// and the merge order are the same functions of the previous graph.  Here the
// leaf-mate rule is applied to every 2-hop id before deduplication (so the
// hash never holds leaf-mates); counter 5 then counts those 2-hop ids.
// ---------------------------------------------------------------------------

constexpr int kJCap = 512;        // candidates a phased-join row may hold
constexpr int kJHashBits = 10;    // per-wave LDS hash: 1024 slots (< 64 own + self + kJCap + 63 inserted)
constexpr int kJBatch = 8;        // neighbour lists whose ids k_jcand loads per batch

// blocks of 4 waves, dealt round-robin to the 8 XCDs: XCD x takes the
// contiguous slice [x*nwork/8, (x+1)*nwork/8) of the work list and its waves
// stride through it together (cluster-mates in flight share that XCD's L2)
// Rows of the wave-per-row join phases: XCD x = block % 8 takes the
// contiguous eighth [x n / 8, (x + 1) n / 8) of the work list (tree 0's leaf
// order), and each wave takes its next row from that eighth's queue counter
// rowq[32 x + slot] (lane 0's atomic, issued a row ahead), so the rows in
// flight on an XCD stay one window of consecutive cluster-mates whose
// neighbour lists, sketches and rows share the XCD's L2.  (With a fixed
// stride per wave the waves drift apart over a launch and the window widens.)
// Slots 0 and 1 belong to k_join; k_sum_counters clears them all after each
// iteration.
constexpr int kQJcand = 2, kQJdot = 3;
struct WaveQueue {
  unsigned long long* ctr;
  int64_t lo, hi;
};
__device__ __forceinline__ WaveQueue wave_queue(int64_t nwork, unsigned long long* rowq, int slot) {
  const int xcd = blockIdx.x & 7;
  return {rowq + 32 * xcd + slot, nwork * xcd / 8, nwork * (xcd + 1) / 8};
}
__device__ __forceinline__ unsigned int wq_claim(const WaveQueue& q) {
  return (threadIdx.x & 63) == 0 ? unsigned(atomicAdd(q.ctr, 1ull)) : 0u;
}
__device__ __forceinline__ int64_t wq_row(const WaveQueue& q, unsigned int t) {
  return q.lo + int64_t(unsigned(__builtin_amdgcn_readfirstlane(int(t))));
}

template <bool REV>   // the reverse-candidate stage compiled in (used when R > 0)
__global__ __launch_bounds__(256, 6) void k_jcand(
    const uint64_t* __restrict__ sk, int d, int K, int M, const int32_t* __restrict__ ids_in,
    const uint64_t* __restrict__ fresh_in, const uint8_t* __restrict__ masks_in,
    const int32_t* __restrict__ order, int64_t nwork, int32_t* __restrict__ cbuf,
    int32_t* __restrict__ ccount, int32_t* __restrict__ ovf, unsigned long long* __restrict__ ovf_n,
    unsigned long long* __restrict__ counters, unsigned long long* __restrict__ rowq,
    const int32_t* __restrict__ ids_rv, const uint8_t* __restrict__ rv_mask, int R) {
  __shared__ int32_t hs_w[4][1 << kJHashBits];   // dedup hash, then (low half) sketch similarities
  __shared__ int32_t cid_w[4][kJCap];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t* hs = hs_w[w];
  int32_t* cid = cid_w[w];
  const int words = d >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long n_u = 0, n_sk = 0, n_lk = 0;   // candidates, sketched candidates, paths dropped
  const WaveQueue wq = wave_queue(nwork, rowq, kQJcand);
  unsigned int tk = wq_claim(wq);
  for (int64_t j = wq_row(wq, tk); j < wq.hi; j = wq_row(wq, tk)) {
    tk = wq_claim(wq);   // the wave's next row, in flight during this one
    const int64_t i = order ? int64_t(order[j]) : j;
    for (int s = lane; s < (1 << kJHashBits); s += 64) hs[s] = -1;
    const int32_t oid = lane < K ? ids_in[i * K + lane] : -1;
    const bool obit = !fresh_in || (lane < K && ((fresh_in[i] >> lane) & 1ull));
    const int32_t omask = (masks_in && lane < K) ? int32_t(masks_in[i * K + lane]) : 0;
    __builtin_amdgcn_wave_barrier();
    if (lane < K) hs_insert(hs, kJHashBits, oid);
    if (lane == 0) hs_insert(hs, kJHashBits, int32_t(i));   // K may be 64: no spare lane
    __builtin_amdgcn_wave_barrier();
    int U = 0;
    // the neighbours' fresh masks, lane k holding neighbour k's (one load per row)
    const uint64_t fnb = (fresh_in && lane < K) ? fresh_in[oid] : ~0ull;
    // 2-hop ids of kJBatch neighbour lists at a time (batch b + 1 loads while
    // batch b is hashed); the leaf-mate rule drops the path i -> n1 -> c when
    // both of its edges share a leaf in the same tree (their mask bytes meet)
    int32_t cn[kJBatch];
    auto load_ids = [&](int k0, int32_t (&dst)[kJBatch]) {
#pragma unroll
      for (int u = 0; u < kJBatch; ++u) {
        const int k = k0 + u;
        dst[u] = -1;
        if (k < K) {
          const int32_t n1 = __shfl(oid, k);
          const bool take = __shfl(int(obit), k) != 0 || ((__shfl(fnb, k) >> lane) & 1ull);
          if (lane < K && take) {
            const int32_t id = ids_in[int64_t(n1) * K + lane];
            const int32_t m = masks_in ? int32_t(masks_in[int64_t(n1) * K + lane]) : 0;
            // -2: a path the leaf-mate rule drops (counted, never hashed)
            dst[u] = (m & __shfl(omask, k)) ? -2 : id;
          }
        }
      }
    };
    load_ids(0, cn);
    for (int k0 = 0; k0 < K && U <= kJCap; k0 += kJBatch) {
      int32_t c[kJBatch];
#pragma unroll
      for (int u = 0; u < kJBatch; ++u) c[u] = cn[u];
      if (k0 + kJBatch < K) load_ids(k0 + kJBatch, cn);
#pragma unroll
      for (int u = 0; u < kJBatch; ++u) {
        n_lk += __popcll(__ballot(c[u] == -2));
        if (U > kJCap) c[u] = -1;   // the row overflows: insert nothing more (wave-uniform)
        const bool nw = c[u] >= 0 && hs_insert(hs, kJHashBits, c[u]);
        const unsigned long long bal = __ballot(nw);
        const int pos = U + __popcll(bal & lt);
        if (nw && pos < kJCap) cid[pos] = c[u];
        U += __popcll(bal);
      }
    }
    // reverse-neighbour candidates (R > 0): the row's reverse entries u (path
    // i <- u), its neighbours' reverse entries w (i -> n1 <- w) and the list
    // entries w of its reverse entries (i <- u -> w); every reverse edge is
    // new, and a path is dropped when its edges' mask bytes meet (i <- u: when
    // the edge's own byte is not 0).  Lane j < R holds reverse entry j; each
    // batch of 256 (entry, lane) pairs is loaded before any of it is hashed.
    if (REV && R > 0) {
      const int32_t rid = lane < R ? ids_rv[i * R + lane] : -1;
      const int32_t rmk = (lane < R && rv_mask) ? int32_t(rv_mask[i * R + lane]) : 0;
      const int nrev = R + 2 * K * R;
      for (int e0 = 0; e0 < nrev && U <= kJCap; e0 += 256) {
        int32_t c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + 64 * u + lane;
          const bool r1 = e < R;
          const int e2 = e - R;
          const bool r2 = !r1 && e2 < K * R;
          const int e3 = e2 - K * R;
          const bool r3 = !r1 && !r2 && e3 < R * K;
          const int k2 = r2 ? e2 / R : 0, j3 = r3 ? e3 / K : 0;
          const int src = r1 ? e : (r2 ? k2 : j3);   // the lane holding this pair's source
          const int32_t s_oid = __shfl(oid, src), s_om = __shfl(omask, src);
          const int32_t s_rid = __shfl(rid, src), s_rmk = __shfl(rmk, src);
          int32_t id = -1;
          if (r1) {
            id = (s_rid >= 0 && s_rmk) ? -2 : s_rid;
          } else if (r2) {
            const int64_t at = int64_t(s_oid) * R + (e2 - k2 * R);
            const int32_t w2 = ids_rv[at];
            id = (w2 >= 0 && rv_mask && (int32_t(rv_mask[at]) & s_om)) ? -2 : w2;
          } else if (r3 && s_rid >= 0) {
            const int64_t at = int64_t(s_rid) * K + (e3 - j3 * K);
            const int32_t w3 = ids_in[at];
            id = (masks_in && (int32_t(masks_in[at]) & s_rmk)) ? -2 : w3;
          }
          c[u] = id;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          n_lk += __popcll(__ballot(c[u] == -2));
          if (U > kJCap) c[u] = -1;
          const bool nw = c[u] >= 0 && hs_insert(hs, kJHashBits, c[u]);
          const unsigned long long bal = __ballot(nw);
          const int pos = U + __popcll(bal & lt);
          if (nw && pos < kJCap) cid[pos] = c[u];
          U += __popcll(bal);
        }
      }
    }
    if (U > kJCap) {   // wave-uniform: k_join runs this row
      if (lane == 0) {
        ccount[j] = -1;
        ovf[atomicAdd(ovf_n, 1ull)] = int32_t(i);
      }
      continue;
    }
    n_u += U;
    __builtin_amdgcn_wave_barrier();
    int S = U;
    constexpr int kPer = kJCap / 64;
    int32_t cv[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) cv[e] = lane + 64 * e < U ? cid[lane + 64 * e] : -1;
    bool keep[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) keep[e] = cv[e] >= 0;
    if (U > M) {
      // sketch similarity h = d - popcount, 8 lanes per candidate (128 B)
      n_sk += U;
      const int sub = lane & 7, grp = lane >> 3;
      const uint64_t o0 = 2 * sub < words ? sk[i * words + 2 * sub] : 0ull;
      const uint64_t o1 = 2 * sub + 1 < words ? sk[i * words + 2 * sub + 1] : 0ull;
      for (int u0 = 0; u0 < U; u0 += 32) {
        uint64_t a[4], b[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = u0 + grp + 8 * r;
          a[r] = b[r] = 0ull;
          if (u < U) {
            const uint64_t* cs = sk + int64_t(cid[u]) * words;
            if (2 * sub < words) a[r] = cs[2 * sub] ^ o0;
            if (2 * sub + 1 < words) b[r] = cs[2 * sub + 1] ^ o1;
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = u0 + grp + 8 * r;
          int pc = __popcll(a[r]) + __popcll(b[r]);
          pc += __shfl_xor(pc, 1);
          pc += __shfl_xor(pc, 2);
          pc += __shfl_xor(pc, 4);
          if (u < U && sub == 0) hs[u] = d - pc;
        }
      }
      __builtin_amdgcn_wave_barrier();
      int hv[kPer];
#pragma unroll
      for (int e = 0; e < kPer; ++e) hv[e] = keep[e] ? hs[lane + 64 * e] : -1;
      // largest t with |{h >= t}| >= M
      int lo = 0, hi = d;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        int cnt = 0;
#pragma unroll
        for (int e = 0; e < kPer; ++e) cnt += __popcll(__ballot(hv[e] >= mid));
        if (cnt >= M)
          lo = mid;
        else
          hi = mid - 1;
      }
#pragma unroll
      for (int e = 0; e < kPer; ++e) keep[e] = keep[e] && hv[e] >= lo;
    }
    // the row's candidates to cbuf, in list order
    int base = 0;
    int32_t* dst = cbuf + j * kJCap;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const unsigned long long bal = __ballot(keep[e]);
      if (keep[e]) dst[base + __popcll(bal & lt)] = cv[e];
      base += __popcll(bal);
    }
    S = base;
    if (lane == 0) ccount[j] = S;
  }
  if (lane == 0) {
    unsigned long long* cc = counters + kCounterFields * (blockIdx.x % kNumCounters);
    atomicAdd(cc + 1, n_u);
    atomicAdd(cc + 4, n_sk);
    atomicAdd(cc + 5, n_lk);
  }
}

template <int DIM>   // DIM > 0: rows of DIM bytes at compile time; 0: any d
__global__ __launch_bounds__(256, 4) void k_jdot_merge(
    const int8_t* __restrict__ q, int d, int K, const int32_t* __restrict__ ids_in,
    const int32_t* __restrict__ sc_in, int32_t* __restrict__ ids_out, int32_t* __restrict__ sc_out,
    uint64_t* __restrict__ fresh_out, const uint8_t* __restrict__ masks_in, uint8_t* __restrict__ masks_out,
    const int32_t* __restrict__ leafkey, int T, int TP, const int32_t* __restrict__ order, int64_t nwork,
    const int32_t* __restrict__ cbuf, const int32_t* __restrict__ ccount,
    unsigned long long* __restrict__ counters, unsigned long long* __restrict__ rowq) {
  __shared__ uint64_t key_w[4][kJCap];   // candidate keys, then the survivors' keys
  __shared__ uint64_t okey_w[4][kMaxK];
  __shared__ int32_t nid_w[4][kMaxK], nsc_w[4][kMaxK], nfr_w[4][kMaxK];
  __shared__ int32_t oleaf_w[4][kMaskTrees];
  __shared__ __attribute__((aligned(16))) int8_t orow_w[4][kMaxDim];   // the row's int8 data
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int8_t* orow = orow_w[w];
  uint64_t* key = key_w[w];
  uint64_t* okey = okey_w[w];
  int32_t* nid = nid_w[w];
  int32_t* nsc = nsc_w[w];
  int32_t* nfr = nfr_w[w];   // 1: from the candidates; else the own entry's mask byte << 1
  int32_t* oleaf = oleaf_w[w];
  if (DIM > 0) d = DIM;
  const int nch = DIM > 0 ? DIM / 16 : d >> 4;
  const int g = lane >> 4, s16 = lane & 15;
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long n_chg = 0, n_ns = 0, n_ins = 0;
  const WaveQueue wq = wave_queue(nwork, rowq, kQJdot);
  unsigned int tk = wq_claim(wq);
  for (int64_t j = wq_row(wq, tk); j < wq.hi; j = wq_row(wq, tk)) {
    tk = wq_claim(wq);   // the wave's next row, in flight during this one
    const int S = ccount[j];
    if (S < 0) continue;   // k_join runs this row
    const int64_t i = order ? int64_t(order[j]) : j;
    const int32_t oid = lane < K ? ids_in[i * K + lane] : -1;
    const int32_t osc = lane < K ? sc_in[i * K + lane] : 0;
    const int32_t om = (masks_in && lane < K) ? int32_t(masks_in[i * K + lane]) : 0;
    const uint64_t ok = lane < K ? pack_key(osc, oid) : ~0ull;
    if (lane < K) okey[lane] = ok;
    if (masks_out && lane < T && lane < kMaskTrees) oleaf[lane] = leafkey[i * TP + lane];
    // the K-th best own key: a candidate below it cannot enter the list
    uint64_t mn = ok;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const uint64_t o = __shfl_xor(mn, m);
      mn = o < mn ? o : mn;
    }
    // ---- int8 dots: 16 lanes per candidate, four candidates per group in flight
    //      (the kernel is bound by rows in flight per CU)
    constexpr int ND = 2;
    // the row's own data in LDS (read again at each step: registers go to rows in flight)
    for (int ch = lane; ch < nch; ch += 64)
      *reinterpret_cast<v4i*>(orow + 16 * ch) = *reinterpret_cast<const v4i*>(q + i * d + 16 * ch);
    __builtin_amdgcn_wave_barrier();
    auto own = [&](int c) -> v4i {
      const int ch = s16 + 16 * c;
      return ch < nch ? *reinterpret_cast<const v4i*>(orow + 16 * ch) : v4i{0, 0, 0, 0};
    };
    const int32_t* cl = cbuf + j * kJCap;
    for (int s0 = 0; s0 < S; s0 += 4 * ND) {
      int32_t id[ND];
      v4i v[ND][4];
#pragma unroll
      for (int r = 0; r < ND; ++r) {
        const int s = s0 + g + 4 * r;
        id[r] = s < S ? cl[s] : -1;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int ch = s16 + 16 * c;
          v[r][c] = (id[r] >= 0 && ch < nch) ? *reinterpret_cast<const v4i*>(q + int64_t(id[r]) * d + 16 * ch)
                                             : v4i{0, 0, 0, 0};
        }
      }
#pragma unroll
      for (int r = 0; r < ND; ++r) {
        int dt = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) dt = dot16_i8(v[r][c], own(c), dt);
#pragma unroll
        for (int m = 8; m >= 1; m >>= 1) dt += __shfl_xor(dt, m);
        const int s = s0 + g + 4 * r;
        if (s16 == 0 && s < S) key[s] = pack_key(dt, id[r]);
      }
    }
    __builtin_amdgcn_wave_barrier();
    // ---- survivors (key above the K-th own key), compacted in place one
    //      64-entry chunk at a time (a chunk is read before any of it is
    //      overwritten, and writes land below the next chunk)
    int P = 0;
    for (int e0 = 0; e0 < S; e0 += 64) {
      const uint64_t kv = e0 + lane < S ? key[e0 + lane] : 0ull;
      __builtin_amdgcn_wave_barrier();
      const bool sv = kv > mn;
      const unsigned long long bal = __ballot(sv);
      if (sv) key[P + __popcll(bal & lt)] = kv;
      P += __popcll(bal);
      __builtin_amdgcn_wave_barrier();
    }
    // ---- ranks in the union (own entries tie-break by slot; candidate ids are
    //      neither own ids nor repeated, so their keys are distinct)
    if (lane < K) {
      int r = 0;
      for (int k = 0; k < K; ++k) {
        const uint64_t o = okey[k];
        r += int(o > ok) | (int(o == ok) & int(k < lane));
      }
      for (int p = 0; p < P && r < K; ++p) r += key[p] > ok ? 1 : 0;
      if (r < K) {
        nid[r] = oid;
        nsc[r] = osc;
        nfr[r] = om << 1;
      }
    }
    for (int p = lane; p < P; p += 64) {
      const uint64_t kp = key[p];
      int r = 0;
      for (int k = 0; k < K; ++k) r += okey[k] > kp ? 1 : 0;
      for (int p2 = 0; p2 < P && r < K; ++p2) r += key[p2] > kp ? 1 : 0;
      if (r < K) {
        nid[r] = key_id(kp);
        nsc[r] = key_score(kp);
        nfr[r] = 1;
      }
    }
    __builtin_amdgcn_wave_barrier();
    // ---- write back (a new entry's mask byte from its leaf keys)
    bool changed = false, fr = false;
    if (lane < K) {
      const int32_t id = nid[lane];
      ids_out[i * K + lane] = id;
      sc_out[i * K + lane] = nsc[lane];
      changed = id != oid;
      fr = (nfr[lane] & 1) != 0;
      if (masks_out)
        masks_out[i * K + lane] = uint8_t(fr ? edge_mask(oleaf, leafkey + int64_t(id) * TP, T) : nfr[lane] >> 1);
    }
    const unsigned long long fm = __ballot(fr);
    if (fresh_out && lane == 0) fresh_out[i] = fm;
    n_chg += __popcll(__ballot(changed));
    n_ins += __popcll(fm);
    n_ns += S;
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) {
    unsigned long long* cc = counters + kCounterFields * (blockIdx.x % kNumCounters);
    atomicAdd(cc + 0, n_chg);
    atomicAdd(cc + 2, n_ns);
    atomicAdd(cc + 3, n_ins);
  }
}

// Sums the per-block counters of a k_join launch into out[0..kCounterFields)
// and leaves the counters and the row-queue counters behind them at zero for
// the next launch (no memset node, so the sequence can be captured in a graph).
__global__ void k_sum_counters(unsigned long long* __restrict__ c,
                               unsigned long long* __restrict__ out) {
  constexpr int F = kCounterFields;
  __shared__ unsigned long long part[F][256];
  unsigned long long a[F] = {};
  for (int j = threadIdx.x; j < kNumCounters; j += 256)
    for (int f = 0; f < F; ++f) {
      a[f] += c[F * j + f];
      c[F * j + f] = 0;
    }
  if (threadIdx.x < kRowQueueWords) c[F * kNumCounters + threadIdx.x] = 0;
  for (int f = 0; f < F; ++f) part[f][threadIdx.x] = a[f];
  __syncthreads();
  if (threadIdx.x < F) {
    unsigned long long s = 0;
    for (int j = 0; j < 256; ++j) s += part[threadIdx.x][j];
    out[threadIdx.x] = s;
  }
}


// ---------------------------------------------------------------------------
// Packed builds: the join's work list is tree 0's leaf order restricted to the
// segments still iterating.  Segment s occupies positions [seg_base[s],
// seg_base[s+1]) of every tree's perm (no node crosses a segment), so the
// list is a concatenation of those ranges.
// ---------------------------------------------------------------------------

// dst[s] = first list position of segment s (active ones), *nwork = list length
__global__ __launch_bounds__(1024) void k_seg_scan(const int32_t* __restrict__ act,
                                                   const int64_t* __restrict__ seg_base, int nseg,
                                                   int64_t* __restrict__ dst, int64_t* __restrict__ nwork) {
  __shared__ int64_t part[1024];
  const int tid = threadIdx.x;
  // each thread owns a contiguous run of segments
  const int per = (nseg + 1023) / 1024, s0 = tid * per, s1 = min(nseg, s0 + per);
  int64_t sum = 0;
  for (int s = s0; s < s1; ++s) sum += act[s] ? seg_base[s + 1] - seg_base[s] : 0;
  part[tid] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {   // inclusive scan of the runs
    const int64_t v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int64_t run = part[tid] - sum;
  for (int s = s0; s < s1; ++s) {
    dst[s] = run;
    run += act[s] ? seg_base[s + 1] - seg_base[s] : 0;
  }
  if (tid == 1023) *nwork = part[1023];
}

__global__ __launch_bounds__(256) void k_seg_gather(const int32_t* __restrict__ perm0, int64_t n,
                                                    const int64_t* __restrict__ seg_base, int nseg,
                                                    const int32_t* __restrict__ act,
                                                    const int64_t* __restrict__ dst,
                                                    int32_t* __restrict__ list) {
  for (int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x; p < n; p += int64_t(gridDim.x) * 256) {
    const int s = seg_of(seg_base, nseg, p);
    if (act[s]) list[dst[s] + (p - seg_base[s])] = perm0[p];
  }
}

// snap[j] = c[j]; c[j] = 0 (per-segment join counters of one iteration)
__global__ __launch_bounds__(256) void k_seg_snapshot(unsigned long long* __restrict__ c,
                                                      unsigned long long* __restrict__ snap, int m) {
  for (int j = blockIdx.x * 256 + threadIdx.x; j < m; j += gridDim.x * 256) {
    snap[j] = c[j];
    c[j] = 0;
  }
}

// Cell forest level on the device: row ids, cell sizes, per-position nodes.
__global__ __launch_bounds__(256) void k_iota(int32_t* __restrict__ v, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    v[i] = int32_t(i);
}

__global__ __launch_bounds__(256) void k_cell_count(const int32_t* __restrict__ cell, int64_t n, int C,
                                                    int32_t* __restrict__ cnt) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
    const int32_t c = cell[i];
    if (c >= 0 && c < C) atomicAdd(&cnt[c], 1);
  }
}

// nodepos[p] = cellnode[cell of the row at position p] (rows sorted by cell)
__global__ __launch_bounds__(256) void k_cell_nodepos(const int32_t* __restrict__ sorted_cell, int64_t n,
                                                      const int32_t* __restrict__ cellnode,
                                                      int32_t* __restrict__ nodepos) {
  for (int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x; p < n; p += int64_t(gridDim.x) * 256)
    nodepos[p] = cellnode[sorted_cell[p]];
}

// ---------------------------------------------------------------------------
// K8: export.  Plain: src/main.cpp:449-457.  Rerank: FP32 dots of the first R
// list entries, one wave per point, best k by (score desc, id asc, slot).
// ---------------------------------------------------------------------------

// rows [row_base, row_base + n) of the graph; ids are written relative to
// row_base (a packed build's own numbering; 0 for a single build)
__global__ void k_export_plain(const int32_t* __restrict__ gid, const int32_t* __restrict__ gsc,
                               int64_t n, int K, int k, int32_t* __restrict__ knns,
                               float* __restrict__ dists, int64_t row_base) {
  const int64_t total = n * k;
  gid += row_base * K;
  gsc += row_base * K;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = e / k;
    const int j = int(e - i * k);
    knns[i * k + (k - 1 - j)] = int32_t(gid[i * K + j] - row_base) + 1;
    dists[i * k + (k - 1 - j)] = float(gsc[i * K + j]);
  }
}

__global__ void k_zero_u64(unsigned long long* __restrict__ v, int m) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) v[j] = 0ull;
}

// waves per rerank block: a block takes this many consecutive rows of its
// XCD's slice per step (one queue atomic per step).  4 and 8 time alike; 16
// (one 1024-thread block per CU at 7 waves per SIMD) is slower.
constexpr int kRerankWaves = 4;

// FP32 rerank in dimension slices, one launch per slice.  Pass s adds the
// FP32 dot products over dims [64 CH s, 64 CH (s + 1)) of the row's first R
// list entries to part[i * R + r]; the last pass ranks the sums
// (((p0 + p1) + p2) + ...) by (score desc, id asc, slot asc) and writes the
// row's best k as 1-based ids, best last.
// Locality: rows are visited in `order` (tree 0's leaf order; null: index
// order), each XCD (blockIdx.x & 7; the grid is a multiple of 8 blocks)
// taking a contiguous eighth of it, and a block takes kRerankWaves
// consecutive rows at a time from its XCD's eighth through the queue counter
// rowq[64 x] (one atomic per step, claimed two steps ahead), so the rows in
// flight on an XCD stay one window of consecutive cluster-mates.  At 6.35M x
// 1024, width 22, a window of 1024 rows touches 0.34 distinct rows per list
// entry.  With a fixed stride per wave instead, the waves drift apart over a
// pass, the window widens beyond the L2, and 64-dim slices fetched 0.74 of
// their gathered bytes from HBM against 0.35 with the queue.  Slicing keeps a
// window's candidate slices within the 4 MB L2 (a one-pass kernel over whole
// 4 KB rows fetched 0.96).
// row_base: a packed build's segment offset (rows, list ids and x are in the
// packed numbering; part, knns and dists in the segment's own, ids written
// relative to the segment); 0 for a single build.
// One wave per row in four groups of 16 lanes: group g scores candidates
// g, g + 4, ... (U per group in flight), lane t of a group holds float4
// chunks t, t + 16, ... (CH of them) of the slice.  d % 64 == 0, so a chunk
// column is either inside the row or past its end.
template <int CH, bool kFirst, bool kLast>
__global__ __launch_bounds__(64 * kRerankWaves) void k_rerank_slice(
    const float* __restrict__ x, int d, int64_t n, int K, int R, int k, int s, const int32_t* __restrict__ gid,
    float* __restrict__ part, int32_t* __restrict__ knns, float* __restrict__ dists,
    const int32_t* __restrict__ order, int64_t row_base, unsigned int* __restrict__ rowq) {
  static_assert(CH == 1 || CH == 2 || CH == 4 || CH == 8, "slice of 64, 128, 256 or 512 dims");
  constexpr int U = CH < 8 ? 8 / CH : 1;   // candidates in flight per 16-lane group
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, t = lane & 15;
  const int xcd = blockIdx.x & 7;
  const int64_t jlo = n * xcd / 8, jhi = n * (xcd + 1) / 8;
  unsigned int* const qctr = rowq + 64 * xcd;
  const int base = s * 64 * CH;
  bool on[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) on[c] = base + 64 * c < d;
  // Claims run two steps ahead in a ring of three: at step t, sq[t % 3] and
  // sq[(t + 1) % 3] hold steps t and t + 1; thread 0 claims step t + 2 and
  // writes it into sq[(t + 2) % 3] after its row of step t, a slot last read
  // in step t - 1, whose closing barrier every wave has passed.  Knowing step
  // t + 1's row, a wave loads its index and list ids while step t's slices
  // are in flight, so a row costs one dependent round trip, not three.
  __shared__ unsigned int sq[3];
  if (threadIdx.x == 0) {
    sq[0] = atomicAdd(qctr, 1u);
    sq[1] = atomicAdd(qctr, 1u);
  }
  __syncthreads();
  auto row_at = [&](int64_t j) -> int64_t { return j < jhi ? (order ? order[j] : j) : int64_t(0); };
  int64_t j = jlo + int64_t(sq[0]) * kRerankWaves + w;
  int64_t i = row_at(j);
  int32_t my_id = (j < jhi && lane < R) ? gid[(row_base + i) * K + lane] : int32_t(row_base);
  for (int step = 0;; ++step) {
    if (jlo + int64_t(sq[step % 3]) * kRerankWaves >= jhi) break;   // block-uniform
    const unsigned int nxt = threadIdx.x == 0 ? atomicAdd(qctr, 1u) : 0u;
    const int64_t j_next = jlo + int64_t(sq[(step + 1) % 3]) * kRerankWaves + w;
    const int64_t i_next = row_at(j_next);
    float my_s = 0.f;
    if (j < jhi) {
      const float4* po = reinterpret_cast<const float4*>(x + (row_base + i) * d + base);
      float4 own[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) own[c] = on[c] ? po[t + 16 * c] : make_float4(0.f, 0.f, 0.f, 0.f);
      for (int r0 = 0; r0 < R; r0 += 4 * U) {
        float4 v[U][CH];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int r = r0 + 4 * u + g;
          const int32_t id = __shfl(my_id, r < R ? r : 0);
          const float4* pc = reinterpret_cast<const float4*>(x + int64_t(id) * d + base);
#pragma unroll
          for (int c = 0; c < CH; ++c)
            v[u][c] = (on[c] && r < R) ? pc[t + 16 * c] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          float sa = 0.f;
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            sa = fmaf(own[c].x, v[u][c].x, sa);
            sa = fmaf(own[c].y, v[u][c].y, sa);
            sa = fmaf(own[c].z, v[u][c].z, sa);
            sa = fmaf(own[c].w, v[u][c].w, sa);
          }
#pragma unroll
          for (int m = 8; m >= 1; m >>= 1) sa += __shfl_xor(sa, m, 16);
          // lane r0 + 4u + g' takes group g''s sum (candidate r0 + 4u + g')
          const int lo = r0 + 4 * u;
          const float got = __shfl(sa, ((lane - lo) & 3) * 16);
          if (lane >= lo && lane < lo + 4) my_s = got;
        }
      }
    }
    const int32_t id_next = (j_next < jhi && lane < R) ? gid[(row_base + i_next) * K + lane] : int32_t(row_base);
    if (j < jhi) {
      if (!kFirst && lane < R) my_s = part[i * R + lane] + my_s;
      if (!kLast) {
        if (lane < R) part[i * R + lane] = my_s;
      } else {
        int rank = 0;
        for (int r = 0; r < R; ++r) {
          const float tv = __shfl(my_s, r);
          const int32_t tid = __shfl(my_id, r);
          rank += (tv > my_s || (tv == my_s && (tid < my_id || (tid == my_id && r < lane)))) ? 1 : 0;
        }
        if (lane < R && rank < k) {
          knns[i * k + (k - 1 - rank)] = int32_t(my_id - row_base) + 1;
          dists[i * k + (k - 1 - rank)] = my_s;
        }
      }
    }
    j = j_next;
    i = i_next;
    my_id = id_next;
    if (threadIdx.x == 0) sq[(step + 2) % 3] = nxt;
    __syncthreads();
  }
}

// stand-alone gathered pair scores (wave per pair)
__global__ __launch_bounds__(256) void k_score_pairs(const int8_t* __restrict__ q,
                                                     const uint64_t* __restrict__ sk, int d,
                                                     const int32_t* __restrict__ a,
                                                     const int32_t* __restrict__ b, int64_t np,
                                                     int32_t* __restrict__ dot,
                                                     int32_t* __restrict__ sim) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  const int words = d >> 6, nch = d >> 4;
  for (int64_t p = wave; p < np; p += nwaves) {
    const int8_t* ra = q + int64_t(a[p]) * d;
    const int8_t* rb = q + int64_t(b[p]) * d;
    int s = 0;
    for (int c = lane; c < nch; c += 64)
      s = dot16_i8(*reinterpret_cast<const v4i*>(ra + 16 * c),
                   *reinterpret_cast<const v4i*>(rb + 16 * c), s);
    s = wave_sum_i(s);
    int pc = lane < words ? __popcll(sk[int64_t(a[p]) * words + lane] ^
                                     sk[int64_t(b[p]) * words + lane])
                          : 0;
    pc = wave_sum_i(pc);
    if (lane == 0) {
      dot[p] = s;
      sim[p] = d - pc;
    }
  }
}

}  // namespace

namespace knng_host {

int launch_quantize(hipStream_t stream, const float* dx, int64_t nrows, int d, int8_t* q,
                    uint64_t* sk, int num_cu, float* xcopy) {
  if (nrows <= 0) return 0;
  int64_t g = (nrows + 3) / 4;
  const int64_t cap = int64_t(num_cu) * 64;
  if (g > cap) g = cap;
  hipLaunchKernelGGL(k_quantize, dim3(g), dim3(256), 0, stream, dx, nrows, d, q, sk, xcopy);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int assign_splits(int64_t nrows, int C, int num_cu, int64_t part_cap) {
  const int64_t g = (nrows + kAssignRows - 1) / kAssignRows;
  const int64_t ntiles = (C + kAssignTile - 1) / kAssignTile;
  // aim at two blocks per CU; every slice keeps at least two tiles
  int64_t S = (2 * int64_t(num_cu) + g - 1) / g;
  S = std::min<int64_t>(S, ntiles / 2);
  if (nrows > 0) S = std::min<int64_t>(S, part_cap / nrows);
  return int(std::max<int64_t>(1, std::min<int64_t>(S, 1024)));
}

int launch_assign(hipStream_t stream, const int8_t* q, int64_t nrows, int d, const int8_t* Z,
                  const float* w, int C, int32_t* out, int num_cu, int64_t row_off,
                  int64_t row_stride, bool exclude_self, float* score, float* pbest,
                  int32_t* pidx, int64_t part_cap) {
  if (nrows <= 0) return 0;
  int64_t g = (nrows + kAssignRows - 1) / kAssignRows;
  const int64_t cap = int64_t(num_cu) * 8;
  if (g > cap) g = cap;
  const int S = (pbest && pidx) ? assign_splits(nrows, C, num_cu, part_cap) : 1;
  // d = 1024: k_assign_wide (two blocks per CU)
  if (d == kMaxDim && !exclude_self)
    hipLaunchKernelGGL((k_assign_wide<false>), dim3(g, S), dim3(256), 0, stream, q, nrows, row_off,
                       row_stride, Z, w, C, out, score, pbest, pidx);
  else if (d == kMaxDim)
    hipLaunchKernelGGL((k_assign_wide<true>), dim3(g, S), dim3(256), 0, stream, q, nrows, row_off,
                       row_stride, Z, w, C, out, score, pbest, pidx);
  else if (exclude_self)
    hipLaunchKernelGGL((k_assign<true, 0>), dim3(g, S), dim3(256), 0, stream, q, nrows, row_off,
                       row_stride, d, Z, w, C, out, score, pbest, pidx);
  else
    hipLaunchKernelGGL((k_assign<false, 0>), dim3(g, S), dim3(256), 0, stream, q, nrows, row_off,
                       row_stride, d, Z, w, C, out, score, pbest, pidx);
  if (S > 1)
    hipLaunchKernelGGL(k_assign_reduce, dim3(int(std::min<int64_t>((nrows + 255) / 256, 4 * num_cu))),
                       dim3(256), 0, stream, pbest, pidx, S, nrows, out, score);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace knng_host

namespace {

using namespace knng_host;

int quantize_rows(knng_ctx* c, const float* dx, int64_t row0, int64_t nrows) {
  if (nrows <= 0) return 0;
  if (launch_quantize(c->stream, dx, nrows, c->d, c->q + row0 * c->d, c->sk + row0 * c->words,
                      c->num_cu))
    return fail(c, KNNG_EHIP, std::string("k_quantize launch: ") + hipGetErrorString(hipGetLastError()));
  return 0;
}

// ---------------------------------------------------------------------------
// Cells.  A centroid is stored as an int8 direction z (largest component
// +-127) with w = 1/|z|; a row goes to argmax_c float(<q, z_c>) * w_c.
// Centroids come from exact integer sums S (k_cell_sum, k_quant_centroids):
// z_k = trunc(S_k * 127 / max_k |S_k|), w = float(1/sqrt(sum z^2)).
//
// One partition g:
//  1. seeds: cells * cell_oversample rows ((2c+1) n / (2C') + 7919 g) mod n,
//     capped at the sample size;
//  2. kmeans_iters Lloyd rounds on the sample rows off + j * cell_sample
//     (off = g mod cell_sample), then a last sample assignment and its sums;
//  3. merging rounds: every live cell finds its nearest other live cell
//     (argmax_k float(<z_j, z_k>) * w_k); mutual pairs whose cosine
//     (that value * w_j) exceeds cell_merge merge into the lower index,
//     summing their sample sums and counts; repeat until no pair merges
//     (at most 64 rounds);
//  4. pruning: cells with fewer sample rows than cell_prune * the median
//     cell's (upper median) are dropped;
//  5. cell_full_iters Lloyd rounds over all rows, then the final assignment.
// Over-seeding makes every dense region start with a seed; merging joins
// the pieces a region was split into; pruning removes the small cells that
// collect stray rows of many regions.
// ---------------------------------------------------------------------------


// stable counting sort of the listed rows by cell: members[offs[c]..offs[c+1])
// in list order; row j of the list is id off + j * stride
void cell_lists(const std::vector<int32_t>& a, int C, int64_t off, int64_t stride,
                std::vector<int32_t>* members, std::vector<int64_t>* offs) {
  offs->assign(size_t(C) + 1, 0);
  for (int32_t c : a) (*offs)[size_t(c) + 1]++;
  for (int c = 0; c < C; ++c) (*offs)[size_t(c) + 1] += (*offs)[size_t(c)];
  std::vector<int64_t> cur(offs->begin(), offs->end() - 1);
  members->resize(a.size());
  for (size_t j = 0; j < a.size(); ++j)
    (*members)[size_t(cur[size_t(a[j])]++)] = int32_t(off + int64_t(j) * stride);
}

struct KmeansBufs {
  int8_t* z = nullptr;       // centroids / merge rows [cap][d]
  float* w = nullptr;        // [cap]
  int32_t* a = nullptr;      // sample assignment [ns] / merge partner [cap]
  float* sc = nullptr;       // merge best value [cap]
  int32_t* members = nullptr;
  int64_t* offs = nullptr;
  int64_t* sums = nullptr;   // [cap][d]
  int64_t* ids = nullptr;    // seed row ids [cap]
  float* pbest = nullptr;    // split k_assign partials [pcap]
  int32_t* pidx = nullptr;
  int64_t pcap = 0;
  int32_t* ccount = nullptr;            // sample rows per cell (device cell lists) [cap]
  int32_t* ccur = nullptr;              // fill cursors of the device cell lists [cap]
  int32_t* partner = nullptr;           // a merging round's partner of each cell [cap]
  int64_t* dcnt = nullptr;              // sample rows per cell during merging [cap]
  unsigned long long* dpairs = nullptr;  // pairs merged since the last check
  unsigned long long* hpairs = nullptr;  // its pinned host copy
};

// exact sums of the listed rows per cell into B.sums (device); counts on the host
int cell_sums(knng_ctx* c, KmeansBufs& B, const std::vector<int32_t>& a, int C, int64_t off,
              int64_t stride, std::vector<int64_t>* cnt) {
  std::vector<int32_t> members;
  std::vector<int64_t> offs;
  cell_lists(a, C, off, stride, &members, &offs);
  HIPCHK(c, hipMemcpyAsync(B.members, members.data(), members.size() * 4, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(B.offs, offs.data(), offs.size() * 8, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_cell_sum, dim3(std::max(1, std::min(C, c->num_cu * 8))), dim3(256), 0,
                     c->stream, c->q, c->d, B.members, B.offs, C, B.sums);
  HIPCHK(c, hipGetLastError());
  cnt->resize(size_t(C));
  for (int k = 0; k < C; ++k) (*cnt)[size_t(k)] = offs[size_t(k) + 1] - offs[size_t(k)];
  HIPCHK(c, hipStreamSynchronize(c->stream));   // members / offs are host temporaries
  return 0;
}

// cell_sums on the device for the assignment a (device, nrows entries): the
// counts stay in B.ccount, the sums in B.sums; no host round trip
int cell_sums_dev(knng_ctx* c, KmeansBufs& B, const int32_t* a, int64_t nrows, int C, int64_t off,
                  int64_t stride) {
  HIPCHK(c, hipMemsetAsync(B.ccount, 0, size_t(C) * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(B.ccur, 0, size_t(C) * 4, c->stream));
  hipLaunchKernelGGL(k_cell_count, dim3(grid_for(c, nrows, 256, 8)), dim3(256), 0, c->stream, a, nrows, C,
                     B.ccount);
  hipLaunchKernelGGL(k_count_offsets, dim3(1), dim3(1024), 0, c->stream, B.ccount, C, B.offs);
  hipLaunchKernelGGL(k_cell_scatter, dim3(grid_for(c, nrows, 256, 8)), dim3(256), 0, c->stream, a, nrows, C,
                     off, stride, B.offs, B.ccur, B.members);
  hipLaunchKernelGGL(k_cell_sum, dim3(std::max(1, std::min(C, c->num_cu * 8))), dim3(256), 0, c->stream, c->q,
                     c->d, B.members, B.offs, C, B.sums);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// requantise the listed cells' centroids (B.z / B.w rows) from B.sums
int requantize(knng_ctx* c, KmeansBufs& B, const std::vector<int32_t>& cells, int8_t* z, float* w) {
  if (cells.empty()) return 0;
  HIPCHK(c, hipMemcpyAsync(B.a, cells.data(), cells.size() * 4, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_quant_centroids<int64_t>, dim3(std::min<int>(int(cells.size()), c->num_cu * 16)),
                     dim3(256), 0, c->stream, B.sums, c->d, B.a, int(cells.size()), z, w);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));   // `cells` is a host temporary
  return 0;
}

int assign_launch(knng_ctx* c, KmeansBufs& B, const int8_t* rows, int64_t nrows, int64_t off,
                  int64_t stride, const int8_t* z, const float* w, int C, int32_t* out, bool excl,
                  float* score) {
  c->st.assign_ops += double(nrows) * double(C) * 2.0 * double(c->d);
  if (launch_assign(c->stream, rows, nrows, c->d, z, w, C, out, c->num_cu, off, stride, excl, score,
                    B.pbest, B.pidx, B.pcap))
    return fail(c, KNNG_EHIP, std::string("k_assign launch: ") + hipGetErrorString(hipGetLastError()));
  return 0;
}

// k-means of partition g; leaves the final assignment of every row in
// c->cassign + g*n and the centroids in *zf / *wf (C_g of them).  Centroids,
// sums and merges stay on the device; the host holds assignments, counts and
// the merge decisions.
int run_kmeans(knng_ctx* c, int g, int Ct, KmeansBufs& B, int8_t* zfin,
               std::vector<int8_t>* zf, std::vector<float>* wf) {
  const int64_t n = c->n;
  const int d = c->d;
  const int64_t stride = std::max(1, c->p.cell_sample);
  const int64_t off = int64_t(g) % stride < n ? int64_t(g) % stride : 0;
  const int64_t ns = (n - off + stride - 1) / stride;
  const int Cs = int(std::max<int64_t>(1, std::min<int64_t>(int64_t(Ct) * c->p.cell_oversample, ns)));
  int rc;
  // 1. seeds: gathered rows, quantised in place
  std::vector<int64_t> ids(Cs);
  for (int k = 0; k < Cs; ++k)
    ids[size_t(k)] = ((2 * int64_t(k) + 1) * n / (2 * int64_t(Cs)) + int64_t(g) * 7919) % n;
  HIPCHK(c, hipMemcpyAsync(B.ids, ids.data(), ids.size() * 8, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(c, int64_t(Cs) * (d >> 4), 256, 8)), dim3(256), 0,
                     c->stream, c->q, d, B.ids, int64_t(Cs), B.z);
  hipLaunchKernelGGL(k_quant_centroids<int8_t>, dim3(std::min(Cs, c->num_cu * 16)), dim3(256), 0,
                     c->stream, B.z, d, nullptr, Cs, B.z, B.w);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  trace_mark(c, "kmeans seeds");
  // 2. Lloyd on the sample, on the device; the counts come back once
  std::vector<int32_t> upd;
  std::vector<int64_t> cnt(static_cast<size_t>(Cs));
  for (int it = 0;; ++it) {
    if ((rc = assign_launch(c, B, c->q, ns, off, stride, B.z, B.w, Cs, B.a, false, nullptr))) return rc;
    trace_mark(c, "kmeans sample assign");
    if ((rc = cell_sums_dev(c, B, B.a, ns, Cs, off, stride))) return rc;
    if (it == c->p.kmeans_iters) break;
    hipLaunchKernelGGL(k_quant_nonempty, dim3(std::min(Cs, c->num_cu * 16)), dim3(256), 0, c->stream, B.sums,
                       d, B.ccount, Cs, B.z, B.w);
    HIPCHK(c, hipGetLastError());
    trace_mark(c, "kmeans sample update");
  }
  {
    std::vector<int32_t> cnt32(static_cast<size_t>(Cs));
    HIPCHK(c, hipMemcpyAsync(cnt32.data(), B.ccount, size_t(Cs) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < Cs; ++k) cnt[size_t(k)] = cnt32[size_t(k)];
  }
  // 3. merging.  Live cells get centroids from their final sample sums; a dead
  // cell's weight is NaN on the device, so it never wins an argmax (float
  // compares with NaN are false) and each round's result equals the argmax
  // over the live cells alone.
  std::vector<int32_t> live, dead;
  for (int k = 0; k < Cs; ++k) (cnt[size_t(k)] ? live : dead).push_back(k);
  if ((rc = requantize(c, B, live, B.z, B.w))) return rc;
  if (!dead.empty()) {
    HIPCHK(c, hipMemcpyAsync(B.a, dead.data(), dead.size() * 4, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_kill_cells, dim3((int(dead.size()) + 255) / 256), dim3(256), 0, c->stream, B.w,
                       B.a, int(dead.size()));
    HIPCHK(c, hipGetLastError());
  }
  // Rounds run on the device: the cells' nearest live cells (k_assign with
  // self exclusion), the round's mutually nearest pairs (k_merge_decide) and
  // their merges (k_merge_apply), counts and liveness (w not NaN) kept on the
  // device.  A round that merges nothing leaves the state as it was, so every
  // later round merges nothing too: the host checks the pair count every 8
  // rounds and stops at the first check that saw none, with the state of the
  // loop that stops at the first empty round.
  if (live.size() >= 2) {
    HIPCHK(c, hipMemcpyAsync(B.dcnt, cnt.data(), size_t(Cs) * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(B.dpairs, 0, sizeof(unsigned long long), c->stream));
    for (int round = 0; round < 64; ++round) {
      if ((rc = assign_launch(c, B, B.z, Cs, 0, 1, B.z, B.w, Cs, B.a, true, B.sc))) return rc;
      hipLaunchKernelGGL(k_merge_decide, dim3((Cs + 255) / 256), dim3(256), 0, c->stream, B.a, B.sc, B.w, Cs,
                         c->p.cell_merge, B.partner, B.dpairs);
      hipLaunchKernelGGL(k_merge_apply, dim3(std::min(Cs, c->num_cu * 16)), dim3(256), 0, c->stream, B.sums,
                         B.dcnt, B.partner, Cs, d, B.z, B.w);
      HIPCHK(c, hipGetLastError());
      if ((round & 7) == 7) {
        HIPCHK(c, hipMemcpyAsync(B.hpairs, B.dpairs, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (*B.hpairs == 0) break;
        HIPCHK(c, hipMemsetAsync(B.dpairs, 0, sizeof(unsigned long long), c->stream));
      }
    }
    // counts and liveness back on the host (live stays in increasing order)
    std::vector<float> wl(static_cast<size_t>(Cs));
    HIPCHK(c, hipMemcpyAsync(cnt.data(), B.dcnt, size_t(Cs) * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(wl.data(), B.w, size_t(Cs) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    live.clear();
    for (int k = 0; k < Cs; ++k)
      if (wl[size_t(k)] == wl[size_t(k)]) live.push_back(k);
  }
  std::vector<float> w(static_cast<size_t>(Cs));
  trace_mark(c, "kmeans merging");
  // 4. pruning
  {
    std::vector<int64_t> sz;
    for (int32_t k : live) sz.push_back(cnt[size_t(k)]);
    std::sort(sz.begin(), sz.end());
    const float cut = c->p.cell_prune * float(sz[sz.size() / 2]);
    std::vector<int32_t> keep;
    for (int32_t k : live)
      if (!(float(cnt[size_t(k)]) < cut)) keep.push_back(k);
    live.swap(keep);
  }
  // 5. Lloyd over all rows with the live cells, compacted in index order
  const int C = int(live.size());
  float* wfin = B.sc;   // C <= Cs floats
  {
    std::vector<int64_t> lids(live.begin(), live.end());
    HIPCHK(c, hipMemcpyAsync(B.ids, lids.data(), lids.size() * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(c, int64_t(C) * (d >> 4), 256, 8)), dim3(256), 0,
                       c->stream, B.z, d, B.ids, int64_t(C), zfin);
    HIPCHK(c, hipGetLastError());
    std::vector<float> wl(static_cast<size_t>(C));
    // current weights (cells merged in the last round were requantised after its download)
    HIPCHK(c, hipMemcpyAsync(w.data(), B.w, size_t(Cs) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int j = 0; j < C; ++j) wl[size_t(j)] = w[size_t(live[size_t(j)])];
    HIPCHK(c, hipMemcpyAsync(wfin, wl.data(), wl.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  int32_t* da = c->cassign + size_t(g) * n;
  std::vector<int32_t> af;
  for (int it = 0;; ++it) {
    if ((rc = assign_launch(c, B, c->q, n, 0, 1, zfin, wfin, C, da, false, nullptr))) return rc;
    trace_mark(c, "kmeans full assign");
    if (it == c->p.cell_full_iters) break;
    af.resize(size_t(n));
    HIPCHK(c, hipMemcpyAsync(af.data(), da, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = cell_sums(c, B, af, C, 0, 1, &cnt))) return rc;
    upd.clear();
    for (int k = 0; k < C; ++k)
      if (cnt[size_t(k)]) upd.push_back(k);
    if ((rc = requantize(c, B, upd, zfin, wfin))) return rc;
    trace_mark(c, "kmeans full update");
  }
  zf->resize(size_t(C) * d);
  wf->resize(size_t(C));
  HIPCHK(c, hipMemcpyAsync(zf->data(), zfin, zf->size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(wf->data(), wfin, wf->size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

void free_kmeans(KmeansBufs& B, int8_t** merge_rows) {
  dfree(&B.z);
  dfree(&B.w);
  dfree(&B.a);
  dfree(&B.sc);
  dfree(&B.members);
  dfree(&B.offs);
  dfree(&B.sums);
  dfree(&B.ids);
  dfree(&B.pbest);
  dfree(&B.pidx);
  dfree(&B.ccount);
  dfree(&B.ccur);
  dfree(&B.partner);
  dfree(&B.dcnt);
  dfree(&B.dpairs);
  if (B.hpairs) (void)hipHostFree(B.hpairs);
  B.hpairs = nullptr;
  dfree(merge_rows);
}

// Seed every tree with its partition's cells: perm = rows by (cell, id);
// cells larger than a leaf become the first level's nodes (depth 1), the
// others leaves.  Returns the node count in *nnodes.
int seed_from_cells(knng_ctx* c, int Ct, int* nnodes) {
  const int64_t n = c->n, T = c->p.num_trees;
  const int d = c->d, leaf = c->p.leaf_size;
  const int G = c->nparts;
  const int64_t stride = std::max(1, c->p.cell_sample);
  const int64_t cap = std::max<int64_t>(
      2, std::min<int64_t>(int64_t(Ct) * c->p.cell_oversample, (n + stride - 1) / stride));
  int rc;
  dfree(&c->cassign);
  if ((rc = dalloc(c, &c->cassign, size_t(G) * n, "cell assignment"))) return rc;
  // one buffer set per partition: the partitions' k-means run concurrently,
  // each on a host thread with a copy of the context that owns its own
  // stream (see below)
  auto alloc_bufs = [&](KmeansBufs& B, int8_t** merge_rows) -> int {
    B.pcap = assign_part_entries(cap, c->num_cu);
    int e;
    if ((e = dalloc(c, &B.z, size_t(cap) * d, "centroids")) ||
        (e = dalloc(c, &B.w, size_t(cap), "centroid norms")) ||
        (e = dalloc(c, &B.a, size_t(std::max<int64_t>(cap, (n + stride - 1) / stride)), "assignment")) ||
        (e = dalloc(c, &B.sc, size_t(cap), "merge values")) ||
        (e = dalloc(c, &B.members, size_t(n), "cell members")) ||
        (e = dalloc(c, &B.offs, size_t(cap) + 1, "cell offsets")) ||
        (e = dalloc(c, &B.sums, size_t(cap) * d, "cell sums")) ||
        (e = dalloc(c, &B.ids, size_t(cap), "seed ids")) ||
        (e = dalloc(c, &B.pbest, size_t(assign_part_entries(cap, c->num_cu)), "assign partials")) ||
        (e = dalloc(c, &B.pidx, size_t(assign_part_entries(cap, c->num_cu)), "assign partials")) ||
        (e = dalloc(c, merge_rows, size_t(cap) * d, "merge rows")) ||
        (e = dalloc(c, &B.ccount, size_t(cap), "cell counts")) ||
        (e = dalloc(c, &B.ccur, size_t(cap), "cell cursors")) ||
        (e = dalloc(c, &B.partner, size_t(cap), "merge partners")) ||
        (e = dalloc(c, &B.dcnt, size_t(cap), "merge counts")) ||
        (e = dalloc(c, &B.dpairs, 1, "merge pair count")))
      return e;
    if (hipHostMalloc(reinterpret_cast<void**>(&B.hpairs), sizeof(unsigned long long), hipHostMallocDefault) !=
        hipSuccess)
      return fail(c, KNNG_ENOMEM, "hipHostMalloc failed (k-means merge pair count)");
    return 0;
  };
  std::vector<std::vector<int8_t>> zs(G);
  std::vector<std::vector<float>> ws(G);
  c->cell_count.assign(size_t(G), 0);
  // The partitions are independent k-means runs (own seeds, buffers and
  // slice of c->cassign; the int8 rows are only read), so each runs on a host
  // thread of its own with a copy of the context whose stream is its own:
  // the many small launches of one partition's merging rounds and its host
  // checks overlap the other partitions' work.  Every partition computes
  // exactly what it computes alone.  KNNG_KMEANS_SERIAL=1 runs them in turn
  // on the context's stream.
  const char* ser = std::getenv("KNNG_KMEANS_SERIAL");
  const bool serial = G < 2 || (ser && ser[0] == '1');
  const int L = serial ? 1 : G;
  std::vector<KmeansBufs> Bs(static_cast<size_t>(L));
  std::vector<int8_t*> merge_rows(size_t(L), nullptr);
  auto free_all = [&]() {
    for (int l = 0; l < L; ++l) free_kmeans(Bs[size_t(l)], &merge_rows[size_t(l)]);
  };
  for (int l = 0; l < L && !rc; ++l) rc = alloc_bufs(Bs[size_t(l)], &merge_rows[size_t(l)]);
  if (rc) {
    free_all();
    return rc;
  }
  if (serial) {
    for (int g = 0; g < G && !rc; ++g) {
      rc = run_kmeans(c, g, Ct, Bs[0], merge_rows[0], &zs[size_t(g)], &ws[size_t(g)]);
      if (!rc) c->cell_count[size_t(g)] = int32_t(ws[size_t(g)].size());
    }
  } else {
    if (hipStreamSynchronize(c->stream) != hipSuccess) {   // the int8 rows are in place
      free_all();
      return fail(c, KNNG_EHIP, "hipStreamSynchronize failed before the cell k-means");
    }
    const double ops0 = c->st.assign_ops;   // each lane adds its own to this count
    std::vector<knng_ctx> lane(size_t(G), *c);
    std::vector<int> rcs(size_t(G), 0);
    for (int g = 0; g < G && !rc; ++g)
      if (hipStreamCreateWithFlags(&lane[size_t(g)].stream, hipStreamNonBlocking) != hipSuccess) {
        lane[size_t(g)].stream = nullptr;
        rc = fail(c, KNNG_EHIP, "hipStreamCreate failed (cell k-means)");
      }
    if (!rc) {
      std::vector<std::thread> th;
      for (int g = 0; g < G; ++g)
        th.emplace_back([&, g] {
          knng_ctx* lc = &lane[size_t(g)];
          rcs[size_t(g)] = hipSetDevice(lc->device) != hipSuccess
                               ? fail(lc, KNNG_EHIP, "hipSetDevice failed (cell k-means)")
                               : run_kmeans(lc, g, Ct, Bs[size_t(g)], merge_rows[size_t(g)], &zs[size_t(g)],
                                            &ws[size_t(g)]);
        });
      for (std::thread& t : th) t.join();
      for (int g = 0; g < G; ++g) {
        c->st.assign_ops += lane[size_t(g)].st.assign_ops - ops0;
        if (rcs[size_t(g)] && !rc) {
          rc = rcs[size_t(g)];
          c->err = lane[size_t(g)].err;
        }
        if (!rcs[size_t(g)]) c->cell_count[size_t(g)] = int32_t(ws[size_t(g)].size());
      }
    }
    for (int g = 0; g < G; ++g)
      if (lane[size_t(g)].stream) {
        (void)hipStreamSynchronize(lane[size_t(g)].stream);
        (void)hipStreamDestroy(lane[size_t(g)].stream);
      }
  }
  free_all();
  if (rc) return rc;
  const int Cm = *std::max_element(c->cell_count.begin(), c->cell_count.end());
  c->ncells = Cm;
  dfree(&c->cz);
  dfree(&c->cw);
  dfree(&c->cell_root);
  if ((rc = dalloc(c, &c->cz, size_t(G) * Cm * d, "centroids")) ||
      (rc = dalloc(c, &c->cw, size_t(G) * Cm, "centroid norms")) ||
      (rc = dalloc(c, &c->cell_root, size_t(T) * Cm, "cell roots")))
    return rc;
  for (int g = 0; g < G; ++g) {
    HIPCHK(c, hipMemcpyAsync(c->cz + size_t(g) * Cm * d, zs[size_t(g)].data(), zs[size_t(g)].size(),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cw + size_t(g) * Cm, ws[size_t(g)].data(), ws[size_t(g)].size() * 4,
                             hipMemcpyHostToDevice, c->stream));
  }
  trace_mark(c, "cells uploaded");
  // perm and nodepos on the device: per partition, a stable radix sort of the
  // rows by cell (inside a cell the rows keep increasing id) and the cell
  // sizes from a histogram; the host builds only the per-cell first-level
  // nodes, small-cell leaves and roots from the sizes
  int32_t *skeys = nullptr, *ivals = nullptr, *svals = nullptr, *dcnt = nullptr, *cellnode = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int bits = 1;
  while ((int64_t(1) << bits) < Cm) ++bits;
  auto free_sort = [&] {
    dfree(&skeys);
    dfree(&ivals);
    dfree(&svals);
    dfree(&dcnt);
    dfree(&cellnode);
    if (tmp) (void)hipFree(tmp);
    tmp = nullptr;
  };
  HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, static_cast<const int32_t*>(nullptr),
                                               static_cast<int32_t*>(nullptr),
                                               static_cast<const int32_t*>(nullptr),
                                               static_cast<int32_t*>(nullptr), int(n), 0, bits, c->stream));
  if ((rc = dalloc(c, &skeys, size_t(G) * n, "sorted cells")) || (rc = dalloc(c, &ivals, size_t(n), "row ids")) ||
      (rc = dalloc(c, &svals, size_t(G) * n, "rows by cell")) ||
      (rc = dalloc(c, &dcnt, size_t(G) * Cm, "cell sizes")) ||
      (rc = dalloc(c, &cellnode, size_t(T) * Cm, "cell nodes"))) {
    free_sort();
    return rc;
  }
  if (hipMalloc(&tmp, tmp_bytes + 256) != hipSuccess) {
    free_sort();
    return fail(c, KNNG_ENOMEM, "hipMalloc failed for the cell sort");
  }
  hipLaunchKernelGGL(k_iota, dim3(grid_for(c, n, 256, 8)), dim3(256), 0, c->stream, ivals, n);
  HIPCHK(c, hipMemsetAsync(dcnt, 0, size_t(G) * Cm * sizeof(int32_t), c->stream));
  for (int g = 0; g < G; ++g) {
    size_t tb = tmp_bytes;
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, c->cassign + size_t(g) * n, skeys + size_t(g) * n,
                                                 ivals, svals + size_t(g) * n, int(n), 0, bits, c->stream));
    hipLaunchKernelGGL(k_cell_count, dim3(grid_for(c, n, 256, 8)), dim3(256), 0, c->stream,
                       c->cassign + size_t(g) * n, n, Cm, dcnt + size_t(g) * Cm);
  }
  HIPCHK(c, hipGetLastError());
  std::vector<int32_t> hcnt(size_t(G) * Cm);
  HIPCHK(c, hipMemcpyAsync(hcnt.data(), dcnt, hcnt.size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int32_t> roots(size_t(T) * Cm, INT32_MIN), cnode(size_t(T) * Cm, -1);
  std::vector<Node> nodes;
  std::vector<int4> leaves;
  for (int t = 0; t < T; ++t) {
    const int g = t % G;
    int64_t s = 0;
    for (int cc = 0; cc < c->cell_count[size_t(g)]; ++cc) {
      const int64_t z = hcnt[size_t(g) * Cm + cc];
      if (z == 0) {
        roots[size_t(t) * Cm + cc] = INT32_MIN;
      } else if (z <= leaf) {
        roots[size_t(t) * Cm + cc] = -int32_t(leaves.size() + 1);
        leaves.push_back(make_int4(t, int(s), int(z), 0));
      } else {
        Node nd;
        nd.tree = t;
        nd.start = int32_t(s);
        nd.size = int32_t(z);
        node_pivots(c->p.seed, t, 1, s, z, &nd.p1, &nd.p2);
        nd.left = nd.right = -1;
        nd.nleft = 0;
        nd.p1id = nd.p2id = -1;
        nd.base = 0;
        roots[size_t(t) * Cm + cc] = int32_t(nodes.size());
        cnode[size_t(t) * Cm + cc] = int32_t(nodes.size());
        nodes.push_back(nd);
      }
      s += z;
    }
  }
  if (int64_t(nodes.size()) > c->node_cap || int64_t(leaves.size()) > c->leaf_cap) {
    free_sort();
    return fail(c, KNNG_ERANGE, "cell seeding exceeds the node/leaf tables");
  }
  HIPCHK(c, hipMemcpyAsync(cellnode, cnode.data(), cnode.size() * 4, hipMemcpyHostToDevice, c->stream));
  for (int t = 0; t < T; ++t) {
    const int g = t % G;
    HIPCHK(c, hipMemcpyAsync(c->perm + size_t(t) * n, svals + size_t(g) * n, size_t(n) * 4,
                             hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_cell_nodepos, dim3(grid_for(c, n, 256, 8)), dim3(256), 0, c->stream,
                       skeys + size_t(g) * n, n, cellnode + size_t(t) * Cm, c->nodepos + size_t(t) * n);
  }
  HIPCHK(c, hipGetLastError());
  if (!nodes.empty())
    HIPCHK(c, hipMemcpyAsync(c->nodes, nodes.data(), nodes.size() * sizeof(Node), hipMemcpyHostToDevice,
                             c->stream));
  if (!leaves.empty())
    HIPCHK(c, hipMemcpyAsync(c->leaves, leaves.data(), leaves.size() * sizeof(int4),
                             hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->cell_root, roots.data(), roots.size() * 4, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));   // host vectors above are sources of the copies
  free_sort();
  int32_t cnt2[2] = {0, int32_t(leaves.size())};
  HIPCHK(c, hipMemcpyAsync(c->dcount, cnt2, sizeof(cnt2), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  trace_mark(c, "cell forest level");
  *nnodes = int(nodes.size());
  return 0;
}

int check_params(const knng_params* p, int64_t n, std::string* why) {
  if (n < 2) return *why = "n must be >= 2", KNNG_EINVAL;
  if (p->dim <= 0 || p->dim % 64 || p->dim > kMaxDim) return *why = "dim must be a multiple of 64 in [64, 1024]", KNNG_ERANGE;
  if (p->k_build < 1 || p->k_build > kMaxK) return *why = "k_build must be in [1, 64]", KNNG_ERANGE;
  if (p->k_eval < 1 || p->k_eval > p->k_build) return *why = "k_eval must be in [1, k_build]", KNNG_EINVAL;
  if (p->num_trees < 1 || p->num_trees > kMaxTrees) return *why = "num_trees must be in [1, 64]", KNNG_ERANGE;
  if (p->leaf_size < 2 || p->leaf_size > kMaxLeaf) return *why = "leaf_size must be in [2, 2048]", KNNG_ERANGE;
  if (p->rerank_mult < 1) return *why = "rerank_mult must be >= 1", KNNG_EINVAL;
  if (p->max_iters < 0 || p->max_iters > 32) return *why = "max_iters must be in [0, 32]", KNNG_ERANGE;
  if (p->rerank_width < 0 || p->rerank_width > p->k_build) return *why = "rerank_width must be in [0, k_build]", KNNG_EINVAL;
  if (p->rerank_width && p->rerank_width < p->k_eval) return *why = "rerank_width must be >= k_eval", KNNG_EINVAL;
  if (p->stop_rule < 0 || p->stop_rule > 2) return *why = "stop_rule must be 0, 1 or 2", KNNG_EINVAL;
  if (p->cells < -1) return *why = "cells must be -1 (auto), 0 (none) or a count", KNNG_EINVAL;
  if (p->kmeans_iters < 0 || p->kmeans_iters > 100) return *why = "kmeans_iters must be in [0, 100]", KNNG_EINVAL;
  if (p->cell_partitions < 1) return *why = "cell_partitions must be >= 1", KNNG_EINVAL;
  if (p->cell_oversample < 1 || p->cell_oversample > 64) return *why = "cell_oversample must be in [1, 64]", KNNG_EINVAL;
  if (p->cell_sample < 1) return *why = "cell_sample must be >= 1", KNNG_EINVAL;
  if (p->cell_full_iters < 0 || p->cell_full_iters > 100) return *why = "cell_full_iters must be in [0, 100]", KNNG_EINVAL;
  if (!(p->cell_merge == p->cell_merge)) return *why = "cell_merge must be a number", KNNG_EINVAL;
  if (!(p->cell_prune >= 0.0f && p->cell_prune <= 1.0f)) return *why = "cell_prune must be in [0, 1]", KNNG_EINVAL;
  if (p->reverse_width < 0 || p->reverse_width > kRevMax) return *why = "reverse_width must be in [0, 8]", KNNG_ERANGE;
  if (p->reverse_width > 0 && p->k_build > 50)
    return *why = "reverse_width needs k_build <= 50 (the block-per-row join's candidate array)", KNNG_EINVAL;
  if (int64_t(p->num_trees) * n >= (int64_t(1) << 31) - 1) return *why = "num_trees * n must fit in int32", KNNG_ERANGE;
  if (n >= (int64_t(1) << 31) - 1) return *why = "n must fit in int32", KNNG_ERANGE;
  return 0;
}


// ---------------------------------------------------------------------------
// k_join shapes and the autotuner.  Every shape computes the same lists (the
// iteration is a deterministic function of the previous graph); they differ
// in blocks per CU and in how many gathers each lane keeps in flight:
//   0: 4 blocks/CU, 2 rows + 8 sketches in flight per lane group (106 VGPRs at d = 1024)
//   1: 3 blocks/CU, 4 rows + 8 sketches (138 VGPRs)
//   2: 4 blocks/CU, 3 rows + 8 sketches (122 VGPRs)
// k_build > 50 has one shape (its 64 KB of LDS allows 2 blocks per CU).
// The first join iteration of a (dim, k_build, rule) shape in a process times
// each shape on a prefix of the work list (same rows, same outputs, so the
// result is unchanged) and keeps the fastest; KNNG_JOIN_VARIANT=v forces v.
// ---------------------------------------------------------------------------

constexpr int kJoinVariants = 3;

// Reverse lists of graph buffer a for the coming iteration (reverse_width R >
// 0): every entry offered when the iteration has no new/old rule (or no join
// iteration ran yet), else the entries that came in during the previous one.
int build_reverse_lists(knng_ctx* c, int a) {
  const int64_t n = c->n;
  const int K = c->p.k_build, R = c->p.reverse_width;
  int rc;
  if (!c->rv_slots) {
    if ((rc = dalloc(c, &c->rv_slots, size_t(n) * R, "reverse-list slots"))) return rc;
    HIPCHK(c, hipMemsetAsync(c->rv_slots, 0, size_t(n) * R * sizeof(unsigned long long), c->stream));
  }
  if ((rc = dalloc(c, &c->rv_id, size_t(n) * R, "reverse lists")) ||
      (rc = dalloc(c, &c->rv_mask, size_t(n) * R, "reverse-list masks")))
    return rc;
  const uint64_t* fin = (c->p.incremental && !c->fresh_all) ? c->fresh[a] : nullptr;
  hipLaunchKernelGGL(k_rev_insert, dim3(grid_for(c, n * K, 256, 16)), dim3(256), 0, c->stream, c->gid[a],
                     c->gsc[a], fin, n, K, R, c->rv_slots);
  hipLaunchKernelGGL(k_rev_final, dim3(grid_for(c, n, 256, 16)), dim3(256), 0, c->stream, c->rv_slots, n, R,
                     c->leaf_valid ? c->leafkey : static_cast<const int32_t*>(nullptr), int(c->p.num_trees),
                     c->leaf_tp, c->rv_id, c->rv_mask);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// reverse-list arguments of the join kernels: (ids, masks, R), or none
#define KNNG_REV_ARGS(c)                                                                   \
  ((c)->p.reverse_width > 0 ? (c)->rv_id : static_cast<const int32_t*>(nullptr)),           \
      ((c)->p.reverse_width > 0 && (c)->leaf_valid ? (c)->rv_mask : static_cast<const uint8_t*>(nullptr)), \
      int((c)->p.reverse_width)

int launch_join(knng_ctx* c, int v, int a, int b, const int32_t* order, const int64_t* nwork_dev,
                const int64_t* seg_base, int nseg, unsigned long long* segc) {
  const int64_t n = c->n;
  const int K = c->p.k_build, M = K * c->p.rerank_mult;
  // grid: a multiple of the 8 XCDs (the schedule needs it), four blocks per
  // CU; blocks beyond what is resident just find their row queue drained
  const int grid = std::max(8, (c->num_cu * 4) / 8 * 8);
  unsigned long long* rowq = c->counters + kCounterFields * kNumCounters;
  const uint64_t* fin = c->p.incremental ? c->fresh[a] : nullptr;
  const int32_t* leafkey = c->leaf_valid ? c->leafkey : nullptr;
  const uint8_t* mk_in = c->leaf_valid ? c->gmask[a] : nullptr;
  uint8_t* mk_out = c->leaf_valid ? c->gmask[b] : nullptr;
#define KNNG_JOIN_ARGS                                                                          \
  c->q, c->sk, c->d, n, K, M, c->gid[a], c->gsc[a], c->gid[b], c->gsc[b], c->counters, rowq, order, \
      fin, c->fresh[b], nwork_dev, seg_base, nseg, segc, static_cast<int32_t*>(nullptr),             \
      static_cast<unsigned long long*>(nullptr), mk_in, mk_out, leafkey, int(c->p.num_trees), c->leaf_tp, \
      KNNG_REV_ARGS(c)
  const bool d1024 = c->d == kMaxDim;   // the rows' length as a compile-time constant
  // reverse candidates: up to K * K + R + 2 K R of them, the 4096-slot shape
  if (c->p.reverse_width > 0 && d1024)
    hipLaunchKernelGGL((k_join<13, 4096, 2, 2, 8, kMaxDim, false, true>), dim3(grid), dim3(256), 0, c->stream,
                       KNNG_JOIN_ARGS);
  else if (c->p.reverse_width > 0)
    hipLaunchKernelGGL((k_join<13, 4096, 2, 2, 8, 0, false, true>), dim3(grid), dim3(256), 0, c->stream,
                       KNNG_JOIN_ARGS);
  else if (K > 50 && d1024)
    hipLaunchKernelGGL((k_join<13, 4096, 2, 2, 8, kMaxDim>), dim3(grid), dim3(256), 0, c->stream, KNNG_JOIN_ARGS);
  else if (K > 50)
    hipLaunchKernelGGL((k_join<13, 4096, 2, 2, 8, 0>), dim3(grid), dim3(256), 0, c->stream, KNNG_JOIN_ARGS);
  else if (v == 1 && d1024)
    hipLaunchKernelGGL((k_join<12, 2560, 3, 4, 8, kMaxDim>), dim3(grid), dim3(256), 0, c->stream, KNNG_JOIN_ARGS);
  else if (v == 1)
    hipLaunchKernelGGL((k_join<12, 2560, 3, 4, 8, 0>), dim3(grid), dim3(256), 0, c->stream, KNNG_JOIN_ARGS);
  else if (v == 2 && d1024)
    hipLaunchKernelGGL((k_join<12, 2560, 4, 3, 8, kMaxDim>), dim3(grid), dim3(256), 0, c->stream, KNNG_JOIN_ARGS);
  else if (v == 2)
    hipLaunchKernelGGL((k_join<12, 2560, 4, 3, 8, 0>), dim3(grid), dim3(256), 0, c->stream, KNNG_JOIN_ARGS);
  else if (d1024)
    hipLaunchKernelGGL((k_join<12, 2560, 4, 2, 8, kMaxDim>), dim3(grid), dim3(256), 0, c->stream, KNNG_JOIN_ARGS);
  else
    hipLaunchKernelGGL((k_join<12, 2560, 4, 2, 8, 0>), dim3(grid), dim3(256), 0, c->stream, KNNG_JOIN_ARGS);
#undef KNNG_JOIN_ARGS
  HIPCHK(c, hipGetLastError());
  return 0;
}

// A join iteration under the new/old rule after the first: the small-row
// variant of k_join over every row, then the full kernel over the rows it
// passed on (device count, no host round trip).  Eight blocks per CU (the
// most a CU holds: 32 waves; 64 VGPRs with some spilling, 15 KB of LDS each)
// measured fastest: the iteration is bound by rows in flight (120 ms at
// 6.35M against 135 / 141 ms at six / five blocks and 193 ms for the full
// kernel at four).
constexpr int kSmallJoinOcc = 8;
int launch_join_small(knng_ctx* c, int a, int b, const int32_t* order) {
  const int64_t n = c->n;
  const int K = c->p.k_build, M = K * c->p.rerank_mult;
  if (!c->ovf) {
    int rc;
    if ((rc = dalloc(c, &c->ovf, size_t(n), "join overflow rows")) ||
        (rc = dalloc(c, &c->ovf_n, 1, "join overflow count")))
      return rc;
  }
  HIPCHK(c, hipMemsetAsync(c->ovf_n, 0, sizeof(unsigned long long), c->stream));
  const int grid = std::max(8, (c->num_cu * kSmallJoinOcc) / 8 * 8);
  const int32_t* leafkey = c->leaf_valid ? c->leafkey : nullptr;
  const uint8_t* mk_in = c->leaf_valid ? c->gmask[a] : nullptr;
  uint8_t* mk_out = c->leaf_valid ? c->gmask[b] : nullptr;
  unsigned long long* rowq = c->counters + kCounterFields * kNumCounters;
#define KNNG_SMALL_ARGS                                                                                 \
  c->q, c->sk, c->d, n, K, M, c->gid[a], c->gsc[a], c->gid[b], c->gsc[b], c->counters, rowq, order,     \
      c->fresh[a], c->fresh[b], static_cast<const int64_t*>(nullptr), static_cast<const int64_t*>(nullptr), \
      0, static_cast<unsigned long long*>(nullptr), c->ovf, c->ovf_n, mk_in, mk_out, leafkey, int(c->p.num_trees), \
      c->leaf_tp, KNNG_REV_ARGS(c)
  if (c->p.reverse_width > 0 && c->d == kMaxDim)
    hipLaunchKernelGGL((k_join<10, 640, kSmallJoinOcc, 1, 8, kMaxDim, true, true>), dim3(grid), dim3(256), 0,
                       c->stream, KNNG_SMALL_ARGS);
  else if (c->p.reverse_width > 0)
    hipLaunchKernelGGL((k_join<10, 640, kSmallJoinOcc, 1, 8, 0, true, true>), dim3(grid), dim3(256), 0, c->stream,
                       KNNG_SMALL_ARGS);
  else if (c->d == kMaxDim)
    hipLaunchKernelGGL((k_join<10, 640, kSmallJoinOcc, 1, 8, kMaxDim, true>), dim3(grid), dim3(256), 0, c->stream,
                       KNNG_SMALL_ARGS);
  else
    hipLaunchKernelGGL((k_join<10, 640, kSmallJoinOcc, 1, 8, 0, true>), dim3(grid), dim3(256), 0, c->stream,
                       KNNG_SMALL_ARGS);
#undef KNNG_SMALL_ARGS
  HIPCHK(c, hipGetLastError());
  return launch_join(c, c->join_variant, a, b, c->ovf, reinterpret_cast<const int64_t*>(c->ovf_n), nullptr, 0,
                     nullptr);
}

// The phased join (k_jcand, k_jdot_merge) over every row, then k_join over
// the rows with more than kJCap candidates (device count, no host round trip).
// blocks of 256 threads one CU holds of kernel f (its registers and LDS), at least 1
template <typename F>
int resident_blocks(F f) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(f), 256, 0) != hipSuccess ||
      nb < 1)
    nb = 1;
  return nb;
}

// a grid of every block the GPU holds at once, a multiple of the 8 XCDs (the
// phased kernels give each wave a fixed share of its XCD's rows: a block
// that waits for a slot would finish its share last)
template <typename F>
int resident_grid(const knng_ctx* c, F f) {
  static std::mutex mu;
  static std::map<const void*, int> per_cu;   // by kernel (instantiations share F)
  int b;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = per_cu.find(reinterpret_cast<const void*>(f));
    if (it == per_cu.end()) it = per_cu.emplace(reinterpret_cast<const void*>(f), resident_blocks(f)).first;
    b = it->second;
  }
  return std::max(8, (c->num_cu * b) / 8 * 8);
}

int launch_join_phased(knng_ctx* c, int a, int b, const int32_t* order) {
  const int64_t n = c->n;
  const int K = c->p.k_build, M = K * c->p.rerank_mult;
  int rc;
  if ((rc = dalloc(c, &c->jcbuf, size_t(n) * kJCap, "join candidates")) ||
      (rc = dalloc(c, &c->jccount, size_t(n), "join candidate counts")) ||
      (rc = dalloc(c, &c->ovf, size_t(n), "join overflow rows")) ||
      (rc = dalloc(c, &c->ovf_n, 1, "join overflow count")))
    return rc;
  HIPCHK(c, hipMemsetAsync(c->ovf_n, 0, sizeof(unsigned long long), c->stream));
  const int32_t* leafkey = c->leaf_valid ? c->leafkey : nullptr;
  const uint8_t* mk_in = c->leaf_valid ? c->gmask[a] : nullptr;
  uint8_t* mk_out = c->leaf_valid ? c->gmask[b] : nullptr;
  const int T = int(c->p.num_trees);
  // every entry still new (no join iteration yet): the new/old rule takes every pair
  const uint64_t* fin = (c->p.incremental && !c->fresh_all) ? c->fresh[a] : nullptr;
  // row-queue counters (zero: k_sum_counters clears them after each iteration)
  unsigned long long* rowq = c->counters + kCounterFields * kNumCounters;
  if (c->p.reverse_width > 0)
    hipLaunchKernelGGL(k_jcand<true>, dim3(resident_grid(c, k_jcand<true>)), dim3(256), 0, c->stream, c->sk, c->d,
                       K, M, c->gid[a], fin, mk_in, order, n, c->jcbuf, c->jccount, c->ovf, c->ovf_n, c->counters,
                       rowq, KNNG_REV_ARGS(c));
  else
    hipLaunchKernelGGL(k_jcand<false>, dim3(resident_grid(c, k_jcand<false>)), dim3(256), 0, c->stream, c->sk, c->d,
                       K, M, c->gid[a], fin, mk_in, order, n, c->jcbuf, c->jccount, c->ovf, c->ovf_n, c->counters,
                       rowq, KNNG_REV_ARGS(c));
  if (c->d == kMaxDim)
    hipLaunchKernelGGL(k_jdot_merge<kMaxDim>, dim3(resident_grid(c, k_jdot_merge<kMaxDim>)), dim3(256), 0, c->stream,
                       c->q, c->d, K, c->gid[a], c->gsc[a], c->gid[b], c->gsc[b], c->fresh[b], mk_in, mk_out, leafkey,
                       T, c->leaf_tp, order, n, c->jcbuf, c->jccount, c->counters, rowq);
  else
    hipLaunchKernelGGL(k_jdot_merge<0>, dim3(resident_grid(c, k_jdot_merge<0>)), dim3(256), 0, c->stream, c->q, c->d,
                       K, c->gid[a], c->gsc[a], c->gid[b], c->gsc[b], c->fresh[b], mk_in, mk_out, leafkey, T,
                       c->leaf_tp, order, n, c->jcbuf, c->jccount, c->counters, rowq);
  HIPCHK(c, hipGetLastError());
  return launch_join(c, 0, a, b, c->ovf, reinterpret_cast<const int64_t*>(c->ovf_n), nullptr, 0, nullptr);
}

bool join_fused_forced() {
  const char* e = std::getenv("KNNG_JOIN_FUSED");
  return e && e[0] == '1';
}

struct TuneCache {
  std::mutex mu;
  std::map<std::tuple<int, int, int>, int> best;
};

TuneCache& tune_cache() {
  static TuneCache tc;
  return tc;
}

// Choose c->join_variant (see above) before the first launch of iteration
// a -> b over order[0 .. nwork).  The timing launches also add to segc
// (packed builds), which the caller clears afterwards.
int ensure_join_variant(knng_ctx* c, int a, int b, const int32_t* order, int64_t nwork,
                        const int64_t* seg_base, int nseg, unsigned long long* segc) {
  if (c->join_variant >= 0) return 0;
  c->st.ms_tune = 0;
  if (c->p.k_build > 50) return c->join_variant = 0, 0;
  if (const char* e = std::getenv("KNNG_JOIN_VARIANT")) {
    const int v = std::atoi(e);
    if (v >= 0 && v < kJoinVariants) return c->join_variant = v, 0;
  }
  const auto key = std::make_tuple(c->d, c->p.k_build, c->p.incremental);
  {
    std::lock_guard<std::mutex> lk(tune_cache().mu);
    auto it = tune_cache().best.find(key);
    if (it != tune_cache().best.end()) return c->join_variant = it->second, 0;
  }
  // too few rows to time reliably: the default shape, not cached
  const int64_t rows = std::max<int64_t>(65536, nwork / 32);
  if (nwork < 2 * rows) return c->join_variant = 0, 0;
  int rc;
  if ((rc = dalloc(c, &c->tune_n, 1, "tuning size"))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->tune_n, &rows, sizeof(rows), hipMemcpyHostToDevice, c->stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(c, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(c, KNNG_EHIP, "hipEventCreate failed (join autotuner)");
  }
  float best_ms = 0, total = 0;
  int best = 0;
  // a failed launch ends the tuning at once: no later timing may pick a shape
  for (int round = 0; round < 2 && !rc; ++round)
    for (int v = 0; v < kJoinVariants && !rc; ++v) {
      (void)hipEventRecord(e0, c->stream);
      if ((rc = launch_join(c, v, a, b, order, c->tune_n, seg_base, nseg, segc))) break;
      (void)hipEventRecord(e1, c->stream);
      hipLaunchKernelGGL(k_sum_counters, dim3(1), dim3(256), 0, c->stream, c->counters, c->csum);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      total += ms;
      if (round == 1 && (v == 0 || ms < best_ms)) {
        best_ms = ms;
        best = v;
      }
    }
  (void)hipStreamSynchronize(c->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  c->st.ms_tune = total;
  c->join_variant = best;
  std::lock_guard<std::mutex> lk(tune_cache().mu);
  tune_cache().best[key] = best;
  return 0;
}

// slices per export at most (64-dim slices of at most kMaxDim dims)
constexpr int kRerankMaxPasses = kMaxDim / 64;

// float4 chunks per lane of one rerank slice (slice = 64 * chunks dims).
// 256-dim slices by default: at 6.35M x 1024 with width 22 one export takes
// 49 ms against 66 (64 dims), 52 (128) and 50 (512) with the row queue.
// KNNG_RERANK_SLICE = 64, 128, 256 or 512 overrides it.
int rerank_chunks(int d) {
  int dims = 256;
  if (const char* e = std::getenv("KNNG_RERANK_SLICE")) {
    const int v = std::atoi(e);
    if (v == 64 || v == 128 || v == 256 || v == 512) dims = v;
  }
  (void)d;
  return dims / 64;
}

// ---------------------------------------------------------------------------
// Leaf-tile shapes for leaves of at most 64 rows and their autotuner.  Both
// shapes write the same keys:
//   0: k_leaf_gram, a block per leaf, each wave one 32 x 32 quadrant, the
//      leaf's rows staged through LDS in 512-dim chunks (four tiles a leaf);
//   1: k_leaf_gram_wave, a wave per leaf, fragments straight from HBM, the
//      symmetric Gram in three tiles (one for leaves of at most 32 rows).
// The first init of a (dim, leaf size) shape in a process times each on a
// prefix of the leaf table (the same keys written twice, so the result is
// unchanged) and keeps the faster; KNNG_LEAF_VARIANT=v forces v.
// ---------------------------------------------------------------------------
constexpr int kLeafVariants = 2;

int launch_leaf_gram(knng_ctx* c, int v, uint64_t* cand, int64_t L0, int64_t nl) {
  if (nl <= 0) return 0;
  const int64_t n = c->n;
  const int T = c->p.num_trees;
  if (v == 1)
    hipLaunchKernelGGL(k_leaf_gram_wave, dim3(grid_for(c, nl, 4, 64)), dim3(256), 0, c->stream, c->q, c->d, n, T,
                       c->perm, c->leaves + L0, nl, cand);
  else
    hipLaunchKernelGGL(k_leaf_gram, dim3(grid_for(c, nl, 1, 64)), dim3(256), 0, c->stream, c->q, c->d, n, T,
                       c->perm, c->leaves + L0, nl, cand);
  HIPCHK(c, hipGetLastError());
  return 0;
}

struct LeafTuneCache {
  std::mutex mu;
  std::map<std::pair<int, int>, int> best;
};

LeafTuneCache& leaf_tune_cache() {
  static LeafTuneCache tc;
  return tc;
}

int ensure_leaf_variant(knng_ctx* c, uint64_t* cand) {
  if (c->leaf_variant >= 0) return 0;
  c->st.ms_tune_leaf = 0;
  if (const char* e = std::getenv("KNNG_LEAF_VARIANT")) {
    const int v = std::atoi(e);
    if (v >= 0 && v < kLeafVariants) return c->leaf_variant = v, 0;
  }
  const auto key = std::make_pair(c->d, int(c->p.leaf_size));
  {
    std::lock_guard<std::mutex> lk(leaf_tune_cache().mu);
    auto it = leaf_tune_cache().best.find(key);
    if (it != leaf_tune_cache().best.end()) return c->leaf_variant = it->second, 0;
  }
  // too few leaves to time reliably: the first shape, not cached
  const int64_t nl = std::min<int64_t>(c->nleaves, std::max<int64_t>(16384, c->nleaves / 16));
  if (nl < 16384) return c->leaf_variant = 0, 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(c, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(c, KNNG_EHIP, "hipEventCreate failed (leaf-tile autotuner)");
  }
  float best_ms = 0, total = 0;
  int best = 0, rc = 0;
  // a failed launch ends the tuning at once: no later timing may pick a shape
  for (int round = 0; round < 2 && !rc; ++round)
    for (int v = 0; v < kLeafVariants && !rc; ++v) {
      (void)hipEventRecord(e0, c->stream);
      if ((rc = launch_leaf_gram(c, v, cand, 0, nl))) break;
      (void)hipEventRecord(e1, c->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      total += ms;
      if (round == 1 && (v == 0 || ms < best_ms)) {
        best_ms = ms;
        best = v;
      }
    }
  (void)hipStreamSynchronize(c->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  c->st.ms_tune_leaf = total;
  c->leaf_variant = best;
  std::lock_guard<std::mutex> lk(leaf_tune_cache().mu);
  leaf_tune_cache().best[key] = best;
  return 0;
}

// FP32 rerank export of rows [base, base + nrows) of graph buffer `cur`: one
// k_rerank_slice pass per slice, so single and packed builds sum the slices
// in the same order.  Stream-ordered on c->stream.
int launch_rerank(knng_ctx* c, int cur, int64_t nrows, int64_t base, int k, int32_t* dk, float* dd,
                  const int32_t* order) {
  const float* x = c->x_ext ? c->x_ext : c->x;
  if (!x) return fail(c, KNNG_ESTATE, "knng_export: rerank needs float32 rows (keep_fp32 or attach)");
  const int K = c->p.k_build;
  int R = c->p.rerank_width ? c->p.rerank_width : K;
  if (R < k) R = k;
  // every block resident at once, so each XCD's rows in flight are one
  // window of consecutive cluster-mates whose candidate slices stay in its L2
  const int ch = rerank_chunks(c->d);
  const int S = (c->d + 64 * ch - 1) / (64 * ch);
  int rc;
  if (S > 1 && (rc = dalloc(c, &c->rr_part, size_t(c->n) * K, "rerank partial sums"))) return rc;
  // each pass's 8 row-queue counters, 256 B apart, cleared by a kernel (no
  // memset node if a caller captures the export)
  if ((rc = dalloc(c, &c->rr_q, size_t(kRerankMaxPasses) * 8 * 64, "rerank row queues"))) return rc;
  if (S > kRerankMaxPasses) return fail(c, KNNG_EINVAL, "knng_export: too many rerank slices");
  hipLaunchKernelGGL(k_zero_u64, dim3(16), dim3(256), 0, c->stream,
                     reinterpret_cast<unsigned long long*>(c->rr_q), S * 8 * 32);
  const int64_t steps = (nrows + kRerankWaves - 1) / kRerankWaves;
  auto run = [&](auto first, auto mid, auto last, auto only) {
    for (int s = 0; s < S; ++s) {
      auto launch = [&](auto kern) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(kern),
                                                         64 * kRerankWaves, 0) != hipSuccess || nb < 1)
          nb = 1;
        // every block resident; no more blocks than steps (rounded up to the 8 XCDs)
        const int64_t want = std::min<int64_t>(int64_t(c->num_cu) * nb, (steps + 7) / 8 * 8);
        const int grid = int(std::max<int64_t>(8, want / 8 * 8));
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kRerankWaves), 0, c->stream, x, c->d, nrows, K, R, k,
                           s, c->gid[cur], c->rr_part, dk, dd, order, base, c->rr_q + size_t(s) * 8 * 64);
      };
      if (S == 1)
        launch(only);
      else if (s == 0)
        launch(first);
      else if (s < S - 1)
        launch(mid);
      else
        launch(last);
    }
  };
  if (ch == 1)
    run(k_rerank_slice<1, true, false>, k_rerank_slice<1, false, false>, k_rerank_slice<1, false, true>,
        k_rerank_slice<1, true, true>);
  else if (ch == 2)
    run(k_rerank_slice<2, true, false>, k_rerank_slice<2, false, false>, k_rerank_slice<2, false, true>,
        k_rerank_slice<2, true, true>);
  else if (ch == 4)
    run(k_rerank_slice<4, true, false>, k_rerank_slice<4, false, false>, k_rerank_slice<4, false, true>,
        k_rerank_slice<4, true, true>);
  else
    run(k_rerank_slice<8, true, false>, k_rerank_slice<8, false, false>, k_rerank_slice<8, false, true>,
        k_rerank_slice<8, true, true>);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free join iteration (knng_join_iteration_noalloc).  Its only
// device memory beyond the caller's inputs and outputs is `scratch`: the
// join's per-block counters and per-XCD row-queue counters, 51,200 bytes per
// GPU, held until the process ends.  Host code never passes the array's name
// to a kernel: each device's address comes from hipGetSymbolAddress, taken
// once per device with that device current.  Every call on a GPU shares that
// GPU's array, so calls that overlap on one GPU (on different streams) race.
// ---------------------------------------------------------------------------

constexpr int kScratchWords = kCounterFields * kNumCounters + kRowQueueWords;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
struct ScratchTable {
  std::mutex mu;
  unsigned long long* ptr[kMaxDevices] = {};
  int num_cu[kMaxDevices] = {};
};

ScratchTable& scratch_table() {
  static ScratchTable t;
  return t;
}

// this device's copy of `scratch` and its CU count (hipGetSymbolAddress on first use)
int scratch_for_device(int dev, unsigned long long** ptr, int* num_cu) {
  if (dev < 0 || dev >= kMaxDevices) return KNNG_ERANGE;
  ScratchTable& t = scratch_table();
  std::lock_guard<std::mutex> lk(t.mu);
  if (!t.ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return KNNG_EHIP;
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
      return KNNG_EHIP;
    t.ptr[dev] = static_cast<unsigned long long*>(p);
    t.num_cu[dev] = cu;
  }
  *ptr = t.ptr[dev];
  *num_cu = t.num_cu[dev];
  return 0;
}

// v[0..m) = 0 (a kernel, not hipMemsetAsync, so a captured graph replays it correctly)
}  // namespace

extern "C" {

int knng_default_params(knng_params* p) {
  if (!p) return KNNG_EINVAL;
  p->dim = 1024;
  p->k_eval = 15;
  p->k_build = 50;
  p->num_trees = 16;
  p->rerank_mult = 8;
  p->leaf_size = 64;
  p->max_iters = 30;
  p->early_stop_rate = 0.04f;
  p->rate_diff_tol = 0.01f;
  p->seed = 1337;
  p->keep_fp32 = 1;
  p->rerank_width = 0;
  p->stop_rule = KNNG_STOP_INSERTED;
  p->insert_stop_rate = 0.01f;
  p->incremental = 0;
  p->cells = -1;
  p->kmeans_iters = 2;
  p->cell_partitions = 2;
  p->cell_oversample = 4;
  p->cell_sample = 8;
  p->cell_merge = 0.6f;
  p->cell_full_iters = 0;
  p->cell_prune = 0.25f;
  p->reverse_width = 0;
  return 0;
}

const char* knng_last_error(const knng_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int knng_create(knng_ctx** out, int32_t device, const knng_params* p, int64_t n) {
  if (!out || !p) return KNNG_EINVAL;
  *out = nullptr;
  std::string why;
  if (int rc = check_params(p, n, &why)) {
    fprintf(stderr, "[knng] knng_create: %s\n", why.c_str());
    return rc;
  }
  knng_ctx* c = new knng_ctx();
  c->device = device;
  c->p = *p;
  c->n = n;
  c->d = p->dim;
  c->words = p->dim / 64;
  auto bail = [&](int rc) {
    fprintf(stderr, "[knng] knng_create: %s\n", c->err.c_str());
    knng_destroy(c);
    return rc;
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return bail(fail(c, KNNG_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
  // a blocking stream: work the caller queued on the null stream (e.g. the
  // kernels that produced a device input) completes before ours starts
  if (hipStreamCreateWithFlags(&c->stream, hipStreamDefault) != hipSuccess)
    return bail(fail(c, KNNG_EHIP, "hipStreamCreate failed"));
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return bail(fail(c, KNNG_EHIP, "hipEventCreate failed"));
  int rc;
  if ((rc = dalloc(c, &c->q, size_t(n) * c->d, "int8 rows"))) return bail(rc);
  if ((rc = dalloc(c, &c->sk, size_t(n) * c->words, "sketches"))) return bail(rc);
  if (p->keep_fp32 && (rc = dalloc(c, &c->x, size_t(n) * c->d, "float32 rows"))) return bail(rc);
  if ((rc = dalloc(c, &c->counters, kCounterFields * kNumCounters + kRowQueueWords, "counters"))) return bail(rc);
  // zero once, on the context's stream (a null-stream memset would wait for
  // every blocking stream of the process, e.g. other contexts' builds);
  // k_sum_counters re-zeroes after every join launch
  if (hipMemsetAsync(c->counters, 0, (kCounterFields * kNumCounters + kRowQueueWords) * sizeof(unsigned long long),
                     c->stream) != hipSuccess)
    return bail(fail(c, KNNG_EHIP, "hipMemsetAsync of the join counters failed"));
  if ((rc = dalloc(c, &c->csum, kCounterFields, "counter sums"))) return bail(rc);
  c->st.n = n;
  c->st.dim = c->d;
  *out = c;
  return 0;
}

int knng_destroy(knng_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dfree(&c->x);
  dfree(&c->q);
  dfree(&c->sk);
  dfree(&c->perm);
  dfree(&c->perm2);
  dfree(&c->nodepos);
  dfree(&c->nodepos2);
  dfree(&c->flags);
  dfree(&c->scan);
  dfree(&c->blocksums);
  dfree(&c->nodes);
  dfree(&c->nodes2);
  dfree(&c->leaves);
  dfree(&c->hier);
  dfree(&c->cz);
  dfree(&c->cw);
  dfree(&c->cassign);
  dfree(&c->cell_root);
  dfree(&c->dcount);
  for (int b = 0; b < 2; ++b) {
    dfree(&c->gid[b]);
    dfree(&c->gsc[b]);
    dfree(&c->fresh[b]);
    dfree(&c->dstage[b]);
    if (b == 0) {
      dfree(&c->ovf);
      dfree(&c->ovf_n);
    }
    if (c->pin[b]) (void)hipHostFree(c->pin[b]);
    if (c->pin_ev[b]) (void)hipEventDestroy(c->pin_ev[b]);
  }
  dfree(&c->counters);
  dfree(&c->csum);
  dfree(&c->leafkey);
  dfree(&c->gmask[0]);
  dfree(&c->gmask[1]);
  dfree(&c->jcbuf);
  dfree(&c->jccount);
  dfree(&c->seg_base);
  dfree(&c->tune_n);
  dfree(&c->ex_knns);
  dfree(&c->ex_dists);
  dfree(&c->rr_part);
  dfree(&c->rr_q);
  dfree(&c->rv_slots);
  dfree(&c->rv_id);
  dfree(&c->rv_mask);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int knng_synchronize(knng_ctx* c) {
  if (!c) return KNNG_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int knng_attach_fp32_device(knng_ctx* c, const float* dev_x) {
  if (!c || !dev_x) return KNNG_EINVAL;
  c->x_ext = dev_x;
  return 0;
}

int knng_host_alloc(knng_ctx* c, void** ptr, int64_t bytes) {
  if (!c || !ptr || bytes <= 0) return KNNG_EINVAL;
  *ptr = nullptr;
  HIPCHK(c, hipSetDevice(c->device));
  if (hipHostMalloc(ptr, size_t(bytes), hipHostMallocDefault) != hipSuccess) {
    *ptr = nullptr;
    return fail(c, KNNG_ENOMEM, "hipHostMalloc failed (" + std::to_string(bytes) + " B)");
  }
  return 0;
}

int knng_host_free(knng_ctx* c, void* ptr) {
  if (!c) return KNNG_EINVAL;
  if (ptr) HIPCHK(c, hipHostFree(ptr));
  return 0;
}

int knng_ingest_device(knng_ctx* c, const float* dev_rows, int64_t row0, int64_t nrows) {
  if (!c || !dev_rows || row0 < 0 || nrows < 0 || row0 + nrows > c->n)
    return fail(c, KNNG_EINVAL, "knng_ingest_device: bad row range");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  if (c->x)
    HIPCHK(c, hipMemcpyAsync(c->x + row0 * c->d, dev_rows, size_t(nrows) * c->d * sizeof(float),
                             hipMemcpyDeviceToDevice, c->stream));
  if (int rc = quantize_rows(c, dev_rows, row0, nrows)) return rc;
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);   // ms stays 0 on failure (a timing stat only)
  c->st.ms_ingest += ms;
  c->ingested += nrows;
  return 0;
}

int knng_ingest_chunk(knng_ctx* c, const float* rows, int64_t row0, int64_t nrows) {
  if (!c || !rows || row0 < 0 || nrows < 0 || row0 + nrows > c->n)
    return fail(c, KNNG_EINVAL, "knng_ingest_chunk: bad row range");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t row_bytes = size_t(c->d) * sizeof(float);
  const int64_t slice_rows = std::max<int64_t>(1, int64_t(c->stage_bytes / row_bytes));
  for (int b = 0; b < 2; ++b) {
    if (!c->pin[b]) {
      HIPCHK(c, hipHostMalloc(&c->pin[b], c->stage_bytes, hipHostMallocDefault));
      HIPCHK(c, hipEventCreateWithFlags(&c->pin_ev[b], hipEventDisableTiming));
      HIPCHK(c, hipEventRecord(c->pin_ev[b], c->stream));
    }
    if (!c->x && !c->dstage[b])
      if (int rc = dalloc(c, &c->dstage[b], c->stage_bytes / sizeof(float), "ingest staging")) return rc;
  }
  // a caller buffer that is already pinned is copied directly
  bool pinned = false;
  {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, rows) == hipSuccess && attr.type == hipMemoryTypeHost)
      pinned = true;
    (void)hipGetLastError();
  }
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  int buf = 0;
  for (int64_t r = 0; r < nrows; r += slice_rows) {
    const int64_t m = std::min(slice_rows, nrows - r);
    const float* src = rows + r * c->d;
    if (!pinned) {
      HIPCHK(c, hipEventSynchronize(c->pin_ev[buf]));   // staging buffer free again
      memcpy(c->pin[buf], src, size_t(m) * row_bytes);
      src = static_cast<const float*>(c->pin[buf]);
    }
    float* dst = c->x ? c->x + (row0 + r) * c->d : c->dstage[buf];
    HIPCHK(c, hipMemcpyAsync(dst, src, size_t(m) * row_bytes, hipMemcpyHostToDevice, c->stream));
    if (int rc = quantize_rows(c, dst, row0 + r, m)) return rc;
    HIPCHK(c, hipEventRecord(c->pin_ev[buf], c->stream));
    buf ^= 1;
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);   // ms stays 0 on failure (a timing stat only)
  c->st.ms_ingest += ms;
  c->ingested += nrows;
  return 0;
}

static int scan_flags(knng_ctx* c, int64_t total) {
  const int64_t nb = (total + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, c->stream, c->flags, total, c->blocksums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, c->stream, c->blocksums, nb, c->dcount + 3);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(256), 0, c->stream, c->flags, total, c->blocksums,
                     c->scan);
  // scan[total] = grand total
  HIPCHK(c, hipMemcpyAsync(c->scan + total, c->dcount + 3, sizeof(int32_t), hipMemcpyDeviceToDevice,
                           c->stream));
  HIPCHK(c, hipGetLastError());
  return 0;
}

int knng_build_forest(knng_ctx* c) {
  if (!c) return KNNG_EINVAL;
  if (c->ingested < c->n) return fail(c, KNNG_ESTATE, "knng_build_forest: not all rows ingested");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t n = c->n, T = c->p.num_trees, total = T * n;
  const int leaf = c->p.leaf_size;
  c->node_cap = T * (n / (leaf + 1) + 1) + 16;
  // every non-empty cell below the leaf size is a leaf of its own
  c->leaf_cap = total / 4 + T + 1024 + T * int64_t(c->nseg) +
                T * std::min<int64_t>(int64_t(cells_for(c->p, n)) * std::max(1, c->p.cell_oversample), n);
  int rc;
  if ((rc = dalloc(c, &c->perm, total, "perm")) || (rc = dalloc(c, &c->perm2, total, "perm2")) ||
      (rc = dalloc(c, &c->nodepos, total, "nodepos")) ||
      (rc = dalloc(c, &c->nodepos2, total, "nodepos2")) ||
      (rc = dalloc(c, &c->flags, total, "flags")) || (rc = dalloc(c, &c->scan, total + 1, "scan")) ||
      (rc = dalloc(c, &c->blocksums, total / kScanTile + 2, "scan blocks")) ||
      (rc = dalloc(c, &c->nodes, c->node_cap, "nodes")) ||
      (rc = dalloc(c, &c->nodes2, c->node_cap, "nodes2")) ||
      (rc = dalloc(c, &c->leaves, c->leaf_cap, "leaves")) || (rc = dalloc(c, &c->dcount, 8, "counts")) ||
      (rc = dalloc(c, &c->hier, c->leaf_cap, "forest hierarchy")))
    return rc;
  c->hier_count = 0;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  HIPCHK(c, hipMemsetAsync(c->dcount, 0, 8 * sizeof(int32_t), c->stream));
  hipLaunchKernelGGL(k_forest_seed, dim3(grid_for(c, total, 256, 16)), dim3(256), 0, c->stream, n,
                     int(T), leaf, c->p.seed, c->perm, c->nodepos, c->nodes);
  HIPCHK(c, hipGetLastError());
  int nnodes = int(T);
  int levels = 0;
  if (c->nseg > 0) {
    // one root per (tree, segment) above the leaf size, else a leaf; pivots
    // as in the segment's own build (start 0, base = its first row)
    const int S = c->nseg;
    std::vector<Node> roots;
    std::vector<int4> lv;
    std::vector<int32_t> segnode(size_t(T) * S, -1);
    for (int t = 0; t < T; ++t)
      for (int sg = 0; sg < S; ++sg) {
        const int64_t b = c->seg_host[size_t(sg)], z = c->seg_host[size_t(sg) + 1] - b;
        if (z > leaf) {
          Node nd;
          nd.tree = t;
          nd.start = int32_t(b);
          nd.size = int32_t(z);
          node_pivots(c->p.seed, t, 0, 0, z, &nd.p1, &nd.p2);
          nd.left = nd.right = -1;
          nd.nleft = 0;
          nd.p1id = nd.p2id = -1;
          nd.base = int32_t(b);
          segnode[size_t(t) * S + sg] = int32_t(roots.size());
          roots.push_back(nd);
        } else {
          lv.push_back(make_int4(t, int(b), int(z), 0));
        }
      }
    if (int64_t(roots.size()) > c->node_cap || int64_t(lv.size()) > c->leaf_cap)
      return fail(c, KNNG_ERANGE, "packed builds exceed the node/leaf tables");
    int32_t* dsegnode = nullptr;
    if ((rc = dalloc(c, &dsegnode, segnode.size(), "segment roots"))) return rc;
    HIPCHK(c, hipMemcpyAsync(dsegnode, segnode.data(), segnode.size() * 4, hipMemcpyHostToDevice,
                             c->stream));
    if (!roots.empty())
      HIPCHK(c, hipMemcpyAsync(c->nodes, roots.data(), roots.size() * sizeof(Node), hipMemcpyHostToDevice,
                               c->stream));
    if (!lv.empty())
      HIPCHK(c, hipMemcpyAsync(c->leaves, lv.data(), lv.size() * sizeof(int4), hipMemcpyHostToDevice,
                               c->stream));
    int32_t cnt[2] = {0, int32_t(lv.size())};
    HIPCHK(c, hipMemcpyAsync(c->dcount, cnt, sizeof(cnt), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_forest_seed_segs, dim3(grid_for(c, total, 256, 16)), dim3(256), 0, c->stream, n,
                       int(T), c->seg_base, S, dsegnode, c->perm, c->nodepos);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(&dsegnode);
    nnodes = int(roots.size());
  } else if (n <= leaf) {
    // every tree is one leaf
    std::vector<int4> lv(T);
    for (int t = 0; t < T; ++t) lv[t] = make_int4(t, 0, int(n), 0);
    HIPCHK(c, hipMemcpyAsync(c->leaves, lv.data(), T * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    int32_t cnt[2] = {0, int32_t(T)};
    HIPCHK(c, hipMemcpyAsync(c->dcount, cnt, sizeof(cnt), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    nnodes = 0;
  }
  c->ncells = 0;
  c->nparts = 0;
  c->cell_count.clear();
  c->st.ms_kmeans = 0;
  c->st.assign_ops = 0;
  const int32_t ncells = c->nseg > 0 ? 0 : cells_for(c->p, n);
  if (nnodes > 0 && ncells >= 2) {
    // k-means cells replace the root split of every tree
    c->nparts = std::max(1, std::min<int>(c->p.cell_partitions, int(T)));
    hipEvent_t k0, k1;
    HIPCHK(c, hipEventCreate(&k0));
    HIPCHK(c, hipEventCreate(&k1));
    HIPCHK(c, hipEventRecord(k0, c->stream));
    trace_mark(c, "forest start");
    rc = seed_from_cells(c, int(std::min<int64_t>(ncells, n)), &nnodes);
    (void)hipEventRecord(k1, c->stream);
    (void)hipEventSynchronize(k1);
    (void)hipEventElapsedTime(&c->st.ms_kmeans, k0, k1);
    (void)hipEventDestroy(k0);
    (void)hipEventDestroy(k1);
    if (rc) return rc;
    levels = 1;
  }
  c->st.cells = 0;
  for (int32_t v : c->cell_count) c->st.cells += v;
  const int nch = c->d / 16;
  while (nnodes > 0) {
    const int run = 16;
    const int fgrid = grid_for(c, (total + run - 1) / run, 4, 32);
    if (nch <= 64)
      hipLaunchKernelGGL(k_forest_flags<1>, dim3(fgrid), dim3(256), 0, c->stream, c->q, c->d, n, total,
                         c->perm, c->nodepos, c->nodes, c->flags, run);
    else
      return fail(c, KNNG_ERANGE, "dim too large for the forest kernel");
    if ((rc = scan_flags(c, total))) return rc;
    HIPCHK(c, hipMemsetAsync(c->dcount, 0, sizeof(int32_t), c->stream));   // next-node count
    hipLaunchKernelGGL(k_forest_nodes, dim3((nnodes + 255) / 256), dim3(256), 0, c->stream, n, leaf,
                       c->p.seed, levels, c->nodes, nnodes, c->scan, c->perm, c->nodes2, c->dcount,
                       c->leaves, c->dcount + 1, c->leaf_cap, c->dcount + 2);
    if (c->hier_count + nnodes > c->leaf_cap) return fail(c, KNNG_ERANGE, "forest hierarchy overflow");
    hipLaunchKernelGGL(k_hier_append, dim3((nnodes + 255) / 256), dim3(256), 0, c->stream, c->nodes,
                       nnodes, c->hier + c->hier_count, c->hier_count + nnodes);
    c->hier_count += nnodes;
    hipLaunchKernelGGL(k_forest_partition, dim3(grid_for(c, total, 256, 16)), dim3(256), 0, c->stream,
                       n, total, c->perm, c->nodepos, c->nodes, c->flags, c->scan, c->perm2,
                       c->nodepos2);
    HIPCHK(c, hipGetLastError());
    int32_t cnt[3];
    HIPCHK(c, hipMemcpyAsync(cnt, c->dcount, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (cnt[2]) return fail(c, KNNG_ERANGE, "leaf table overflow");
    std::swap(c->perm, c->perm2);
    std::swap(c->nodepos, c->nodepos2);
    std::swap(c->nodes, c->nodes2);
    nnodes = cnt[0];
    ++levels;
    if (levels > 4096) return fail(c, KNNG_ERANGE, "forest did not terminate");
  }
  trace_mark(c, "forest levels");
  int32_t nl = 0;
  HIPCHK(c, hipMemcpyAsync(&nl, c->dcount + 1, sizeof(nl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);   // ms stays 0 on failure (a timing stat only)
  c->st.ms_forest = ms;
  c->st.forest_levels = levels;
  c->st.num_leaves = nl;
  c->nleaves = nl;
  c->have_forest = true;
  c->have_graph = false;
  return 0;
}

int knng_build_init(knng_ctx* c) {
  if (!c) return KNNG_EINVAL;
  if (!c->have_forest) return fail(c, KNNG_ESTATE, "knng_build_init: no forest");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t n = c->n;
  const int T = c->p.num_trees, K = c->p.k_build;
  int rc;
  for (int b = 0; b < 2; ++b)
    if ((rc = dalloc(c, &c->gid[b], size_t(n) * K, "graph ids")) ||
        (rc = dalloc(c, &c->gsc[b], size_t(n) * K, "graph scores")) ||
        (rc = dalloc(c, &c->fresh[b], size_t(n), "fresh masks")) ||
        (rc = dalloc(c, &c->gmask[b], size_t(n) * K, "leaf mask bytes")))
      return rc;
  HIPCHK(c, hipMemsetAsync(c->fresh[0], 0xFF, size_t(n) * sizeof(uint64_t), c->stream));   // all fresh
  c->fresh_all = true;
  // scratch of this call, freed on every return
  ScopedDev<uint64_t> cand_buf;
  ScopedDev<int4> items_buf;
  ScopedDev<int32_t> scores_buf;
  uint64_t*& cand = cand_buf.p;
  if ((rc = dalloc(c, &cand, size_t(n) * T * kLeafSlots, "init candidates"))) return rc;
  c->st.init_tile_ops = 0;
  if ((rc = ensure_leaf_variant(c, cand))) return rc;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  if ((rc = launch_leaf_gram(c, c->leaf_variant, cand, 0, c->nleaves))) return rc;
  if (c->p.leaf_size > kLeafSlots) {
    // leaves above 64 rows: work items of 128 members each
    std::vector<int4> lv(size_t(c->nleaves));
    HIPCHK(c, hipMemcpyAsync(lv.data(), c->leaves, lv.size() * sizeof(int4), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<int4> items;
    int mmax = 0;
    for (const int4& l : lv)
      if (l.z > kLeafSlots) {
        c->st.init_tile_ops += double(l.z) * double(l.z) * 2.0 * double(c->d);
        mmax = std::max(mmax, int(l.z));
        for (int r0 = 0; r0 < l.z; r0 += kBigRows) items.push_back(make_int4(l.x, l.y, l.z, r0));
      }
    if (mmax > kMaxLeaf) return fail(c, KNNG_ERANGE, "leaf larger than the big-leaf kernel allows");
    if (!items.empty()) {
      // items in chunks whose score rows fit a bounded scratch (<= 4 GB):
      // scores on MFMA (one wave per SIMD), then the top-64 selection with
      // one wave per row at full occupancy
      const int mpad = (mmax + 63) & ~63;
      const int64_t per_item = int64_t(kBigRows) * mpad;
      const int64_t chunk = std::max<int64_t>(
          1, std::min<int64_t>(int64_t(items.size()), (int64_t(1) << 30) / per_item));
      int4*& ditems = items_buf.p;
      int32_t*& scratch = scores_buf.p;
      if ((rc = dalloc(c, &ditems, items.size(), "big-leaf items")) ||
          (rc = dalloc(c, &scratch, size_t(chunk * per_item + kMaxLeaf), "big-leaf scores")))
        return rc;
      HIPCHK(c, hipMemcpyAsync(ditems, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice,
                               c->stream));
      // rows the selection passes on to k_leaf_select_slow, per chunk
      ScopedDev<int32_t> slow_buf;
      ScopedDev<unsigned int> nslow_buf;
      int32_t*& slow = slow_buf.p;
      unsigned int*& nslow = nslow_buf.p;
      if ((rc = dalloc(c, &slow, size_t(chunk * kBigRows), "leaf selection slow rows")) ||
          (rc = dalloc(c, &nslow, 1, "leaf selection slow count")))
        return rc;
      hipError_t e = hipSuccess;
      for (int64_t i0 = 0; i0 < int64_t(items.size()) && e == hipSuccess; i0 += chunk) {
        const int64_t ni = std::min<int64_t>(chunk, int64_t(items.size()) - i0);
        HIPCHK(c, hipMemsetAsync(nslow, 0, sizeof(unsigned int), c->stream));
        const int sgrid = int(std::min<int64_t>(ni, int64_t(c->num_cu) * 2));
        const int wgrid = int(std::min<int64_t>(ni, int64_t(c->num_cu) * 4));   // two resident per CU
        if (c->d == kMaxDim)
          hipLaunchKernelGGL(k_leaf_scores_wide, dim3(wgrid), dim3(256), 0, c->stream, c->q, n, c->perm, ditems,
                             i0, ni, scratch, mpad);
        else
          hipLaunchKernelGGL((k_leaf_scores<0>), dim3(sgrid), dim3(256), 0, c->stream, c->q, c->d, n, c->perm,
                             ditems, i0, ni, scratch, mpad);
        hipLaunchKernelGGL(k_leaf_select, dim3(grid_for(c, ni * kBigRows, 4, 32)), dim3(256), 0, c->stream, n, T,
                           c->perm, ditems, i0, ni, scratch, mpad, cand, slow, nslow);
        hipLaunchKernelGGL(k_leaf_select_slow, dim3(c->num_cu), dim3(256), 0, c->stream, n, T, c->perm, ditems,
                           i0, scratch, mpad, cand, slow, nslow);
        e = hipGetLastError();
      }
      HIPCHK(c, hipStreamSynchronize(c->stream));
      dfree(&ditems);
      dfree(&scratch);
      if (e != hipSuccess) return fail(c, KNNG_EHIP, std::string("big-leaf launch: ") + hipGetErrorString(e));
    }
  }
  const int slots = T * kLeafSlots;
  const int mgrid = grid_for(c, n, 1, 32);
  c->cur = 0;
  const int64_t* sb = c->nseg ? c->seg_base : nullptr;
  // one wave per row up to 1024 slots (16 trees), one block per row beyond
  const int wgrid = grid_for(c, n, 4, 16);
  if (slots <= 256)
    hipLaunchKernelGGL(k_init_merge_wave<256>, dim3(wgrid), dim3(256), 0, c->stream, c->q, c->d, n, T, K,
                       cand, c->gid[0], c->gsc[0], sb, c->nseg);
  else if (slots <= 512)
    hipLaunchKernelGGL(k_init_merge_wave<512>, dim3(wgrid), dim3(256), 0, c->stream, c->q, c->d, n, T, K,
                       cand, c->gid[0], c->gsc[0], sb, c->nseg);
  else if (slots <= 1024)
    hipLaunchKernelGGL(k_init_merge_wave<1024>, dim3(wgrid), dim3(256), 0, c->stream, c->q, c->d, n, T, K,
                       cand, c->gid[0], c->gsc[0], sb, c->nseg);
  else if (slots <= 2048)
    hipLaunchKernelGGL(k_init_merge<2048>, dim3(mgrid), dim3(256), 0, c->stream, c->q, c->d, n, T, K,
                       cand, c->gid[0], c->gsc[0], sb, c->nseg);
  else
    hipLaunchKernelGGL(k_init_merge<4096>, dim3(mgrid), dim3(256), 0, c->stream, c->q, c->d, n, T, K,
                       cand, c->gid[0], c->gsc[0], sb, c->nseg);
  HIPCHK(c, hipGetLastError());
  // every row's leaf in every tree, for the join's leaf-mate rule
  c->leaf_tp = (T + 3) & ~3;
  c->leaf_valid = false;
  if (!c->leafkey && (rc = dalloc(c, &c->leafkey, size_t(n) * c->leaf_tp, "leaf keys"))) return rc;
  hipLaunchKernelGGL(k_leaf_keys, dim3(grid_for(c, c->nleaves, 4, 16)), dim3(256), 0, c->stream, c->leaves,
                     c->nleaves, c->perm, n, c->leaf_tp, c->leafkey);
  hipLaunchKernelGGL(k_edge_masks, dim3(rerank_grid(c, n)), dim3(256), 0, c->stream, c->gid[0], n, K, c->leafkey,
                     T, c->leaf_tp, c->gmask[0], c->perm);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);   // ms stays 0 on failure (a timing stat only)
  c->st.ms_init = ms;
  dfree(&cand);
  c->have_graph = true;
  c->leaf_valid = true;
  return 0;
}

int knng_join_iteration(knng_ctx* c, int64_t* updates) {
  if (!c) return KNNG_EINVAL;
  if (!c->have_graph) return fail(c, KNNG_ESTATE, "knng_join_iteration: no graph");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t n = c->n;
  const int K = c->p.k_build, M = K * c->p.rerank_mult;
  const int a = c->cur, b = c->cur ^ 1;
  const int32_t* order = c->have_forest ? c->perm : nullptr;   // tree 0's leaf order
  // default: the phased join, k_join for the rows it cannot hold; with
  // KNNG_JOIN_FUSED=1 the block-per-row k_join (and, under the new/old rule
  // after the first iteration, its small-row variant) over every row
  const bool phased = !join_fused_forced();
  if (phased) {
    if (c->join_variant < 0) c->join_variant = 0;
  } else if (int rc = ensure_join_variant(c, a, b, order, n, nullptr, 0, nullptr)) {
    return rc;
  }
  const bool small = !phased && c->p.incremental && !c->fresh_all && K <= 50;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  if (c->p.reverse_width > 0)
    if (int rc = build_reverse_lists(c, a)) return rc;
  if (phased) {
    if (int rc = launch_join_phased(c, a, b, order)) return rc;
  } else if (small) {
    if (int rc = launch_join_small(c, a, b, order)) return rc;
  } else if (int rc = launch_join(c, c->join_variant, a, b, order, nullptr, nullptr, 0, nullptr)) {
    return rc;
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  c->fresh_all = false;
  hipLaunchKernelGGL(k_sum_counters, dim3(1), dim3(256), 0, c->stream, c->counters, c->csum);
  unsigned long long sums[kCounterFields];
  unsigned long long passed = 0;
  HIPCHK(c, hipMemcpyAsync(sums, c->csum, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
  if (small || phased) HIPCHK(c, hipMemcpyAsync(&passed, c->ovf_n, sizeof(passed), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);   // ms stays 0 on failure (a timing stat only)
  c->cur = b;
  const int it = c->st.iterations;
  if (it < 32) {
    c->st.updates[it] = int64_t(sums[0]);
    c->st.rates[it] = float(sums[0]) / float(size_t(n) * size_t(K));
    c->st.ms_join_iter[it] = ms;
    c->st.join_candidates[it] = int64_t(sums[1]);
    c->st.join_rescored[it] = int64_t(sums[2]);
    c->st.join_inserted[it] = int64_t(sums[3]);
    c->st.insert_rates[it] = float(sums[3]) / float(size_t(n) * size_t(K));
    c->st.join_passed_on[it] = (small || phased) ? int64_t(passed) : -1;
    c->st.join_leafmates[it] = int64_t(sums[5]);
    // algorithmic bytes: own list + own row/sketch, K neighbour id rows, one
    // sketch per candidate of a row that gathers them, one row per re-scored
    // candidate, writes; with the leaf-mate rule also the mask bytes (own,
    // neighbours', written) and one leaf-key row per new entry
    const double lk_bytes =
        c->leaf_valid ? double(n) * (2.0 * K + double(K) * K) + double(sums[3]) * 4.0 * c->leaf_tp : 0.0;
    c->st.join_bytes[it] = double(n) * (8.0 * K + c->d + 8.0 * c->words + 4.0 * K * K + 8.0 * K) + lk_bytes +
                           double(sums[4]) * 8.0 * c->words + double(sums[2]) * c->d;
    // reverse lists: offered edges (ids, scores, fresh masks; R slot loads
    // each), the lists written, and the candidate stage's reads (own and
    // neighbours' reverse entries, the lists of the own reverse entries; an
    // upper bound: every reverse slot taken as filled)
    if (const int R = c->p.reverse_width) {
      const double offered = (c->p.incremental && it > 0) ? double(c->st.join_inserted[it - 1]) : double(n) * K;
      c->st.join_bytes[it] += double(n) * (8.0 * K + 8.0) + offered * 8.0 * R + double(n) * R * (16.0 + 5.0) +
                              double(n) * (5.0 * R + 5.0 * K * R + 5.0 * K * R);
    }
  }
  c->st.iterations = it + 1;
  c->st.ms_join += ms;
  if (updates) *updates = int64_t(sums[0]);
  return 0;
}

int knng_build(knng_ctx* c) {
  if (!c) return KNNG_EINVAL;
  if (c->nseg > 0) return fail(c, KNNG_ESTATE, "knng_build: packed builds run through knng_build_batch");
  memset(c->st.updates, 0, sizeof(c->st.updates));
  c->st.iterations = 0;
  c->st.ms_join = 0;
  int rc;
  if ((rc = knng_build_forest(c))) return rc;
  if ((rc = knng_build_init(c))) return rc;
  float prev = 1.0f;
  for (int it = 0; it < c->p.max_iters; ++it) {
    int64_t upd = 0;
    if ((rc = knng_join_iteration(c, &upd))) return rc;
    // KNNG_STOP_REFERENCE: src/main.cpp:432-439 on the rate of slots whose id
    // changed (a shift counts).  That rate stays near 1 while lists improve
    // anywhere near their top, so from a weak init it stops after two
    // iterations; KNNG_STOP_INSERTED instead stops once fewer than
    // insert_stop_rate of all slots received a new id.
    const int j = c->st.iterations - 1 < 31 ? c->st.iterations - 1 : 31;
    if (c->p.stop_rule == KNNG_STOP_NONE) continue;
    if (c->p.stop_rule == KNNG_STOP_INSERTED) {
      if (c->st.insert_rates[j] < c->p.insert_stop_rate) break;
      continue;
    }
    const float rate = float(upd) / float(size_t(c->n) * size_t(c->p.k_build));
    if (rate < c->p.early_stop_rate || (it > 0 && (prev - rate) < c->p.rate_diff_tol)) break;
    prev = rate;
  }
  c->st.ms_build = c->st.ms_forest + c->st.ms_init + c->st.ms_join;
  return 0;
}

int knng_export_device(knng_ctx* c, int32_t* dk, float* dd, int32_t k, int32_t rerank) {
  if (!c || !dk || !dd) return KNNG_EINVAL;
  if (!c->have_graph) return fail(c, KNNG_ESTATE, "knng_export: no graph");
  if (k < 1 || k > c->p.k_build) return fail(c, KNNG_EINVAL, "knng_export: k must be in [1, k_build]");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t n = c->n;
  const int K = c->p.k_build;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  if (!rerank) {
    hipLaunchKernelGGL(k_export_plain, dim3(grid_for(c, n * k, 256, 16)), dim3(256), 0, c->stream,
                       c->gid[c->cur], c->gsc[c->cur], n, K, k, dk, dd, int64_t(0));
  } else {
    const int32_t* order = c->have_forest ? c->perm : nullptr;   // tree 0's leaf order
    if (int rc = launch_rerank(c, c->cur, n, 0, k, dk, dd, order)) return rc;
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);   // ms stays 0 on failure (a timing stat only)
  c->st.ms_export = ms;
  return 0;
}

int knng_export(knng_ctx* c, int32_t* knns, float* dists, int32_t k, int32_t rerank) {
  if (!c || !knns || !dists) return KNNG_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t need = c->n * k;
  if (c->ex_cap < need) {
    dfree(&c->ex_knns);
    dfree(&c->ex_dists);
    int rc;
    if ((rc = dalloc(c, &c->ex_knns, need, "export ids")) || (rc = dalloc(c, &c->ex_dists, need, "export dists")))
      return rc;
    c->ex_cap = need;
  }
  if (int rc = knng_export_device(c, c->ex_knns, c->ex_dists, k, rerank)) return rc;
  HIPCHK(c, hipMemcpyAsync(knns, c->ex_knns, need * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(dists, c->ex_dists, need * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int knng_get_stats(knng_ctx* c, knng_stats* st) {
  if (!c || !st) return KNNG_EINVAL;
  c->st.join_variant = c->join_variant;
  c->st.leaf_variant = c->leaf_variant;
  *st = c->st;
  return 0;
}

int knng_copy_quantized(knng_ctx* c, int8_t* q, uint64_t* sk) {
  if (!c) return KNNG_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  if (q) HIPCHK(c, hipMemcpyAsync(q, c->q, size_t(c->n) * c->d, hipMemcpyDeviceToHost, c->stream));
  if (sk)
    HIPCHK(c, hipMemcpyAsync(sk, c->sk, size_t(c->n) * c->words * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int knng_copy_forest(knng_ctx* c, int32_t* perm, int64_t* leaves, int64_t cap, int64_t* nleaf) {
  if (!c) return KNNG_EINVAL;
  if (!c->have_forest) return fail(c, KNNG_ESTATE, "knng_copy_forest: no forest");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t total = int64_t(c->p.num_trees) * c->n;
  if (perm) HIPCHK(c, hipMemcpyAsync(perm, c->perm, total * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  std::vector<int4> lv(c->nleaves);
  if (c->nleaves)
    HIPCHK(c, hipMemcpyAsync(lv.data(), c->leaves, c->nleaves * sizeof(int4), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::sort(lv.begin(), lv.end(), [](const int4& a, const int4& b) {
    return a.x != b.x ? a.x < b.x : a.y < b.y;
  });
  if (nleaf) *nleaf = c->nleaves;
  if (leaves) {
    if (cap < c->nleaves) return fail(c, KNNG_EINVAL, "knng_copy_forest: leaf buffer too small");
    for (int64_t l = 0; l < c->nleaves; ++l) {
      leaves[3 * l] = lv[l].x;
      leaves[3 * l + 1] = lv[l].y;
      leaves[3 * l + 2] = lv[l].z;
    }
  }
  return 0;
}

int knng_copy_cells(knng_ctx* c, int32_t part, int32_t* count, int32_t* assign, int8_t* z,
                    float* w) {
  if (!c) return KNNG_EINVAL;
  if (!c->have_forest) return fail(c, KNNG_ESTATE, "knng_copy_cells: no forest");
  if (c->nparts == 0) {
    if (count) *count = 0;
    return 0;
  }
  if (part < 0 || part >= c->nparts) return fail(c, KNNG_EINVAL, "knng_copy_cells: no such partition");
  HIPCHK(c, hipSetDevice(c->device));
  const int32_t C = c->cell_count[size_t(part)];
  if (count) *count = C;
  if (assign)
    HIPCHK(c, hipMemcpyAsync(assign, c->cassign + size_t(part) * c->n, size_t(c->n) * 4,
                             hipMemcpyDeviceToHost, c->stream));
  if (z)
    HIPCHK(c, hipMemcpyAsync(z, c->cz + size_t(part) * c->ncells * c->d, size_t(C) * c->d,
                             hipMemcpyDeviceToHost, c->stream));
  if (w)
    HIPCHK(c, hipMemcpyAsync(w, c->cw + size_t(part) * c->ncells, size_t(C) * 4, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int knng_copy_graph(knng_ctx* c, int32_t* ids, int32_t* scores) {
  if (!c) return KNNG_EINVAL;
  if (!c->have_graph) return fail(c, KNNG_ESTATE, "knng_copy_graph: no graph");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t bytes = size_t(c->n) * c->p.k_build * sizeof(int32_t);
  if (ids) HIPCHK(c, hipMemcpyAsync(ids, c->gid[c->cur], bytes, hipMemcpyDeviceToHost, c->stream));
  if (scores) HIPCHK(c, hipMemcpyAsync(scores, c->gsc[c->cur], bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int knng_load_graph(knng_ctx* c, const int32_t* ids, const int32_t* scores) {
  if (!c || !ids || !scores) return KNNG_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t n = c->n;
  const int K = c->p.k_build;
  for (int64_t j = 0; j < n * K; ++j)
    if (ids[j] < 0 || ids[j] >= n) return fail(c, KNNG_EINVAL, "knng_load_graph: id out of range");
  int rc;
  for (int b = 0; b < 2; ++b)
    if ((rc = dalloc(c, &c->gid[b], size_t(n) * K, "graph ids")) ||
        (rc = dalloc(c, &c->gsc[b], size_t(n) * K, "graph scores")) ||
        (rc = dalloc(c, &c->fresh[b], size_t(n), "fresh masks")) ||
        (rc = dalloc(c, &c->gmask[b], size_t(n) * K, "leaf mask bytes")))
      return rc;
  c->cur = 0;
  const size_t bytes = size_t(n) * K * sizeof(int32_t);
  HIPCHK(c, hipMemcpyAsync(c->gid[0], ids, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->gsc[0], scores, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->fresh[0], 0xFF, size_t(n) * sizeof(uint64_t), c->stream));   // all fresh
  c->fresh_all = true;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_graph = true;
  c->leaf_valid = false;   // a graph from elsewhere: its lists do not bound leaf-mates
  return 0;
}

int knng_quantize_device(const float* dx, int64_t n, int32_t d, int8_t* dq, uint64_t* dsk) {
  if (!dx || !dq || !dsk || n < 0 || d <= 0 || d % 64 || d > kMaxDim) return KNNG_EINVAL;
  if (n == 0) return 0;
  int64_t g = (n + 3) / 4;
  if (g > 16384) g = 16384;
  hipLaunchKernelGGL(k_quantize, dim3(g), dim3(256), 0, 0, dx, n, d, dq, dsk, static_cast<float*>(nullptr));
  if (hipGetLastError() != hipSuccess) return KNNG_EHIP;
  return hipDeviceSynchronize() == hipSuccess ? 0 : KNNG_EHIP;
}

int knng_join_iteration_noalloc(const int8_t* dq, const uint64_t* dsk, int32_t d, int64_t n,
                                int32_t k_build, int32_t rerank_mult, const int32_t* ids_in,
                                const int32_t* sc_in, int32_t* ids_out, int32_t* sc_out,
                                const int32_t* order, uint64_t* counts, void* stream) {
  if (!dq || !dsk || !ids_in || !sc_in || !ids_out || !sc_out || !counts) return KNNG_EINVAL;
  if (ids_in == ids_out || sc_in == sc_out) {
    fprintf(stderr, "[knng] knng_join_iteration_noalloc: input and output lists must differ\n");
    return KNNG_EINVAL;
  }
  if (d <= 0 || d % 64 || d > kMaxDim || k_build < 1 || k_build > kMaxK || rerank_mult < 1 || n < 2 ||
      n >= (int64_t(1) << 31) - 1) {
    fprintf(stderr,
            "[knng] knng_join_iteration_noalloc: sizes beyond its limits (d a multiple of 64 up to %d, "
            "1 <= k_build <= %d, rerank_mult >= 1, 2 <= n < 2^31 - 1)\n",
            kMaxDim, kMaxK);
    return KNNG_ERANGE;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return KNNG_EHIP;
  unsigned long long* ctr = nullptr;
  int num_cu = 0;
  if (int rc = scratch_for_device(dev, &ctr, &num_cu)) {
    fprintf(stderr, "[knng] knng_join_iteration_noalloc: no scratch address on device %d\n", dev);
    return rc;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int K = k_build, M = k_build * rerank_mult;
  const int grid = std::max(8, (num_cu * 4) / 8 * 8);   // a multiple of the 8 XCDs, as launch_join
  unsigned long long* rowq = ctr + kCounterFields * kNumCounters;
  unsigned long long* out = reinterpret_cast<unsigned long long*>(counts);
  hipLaunchKernelGGL(k_zero_u64, dim3((kScratchWords + 255) / 256), dim3(256), 0, st, ctr, kScratchWords);
#define KNNG_NOALLOC_ARGS                                                                                  \
  dq, dsk, d, n, K, M, ids_in, sc_in, ids_out, sc_out, ctr, rowq, order, static_cast<const uint64_t*>(nullptr), \
      static_cast<uint64_t*>(nullptr), static_cast<const int64_t*>(nullptr),                                   \
      static_cast<const int64_t*>(nullptr), 0, static_cast<unsigned long long*>(nullptr),                      \
      static_cast<int32_t*>(nullptr), static_cast<unsigned long long*>(nullptr),                               \
      static_cast<const uint8_t*>(nullptr), static_cast<uint8_t*>(nullptr), static_cast<const int32_t*>(nullptr), \
      0, 4, static_cast<const int32_t*>(nullptr), static_cast<const uint8_t*>(nullptr), 0
  if (K > 50 && d == kMaxDim)
    hipLaunchKernelGGL((k_join<13, 4096, 2, 2, 8, kMaxDim>), dim3(grid), dim3(256), 0, st, KNNG_NOALLOC_ARGS);
  else if (K > 50)
    hipLaunchKernelGGL((k_join<13, 4096, 2, 2, 8, 0>), dim3(grid), dim3(256), 0, st, KNNG_NOALLOC_ARGS);
  else if (d == kMaxDim)
    hipLaunchKernelGGL((k_join<12, 2560, 4, 2, 8, kMaxDim>), dim3(grid), dim3(256), 0, st, KNNG_NOALLOC_ARGS);
  else
    hipLaunchKernelGGL((k_join<12, 2560, 4, 2, 8, 0>), dim3(grid), dim3(256), 0, st, KNNG_NOALLOC_ARGS);
#undef KNNG_NOALLOC_ARGS
  // sums into the caller's counts and leaves the counters at zero
  hipLaunchKernelGGL(k_sum_counters, dim3(1), dim3(256), 0, st, ctr, out);
  return hipGetLastError() == hipSuccess ? 0 : KNNG_EHIP;
}

int knng_score_pairs_device(const int8_t* dq, const uint64_t* dsk, int32_t d, const int32_t* da,
                            const int32_t* db, int64_t np, int32_t* ddot, int32_t* dsim) {
  if (!dq || !dsk || !da || !db || !ddot || !dsim || np < 0 || d <= 0 || d % 64 || d > kMaxDim)
    return KNNG_EINVAL;
  if (np == 0) return 0;
  int64_t g = (np + 3) / 4;
  if (g > 16384) g = 16384;
  hipLaunchKernelGGL(k_score_pairs, dim3(g), dim3(256), 0, 0, dq, dsk, d, da, db, np, ddot, dsim);
  if (hipGetLastError() != hipSuccess) return KNNG_EHIP;
  return hipDeviceSynchronize() == hipSuccess ? 0 : KNNG_EHIP;
}

}  // extern "C"

namespace knng_host {

int set_segments(knng_ctx* c, const std::vector<int64_t>& seg) {
  if (!c || seg.size() < 2 || seg.front() != 0 || seg.back() != c->n) return KNNG_EINVAL;
  for (size_t s = 0; s + 1 < seg.size(); ++s)
    if (seg[s + 1] - seg[s] < 2) return fail(c, KNNG_EINVAL, "every packed build needs >= 2 rows");
  if (c->p.cells != 0) return fail(c, KNNG_EINVAL, "packed builds use plain RP trees (cells = 0)");
  HIPCHK(c, hipSetDevice(c->device));
  c->nseg = int32_t(seg.size() - 1);
  c->seg_host = seg;
  dfree(&c->seg_base);
  if (int rc = dalloc(c, &c->seg_base, seg.size(), "segment bases")) return rc;
  HIPCHK(c, hipMemcpy(c->seg_base, seg.data(), seg.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  return 0;
}

namespace {

// buffers of one packed join loop
struct PackedLoop {
  int32_t* act = nullptr;          // device: segment still iterating
  int64_t* dst = nullptr;          // device: list position of each segment
  int64_t* nwork = nullptr;        // device: list length
  int32_t* list = nullptr;         // device: rows to process this iteration
  unsigned long long* segc = nullptr;   // device: per-segment (changed, new) counts
  unsigned long long* snap = nullptr;   // device: their snapshot
  int32_t* act_h = nullptr;        // pinned host copies
  unsigned long long* snap_h = nullptr;
  unsigned long long* csum_h = nullptr;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
};

void free_loop(PackedLoop& L) {
  dfree(&L.act);
  dfree(&L.dst);
  dfree(&L.nwork);
  dfree(&L.list);
  dfree(&L.segc);
  dfree(&L.snap);
  if (L.act_h) (void)hipHostFree(L.act_h);
  if (L.snap_h) (void)hipHostFree(L.snap_h);
  if (L.csum_h) (void)hipHostFree(L.csum_h);
  for (auto& e : L.exec)
    if (e) (void)hipGraphExecDestroy(e);
}

// enqueue one join iteration a -> b of the packed build on c->stream
int enqueue_packed_iteration(knng_ctx* c, PackedLoop& L, int a, int b) {
  const int S = c->nseg;
  HIPCHK(c, hipMemcpyAsync(L.act, L.act_h, size_t(S) * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(1024), 0, c->stream, L.act, c->seg_base, S, L.dst, L.nwork);
  hipLaunchKernelGGL(k_seg_gather, dim3(grid_for(c, c->n, 256, 8)), dim3(256), 0, c->stream, c->perm, c->n,
                     c->seg_base, S, L.act, L.dst, L.list);
  if (int rc = launch_join(c, c->join_variant, a, b, L.list, L.nwork, c->seg_base, S, L.segc)) return rc;
  hipLaunchKernelGGL(k_sum_counters, dim3(1), dim3(256), 0, c->stream, c->counters, c->csum);
  hipLaunchKernelGGL(k_seg_snapshot, dim3(std::max(1, std::min(1024, (2 * S + 255) / 256))), dim3(256), 0,
                     c->stream, L.segc, L.snap, 2 * S);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(L.snap_h, L.snap, size_t(2 * S) * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(L.csum_h, c->csum, kCounterFields * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->stream));
  return 0;
}

}  // namespace

int build_packed(knng_ctx* c, std::vector<int32_t>* iters, std::vector<int32_t>* cur, bool use_graph,
                 bool* graph_used) {
  if (!c || c->nseg < 1) return KNNG_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  const int S = c->nseg;
  const int K = c->p.k_build;
  memset(c->st.updates, 0, sizeof(c->st.updates));
  c->st.iterations = 0;
  c->st.ms_join = 0;
  int rc;
  if ((rc = knng_build_forest(c))) return rc;
  if ((rc = knng_build_init(c))) return rc;
  PackedLoop L;
  auto bail = [&](int code) {
    free_loop(L);
    return code;
  };
  if ((rc = dalloc(c, &L.act, size_t(S), "segment flags")) || (rc = dalloc(c, &L.dst, size_t(S), "segment offsets")) ||
      (rc = dalloc(c, &L.nwork, 1, "work size")) || (rc = dalloc(c, &L.list, size_t(c->n), "work list")) ||
      (rc = dalloc(c, &L.segc, size_t(2 * S), "segment counters")) ||
      (rc = dalloc(c, &L.snap, size_t(2 * S), "segment counters")))
    return bail(rc);
  if (hipHostMalloc(&L.act_h, size_t(S) * sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&L.snap_h, size_t(2 * S) * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&L.csum_h, kCounterFields * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
    return bail(fail(c, KNNG_ENOMEM, "hipHostMalloc failed for the packed join loop"));
  HIPCHK(c, hipMemsetAsync(L.segc, 0, size_t(2 * S) * sizeof(unsigned long long), c->stream));
  for (int s = 0; s < S; ++s) L.act_h[s] = 1;
  iters->assign(size_t(S), 0);
  cur->assign(size_t(S), 0);
  std::vector<float> prev(size_t(S), 1.0f);
  c->cur = 0;
  // autotune on the first iteration's full work list (every row of tree 0's order)
  if ((rc = ensure_join_variant(c, 0, 1, c->perm, c->n, c->seg_base, S, L.segc))) return bail(rc);
  HIPCHK(c, hipMemsetAsync(L.segc, 0, size_t(2 * S) * sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // one graph per buffer direction
  bool graphs = use_graph;
  for (int dir = 0; dir < 2 && graphs; ++dir) {
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      graphs = false;
      break;
    }
    const int erc = enqueue_packed_iteration(c, L, dir, dir ^ 1);
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (erc || ec != hipSuccess || !g ||
        hipGraphInstantiate(&L.exec[dir], g, nullptr, nullptr, 0) != hipSuccess)
      graphs = false;
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
  }
  if (!graphs)
    for (auto& e : L.exec)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  *graph_used = graphs;
  int active = S;
  for (int it = 0; it < c->p.max_iters && active > 0; ++it) {
    const int a = c->cur, b = a ^ 1;
    int64_t rows = 0;   // rows this iteration processes
    for (int s = 0; s < S; ++s)
      if (L.act_h[s]) rows += c->seg_host[size_t(s) + 1] - c->seg_host[size_t(s)];
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (graphs) {
      HIPCHK(c, hipGraphLaunch(L.exec[a], c->stream));
    } else if ((rc = enqueue_packed_iteration(c, L, a, b))) {
      return bail(rc);
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->cur = b;
    if (it < 32) {
      c->st.updates[it] = int64_t(L.csum_h[0]);
      c->st.rates[it] = float(L.csum_h[0]) / float(size_t(c->n) * size_t(K));
      c->st.ms_join_iter[it] = ms;
      c->st.join_candidates[it] = int64_t(L.csum_h[1]);
      c->st.join_rescored[it] = int64_t(L.csum_h[2]);
      c->st.join_inserted[it] = int64_t(L.csum_h[3]);
      c->st.join_passed_on[it] = -1;
      c->st.join_leafmates[it] = int64_t(L.csum_h[5]);
      c->st.insert_rates[it] = float(L.csum_h[3]) / float(size_t(c->n) * size_t(K));
      // algorithmic bytes as knng_join_iteration counts them, over the rows processed
      const double lk_bytes =
          c->leaf_valid ? double(rows) * (2.0 * K + double(K) * K) + double(L.csum_h[3]) * 4.0 * c->leaf_tp : 0.0;
      c->st.join_bytes[it] = double(rows) * (8.0 * K + c->d + 8.0 * c->words + 4.0 * K * K + 8.0 * K) +
                             lk_bytes + double(L.csum_h[4]) * 8.0 * c->words + double(L.csum_h[2]) * c->d;
    }
    c->st.iterations = it + 1;
    c->st.ms_join += ms;
    // each segment's stop rule, exactly as knng_build applies it to a build alone
    for (int s = 0; s < S; ++s) {
      if (!L.act_h[s]) continue;
      const int64_t ns = c->seg_host[size_t(s) + 1] - c->seg_host[size_t(s)];
      const int j = (*iters)[size_t(s)];
      (*iters)[size_t(s)] = j + 1;
      (*cur)[size_t(s)] = b;
      bool stop = false;
      if (c->p.stop_rule == KNNG_STOP_INSERTED) {
        stop = float(L.snap_h[2 * s + 1]) / float(size_t(ns) * size_t(K)) < c->p.insert_stop_rate;
      } else if (c->p.stop_rule == KNNG_STOP_REFERENCE) {
        const float rate = float(L.snap_h[2 * s]) / float(size_t(ns) * size_t(K));
        stop = rate < c->p.early_stop_rate || (j > 0 && (prev[size_t(s)] - rate) < c->p.rate_diff_tol);
        prev[size_t(s)] = rate;
      }
      if (stop) {
        L.act_h[s] = 0;
        --active;
      }
    }
  }
  c->st.ms_build = c->st.ms_forest + c->st.ms_init + c->st.ms_join;
  free_loop(L);
  return 0;
}

int export_segment(knng_ctx* c, int s, int cur, int k, int rerank, int32_t* dk, float* dd) {
  if (!c || s < 0 || s >= c->nseg || k < 1 || k > c->p.k_build) return KNNG_EINVAL;
  const int64_t base = c->seg_host[size_t(s)], ns = c->seg_host[size_t(s) + 1] - base;
  const int K = c->p.k_build;
  if (!rerank) {
    hipLaunchKernelGGL(k_export_plain, dim3(grid_for(c, ns * k, 256, 16)), dim3(256), 0, c->stream,
                       c->gid[cur], c->gsc[cur], ns, K, k, dk, dd, base);
  } else {
    if (int rc = launch_rerank(c, cur, ns, base, k, dk, dd, nullptr)) return rc;
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

}  // namespace knng_host

