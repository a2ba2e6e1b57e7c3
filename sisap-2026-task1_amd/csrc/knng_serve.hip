// knng_serve.hip -- device-resident kNN serving on a built graph (the
// knng_server_* entry points of include/knng.h).  The reference engine has
// no query path (src/main.cpp:479 writes querytime = 0); this is new.
//
// Index in HBM: int8 rows + sketches + float32 rows + the K-list graph, with
// room for insert_capacity more rows.  Search (k_search, one 256-thread block
// per query):
//   1. quantise the query with K1 (same int8 rules as the build);
//   2. descend `entry_trees` RP trees of the build's forest (same two-pivot
//      rule) and take the leaves' members as entry candidates;
//   3. beam search: the best unexpanded beam entries (4 per step, one per
//      wave) are expanded, their K neighbours filtered through an LDS visited
//      set, scored with int8 dots (16 lanes per 1 KB row gather) and merged
//      into the beam by rank in (score desc, id asc) order;
//   4. optional FP32 re-score of the final beam, best k out.
// Results depend only on the index contents, never on what else is running.
//
// Scheduler (one host thread): requests are split into max_batch pieces;
// each piece runs on a "slot" (HIP stream + buffers).  Bulk slots are created
// on demand while fewer than max_slots exist and hipMemGetInfo leaves
// reserve_bytes free; pieces of different requests overlap on the device.
// Query requests of at most express_rows rows go to a second queue served by
// express slots, created with the server; with express_cus > 0 their streams
// carry a CU mask of express_cus CUs spread over the chip and the bulk
// streams carry the complement, so a small request's blocks start at once
// instead of waiting for bulk blocks to drain from the CUs.
// Inserts append rows, search their neighbours and merge reverse edges into
// the neighbours' lists.  By default they run on an insert thread and stream
// of their own beside the searches (pending inserts are applied together, up
// to the insert scratch size): a query sees the inserts finished before it
// was dispatched (knng_server_wait_info reports that index size), and one
// that overlaps an insert may or may not see its rows -- every list word is
// an aligned 4-byte store, so a search reads old or new ids, always valid
// rows.  With strict_order = 1 an insert instead waits until the queries
// submitted before it have finished and a query runs only after the inserts
// submitted before it (its epoch), whichever lane it takes.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "knng.h"
#include "knng_device.h"
#include "knng_internal.h"

using namespace knng_dev;

namespace {

constexpr int kBeamMax = 64;
constexpr int kExpand = 4;        // beam entries expanded per step (one per wave)
constexpr int kCandCap = 256;     // >= kExpand * kMaxK and >= 4 entry leaves of 64
constexpr int kVisitBits = 12;    // visited set: 4096 slots, reset when 3/4 full
constexpr int kExpressRouteCUs = 16;   // express routing: ~32 blocks per partition

__device__ __forceinline__ bool ahead(int32_t sa, int32_t ia, int32_t sb, int32_t ib) {
  return sa != sb ? sa > sb : ia < ib;   // (score desc, id asc)
}

// Cell routing of a small query batch: each query's best cell in every
// partition, argmax_c float(<q, z_c>) * w_c with ties to the smaller c -- the
// same values and the same order as k_assign / k_assign_wide, so routing (and
// every search result) does not depend on which kernel ran.  One wave per
// block (64 threads, 16 KB of LDS, few VGPRs), so its blocks start on CUs that
// resident bulk search blocks already occupy instead of waiting for them to
// drain (the two-block-per-CU k_assign_wide cannot fit beside them), and all
// partitions go in one launch.  Block (s, qg, g): queries [16 qg, 16 qg + 16)
// (staged in LDS) against the 16-cell tiles [8 s, 8 s + 8) of partition g,
// with v_mfma_i32_16x16x64_i8: A = 16 queries, B = 16 cells, lane l holding
// row / cell column l % 16 and the same 16-dim quarter l / 16 of every 64-dim
// step in both operands, so the exact int32 dot is summed over all dims; the
// accumulator holds queries 4 (l / 16) + j of cell column l % 16.  The last
// block of a (g, qg) to finish (an agent-scope counter, release / acquire
// fences around it) folds the slices' partial bests into out[g * nq + q] and
// re-arms the counter for the slot's next launch.
constexpr int kRouteQ = 16;       // queries per block
constexpr int kRouteTiles = 8;    // 16-cell tiles per block
constexpr int kRouteMaxQ = 256;   // queries per k_route_small launch
constexpr int kRouteGroups = kRouteMaxQ / kRouteQ;

__global__ __launch_bounds__(64) void k_route_small(const int8_t* __restrict__ q, int nq, int d,
                                                    const int8_t* __restrict__ Z, const float* __restrict__ W,
                                                    int64_t ncells, const int32_t* __restrict__ cells_of,
                                                    float* __restrict__ pbest, int32_t* __restrict__ pidx,
                                                    int32_t* __restrict__ ctr, int32_t* __restrict__ out,
                                                    int express) {
  __shared__ __attribute__((aligned(16))) int8_t qs[kRouteQ * kMaxDim];
  // an express launch's waves take issue priority over resident bulk waves
  if (express) __builtin_amdgcn_s_setprio(2);
  const int lane = threadIdx.x;
  const int s = blockIdx.x, qg = blockIdx.y, g = blockIdx.z, smax = gridDim.x;
  const int C = cells_of[g];
  const int tiles = (C + 15) >> 4;
  const int S = max(1, (tiles + kRouteTiles - 1) / kRouteTiles);
  if (s >= S) return;
  const int q0 = qg * kRouteQ, ks = d >> 6, cpr = d >> 4;
  for (int e = lane; e < kRouteQ * cpr; e += 64) {
    const int r = e / cpr, c = e - r * cpr;
    const v4i v = q0 + r < nq ? *reinterpret_cast<const v4i*>(q + int64_t(q0 + r) * d + 16 * c)
                              : v4i{0, 0, 0, 0};
    *reinterpret_cast<v4i*>(qs + r * d + 16 * c) = v;
  }
  __syncthreads();
  const int col = lane & 15, kq = lane >> 4;
  const int8_t* zg = Z + int64_t(g) * ncells * d;
  const float* wg = W + int64_t(g) * ncells;
  float best[4];
  int32_t bidx[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    best[j] = -INFINITY;
    bidx[j] = 0x7FFFFFFF;
  }
  const int t_hi = min(tiles, (s + 1) * kRouteTiles);
  for (int t = s * kRouteTiles; t < t_hi; ++t) {
    const int cell = 16 * t + col;
    const bool ok = cell < C;
    const int8_t* zr = zg + int64_t(ok ? cell : 0) * d + 16 * kq;
    const float wc = ok ? wg[cell] : 0.f;
    v4i acc = {0, 0, 0, 0};
    // the cell rows in two halves of 8 steps (32 VGPRs of loads in flight),
    // so the block stays small enough to sit beside 4 resident search waves
#pragma unroll
    for (int h = 0; h < kMaxDim / 512; ++h) {
      v4i b[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        b[k] = 8 * h + k < ks ? *reinterpret_cast<const v4i*>(zr + 64 * (8 * h + k)) : v4i{0, 0, 0, 0};
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (8 * h + k < ks)
          acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(
              *reinterpret_cast<const v4i*>(qs + col * d + 64 * (8 * h + k) + 16 * kq), b[k], acc, 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float sc = ok ? float(acc[j]) * wc : -INFINITY;
      const bool take = sc > best[j];   // cells arrive in increasing order: ties keep the smaller
      best[j] = take ? sc : best[j];
      bidx[j] = take ? cell : bidx[j];
    }
  }
  const int64_t pbase = (int64_t(g) * smax + s) * nq;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) {   // over the 16 lanes of this quarter
      const float ov = __shfl_xor(best[j], m);
      const int32_t oi = __shfl_xor(bidx[j], m);
      if (ov > best[j] || (ov == best[j] && oi < bidx[j])) {
        best[j] = ov;
        bidx[j] = oi;
      }
    }
    const int qq = q0 + 4 * kq + j;
    if (col == 0 && qq < nq) {
      pbest[pbase + qq] = best[j];
      pidx[pbase + qq] = bidx[j];
    }
  }
  __threadfence();   // release this block's partials
  int prev = 0;
  if (lane == 0) prev = atomicAdd(&ctr[g * kRouteGroups + qg], 1);
  prev = __shfl(prev, 0);
  if (prev != S - 1) return;
  __threadfence();   // acquire the other slices' partials
  if (lane < kRouteQ && q0 + lane < nq) {
    const int qq = q0 + lane;
    const int64_t base = int64_t(g) * smax * nq + qq;
    float bv = pbest[base];
    int32_t bi = pidx[base];
    for (int y = 1; y < S; ++y) {
      const float v = pbest[base + int64_t(y) * nq];
      const int32_t vi = pidx[base + int64_t(y) * nq];
      if (v > bv || (v == bv && vi < bi)) {
        bv = v;
        bi = vi;
      }
    }
    out[int64_t(g) * nq + qq] = bi;
  }
  if (lane == 0) ctr[g * kRouteGroups + qg] = 0;   // re-armed for the slot's next launch
}

__global__ __launch_bounds__(256, 6) void k_search(
    const int8_t* __restrict__ q, const float* __restrict__ x, int d,
    const int32_t* __restrict__ gid, int K, const Node* __restrict__ hier, int64_t hier_count,
    const int4* __restrict__ leaves, const int32_t* __restrict__ perm, int64_t n_forest,
    const int32_t* __restrict__ cell_root, const int32_t* __restrict__ qcell, int ncells,
    int nparts, int entry_trees, const int8_t* __restrict__ qq, const float* __restrict__ qx,
    int nq, int beam, int k, int rerank, int max_steps, int32_t* __restrict__ out_ids,
    float* __restrict__ out_sc, int express) {
  // express launches raise their waves' issue priority, so on a SIMD shared
  // with bulk search waves theirs issue first (results are unaffected)
  if (express) __builtin_amdgcn_s_setprio(2);
  constexpr int HS = 1 << kVisitBits;
  __shared__ int32_t hs[HS];
  __shared__ __attribute__((aligned(16))) int8_t qrow[kMaxDim];
  __shared__ int32_t cid[kCandCap], csc[kCandCap], surv[kCandCap];
  __shared__ int32_t bid[kBeamMax], bsc[kBeamMax], bexp[kBeamMax];
  __shared__ int32_t nid[kBeamMax], nsc[kBeamMax], nexp[kBeamMax];
  __shared__ float fsc[kBeamMax];
  __shared__ int32_t sel[kExpand];
  __shared__ int32_t cnt[8];   // 0 candidates, 1 beam size, 2 visited used, 3 selected, 4 survivors
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int g = lane >> 4, s16 = lane & 15;
  const int nch = d >> 4;

  for (int qi = blockIdx.x; qi < nq; qi += gridDim.x) {
    for (int j = tid; j < HS; j += 256) hs[j] = -1;
    if (tid < 8) cnt[tid] = 0;
    for (int c = tid; c < nch; c += 256)
      *reinterpret_cast<v4i*>(qrow + 16 * c) = *reinterpret_cast<const v4i*>(qq + int64_t(qi) * d + 16 * c);
    __syncthreads();
    // the query's own chunks are read from LDS at each use: held in registers
    // they would cost 16 VGPRs for the whole query
    auto own = [&](int c) -> v4i {
      const int chn = s16 + 16 * c;
      return chn < nch ? *reinterpret_cast<const v4i*>(qrow + 16 * chn) : v4i{0, 0, 0, 0};
    };

    // ---- entry points: descend trees w, w+4, ... to a leaf; with cells, each
    // tree starts at the subtree of the query's cell in that tree's partition
    for (int t = w; t < entry_trees; t += 4) {
      int leaf_idx = -1;
      int node = t;
      bool skip = false;
      if (ncells > 0) {
        const int32_t r = cell_root[int64_t(t) * ncells + qcell[int64_t(t % nparts) * nq + qi]];
        if (r == INT32_MIN)
          skip = true;   // empty cell
        else if (r < 0)
          leaf_idx = -r - 1;
        else
          node = r;
      } else if (hier_count == 0) {
        leaf_idx = t;   // every tree is a single leaf
      }
      if (!skip && leaf_idx < 0) {
        for (int depth = 0; depth < 4096 && leaf_idx < 0; ++depth) {
          const Node nd = hier[node];
          const int8_t* ra = q + int64_t(nd.p1id) * d;
          const int8_t* rb = q + int64_t(nd.p2id) * d;
          int s = 0;
          for (int c = lane; c < nch; c += 64) {
            const v4i v = *reinterpret_cast<const v4i*>(qrow + 16 * c);
            s = dot16_i8(v, *reinterpret_cast<const v4i*>(ra + 16 * c), s);
            s -= dot16_i8(v, *reinterpret_cast<const v4i*>(rb + 16 * c), 0);
          }
          s = wave_sum_i(s);
          const int child = s > 0 ? nd.left : nd.right;
          if (child < 0)
            leaf_idx = -child - 1;
          else
            node = child;
        }
      }
      if (leaf_idx >= 0) {
        const int4 lf = leaves[leaf_idx];
        const int m = lf.z < kLeafSlots ? lf.z : kLeafSlots;   // a big leaf enters its first 64 rows
        for (int j = lane; j < m; j += 64) {
          const int32_t id = perm[int64_t(lf.x) * n_forest + lf.y + j];
          if (hs_insert(hs, kVisitBits, id)) {
            atomicAdd(&cnt[2], 1);
            const int slot = atomicAdd(&cnt[0], 1);
            if (slot < kCandCap) cid[slot] = id;
          }
        }
      }
    }
    __syncthreads();

    for (int step = 0;; ++step) {
      // ---- score the candidates: 16 lanes per 1 KB row, 2 rows per lane group
      const int nc = cnt[0] < kCandCap ? cnt[0] : kCandCap;
      for (int s0 = 0; s0 < nc; s0 += 32) {
        const int sa = s0 + w * 4 + g, sb = sa + 16;
        const int8_t* ra = sa < nc ? q + int64_t(cid[sa]) * d : nullptr;
        const int8_t* rb = sb < nc ? q + int64_t(cid[sb]) * d : nullptr;
        int da = 0, db = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int chn = s16 + 16 * c;
          if (chn < nch) {
            const v4i oc = own(c);
            if (ra) da = dot16_i8(*reinterpret_cast<const v4i*>(ra + 16 * chn), oc, da);
            if (rb) db = dot16_i8(*reinterpret_cast<const v4i*>(rb + 16 * chn), oc, db);
          }
        }
#pragma unroll
        for (int m = 8; m >= 1; m >>= 1) {
          da += __shfl_xor(da, m);
          db += __shfl_xor(db, m);
        }
        if (s16 == 0) {
          if (sa < nc) csc[sa] = da;
          if (sb < nc) csc[sb] = db;
        }
      }
      if (tid == 0) cnt[4] = 0;
      __syncthreads();
      // ---- merge candidates into the beam by rank
      const int nb = cnt[1];
      for (int u = tid; u < nc; u += 256) {
        int r = 0;
        for (int j = 0; j < nb; ++j) r += ahead(bsc[j], bid[j], csc[u], cid[u]) ? 1 : 0;
        if (r < beam) surv[atomicAdd(&cnt[4], 1)] = u;
      }
      __syncthreads();
      const int P = cnt[4];
      for (int e = tid; e < nb + P; e += 256) {
        int32_t sc, id, ex;
        int r;
        if (e < nb) {
          sc = bsc[e];
          id = bid[e];
          ex = bexp[e];
          r = e;
        } else {
          const int u = surv[e - nb];
          sc = csc[u];
          id = cid[u];
          ex = 0;
          r = 0;
          for (int j = 0; j < nb; ++j) r += ahead(bsc[j], bid[j], sc, id) ? 1 : 0;
        }
        for (int p = 0; p < P && r < beam; ++p) {
          const int u = surv[p];
          r += ahead(csc[u], cid[u], sc, id) ? 1 : 0;
        }
        if (r < beam) {
          nid[r] = id;
          nsc[r] = sc;
          nexp[r] = ex;
        }
      }
      __syncthreads();
      const int nnew = (nb + P) < beam ? (nb + P) : beam;
      if (tid < nnew) {
        bid[tid] = nid[tid];
        bsc[tid] = nsc[tid];
        bexp[tid] = nexp[tid];
      }
      __syncthreads();
      // ---- pick the best unexpanded entries (wave 0)
      if (w == 0) {
        unsigned long long m = __ballot(lane < nnew && !bexp[lane]);
        int ns = 0;
        for (int e = 0; e < kExpand && m; ++e) {
          const int j = __ffsll(m) - 1;
          m &= m - 1;
          if (lane == 0) {
            sel[e] = bid[j];
            bexp[j] = 1;
          }
          ++ns;
        }
        if (lane == 0) {
          cnt[1] = nnew;
          cnt[3] = (step < max_steps) ? ns : 0;
          cnt[0] = 0;
        }
      }
      __syncthreads();
      const int ns = cnt[3];
      if (ns == 0) break;
      // ---- keep the visited set below 3/4 full: reset it to the beam
      if (cnt[2] + kExpand * K > (HS * 3) / 4) {
        for (int j = tid; j < HS; j += 256) hs[j] = -1;
        __syncthreads();
        if (tid < nnew) hs_insert(hs, kVisitBits, bid[tid]);
        if (tid == 0) cnt[2] = nnew;
        __syncthreads();
      }
      // ---- expand: wave w takes selected entry w
      if (w < ns && lane < K) {
        const int32_t id = gid[int64_t(sel[w]) * K + lane];
        if (hs_insert(hs, kVisitBits, id)) {
          atomicAdd(&cnt[2], 1);
          cid[atomicAdd(&cnt[0], 1)] = id;
        }
      }
      __syncthreads();
    }

    // ---- output
    const int nb = cnt[1];
    if (rerank) {
      float4 ox[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int chn = lane + 64 * c;
        ox[c] = chn < (d >> 2) ? reinterpret_cast<const float4*>(qx + int64_t(qi) * d)[chn]
                               : make_float4(0.f, 0.f, 0.f, 0.f);
      }
      for (int j = w; j < nb; j += 4) {
        const float4* row = reinterpret_cast<const float4*>(x + int64_t(bid[j]) * d);
        float s = 0.f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int chn = lane + 64 * c;
          if (chn < (d >> 2)) {
            const float4 v = row[chn];
            s = fmaf(ox[c].x, v.x, s);
            s = fmaf(ox[c].y, v.y, s);
            s = fmaf(ox[c].z, v.z, s);
            s = fmaf(ox[c].w, v.w, s);
          }
        }
        s = wave_sum_f(s);
        if (lane == 0) fsc[j] = s;
      }
      __syncthreads();
      if (tid < nb) {
        const float s = fsc[tid];
        const int32_t id = bid[tid];
        int r = 0;
        for (int j = 0; j < nb; ++j) r += (fsc[j] > s || (fsc[j] == s && bid[j] < id)) ? 1 : 0;
        if (r < k) {
          out_ids[int64_t(qi) * k + r] = id;
          out_sc[int64_t(qi) * k + r] = s;
        }
      }
    } else if (tid < nb && tid < k) {
      out_ids[int64_t(qi) * k + tid] = bid[tid];
      out_sc[int64_t(qi) * k + tid] = float(bsc[tid]);
    }
    if (tid >= nb && tid < k) {
      out_ids[int64_t(qi) * k + tid] = -1;
      out_sc[int64_t(qi) * k + tid] = -INFINITY;
    }
    __syncthreads();
  }
}

// Merge reverse-edge proposals (new row p with score s) into target rows'
// lists; one block per target.  The K kept entries are the forced proposals
// first (a new row always enters the lists of its nearest neighbours, so a
// search that reaches them can reach it), then the best of the rest; the
// kept entries are stored in (score desc, id asc) order.
__global__ __launch_bounds__(64) void k_merge_reverse(int32_t* __restrict__ gid,
                                                      int32_t* __restrict__ gsc, int K,
                                                      const int32_t* __restrict__ targets,
                                                      const int32_t* __restrict__ offs,
                                                      const int32_t* __restrict__ pid,
                                                      const int32_t* __restrict__ psc,
                                                      const int32_t* __restrict__ pforced,
                                                      int ntargets) {
  __shared__ int32_t li[2 * kMaxK], ls[2 * kMaxK], lf[2 * kMaxK], keep[2 * kMaxK];
  __shared__ int32_t oi[kMaxK], os[kMaxK];
  const int lane = threadIdx.x;
  for (int t = blockIdx.x; t < ntargets; t += gridDim.x) {
    const int64_t row = targets[t];
    const int p0 = offs[t], np = offs[t + 1] - offs[t];
    const int m = np < kMaxK ? np : kMaxK;
    if (lane < K) {
      li[lane] = gid[row * K + lane];
      ls[lane] = gsc[row * K + lane];
      lf[lane] = 0;
    }
    if (lane < m) {
      li[K + lane] = pid[p0 + lane];
      ls[K + lane] = psc[p0 + lane];
      lf[K + lane] = pforced[p0 + lane];
    }
    __syncthreads();
    auto first = [&](int a, int b, bool use_forced) {   // a ahead of b
      if (use_forced && lf[a] != lf[b]) return lf[a] > lf[b];
      if (ls[a] != ls[b]) return ls[a] > ls[b];
      if (li[a] != li[b]) return li[a] < li[b];
      return a < b;
    };
    for (int e = lane; e < K + m; e += 64) {   // which K survive
      int r = 0;
      for (int j = 0; j < K + m; ++j) r += first(j, e, true) ? 1 : 0;
      keep[e] = r < K;
    }
    __syncthreads();
    for (int e = lane; e < K + m; e += 64) {   // their order
      if (!keep[e]) continue;
      int r = 0;
      for (int j = 0; j < K + m; ++j) r += (keep[j] && first(j, e, false)) ? 1 : 0;
      oi[r] = li[e];
      os[r] = ls[e];
    }
    __syncthreads();
    if (lane < K) {
      gid[row * K + lane] = oi[lane];
      gsc[row * K + lane] = os[lane];
    }
    __syncthreads();
  }
}

using clk = std::chrono::steady_clock;

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  float* dx = nullptr;
  int8_t* dq = nullptr;
  uint64_t* dsk = nullptr;
  int32_t* did = nullptr;
  float* dsc = nullptr;
  int32_t* dcell = nullptr;   // query cells, one row of max_batch per partition
  float* pbest = nullptr;     // split cell-routing partials
  int32_t* pidx = nullptr;
  int64_t pcap = 0;
  float* rbest = nullptr;     // k_route_small partials [nparts][route_smax][kRouteMaxQ]
  int32_t* ridx = nullptr;
  int32_t* rctr = nullptr;    // its last-block counters [nparts][kRouteGroups], zero between launches
  float* hx = nullptr;        // pinned staging
  int32_t* hid = nullptr;
  float* hsc = nullptr;
  bool busy = false;
  bool express = false;
  // the rows of this launch: part j is rows [row0, row0 + nrows) of request
  // `ticket`, staged at launch row `off` (bulk: one part; express: every
  // small request that was waiting, packed into one launch)
  struct Part {
    int64_t ticket;
    int32_t row0, nrows, off;
  };
  std::vector<Part> parts;
};

struct Request {
  int64_t ticket = 0;
  int kind = 0;
  int32_t nrows = 0;
  std::vector<float> rows;
  std::vector<int32_t> ids;
  std::vector<float> scores;
  int32_t next_row = 0;
  int32_t running = 0;
  int state = 0;   // 0 queued, 1 running, 2 done
  int error = 0;
  int64_t epoch = 0;   // inserts submitted before this request
  int64_t rows_seen = -1;   // index rows when the first piece was dispatched
  int64_t first_id = 0;     // inserts: id of the first new row
  clk::time_point t_submit, t_done;
};

}  // namespace

struct knng_server {
  knng_ctx* ctx = nullptr;
  knng_server_params sp{};
  int device = 0, num_cu = 256, d = 0, words = 0, K = 0;
  std::atomic<int64_t> n{0};   // rows in the index (grows with inserts)
  int64_t cap = 0, n_forest = 0;
  int8_t* q = nullptr;
  uint64_t* sk = nullptr;
  float* x = nullptr;
  int32_t* gid = nullptr;
  int32_t* gsc = nullptr;
  hipStream_t ins_stream = nullptr;
  std::vector<Slot*> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request*> queue;     // bulk queries and inserts, in submission order
  std::deque<Request*> express;   // query requests of at most express_rows rows
  int64_t inserts_submitted = 0, inserts_done = 0;
  bool cu_masked = false;         // the express / bulk CU split is in force
  int bulk_blocks = 0;            // search blocks per bulk launch
  int32_t* route_cells = nullptr; // cells of each partition (device copy of ctx->cell_count)
  int route_smax = 0;             // k_route_small slices per partition (16-cell tiles / 8)
  std::string mask_err;           // why the runtime refused the CU masks, if it did
  // concurrent inserts: queue, thread and preallocated scratch for ins_cap rows
  std::deque<Request*> ins_queue;
  std::thread ins_th;
  int32_t ins_cap = 0;
  int32_t *ins_did = nullptr, *ins_dcell = nullptr, *ins_pidx = nullptr;
  float *ins_dsc = nullptr, *ins_pbest = nullptr;
  int32_t *ins_t = nullptr, *ins_off = nullptr, *ins_pi = nullptr, *ins_ps = nullptr, *ins_pf = nullptr;
  std::map<int64_t, Request*> reqs;
  int64_t next_ticket = 1;
  bool stop = false;
  std::thread th, th_express;   // bulk-lane and express-lane schedulers
  std::string err;
};

namespace {

int sfail(knng_server* s, int code, const std::string& m) {
  if (s) s->err = m;
  return code;
}

#define SCHK(srv, expr)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return sfail(srv, KNNG_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

size_t slot_device_bytes(const knng_server* s) {
  const size_t b = size_t(s->sp.max_batch);
  return b * (size_t(s->d) * 5 + size_t(s->words) * 8 + size_t(s->sp.k) * 8) + 4096;
}

void free_slot(Slot* sl) {
  if (!sl) return;
  if (sl->stream) (void)hipStreamSynchronize(sl->stream);
  for (void* p : {static_cast<void*>(sl->dx), static_cast<void*>(sl->dq),
                  static_cast<void*>(sl->dsk), static_cast<void*>(sl->did),
                  static_cast<void*>(sl->dsc), static_cast<void*>(sl->dcell),
                  static_cast<void*>(sl->pbest), static_cast<void*>(sl->pidx),
                  static_cast<void*>(sl->rbest), static_cast<void*>(sl->ridx), static_cast<void*>(sl->rctr)})
    if (p) (void)hipFree(p);
  for (void* p : {static_cast<void*>(sl->hx), static_cast<void*>(sl->hid), static_cast<void*>(sl->hsc)})
    if (p) (void)hipHostFree(p);
  if (sl->done) (void)hipEventDestroy(sl->done);
  if (sl->stream) (void)hipStreamDestroy(sl->stream);
  delete sl;
}

// CU mask of the express streams (express = true) or of the bulk streams:
// express_cus CUs at even spacing over the chip (so each XCD gives up a few)
std::vector<uint32_t> cu_mask(const knng_server* s, bool express) {
  const int ncu = s->num_cu;
  std::vector<uint32_t> m(size_t((ncu + 31) / 32), 0u);
  const int e = std::min(s->sp.express_cus, ncu - 1);
  std::vector<char> res(size_t(ncu), 0);
  for (int j = 0; j < e; ++j) res[size_t(int64_t(j) * ncu / e)] = 1;
  for (int c = 0; c < ncu; ++c)
    if (bool(res[size_t(c)]) == express) m[size_t(c / 32)] |= 1u << (c % 32);
  return m;
}

// A slot stream.  Where the runtime refuses CU masks (checked on the first
// express stream, before any bulk stream exists) the server runs unmasked.
hipError_t make_stream(knng_server* s, bool express, hipStream_t* st) {
  if (s->cu_masked) {
    const std::vector<uint32_t> m = cu_mask(s, express);
    const hipError_t e = hipExtStreamCreateWithCUMask(st, uint32_t(m.size()), m.data());
    if (e == hipSuccess || s->slots.size() > 0) return e;
    (void)hipGetLastError();
    s->cu_masked = false;
    s->mask_err = hipGetErrorString(e);
  }
  return hipStreamCreateWithFlags(st, hipStreamNonBlocking);
}

// k_route_small buffers of a slot: partials and counters for every partition
size_t route_entries(const knng_server* s) {
  return size_t(std::max(1, s->ctx->nparts)) * size_t(std::max(1, s->route_smax)) * kRouteMaxQ;
}
size_t route_counters(const knng_server* s) { return size_t(std::max(1, s->ctx->nparts)) * kRouteGroups; }

Slot* new_slot(knng_server* s, bool express) {
  auto* sl = new Slot();
  sl->express = express;
  const size_t B = size_t(s->sp.max_batch);
  bool ok = make_stream(s, express, &sl->stream) == hipSuccess &&
            hipEventCreateWithFlags(&sl->done, hipEventDisableTiming) == hipSuccess &&
            hipMalloc(&sl->dx, B * s->d * sizeof(float)) == hipSuccess &&
            hipMalloc(&sl->dq, B * s->d) == hipSuccess &&
            hipMalloc(&sl->dsk, B * s->words * sizeof(uint64_t)) == hipSuccess &&
            hipMalloc(&sl->did, B * s->sp.k * sizeof(int32_t)) == hipSuccess &&
            hipMalloc(&sl->dsc, B * s->sp.k * sizeof(float)) == hipSuccess &&
            hipMalloc(&sl->dcell, B * std::max(1, s->ctx->nparts) * sizeof(int32_t)) == hipSuccess &&
            hipMalloc(&sl->pbest, size_t(knng_host::assign_part_entries(int64_t(B), s->num_cu)) *
                                      sizeof(float)) == hipSuccess &&
            hipMalloc(&sl->pidx, size_t(knng_host::assign_part_entries(int64_t(B), s->num_cu)) *
                                     sizeof(int32_t)) == hipSuccess &&
            hipMalloc(&sl->rbest, route_entries(s) * sizeof(float)) == hipSuccess &&
            hipMalloc(&sl->ridx, route_entries(s) * sizeof(int32_t)) == hipSuccess &&
            hipMalloc(&sl->rctr, route_counters(s) * sizeof(int32_t)) == hipSuccess &&
            hipMemset(sl->rctr, 0, route_counters(s) * sizeof(int32_t)) == hipSuccess &&
            hipHostMalloc(&sl->hx, B * s->d * sizeof(float), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc(&sl->hid, B * s->sp.k * sizeof(int32_t), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc(&sl->hsc, B * s->sp.k * sizeof(float), hipHostMallocDefault) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    free_slot(sl);
    return nullptr;
  }
  sl->pcap = knng_host::assign_part_entries(int64_t(B), s->num_cu);
  return sl;
}

// queue one search launch on `stream`: device queries dx (float32) -> ids/scores
// dcell: nparts * nq int32 scratch for the query's cell in each partition;
// pbest/pidx: pcap-entry partials that let the routing of a few queries use
// many blocks (cells split into slices)
// max_blocks caps the search grid (blocks then take several queries each);
// route_cu sizes the cell routing as if the chip had that many CUs (two
// blocks per CU per partition), so a small launch's routing fits the block
// slots that bulk launches leave free.
int launch_search(knng_server* s, hipStream_t stream, const float* dx, int8_t* dq, uint64_t* dsk,
                  int32_t* dcell, float* pbest, int32_t* pidx, int64_t pcap, int32_t nq, int beam,
                  int k, int rerank, int32_t* did, float* dsc, int max_blocks, int route_cu,
                  const Slot* sl, const float* hsrc = nullptr) {
  knng_ctx* c = s->ctx;
  // hsrc: the queries in pinned host memory, quantised from there and copied
  // to dx by the same kernel (no separate upload)
  if (knng_host::launch_quantize(stream, hsrc ? hsrc : dx, nq, s->d, dq, dsk, s->num_cu,
                                 hsrc ? const_cast<float*>(dx) : nullptr))
    return sfail(s, KNNG_EHIP, "quantise launch failed");
  // route the queries to cells: small batches (a slot's launch of at most
  // kRouteMaxQ queries) in one k_route_small launch over all partitions, the
  // rest through k_assign per partition
  if (c->nparts > 0 && sl && nq <= kRouteMaxQ) {
    hipLaunchKernelGGL(k_route_small, dim3(s->route_smax, (nq + kRouteQ - 1) / kRouteQ, c->nparts), dim3(64),
                       0, stream, dq, nq, s->d, c->cz, c->cw, int64_t(c->ncells), s->route_cells, sl->rbest,
                       sl->ridx, sl->rctr, dcell, sl->express ? 1 : 0);
    SCHK(s, hipGetLastError());
  } else {
    for (int g = 0; g < c->nparts; ++g)
      if (knng_host::launch_assign(stream, dq, nq, s->d, c->cz + size_t(g) * c->ncells * s->d,
                                   c->cw + size_t(g) * c->ncells, c->cell_count[size_t(g)],
                                   dcell + size_t(g) * nq, route_cu, 0, 1, false, nullptr,
                                   pbest, pidx, pcap))
        return sfail(s, KNNG_EHIP, "cell routing launch failed");
  }
  const int grid = std::max(1, std::min<int>(nq, max_blocks));
  // at most 4 trees: their leaves (<= 64 rows each) fill the candidate buffer
  const int et = std::min(std::min<int>(s->sp.entry_trees, c->p.num_trees), kCandCap / 64);
  hipLaunchKernelGGL(k_search, dim3(grid), dim3(256), 0, stream, s->q, s->x, s->d, s->gid, s->K,
                     c->hier, c->hier_count, c->leaves, c->perm, s->n_forest, c->cell_root, dcell,
                     c->ncells, std::max(1, c->nparts), et, dq, dx, nq, beam, k, rerank,
                     s->sp.max_steps, did, dsc, (sl && sl->express) ? 1 : 0);
  SCHK(s, hipGetLastError());
  return 0;
}

// Append m rows (host float32) to the index as ids base .. base+m-1, base =
// the current size: copy and quantise them, find their K nearest rows with
// the search kernel (int8 scores, no rerank), write their lists, and merge
// reverse edges into those neighbours' lists.  Runs on ins_stream with the
// preallocated scratch (m <= ins_cap); only one insert runs at a time.
int do_insert_rows(knng_server* s, const float* rows, int32_t m, int64_t* first_id) {
  const int K = s->K;
  const int64_t base = s->n.load();
  if (m > s->ins_cap) return sfail(s, KNNG_ERANGE, "insert batch exceeds the insert scratch");
  if (base + m > s->cap) return sfail(s, KNNG_ERANGE, "insert exceeds insert_capacity");
  hipStream_t st = s->ins_stream;
  SCHK(s, hipMemcpyAsync(s->x + base * s->d, rows, size_t(m) * s->d * sizeof(float),
                         hipMemcpyHostToDevice, st));
  // neighbours of the new rows among the current ones (int8 scores, no rerank)
  const int beam = std::max(s->sp.beam, K);
  int rc = launch_search(s, st, s->x + base * s->d, s->q + base * s->d, s->sk + base * s->words,
                         s->ins_dcell, s->ins_pbest, s->ins_pidx,
                         knng_host::assign_part_entries(s->ins_cap, s->num_cu), m,
                         std::min(beam, kBeamMax), K, 0, s->ins_did, s->ins_dsc, s->num_cu * 8, s->num_cu,
                         nullptr);
  if (rc) return rc;
  std::vector<int32_t> ids(size_t(m) * K);
  std::vector<float> sc(size_t(m) * K);
  SCHK(s, hipMemcpyAsync(ids.data(), s->ins_did, ids.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  SCHK(s, hipMemcpyAsync(sc.data(), s->ins_dsc, sc.size() * sizeof(float), hipMemcpyDeviceToHost, st));
  SCHK(s, hipStreamSynchronize(st));
  // lists of the new rows; reverse proposals (target, new id, score); a new
  // row is forced into the lists of its kForced nearest neighbours
  constexpr int kForced = 16;
  struct Prop {
    int32_t target, score, self, forced;
  };
  std::vector<int32_t> gi(size_t(m) * K), gs(size_t(m) * K);
  std::vector<Prop> prop;
  prop.reserve(size_t(m) * K);
  for (int32_t j = 0; j < m; ++j) {
    const int32_t self = int32_t(base + j);
    for (int e = 0; e < K; ++e) {
      int32_t id = ids[size_t(j) * K + e];
      const int32_t scv = id >= 0 ? int32_t(sc[size_t(j) * K + e]) : INT32_MIN;
      if (id < 0) id = ids[size_t(j) * K] >= 0 ? ids[size_t(j) * K] : 0;
      gi[size_t(j) * K + e] = id;
      gs[size_t(j) * K + e] = scv;
      if (scv != INT32_MIN) prop.push_back(Prop{id, scv, self, e < kForced ? 1 : 0});
    }
  }
  SCHK(s, hipMemcpyAsync(s->gid + base * K, gi.data(), gi.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, st));
  SCHK(s, hipMemcpyAsync(s->gsc + base * K, gs.data(), gs.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, st));
  std::sort(prop.begin(), prop.end(), [](const Prop& a, const Prop& b) {
    if (a.target != b.target) return a.target < b.target;
    if (a.forced != b.forced) return a.forced > b.forced;
    if (a.score != b.score) return a.score > b.score;
    return a.self < b.self;
  });
  std::vector<int32_t> targets, offs, pid, psc, pfo;
  for (size_t a = 0; a < prop.size();) {
    size_t b = a;
    while (b < prop.size() && prop[b].target == prop[a].target) ++b;
    targets.push_back(prop[a].target);
    offs.push_back(int32_t(pid.size()));
    for (size_t e = a; e < b && e < a + size_t(kMaxK); ++e) {   // first kMaxK per target
      pid.push_back(prop[e].self);
      psc.push_back(prop[e].score);
      pfo.push_back(prop[e].forced);
    }
    a = b;
  }
  offs.push_back(int32_t(pid.size()));
  if (!targets.empty()) {
    SCHK(s, hipMemcpyAsync(s->ins_t, targets.data(), targets.size() * 4, hipMemcpyHostToDevice, st));
    SCHK(s, hipMemcpyAsync(s->ins_off, offs.data(), offs.size() * 4, hipMemcpyHostToDevice, st));
    SCHK(s, hipMemcpyAsync(s->ins_pi, pid.data(), pid.size() * 4, hipMemcpyHostToDevice, st));
    SCHK(s, hipMemcpyAsync(s->ins_ps, psc.data(), psc.size() * 4, hipMemcpyHostToDevice, st));
    SCHK(s, hipMemcpyAsync(s->ins_pf, pfo.data(), pfo.size() * 4, hipMemcpyHostToDevice, st));
    const int grid = std::max(1, std::min<int>(int(targets.size()), s->num_cu * 16));
    hipLaunchKernelGGL(k_merge_reverse, dim3(grid), dim3(64), 0, st, s->gid, s->gsc, K, s->ins_t,
                       s->ins_off, s->ins_pi, s->ins_ps, s->ins_pf, int(targets.size()));
    SCHK(s, hipGetLastError());
  }
  SCHK(s, hipStreamSynchronize(st));   // host vectors above are read by the copies
  s->n.store(base + m);
  *first_id = base;
  return 0;
}

// An insert request of any size, in pieces of at most ins_cap rows.
int insert_rows(knng_server* s, const float* rows, int32_t m, int64_t* first_id) {
  for (int32_t r = 0; r < m; r += s->ins_cap) {
    int64_t f = 0;
    const int rc = do_insert_rows(s, rows + size_t(r) * s->d, std::min(s->ins_cap, m - r), &f);
    if (rc) return rc;
    if (r == 0) *first_id = f;
  }
  return 0;
}

// Insert thread (concurrent mode): applies pending inserts in submission
// order, several requests per pass while their rows fit the scratch.
void inserter(knng_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  std::vector<float> rows;
  while (true) {
    s->cv.wait(lk, [&] { return s->stop || !s->ins_queue.empty(); });
    if (s->ins_queue.empty()) break;   // stop requested and nothing pending
    std::vector<Request*> batch;
    int32_t m = 0;
    while (!s->ins_queue.empty()) {
      Request* r = s->ins_queue.front();
      if (!batch.empty() && m + r->nrows > s->ins_cap) break;
      batch.push_back(r);
      m += r->nrows;
      r->state = 1;
      s->ins_queue.pop_front();
    }
    lk.unlock();
    int rc = 0;
    int64_t first = 0;
    if (batch.size() == 1) {
      rc = insert_rows(s, batch[0]->rows.data(), m, &first);
    } else {
      rows.clear();
      for (Request* r : batch) rows.insert(rows.end(), r->rows.begin(), r->rows.end());
      rc = insert_rows(s, rows.data(), m, &first);
    }
    lk.lock();
    for (Request* r : batch) {
      r->error = rc;
      r->first_id = first;
      first += r->nrows;
      r->state = 2;
      r->t_done = clk::now();
    }
    s->inserts_done += int64_t(batch.size());
    s->cv.notify_all();
  }
}

int count_slots(const knng_server* s, bool express) {
  int k = 0;
  for (const Slot* sl : s->slots) k += sl->express == express ? 1 : 0;
  return k;
}

// A free bulk slot, or a new one while fewer than max_slots exist and HBM
// beyond reserve_bytes is free.  The first bulk slot ignores the reserve (some
// progress is always possible); *fatal is set when even that cannot be
// allocated.  Express slots all exist from the start.
Slot* acquire_slot(knng_server* s, bool express, bool* fatal) {
  *fatal = false;
  for (Slot* sl : s->slots)
    if (!sl->busy && sl->express == express) return sl;
  if (express) return nullptr;
  const int nb = count_slots(s, false);
  if (nb >= s->sp.max_slots) return nullptr;
  if (nb > 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return nullptr;
    if (free_b < slot_device_bytes(s) + size_t(s->sp.reserve_bytes)) return nullptr;   // admission
  }
  Slot* sl = new_slot(s, false);
  if (sl)
    s->slots.push_back(sl);
  else if (nb == 0)
    *fatal = true;
  return sl;
}

// Launch `parts` (request rows, see Slot::Part) on slot sl.  Called with
// the lock held and the requests' bookkeeping already updated; the host
// staging copy and the launches run with the lock released (the slot is
// marked busy, so no other thread touches it), so a lane's thread never
// holds up the other lane, the insert thread or the callers.
void launch_parts(knng_server* s, Slot* sl, std::unique_lock<std::mutex>& lk) {
  sl->busy = true;
  int32_t m = 0;
  std::vector<const float*> src;
  for (const Slot::Part& pt : sl->parts) {
    Request* r = s->reqs[pt.ticket];
    src.push_back(r->rows.data() + size_t(pt.row0) * s->d);
    m = pt.off + pt.nrows;
  }
  lk.unlock();
  for (size_t j = 0; j < sl->parts.size(); ++j)
    memcpy(sl->hx + size_t(sl->parts[j].off) * s->d, src[j],
           size_t(sl->parts[j].nrows) * s->d * sizeof(float));
  const int blocks = sl->express ? s->num_cu * 8 : s->bulk_blocks;
  const int route_cu = sl->express ? kExpressRouteCUs : s->num_cu;
  bool ok;
  if (sl->express) {
    // zero-copy: the quantise kernel reads the pinned staging rows and the
    // search writes ids and scores straight into the pinned result buffers,
    // so the chain holds no copy commands (each one a dispatch that waits for
    // free CUs behind the resident bulk searches)
    ok = !launch_search(s, sl->stream, sl->dx, sl->dq, sl->dsk, sl->dcell, sl->pbest, sl->pidx, sl->pcap, m,
                        s->sp.beam, s->sp.k, s->sp.rerank, sl->hid, sl->hsc, blocks, route_cu, sl, sl->hx);
  } else {
    ok = hipMemcpyAsync(sl->dx, sl->hx, size_t(m) * s->d * sizeof(float), hipMemcpyHostToDevice, sl->stream) ==
         hipSuccess;
    ok = ok && !launch_search(s, sl->stream, sl->dx, sl->dq, sl->dsk, sl->dcell, sl->pbest, sl->pidx,
                              sl->pcap, m, s->sp.beam, s->sp.k, s->sp.rerank, sl->did, sl->dsc, blocks,
                              route_cu, sl);
    ok = ok && hipMemcpyAsync(sl->hid, sl->did, size_t(m) * s->sp.k * sizeof(int32_t),
                              hipMemcpyDeviceToHost, sl->stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(sl->hsc, sl->dsc, size_t(m) * s->sp.k * sizeof(float),
                              hipMemcpyDeviceToHost, sl->stream) == hipSuccess;
  }
  ok = ok && hipEventRecord(sl->done, sl->stream) == hipSuccess;
  lk.lock();
  if (!ok)
    for (const Slot::Part& pt : sl->parts) s->reqs[pt.ticket]->error = KNNG_EHIP;
}

// Add the next max_batch rows (at most) of r to slot sl's parts.
void take_rows(knng_server* s, Request* r, Slot* sl, int32_t limit) {
  const int32_t off = sl->parts.empty() ? 0 : sl->parts.back().off + sl->parts.back().nrows;
  const int32_t m = std::min(limit - off, r->nrows - r->next_row);
  if (r->rows_seen < 0) r->rows_seen = s->n.load();
  sl->parts.push_back(Slot::Part{r->ticket, r->next_row, m, off});
  r->next_row += m;
  r->running += 1;
  r->state = 1;
}

// One thread per lane.  The express thread serves the express queue on the
// express slots; the bulk thread serves the bulk queue (and, with
// strict_order, the inserts in it) on the bulk slots.
void scheduler(knng_server* s, bool express_lane) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    bool progress = false, busy = false;
    for (size_t j = 0; j < s->slots.size(); ++j) {
      Slot* sl = s->slots[j];
      if (sl->express != express_lane || !sl->busy) continue;
      const hipError_t e = hipEventQuery(sl->done);
      if (e == hipErrorNotReady) {
        busy = true;
        continue;
      }
      const size_t kk = size_t(s->sp.k);
      for (const Slot::Part& pt : sl->parts) {
        Request* r = s->reqs[pt.ticket];
        if (e != hipSuccess) r->error = KNNG_EHIP;
        memcpy(r->ids.data() + size_t(pt.row0) * kk, sl->hid + size_t(pt.off) * kk,
               size_t(pt.nrows) * kk * 4);
        memcpy(r->scores.data() + size_t(pt.row0) * kk, sl->hsc + size_t(pt.off) * kk,
               size_t(pt.nrows) * kk * 4);
        r->running -= 1;
        if (r->running == 0 && r->next_row == r->nrows) {
          r->state = 2;
          r->t_done = clk::now();
          s->cv.notify_all();
        }
      }
      sl->parts.clear();
      sl->busy = false;
      progress = true;
    }
    if (express_lane) {
      // continuous batching: every waiting small request that fits goes into
      // one launch on a free express slot
      while (!s->express.empty()) {
        bool fatal = false;
        Slot* sl = acquire_slot(s, true, &fatal);
        if (!sl) break;
        int32_t rows = 0;
        while (!s->express.empty()) {
          Request* r = s->express.front();
          if (s->sp.strict_order && r->epoch > s->inserts_done) break;   // an earlier insert is pending
          if (rows + r->nrows > s->sp.max_batch) break;
          s->express.pop_front();
          take_rows(s, r, sl, s->sp.max_batch);
          rows += r->nrows;
        }
        if (sl->parts.empty()) break;
        launch_parts(s, sl, lk);
        progress = true;
        busy = true;
      }
    } else {
      // bulk_yield: while express requests wait or run, at most that many
      // bulk slots stay busy, so small launches find more free CUs
      int bulk_busy = 0;
      bool express_active = !s->express.empty();
      for (const Slot* x : s->slots) {
        if (!x->busy) continue;
        if (x->express)
          express_active = true;
        else
          ++bulk_busy;
      }
      const bool hold = s->sp.bulk_yield > 0 && express_active;
      while (!s->queue.empty()) {
        Request* r = s->queue.front();
        if (r->kind == 1) {   // strict_order: the insert runs here, alone
          bool any = false;
          for (Slot* sl : s->slots) any = any || sl->busy;
          // drain the queries submitted before the insert, in either lane
          if (!s->express.empty() && s->express.front()->ticket < r->ticket) any = true;
          if (any) break;
          r->state = 1;
          lk.unlock();
          int64_t first = 0;
          const int rc = insert_rows(s, r->rows.data(), r->nrows, &first);
          lk.lock();
          r->error = rc;
          r->first_id = first;
          r->state = 2;
          r->t_done = clk::now();
          s->inserts_done += 1;
          s->queue.pop_front();
          s->cv.notify_all();
          progress = true;
          continue;
        }
        bool fatal = false;
        while (r->next_row < r->nrows) {
          if (hold && bulk_busy >= s->sp.bulk_yield) break;
          Slot* sl = acquire_slot(s, false, &fatal);
          if (!sl) break;
          take_rows(s, r, sl, s->sp.max_batch);
          launch_parts(s, sl, lk);
          ++bulk_busy;
          progress = true;
          busy = true;
        }
        if (fatal) {   // not even one slot fits in HBM: fail the request
          s->err = "cannot allocate a search slot";
          r->error = KNNG_ENOMEM;
          r->next_row = r->nrows;
          if (r->running == 0) {
            r->state = 2;
            r->t_done = clk::now();
            s->cv.notify_all();
          }
          s->queue.pop_front();
          progress = true;
          continue;
        }
        if (r->next_row < r->nrows) break;
        s->queue.pop_front();
      }
    }
    const bool idle = express_lane ? s->express.empty() : s->queue.empty();
    if (s->stop && idle && !busy) break;
    if (!progress) s->cv.wait_for(lk, std::chrono::microseconds(busy ? 20 : 500));
  }
}

}  // namespace

extern "C" {

int knng_server_default_params(knng_server_params* sp) {
  if (!sp) return KNNG_EINVAL;
  sp->beam = 64;
  sp->k = 15;
  sp->entry_trees = 4;
  sp->rerank = 1;
  sp->max_steps = 128;
  sp->max_batch = 1024;
  sp->max_slots = 8;
  sp->pad0 = 0;
  sp->insert_capacity = 1 << 20;
  sp->reserve_bytes = int64_t(2) << 30;
  sp->express_rows = 64;
  sp->express_slots = 2;
  sp->express_cus = 0;
  sp->strict_order = 0;
  sp->bulk_yield = 0;
  return 0;
}

const char* knng_server_last_error(const knng_server* s) { return s ? s->err.c_str() : "null server"; }

int knng_server_create(knng_server** out, knng_ctx* c, const knng_server_params* sp) {
  if (!out || !c || !sp) return KNNG_EINVAL;
  *out = nullptr;
  if (!c->have_graph || !c->have_forest) return KNNG_ESTATE;
  if (sp->beam < 1 || sp->beam > kBeamMax || sp->k < 1 || sp->k > sp->beam || sp->entry_trees < 1 ||
      sp->max_batch < 1 || sp->max_slots < 1 || sp->max_steps < 0 || sp->insert_capacity < 0 ||
      sp->express_rows < 0 || sp->express_slots < 0 || sp->express_cus < 0 || sp->bulk_yield < 0)
    return KNNG_EINVAL;
  const float* xsrc = c->x_ext ? c->x_ext : c->x;
  if (!xsrc) return KNNG_ESTATE;   // the index keeps float32 rows (rerank, inserts)
  auto* s = new knng_server();
  s->ctx = c;
  s->sp = *sp;
  s->device = c->device;
  s->num_cu = c->num_cu;
  s->d = c->d;
  s->words = c->words;
  s->K = c->p.k_build;
  s->n.store(c->n);
  s->n_forest = c->n;
  s->cap = c->n + sp->insert_capacity;
  auto bail = [&](int rc) {
    knng_server_destroy(s);
    return rc;
  };
  if (hipSetDevice(s->device) != hipSuccess) return bail(KNNG_EHIP);
  const size_t cap = size_t(s->cap);
  if (hipMalloc(&s->q, cap * s->d) != hipSuccess || hipMalloc(&s->sk, cap * s->words * 8) != hipSuccess ||
      hipMalloc(&s->x, cap * s->d * sizeof(float)) != hipSuccess ||
      hipMalloc(&s->gid, cap * s->K * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&s->gsc, cap * s->K * sizeof(int32_t)) != hipSuccess)
    return bail(KNNG_ENOMEM);
  s->sp.express_rows = std::min(s->sp.express_rows, s->sp.max_batch);
  // bulk launches are not capped below the chip's block slots: a cap that left
  // room for express blocks cost more bulk throughput than it saved express
  // latency at 880K queries/s
  s->bulk_blocks = s->num_cu * 8;
  if (s->sp.express_slots == 0) s->sp.express_rows = 0;
  s->cu_masked = s->sp.express_rows > 0 && s->sp.express_cus > 0 && s->sp.express_cus < s->num_cu;
  if (hipStreamCreateWithFlags(&s->ins_stream, hipStreamNonBlocking) != hipSuccess) return bail(KNNG_EHIP);
  const size_t n = size_t(c->n);
  if (hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(s->q, c->q, n * s->d, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(s->sk, c->sk, n * s->words * 8, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(s->x, xsrc, n * s->d * sizeof(float), hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(s->gid, c->gid[c->cur], n * s->K * 4, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(s->gsc, c->gsc[c->cur], n * s->K * 4, hipMemcpyDeviceToDevice) != hipSuccess)
    return bail(KNNG_EHIP);
  // insert scratch (inserts run in pieces of at most ins_cap rows)
  s->ins_cap = std::max(s->sp.max_batch, 4096);
  {
    const size_t m = size_t(s->ins_cap), mk = m * size_t(s->K);
    const size_t pe = size_t(knng_host::assign_part_entries(s->ins_cap, s->num_cu));
    if (hipMalloc(&s->ins_did, mk * 4) != hipSuccess || hipMalloc(&s->ins_dsc, mk * 4) != hipSuccess ||
        hipMalloc(&s->ins_dcell, m * size_t(std::max(1, c->nparts)) * 4) != hipSuccess ||
        hipMalloc(&s->ins_pbest, pe * 4) != hipSuccess || hipMalloc(&s->ins_pidx, pe * 4) != hipSuccess ||
        hipMalloc(&s->ins_t, mk * 4) != hipSuccess || hipMalloc(&s->ins_off, (mk + 1) * 4) != hipSuccess ||
        hipMalloc(&s->ins_pi, mk * 4) != hipSuccess || hipMalloc(&s->ins_ps, mk * 4) != hipSuccess ||
        hipMalloc(&s->ins_pf, mk * 4) != hipSuccess)
      return bail(KNNG_ENOMEM);
  }
  // per-partition cell counts for k_route_small, and its slices per partition
  if (c->nparts > 0) {
    if (hipMalloc(&s->route_cells, size_t(c->nparts) * sizeof(int32_t)) != hipSuccess ||
        hipMemcpy(s->route_cells, c->cell_count.data(), size_t(c->nparts) * sizeof(int32_t),
                  hipMemcpyHostToDevice) != hipSuccess)
      return bail(KNNG_ENOMEM);
    s->route_smax = std::max(1, ((c->ncells + 15) / 16 + kRouteTiles - 1) / kRouteTiles);
  }
  // express slots up front; one bulk slot (more as requests and HBM allow)
  for (int j = 0; j < (s->sp.express_rows > 0 ? s->sp.express_slots : 0); ++j) {
    Slot* sl = new_slot(s, true);
    if (!sl) return bail(KNNG_ENOMEM);
    s->slots.push_back(sl);
  }
  {
    Slot* sl = new_slot(s, false);
    if (!sl) return bail(KNNG_ENOMEM);
    s->slots.push_back(sl);
  }
  s->th = std::thread(scheduler, s, false);
  if (s->sp.express_rows > 0) s->th_express = std::thread(scheduler, s, true);
  if (!s->sp.strict_order) s->ins_th = std::thread(inserter, s);
  *out = s;
  return 0;
}

int knng_server_destroy(knng_server* s) {
  if (!s) return 0;
  if (s->th.joinable()) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv.notify_all();
    s->th.join();
    if (s->th_express.joinable()) s->th_express.join();
    if (s->ins_th.joinable()) s->ins_th.join();
  }
  (void)hipSetDevice(s->device);
  for (Slot* sl : s->slots) free_slot(sl);
  for (auto& kv : s->reqs) delete kv.second;
  for (void* p : {static_cast<void*>(s->q), static_cast<void*>(s->sk), static_cast<void*>(s->x),
                  static_cast<void*>(s->gid), static_cast<void*>(s->gsc)})
    if (p) (void)hipFree(p);
  for (void* p : {static_cast<void*>(s->ins_did), static_cast<void*>(s->ins_dsc),
                  static_cast<void*>(s->ins_dcell), static_cast<void*>(s->ins_pbest),
                  static_cast<void*>(s->ins_pidx), static_cast<void*>(s->ins_t),
                  static_cast<void*>(s->ins_off), static_cast<void*>(s->ins_pi),
                  static_cast<void*>(s->ins_ps), static_cast<void*>(s->ins_pf),
                  static_cast<void*>(s->route_cells)})
    if (p) (void)hipFree(p);
  if (s->ins_stream) (void)hipStreamDestroy(s->ins_stream);
  delete s;
  return 0;
}

int knng_server_submit(knng_server* s, int32_t kind, const float* rows, int32_t nrows,
                       int64_t* ticket) {
  if (!s || !rows || nrows < 1 || (kind != 0 && kind != 1) || !ticket) return KNNG_EINVAL;
  auto* r = new Request();
  r->kind = kind;
  r->nrows = nrows;
  r->rows.assign(rows, rows + size_t(nrows) * s->d);
  if (kind == 0) {
    r->ids.assign(size_t(nrows) * s->sp.k, -1);
    r->scores.assign(size_t(nrows) * s->sp.k, 0.f);
  }
  r->t_submit = clk::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    r->epoch = s->inserts_submitted;
    s->reqs[r->ticket] = r;
    if (kind == 1) s->inserts_submitted += 1;
    if (kind == 0 && nrows <= s->sp.express_rows)
      s->express.push_back(r);
    else if (kind == 1 && !s->sp.strict_order)
      s->ins_queue.push_back(r);
    else
      s->queue.push_back(r);
  }
  s->cv.notify_all();
  *ticket = r->ticket;
  return 0;
}

int knng_server_poll(knng_server* s, int64_t ticket, int32_t* state) {
  if (!s || !state) return KNNG_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return KNNG_EINVAL;
  *state = it->second->state;
  return 0;
}

int knng_server_wait(knng_server* s, int64_t ticket, int32_t* ids, float* scores, double* latency_ms) {
  return knng_server_wait_info(s, ticket, ids, scores, latency_ms, nullptr);
}

int knng_server_wait_info(knng_server* s, int64_t ticket, int32_t* ids, float* scores,
                          double* latency_ms, int64_t* index_rows) {
  if (!s) return KNNG_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return KNNG_EINVAL;
  Request* r = it->second;
  s->cv.wait(lk, [&] { return r->state == 2; });
  s->reqs.erase(it);
  lk.unlock();
  const int rc = r->error;
  if (!rc) {
    if (ids && r->kind == 0) memcpy(ids, r->ids.data(), r->ids.size() * sizeof(int32_t));
    if (ids && r->kind == 1)
      for (int32_t j = 0; j < r->nrows; ++j) ids[j] = int32_t(r->first_id + j);
    if (scores && r->kind == 0) memcpy(scores, r->scores.data(), r->scores.size() * sizeof(float));
  } else if (s->err.empty()) {
    s->err = "request failed";
  }
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (index_rows) *index_rows = r->kind == 0 ? r->rows_seen : r->first_id + r->nrows;
  delete r;
  return rc;
}

int knng_server_info(const knng_server* s, int32_t* cu_masked, int32_t* bulk_slots,
                     int32_t* express_slots) {
  if (!s) return KNNG_EINVAL;
  std::lock_guard<std::mutex> lk(const_cast<knng_server*>(s)->mu);
  if (cu_masked) *cu_masked = s->cu_masked ? 1 : 0;
  if (bulk_slots) *bulk_slots = count_slots(s, false);
  if (express_slots) *express_slots = count_slots(s, true);
  return 0;
}

const char* knng_server_mask_error(const knng_server* s) { return s ? s->mask_err.c_str() : ""; }

int64_t knng_server_size(const knng_server* s) {
  if (!s) return -1;
  return s->n.load();
}

int knng_server_search(knng_server* s, const float* queries, int32_t nq, int32_t* ids, float* scores) {
  int64_t t = 0;
  int rc = knng_server_submit(s, 0, queries, nq, &t);
  if (rc) return rc;
  return knng_server_wait(s, t, ids, scores, nullptr);
}

}  // extern "C"
