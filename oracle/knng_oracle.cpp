// knng_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference kNN-graph engine (src/main.cpp of
// Aftab5550/SISAP-2026-task1) used as the checker for the MI355X build.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library; the product path (libknng.so, knng_main) never links it.
//
// Two modes live here:
//
//  * reference mode  (orc_build_reference): the reference pipeline with its
//    own RNG (std::mt19937 seeded 1337+t, uniform_int_distribution pivots),
//    its in-place (Gauss-Seidel) refinement with nth_element / partial_sort
//    tie orders, and its stop rule.  At one thread the reference is
//    deterministic, so this mode is pinned byte-for-byte against the compiled
//    reference's golden outputs (tests/golden/).
//
//  * device mode (orc_forest_device / orc_init_device / orc_join_device /
//    orc_rerank_device): the deterministic semantics the GPU kernels
//    implement (hash-chosen pivots, Jacobi refinement, (score desc, id asc)
//    order, dedup before the sketch prefilter).  The GPU is checked
//    bit-exactly against this mode stage by stage.
//
// Shared primitives (quantise, sketch, popcount similarity, int8 dot,
// xorshift32 padding) are the same functions in both modes.
//
// Pinning (what checks each function independently of this file):
//   orc_quantize         closed form + edge rows + reference golden dists
//   orc_dot_i8           closed form (exact int32) + golden dists
//   orc_sketch_sim       closed form; golden knns (top-400 choice uses it)
//   orc_build_reference  golden knns/dists from the compiled reference
//   orc_forest_device    invariants (partition, stable order, leaf bound) +
//                        brute-force recursion in tests/test_oracle.py
//   orc_forest_cells_device  same invariants + cells as contiguous runs
//   orc_assign_device    float brute-force argmax in numpy (tests)
//   orc_kmeans_device    fixed points: iters=0 gives the quantised seed rows;
//                        separated clusters are recovered exactly; each
//                        centroid equals the requantised sum of its cell
//   orc_init_device      brute-force restatement in numpy (tests)
//   orc_join_device      brute-force restatement in numpy (tests), with and
//                        without the leaf-mate rule and the reverse lists
//   orc_rerank_device    FP64 numpy dot products
//   orc_export           reference golden outputs (through reference mode)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

namespace {

constexpr int32_t kScoreMin = std::numeric_limits<int32_t>::min();
constexpr float kFltMin = 1.17549435e-38f;
constexpr int32_t kMaxDepth = 64;   // device-mode forest depth cap

struct Edge {              // one graph slot: src/main.cpp:31-37
  int32_t id;
  int32_t score;
};

// xorshift32 step used for init padding: src/main.cpp:273-275
inline uint32_t xorshift32(uint32_t s) {
  s ^= s << 13;
  s ^= s >> 17;
  s ^= s << 5;
  return s;
}

// Quantise one element: src/main.cpp:184-186 (multiply, min(127), max(-128),
// truncating cast).  The reference is built with -ffast-math, so x86 runs it
// with DAZ/FTZ set: a denormal input reads as zero.
inline void quantize_elem(float x, int8_t* q, bool* bit) {
  if (std::fabs(x) < kFltMin) x = 0.0f;   // DAZ
  float scaled = x * 127.0f;
  float hi = (scaled < 127.0f) ? scaled : 127.0f;
  float c = (hi > -128.0f) ? hi : -128.0f;
  *q = static_cast<int8_t>(c);
  *bit = c > 0.0f;                        // src/main.cpp:188-190
}

inline int32_t dot_i8(const int8_t* a, const int8_t* b, int d) {
  int32_t s = 0;
  for (int k = 0; k < d; ++k) s += int32_t(a[k]) * int32_t(b[k]);
  return s;
}

// 1024 - sum of popcount(a ^ b): src/main.cpp:354-370, generalised to d bits.
inline int32_t sketch_sim(const uint64_t* a, const uint64_t* b, int words) {
  int32_t h = words * 64;
  for (int w = 0; w < words; ++w) h -= __builtin_popcountll(a[w] ^ b[w]);
  return h;
}

// splitmix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014).  Device-mode pivot choice only.
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

extern "C" {

struct orc_params {
  int32_t k_build;        // src/main.cpp:22
  int32_t num_trees;      // :23
  int32_t rerank_mult;    // :24
  int32_t leaf_size;      // :25
  int32_t max_iters;      // :27
  float early_stop_rate;  // :28
  float rate_diff_tol;    // :29
};

struct orc_stats {
  int32_t iterations;
  int32_t num_leaves;
  int64_t updates[64];
  float rates[64];
};

// Device-mode pivot positions of a node: p1 in [0,size), p2 in [0,size)\{p1}.
void orc_pivots(uint64_t seed, int32_t tree, int32_t depth, int64_t start,
                int64_t size, int64_t* p1pos, int64_t* p2pos) {
  uint64_t key = (uint64_t(uint32_t(tree)) << 48) ^ (uint64_t(uint32_t(depth)) << 40) ^
                 uint64_t(start);
  uint64_t h1 = mix64(seed ^ mix64(key));
  uint64_t h2 = mix64(h1);
  int64_t a = int64_t(h1 % uint64_t(size));
  int64_t b = int64_t(h2 % uint64_t(size - 1));
  if (b >= a) ++b;
  *p1pos = a;
  *p2pos = b;
}

// Quantise + sketch rows: src/main.cpp:179-192.  d % 64 == 0.
int orc_quantize(const float* x, int64_t n, int32_t d, int8_t* q, uint64_t* sk) {
  if (d <= 0 || d % 64) return -1;
  const int words = d / 64;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t* s = sk + i * words;
    for (int w = 0; w < words; ++w) s[w] = 0;
    for (int k = 0; k < d; ++k) {
      bool bit;
      quantize_elem(x[i * d + k], &q[i * d + k], &bit);
      if (bit) s[k / 64] |= 1ull << (k % 64);
    }
  }
  return 0;
}

int32_t orc_dot_i8(const int8_t* a, const int8_t* b, int32_t d) { return dot_i8(a, b, d); }

int32_t orc_sketch_sim(const uint64_t* a, const uint64_t* b, int32_t words) {
  return sketch_sim(a, b, words);
}

// ---------------------------------------------------------------------------
// Reference mode
// ---------------------------------------------------------------------------

// Two-pivot RP tree, src/main.cpp:47-99.  Explicit stack, left subtree first,
// so leaves come out in the reference's completion order and the shared
// mt19937 is advanced in the same depth-first order.
static void ref_tree(std::vector<int>& root, std::vector<std::vector<int>>& leaves,
                     std::mt19937& rng, const int8_t* q, int d, int leaf_size) {
  std::vector<std::vector<int>> stack;
  stack.push_back(std::move(root));
  while (!stack.empty()) {
    std::vector<int> set = std::move(stack.back());
    stack.pop_back();
    if (set.size() <= size_t(leaf_size)) {
      leaves.push_back(std::move(set));
      continue;
    }
    std::uniform_int_distribution<int> pick(0, int(set.size()) - 1);
    const int a = set[pick(rng)];
    int b = set[pick(rng)];
    while (b == a) b = set[pick(rng)];
    std::vector<int> lo, hi;
    lo.reserve(set.size());
    hi.reserve(set.size());
    const int8_t* pa = q + size_t(a) * d;
    const int8_t* pb = q + size_t(b) * d;
    for (int id : set) {
      if (id == a) { lo.push_back(id); continue; }
      if (id == b) { hi.push_back(id); continue; }
      const int8_t* v = q + size_t(id) * d;
      (dot_i8(v, pa, d) > dot_i8(v, pb, d) ? lo : hi).push_back(id);
    }
    // push right first so the left child is processed (and consumes the RNG) first
    stack.push_back(std::move(hi));
    stack.push_back(std::move(lo));
  }
}

int orc_build_reference(const int8_t* q, const uint64_t* sk, int64_t n64, int32_t d,
                        const orc_params* p, int32_t* out_ids, int32_t* out_scores,
                        orc_stats* st) {
  if (n64 < 2 || d <= 0 || d % 64) return -1;
  const size_t n = size_t(n64);
  const int K = p->k_build, T = p->num_trees, words = d / 64;
  const int M = K * p->rerank_mult;

  // forest: src/main.cpp:205-216 (trees in order at one thread)
  std::vector<std::vector<int>> leaves;
  for (int t = 0; t < T; ++t) {
    std::mt19937 rng(1337 + t);
    std::vector<int> all(n);
    for (size_t i = 0; i < n; ++i) all[i] = int(i);
    ref_tree(all, leaves, rng, q, d, p->leaf_size);
  }

  // leaf pointers: src/main.cpp:222-232
  std::vector<int> leaf_of(n * T, -1);
  for (size_t l = 0; l < leaves.size(); ++l)
    for (int id : leaves[l]) {
      int* slot = &leaf_of[size_t(id) * T];
      int t = 0;
      while (t < T && slot[t] != -1) ++t;
      if (t < T) slot[t] = int(l);
    }

  std::vector<Edge> g(n * K, Edge{-1, kScoreMin});
  auto by_score = [](const Edge& x, const Edge& y) { return x.score > y.score; };

  // init: src/main.cpp:234-314
  std::vector<int> cands;
  std::vector<Edge> pool;
  for (size_t i = 0; i < n; ++i) {
    cands.clear();
    pool.clear();
    for (int t = 0; t < T; ++t) {
      const int l = leaf_of[i * T + t];
      if (l < 0) continue;
      for (int m : leaves[l])
        if (size_t(m) != i) cands.push_back(m);
    }
    std::sort(cands.begin(), cands.end());
    cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
    const int8_t* qi = q + i * d;
    for (int c : cands) pool.push_back(Edge{c, dot_i8(qi, q + size_t(c) * d, d)});
    uint32_t s = uint32_t(1337 + i);
    while (pool.size() < size_t(K)) {
      s = xorshift32(s);
      const int r = int(s % n);
      if (size_t(r) == i) continue;
      pool.push_back(Edge{r, dot_i8(qi, q + size_t(r) * d, d)});
    }
    if (size_t(K) < pool.size())
      std::partial_sort(pool.begin(), pool.begin() + K, pool.end(), by_score);
    else
      std::sort(pool.begin(), pool.end(), by_score);
    int kept = 0, prev = -1;
    for (size_t j = 0; j < pool.size() && kept < K; ++j)
      if (pool[j].id != prev) {
        g[i * K + kept++] = pool[j];
        prev = pool[j].id;
      }
    while (kept < K) {
      s = xorshift32(s);
      const int r = int(s % n);
      if (size_t(r) != i) g[i * K + kept++] = Edge{r, kScoreMin};
    }
  }

  // refinement: src/main.cpp:320-440 (in place, row order, one thread)
  std::vector<std::pair<int, int32_t>> sk_c(size_t(K) * K);
  std::vector<Edge> live(K);
  float prev_rate = 1.0f;
  int iters = 0;
  for (int it = 0; it < p->max_iters; ++it) {
    int64_t updates = 0;
    for (size_t i = 0; i < n; ++i) {
      Edge* row = &g[i * K];
      for (int k = 0; k < K; ++k) live[k] = row[k];
      const uint64_t* si = sk + i * words;
      int cnt = 0;
      for (int k = 0; k < K; ++k) {
        const int n1 = row[k].id;
        for (int k2 = 0; k2 < K; ++k2) {
          const int n2 = g[size_t(n1) * K + k2].id;
          if (size_t(n2) == i) continue;
          sk_c[cnt++] = {n2, sketch_sim(si, sk + size_t(n2) * words, words)};
        }
      }
      const int top = cnt < M ? cnt : M;
      if (top > 0 && top < cnt)
        std::nth_element(sk_c.begin(), sk_c.begin() + top, sk_c.begin() + cnt,
                         [](const std::pair<int, int32_t>& x, const std::pair<int, int32_t>& y) {
                           return x.second > y.second;
                         });
      const int8_t* qi = q + i * d;
      for (int m = 0; m < top; ++m) {
        const int c = sk_c[m].first;
        bool present = false;
        for (int k = 0; k < K && !present; ++k) present = live[k].id == c;
        if (present) continue;
        const int32_t sc = dot_i8(qi, q + size_t(c) * d, d);
        if (sc <= live[K - 1].score) continue;
        int pos = K - 1;                       // insert after every score >= sc
        for (; pos > 0 && live[pos - 1].score < sc; --pos) live[pos] = live[pos - 1];
        live[pos] = Edge{c, sc};
      }
      int changed = 0;
      for (int k = 0; k < K; ++k)
        if (row[k].id != live[k].id) {
          row[k] = live[k];
          ++changed;
        }
      updates += changed;
    }
    const float rate = float(updates) / float(n * size_t(K));
    if (st && it < 64) {
      st->updates[it] = updates;
      st->rates[it] = rate;
    }
    iters = it + 1;
    if (rate < p->early_stop_rate || (it > 0 && (prev_rate - rate) < p->rate_diff_tol)) break;
    prev_rate = rate;
  }
  if (st) {
    st->iterations = iters;
    st->num_leaves = int32_t(leaves.size());
  }
  for (size_t j = 0; j < n * K; ++j) {
    out_ids[j] = g[j].id;
    out_scores[j] = g[j].score;
  }
  return 0;
}

// Export best k of each K-list: id+1, float(score), row reversed.
// src/main.cpp:449-457.
int orc_export(const int32_t* ids, const int32_t* scores, int64_t n, int32_t K,
               int32_t k_eval, int32_t* knns, float* dists) {
  if (k_eval > K) return -1;
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < k_eval; ++k) {
      knns[i * k_eval + (k_eval - 1 - k)] = ids[i * K + k] + 1;
      dists[i * k_eval + (k_eval - 1 - k)] = float(scores[i * K + k]);
    }
  return 0;
}

// ---------------------------------------------------------------------------
// Device mode
// ---------------------------------------------------------------------------

// Forest with hash-chosen pivots, all trees.  perm is T x n; leaves are
// written as (tree, start, size) triples in (tree, start) order.
// Returns the number of leaves, or -1.
}  // extern "C"

namespace {

struct Split {             // an internal node: pivot ids and left-child size
  int32_t p1id, p2id;
  int64_t nleft;
};

inline uint64_t node_key(int32_t tree, int32_t depth, int64_t start) {
  return (uint64_t(uint32_t(tree)) << 48) ^ (uint64_t(uint32_t(depth)) << 40) ^ uint64_t(start);
}

// Cell seeding of the forest: `assign` holds G partitions' cell of every row
// ([G][n]; partition g has cells 0..counts[g]-1); tree t starts from
// partition t % G.
struct CellSeed {
  const int32_t* assign;
  const int32_t* counts;
  int32_t G;
};

// stable counting sort of rows by cell; offs has C + 1 entries
void cell_order(const int32_t* a, int64_t n, int32_t C, int32_t* rows, std::vector<int64_t>* offs) {
  offs->assign(size_t(C) + 1, 0);
  for (int64_t i = 0; i < n; ++i) (*offs)[size_t(a[i]) + 1]++;
  for (int32_t c = 0; c < C; ++c) (*offs)[size_t(c) + 1] += (*offs)[size_t(c)];
  std::vector<int64_t> cur(offs->begin(), offs->end() - 1);
  for (int64_t i = 0; i < n; ++i) rows[cur[size_t(a[i])]++] = int32_t(i);
}

int64_t forest_impl(const int8_t* q, int64_t n, int32_t d, int32_t T, int32_t leaf_size,
                    uint64_t seed, int32_t* perm, int64_t* leaf_tri, int64_t leaf_cap,
                    std::unordered_map<uint64_t, Split>* rec, const CellSeed* cs = nullptr) {
  if (n < 2 || leaf_size < 2) return -1;
  int64_t nleaf = 0;
  std::vector<int32_t> lo, hi;
  struct Node { int64_t start, size; int32_t depth; };
  // cells replace the root split (as the library does) unless the whole set is one leaf
  const bool cells = cs && n > leaf_size;
  for (int t = 0; t < T; ++t) {
    int32_t* P = perm + int64_t(t) * n;
    std::vector<Node> stack;
    if (cells) {
      std::vector<int64_t> offs;
      const int32_t C = cs->counts[t % cs->G];
      cell_order(cs->assign + int64_t(t % cs->G) * n, n, C, P, &offs);
      for (int32_t c = C - 1; c >= 0; --c)
        stack.push_back(Node{offs[size_t(c)], offs[size_t(c) + 1] - offs[size_t(c)], 1});
    } else {
      for (int64_t i = 0; i < n; ++i) P[i] = int32_t(i);
      stack.push_back(Node{0, n, 0});
    }
    while (!stack.empty()) {
      Node nd = stack.back();
      stack.pop_back();
      if (nd.size <= leaf_size || nd.depth >= kMaxDepth) {
        // a node still too large at depth kMaxDepth (degenerate input, e.g.
        // many identical rows) is cut into consecutive leaves
        for (int64_t s0 = 0; s0 < nd.size; s0 += leaf_size) {
          if (leaf_tri) {
            if (nleaf >= leaf_cap) return -1;
            leaf_tri[3 * nleaf] = t;
            leaf_tri[3 * nleaf + 1] = nd.start + s0;
            leaf_tri[3 * nleaf + 2] = std::min<int64_t>(leaf_size, nd.size - s0);
          }
          ++nleaf;
        }
        continue;
      }
      int64_t a, b;
      orc_pivots(seed, t, nd.depth, nd.start, nd.size, &a, &b);
      const int32_t pa = P[nd.start + a], pb = P[nd.start + b];
      lo.clear();
      hi.clear();
      for (int64_t j = 0; j < nd.size; ++j) {
        const int32_t id = P[nd.start + j];
        if (j == a) { lo.push_back(id); continue; }
        if (j == b) { hi.push_back(id); continue; }
        const int8_t* v = q + int64_t(id) * d;
        const int32_t diff = dot_i8(v, q + int64_t(pa) * d, d) - dot_i8(v, q + int64_t(pb) * d, d);
        (diff > 0 ? lo : hi).push_back(id);
      }
      std::copy(lo.begin(), lo.end(), P + nd.start);
      std::copy(hi.begin(), hi.end(), P + nd.start + int64_t(lo.size()));
      const int64_t nl = int64_t(lo.size());
      if (rec) (*rec)[node_key(t, nd.depth, nd.start)] = Split{pa, pb, nl};
      stack.push_back(Node{nd.start + nl, nd.size - nl, nd.depth + 1});
      stack.push_back(Node{nd.start, nl, nd.depth + 1});
    }
  }
  return nleaf;
}

// Centroid of a cell from its exact int64 member sum S: the int8 direction
// z_k = trunc(S_k * 127 / max_k |S_k|) and w = float(1 / sqrt(sum z_k^2));
// an all-zero sum gives z = 0, w = 0.
void quantize_centroid(const int64_t* S, int32_t d, int8_t* z, float* w) {
  int64_t M = 0;
  for (int32_t k = 0; k < d; ++k) M = std::max<int64_t>(M, S[k] < 0 ? -S[k] : S[k]);
  if (M == 0) {
    std::fill(z, z + d, int8_t(0));
    *w = 0.0f;
    return;
  }
  double ss = 0.0;
  for (int32_t k = 0; k < d; ++k) {
    z[k] = int8_t((S[k] * 127) / M);
    ss += double(z[k]) * double(z[k]);
  }
  *w = float(1.0 / std::sqrt(ss));
}

// cell of rows off + j * stride: argmax_c float(<q_r, z_c>) * w_c, the first
// (smallest c) maximum; excl skips c == j; score[j] (if given) = the maximum
void assign_rows(const int8_t* q, int64_t nrows, int64_t off, int64_t stride, int32_t d,
                 const int8_t* z, const float* w, int32_t C, bool excl, int32_t* out, float* score) {
  for (int64_t j = 0; j < nrows; ++j) {
    const int8_t* v = q + (off + j * stride) * d;
    float best = -std::numeric_limits<float>::infinity();
    int32_t bi = std::numeric_limits<int32_t>::max();
    for (int32_t c = 0; c < C; ++c) {
      if (excl && c == j) continue;
      const float s = float(dot_i8(v, z + int64_t(c) * d, d)) * w[c];
      if (s > best) {
        best = s;
        bi = c;
      }
    }
    out[j] = bi;
    if (score) score[j] = best;
  }
}

// exact member sums per cell of rows off + j * stride with cells a[j]
void cell_sums(const int8_t* q, int32_t d, const int32_t* a, int64_t nrows, int64_t off, int64_t stride,
               int32_t C, std::vector<int64_t>* S, std::vector<int64_t>* cnt) {
  S->assign(size_t(C) * d, 0);
  cnt->assign(size_t(C), 0);
  for (int64_t j = 0; j < nrows; ++j) {
    int64_t* s = &(*S)[size_t(a[j]) * d];
    const int8_t* v = q + (off + j * stride) * d;
    for (int32_t k = 0; k < d; ++k) s[k] += v[k];
    ++(*cnt)[size_t(a[j])];
  }
}

}  // namespace

extern "C" {

int64_t orc_forest_device(const int8_t* q, int64_t n, int32_t d, int32_t T, int32_t leaf_size,
                          uint64_t seed, int32_t* perm, int64_t* leaf_tri, int64_t leaf_cap) {
  return forest_impl(q, n, d, T, leaf_size, seed, perm, leaf_tri, leaf_cap, nullptr);
}

int32_t orc_assign_device(const int8_t* q, int64_t n, int32_t d, const int8_t* z, const float* w,
                          int32_t C, int32_t* out) {
  if (C < 1) return -1;
  assign_rows(q, n, 0, 1, d, z, w, C, false, out, nullptr);
  return 0;
}

// Cells of partition g (device semantics of knng.hip's run_kmeans):
//  seeds: Cs = min(Ct * over, ns) rows ((2c + 1) n / (2 Cs) + 7919 g) mod n,
//    ns = sample rows off + j * sample, off = g mod sample;
//  iters Lloyd rounds on the sample (an empty cell keeps its centroid), a
//    last sample assignment and its exact sums;
//  merging: live (non-empty) cells, centroids requantised from their sums;
//    per round each live cell j finds k = argmax_{k != j} float(<z_j, z_k>) *
//    w_k (value b_j); mutual pairs j < k with b_j * w_j > merge fold k into j
//    (sums and counts added, centroid requantised); rounds until none merges
//    (<= 64);
//  pruning: cells with float(count) < prune * float(upper-median count) go;
//  full_iters Lloyd rounds over all rows with the live cells in index order,
//    then the final assignment.
// Returns the cell count C_g (z / w hold C_g centroids; room for Cs needed).
int32_t orc_kmeans_device(const int8_t* q, int64_t n, int32_t d, int32_t Ct, int32_t g,
                          int32_t iters, int32_t over, int32_t sample, float merge,
                          int32_t full_iters, float prune, int8_t* zf, float* wf,
                          int32_t* assign) {
  if (Ct < 1 || n < 1 || over < 1 || sample < 1) return -1;
  const int64_t stride = sample;
  const int64_t off = int64_t(g) % stride < n ? int64_t(g) % stride : 0;
  const int64_t ns = (n - off + stride - 1) / stride;
  const int32_t Cs = int32_t(std::max<int64_t>(1, std::min<int64_t>(int64_t(Ct) * over, ns)));
  std::vector<int8_t> z(size_t(Cs) * d);
  std::vector<float> w(static_cast<size_t>(Cs));
  {
    std::vector<int64_t> S(static_cast<size_t>(d));
    for (int32_t c = 0; c < Cs; ++c) {
      const int64_t id = ((2 * int64_t(c) + 1) * n / (2 * int64_t(Cs)) + int64_t(g) * 7919) % n;
      for (int32_t k = 0; k < d; ++k) S[size_t(k)] = q[id * d + k];
      quantize_centroid(S.data(), d, &z[size_t(c) * d], &w[size_t(c)]);
    }
  }
  std::vector<int32_t> a(static_cast<size_t>(ns));
  std::vector<int64_t> S, cnt;
  for (int32_t it = 0;; ++it) {
    assign_rows(q, ns, off, stride, d, z.data(), w.data(), Cs, false, a.data(), nullptr);
    cell_sums(q, d, a.data(), ns, off, stride, Cs, &S, &cnt);
    if (it == iters) break;
    for (int32_t c = 0; c < Cs; ++c)
      if (cnt[size_t(c)]) quantize_centroid(&S[size_t(c) * d], d, &z[size_t(c) * d], &w[size_t(c)]);
  }
  std::vector<int32_t> live;
  for (int32_t c = 0; c < Cs; ++c)
    if (cnt[size_t(c)]) {
      live.push_back(c);
      quantize_centroid(&S[size_t(c) * d], d, &z[size_t(c) * d], &w[size_t(c)]);
    }
  for (int round = 0; round < 64 && live.size() >= 2; ++round) {
    const int32_t m = int32_t(live.size());
    std::vector<int8_t> zl(size_t(m) * d);
    std::vector<float> wl(static_cast<size_t>(m)), bval(static_cast<size_t>(m));
    std::vector<int32_t> best(static_cast<size_t>(m));
    for (int32_t j = 0; j < m; ++j) {
      std::copy_n(&z[size_t(live[size_t(j)]) * d], d, &zl[size_t(j) * d]);
      wl[size_t(j)] = w[size_t(live[size_t(j)])];
    }
    assign_rows(zl.data(), m, 0, 1, d, zl.data(), wl.data(), m, true, best.data(), bval.data());
    std::vector<char> gone(size_t(m), 0);
    int merged = 0;
    for (int32_t j = 0; j < m; ++j) {
      const int32_t k = best[size_t(j)];
      if (k <= j || k >= m || best[size_t(k)] != j) continue;
      if (!(bval[size_t(j)] * wl[size_t(j)] > merge)) continue;
      int64_t* sj = &S[size_t(live[size_t(j)]) * d];
      const int64_t* sk = &S[size_t(live[size_t(k)]) * d];
      for (int32_t t = 0; t < d; ++t) sj[t] += sk[t];
      cnt[size_t(live[size_t(j)])] += cnt[size_t(live[size_t(k)])];
      quantize_centroid(sj, d, &z[size_t(live[size_t(j)]) * d], &w[size_t(live[size_t(j)])]);
      gone[size_t(k)] = 1;
      ++merged;
    }
    if (!merged) break;
    std::vector<int32_t> next;
    for (int32_t j = 0; j < m; ++j)
      if (!gone[size_t(j)]) next.push_back(live[size_t(j)]);
    live.swap(next);
  }
  {
    std::vector<int64_t> sz;
    for (int32_t k : live) sz.push_back(cnt[size_t(k)]);
    std::sort(sz.begin(), sz.end());
    const float cut = prune * float(sz[sz.size() / 2]);
    std::vector<int32_t> keep;
    for (int32_t k : live)
      if (!(float(cnt[size_t(k)]) < cut)) keep.push_back(k);
    live.swap(keep);
  }
  const int32_t C = int32_t(live.size());
  for (int32_t j = 0; j < C; ++j) {
    std::copy_n(&z[size_t(live[size_t(j)]) * d], d, zf + int64_t(j) * d);
    wf[j] = w[size_t(live[size_t(j)])];
  }
  for (int32_t it = 0;; ++it) {
    assign_rows(q, n, 0, 1, d, zf, wf, C, false, assign, nullptr);
    if (it == full_iters) break;
    cell_sums(q, d, assign, n, 0, 1, C, &S, &cnt);
    for (int32_t c = 0; c < C; ++c)
      if (cnt[size_t(c)]) quantize_centroid(&S[size_t(c) * d], d, zf + int64_t(c) * d, wf + c);
  }
  return C;
}

// Forest whose trees start from cells: assign = [G][n] cells of the rows,
// counts = cells per partition.
int64_t orc_forest_cells_device(const int8_t* q, int64_t n, int32_t d, int32_t T, int32_t leaf_size,
                                uint64_t seed, const int32_t* assign, const int32_t* counts,
                                int32_t G, int32_t* perm, int64_t* leaf_tri, int64_t leaf_cap) {
  if (G < 1) return -1;
  const CellSeed cs{assign, counts, G};
  return forest_impl(q, n, d, T, leaf_size, seed, perm, leaf_tri, leaf_cap, nullptr, &cs);
}

// Beam search over a K-list graph (device semantics of the serving kernel):
// entry points = members of the leaf reached by descending each of the first
// min(entry_trees, T, 4) trees (left iff <q,p1> - <q,p2> > 0; a depth-capped
// node enters its first <= leaf_size rows, and any leaf at most its first 64;
// with cells, tree t starts at the query's cell of partition t % G); then, per step, merge the scored
// candidates into the beam (best `beam` by (score desc, id asc), expanded
// flags kept), expand the first 4 unexpanded beam entries in beam order (at
// most max_steps times), taking their neighbours that are not in the visited
// set; the visited set is reset to the beam's ids when its insert count plus
// 4K would exceed 3072.  Output: best k of the beam (ids, int8 scores),
// padded with -1 / INT32_MIN.
int orc_search_device(const int8_t* q, int64_t n, int32_t d, const int32_t* gid, int32_t K,
                      int32_t T, int32_t leaf_size, uint64_t seed, int32_t entry_trees,
                      int32_t beam, int32_t k, int32_t max_steps, const int8_t* qq, int64_t nq,
                      const int32_t* assign, const int32_t* counts, int32_t G,
                      const int32_t* qcell, int32_t* out_ids, int32_t* out_sc) {
  std::vector<int32_t> perm(size_t(T) * n);
  std::unordered_map<uint64_t, Split> rec;
  const CellSeed cs{assign, counts, G};
  const bool cells = assign && G >= 1 && n > leaf_size;
  if (forest_impl(q, n, d, T, leaf_size, seed, perm.data(), nullptr, 0, &rec, cells ? &cs : nullptr) < 0)
    return -1;
  std::vector<std::vector<int64_t>> coffs(cells ? G : 0);
  for (int32_t g = 0; g < int32_t(coffs.size()); ++g) {
    std::vector<int32_t> rows(n);
    cell_order(assign + int64_t(g) * n, n, counts[g], rows.data(), &coffs[g]);
  }
  const int et = std::min(std::min(entry_trees, T), 4);
  constexpr int kExpand = 4;
  struct B { int32_t id, sc, ex; };
  for (int64_t qi = 0; qi < nq; ++qi) {
    const int8_t* y = qq + qi * d;
    std::unordered_set<int32_t> vis;
    int64_t used = 0;
    std::vector<int32_t> cands;
    for (int t = 0; t < et; ++t) {
      int64_t start = 0, size = n;
      int32_t depth = 0;
      if (cells) {   // enter at the query's cell of this tree's partition
        const int32_t g = t % G, c = qcell[int64_t(g) * nq + qi];
        start = coffs[g][size_t(c)];
        size = coffs[g][size_t(c) + 1] - start;
        depth = 1;
        if (size == 0) continue;
      }
      while (size > leaf_size && depth < kMaxDepth) {
        const Split& sp = rec.at(node_key(t, depth, start));
        const int32_t diff = dot_i8(y, q + int64_t(sp.p1id) * d, d) - dot_i8(y, q + int64_t(sp.p2id) * d, d);
        if (diff > 0) {
          size = sp.nleft;
        } else {
          start += sp.nleft;
          size -= sp.nleft;
        }
        ++depth;
      }
      const int64_t m = std::min<int64_t>(std::min<int64_t>(size, leaf_size), 64);
      for (int64_t j = 0; j < m; ++j) {
        const int32_t id = perm[size_t(t) * n + start + j];
        if (vis.insert(id).second) {
          ++used;
          cands.push_back(id);
        }
      }
    }
    std::vector<B> bm, all;
    for (int step = 0;; ++step) {
      all = bm;
      for (int32_t c : cands) all.push_back(B{c, dot_i8(y, q + int64_t(c) * d, d), 0});
      std::sort(all.begin(), all.end(), [](const B& a, const B& b) {
        return a.sc != b.sc ? a.sc > b.sc : a.id < b.id;
      });
      if (int(all.size()) > beam) all.resize(beam);
      bm = all;
      std::vector<int32_t> sel;
      if (step < max_steps)
        for (auto& e : bm)
          if (!e.ex && int(sel.size()) < kExpand) {
            e.ex = 1;
            sel.push_back(e.id);
          }
      if (sel.empty()) break;
      cands.clear();
      if (used + int64_t(kExpand) * K > 3072) {
        vis.clear();
        for (auto& e : bm) vis.insert(e.id);
        used = int64_t(bm.size());
      }
      for (int32_t sid : sel)
        for (int e = 0; e < K; ++e) {
          const int32_t id = gid[int64_t(sid) * K + e];
          if (vis.insert(id).second) {
            ++used;
            cands.push_back(id);
          }
        }
    }
    for (int j = 0; j < k; ++j) {
      const bool ok = j < int(bm.size());
      out_ids[qi * k + j] = ok ? bm[j].id : -1;
      out_sc[qi * k + j] = ok ? bm[j].sc : kScoreMin;
    }
  }
  return 0;
}

static inline bool edge_before(const Edge& x, const Edge& y) {
  return x.score != y.score ? x.score > y.score : x.id < y.id;
}

// Init from the forest: union of leaf-mates, int8 dots, top-K by
// (score desc, id asc) with duplicate ids dropped, xorshift32 padding as the
// reference (src/main.cpp:271-312) when fewer than K candidates exist.
int orc_init_device(const int8_t* q, int64_t n, int32_t d, int32_t T, int32_t K,
                    const int32_t* perm, const int64_t* leaf_tri, int64_t nleaf,
                    int32_t* ids, int32_t* scores) {
  std::vector<int64_t> leaf_of(size_t(n) * T, -1);   // point -> leaf index per tree
  for (int64_t l = 0; l < nleaf; ++l) {
    const int64_t t = leaf_tri[3 * l], s = leaf_tri[3 * l + 1], z = leaf_tri[3 * l + 2];
    for (int64_t j = 0; j < z; ++j) leaf_of[size_t(perm[t * n + s + j]) * T + t] = l;
  }
  std::vector<int32_t> cands;
  std::vector<Edge> pool;
  for (int64_t i = 0; i < n; ++i) {
    cands.clear();
    pool.clear();
    for (int t = 0; t < T; ++t) {
      const int64_t l = leaf_of[size_t(i) * T + t];
      if (l < 0) return -1;
      const int64_t s = leaf_tri[3 * l + 1], z = leaf_tri[3 * l + 2];
      for (int64_t j = 0; j < z; ++j) {
        const int32_t m = perm[int64_t(t) * n + s + j];
        if (m != i) cands.push_back(m);
      }
    }
    std::sort(cands.begin(), cands.end());
    cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
    const int8_t* qi = q + i * d;
    for (int32_t c : cands) pool.push_back(Edge{c, dot_i8(qi, q + int64_t(c) * d, d)});
    uint32_t s = uint32_t(1337 + i);
    while (pool.size() < size_t(K)) {
      s = xorshift32(s);
      const int32_t r = int32_t(s % uint64_t(n));
      if (r == i) continue;
      pool.push_back(Edge{r, dot_i8(qi, q + int64_t(r) * d, d)});
    }
    std::sort(pool.begin(), pool.end(), edge_before);
    int kept = 0, prev = -1;
    for (size_t j = 0; j < pool.size() && kept < K; ++j)
      if (pool[j].id != prev) {
        ids[i * K + kept] = pool[j].id;
        scores[i * K + kept] = pool[j].score;
        prev = pool[j].id;
        ++kept;
      }
    while (kept < K) {
      s = xorshift32(s);
      const int32_t r = int32_t(s % uint64_t(n));
      if (r != i) {
        ids[i * K + kept] = r;
        scores[i * K + kept] = kScoreMin;
        ++kept;
      }
    }
  }
  return 0;
}

// Sketch-prefilter threshold: the largest t with |{h >= t}| >= M (all ties
// at t are kept).  Returns 0 when fewer than M candidates exist.
int32_t orc_sketch_threshold(const int32_t* h, int64_t cnt, int32_t M, int32_t bits) {
  if (cnt <= M) return 0;
  std::vector<int64_t> hist(size_t(bits) + 1, 0);
  for (int64_t j = 0; j < cnt; ++j) hist[size_t(h[j])]++;
  int64_t acc = 0;
  for (int32_t t = bits; t >= 0; --t) {
    acc += hist[size_t(t)];
    if (acc >= M) return t;
  }
  return 0;
}

// One Jacobi refinement iteration (device semantics).  Returns the number of
// slots whose id changed (src/main.cpp:421-427 counting rule), or -1.
//
// fresh_in (optional, one u64 per row, bit k = entry k entered the list in
// the previous iteration): incremental join -- the pair (i -> n1 -> c) is a
// candidate only if n1 is fresh in i's list or c is fresh in n1's list.
// fresh_out (optional) receives the same masks for the new lists.
// leafkey (optional, [n][T]: row r's leaf in tree t, as its first position in
// the tree's permutation) when the lists descend from that forest's leaf init:
// the leaf-mate rule.  The mask of an edge r -> e has bit t (trees 0..7) set
// when r and e share tree t's leaf; the path i -> n1 -> c is dropped when the
// masks of i -> n1 and n1 -> c share a bit (then c shares that leaf with i:
// the init scored it exactly, so it cannot enter i's list).  c is a candidate
// when some path to it survives.
// R > 0 (reverse_width): reverse-neighbour candidates.  RV(v) = the R best
// rows u by (score desc, u asc) whose lists hold v as an entry this iteration
// treats as new (every entry without fresh_in, else the fresh ones).  Row i
// also takes u in RV(i) (path i <- u), w in RV(n1) for its neighbours n1
// (path i -> n1 <- w) and the entries w of u's list for u in RV(i) (path
// i <- u -> w); every reverse edge is new, and each path is dropped under the
// leaf-mate rule when its two edges' masks share a bit (for i <- u: when the
// edge's mask is not 0).
int64_t orc_join_device(const int8_t* q, const uint64_t* sk, int64_t n, int32_t d, int32_t K,
                        int32_t M, const int32_t* ids_in, const int32_t* sc_in,
                        int32_t* ids_out, int32_t* sc_out, int64_t* inserted,
                        const uint64_t* fresh_in, uint64_t* fresh_out, const int32_t* leafkey,
                        int32_t T, int32_t R) {
  int64_t ins = 0;
  const int words = d / 64;
  int64_t updates = 0;
  std::unordered_set<int32_t> seen;
  std::vector<int32_t> cand, hval;
  struct Slot { Edge e; int32_t slot; };
  std::vector<Slot> all;
  const int tm = T < 8 ? T : 8;
  auto edge_mask = [&](int64_t r, int64_t e) {
    uint32_t m = 0;
    for (int t = 0; t < tm; ++t) m |= uint32_t(leafkey[r * T + t] == leafkey[e * T + t]) << t;
    return m;
  };
  // reverse lists: every new edge u -> v offered to v, the R best kept
  std::vector<int32_t> rv(size_t(R > 0 ? n * R : 0), -1);
  if (R > 0) {
    std::vector<std::vector<Edge>> in(static_cast<size_t>(n));
    for (int64_t u = 0; u < n; ++u)
      for (int k = 0; k < K; ++k)
        if (!fresh_in || ((fresh_in[u] >> k) & 1ull)) {
          const int32_t v = ids_in[u * K + k];
          if (v >= 0 && v < n) in[size_t(v)].push_back(Edge{int32_t(u), sc_in[u * K + k]});
        }
    for (int64_t v = 0; v < n; ++v) {
      std::vector<Edge>& e = in[size_t(v)];
      std::sort(e.begin(), e.end(), edge_before);
      for (size_t j = 0; j < e.size() && j < size_t(R); ++j) rv[size_t(v * R) + j] = e[j].id;
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* own = ids_in + i * K;
    seen.clear();
    cand.clear();
    hval.clear();
    seen.insert(int32_t(i));
    for (int k = 0; k < K; ++k) seen.insert(own[k]);
    for (int k = 0; k < K; ++k) {
      const int32_t* nb = ids_in + int64_t(own[k]) * K;
      const bool own_fresh = !fresh_in || ((fresh_in[i] >> k) & 1ull);
      const uint32_t m1 = leafkey ? edge_mask(i, own[k]) : 0u;
      for (int k2 = 0; k2 < K; ++k2) {
        if (!own_fresh && !((fresh_in[own[k]] >> k2) & 1ull)) continue;
        if (leafkey && (m1 & edge_mask(own[k], nb[k2]))) continue;   // leaf-mate path
        if (seen.insert(nb[k2]).second) cand.push_back(nb[k2]);
      }
    }
    if (R > 0) {
      const int32_t* ri = rv.data() + i * R;
      for (int j = 0; j < R; ++j) {   // i <- u
        const int32_t u = ri[j];
        if (u < 0 || (leafkey && edge_mask(u, i))) continue;
        if (seen.insert(u).second) cand.push_back(u);
      }
      for (int k = 0; k < K; ++k) {   // i -> n1 <- w
        const int32_t n1 = own[k];
        const uint32_t m1 = leafkey ? edge_mask(i, n1) : 0u;
        for (int j = 0; j < R; ++j) {
          const int32_t w = rv[size_t(int64_t(n1) * R + j)];
          if (w < 0 || (leafkey && (m1 & edge_mask(w, n1)))) continue;
          if (seen.insert(w).second) cand.push_back(w);
        }
      }
      for (int j = 0; j < R; ++j) {   // i <- u -> w
        const int32_t u = ri[j];
        if (u < 0) continue;
        const uint32_t mu = leafkey ? edge_mask(u, i) : 0u;
        for (int k = 0; k < K; ++k) {
          const int32_t w = ids_in[int64_t(u) * K + k];
          if (leafkey && (mu & edge_mask(u, w))) continue;
          if (seen.insert(w).second) cand.push_back(w);
        }
      }
    }
    const uint64_t* si = sk + i * words;
    for (int32_t c : cand) hval.push_back(sketch_sim(si, sk + int64_t(c) * words, words));
    const int32_t thr = orc_sketch_threshold(hval.data(), int64_t(cand.size()), M, d);
    all.clear();
    for (int k = 0; k < K; ++k) all.push_back(Slot{Edge{own[k], sc_in[i * K + k]}, k});
    const int8_t* qi = q + i * d;
    for (size_t j = 0; j < cand.size(); ++j)
      if (hval[j] >= thr)
        all.push_back(Slot{Edge{cand[j], dot_i8(qi, q + int64_t(cand[j]) * d, d)}, K});
    std::sort(all.begin(), all.end(), [](const Slot& x, const Slot& y) {
      if (x.e.score != y.e.score) return x.e.score > y.e.score;
      if (x.e.id != y.e.id) return x.e.id < y.e.id;
      return x.slot < y.slot;
    });
    uint64_t fm = 0;
    for (int k = 0; k < K; ++k) {
      if (all[k].e.id != own[k]) ++updates;
      if (all[k].slot == K) {   // a candidate (new id) entered the list
        ++ins;
        fm |= 1ull << k;
      }
      ids_out[i * K + k] = all[k].e.id;
      sc_out[i * K + k] = all[k].e.score;
    }
    if (fresh_out) fresh_out[i] = fm;
  }
  if (inserted) *inserted = ins;
  return updates;
}

// FP32-rerank export (device semantics): the first R list entries are
// re-scored with exact (FP64) dot products of the original rows, the best
// k_eval by (score desc, id asc) are emitted reversed, as orc_export.
int orc_rerank_device(const float* x, int64_t n, int32_t d, int32_t K, int32_t R,
                      int32_t k_eval, const int32_t* ids, int32_t* knns, double* dists) {
  if (R > K || k_eval > R) return -1;
  struct Cand { double s; int32_t id; };
  std::vector<Cand> c(R);
  for (int64_t i = 0; i < n; ++i) {
    const float* xi = x + i * d;
    for (int r = 0; r < R; ++r) {
      const int32_t id = ids[i * K + r];
      const float* xc = x + int64_t(id) * d;
      double acc = 0.0;
      for (int k = 0; k < d; ++k) acc += double(xi[k]) * double(xc[k]);
      c[r] = Cand{acc, id};
    }
    std::sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) {
      return a.s != b.s ? a.s > b.s : a.id < b.id;
    });
    for (int k = 0; k < k_eval; ++k) {
      knns[i * k_eval + (k_eval - 1 - k)] = c[k].id + 1;
      dists[i * k_eval + (k_eval - 1 - k)] = c[k].s;
    }
  }
  return 0;
}

}  // extern "C"
