// knng_batch.cpp -- knng_build_batch: many independent kNN-graph builds of
// uneven size on one GPU (include/knng.h).  No reference counterpart: the
// reference builds one graph of DIM = 1024 (src/main.cpp:20).
//
// Jobs whose builds use plain RP trees (no k-means cells at their size) and
// share every build parameter except k_eval are packed into ONE context: their
// rows are concatenated into segments, and every stage runs once for all of
// them -- one forest pass (each segment's trees start from a root of their
// own and never cross segments), one leaf-tile init, and join iterations
// whose work list (a device array rebuilt each iteration from the segments
// still iterating) feeds a persistent k_join through its row queues.  Each
// job keeps its own stop rule, evaluated on its own per-segment counters, so
// its lists are exactly those of its build alone.  The iteration (work list,
// join, counter snapshot, copies) is captured once per buffer direction as a
// hipGraph and replayed, and the join's shape comes from the autotuner.
// Groups run one after another; jobs that need cells (or that a group cannot
// take) run whole builds on worker threads, one context and stream each,
// beside the groups.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <tuple>
#include <vector>

#include "knng.h"
#include "knng_internal.h"

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point a) {
  return std::chrono::duration<double, std::milli>(clk::now() - a).count();
}

int run_job(knng_job* j, int device) {
  knng_params p = j->params;
  p.keep_fp32 = j->rerank && !j->rows_on_device;
  knng_ctx* ctx = nullptr;
  int rc = knng_create(&ctx, device, &p, j->n);
  if (rc) return rc;
  if (j->rows_on_device) {
    rc = knng_ingest_device(ctx, j->rows, 0, j->n);
    if (!rc && j->rerank) rc = knng_attach_fp32_device(ctx, j->rows);
  } else {
    rc = knng_ingest_chunk(ctx, j->rows, 0, j->n);
  }
  if (!rc) rc = knng_build(ctx);
  if (!rc) rc = knng_export(ctx, j->knns, j->dists, p.k_eval, j->rerank);
  if (!rc) {
    knng_stats st;
    knng_get_stats(ctx, &st);
    j->iterations = st.iterations;
  }
  knng_destroy(ctx);
  return rc;
}

// every build parameter a packed group must share (k_eval and keep_fp32 may differ)
using GroupKey = std::tuple<int32_t, int32_t, int32_t, int32_t, int32_t, int32_t, uint32_t, uint32_t, uint64_t,
                            int32_t, int32_t, uint32_t, int32_t>;

uint32_t fbits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

GroupKey group_key(const knng_params& p) {
  return GroupKey(p.dim, p.k_build, p.num_trees, p.rerank_mult, p.leaf_size, p.max_iters,
                  fbits(p.early_stop_rate), fbits(p.rate_diff_tol), p.seed, p.rerank_width, p.stop_rule,
                  fbits(p.insert_stop_rate), p.incremental);
}

// One packed group: all jobs in `ids` in one context.
int run_group(knng_job* jobs, const std::vector<int32_t>& ids, int device) {
  const auto t0 = clk::now();
  knng_params p = jobs[ids[0]].params;
  p.cells = 0;
  p.keep_fp32 = 0;
  for (int32_t j : ids) p.keep_fp32 |= jobs[j].rerank ? 1 : 0;
  std::vector<int64_t> seg(1, 0);
  for (int32_t j : ids) seg.push_back(seg.back() + jobs[j].n);
  knng_ctx* ctx = nullptr;
  int rc = knng_create(&ctx, device, &p, seg.back());
  auto finish = [&](int code) {
    for (int32_t j : ids) {
      if (!jobs[j].status) jobs[j].status = code;
      jobs[j].build_ms = ms_since(t0);
    }
    if (ctx) knng_destroy(ctx);
    return code;
  };
  if (rc) return finish(rc);
  if ((rc = knng_host::set_segments(ctx, seg))) return finish(rc);
  for (size_t s = 0; s < ids.size() && !rc; ++s) {
    const knng_job& j = jobs[ids[s]];
    rc = j.rows_on_device ? knng_ingest_device(ctx, j.rows, seg[s], j.n)
                          : knng_ingest_chunk(ctx, j.rows, seg[s], j.n);
  }
  if (rc) return finish(rc);
  std::vector<int32_t> iters, cur;
  bool graph_used = false;
  const char* ng = std::getenv("KNNG_BATCH_NO_GRAPH");
  const bool use_graph = !(ng && ng[0] == '1');
  if ((rc = knng_host::build_packed(ctx, &iters, &cur, use_graph, &graph_used))) return finish(rc);
  // exports: one device buffer pair sized for the largest job, copied out per job
  int64_t cap = 0;
  for (int32_t j : ids) cap = std::max<int64_t>(cap, jobs[j].n * int64_t(jobs[j].params.k_eval));
  int32_t* dk = nullptr;
  float* dd = nullptr;
  if (hipMalloc(&dk, size_t(cap) * 4) != hipSuccess || hipMalloc(&dd, size_t(cap) * 4) != hipSuccess) {
    if (dk) (void)hipFree(dk);
    return finish(KNNG_ENOMEM);
  }
  for (size_t s = 0; s < ids.size() && !rc; ++s) {
    knng_job& j = jobs[ids[s]];
    const int k = j.params.k_eval;
    rc = knng_host::export_segment(ctx, int(s), cur[s], k, j.rerank, dk, dd);
    if (!rc && (hipMemcpyAsync(j.knns, dk, size_t(j.n) * k * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipMemcpyAsync(j.dists, dd, size_t(j.n) * k * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess))
      rc = KNNG_EHIP;
    j.iterations = iters[s];
  }
  (void)hipFree(dk);
  (void)hipFree(dd);
  return finish(rc);
}

}  // namespace

extern "C" int knng_build_batch(knng_job* jobs, int32_t njobs, int32_t device, int32_t workers) {
  if (!jobs || njobs < 0 || workers < 1) return KNNG_EINVAL;
  for (int32_t j = 0; j < njobs; ++j) {
    jobs[j].status = 0;
    jobs[j].iterations = 0;
    jobs[j].build_ms = 0;
    if (!jobs[j].rows || jobs[j].n < 2 || !jobs[j].knns || !jobs[j].dists) return KNNG_EINVAL;
    if (jobs[j].params.reverse_width != 0) return KNNG_EINVAL;   // reverse lists: single builds only
  }
  // packed groups: plain-tree jobs with equal build parameters (a job of
  // <= leaf_size rows has a single leaf per tree, which packs as well)
  std::map<GroupKey, std::vector<int32_t>> groups;
  std::vector<int32_t> single;
  const char* np = std::getenv("KNNG_BATCH_NO_PACK");
  const bool pack = !(np && np[0] == '1');
  for (int32_t j = 0; j < njobs; ++j) {
    if (pack && knng_host::cells_for(jobs[j].params, jobs[j].n) == 0)
      groups[group_key(jobs[j].params)].push_back(j);
    else
      single.push_back(j);
  }
  std::sort(single.begin(), single.end(), [&](int32_t a, int32_t b) {
    return jobs[a].n != jobs[b].n ? jobs[a].n > jobs[b].n : a < b;
  });
  std::vector<std::vector<int32_t>> glist;
  for (auto& kv : groups) glist.push_back(kv.second);
  // the groups on one thread, in turn (each fills the GPU); whole builds on the others
  std::thread gth([&] {
    for (const auto& g : glist) run_group(jobs, g, device);
  });
  std::atomic<int32_t> next{0};
  auto work = [&] {
    for (;;) {
      const int32_t t = next.fetch_add(1);
      if (t >= int32_t(single.size())) return;
      knng_job* j = &jobs[single[t]];
      const auto a = clk::now();
      j->status = run_job(j, device);
      j->build_ms = ms_since(a);
    }
  };
  std::vector<std::thread> th;
  const int nw = std::min<int32_t>(std::max(1, workers - (glist.empty() ? 0 : 1)),
                                   std::max<int32_t>(int32_t(single.size()), 1));
  for (int w = 0; w < nw && !single.empty(); ++w) th.emplace_back(work);
  for (auto& t : th) t.join();
  gth.join();
  int rc = 0;
  for (int32_t j = 0; j < njobs; ++j)
    if (jobs[j].status) rc = jobs[j].status;
  return rc;
}
