// knng_batch.cpp -- knng_build_batch: many independent kNN-graph builds of
// uneven size on one GPU (include/knng.h).  Worker threads each own one
// context (and so one HIP stream) per job and take jobs largest-first from a
// shared counter; kernels of different jobs run concurrently on the device,
// which keeps the GPU busy while small jobs are in their launch-bound phases.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <numeric>
#include <thread>
#include <vector>

#include "knng.h"

namespace {

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

}  // namespace

extern "C" int knng_build_batch(knng_job* jobs, int32_t njobs, int32_t device, int32_t workers) {
  if (!jobs || njobs < 0 || workers < 1) return KNNG_EINVAL;
  std::vector<int32_t> order(njobs);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return jobs[a].n != jobs[b].n ? jobs[a].n > jobs[b].n : a < b;
  });
  std::atomic<int32_t> next{0};
  auto work = [&] {
    for (;;) {
      const int32_t t = next.fetch_add(1);
      if (t >= njobs) return;
      knng_job* j = &jobs[order[t]];
      const auto a = std::chrono::steady_clock::now();
      j->status = run_job(j, device);
      j->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
    }
  };
  std::vector<std::thread> th;
  const int nw = std::min<int32_t>(workers, std::max<int32_t>(njobs, 1));
  for (int w = 0; w < nw; ++w) th.emplace_back(work);
  for (auto& t : th) t.join();
  int rc = 0;
  for (int32_t j = 0; j < njobs; ++j)
    if (jobs[j].status) rc = jobs[j].status;
  return rc;
}
