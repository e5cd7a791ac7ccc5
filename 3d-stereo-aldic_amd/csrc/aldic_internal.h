// aldic_internal.h -- objects behind the opaque handles of include/aldic.h
// and the error helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <string>
#include <utility>
#include <atomic>
#include <memory>
#include <vector>

#include "../../include/aldic.h"

namespace aldic {
int set_error(int code, const std::string& msg);
// aldic_subpb2_dev_async with a device flag: nonzero *d_skip (may be null)
// ends the CG before its first chunk (d_Uhat then holds the start)
int subpb2_async(aldic_ctx* ctx, aldic_fem* f, double beta, double mu, double alpha,
                 double winstepsize, const double* d_U, const double* d_F, double* d_Uhat,
                 int32_t max_iter, int32_t* d_iterations, const int* d_skip);
int64_t fem_nodes(const aldic_fem* f);
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return aldic::set_error(e_ == hipErrorOutOfMemory ? ALDIC_ENOMEM : ALDIC_EDEVICE, \
                              std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define RC_TRY(expr)                 \
  do {                               \
    const int rc_ = (expr);          \
    if (rc_ != ALDIC_OK) return rc_; \
  } while (0)

struct aldic_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // reusable staging buffers of the host-pointer entry points
  std::vector<void*> buf;
  std::vector<size_t> cap;
  // per-stage kernel timing (HIP events on the context stream)
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  double stage_ms[ALDIC_NSTAGES] = {0};
  int stage_cnt[ALDIC_NSTAGES] = {0};
  // cross-context ordering points (aldic_ctx_record / aldic_ctx_wait)
  hipEvent_t marks[ALDIC_NMARKS] = {nullptr};
  // a second stream (default priority) for the small size classes of a
  // solve, forked from and joined back to `stream` with two events
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // graphs captured on this context that are still alive, plus one while a
  // capture is open: staging buffers must not move while it is non-zero
  // (shared with the graphs so either may be destroyed first)
  std::shared_ptr<std::atomic<int>> pinned = std::make_shared<std::atomic<int>>(0);
  // integer search lag tiles: ALDIC_SEARCH_AUTO, _DIRECT or _FFT
  int search_method = 2;
};

struct aldic_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::shared_ptr<std::atomic<int>> pinned;  // of the capturing context
};

// Handles record their device so they can be destroyed after their context.
struct aldic_images {
  aldic_ctx* ctx = nullptr;
  int device = 0;
  int64_t M = 0, N = 0;
  float* fm = nullptr;
  float* g = nullptr;
  float* mask = nullptr;
  float* fx = nullptr;
  float* fy = nullptr;
  // interpolant of g: 0 the reference's local kernel on g, 1 the
  // interpolating cubic B-spline on gc (its coefficients, made on request)
  int interp = 0;
  float* gc = nullptr;
};

struct aldic_plan {
  aldic_ctx* ctx = nullptr;
  int device = 0;
  int64_t n = 0;
  int rect = 0;
  int* lists[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int cnt[5] = {0, 0, 0, 0, 0};
  // class c's subsets are all class_side(c) x class_side(c)
  int full[5] = {0, 0, 0, 0, 0};
  int* defer_cnt = nullptr;
  int* defer_list = nullptr;
  // hand-on list of the classes that run on the side stream (solved there
  // while the largest class still runs on the context stream)
  int* defer_cnt2 = nullptr;
  int* defer_list2 = nullptr;
  // per-class lists of a node subset (IC-GN of listed nodes): capacity
  // cnt[c] for c < 4, n for list 4 (the workgroup kernel, which also takes
  // the overflow of a full class list)
  int* sub_lists[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int* sub_cnt = nullptr;  // 5 counts + the number of entries dropped (list 4 full)
  // launch shape per size class and solver (0 IC-GN, 1 Subpb1)
  int var[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
  // IC-GN stopping rule of the plan's solves (aldic_plan_set_stopping)
  int stop_rule = 0;
  float tol_dp = 0.0f;
  // the fast classes as one persistent launch (aldic_plan_set_persistent),
  // with its two work counters
  int persistent = 0;
  int* pqueue = nullptr;
};

// Stage timing: a pair of events around each launch of a stage on the
// context stream, resolved when the caller asks for the totals.
inline hipEvent_t take_event(aldic_ctx* ctx) {
  if (!ctx->ev_pool.empty()) {
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Names of the stages in profiler traces (roctx ranges, "aldic:<stage>").
inline const char* stage_name(int id) {
  static const char* const names[ALDIC_NSTAGES] = {
      "aldic:search",        "aldic:grid_to_nodes",  "aldic:icgn_c0",        "aldic:icgn_c1",
      "aldic:icgn_c2",       "aldic:icgn_c3",        "aldic:icgn_generic",   "aldic:icgn_handed_on",
      "aldic:subpb1_c0",     "aldic:subpb1_c1",      "aldic:subpb1_c2",      "aldic:subpb1_c3",
      "aldic:subpb1_generic", "aldic:subpb1_handed_on", "aldic:cnn",          "aldic:fallback",
      "aldic:subpb2"};
  return id >= 0 && id < ALDIC_NSTAGES ? names[id] : "aldic:stage";
}

// A stage's launches: a roctx range on the host thread (shown by
// rocprofv3 --marker-trace; a no-op without a tool) and, with timing on, a
// pair of events on the stream.
struct StageTimer {
  aldic_ctx* ctx;
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t st;
  StageTimer(aldic_ctx* c, int stage, hipStream_t on = nullptr)
      : ctx(c), id(stage), st(on ? on : c->stream) {
    roctxRangePushA(stage_name(stage));
    if (!ctx->timing) return;
    a = take_event(ctx);
    b = take_event(ctx);
    if (a) (void)hipEventRecord(a, st);
  }
  ~StageTimer() {
    roctxRangePop();
    if (!ctx->timing || !a || !b) return;
    (void)hipEventRecord(b, st);
    ctx->pending.push_back({id, {a, b}});
  }
};
