// capi.hip -- the C ABI of include/aldic.h: contexts, resident image planes,
// plans (size-class node lists + hand-on scratch) and the host-pointer
// drop-in entry points that copy in, launch and copy out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "aldic_internal.h"
#include "aldic_kernels.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

namespace aldic {
int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
}  // namespace aldic

static int fail(int code, const std::string& msg) { return aldic::set_error(code, msg); }

extern "C" const char* aldic_last_error(void) { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Scratch of the no-allocation search: a file-scope device array, one copy
// per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol): its device address on each GPU is looked up once, with that GPU
// current, and kept in scratch_dev[device].
// ---------------------------------------------------------------------------
namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = ALDIC_NOALLOC_SCRATCH_BYTES;  // 32 MiB per GPU
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
std::mutex scratch_mu;
void* scratch_dev[NOALLOC_MAX_DEVICES] = {nullptr};

// device address of `scratch` on `device`, which must be current
int scratch_address(int device, void** out) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    return fail(ALDIC_EINVAL, "no-allocation scratch: device index beyond " +
                                  std::to_string(NOALLOC_MAX_DEVICES - 1));
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ALDIC_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
    scratch_dev[device] = p;
  }
  *out = scratch_dev[device];
  return ALDIC_OK;
}
}  // namespace

namespace {

__global__ void zero_ints(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int set_device(aldic_ctx* ctx) {
  HIP_TRY(hipSetDevice(ctx->device));
  return ALDIC_OK;
}

// staging buffer `slot` of at least `bytes`
int staging(aldic_ctx* ctx, int slot, size_t bytes, void** out) {
  if ((int)ctx->buf.size() <= slot) {
    ctx->buf.resize(slot + 1, nullptr);
    ctx->cap.resize(slot + 1, 0);
  }
  if (ctx->cap[slot] < bytes) {
    if (ctx->pinned->load() > 0)
      return fail(ALDIC_EINVAL, "staging buffer " + std::to_string(slot) +
                                    " would grow while graphs captured on this context are alive "
                                    "(run the sequence once before capturing)");
    if (ctx->buf[slot]) HIP_TRY(hipFree(ctx->buf[slot]));
    ctx->buf[slot] = nullptr;
    ctx->cap[slot] = 0;
    const size_t want = std::max<size_t>(bytes, 256);
    HIP_TRY(hipMalloc(&ctx->buf[slot], want));
    ctx->cap[slot] = want;
  }
  *out = ctx->buf[slot];
  return ALDIC_OK;
}

template <typename T>
int stage(aldic_ctx* ctx, int slot, size_t count, T** out) {
  void* p = nullptr;
  const int rc = staging(ctx, slot, count * sizeof(T), &p);
  *out = static_cast<T*>(p);
  return rc;
}


// staging slots of the device-resident search entry points
constexpr int SLOT_SEARCH = 20;
constexpr int SLOT_SEARCH_NODES = 21;
constexpr int SLOT_SEARCH_GRID_NODES = 22;

int build_plan(aldic_ctx* ctx, const int32_t* winsize, int64_t n, int per_node, int rect,
               aldic_plan* p) {
  p->ctx = ctx;
  p->n = n;
  p->rect = rect;
  std::vector<int> lists[5];
  bool full[5] = {true, true, true, true, false};
  for (int64_t i = 0; i < n; i++) {
    int wx, wy;
    if (rect) {
      wx = per_node ? winsize[i] : winsize[0];
      wy = per_node ? winsize[i + n] : winsize[1];
    } else {
      wx = wy = per_node ? winsize[i] : winsize[0];
    }
    if (wx < 0 || wy < 0 || (wx % 2) || (wy % 2))
      return fail(ALDIC_EINVAL, "winsize must be even and non-negative (validateDICpara.m:22-28)");
    const int c = aldic::size_class(wx + 1, wy + 1);
    lists[c].push_back((int)i);
    if (c < 4 && (wx + 1 != aldic::class_side(c) || wy + 1 != aldic::class_side(c))) full[c] = false;
  }
  for (int c = 0; c < 5; c++) p->full[c] = full[c] && !lists[c].empty();
  for (int c = 0; c < 5; c++) {
    p->cnt[c] = (int)lists[c].size();
    if (p->cnt[c] > 0) {
      HIP_TRY(hipMalloc(&p->lists[c], sizeof(int) * lists[c].size()));
      HIP_TRY(hipMemcpy(p->lists[c], lists[c].data(), sizeof(int) * lists[c].size(),
                        hipMemcpyHostToDevice));
    }
  }
  HIP_TRY(hipMalloc(&p->defer_cnt, sizeof(int)));
  HIP_TRY(hipMalloc(&p->defer_list, sizeof(int) * std::max<int64_t>(n, 1)));
  HIP_TRY(hipMalloc(&p->defer_cnt2, sizeof(int)));
  HIP_TRY(hipMalloc(&p->defer_list2, sizeof(int) * std::max<int64_t>(n, 1)));
  // the per-class lists of a node subset hold at most the class's plan
  // count; list 4 (the workgroup kernel) holds every node, so that a listed
  // node whose class list is full still has a place
  for (int c = 0; c < 4; c++)
    if (p->cnt[c] > 0) HIP_TRY(hipMalloc(&p->sub_lists[c], sizeof(int) * p->cnt[c]));
  HIP_TRY(hipMalloc(&p->sub_lists[4], sizeof(int) * std::max<int64_t>(n, 1)));
  HIP_TRY(hipMalloc(&p->sub_cnt, sizeof(int) * 6));
  HIP_TRY(hipMalloc(&p->pqueue, sizeof(int) * 2));
  return ALDIC_OK;
}

void free_plan(aldic_plan* p) {
  for (int c = 0; c < 5; c++) {
    if (p->lists[c]) (void)hipFree(p->lists[c]);
    if (p->sub_lists[c]) (void)hipFree(p->sub_lists[c]);
  }
  if (p->sub_cnt) (void)hipFree(p->sub_cnt);
  if (p->defer_cnt) (void)hipFree(p->defer_cnt);
  if (p->defer_list) (void)hipFree(p->defer_list);
  if (p->defer_cnt2) (void)hipFree(p->defer_cnt2);
  if (p->defer_list2) (void)hipFree(p->defer_list2);
  if (p->pqueue) (void)hipFree(p->pqueue);
}

// Launch sequence shared by IC-GN (mode 0) and Subpb1 (mode 1).
// The size classes of one solve are independent: the largest class runs on
// the context stream, the others on the side stream at the same time.  Each
// stream hands its subsets on to a list of its own and runs the workgroup
// kernel over it, so the side stream's handed-on subsets (the small classes
// near masked regions) are solved while the largest class still runs; the
// side stream joins the context stream at the end.
int side_stream(aldic_ctx* ctx) {
  if (ctx->side) return ALDIC_OK;
  HIP_TRY(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming));
  return ALDIC_OK;
}

int largest_class(const int cnt[5]) {
  int big = 0;
  for (int c = 1; c < 5; c++)
    if (cnt[c] > cnt[big]) big = c;
  return big;
}

// The persistent variant: classes 0..3 in one launch on the context stream
// (timed as the 65 px class's stage), then the workgroup kernel over the
// > 65 px class and over the subsets the fast classes handed on.
int run_solver_persistent(aldic_ctx* ctx, int mode, SolveArgs a, const aldic_plan* p) {
  hipStream_t s = ctx->stream;
  const int base = mode == 0 ? ALDIC_STAGE_ICGN_C0 : ALDIC_STAGE_SUBPB1_C0;
  zero_ints<<<1, 64, 0, s>>>(p->defer_cnt, 1);
  zero_ints<<<1, 64, 0, s>>>(p->pqueue, 2);
  HIP_TRY(hipGetLastError());
  a.defer_cnt = p->defer_cnt;
  a.defer_list = p->defer_list;
  aldic::PersistArgs P{};
  P.a = a;
  int64_t blocks = 0, waves = 0;
  for (int c = 0; c < 4; c++) {
    P.list[c] = p->lists[c];
    P.cnt[c] = p->cnt[c];
    P.full[c] = p->full[c];
  }
  P.queue = p->pqueue;
  blocks = (int64_t)p->cnt[3] + p->cnt[2];
  waves = (int64_t)p->cnt[1] + p->cnt[0];
  // the resident blocks; no more blocks than there is work for
  int dev_cus = 0;
  (void)hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  const int64_t resident = aldic::PERSIST_BLOCKS_PER_CU * (int64_t)std::max(dev_cus, 1);
  const int grid = (int)std::min<int64_t>(resident, std::max<int64_t>(blocks, (waves + 3) / 4));
  {
    StageTimer t(ctx, base + 3);
    HIP_TRY(aldic::launch_persistent(mode, P, grid, s));
  }
  if (p->cnt[4] > 0) {
    StageTimer t(ctx, base + 4);
    const int g4 = std::min(p->cnt[4], 2048);
    HIP_TRY(aldic::launch_generic(mode, a, p->lists[4], nullptr, p->cnt[4], g4, s));
  }
  StageTimer t(ctx, base + 5);
  const int g5 = (int)std::min<int64_t>(std::max<int64_t>(blocks + waves, 1), 1024);
  HIP_TRY(aldic::launch_generic(mode, a, p->defer_list, p->defer_cnt, (int)p->n, g5, s));
  return ALDIC_OK;
}

int run_solver(aldic_ctx* ctx, int mode, SolveArgs a, const aldic_plan* p) {
  hipStream_t s = ctx->stream;
  RC_TRY(side_stream(ctx));
  a.n = (int)p->n;
  if (mode == 0) {
    a.stop_rule = p->stop_rule;
    a.tol_dp = p->tol_dp;
  }
  if (p->persistent) return run_solver_persistent(ctx, mode, a, p);
  const int base = mode == 0 ? ALDIC_STAGE_ICGN_C0 : ALDIC_STAGE_SUBPB1_C0;
  zero_ints<<<1, 64, 0, s>>>(p->defer_cnt, 1);
  zero_ints<<<1, 64, 0, s>>>(p->defer_cnt2, 1);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ctx->fork_ev, s));
  HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
  const int big = largest_class(p->cnt);
  int64_t side_total = 0;
  for (int c = 0; c < 5; c++) {
    if (p->cnt[c] == 0) continue;
    const bool main = c == big;
    hipStream_t on = main ? s : ctx->side;
    if (!main) side_total += p->cnt[c];
    a.defer_cnt = main ? p->defer_cnt : p->defer_cnt2;
    a.defer_list = main ? p->defer_list : p->defer_list2;
    StageTimer t(ctx, base + c, on);
    if (c < 4) {
      a.list = p->lists[c];
      a.cnt = p->cnt[c];
      a.fixed_side = p->full[c];
      HIP_TRY(aldic::launch_fast(mode, c, p->var[mode][c], a, on));
    } else {
      const int grid = std::min(p->cnt[4], 2048);
      HIP_TRY(aldic::launch_generic(mode, a, p->lists[4], nullptr, p->cnt[4], grid, on));
    }
  }
  // subsets handed on by the fast kernels, per stream; the grid-stride loop
  // reads the count on the device, so no host round trip is needed.  The
  // side stream's pass runs beside the largest class: a smaller grid keeps
  // its idle blocks (64 KB of LDS each) from taking that class's slots.
  if (side_total > 0) {
    StageTimer t(ctx, base + 5, ctx->side);
    const int g2 = (int)std::min<int64_t>(side_total, 256);
    HIP_TRY(aldic::launch_generic(mode, a, p->defer_list2, p->defer_cnt2, (int)p->n, g2,
                                  ctx->side));
  }
  HIP_TRY(hipEventRecord(ctx->join_ev, ctx->side));
  HIP_TRY(hipStreamWaitEvent(s, ctx->join_ev, 0));
  StageTimer t(ctx, base + 5);
  const int grid = (int)std::min<int64_t>(std::max<int64_t>(p->cnt[big], 1), 1024);
  HIP_TRY(aldic::launch_generic(mode, a, p->defer_list, p->defer_cnt, (int)p->n, grid, s));
  return ALDIC_OK;
}

bool images_ok(const aldic_images* im) {
  return im && im->fm && im->g && im->mask && im->fx && im->fy && im->M >= 7 && im->N >= 7;
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int aldic_ctx_create(int device, aldic_ctx** out) {
  if (!out) return fail(ALDIC_EINVAL, "aldic_ctx_create: out is NULL");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(ALDIC_EINVAL, "aldic_ctx_create: no such device");
  auto* c = new aldic_ctx();
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(ALDIC_EDEVICE, std::string("aldic_ctx_create: ") + hipGetErrorString(e));
  }
  // set-up of the no-allocation entry point's scratch on this device (its
  // first call retries and reports the error if this lookup failed)
  void* sp = nullptr;
  (void)scratch_address(device, &sp);
  *out = c;
  return ALDIC_OK;
}

extern "C" int aldic_ctx_set_timing(aldic_ctx* ctx, int32_t enable) {
  if (!ctx) return fail(ALDIC_EINVAL, "NULL context");
  ctx->timing = enable != 0;
  return ALDIC_OK;
}

extern "C" int aldic_ctx_set_search_method(aldic_ctx* ctx, int32_t method) {
  if (!ctx) return fail(ALDIC_EINVAL, "NULL context");
  if (method != ALDIC_SEARCH_DIRECT && method != ALDIC_SEARCH_FFT && method != ALDIC_SEARCH_AUTO)
    return fail(ALDIC_EINVAL, "search method must be ALDIC_SEARCH_AUTO, _DIRECT or _FFT");
  ctx->search_method = method;
  return ALDIC_OK;
}

extern "C" int aldic_ctx_stage_times(aldic_ctx* ctx, double* ms, int32_t* launches, int32_t reset) {
  if (!ctx) return fail(ALDIC_EINVAL, "NULL context");
  RC_TRY(set_device(ctx));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (auto& pr : ctx->pending) {
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, pr.second.first, pr.second.second));
    ctx->stage_ms[pr.first] += t;
    ctx->stage_cnt[pr.first] += 1;
    ctx->ev_pool.push_back(pr.second.first);
    ctx->ev_pool.push_back(pr.second.second);
  }
  ctx->pending.clear();
  for (int k = 0; k < ALDIC_NSTAGES; k++) {
    if (ms) ms[k] = ctx->stage_ms[k];
    if (launches) launches[k] = ctx->stage_cnt[k];
    if (reset) {
      ctx->stage_ms[k] = 0.0;
      ctx->stage_cnt[k] = 0;
    }
  }
  return ALDIC_OK;
}

extern "C" int aldic_ctx_destroy(aldic_ctx* ctx) {
  if (!ctx) return ALDIC_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& pr : ctx->pending) {
    ctx->ev_pool.push_back(pr.second.first);
    ctx->ev_pool.push_back(pr.second.second);
  }
  for (hipEvent_t e : ctx->ev_pool)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->marks)
    if (e) (void)hipEventDestroy(e);
  for (void* p : ctx->buf)
    if (p) (void)hipFree(p);
  if (ctx->side) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
  }
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return ALDIC_OK;
}

extern "C" int aldic_ctx_record(aldic_ctx* ctx, int32_t mark) {
  if (!ctx || mark < 0 || mark >= ALDIC_NMARKS) return fail(ALDIC_EINVAL, "ctx_record: bad argument");
  RC_TRY(set_device(ctx));
  if (!ctx->marks[mark])
    HIP_TRY(hipEventCreateWithFlags(&ctx->marks[mark], hipEventDisableTiming));
  HIP_TRY(hipEventRecord(ctx->marks[mark], ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_ctx_wait(aldic_ctx* ctx, aldic_ctx* other, int32_t mark) {
  if (!ctx || !other || mark < 0 || mark >= ALDIC_NMARKS)
    return fail(ALDIC_EINVAL, "ctx_wait: bad argument");
  if (!other->marks[mark]) return ALDIC_OK;  // never recorded: nothing to wait for
  RC_TRY(set_device(ctx));
  HIP_TRY(hipStreamWaitEvent(ctx->stream, other->marks[mark], 0));
  return ALDIC_OK;
}

// ---------------------------------------------------------------------------
// HIP graphs of a context's work: everything submitted to ctx between
// capture_begin and capture_end (kernels only: no entry point on the
// device-resident path issues a memset or an allocation once its staging
// buffers exist) becomes one graph, replayed with its arguments as captured.
// ---------------------------------------------------------------------------
extern "C" int aldic_ctx_capture_begin(aldic_ctx* ctx) {
  if (!ctx) return fail(ALDIC_EINVAL, "ctx_capture_begin: NULL context");
  if (ctx->timing) return fail(ALDIC_EINVAL, "ctx_capture_begin: stage timing is on");
  RC_TRY(set_device(ctx));
  HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
  ctx->pinned->fetch_add(1);
  return ALDIC_OK;
}

extern "C" int aldic_ctx_capture_end(aldic_ctx* ctx, aldic_graph** out) {
  if (!ctx || !out) return fail(ALDIC_EINVAL, "ctx_capture_end: NULL argument");
  RC_TRY(set_device(ctx));
  *out = nullptr;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(ctx->stream, &st));
  if (st == hipStreamCaptureStatusNone)
    return fail(ALDIC_EINVAL, "ctx_capture_end: no capture is open on this context");
  auto* g = new aldic_graph();
  g->device = ctx->device;
  g->pinned = ctx->pinned;  // the open capture's count passes to the graph
  hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    ctx->pinned->fetch_sub(1);
    delete g;
    return fail(ALDIC_EDEVICE, std::string("ctx_capture_end: ") + hipGetErrorString(e));
  }
  *out = g;
  return ALDIC_OK;
}

extern "C" int aldic_graph_launch(aldic_graph* g, aldic_ctx* ctx) {
  if (!g || !ctx) return fail(ALDIC_EINVAL, "graph_launch: NULL argument");
  RC_TRY(set_device(ctx));
  HIP_TRY(hipGraphLaunch(g->exec, ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_graph_destroy(aldic_graph* g) {
  if (!g) return ALDIC_OK;
  (void)hipSetDevice(g->device);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->pinned) g->pinned->fetch_sub(1);
  delete g;
  return ALDIC_OK;
}

extern "C" void* aldic_ctx_stream(aldic_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int aldic_ctx_synchronize(aldic_ctx* ctx) {
  if (!ctx) return fail(ALDIC_EINVAL, "NULL context");
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return ALDIC_OK;
}

// ---------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------
extern "C" int aldic_dev_malloc(aldic_ctx* ctx, int64_t bytes, void** out) {
  if (!ctx || !out || bytes < 0) return fail(ALDIC_EINVAL, "dev_malloc: bad argument");
  RC_TRY(set_device(ctx));
  *out = nullptr;
  HIP_TRY(hipMalloc(out, (size_t)std::max<int64_t>(bytes, 1)));
  return ALDIC_OK;
}

extern "C" int aldic_dev_free(aldic_ctx* ctx, void* p) {
  if (!ctx) return fail(ALDIC_EINVAL, "dev_free: NULL context");
  if (!p) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipFree(p));
  return ALDIC_OK;
}

extern "C" int aldic_memcpy_htod(aldic_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes) || bytes < 0)
    return fail(ALDIC_EINVAL, "memcpy_htod: bad argument");
  if (bytes == 0) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_memcpy_dtoh(aldic_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes) || bytes < 0)
    return fail(ALDIC_EINVAL, "memcpy_dtoh: bad argument");
  if (bytes == 0) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return ALDIC_OK;
}

// ---------------------------------------------------------------------------
// ba_interp2_spline
// ---------------------------------------------------------------------------
template <typename T>
static int interp_host(aldic_ctx* ctx, const T* F, int64_t M, int64_t N, int64_t P, const T* X,
                       const T* Y, int64_t ND, T* Z) {
  if (!ctx || !F || !X || !Y || !Z) return fail(ALDIC_EINVAL, "ba_interp2_spline: NULL argument");
  if (M < 1 || N < 1 || P < 1 || ND < 0)
    return fail(ALDIC_EINVAL, "ba_interp2_spline: bad dimensions");
  if (ND == 0) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  T *dF, *dX, *dY, *dZ;
  RC_TRY(stage(ctx, 0, (size_t)(M * N * P), &dF));
  RC_TRY(stage(ctx, 1, (size_t)ND, &dX));
  RC_TRY(stage(ctx, 2, (size_t)ND, &dY));
  RC_TRY(stage(ctx, 3, (size_t)(ND * P), &dZ));
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(dF, F, sizeof(T) * M * N * P, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dX, X, sizeof(T) * ND, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dY, Y, sizeof(T) * ND, hipMemcpyHostToDevice, s));
  if constexpr (sizeof(T) == 8)
    HIP_TRY(aldic::launch_interp_f64(dF, M, N, P, dX, dY, ND, dZ, s));
  else
    HIP_TRY(aldic::launch_interp_f32(dF, M, N, P, dX, dY, ND, dZ, s));
  HIP_TRY(hipMemcpyAsync(Z, dZ, sizeof(T) * ND * P, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return ALDIC_OK;
}

extern "C" int aldic_interp_spline_f64(aldic_ctx* ctx, const double* F, int64_t M, int64_t N,
                                       int64_t P, const double* X, const double* Y, int64_t ND,
                                       double* Z) {
  return interp_host<double>(ctx, F, M, N, P, X, Y, ND, Z);
}

extern "C" int aldic_interp_spline_f32(aldic_ctx* ctx, const float* F, int64_t M, int64_t N,
                                       int64_t P, const float* X, const float* Y, int64_t ND,
                                       float* Z) {
  return interp_host<float>(ctx, F, M, N, P, X, Y, ND, Z);
}

// ---------------------------------------------------------------------------
// funImgGradient
// ---------------------------------------------------------------------------
extern "C" int aldic_img_gradient(aldic_ctx* ctx, const double* ImgRef, const double* mask,
                                  int64_t M, int64_t N, double* DfDx, double* DfDy) {
  if (!ctx || !ImgRef || !DfDx || !DfDy) return fail(ALDIC_EINVAL, "funImgGradient: NULL argument");
  if (M < 7 || N < 7) return fail(ALDIC_EINVAL, "funImgGradient: image smaller than 7 x 7");
  RC_TRY(set_device(ctx));
  double *dI, *dm = nullptr, *dx, *dy;
  const int64_t MN = M * N, MC = (M - 6) * (N - 6);
  RC_TRY(stage(ctx, 0, (size_t)MN, &dI));
  if (mask) RC_TRY(stage(ctx, 1, (size_t)MN, &dm));
  RC_TRY(stage(ctx, 2, (size_t)MC, &dx));
  RC_TRY(stage(ctx, 3, (size_t)MC, &dy));
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(dI, ImgRef, sizeof(double) * MN, hipMemcpyHostToDevice, s));
  if (mask) HIP_TRY(hipMemcpyAsync(dm, mask, sizeof(double) * MN, hipMemcpyHostToDevice, s));
  HIP_TRY(aldic::launch_gradient_out_f64(dI, dm, M, N, dx, dy, s));
  HIP_TRY(hipMemcpyAsync(DfDx, dx, sizeof(double) * MC, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(DfDy, dy, sizeof(double) * MC, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return ALDIC_OK;
}

// ---------------------------------------------------------------------------
// images
// ---------------------------------------------------------------------------
static int alloc_planes(aldic_images* im) {
  const size_t bytes = sizeof(float) * (size_t)(im->M * im->N);
  HIP_TRY(hipMalloc(&im->fm, bytes));
  HIP_TRY(hipMalloc(&im->g, bytes));
  HIP_TRY(hipMalloc(&im->mask, bytes));
  HIP_TRY(hipMalloc(&im->fx, bytes));
  HIP_TRY(hipMalloc(&im->fy, bytes));
  return ALDIC_OK;
}

extern "C" int aldic_images_destroy(aldic_images* im) {
  if (!im) return ALDIC_OK;
  (void)hipSetDevice(im->device);
  float* ps[6] = {im->fm, im->g, im->mask, im->fx, im->fy, im->gc};
  for (float* p : ps)
    if (p) (void)hipFree(p);
  delete im;
  return ALDIC_OK;
}

extern "C" int aldic_images_create(aldic_ctx* ctx, const double* ImgRef, const double* ImgDef,
                                   const double* mask, const double* DfDx, const double* DfDy,
                                   int64_t M, int64_t N, aldic_images** out) {
  if (!ctx || !ImgRef || !ImgDef || !out) return fail(ALDIC_EINVAL, "images: NULL argument");
  if (M < 7 || N < 7) return fail(ALDIC_EINVAL, "images: smaller than 7 x 7");
  if ((DfDx == nullptr) != (DfDy == nullptr))
    return fail(ALDIC_EINVAL, "images: give both DfDx and DfDy or neither");
  RC_TRY(set_device(ctx));
  auto* im = new aldic_images();
  im->ctx = ctx;
  im->device = ctx->device;
  im->M = M;
  im->N = N;
  int rc = alloc_planes(im);
  if (rc != ALDIC_OK) {
    aldic_images_destroy(im);
    return rc;
  }
  const int64_t MN = M * N;
  double *dr, *dd, *dm = nullptr;
  hipStream_t s = ctx->stream;
  auto body = [&]() -> int {
    RC_TRY(stage(ctx, 0, (size_t)MN, &dr));
    RC_TRY(stage(ctx, 1, (size_t)MN, &dd));
    if (mask) RC_TRY(stage(ctx, 2, (size_t)MN, &dm));
    HIP_TRY(hipMemcpyAsync(dr, ImgRef, sizeof(double) * MN, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dd, ImgDef, sizeof(double) * MN, hipMemcpyHostToDevice, s));
    if (mask) HIP_TRY(hipMemcpyAsync(dm, mask, sizeof(double) * MN, hipMemcpyHostToDevice, s));
    HIP_TRY(aldic::launch_prep_planes(dr, dd, dm, M, N, im->fm, im->g, im->mask, s));
    if (DfDx) {
      const int64_t MC = (M - 6) * (N - 6);
      double *dgx, *dgy;
      RC_TRY(stage(ctx, 3, (size_t)MC, &dgx));
      RC_TRY(stage(ctx, 4, (size_t)MC, &dgy));
      HIP_TRY(hipMemcpyAsync(dgx, DfDx, sizeof(double) * MC, hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(dgy, DfDy, sizeof(double) * MC, hipMemcpyHostToDevice, s));
      HIP_TRY(aldic::launch_place_df(dgx, dgy, M, N, im->fx, im->fy, s));
    } else {
      HIP_TRY(aldic::launch_gradient(dr, dm, M, N, im->fx, im->fy, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return ALDIC_OK;
  };
  rc = body();
  if (rc != ALDIC_OK) {
    aldic_images_destroy(im);
    return rc;
  }
  *out = im;
  return ALDIC_OK;
}

// ---------------------------------------------------------------------------
// raw 8/16-bit images: funNormalizeImg on the device
// ---------------------------------------------------------------------------
namespace {

// ROI of funNormalizeImg.m:22-28, clamped to the image; 0-based inclusive
int clamp_roi(const int64_t gridx[2], const int64_t gridy[2], int64_t M, int64_t N, int64_t r[4]) {
  if (!gridx || !gridy) return fail(ALDIC_EINVAL, "normalize: NULL ROI");
  int64_t a = gridx[0], b = gridx[1], c = gridy[0], d = gridy[1];
  if (a < 1) a = 1;
  if (b > M) b = M;
  if (c < 1) c = 1;
  if (d > N) d = N;
  if (a > b || c > d) return fail(ALDIC_EINVAL, "normalize: empty ROI");
  r[0] = a - 1;
  r[1] = b - 1;
  r[2] = c - 1;
  r[3] = d - 1;
  return ALDIC_OK;
}

// mean and std (n - 1 normalisation) from exact integer sums: the mean is
// the correctly rounded S / n, the variance (n Q - S^2) / (n (n - 1)) is
// formed exactly before one division
void mean_std(unsigned long long S, unsigned long long Q, int64_t n, double* mean, double* sd) {
  *mean = (double)S / (double)n;
  if (n < 2) {
    *sd = 0.0;
    return;
  }
  const __int128 num = (__int128)n * (__int128)Q - (__int128)S * (__int128)S;
  const long double var = (long double)num / ((long double)n * (long double)(n - 1));
  *sd = (double)sqrtl(var);
}

int normalize_raw(aldic_ctx* ctx, const void* d_raw, int bits, int64_t M, int64_t N,
                  const int64_t r[4], double* d_out, double* ms) {
  hipStream_t s = ctx->stream;
  unsigned long long* sums = nullptr;
  RC_TRY(stage(ctx, 23, 2, &sums));
  HIP_TRY(aldic::launch_raw_sums(d_raw, bits, M, r[0], r[1], r[2], r[3], sums, s));
  unsigned long long h[2];
  HIP_TRY(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  double mean, sd;
  mean_std(h[0], h[1], (r[1] - r[0] + 1) * (r[3] - r[2] + 1), &mean, &sd);
  HIP_TRY(aldic::launch_raw_normalize(d_raw, bits, M * N, mean, sd, d_out, s));
  if (ms) {
    ms[0] = mean;
    ms[1] = sd;
  }
  return ALDIC_OK;
}

}  // namespace

extern "C" int aldic_images_planes(const aldic_images* im, const float** fm, const float** g,
                                   const float** mask, const float** fx, const float** fy) {
  if (!im) return fail(ALDIC_EINVAL, "images_planes: NULL images");
  if (fm) *fm = im->fm;
  if (g) *g = im->g;
  if (mask) *mask = im->mask;
  if (fx) *fx = im->fx;
  if (fy) *fy = im->fy;
  return ALDIC_OK;
}

extern "C" int aldic_bspline_prefilter_dev(aldic_ctx* ctx, const float* d_in, float* d_out,
                                           int64_t M, int64_t N) {
  if (!ctx || !d_in || !d_out) return fail(ALDIC_EINVAL, "bspline_prefilter_dev: NULL argument");
  if (M < 1 || N < 1 || M > INT32_MAX / 2 || N > INT32_MAX / 2)
    return fail(ALDIC_EINVAL, "bspline_prefilter_dev: bad size");
  const int64_t MN = M * N;
  if (d_in < d_out + MN && d_out < d_in + MN)
    return fail(ALDIC_EINVAL, "bspline_prefilter_dev: input and output overlap");
  RC_TRY(set_device(ctx));
  HIP_TRY(aldic::launch_bspline_prefilter(d_in, d_out, M, N, ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_images_set_interp(aldic_images* im, int32_t mode) {
  if (!im || !im->ctx) return fail(ALDIC_EINVAL, "images_set_interp: NULL images");
  if (mode != ALDIC_INTERP_LOCAL && mode != ALDIC_INTERP_BSPLINE)
    return fail(ALDIC_EINVAL, "images_set_interp: mode must be ALDIC_INTERP_LOCAL or _BSPLINE");
  RC_TRY(set_device(im->ctx));
  if (mode == ALDIC_INTERP_BSPLINE) {
    if (!im->gc) HIP_TRY(hipMalloc(&im->gc, sizeof(float) * (size_t)(im->M * im->N)));
    HIP_TRY(aldic::launch_bspline_prefilter(im->g, im->gc, im->M, im->N, im->ctx->stream));
    // solvers may run on another context's stream: the plane is complete
    // when this returns
    HIP_TRY(hipStreamSynchronize(im->ctx->stream));
  }
  im->interp = mode;
  return ALDIC_OK;
}

extern "C" int aldic_normalize_raw_dev(aldic_ctx* ctx, const void* d_raw, int32_t bits, int64_t M,
                                       int64_t N, const int64_t gridx[2], const int64_t gridy[2],
                                       double* d_out, double* mean_std_out) {
  if (!ctx || !d_raw || !d_out || M < 1 || N < 1)
    return fail(ALDIC_EINVAL, "normalize_raw_dev: bad argument");
  if (bits != 8 && bits != 16) return fail(ALDIC_EINVAL, "normalize_raw_dev: bits must be 8 or 16");
  RC_TRY(set_device(ctx));
  int64_t r[4];
  RC_TRY(clamp_roi(gridx, gridy, M, N, r));
  return normalize_raw(ctx, d_raw, bits, M, N, r, d_out, mean_std_out);
}

extern "C" int aldic_images_create_raw(aldic_ctx* ctx, const void* ImgRef, const void* ImgDef,
                                       int32_t bits, const double* mask, int64_t M, int64_t N,
                                       const int64_t gridx[2], const int64_t gridy[2],
                                       aldic_images** out, double* stats) {
  if (!ctx || !ImgRef || !ImgDef || !out) return fail(ALDIC_EINVAL, "images_raw: NULL argument");
  if (bits != 8 && bits != 16) return fail(ALDIC_EINVAL, "images_raw: bits must be 8 or 16");
  if (M < 7 || N < 7) return fail(ALDIC_EINVAL, "images_raw: smaller than 7 x 7");
  RC_TRY(set_device(ctx));
  int64_t r[4];
  RC_TRY(clamp_roi(gridx, gridy, M, N, r));
  auto* im = new aldic_images();
  im->ctx = ctx;
  im->device = ctx->device;
  im->M = M;
  im->N = N;
  int rc = alloc_planes(im);
  if (rc != ALDIC_OK) {
    aldic_images_destroy(im);
    return rc;
  }
  const int64_t MN = M * N;
  const size_t rb = (size_t)MN * (bits / 8);
  hipStream_t s = ctx->stream;
  auto body = [&]() -> int {
    unsigned char *rr, *rd;
    double *dr, *dd, *dm = nullptr;
    RC_TRY(stage(ctx, 5, rb, &rr));
    RC_TRY(stage(ctx, 6, rb, &rd));
    RC_TRY(stage(ctx, 0, (size_t)MN, &dr));
    RC_TRY(stage(ctx, 1, (size_t)MN, &dd));
    if (mask) RC_TRY(stage(ctx, 2, (size_t)MN, &dm));
    HIP_TRY(hipMemcpyAsync(rr, ImgRef, rb, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(rd, ImgDef, rb, hipMemcpyHostToDevice, s));
    if (mask) HIP_TRY(hipMemcpyAsync(dm, mask, sizeof(double) * MN, hipMemcpyHostToDevice, s));
    RC_TRY(normalize_raw(ctx, rr, bits, M, N, r, dr, stats));
    RC_TRY(normalize_raw(ctx, rd, bits, M, N, r, dd, stats ? stats + 2 : nullptr));
    HIP_TRY(aldic::launch_prep_planes(dr, dd, dm, M, N, im->fm, im->g, im->mask, s));
    HIP_TRY(aldic::launch_gradient(dr, dm, M, N, im->fx, im->fy, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ALDIC_OK;
  };
  rc = body();
  if (rc != ALDIC_OK) {
    aldic_images_destroy(im);
    return rc;
  }
  *out = im;
  return ALDIC_OK;
}

namespace {
__global__ void fill_ones(float* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 1.0f;
}
}  // namespace

extern "C" int aldic_images_create_dev(aldic_ctx* ctx, const float* fm, const float* g,
                                       const float* mask, const float* fx, const float* fy,
                                       int64_t M, int64_t N, aldic_images** out) {
  if (!ctx || !fm || !g || !out) return fail(ALDIC_EINVAL, "images_dev: NULL argument");
  if (M < 7 || N < 7) return fail(ALDIC_EINVAL, "images_dev: smaller than 7 x 7");
  if ((fx == nullptr) != (fy == nullptr))
    return fail(ALDIC_EINVAL, "images_dev: give both fx and fy or neither");
  RC_TRY(set_device(ctx));
  auto* im = new aldic_images();
  im->ctx = ctx;
  im->device = ctx->device;
  im->M = M;
  im->N = N;
  int rc = alloc_planes(im);
  if (rc != ALDIC_OK) {
    aldic_images_destroy(im);
    return rc;
  }
  const size_t bytes = sizeof(float) * (size_t)(M * N);
  hipStream_t s = ctx->stream;
  auto body = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(im->fm, fm, bytes, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(im->g, g, bytes, hipMemcpyDeviceToDevice, s));
    if (mask) {
      HIP_TRY(hipMemcpyAsync(im->mask, mask, bytes, hipMemcpyDeviceToDevice, s));
    } else {
      fill_ones<<<1024, 256, 0, s>>>(im->mask, M * N);
      HIP_TRY(hipGetLastError());
    }
    if (fx) {
      HIP_TRY(hipMemcpyAsync(im->fx, fx, bytes, hipMemcpyDeviceToDevice, s));
      HIP_TRY(hipMemcpyAsync(im->fy, fy, bytes, hipMemcpyDeviceToDevice, s));
    } else {
      HIP_TRY(aldic::launch_gradient_f32(im->fm, nullptr, M, N, im->fx, im->fy, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return ALDIC_OK;
  };
  rc = body();
  if (rc != ALDIC_OK) {
    aldic_images_destroy(im);
    return rc;
  }
  *out = im;
  return ALDIC_OK;
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
extern "C" int aldic_plan_create(aldic_ctx* ctx, const int32_t* winsize, int64_t n,
                                 int32_t winsize_per_node, int32_t rect, aldic_plan** out) {
  if (!ctx || !winsize || !out || n < 0) return fail(ALDIC_EINVAL, "plan: bad argument");
  if (n > 0x7fffffff) return fail(ALDIC_EINVAL, "plan: more than 2^31-1 nodes");
  RC_TRY(set_device(ctx));
  auto* p = new aldic_plan();
  p->device = ctx->device;
  const int rc = build_plan(ctx, winsize, n, winsize_per_node, rect, p);
  if (rc != ALDIC_OK) {
    free_plan(p);
    delete p;
    return rc;
  }
  *out = p;
  return ALDIC_OK;
}

extern "C" int aldic_plan_destroy(aldic_plan* plan) {
  if (!plan) return ALDIC_OK;
  (void)hipSetDevice(plan->device);
  free_plan(plan);
  delete plan;
  return ALDIC_OK;
}

// ---------------------------------------------------------------------------
// LocalICGN / Subpb1, device-resident
// ---------------------------------------------------------------------------
namespace {

// The resident planes a solver reads; g is the B-spline coefficient plane
// when the images were switched to that interpolant.
void image_args(SolveArgs& a, const aldic_images* im) {
  a.fm = im->fm;
  a.fx = im->fx;
  a.fy = im->fy;
  a.g = im->interp ? im->gc : im->g;
  a.basis = im->interp;
  a.mask = im->mask;
  a.M = (int)im->M;
  a.N = (int)im->N;
}

SolveArgs icgn_args(const aldic_images* im, int64_t n, const double* d_coords, const float* d_U0,
                    const float* d_F0, const int32_t* d_winsize, double tol, float* d_U,
                    float* d_F, float* d_H21, int32_t* d_convIt, uint8_t* d_hole, float* d_zncc) {
  SolveArgs a{};
  image_args(a, im);
  a.n = (int)n;
  a.coords = d_coords;
  a.wsx = d_winsize;
  a.wsy = nullptr;
  a.ws_per_node = 1;
  a.U0 = d_U0;
  a.FOld = d_F0;
  a.tol = (float)tol;
  a.U = d_U;
  a.F = d_F;
  a.H21 = d_H21;
  a.conv = d_convIt;
  a.zncc = d_zncc;
  a.hole = d_hole;
  return a;
}

// Gate of the learned starting point: nodes whose IC-GN did not converge
// within max_iter evaluations or ends below min_zncc are appended to list
__global__ void gate_kernel(int n, const int32_t* conv, const float* zncc, int max_iter,
                            float min_zncc, int* list, int* count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int st = conv[i];
  const float z = zncc ? zncc[i] : 1.0f;
  // NaN ZNCC fails the gate too
  if (st < 1 || st > max_iter || !(z >= min_zncc)) list[atomicAdd(count, 1)] = i;
}

}  // namespace

extern "C" int aldic_local_icgn_f0_dev(aldic_ctx* ctx, const aldic_images* im,
                                       const aldic_plan* plan, const double* d_coords,
                                       const float* d_U0, const float* d_F0,
                                       const int32_t* d_winsize, double tol, int32_t max_iter,
                                       float* d_U, float* d_F, float* d_H21, int32_t* d_convIt,
                                       uint8_t* d_hole, float* d_zncc) {
  if (max_iter < 0 || max_iter > 100)
    return fail(ALDIC_EINVAL, "local_icgn_f0_dev: max_iter must be in [0, 100]");
  if (!ctx || !images_ok(im) || !plan || plan->rect)
    return fail(ALDIC_EINVAL, "local_icgn_dev: bad context/images/plan");
  if (plan->n == 0) return ALDIC_OK;
  if (!d_coords || !d_U0 || !d_winsize || !d_U || !d_convIt)
    return fail(ALDIC_EINVAL, "local_icgn_dev: NULL array");
  RC_TRY(set_device(ctx));
  SolveArgs a = icgn_args(im, plan->n, d_coords, d_U0, d_F0, d_winsize, tol, d_U, d_F, d_H21,
                          d_convIt, d_hole, d_zncc);
  a.maxit = max_iter;
  return run_solver(ctx, 0, a, plan);
}

extern "C" int aldic_gate_dev(aldic_ctx* ctx, int64_t n, const int32_t* d_convIt,
                              const float* d_zncc, int32_t max_iter, double min_zncc,
                              int32_t* d_list, int32_t* d_count) {
  if (!ctx || !d_convIt || !d_list || !d_count || n < 0)
    return fail(ALDIC_EINVAL, "gate_dev: NULL argument");
  RC_TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  StageTimer t(ctx, ALDIC_STAGE_FALLBACK);
  zero_ints<<<1, 64, 0, s>>>(d_count, 1);
  HIP_TRY(hipGetLastError());
  if (n == 0) return ALDIC_OK;
  gate_kernel<<<(int)((n + 255) / 256), 256, 0, s>>>((int)n, d_convIt, d_zncc, max_iter,
                                                      (float)min_zncc, d_list, d_count);
  HIP_TRY(hipGetLastError());
  return ALDIC_OK;
}

extern "C" int aldic_search_nodes_dev(aldic_ctx* ctx, const aldic_images* im,
                                      const double* d_coords, int64_t n, const int32_t* d_list,
                                      const int32_t* d_count, int64_t max_count, int32_t winsize,
                                      const int64_t region[4], float* d_U0, float* d_F0) {
  if (!ctx || !images_ok(im) || !d_coords || !d_list || !d_count || !d_U0 || !region)
    return fail(ALDIC_EINVAL, "search_nodes_dev: NULL argument");
  if (winsize < 0 || (winsize % 2)) return fail(ALDIC_EINVAL, "search_nodes_dev: winsize must be even");
  for (int k = 0; k < 4; k++)
    if (region[k] < 0) return fail(ALDIC_EINVAL, "search_nodes_dev: negative search region");
  if (aldic::search_smem_bytes(winsize) > 160 * 1024)
    return fail(ALDIC_EINVAL, "search_nodes_dev: winsize too large for one workgroup's LDS");
  if (max_count <= 0 || n == 0) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  if (max_count > n) max_count = n;
  const int rg[4] = {(int)region[0], (int)region[1], (int)region[2], (int)region[3]};
  void* scratch;
  RC_TRY(staging(ctx, SLOT_SEARCH_NODES,
                 aldic::search_scratch_bytes((int)max_count, winsize, rg), &scratch));
  StageTimer t(ctx, ALDIC_STAGE_FALLBACK);
  HIP_TRY(aldic::launch_search_nodes(im->fm, im->g, (int)im->M, (int)im->N, d_coords, (int)n,
                                     d_list, d_count, (int)max_count, winsize, rg, scratch, d_U0,
                                     d_F0, ctx->stream, ctx->search_method));
  return ALDIC_OK;
}

// IC-GN over the listed nodes d_list[0 .. *d_count) of a plan: the nodes
// are sorted into per-size-class lists on the device and run through the
// same wave kernels as the whole plan (a bounded grid walks each class's
// list, the count read on the device), so no host round trip is needed
extern "C" int aldic_local_icgn_list_dev(aldic_ctx* ctx, const aldic_images* im,
                                         const aldic_plan* plan, const double* d_coords,
                                         const float* d_U0, const float* d_F0,
                                         const int32_t* d_winsize, const int32_t* d_list,
                                         const int32_t* d_count, double tol, float* d_U,
                                         float* d_F, float* d_H21, int32_t* d_convIt,
                                         uint8_t* d_hole, float* d_zncc) {
  if (!ctx || !images_ok(im) || !plan || plan->rect)
    return fail(ALDIC_EINVAL, "local_icgn_list_dev: bad context/images/plan");
  if (!d_coords || !d_U0 || !d_winsize || !d_list || !d_count || !d_U || !d_convIt)
    return fail(ALDIC_EINVAL, "local_icgn_list_dev: NULL array");
  if (plan->n == 0) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  SolveArgs a = icgn_args(im, plan->n, d_coords, d_U0, d_F0, d_winsize, tol, d_U, d_F, d_H21,
                          d_convIt, d_hole, d_zncc);
  a.defer_cnt = plan->defer_cnt;
  a.defer_list = plan->defer_list;
  a.stop_rule = plan->stop_rule;
  a.tol_dp = plan->tol_dp;
  StageTimer t(ctx, ALDIC_STAGE_FALLBACK);
  aldic::ClassLists cl;
  for (int c = 0; c < 5; c++) {
    cl.list[c] = plan->sub_lists[c];
    cl.cap[c] = c < 4 ? plan->cnt[c] : (int)plan->n;
  }
  cl.counts = plan->sub_cnt;
  HIP_TRY(aldic::launch_bucket(a, d_list, d_count, (int)plan->n, cl, s));
  zero_ints<<<1, 64, 0, s>>>(plan->defer_cnt, 1);
  HIP_TRY(hipGetLastError());
  // one stream: the lists are short, and a caller pipelining frames over
  // several contexts already fills the hardware queues
  for (int c = 0; c < 4; c++) {
    if (plan->cnt[c] == 0) continue;
    a.list = plan->sub_lists[c];
    a.cnt = plan->cnt[c];
    a.dcnt = plan->sub_cnt + c;
    a.fixed_side = plan->full[c];
    HIP_TRY(aldic::launch_fast(0, c, plan->var[0][c], a, s));
  }
  a.dcnt = nullptr;
  // list 4: the plan's > 65 px subsets and any node whose class list was
  // full (the count is read on the device, clamped to the capacity n)
  HIP_TRY(aldic::launch_generic(0, a, plan->sub_lists[4], plan->sub_cnt + 4, (int)plan->n,
                                std::max(1, std::min(plan->cnt[4], 2048)), s));
  const int grid = (int)std::min<int64_t>(std::max<int64_t>(plan->n, 1), 1024);
  HIP_TRY(aldic::launch_generic(0, a, plan->defer_list, plan->defer_cnt, (int)plan->n, grid, s));
  return ALDIC_OK;
}

extern "C" int aldic_local_icgn_dev(aldic_ctx* ctx, const aldic_images* im, const aldic_plan* plan,
                                    const double* d_coords, const float* d_U0,
                                    const int32_t* d_winsize, double tol, float* d_U, float* d_F,
                                    float* d_H21, int32_t* d_convIt, uint8_t* d_hole,
                                    float* d_zncc) {
  if (!ctx || !images_ok(im) || !plan || plan->rect)
    return fail(ALDIC_EINVAL, "local_icgn_dev: bad context/images/plan");
  if (plan->n == 0) return ALDIC_OK;
  if (!d_coords || !d_U0 || !d_winsize || !d_U || !d_convIt)
    return fail(ALDIC_EINVAL, "local_icgn_dev: NULL array");
  RC_TRY(set_device(ctx));
  SolveArgs a{};
  image_args(a, im);
  a.coords = d_coords;
  a.wsx = d_winsize;
  a.wsy = nullptr;
  // a plan over one winsize value still carries a per-node flag of 0
  a.ws_per_node = 1;
  a.U0 = d_U0;
  a.tol = (float)tol;
  a.U = d_U;
  a.F = d_F;
  a.H21 = d_H21;
  a.conv = d_convIt;
  a.zncc = d_zncc;
  a.hole = d_hole;
  return run_solver(ctx, 0, a, plan);
}

// Autotuner: time every launch shape of every size class on the caller's
// own nodes and keep the fastest in the plan.  The solve it runs is a real
// one (outputs are written); subsets handed on are not part of the timing.
extern "C" int aldic_local_icgn_autotune_dev(aldic_ctx* ctx, const aldic_images* im,
                                             aldic_plan* plan, const double* d_coords,
                                             const float* d_U0, const int32_t* d_winsize,
                                             double tol, float* d_U, float* d_F,
                                             int32_t* d_convIt, float* d_zncc, int32_t reps,
                                             int32_t* chosen) {
  if (!ctx || !images_ok(im) || !plan || plan->rect || reps < 1)
    return fail(ALDIC_EINVAL, "autotune: bad context/images/plan/reps");
  if (!d_coords || !d_U0 || !d_winsize || !d_U || !d_convIt)
    return fail(ALDIC_EINVAL, "autotune: NULL array");
  RC_TRY(set_device(ctx));
  SolveArgs a{};
  image_args(a, im);
  a.n = (int)plan->n;
  a.coords = d_coords;
  a.wsx = d_winsize;
  a.ws_per_node = 1;
  a.U0 = d_U0;
  a.tol = (float)tol;
  a.U = d_U;
  a.F = d_F;
  a.conv = d_convIt;
  a.zncc = d_zncc;
  a.defer_cnt = plan->defer_cnt;
  a.defer_list = plan->defer_list;
  a.stop_rule = plan->stop_rule;
  a.tol_dp = plan->tol_dp;
  hipStream_t s = ctx->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int rc = ALDIC_OK;
  for (int c = 0; c < 4 && rc == ALDIC_OK; c++) {
    if (plan->cnt[c] == 0) continue;
    a.list = plan->lists[c];
    a.cnt = plan->cnt[c];
    a.fixed_side = plan->full[c];
    float best = 1e30f;
    int bestv = 0;
    for (int v = 0; v < aldic::num_variants(); v++) {
      // warm once, then time reps launches
      zero_ints<<<1, 64, 0, s>>>(plan->defer_cnt, 1);
      if (aldic::launch_fast(0, c, v, a, s) != hipSuccess) {
        rc = fail(ALDIC_EDEVICE, "autotune: launch failed");
        break;
      }
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < reps; r++) {
        zero_ints<<<1, 64, 0, s>>>(plan->defer_cnt, 1);
        (void)aldic::launch_fast(0, c, v, a, s);
      }
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) {
        rc = fail(ALDIC_EDEVICE, "autotune: synchronize failed");
        break;
      }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        bestv = v;
      }
    }
    plan->var[0][c] = bestv;
    plan->var[1][c] = bestv;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (chosen)
    for (int c = 0; c < 4; c++) chosen[c] = plan->var[0][c];
  return rc;
}

extern "C" int aldic_plan_set_variant(aldic_plan* plan, int32_t solver, int32_t cls,
                                      int32_t variant) {
  if (!plan || solver < 0 || solver > 1 || cls < 0 || cls > 3 || variant < 0 ||
      variant >= aldic::num_variants())
    return fail(ALDIC_EINVAL, "plan_set_variant: out of range");
  plan->var[solver][cls] = variant;
  return ALDIC_OK;
}

extern "C" int aldic_plan_set_persistent(aldic_plan* plan, int32_t on) {
  if (!plan || (on != 0 && on != 1))
    return fail(ALDIC_EINVAL, "plan_set_persistent: on must be 0 or 1");
  plan->persistent = on;
  return ALDIC_OK;
}

extern "C" int aldic_plan_set_stopping(aldic_plan* plan, int32_t rule, double tol_dp) {
  if (!plan || rule < 0 || rule > 2)
    return fail(ALDIC_EINVAL,
                "plan_set_stopping: rule must be 0 (reference), 1 (step size) or 2 (a-posteriori)");
  if (rule >= 1 && !(tol_dp > 0.0 && tol_dp < 1e3))
    return fail(ALDIC_EINVAL, "plan_set_stopping: the step-size threshold must be positive");
  plan->stop_rule = rule;
  plan->tol_dp = rule >= 1 ? (float)tol_dp : 0.0f;
  return ALDIC_OK;
}

extern "C" int aldic_subpb1_dev(aldic_ctx* ctx, const aldic_images* im, const aldic_plan* plan,
                                const double* d_coords, const float* d_UOld, const float* d_FOld,
                                const float* d_vdual, const int32_t* d_winsize_xy, double mu,
                                double tol, float* d_U, int32_t* d_convIt) {
  if (!ctx || !images_ok(im) || !plan || !plan->rect)
    return fail(ALDIC_EINVAL, "subpb1_dev: bad context/images/plan (plan must be rect)");
  if (plan->n == 0) return ALDIC_OK;
  if (!d_coords || !d_UOld || !d_FOld || !d_vdual || !d_winsize_xy || !d_U || !d_convIt)
    return fail(ALDIC_EINVAL, "subpb1_dev: NULL array");
  RC_TRY(set_device(ctx));
  SolveArgs a{};
  image_args(a, im);
  a.coords = d_coords;
  a.wsx = d_winsize_xy;
  a.wsy = d_winsize_xy + plan->n;
  a.ws_per_node = 1;
  a.UOld = d_UOld;
  a.FOld = d_FOld;
  a.vdual = d_vdual;
  a.mu = (float)mu;
  a.tol = (float)tol;
  a.U = d_U;
  a.conv = d_convIt;
  return run_solver(ctx, 1, a, plan);
}

// ---------------------------------------------------------------------------
// LocalICGN / Subpb1, host pointers
// ---------------------------------------------------------------------------
extern "C" int aldic_local_icgn(aldic_ctx* ctx, const aldic_images* im, const double* coords,
                                const double* U0, int64_t n, const int32_t* winsize,
                                int32_t winsize_per_node, double tol, double* U, double* F,
                                double* H21, int32_t* convIt, uint8_t* markHole, double* zncc) {
  if (!ctx || !images_ok(im) || !coords || !U0 || !winsize || !U || !F || !convIt)
    return fail(ALDIC_EINVAL, "LocalICGNQuadtree: NULL argument");
  if (n < 0) return fail(ALDIC_EINVAL, "LocalICGNQuadtree: n < 0");
  if (n == 0) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  // per-node winsize array on both sides
  std::vector<int32_t> ws((size_t)n);
  for (int64_t i = 0; i < n; i++) ws[i] = winsize_per_node ? winsize[i] : winsize[0];
  aldic_plan plan;
  int rc = build_plan(ctx, ws.data(), n, 1, 0, &plan);
  if (rc != ALDIC_OK) {
    free_plan(&plan);
    return rc;
  }
  std::vector<float> U0f((size_t)(2 * n));
  for (int64_t i = 0; i < 2 * n; i++) U0f[i] = (float)U0[i];
  auto body = [&]() -> int {
    double* dc;
    float *dU0, *dU, *dF, *dH, *dz;
    int32_t *dws, *dconv;
    uint8_t* dhole;
    RC_TRY(stage(ctx, 0, (size_t)(2 * n), &dc));
    RC_TRY(stage(ctx, 1, (size_t)(2 * n), &dU0));
    RC_TRY(stage(ctx, 2, (size_t)n, &dws));
    RC_TRY(stage(ctx, 3, (size_t)(2 * n), &dU));
    RC_TRY(stage(ctx, 4, (size_t)(4 * n), &dF));
    RC_TRY(stage(ctx, 5, (size_t)(21 * n), &dH));
    RC_TRY(stage(ctx, 6, (size_t)n, &dconv));
    RC_TRY(stage(ctx, 7, (size_t)n, &dhole));
    RC_TRY(stage(ctx, 8, (size_t)n, &dz));
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(dc, coords, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dU0, U0f.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dws, ws.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    RC_TRY(aldic_local_icgn_dev(ctx, im, &plan, dc, dU0, dws, tol, dU, dF, dH, dconv, dhole, dz));
    std::vector<float> hU((size_t)(2 * n)), hF((size_t)(4 * n)), hH((size_t)(21 * n)),
        hz((size_t)n);
    HIP_TRY(hipMemcpyAsync(hU.data(), dU, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hF.data(), dF, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hH.data(), dH, sizeof(float) * 21 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hz.data(), dz, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(convIt, dconv, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    std::vector<uint8_t> hh((size_t)n);
    HIP_TRY(hipMemcpyAsync(hh.data(), dhole, n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int64_t i = 0; i < 2 * n; i++) U[i] = hU[i];
    for (int64_t i = 0; i < 4 * n; i++) F[i] = hF[i];
    if (H21)
      for (int64_t i = 0; i < 21 * n; i++) H21[i] = hH[i];
    if (zncc)
      for (int64_t i = 0; i < n; i++) zncc[i] = hz[i];
    if (markHole) std::memcpy(markHole, hh.data(), (size_t)n);
    return ALDIC_OK;
  };
  rc = body();
  (void)hipStreamSynchronize(ctx->stream);
  free_plan(&plan);
  return rc;
}

extern "C" int aldic_subpb1(aldic_ctx* ctx, const aldic_images* im, const double* coords,
                            const double* UOld, const double* FOld, const double* udual,
                            const double* vdual, int64_t n, const int32_t* winsize_xy, double mu,
                            double tol, double* U, int32_t* convIt) {
  (void)udual;  // read by the reference, unused by its first-order update
  if (!ctx || !images_ok(im) || !coords || !UOld || !FOld || !vdual || !winsize_xy || !U ||
      !convIt)
    return fail(ALDIC_EINVAL, "Subpb1Quadtree: NULL argument");
  if (n < 0) return fail(ALDIC_EINVAL, "Subpb1Quadtree: n < 0");
  if (n == 0) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  aldic_plan plan;
  int rc = build_plan(ctx, winsize_xy, n, 1, 1, &plan);
  if (rc != ALDIC_OK) {
    free_plan(&plan);
    return rc;
  }
  std::vector<float> hUO((size_t)(2 * n)), hFO((size_t)(4 * n)), hvd((size_t)(2 * n));
  for (int64_t i = 0; i < 2 * n; i++) {
    hUO[i] = (float)UOld[i];
    hvd[i] = (float)vdual[i];
  }
  for (int64_t i = 0; i < 4 * n; i++) hFO[i] = (float)FOld[i];
  auto body = [&]() -> int {
    double* dc;
    float *dUO, *dFO, *dvd, *dU;
    int32_t *dws, *dconv;
    RC_TRY(stage(ctx, 0, (size_t)(2 * n), &dc));
    RC_TRY(stage(ctx, 1, (size_t)(2 * n), &dUO));
    RC_TRY(stage(ctx, 2, (size_t)(4 * n), &dFO));
    RC_TRY(stage(ctx, 3, (size_t)(2 * n), &dvd));
    RC_TRY(stage(ctx, 4, (size_t)(2 * n), &dws));
    RC_TRY(stage(ctx, 5, (size_t)(2 * n), &dU));
    RC_TRY(stage(ctx, 6, (size_t)n, &dconv));
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(dc, coords, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dUO, hUO.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dFO, hFO.data(), sizeof(float) * 4 * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dvd, hvd.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dws, winsize_xy, sizeof(int32_t) * 2 * n, hipMemcpyHostToDevice, s));
    RC_TRY(aldic_subpb1_dev(ctx, im, &plan, dc, dUO, dFO, dvd, dws, mu, tol, dU, dconv));
    std::vector<float> hU((size_t)(2 * n));
    HIP_TRY(hipMemcpyAsync(hU.data(), dU, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(convIt, dconv, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int64_t i = 0; i < 2 * n; i++) U[i] = hU[i];
    return ALDIC_OK;
  };
  rc = body();
  (void)hipStreamSynchronize(ctx->stream);
  free_plan(&plan);
  return rc;
}

// ---------------------------------------------------------------------------
// integer search
// ---------------------------------------------------------------------------
namespace {

// XList = gridx(1):step:gridx(2)-ws-step (funIntegerSearch.m:106-107)
int64_t grid_count(int64_t a, int64_t b, int64_t ws, int64_t step) {
  const int64_t last = b - ws - step;
  if (step <= 0 || last < a) return 0;
  return (last - a) / step + 1;
}

int search_checks(int64_t M, int64_t N, const int64_t gridx[2], const int64_t gridy[2],
                  int32_t ws, int32_t step, const int64_t region[4]) {
  if (!gridx || !gridy || !region) return fail(ALDIC_EINVAL, "integer search: NULL argument");
  if (ws < 0 || (ws % 2)) return fail(ALDIC_EINVAL, "integer search: winsize must be even");
  if (step <= 0) return fail(ALDIC_EINVAL, "integer search: winstepsize must be positive");
  for (int k = 0; k < 4; k++)
    if (region[k] < 0) return fail(ALDIC_EINVAL, "integer search: negative search region");
  if (gridx[0] < 1 || gridy[0] < 1 || gridx[1] > M || gridy[1] > N)
    return fail(ALDIC_EINVAL, "integer search: ROI outside the image (funNormalizeImg.m:24-28)");
  if (M > 0x7fffffff || N > 0x7fffffff) return fail(ALDIC_EINVAL, "integer search: image too large");
  if (aldic::search_smem_bytes(ws) > 160 * 1024)
    return fail(ALDIC_EINVAL, "integer search: winsize too large for one workgroup's LDS");
  return ALDIC_OK;
}

}  // namespace

// the device search of the grid; d_q (may be null) receives the raw
// (PCE, PPE) pair of every point
static int search_dev(aldic_ctx* ctx, const float* d_f, const float* d_g, int64_t M, int64_t N,
                      const int64_t gridx[2], const int64_t gridy[2], int32_t winsize,
                      int32_t winstepsize, const int64_t region[4], double* d_x, double* d_y,
                      double* d_u, double* d_v, double* d_ccmax, int32_t* d_ipr, int32_t* d_ipc,
                      double* d_q) {
  if (!ctx || !d_f || !d_g) return fail(ALDIC_EINVAL, "integer search: NULL argument");
  RC_TRY(search_checks(M, N, gridx, gridy, winsize, winstepsize, region));
  const int64_t nX = grid_count(gridx[0], gridx[1], winsize, winstepsize);
  const int64_t nY = grid_count(gridy[0], gridy[1], winsize, winstepsize);
  if (nX * nY == 0) return ALDIC_OK;
  if (!d_x || !d_y || !d_u || !d_v || !d_ccmax || !d_ipr || !d_ipc)
    return fail(ALDIC_EINVAL, "integer search: NULL output");
  RC_TRY(set_device(ctx));
  const int gx[2] = {(int)gridx[0], (int)gridx[1]}, gy[2] = {(int)gridy[0], (int)gridy[1]};
  const int rg[4] = {(int)region[0], (int)region[1], (int)region[2], (int)region[3]};
  void* scratch;
  RC_TRY(staging(ctx, SLOT_SEARCH, aldic::search_scratch_bytes((int)(nX * nY), winsize, rg),
                 &scratch));
  StageTimer t(ctx, ALDIC_STAGE_SEARCH);
  HIP_TRY(aldic::launch_search(d_f, d_g, (int)M, (int)N, (int)nX, (int)nY, gx, gy, winsize,
                               winstepsize, rg, scratch, d_x, d_y, d_u, d_v, d_ccmax, d_ipr,
                               d_ipc, ctx->stream, d_q, ctx->search_method));
  return ALDIC_OK;
}

extern "C" int aldic_integer_search_dev(aldic_ctx* ctx, const float* d_f, const float* d_g,
                                        int64_t M, int64_t N, const int64_t gridx[2],
                                        const int64_t gridy[2], int32_t winsize,
                                        int32_t winstepsize, const int64_t region[4], double* d_x,
                                        double* d_y, double* d_u, double* d_v, double* d_ccmax,
                                        int32_t* d_ipr, int32_t* d_ipc) {
  return search_dev(ctx, d_f, d_g, M, N, gridx, gridy, winsize, winstepsize, region, d_x, d_y,
                    d_u, d_v, d_ccmax, d_ipr, d_ipc, nullptr);
}

// The search-init start of the listed nodes: the grid points their
// bilinear transfer reads are found on the device, searched (grid-list
// mode) and interpolated onto the listed nodes only.
extern "C" int aldic_search_grid_nodes_dev(aldic_ctx* ctx, const float* d_f, const float* d_g,
                                           int64_t M, int64_t N, const int64_t gridx[2],
                                           const int64_t gridy[2], int32_t winsize,
                                           int32_t winstepsize, const int64_t region[4],
                                           const double* d_coords, int64_t n,
                                           const int32_t* d_list, const int32_t* d_count,
                                           float* d_U0, float* d_F0) {
  if (!ctx || !d_f || !d_g || !d_coords || !d_list || !d_count || !d_U0)
    return fail(ALDIC_EINVAL, "search_grid_nodes_dev: NULL argument");
  RC_TRY(search_checks(M, N, gridx, gridy, winsize, winstepsize, region));
  const int64_t nX = grid_count(gridx[0], gridx[1], winsize, winstepsize);
  const int64_t nY = grid_count(gridy[0], gridy[1], winsize, winstepsize);
  if (nX * nY == 0 || n == 0) return ALDIC_OK;
  if (n > 0x7fffffff) return fail(ALDIC_EINVAL, "search_grid_nodes_dev: too many nodes");
  RC_TRY(set_device(ctx));
  const int m = (int)(nX * nY);
  const int gx[2] = {(int)gridx[0], (int)gridx[1]}, gy[2] = {(int)gridy[0], (int)gridy[1]};
  const int rg[4] = {(int)region[0], (int)region[1], (int)region[2], (int)region[3]};
  // grid-sized work arrays: flags, point list, count, x y u v cc (double), peaks
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t ib = up(sizeof(int) * m), db = up(sizeof(double) * m);
  const size_t work = 2 * ib + 256 + 5 * db + 2 * ib;
  const size_t sb = aldic::search_scratch_bytes(m, winsize, rg);
  void* base;
  RC_TRY(staging(ctx, SLOT_SEARCH_GRID_NODES, work + sb, &base));
  char* p = static_cast<char*>(base);
  int* flags = reinterpret_cast<int*>(p);
  int* points = reinterpret_cast<int*>(p + ib);
  int* npoints = reinterpret_cast<int*>(p + 2 * ib);
  double* gd = reinterpret_cast<double*>(p + 2 * ib + 256);
  double *x = gd, *y = gd + db / 8, *u = gd + 2 * (db / 8), *v = gd + 3 * (db / 8),
         *cc = gd + 4 * (db / 8);
  int* ipr = reinterpret_cast<int*>(p + 2 * ib + 256 + 5 * db);
  int* ipc = reinterpret_cast<int*>(p + 3 * ib + 256 + 5 * db);
  void* scratch = p + work;
  const double x0 = gridx[0] + winsize / 2.0, y0 = gridy[0] + winsize / 2.0;
  hipStream_t s = ctx->stream;
  StageTimer t(ctx, ALDIC_STAGE_FALLBACK);
  HIP_TRY(aldic::launch_mark_cells((int)nX, (int)nY, x0, y0, (double)winstepsize, d_coords,
                                   (int)n, d_list, d_count, flags, points, npoints, s));
  HIP_TRY(aldic::launch_search_points(d_f, d_g, (int)M, (int)N, (int)nX, (int)nY, gx, gy, winsize,
                                      winstepsize, rg, points, npoints, scratch, x, y, u, v, cc,
                                      ipr, ipc, s, ctx->search_method));
  HIP_TRY(aldic::launch_grid_to_listed_nodes(u, v, (int)nX, (int)nY, x0, y0,
                                             (double)winstepsize, d_coords, (int)n, d_list,
                                             d_count, d_U0, d_F0, s));
  return ALDIC_OK;
}

extern "C" int aldic_grid_to_nodes_dev(aldic_ctx* ctx, const double* d_u, const double* d_v,
                                       int64_t nX, int64_t nY, double x0, double y0, double step,
                                       const double* d_coords, int64_t n, float* d_U0) {
  if (!ctx || !d_u || !d_v || !d_coords || !d_U0)
    return fail(ALDIC_EINVAL, "grid_to_nodes: NULL argument");
  if (nX < 1 || nY < 1 || n < 0 || step <= 0 || n > 0x7fffffff)
    return fail(ALDIC_EINVAL, "grid_to_nodes: bad sizes");
  RC_TRY(set_device(ctx));
  StageTimer t(ctx, ALDIC_STAGE_GRID2NODES);
  HIP_TRY(aldic::launch_grid_to_nodes(d_u, d_v, (int)nX, (int)nY, x0, y0, step, d_coords, (int)n,
                                      d_U0, ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_field_to_nodes_dev(aldic_ctx* ctx, const float* d_Ug, const float* d_Fg,
                                        int64_t nX, int64_t nY, double x0, double y0, double step,
                                        const double* d_coords, int64_t n, float* d_U0,
                                        float* d_F0) {
  if (!ctx || !d_Ug || !d_coords || !d_U0 || (d_F0 && !d_Fg))
    return fail(ALDIC_EINVAL, "field_to_nodes: NULL argument");
  if (nX < 1 || nY < 1 || n < 0 || step <= 0 || n > 0x7fffffff || nX * nY > 0x7fffffff)
    return fail(ALDIC_EINVAL, "field_to_nodes: bad sizes");
  RC_TRY(set_device(ctx));
  StageTimer t(ctx, ALDIC_STAGE_GRID2NODES);
  HIP_TRY(aldic::launch_field_to_nodes(d_Ug, d_Fg, (int)nX, (int)nY, x0, y0, step, d_coords,
                                       (int)n, d_U0, d_F0, ctx->stream));
  return ALDIC_OK;
}

// The same search with its per-point scratch in the file-scope array
// `scratch` of this GPU instead of a context buffer: no allocation at all.
extern "C" int aldic_integer_search_noalloc_dev(aldic_ctx* ctx, const float* d_f,
                                                const float* d_g, int64_t M, int64_t N,
                                                const int64_t gridx[2], const int64_t gridy[2],
                                                int32_t winsize, int32_t winstepsize,
                                                const int64_t region[4], double* d_x,
                                                double* d_y, double* d_u, double* d_v,
                                                double* d_ccmax, int32_t* d_ipr,
                                                int32_t* d_ipc) {
  if (!ctx || !d_f || !d_g) return fail(ALDIC_EINVAL, "integer search (no alloc): NULL argument");
  RC_TRY(search_checks(M, N, gridx, gridy, winsize, winstepsize, region));
  const int64_t nX = grid_count(gridx[0], gridx[1], winsize, winstepsize);
  const int64_t nY = grid_count(gridy[0], gridy[1], winsize, winstepsize);
  if (nX * nY == 0) return ALDIC_OK;
  if (!d_x || !d_y || !d_u || !d_v || !d_ccmax || !d_ipr || !d_ipc)
    return fail(ALDIC_EINVAL, "integer search (no alloc): NULL output");
  const int gx[2] = {(int)gridx[0], (int)gridx[1]}, gy[2] = {(int)gridy[0], (int)gridy[1]};
  const int rg[4] = {(int)region[0], (int)region[1], (int)region[2], (int)region[3]};
  const size_t need = aldic::search_scratch_bytes((int)(nX * nY), winsize, rg);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(ALDIC_EINVAL,
                "integer search (no alloc): " + std::to_string(nX * nY) + " points at winsize " +
                    std::to_string(winsize) + " need " + std::to_string(need) +
                    " bytes of scratch, more than the " +
                    std::to_string(NOALLOC_SCRATCH_BYTES) +
                    " this entry point holds; split the grid or call aldic_integer_search_dev");
  RC_TRY(set_device(ctx));
  void* sp = nullptr;
  RC_TRY(scratch_address(ctx->device, &sp));
  // the prep kernel writes every point's clamps and template and the tile
  // kernel every (point, tile) best before they are read: nothing in the
  // scratch needs clearing between calls
  HIP_TRY(aldic::launch_search(d_f, d_g, (int)M, (int)N, (int)nX, (int)nY, gx, gy, winsize,
                               winstepsize, rg, sp, d_x, d_y, d_u, d_v, d_ccmax, d_ipr, d_ipc,
                               ctx->stream, nullptr, ctx->search_method));
  return ALDIC_OK;
}

extern "C" int aldic_integer_search(aldic_ctx* ctx, const double* f, const double* g, int64_t M,
                                    int64_t N, const int64_t gridx[2], const int64_t gridy[2],
                                    int32_t winsize, int32_t winstepsize, const int64_t region[4],
                                    double* x, double* y, double* u, double* v, double* ccmax,
                                    int32_t* ipr, int32_t* ipc, double* qfactors,
                                    int64_t* nX_out, int64_t* nY_out) {
  if (!ctx) return fail(ALDIC_EINVAL, "integer search: NULL context");
  RC_TRY(search_checks(M, N, gridx, gridy, winsize, winstepsize, region));
  const int64_t nX = grid_count(gridx[0], gridx[1], winsize, winstepsize);
  const int64_t nY = grid_count(gridy[0], gridy[1], winsize, winstepsize);
  if (nX_out) *nX_out = nX;
  if (nY_out) *nY_out = nY;
  if (!x && !y && !u && !v) return ALDIC_OK;  // size query
  if (!f || !g || !x || !y || !u || !v || !ccmax)
    return fail(ALDIC_EINVAL, "integer search: NULL argument");
  const int64_t np = nX * nY;
  if (np == 0) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  const int64_t MN = M * N;
  double *df, *dg, *dx, *dy, *du, *dv, *dc;
  float *ff, *fg;
  int32_t *dpr, *dpc;
  RC_TRY(stage(ctx, 0, (size_t)MN, &df));
  RC_TRY(stage(ctx, 1, (size_t)MN, &dg));
  RC_TRY(stage(ctx, 2, (size_t)MN, &ff));
  RC_TRY(stage(ctx, 3, (size_t)MN, &fg));
  RC_TRY(stage(ctx, 4, (size_t)np, &dx));
  RC_TRY(stage(ctx, 5, (size_t)np, &dy));
  RC_TRY(stage(ctx, 6, (size_t)np, &du));
  RC_TRY(stage(ctx, 7, (size_t)np, &dv));
  RC_TRY(stage(ctx, 8, (size_t)np, &dc));
  RC_TRY(stage(ctx, 9, (size_t)np, &dpr));
  RC_TRY(stage(ctx, 10, (size_t)np, &dpc));
  double* dq = nullptr;
  if (qfactors) RC_TRY(stage(ctx, 11, (size_t)(2 * np), &dq));
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(df, f, sizeof(double) * MN, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dg, g, sizeof(double) * MN, hipMemcpyHostToDevice, s));
  HIP_TRY(aldic::launch_prep_planes(df, dg, nullptr, M, N, ff, fg, nullptr, s));
  RC_TRY(search_dev(ctx, ff, fg, M, N, gridx, gridy, winsize, winstepsize, region, dx, dy, du,
                    dv, dc, dpr, dpc, dq));
  std::vector<double> raw;
  if (qfactors) {
    raw.resize(2 * np);
    HIP_TRY(hipMemcpyAsync(raw.data(), dq, sizeof(double) * 2 * np, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipMemcpyAsync(x, dx, sizeof(double) * np, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(y, dy, sizeof(double) * np, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(u, du, sizeof(double) * np, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(v, dv, sizeof(double) * np, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(ccmax, dc, sizeof(double) * np, hipMemcpyDeviceToHost, s));
  if (ipr) HIP_TRY(hipMemcpyAsync(ipr, dpr, sizeof(int32_t) * np, hipMemcpyDeviceToHost, s));
  if (ipc) HIP_TRY(hipMemcpyAsync(ipc, dpc, sizeof(int32_t) * np, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (qfactors) {
    // cc.qfactors(:,k) = (q - min q) / max(q - min q) over the grid, NaN
    // left out of min and max (funIntegerSearch.m:184-187); an Inf (a point
    // whose normxcorr2 threw) makes the others 0 and itself NaN, as there
    for (int k = 0; k < 2; k++) {
      double mn = NAN, mx = NAN;
      for (int64_t t = 0; t < np; t++) {
        const double q = raw[2 * t + k];
        if (!std::isnan(q) && (std::isnan(mn) || q < mn)) mn = q;
      }
      for (int64_t t = 0; t < np; t++) {
        const double q = raw[2 * t + k] - mn;
        if (!std::isnan(q) && (std::isnan(mx) || q > mx)) mx = q;
      }
      for (int64_t t = 0; t < np; t++) qfactors[t + k * np] = (raw[2 * t + k] - mn) / mx;
    }
  }
  return ALDIC_OK;
}

// ---------------------------------------------------------------------------
// One iteration of the augmented-Lagrangian loop on the device
// (TemporalMatch_quadtree_ST1.m:337-387): Subpb1 from the global estimate,
// the bad-point rule of Subpb1Quadtree.m:167-182, Subpb2, the hole-edge
// copy, the UpdateY norms, the duals and the early exit, with no host
// read-back, so that a frame's iterations can be recorded as one graph.
// ---------------------------------------------------------------------------
namespace {

constexpr int SLOT_ALM = 30;  // .. 36

__global__ void alm_start_kernel(int first, int32_t* state, double* update) {
  if (first) {
    state[0] = 0;           // done
    update[0] = 1e4;        // UpdateY before any is computed (:333)
    update[1] = __builtin_nan("");
  }
  if (!state[0]) state[1] = 0;  // bad Subpb1 nodes of this iteration
}

__global__ void alm_prep_kernel(int64_t n, const int32_t* state, const double* U2,
                                const double* F1, const double* vdual, float* UOld, float* FOld,
                                float* vd) {
  if (state[0]) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 4 * n;
       i += (int64_t)gridDim.x * blockDim.x) {
    FOld[i] = (float)F1[i];
    if (i < 2 * n) {
      UOld[i] = (float)U2[i];
      vd[i] = (float)vdual[i];
    }
  }
}

// the Subpb1 bad-point threshold max(mean + 0.25 std, 10) over the counts in
// [0, 99] (unbiased std; 10 when there are none), one workgroup
__global__ __launch_bounds__(1024) void alm_badstat_kernel(int64_t n, const int32_t* state,
                                                           const int32_t* conv, double* thr) {
  __shared__ double s1[1024], s2[1024], sc[1024];
  if (state[0]) return;
  double a = 0.0, b = 0.0, c = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) {
    const int v = conv[i];
    if (v >= 0 && v <= 99) {
      a += v;
      b += (double)v * v;
      c += 1.0;
    }
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      s1[threadIdx.x] += s1[threadIdx.x + h];
      s2[threadIdx.x] += s2[threadIdx.x + h];
      sc[threadIdx.x] += sc[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double cnt = sc[0];
    double t = 10.0;
    if (cnt > 0.0) {
      const double m = s1[0] / cnt;
      const double var = cnt > 1.0 ? fmax((s2[0] - cnt * m * m) / (cnt - 1.0), 0.0) : 0.0;
      t = fmax(m + 0.25 * sqrt(var), 10.0);
    }
    thr[0] = t;
  }
}

// USubpb1: the Subpb1 result in FP64; a bad or thrown (NaN) node takes the
// global step's estimate it started from
__global__ void alm_fill_kernel(int64_t n, int32_t* state, const int32_t* conv, const double* thr,
                                const float* U1f, const double* U2, double* U1) {
  if (state[0]) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = conv[i];
    const float u = U1f[2 * i], v = U1f[2 * i + 1];
    const bool bad = c < 0 || c > 99 || (double)c > thr[0] || u != u || v != v;
    U1[2 * i] = bad ? U2[2 * i] : (double)u;
    U1[2 * i + 1] = bad ? U2[2 * i + 1] : (double)v;
    if (bad) atomicAdd(&state[1], 1);
  }
}

__global__ void alm_hole_kernel(int64_t n, const int32_t* state, const uint8_t* hole,
                                const double* U1, double* U2n) {
  if (state[0]) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (hole[i]) {
      U2n[2 * i] = U1[2 * i];
      U2n[2 * i + 1] = U1[2 * i + 1];
    }
}

// UpdateY = ||U2_prev - U2|| / sqrt(2n), UpdateY2 = ||U1_prev - U1|| / sqrt(2n)
// (:364-366), one workgroup
__global__ __launch_bounds__(1024) void alm_norm_kernel(int64_t n, const int32_t* state,
                                                        int check, const double* U2p,
                                                        const double* U2n, const double* U1p,
                                                        const double* U1, double* update) {
  __shared__ double sa[1024], sb[1024];
  if (state[0] || !check) return;
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < 2 * n; i += 1024) {
    const double d2 = U2p[i] - U2n[i], d1 = U1p[i] - U1[i];
    a += d2 * d2;
    b += d1 * d1;
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      sa[threadIdx.x] += sa[threadIdx.x + h];
      sb[threadIdx.x] += sb[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    update[0] = sqrt(sa[0]) / sqrt((double)(2 * n));
    update[1] = sqrt(sb[0]) / sqrt((double)(2 * n));
  }
}

// the iteration's results and the duals vdual = U2 - U1 (:371, :379)
__global__ void alm_commit_kernel(int64_t n, const int32_t* state, const double* U1,
                                  const double* U2n, double* dU1, double* dU2, double* vdual) {
  if (state[0]) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 2 * n;
       i += (int64_t)gridDim.x * blockDim.x) {
    dU1[i] = U1[i];
    dU2[i] = U2n[i];
    vdual[i] = U2n[i] - U1[i];
  }
}

// the early exit of :381-384 (UpdateY2 exists only where UpdateY was
// computed, so without the norms the loop does not stop)
__global__ void alm_exit_kernel(int32_t* state, int check, double tol2, const double* update) {
  if (state[0] || !check) return;
  if (update[0] < tol2 || update[1] < tol2) state[0] = 1;
}

int grid_for(int64_t work) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, 2048));
}

}  // namespace

extern "C" int aldic_alm_iteration_dev(aldic_ctx* ctx, const aldic_images* im,
                                       const aldic_plan* plan, aldic_fem* fem,
                                       const double* d_coords, const int32_t* d_winsize_xy,
                                       double mu, double beta, double alpha, double winstepsize,
                                       double tol, double tol2, int32_t check_update,
                                       int32_t first, const double* d_F1, const uint8_t* d_hole,
                                       double* d_U1, double* d_U2, double* d_vdual,
                                       double* d_update, int32_t* d_state, int32_t cg_max_iter) {
  if (!ctx || !images_ok(im) || !plan || !plan->rect || !fem)
    return fail(ALDIC_EINVAL, "alm_iteration_dev: bad context/images/plan (rect)/mesh");
  const int64_t n = plan->n;
  if (aldic::fem_nodes(fem) != n)
    return fail(ALDIC_EINVAL, "alm_iteration_dev: the plan and the mesh hold different node counts");
  if (!d_coords || !d_winsize_xy || !d_F1 || !d_U1 || !d_U2 || !d_vdual || !d_update || !d_state)
    return fail(ALDIC_EINVAL, "alm_iteration_dev: NULL array");
  if (cg_max_iter < 1 || cg_max_iter > (1 << 16))
    return fail(ALDIC_EINVAL, "alm_iteration_dev: cg_max_iter must be in 1 .. 65536");
  if (n == 0) return ALDIC_OK;
  RC_TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  float *UOld, *FOld, *vd, *U1f;
  int32_t* conv;
  double *U1, *U2n, *thr;
  RC_TRY(stage(ctx, SLOT_ALM, (size_t)(2 * n), &UOld));
  RC_TRY(stage(ctx, SLOT_ALM + 1, (size_t)(4 * n), &FOld));
  RC_TRY(stage(ctx, SLOT_ALM + 2, (size_t)(2 * n), &vd));
  RC_TRY(stage(ctx, SLOT_ALM + 3, (size_t)(2 * n), &U1f));
  RC_TRY(stage(ctx, SLOT_ALM + 4, (size_t)n, &conv));
  RC_TRY(stage(ctx, SLOT_ALM + 5, (size_t)(4 * n + 1), &U1));
  U2n = U1 + 2 * n;
  thr = U1 + 4 * n;
  const int g2 = grid_for(4 * n);
  alm_start_kernel<<<1, 1, 0, s>>>(first ? 1 : 0, d_state, d_update);
  alm_prep_kernel<<<g2, 256, 0, s>>>(n, d_state, d_U2, d_F1, d_vdual, UOld, FOld, vd);
  HIP_TRY(hipGetLastError());
  // Subpb1 from [FSubpb2 USubpb2] (Subpb1Quadtree.m:71-165); skipped once done
  SolveArgs a{};
  image_args(a, im);
  a.coords = d_coords;
  a.wsx = d_winsize_xy;
  a.wsy = d_winsize_xy + n;
  a.ws_per_node = 1;
  a.UOld = UOld;
  a.FOld = FOld;
  a.vdual = vd;
  a.mu = (float)mu;
  a.tol = (float)tol;
  a.U = U1f;
  a.conv = conv;
  a.skip = d_state;
  RC_TRY(run_solver(ctx, 1, a, plan));
  alm_badstat_kernel<<<1, 1024, 0, s>>>(n, d_state, conv, thr);
  alm_fill_kernel<<<grid_for(n), 256, 0, s>>>(n, d_state, conv, thr, U1f, d_U2, U1);
  HIP_TRY(hipGetLastError());
  // Subpb2 from USubpb1 (FSubpb1 = FSubpb2, the duals' F part is zero)
  RC_TRY(aldic::subpb2_async(ctx, fem, beta, mu, alpha, winstepsize, U1, d_F1, U2n, cg_max_iter,
                             nullptr, d_state));
  if (d_hole) alm_hole_kernel<<<grid_for(n), 256, 0, s>>>(n, d_state, d_hole, U1, U2n);
  alm_norm_kernel<<<1, 1024, 0, s>>>(n, d_state, check_update ? 1 : 0, d_U2, U2n, d_U1, U1,
                                     d_update);
  alm_commit_kernel<<<grid_for(2 * n), 256, 0, s>>>(n, d_state, U1, U2n, d_U1, d_U2, d_vdual);
  alm_exit_kernel<<<1, 1, 0, s>>>(d_state, check_update ? 1 : 0, tol2, d_update);
  HIP_TRY(hipGetLastError());
  return ALDIC_OK;
}
