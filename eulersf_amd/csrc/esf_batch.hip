// esf_batch.hip -- esf_step_batch: one coupled substep of several scenes,
// every kernel of the step launched once over the scenes that reach it
// together.
//
// Each scene's esf_step runs in a fiber (esf_fiber.h) on the calling
// thread.  A fiber runs until its step launches a kernel (it deposits the
// launch and yields) or waits for the device (it yields until its event
// has completed).  When every fiber has deposited or is waiting, the
// scheduler groups the deposits by kernel instance and block size and
// launches each group as one kMany over all its scenes on the batch
// stream, with the scenes' argument records staged through a pinned ring
// and one copy per launch.  The scenes' steps run the same code, so they
// reach the same launches together: one launch per kernel of the step,
// and per Krylov iteration while several scenes iterate.  A scene's
// kernels compute what they would in its own launches (esf_launch.h), and
// every host decision of its step reads the same device values, so each
// scene's results are bitwise those of esf_step on that scene alone.
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../host/esf_fiber.h"
#include "esf_internal.h"

namespace esf {

struct SceneFiber {
  Fiber f;
  esf_ctx* ctx = nullptr;
  enum State { RUNNABLE, DEPOSITED, WAITING, FINISHED } state = RUNNABLE;
  hipEvent_t wait = nullptr;  // WAITING: the event it waits for
  hipError_t waitResult = hipSuccess;
  hipEvent_t own = nullptr;   // recorded by streamSync
  double alpha = -1;
  esf_solve_stats stats{};
  esf_status status = ESF_OK;
  FiberOp op;
};

constexpr size_t kStageChunk = 4u << 20;  // bytes per chunk of the staging ring
constexpr int kStageChunks = 4;
constexpr size_t kFiberStack = 2u << 20;

struct BatchRun {
  hipStream_t stream = nullptr;
  unsigned char* host = nullptr;  // pinned staging ring
  unsigned char* dev = nullptr;   // its device mirror
  size_t off = 0;
  int chunk = 0;
  hipEvent_t chunkDone[kStageChunks] = {};
  bool chunkPending[kStageChunks] = {};
  hipError_t err = hipSuccess;
  long launches = 0, scenesLaunched = 0;
};

// per device: the staging ring and the fibers, kept between calls; one
// batch at a time per device
struct BatchPool {
  std::mutex m;
  BatchRun run;
  std::vector<std::unique_ptr<SceneFiber>> fibers;
  bool ready = false;
};

static std::mutex gPoolsMutex;
static std::unordered_map<int, std::unique_ptr<BatchPool>> gPools;

static BatchPool* poolFor(int device) {
  std::lock_guard<std::mutex> g(gPoolsMutex);
  auto& p = gPools[device];
  if (!p) p.reset(new BatchPool);
  return p.get();
}

// ---------------------------------------------------------------------------
// the fibers' side
// ---------------------------------------------------------------------------
FiberOp* fiberOp(esf_ctx* ctx) { return &ctx->fiber->op; }

void fiberDeposit(esf_ctx* ctx) {
  SceneFiber* s = ctx->fiber;
  s->state = SceneFiber::DEPOSITED;
  esf_fiber_yield(&s->f);
}

hipError_t fiberWait(esf_ctx* ctx, hipEvent_t ev) {
  const hipError_t q = hipEventQuery(ev);
  if (q != hipErrorNotReady) return q;
  SceneFiber* s = ctx->fiber;
  s->wait = ev;
  s->state = SceneFiber::WAITING;
  esf_fiber_yield(&s->f);
  return s->waitResult;
}

hipEvent_t fiberEvent(esf_ctx* ctx) { return ctx->fiber->own; }

// ---------------------------------------------------------------------------
// staging ring
// ---------------------------------------------------------------------------
unsigned char* stageAlloc(BatchRun& run, size_t bytes, const unsigned char** dev) {
  bytes = (bytes + 255) & ~size_t(255);
  if (bytes > kStageChunk) {
    // more scenes in one launch than a chunk holds (thousands): refused in
    // esf_step_batch, never reached
    run.err = hipErrorInvalidValue;
    bytes = kStageChunk;
  }
  if (run.off + bytes > (size_t)(run.chunk + 1) * kStageChunk) {
    // the copies out of this chunk end before the event recorded here;
    // the next chunk is reused once the copies out of it have ended
    hipEventRecord(run.chunkDone[run.chunk], run.stream);
    run.chunkPending[run.chunk] = true;
    run.chunk = (run.chunk + 1) % kStageChunks;
    run.off = (size_t)run.chunk * kStageChunk;
    if (run.chunkPending[run.chunk]) {
      hipEventSynchronize(run.chunkDone[run.chunk]);
      run.chunkPending[run.chunk] = false;
    }
  }
  unsigned char* h = run.host + run.off;
  *dev = run.dev + run.off;
  run.off += bytes;
  return h;
}

void stageCopy(BatchRun& run, const unsigned char* host, const unsigned char* dev, size_t bytes) {
  const hipError_t e =
      hipMemcpyAsync((void*)dev, host, bytes, hipMemcpyHostToDevice, run.stream);
  if (e != hipSuccess) run.err = e;
}

hipStream_t batchStream(BatchRun& run) { return run.stream; }

void batchLaunchFailed(BatchRun& run, hipError_t e) { run.err = e; }

// ---------------------------------------------------------------------------
// the scheduler
// ---------------------------------------------------------------------------
static void sceneMain(void* p) {
  SceneFiber* s = (SceneFiber*)p;
  s->status = esf_step(s->ctx, s->alpha, &s->stats);
}

static void schedule(BatchRun& run, const std::vector<SceneFiber*>& fs) {
  std::vector<SceneFiber*> dep;
  std::vector<FiberOp*> group;
  std::vector<char> used;
  for (;;) {
    for (SceneFiber* s : fs)
      if (s->state == SceneFiber::RUNNABLE) {
        esf_fiber_resume(&s->f);
        if (s->f.finished) s->state = SceneFiber::FINISHED;
      }
    dep.clear();
    for (SceneFiber* s : fs)
      if (s->state == SceneFiber::DEPOSITED) dep.push_back(s);
    if (!dep.empty()) {
      // one launch per (kernel instance, block size), scenes in order
      used.assign(dep.size(), 0);
      for (size_t i = 0; i < dep.size(); i++) {
        if (used[i]) continue;
        group.clear();
        for (size_t j = i; j < dep.size(); j++)
          if (!used[j] && dep[j]->op.key == dep[i]->op.key &&
              dep[j]->op.threads == dep[i]->op.threads) {
            group.push_back(&dep[j]->op);
            used[j] = 1;
          }
        group[0]->submit(run, group.data(), (int)group.size());
        run.launches++;
        run.scenesLaunched += (long)group.size();
      }
      for (SceneFiber* s : dep) s->state = SceneFiber::RUNNABLE;
      continue;
    }
    // nothing deposited: every live fiber waits for the device
    SceneFiber* oldest = nullptr;
    bool woke = false;
    for (SceneFiber* s : fs) {
      if (s->state != SceneFiber::WAITING) continue;
      const hipError_t q = hipEventQuery(s->wait);
      if (q == hipErrorNotReady) {
        if (!oldest) oldest = s;
        continue;
      }
      s->waitResult = q;
      s->state = SceneFiber::RUNNABLE;
      woke = true;
    }
    if (woke) continue;
    if (!oldest) return;  // all finished
    oldest->waitResult = hipEventSynchronize(oldest->wait);
    oldest->state = SceneFiber::RUNNABLE;
  }
}

static esf_status prepare(BatchPool& pool, int n, std::string& why) {
  BatchRun& run = pool.run;
  if (!pool.ready) {
    if (hipHostMalloc((void**)&run.host, kStageChunk * kStageChunks, hipHostMallocDefault) !=
            hipSuccess ||
        hipMalloc((void**)&run.dev, kStageChunk * kStageChunks) != hipSuccess) {
      why = "batch staging: out of memory";
      return ESF_ERR_OOM;
    }
    for (int k = 0; k < kStageChunks; k++)
      if (hipEventCreateWithFlags(&run.chunkDone[k], hipEventDisableTiming) != hipSuccess) {
        why = "batch staging: event";
        return ESF_ERR_DEVICE;
      }
    pool.ready = true;
  }
  while ((int)pool.fibers.size() < n) {
    std::unique_ptr<SceneFiber> s(new SceneFiber);
    if (!esf_fiber_create(&s->f, kFiberStack, sceneMain, s.get())) {
      why = "batch: cannot map a fiber stack";
      return ESF_ERR_OOM;
    }
    if (hipEventCreateWithFlags(&s->own, hipEventDisableTiming) != hipSuccess) {
      why = "batch: event";
      return ESF_ERR_DEVICE;
    }
    pool.fibers.push_back(std::move(s));
  }
  return ESF_OK;
}

}  // namespace esf

using namespace esf;

extern "C" esf_status esf_step_batch(esf_ctx** ctxs, int n, double alpha,
                                     esf_solve_stats* stats, esf_status* status) {
  if (!ctxs || n < 1) return ESF_ERR_INVALID;
  for (int i = 0; i < n; i++) {
    if (!ctxs[i]) return ESF_ERR_INVALID;
    if (ctxs[i]->fiber) {
      ctxs[i]->err = "context is already stepping in a batch";
      return ESF_ERR_STATE;
    }
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) {
        ctxs[i]->err = "a context appears twice in one batch";
        return ESF_ERR_INVALID;
      }
  }
  // the argument records of one launch must fit one staging chunk
  if ((size_t)n * (kMaxPack + 8) + 256 > kStageChunk) {
    ctxs[0]->err = "too many scenes in one batch";
    return ESF_ERR_INVALID;
  }
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return ESF_ERR_DEVICE;
  BatchPool* pool = poolFor(device);
  std::lock_guard<std::mutex> g(pool->m);
  std::string why;
  esf_status st = prepare(*pool, n, why);
  if (st) {
    ctxs[0]->err = why;
    return st;
  }
  BatchRun& run = pool->run;
  run.stream = ctxs[0]->stream;
  run.off = 0;
  run.chunk = 0;
  for (bool& p : run.chunkPending) p = false;
  run.err = hipSuccess;
  // all scenes' work on the batch stream; timing off for the batch
  struct Saved {
    hipStream_t stream;
    bool kernelTiming, stageTiming;
  };
  std::vector<Saved> saved(n);
  std::vector<SceneFiber*> fs(n);
  for (int i = 0; i < n; i++) {
    esf_ctx* c = ctxs[i];
    saved[i] = Saved{c->stream, c->kernelTiming, c->stageTiming};
    SceneFiber* s = pool->fibers[i].get();
    esf_fiber_reset(&s->f, sceneMain, s);
    s->ctx = c;
    s->alpha = alpha;
    s->state = SceneFiber::RUNNABLE;
    s->stats = esf_solve_stats{};
    s->status = ESF_OK;
    c->stream = run.stream;
    c->kernelTiming = false;
    c->stageTiming = false;
    c->fiber = s;
    fs[i] = s;
  }
  // the batch stream starts after the work already queued on each
  // context's own stream, and each stream's later work after the batch
  for (int i = 1; i < n; i++)
    if (saved[i].stream != run.stream) {
      hipEventRecord(fs[i]->own, saved[i].stream);
      hipStreamWaitEvent(run.stream, fs[i]->own, 0);
    }
  schedule(run, fs);
  esf_status worst = ESF_OK;
  for (int i = 0; i < n; i++) {
    esf_ctx* c = ctxs[i];
    if (saved[i].stream != run.stream) {
      hipEventRecord(fs[i]->own, run.stream);
      hipStreamWaitEvent(saved[i].stream, fs[i]->own, 0);
    }
    c->stream = saved[i].stream;
    c->kernelTiming = saved[i].kernelTiming;
    c->stageTiming = saved[i].stageTiming;
    c->fiber = nullptr;
    if (stats) stats[i] = fs[i]->stats;
    if (status) status[i] = fs[i]->status;
    const esf_status si = fs[i]->status;
    if (si < 0 && (worst >= 0 || si < worst)) worst = si;
    if (si > 0 && worst == ESF_OK) worst = si;
  }
  if (run.err != hipSuccess) {
    ctxs[0]->err = std::string("batched launch: ") + hipGetErrorString(run.err);
    return ESF_ERR_DEVICE;
  }
  return worst;
}
