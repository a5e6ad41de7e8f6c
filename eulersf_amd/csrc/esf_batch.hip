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
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../host/esf_fiber.h"
#include "esf_internal.h"

namespace esf {

const char kMarkKeyObj = 0;

struct BatchRun;

struct SceneFiber {
  Fiber f;
  esf_ctx* ctx = nullptr;
  BatchRun* run = nullptr;
  enum State { RUNNABLE, DEPOSITED, WAITING, BARRIER, FINISHED } state = RUNNABLE;
  long depRound = 0;          // DEPOSITED: the round of the deposit
  hipEvent_t wait = nullptr;  // WAITING: the event it waits for
  hipError_t waitResult = hipSuccess;
  hipEvent_t own = nullptr;   // orders the batch stream with the context's own
  double alpha = -1;
  esf_solve_stats stats{};
  esf_status status = ESF_OK;
  FiberOp op;
};

constexpr size_t kStageChunk = 4u << 20;  // bytes per chunk of the staging ring
constexpr int kStageChunks = 4;
constexpr size_t kFiberStack = 2u << 20;
constexpr int kMarkEvents = 256;  // events the scheduler records for the fibers' marks

struct BatchRun {
  hipStream_t stream = nullptr;
  unsigned char* host = nullptr;  // pinned staging ring
  unsigned char* dev = nullptr;   // its device mirror
  size_t off = 0;
  int chunk = 0;
  hipEvent_t chunkDone[kStageChunks] = {};
  bool chunkPending[kStageChunks] = {};
  hipError_t err = hipSuccess;
  // mark events, recorded round-robin: re-recording one that a fiber still
  // holds only makes that fiber wait longer, never less; `known` caches the
  // last query of each (round, result), reset when it is recorded again
  hipEvent_t marks[kMarkEvents] = {};
  long markRound[kMarkEvents] = {};
  hipError_t markKnown[kMarkEvents] = {};
  int nextMark = 0;
  std::unordered_map<hipEvent_t, int> markIndex;
  // host time of the batch (ESF_BATCH_TRACE=1 prints it per call)
  long rounds = 0, launches = 0, scenesLaunched = 0, waits = 0, markRecords = 0;
  double tScenes = 0, tSubmit = 0, tWait = 0;
};

static double nowS() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// the staging ring and the fibers of one batch in flight, kept between
// calls; each batch in flight (a calling thread) takes a pool of its device
// that no other batch holds
struct BatchPool {
  int device = 0;
  bool inUse = false;
  BatchRun run;
  std::vector<std::unique_ptr<SceneFiber>> fibers;
  bool ready = false;
};

static std::mutex gPoolsMutex;
static std::vector<std::unique_ptr<BatchPool>> gPools;

static BatchPool* takePool(int device) {
  std::lock_guard<std::mutex> g(gPoolsMutex);
  for (auto& p : gPools)
    if (p->device == device && !p->inUse) {
      p->inUse = true;
      return p.get();
    }
  gPools.emplace_back(new BatchPool);
  gPools.back()->device = device;
  gPools.back()->inUse = true;
  return gPools.back().get();
}

static void releasePool(BatchPool* p) {
  std::lock_guard<std::mutex> g(gPoolsMutex);
  p->inUse = false;
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

// the result of querying `ev` in this round (one query per mark event per
// round; a completed mark stays completed until it is recorded again)
static hipError_t cachedQuery(BatchRun& run, hipEvent_t ev) {
  auto it = run.markIndex.find(ev);
  if (it == run.markIndex.end()) return hipEventQuery(ev);
  const int k = it->second;
  if (run.markKnown[k] == hipSuccess || run.markRound[k] == run.rounds) return run.markKnown[k];
  run.markKnown[k] = hipEventQuery(ev);
  run.markRound[k] = run.rounds;
  return run.markKnown[k];
}

hipError_t fiberQuery(esf_ctx* ctx, hipEvent_t ev) { return cachedQuery(*ctx->fiber->run, ev); }

hipError_t fiberWait(esf_ctx* ctx, hipEvent_t ev) {
  const hipError_t q = fiberQuery(ctx, ev);
  if (q != hipErrorNotReady) return q;
  SceneFiber* s = ctx->fiber;
  s->wait = ev;
  s->state = SceneFiber::WAITING;
  esf_fiber_yield(&s->f);
  return s->waitResult;
}

void fiberBarrier(esf_ctx* ctx) {
  SceneFiber* s = ctx->fiber;
  s->state = SceneFiber::BARRIER;
  esf_fiber_yield(&s->f);
}

hipEvent_t fiberMark(esf_ctx* ctx) {
  SceneFiber* s = ctx->fiber;
  s->op.key = kMarkKey;
  s->op.threads = 0;
  s->op.blocks = 0;
  s->op.mark = nullptr;
  s->op.name = "mark";
  fiberDeposit(ctx);
  return s->op.mark;
}

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

static int traceLevel() {
  static const int t = getenv("ESF_BATCH_TRACE") ? atoi(getenv("ESF_BATCH_TRACE")) : 0;
  return t;
}

// ESF_BATCH_TRACE=2: one line per launch group
static void traceGroup(const std::vector<SceneFiber*>& group, size_t held, long round) {
  const char* nm = group[0]->op.name;
  const char* f = strstr(nm, "FH = ");
  std::string k = f ? std::string(f + 5, strcspn(f + 5, ",;]")) : std::string(nm);
  fprintf(stderr, "  round %ld: %zu scenes (%zu held)  %s\n", round, group.size(), held,
          k.c_str());
}

// One group per round.  Of the deposits held, the group of equal deposits
// (same kernel instance and block size, or mark requests) whose oldest
// deposit is the most recent goes first, the larger of equally recent ones:
// fibers that ran on while others did extra work (more solids, a wait for
// their own poll) hold their deposits until the others catch up, so the
// scenes' launches fall back into step instead of splitting.  A fiber that
// waits is woken once its event has completed; when no fiber holds a
// deposit the scheduler blocks on the oldest wait.  The step's barriers
// (batchBarrier) release once every live fiber has reached one.
static void schedule(BatchRun& run, const std::vector<SceneFiber*>& fs) {
  std::vector<SceneFiber*> dep, group;
  std::vector<FiberOp*> ops;
  for (;;) {
    double t0 = nowS();
    for (SceneFiber* s : fs)
      if (s->state == SceneFiber::RUNNABLE) {
        esf_fiber_resume(&s->f);
        if (s->f.finished) s->state = SceneFiber::FINISHED;
        if (s->state == SceneFiber::DEPOSITED) s->depRound = run.rounds;
      }
    double t1 = nowS();
    run.tScenes += t1 - t0;
    run.rounds++;
    // waits whose events have completed
    SceneFiber* oldest = nullptr;
    bool woke = false;
    for (SceneFiber* s : fs) {
      if (s->state != SceneFiber::WAITING) continue;
      const hipError_t q = cachedQuery(run, s->wait);
      if (q == hipErrorNotReady) {
        if (!oldest) oldest = s;
        continue;
      }
      s->waitResult = q;
      s->state = SceneFiber::RUNNABLE;
      woke = true;
    }
    if (woke) {
      // the woken fibers deposit before a group is chosen
      run.tWait += nowS() - t1;
      continue;
    }
    dep.clear();
    for (SceneFiber* s : fs)
      if (s->state == SceneFiber::DEPOSITED) dep.push_back(s);
    if (dep.empty()) {
      if (!oldest) {
        // every live fiber is at a barrier (or none is left)
        bool any = false;
        for (SceneFiber* s : fs)
          if (s->state == SceneFiber::BARRIER) {
            s->state = SceneFiber::RUNNABLE;
            any = true;
          }
        if (!any) return;
        continue;
      }
      oldest->waitResult = hipEventSynchronize(oldest->wait);
      oldest->state = SceneFiber::RUNNABLE;
      if (traceLevel() > 1) fprintf(stderr, "  round %ld: blocking wait\n", run.rounds);
      run.waits++;
      run.tWait += nowS() - t1;
      continue;
    }
    // the group whose oldest deposit is the most recent; the larger on a tie
    size_t best = 0, bestSize = 0;
    long bestFirst = -1;
    for (size_t i = 0; i < dep.size(); i++) {
      size_t k = 0;
      long first = dep[i]->depRound;
      for (size_t j = 0; j < dep.size(); j++)
        if (dep[j]->op.key == dep[i]->op.key && dep[j]->op.threads == dep[i]->op.threads) {
          k++;
          first = std::min(first, dep[j]->depRound);
        }
      if (first > bestFirst || (first == bestFirst && k > bestSize)) {
        bestFirst = first;
        bestSize = k;
        best = i;
      }
    }
    group.clear();
    for (SceneFiber* s : dep)
      if (s->op.key == dep[best]->op.key && s->op.threads == dep[best]->op.threads)
        group.push_back(s);
    if (traceLevel() > 1) traceGroup(group, dep.size() - group.size(), run.rounds);
    if (group[0]->op.key == kMarkKey) {
      const int k = run.nextMark;
      run.nextMark = (k + 1) % kMarkEvents;
      const hipError_t e = hipEventRecord(run.marks[k], run.stream);
      if (e != hipSuccess) run.err = e;
      run.markKnown[k] = hipErrorNotReady;
      run.markRound[k] = -1;
      for (SceneFiber* s : group) s->op.mark = run.marks[k];
      run.markRecords++;
    } else {
      ops.clear();
      for (SceneFiber* s : group) ops.push_back(&s->op);
      ops[0]->submit(run, ops.data(), (int)ops.size());
      run.launches++;
      run.scenesLaunched += (long)ops.size();
    }
    for (SceneFiber* s : group) s->state = SceneFiber::RUNNABLE;
    run.tSubmit += nowS() - t1;
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
    for (int k = 0; k < kMarkEvents; k++) {
      if (hipEventCreateWithFlags(&run.marks[k], hipEventDisableTiming) != hipSuccess) {
        why = "batch: event";
        return ESF_ERR_DEVICE;
      }
      run.markIndex[run.marks[k]] = k;
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
  // a batch of one is esf_step itself (nothing to merge its launches with)
  if (n == 1) {
    const esf_status r = esf_step(ctxs[0], alpha, stats);
    if (status) status[0] = r;
    return r;
  }
  // the argument records of one launch must fit one staging chunk
  if ((size_t)n * (kMaxPack + 8) + 256 > kStageChunk) {
    ctxs[0]->err = "too many scenes in one batch";
    return ESF_ERR_INVALID;
  }
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return ESF_ERR_DEVICE;
  BatchPool* pool = takePool(device);
  struct Release {
    BatchPool* p;
    ~Release() { releasePool(p); }
  } release{pool};
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
  run.rounds = run.launches = run.scenesLaunched = run.waits = run.markRecords = 0;
  for (int k = 0; k < kMarkEvents; k++) {
    run.markKnown[k] = hipSuccess;
    run.markRound[k] = -1;
  }
  run.tScenes = run.tSubmit = run.tWait = 0;
  const double tStart = nowS();
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
    s->run = &run;
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
  if (traceLevel() > 0)
    fprintf(stderr,
            "[esf_step_batch] %d scenes: %.3f ms, %ld rounds, %ld launches (%.1f scenes each), "
            "%ld marks, %ld blocking waits; host ms: scenes %.3f, submit %.3f, waits %.3f\n",
            n, 1e3 * (nowS() - tStart), run.rounds, run.launches,
            run.launches ? (double)run.scenesLaunched / run.launches : 0.0, run.markRecords,
            run.waits,
            1e3 * run.tScenes, 1e3 * run.tSubmit, 1e3 * run.tWait);
  if (run.err != hipSuccess) {
    ctxs[0]->err = std::string("batched launch: ") + hipGetErrorString(run.err);
    return ESF_ERR_DEVICE;
  }
  return worst;
}
