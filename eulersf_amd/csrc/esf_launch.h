// esf_launch.h -- kernels that run as their own launch or as one scene's
// share of a launch over several scenes.
//
// Every kernel of the step is written as a device function template
//     template <class G> __device__ void kFoo(const G& wg, params...)
// that takes its work-group index and count from `wg`:
//   HwGrid    : its own launch (the hardware indices, XCD-renumbered by
//               xcdBlock());
//   SceneGrid : one scene's blocks inside a launch over several scenes
//               (the index and count the scene's own launch would have).
// A kernel computes the same values for the same (index, count), so a
// scene's results do not depend on which launch carried its blocks.
// launch<ESF_K(kFoo)>(ctx, grid, block, args...) launches kOne<kFoo> on the
// context's stream -- or, while the context's step runs as a fiber of
// esf_step_batch, hands the launch to the batch scheduler, which puts the
// launches that the scenes reach together into one kMany<kFoo> launch.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <utility>

namespace esf {

struct HwGrid {
  __device__ __forceinline__ int block() const { return xcdBlock(); }
  __device__ __forceinline__ int blocks() const { return (int)gridDim.x; }
};
struct SceneGrid {
  int b, n;
  __device__ __forceinline__ int block() const { return b; }
  __device__ __forceinline__ int blocks() const { return n; }
};

// a kernel's arguments in one trivially copyable record, laid out the same
// by the host and the device compilers (members in order, natural alignment)
template <typename... P>
struct ArgPack {};
template <typename H, typename... T>
struct ArgPack<H, T...> {
  H h;
  ArgPack<T...> t;
};

inline ArgPack<> makePack() { return {}; }
template <typename H>
inline ArgPack<H> makePack(const H& h) {
  ArgPack<H> p;
  std::memcpy((void*)&p.h, (const void*)&h, sizeof(H));
  return p;
}
template <typename H, typename N, typename... T>
inline ArgPack<H, N, T...> makePack(const H& h, const N& n, const T&... t) {
  ArgPack<H, N, T...> p;
  std::memcpy((void*)&p.h, (const void*)&h, sizeof(H));
  p.t = makePack<N, T...>(n, t...);
  return p;
}

// the packs a launch over several scenes reads are constant for the whole
// kernel: read through the constant address space, so a by-value argument
// copied out of a pack needs no private copy (the compiler reads the pack
// in place, as it reads a kernel's own arguments)
#define ESF_CONST __attribute__((address_space(4)))

template <std::size_t I, typename H, typename... T>
__device__ __forceinline__ auto packGet(const ESF_CONST ArgPack<H, T...>& p) {
  if constexpr (I == 0) {
    H v;
    __builtin_memcpy(&v, (const ESF_CONST void*)&p.h, sizeof(H));
    return v;
  } else {
    return packGet<I - 1>(p.t);
  }
}

template <auto F, typename G, typename... P, std::size_t... I>
__device__ __forceinline__ void callPack(const G& g, const ESF_CONST ArgPack<P...>& p,
                                         std::index_sequence<I...>) {
  F(g, packGet<I>(p)...);
}

template <typename... P>
struct TypeList {};
template <typename G, typename... P>
TypeList<P...> paramsOf(void (*)(const G&, P...));

// the kernel alone
template <auto F, int LB, typename... P>
__global__ void __launch_bounds__(LB) kOne(P... p) {
  F(HwGrid{}, p...);
}

// several scenes' launches of one kernel: scene k has blocks
// [start[k], start[k] + count[k]) of this grid and its arguments in pack k.
// tab: start[n], count[n], then the packs from byte packOff, `stride` apart.
template <auto F, int LB, typename... P>
__global__ void __launch_bounds__(LB) kMany(const unsigned char* __restrict__ tab, int n,
                                            int packOff, int stride) {
  const int g = xcdBlock();
  const int* start = (const int*)tab;
  const int lane = threadIdx.x & 63;
  // the scene is the last k with start[k] <= g: count the starts at or
  // below g, 64 at a time with a ballot (the same in every wave)
  int s = -1;
  for (int base = 0; base < n; base += 64) {
    const int i = base + lane;
    const bool le = i < n && start[i] <= g;
    s += __popcll(__ballot(le));
  }
  s = __builtin_amdgcn_readfirstlane(s);
  const int first = start[s], count = start[n + s];
  const ESF_CONST ArgPack<P...>& pk =
      *(const ESF_CONST ArgPack<P...>*)(tab + packOff + (long)s * stride);
  callPack<F>(SceneGrid{g - first, count}, pk, std::index_sequence_for<P...>{});
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
constexpr int kMaxPack = 1024;  // bytes of one launch's arguments

struct BatchRun;
// what a scene's fiber hands to the scheduler: a launch, or (key ==
// kMarkKey) a request for an event that marks its work queued so far
struct FiberOp {
  const void* key;  // the kMany instance
  int blocks, threads;
  int packBytes;
  // launch kMany for `n` scenes' ops of the same key and block size
  void (*submit)(BatchRun& run, FiberOp* const* ops, int n);
  hipEvent_t mark;  // kMarkKey: the event the scheduler recorded
  const char* name; // the launch's signature (diagnostics)
  alignas(16) unsigned char pack[kMaxPack];
};
extern const char kMarkKeyObj;
#define kMarkKey ((const void*)&::esf::kMarkKeyObj)

// esf_batch.hip: the scheduler's side of a deposit and of a wait
FiberOp* fiberOp(esf_ctx* ctx);
void fiberDeposit(esf_ctx* ctx);
hipError_t fiberWait(esf_ctx* ctx, hipEvent_t ev);
hipEvent_t fiberMark(esf_ctx* ctx);
void fiberBarrier(esf_ctx* ctx);
hipError_t fiberQuery(esf_ctx* ctx, hipEvent_t ev);
// staging of the launch tables: `bytes` of pinned host memory whose
// device copy is at *dev once stageCopy has been queued
unsigned char* stageAlloc(BatchRun& run, size_t bytes, const unsigned char** dev);
void stageCopy(BatchRun& run, const unsigned char* host, const unsigned char* dev, size_t bytes);
hipStream_t batchStream(BatchRun& run);
void batchLaunchFailed(BatchRun& run, hipError_t e);

template <auto F, int LB, typename... P>
void submitMany(BatchRun& run, FiberOp* const* ops, int n) {
  using Pk = ArgPack<P...>;
  const int stride = (int)((sizeof(Pk) + 15) & ~size_t(15));
  const int packOff = (8 * n + 15) & ~15;
  const size_t bytes = (size_t)packOff + (size_t)stride * n;
  const unsigned char* dev = nullptr;
  unsigned char* h = stageAlloc(run, bytes, &dev);
  int* start = (int*)h;
  long total = 0;
  for (int k = 0; k < n; k++) {
    start[k] = (int)total;
    start[n + k] = ops[k]->blocks;
    total += ops[k]->blocks;
    std::memcpy(h + packOff + (size_t)k * stride, ops[k]->pack, sizeof(Pk));
  }
  stageCopy(run, h, dev, bytes);
  hipLaunchKernelGGL((kMany<F, LB, P...>), dim3((unsigned)total), dim3(ops[0]->threads), 0,
                     batchStream(run), dev, n, packOff, stride);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) batchLaunchFailed(run, e);
}

template <auto FH, auto FS, int LB, typename... P, typename... A>
inline void launchImpl(esf_ctx* ctx, hipStream_t stream, int cls, dim3 grid, dim3 block,
                       TypeList<P...>, A&&... a) {
  static_assert(sizeof(ArgPack<P...>) <= kMaxPack, "kernel arguments exceed kMaxPack");
  if (grid.x == 0) return;
  if (ctx->fiber && stream == ctx->stream) {
    FiberOp* op = fiberOp(ctx);
    const ArgPack<P...> pk = makePack<P...>(P(a)...);
    std::memcpy(op->pack, &pk, sizeof(pk));
    op->packBytes = (int)sizeof(pk);
    op->key = (const void*)&kMany<FS, LB, P...>;
    op->blocks = (int)grid.x;
    op->threads = (int)block.x;
    op->submit = &submitMany<FS, LB, P...>;
    op->name = __PRETTY_FUNCTION__;
    fiberDeposit(ctx);
    return;
  }
  if (cls < 0 || !ctx->kernelTiming) {
    hipLaunchKernelGGL((kOne<FH, LB, P...>), grid, block, 0, stream, P(a)...);
    return;
  }
  if (ctx->timedUsed == ctx->timed.size()) {
    esf_ctx::TimedLaunch t{};
    (void)hipEventCreate(&t.a);
    (void)hipEventCreate(&t.b);
    ctx->timed.push_back(t);
  }
  esf_ctx::TimedLaunch* t = &ctx->timed[ctx->timedUsed++];
  t->cls = cls;
  // the event pair rides in the launch's own dispatch, so it brackets the
  // kernel's execution alone
  hipExtLaunchKernelGGL((kOne<FH, LB, P...>), grid, block, 0, stream, t->a, t->b, 0, P(a)...);
}

// launch<ESF_K(kFoo)>(ctx, grid, block, args...)            on ctx->stream
// launchTimed<ESF_K(kFoo)>(ctx, cls, grid, block, args...)  the same, timed as
//     kernel class `cls` while kernel timing is on
// launchOn<ESF_K(kFoo)>(ctx, stream, grid, block, args...)  on another stream
#define ESF_K(k) &k<::esf::HwGrid>, &k<::esf::SceneGrid>, k##_LB
#define ESF_KT(k, ...) &k<__VA_ARGS__, ::esf::HwGrid>, &k<__VA_ARGS__, ::esf::SceneGrid>, k##_LB
#define ESF_BOUNDS(k, n) constexpr int k##_LB = (n)

template <auto FH, auto FS, int LB, typename... A>
inline void launch(esf_ctx* ctx, dim3 grid, dim3 block, A&&... a) {
  launchImpl<FH, FS, LB>(ctx, ctx->stream, -1, grid, block, decltype(paramsOf(FH)){},
                         std::forward<A>(a)...);
}
template <auto FH, auto FS, int LB, typename... A>
inline void launchTimed(esf_ctx* ctx, int cls, dim3 grid, dim3 block, A&&... a) {
  launchImpl<FH, FS, LB>(ctx, ctx->stream, cls, grid, block, decltype(paramsOf(FH)){},
                         std::forward<A>(a)...);
}
template <auto FH, auto FS, int LB, typename... A>
inline void launchOn(esf_ctx* ctx, hipStream_t s, dim3 grid, dim3 block, A&&... a) {
  launchImpl<FH, FS, LB>(ctx, s, -1, grid, block, decltype(paramsOf(FH)){},
                         std::forward<A>(a)...);
}

// Events and waits.  Alone, markEvent records `own` on the context's
// stream and returns it; in a batched step it returns an event the
// scheduler records once for all scenes that ask in the same round (later
// than the scene's own work, never earlier).  The waits and queries let the
// other scenes of a batch go on meanwhile.
inline hipEvent_t markEvent(esf_ctx* ctx, hipEvent_t own) {
  if (ctx->fiber) return fiberMark(ctx);
  (void)hipEventRecord(own, ctx->stream);
  return own;
}
inline hipError_t queryEvent(esf_ctx* ctx, hipEvent_t ev) {
  return ctx->fiber ? fiberQuery(ctx, ev) : hipEventQuery(ev);
}
inline hipError_t eventSync(esf_ctx* ctx, hipEvent_t ev) {
  if (!ctx->fiber) return hipEventSynchronize(ev);
  return fiberWait(ctx, ev);
}
// a step's stage boundary: in a batched step the scene waits here until
// every scene of the batch has reached it (so the next stage's launches
// start together); alone it does nothing
inline void batchBarrier(esf_ctx* ctx) {
  if (ctx->fiber) fiberBarrier(ctx);
}

inline hipError_t streamSync(esf_ctx* ctx) {
  if (!ctx->fiber) return hipStreamSynchronize(ctx->stream);
  return fiberWait(ctx, fiberMark(ctx));
}

// `bytes` (a multiple of 4) of device memory into pinned host memory,
// queued on the context's stream: a kernel, so a batched step puts the
// scenes' copies into one launch
template <class GridT>
__device__ __forceinline__ void kToHost(const GridT& wg, const unsigned* src, unsigned* dst,
                                        long words) {
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < words;
       i += (long)wg.blocks() * blockDim.x)
    dst[i] = src[i];
}
ESF_BOUNDS(kToHost, 256);

inline void toHostAsync(esf_ctx* ctx, void* hostPinned, const void* dev, size_t bytes) {
  const long words = (long)(bytes / 4);
  const int nb = (int)std::min<long>(std::max<long>((words + 255) / 256, 1), 64);
  launch<ESF_K(kToHost)>(ctx, dim3(nb), dim3(256), (const unsigned*)dev, (unsigned*)hostPinned,
                         words);
}

}  // namespace esf
