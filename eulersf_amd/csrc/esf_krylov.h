// esf_krylov.h -- pieces shared by the FP64/FP32 pressure solve and the
// solid solve: the CR direction update and the host-side done-flag poller.
#pragma once
#include "esf_internal.h"

namespace esf {
namespace {

// p = beta p + r ; Ap = beta Ap + s ; partial Ap . (Ap / D) -> alpha
// (CONJUGATE_RESIDUAL.inl:139-160: _direction *= beta; += _residual; same
// for _Ap; the next iteration's q = Ap / D and alpha = Ap . q)
template <typename T>
__global__ void __launch_bounds__(256) kCrDirection(long n, T* p, T* Ap, const T* r,
                                                    const T* s, const T* D,
                                                    CrScalars* sc, double* partials,
                                                    unsigned* counter) {
  if (sc->done) return;
  __shared__ double red[8];
  const T beta = (T)sc->beta;
  double part = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const T pn = p[i] * beta + r[i];
    const T an = Ap[i] * beta + s[i];
    p[i] = pn;
    Ap[i] = an;
    const T q = an / D[i];
    part += (double)an * (double)q;
  }
  double mine[1] = {part}, tot[1];
  if (lastBlockReduce<1>(mine, partials, counter, red, tot) && threadIdx.x == 0) {
    const double apq = tot[0];
    sc->alpha = (fabs(apq) > 0.0) ? sc->delta / apq : apq;
  }
}

}  // namespace

// Poll the device "done" flag of scalar block `which` without stalling the
// queue: after each convergence check an async copy of the flag lands in a
// pinned ring slot behind an event; the host keeps one check interval
// queued ahead and waits on the previous one, so the device never idles
// and at most one interval of early-exiting launches follows convergence.
struct DonePoller {
  esf_ctx* ctx;
  int which;
  static constexpr int kRing = 8;
  static constexpr int kLook = 1;
  int issued = 0, waited = 0;
  hipEvent_t ev[kRing];
  explicit DonePoller(esf_ctx* c, int w) : ctx(c), which(w) {
    for (int i = 0; i < kRing; i++) ev[i] = ctx->ev[8 + i];
  }
  // returns true when the solve is known to be done
  bool mark() {
    const int slot = issued % kRing;
    hipMemcpyAsync(&ctx->hostFlags[slot], &ctx->sc[which].done, sizeof(int),
                   hipMemcpyDeviceToHost, ctx->stream);
    hipEventRecord(ev[slot], ctx->stream);
    issued++;
    while (issued - waited > kLook) {
      const int ws = waited % kRing;
      hipEventSynchronize(ev[ws]);
      waited++;
      if (ctx->hostFlags[ws]) return true;
    }
    // opportunistic non-blocking checks
    while (waited < issued) {
      const int ws = waited % kRing;
      if (hipEventQuery(ev[ws]) != hipSuccess) break;
      waited++;
      if (ctx->hostFlags[ws]) return true;
    }
    return false;
  }
};

}  // namespace esf
