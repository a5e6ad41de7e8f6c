// esf_krylov.h -- pieces shared by the FP64/FP32 pressure solve and the
// solid solve: the reduction finishers, the scalar-block kernels and the
// host-side done-flag poller.
#pragma once
#include <algorithm>

#include "esf_internal.h"

namespace esf {
namespace {

// What the finisher does with the totals (t0, t1, ...) of the producer
// that ran just before it.  The scalar updates follow CONJUGATE_RESIDUAL.inl
// (:118-131 setup, :141-160 loop, :166-176 convergence test); alpha stays
// the denominator when that is zero, as there.
enum FinishOp {
  F_BNORM = 0,    // |b| = sqrt(t0)
  F_SETUP_S = 1,  // delta = t0 (r.s); alpha = delta / t1 (Ap.D^-1 Ap)
  F_CHECK = 3,    // |Ax - b| = sqrt(t0); done when below eps |b|
  F_IR = 4,       // outer residual sqrt(t0); |b| = sqrt(t1) on the first call
  F_CR = 6,       // delta, beta from t0; alpha from t1 + 2 beta t2 + beta^2 t3
  F_WARM = 7,     // |b| = sqrt(t0), |b - A x0| = sqrt(t1); done at once when that
                  // is already below eps |b| (a warm start that needs no iteration)
};

// hostOut (pinned host memory, may be null): the scalar block after the
// update, written by the kernel itself, so a host poll needs no copy
template <int OP, int K, class GridT>
__device__ __forceinline__ void kFinish(const GridT& wg, int nb, const double* partials,
                                        CrScalars* sc, int iter, int first, CrScalars* hostOut) {
  // the sums' loads go out together with the done flag's (the partials
  // of a finished solve's early-exit launches are never used)
  const bool done = sc->done != 0;
  __shared__ double red[16 * K];
  double t[K];
  finishSums<K>(nb, partials, red, t);
  if (threadIdx.x != 0) return;
  if (done) {
    if (hostOut) *hostOut = *sc;
    return;
  }
  if constexpr (OP == F_BNORM) {
    sc->bnorm = sqrt(t[0]);
  } else if constexpr (OP == F_WARM) {
    sc->bnorm = sqrt(t[0]);
    sc->resnorm = sqrt(t[1]);
    if (sc->resnorm < sc->eps * sc->bnorm) {
      sc->done = 1;
      sc->iter = 0;
    }
  } else if constexpr (OP == F_SETUP_S) {
    sc->delta = t[0];
    sc->beta = 0.0;
    sc->alpha = (fabs(t[K - 1]) > 0.0) ? t[0] / t[K - 1] : t[K - 1];
  } else if constexpr (OP == F_CHECK) {
    // first = 1: the check rode along with the next iteration, whose x
    // update the solve undoes at its end (aux)
    const double rn = sqrt(t[0]);
    sc->resnorm = rn;
    if (rn < sc->eps * sc->bnorm) {
      sc->done = 1;
      sc->iter = iter;
      sc->aux = first;
    }
  } else if constexpr (OP == F_IR) {
    sc->resnorm = sqrt(t[0]);
    if (first) sc->bnorm = sqrt(t[K - 1]);
    if (sc->resnorm < sc->eps * sc->bnorm) sc->done = 1;
  } else if constexpr (OP == F_CR) {
    // r.s, s.D^-1 s, Ap.D^-1 s, Ap.D^-1 Ap of the iteration just done; alpha
    // for the next direction Ap' = s + beta Ap (CONJUGATE_RESIDUAL.inl:
    // 141-160) with (Ap'.D^-1 Ap') expanded; alpha stays the denominator
    // when that is zero, as there.  first = 1: done when the fifth sum, the
    // recursively carried |b - A x|^2, is below (eps |b|)^2
    static_assert(K >= 4, "CR finisher takes four sums");
    if constexpr (K >= 5) {
      if (first) {
        const double rn = sqrt(t[4]);
        sc->resnorm = rn;
        if (rn < sc->eps * sc->bnorm) {
          sc->done = 1;
          sc->iter = iter + 1;
        }
      }
    }
    const double dNew = t[0];
    const double beta = dNew / sc->delta;
    const double apq = (t[1] + (2.0 * beta) * t[2]) + (beta * beta) * t[3];
    sc->delta = dNew;
    sc->beta = beta;
    sc->alpha = (fabs(apq) > 0.0) ? dNew / apq : apq;
  }
  if (hostOut) *hostOut = *sc;
}
ESF_BOUNDS(kFinish, 1024);

// partials: the producer's block partials (default: the context's buffer);
// hostOut: pinned slot that receives the updated scalar block
template <int OP, int K>
inline void launchFinish(esf_ctx* ctx, int nb, CrScalars* sc, int iter = 0, int first = 0,
                         const double* partials = nullptr, CrScalars* hostOut = nullptr) {
  launch<ESF_KT(kFinish, OP, K)>(ctx, dim3(1), dim3(1024), nb,
                                 partials ? partials : (const double*)ctx->partials, sc, iter,
                                 first, hostOut);
}

// a solve's scalar block at its start: tolerance eps, nothing done yet
template <class GridT>
__device__ __forceinline__ void kScalarsInit(const GridT& wg, CrScalars* sc, double eps) {
  if (threadIdx.x != 0) return;
  initScalarBlock(sc, eps);
}
ESF_BOUNDS(kScalarsInit, 1024);

// a scalar block into pinned host memory (the report of a queued solve)
template <class GridT>
__device__ __forceinline__ void kScalarsCopy(const GridT& wg, const CrScalars* sc, CrScalars* out) {
  if (threadIdx.x == 0) *out = *sc;
}
ESF_BOUNDS(kScalarsCopy, 1024);

inline void initScalarsAsync(esf_ctx* ctx, CrScalars* sc, double eps) {
  launch<ESF_K(kScalarsInit)>(ctx, dim3(1), dim3(64), sc, eps);
}

inline void copyScalarsAsync(esf_ctx* ctx, const CrScalars* sc, CrScalars* out) {
  launch<ESF_K(kScalarsCopy)>(ctx, dim3(1), dim3(64), sc, out);
}

}  // namespace

// Poll the scalar block `which` of a Krylov solve without stalling the
// queue: the finisher of each convergence check writes the block into the
// pinned ring slot slot() and mark(it) records an event behind it (it: the
// iteration just queued).  After queueing each iteration the loop calls
// drain(it): pending checks are queried without blocking, and a check is
// waited for once `lag` further iterations are queued behind it, so the
// device never idles on the host and at most `lag` iterations of
// early-exiting launches follow convergence.
// With `stagnation` set (the FP32 inner solve of the mixed form), the poll
// also ends the solve when `flatMax` consecutive polls set no new minimum
// of the polled residual (4 checks of 5 iterations, or 4 iterations when
// every iteration is polled): FP32 rounding puts a noise floor under the
// inner residual (a slowly converging solve still sets new minima), and the
// FP64 refinement around it takes over from there.
struct DonePoller {
  esf_ctx* ctx;
  int which;
  bool stagnation;
  static constexpr int kRing = 8;
  int lag;      // iterations queued behind a check before the host waits for it
  int flatMax;  // polls without a new minimum that end a stagnating solve
  int issued = 0, waited = 0, flat = 0;
  int markIt[kRing];
  double best = 0;
  bool stagnated = false;
  int stopIt = -1;  // stagnated: the last iteration the solve runs
  hipEvent_t own[kRing];  // the context's events for the marks
  hipEvent_t ev[kRing];   // the event marking each check (own[] alone)
  explicit DonePoller(esf_ctx* c, int w, bool stag = false, int lagIters = 2, int flatPolls = 4)
      : ctx(c), which(w), stagnation(stag), lag(lagIters), flatMax(flatPolls) {
    for (int i = 0; i < kRing; i++) own[i] = ev[i] = ctx->ev[8 + i];
  }
  bool seen(int ws) {
    const CrScalars& h = ctx->hostRing[ws];
    if (h.done) return true;
    if (!stagnation) return false;
    if (best == 0 || h.resnorm < best) {
      best = h.resnorm;
      flat = 0;
    } else if (++flat >= flatMax) {
      stagnated = true;
      stopIt = markIt[ws] + lag;
      return true;
    }
    return false;
  }
  // the pinned slot the next check's finisher writes
  CrScalars* slot() { return &ctx->hostRing[issued % kRing]; }
  // after that finisher is queued with iteration `it`
  void mark(int it) {
    const int slot = issued % kRing;
    markIt[slot] = it;
    ev[slot] = markEvent(ctx, own[slot]);
    issued++;
  }
  // after iteration `it` is queued: true when the solve is known to be done
  // (or stagnated)
  bool drain(int it) {
    while (waited < issued) {
      const int ws = waited % kRing;
      if (markIt[ws] + lag <= it || issued - waited >= kRing) {
        (void)eventSync(ctx, ev[ws]);
      } else if (queryEvent(ctx, ev[ws]) != hipSuccess) {
        break;
      }
      waited++;
      if (seen(ws)) return true;
    }
    return false;
  }
};

}  // namespace esf
