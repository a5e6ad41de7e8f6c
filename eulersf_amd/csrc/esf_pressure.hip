// esf_pressure.hip -- the Schur pressure system S = J M^-1 J^T on the
// collocated Q1-P0 grid of the reference, and its Krylov solvers.
//
// S x (COUPLED_INTEGRATOR_3D.cpp:638-660) is never assembled.  One kernel
// streams x through LDS in 32x8-cell tiles marching along z, the next layer
// prefetched into registers: for each node layer it forms w = minv * (J^T x)
// (8 cells, signs from the corner bits, COUPLED_INTEGRATOR_3D.cpp:239-249),
// for each cell layer it forms y = J w (8 corners x 3 components).  Sums run
// in the reference's sparse row/column order, so the FP64 result is bitwise
// the reference's CSR product.  Bytes per apply: x once, y once, minv once
// (8 B each in FP64).
//
// The Krylov loop is CONJUGATE_RESIDUAL::solvePCR (CONJUGATE_RESIDUAL.inl:
// 109-192), one SpMV launch per iteration with the direction update of the
// previous iteration folded in:
//   Ap = beta Ap + s, p = beta p + r (:157-160), x += alpha p,
//   r -= alpha D^-1 Ap (recomputed on the halo), s = S r; partial sums r.s,
//   s.D^-1 s, Ap.D^-1 s, Ap.D^-1 Ap
// and one finisher: delta and beta from r.s, alpha = delta / (Ap.D^-1 Ap)
// for the next direction Ap = s + beta Ap with the square expanded, so a
// single reduction per iteration provides both (only alpha's rounding
// differs from the reference's evaluation).  s and Ap alternate between two
// buffers (halo cells read the previous values while owners write the new
// ones), as r does.  A true-residual check every 5 iterations, as the
// reference does.  Block partials are summed by a one-block finisher (fixed
// order, deterministic).
#include "esf_internal.h"
#include "esf_krylov.h"

namespace esf {

enum SpmvMode {
  M_APPLY = 0,
  M_SETUP_R = 1,
  M_SETUP_S = 2,
  M_UPDATE = 3,
  M_CHECK = 4,
  M_IR_RESID = 5,
};

template <typename T>
struct SpmvArgs {
  const T* in;       // input vector (x, r, or r_old / x_old)
  const T* minv;     // per node
  T* out;            // M_APPLY output
  const T* b;        // right-hand side
  const T* D;        // Jacobi diagonal
  T* x;              // solution (M_UPDATE, M_IR_RESID)
  T* r;              // residual
  T* p;              // direction
  T* s;              // S r (M_UPDATE: the previous s, read)
  T* Ap;             // A p (M_UPDATE: the previous Ap, read)
  T* sOut;           // M_UPDATE: the new s
  T* ApOut;          // M_UPDATE: the new Ap
  const float* e32;  // M_IR_RESID: correction added to x (may be null)
  float* r32;        // M_IR_RESID: residual handed to the FP32 solve
  CrScalars* sc;
  double* partials;
  int iter;
  int first;         // M_IR_RESID: also compute |b|
  int recursive;     // M_UPDATE: the finisher tests the recursive residual
  CrScalars* hostOut;  // M_CHECK (M_UPDATE when recursive): pinned slot for the scalar block
};

constexpr int TX = 32, TY = 8, TZ = 8;
constexpr int HALO = (TX + 2) * (TY + 2);  // cells of one layer incl. halo
constexpr int NODES = (TX + 1) * (TY + 1); // nodes of one layer
constexpr int HPT = (HALO + 255) / 256;    // halo cells per thread
constexpr int NPT = (NODES + 255) / 256;   // nodes per thread

// raw global values of one halo cell, fetched a layer ahead and turned
// into the LDS value when stored (so the loads overlap the previous
// layer's arithmetic)
template <typename T>
struct Fetched {
  T in, ap, s, d, x, p;
  float e;
  bool valid, own;
  long c;
};

template <typename T, int MODE>
__global__ void __launch_bounds__(256) kSpmv(Dims d, SpmvArgs<T> a, int tilesX,
                                             int tilesY, int tilesZ) {
  if (MODE != M_APPLY && a.sc && a.sc->done) return;
  __shared__ T xs[2][HALO];
  __shared__ T ws[2][NODES * 3];
  __shared__ double red[16];

  // XCD-aware tile order: blocks are dealt round-robin to the 8 XCDs, so
  // give each XCD a contiguous run of tiles (shared halos stay in one L2).
  const int nb = tilesX * tilesY * tilesZ;
  int bid = blockIdx.x;
  int tile = (nb % 8 == 0) ? (bid % 8) * (nb / 8) + bid / 8 : bid;
  const int tx = tile % tilesX, ty = (tile / tilesX) % tilesY,
            tz = tile / (tilesX * tilesY);
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  const int zEnd = min(z0 + TZ, d.CZ);

  T alpha = 0, beta = 0;
  if (MODE == M_UPDATE) {
    alpha = (T)a.sc->alpha;
    beta = (T)a.sc->beta;
  }

  double part0 = 0, part1 = 0, part2 = 0, part3 = 0, part4 = 0;

  Fetched<T> f[HPT];
  auto fetch = [&](int cz) {
#pragma unroll
    for (int h = 0; h < HPT; h++) {
      const int t = threadIdx.x + 256 * h;
      Fetched<T>& q = f[h];
      q.valid = false;
      if (t >= HALO) continue;
      const int lx = t % (TX + 2), ly = t / (TX + 2);
      const int cx = x0 - 1 + lx, cy = y0 - 1 + ly;
      if (cx < 0 || cx >= d.CX || cy < 0 || cy >= d.CY || cz < 0 || cz >= d.CZ) continue;
      q.valid = true;
      q.c = cx + (long)cy * d.CX + (long)cz * d.cslab;
      q.own = lx >= 1 && lx <= TX && ly >= 1 && ly <= TY && cz >= z0 && cz < zEnd;
      q.in = a.in[q.c];
      if (MODE == M_UPDATE) {
        q.ap = a.Ap[q.c];
        q.s = a.s[q.c];
        q.d = a.D[q.c];
        if (q.own) {
          q.x = a.x[q.c];
          q.p = a.p[q.c];
        }
      } else if (MODE == M_IR_RESID) {
        if (a.e32) q.e = a.e32[q.c];
      }
    }
  };
  auto store = [&](int cz) {
    T* L = xs[cz & 1];
#pragma unroll
    for (int h = 0; h < HPT; h++) {
      const int t = threadIdx.x + 256 * h;
      if (t >= HALO) continue;
      const Fetched<T>& q = f[h];
      T v = 0;
      if (q.valid) {
        if (MODE == M_UPDATE) {
          // r_old, Ap_old, s_old read on the halo; r_new (and at own cells
          // Ap, p, x) written to the other buffers
          const T apn = q.ap * beta + q.s;
          const T qq = apn / q.d;
          v = q.in - alpha * qq;
          if (q.own) {
            const T pn = q.p * beta + q.in;
            a.r[q.c] = v;
            a.ApOut[q.c] = apn;
            a.p[q.c] = pn;
            a.x[q.c] = q.x + alpha * pn;
          }
        } else if (MODE == M_IR_RESID) {
          // x_old (+ e) read on the halo, x_new written to the other buffer
          v = q.in;
          if (a.e32) {
            v = v + (T)q.e;
            if (q.own) a.x[q.c] = v;
          }
        } else {
          v = q.in;
        }
      }
      L[t] = v;
    }
  };

  // minv of the next node layer is loaded into one of two register sets
  // while the other feeds the current layer (selected by layer parity)
  T mvA[NPT], mvB[NPT];
  auto fetchMinv = [&](int k, T* mv) {
#pragma unroll
    for (int h = 0; h < NPT; h++) {
      const int t = threadIdx.x + 256 * h;
      mv[h] = 0;
      if (t >= NODES) continue;
      const int i = x0 + t % (TX + 1), j = y0 + t / (TX + 1);
      if (i >= 1 && i <= d.X - 2 && j >= 1 && j <= d.Y - 2 && k >= 1 && k <= d.Z - 2)
        mv[h] = a.minv[i + (long)j * d.X + (long)k * d.slab];
    }
  };
  auto nodeLayer = [&](int k, const T* mv) {
    T* W = ws[k & 1];
    const T* L0 = xs[(k - 1) & 1];
    const T* L1 = xs[k & 1];
#pragma unroll
    for (int h = 0; h < NPT; h++) {
      const int t = threadIdx.x + 256 * h;
      if (t >= NODES) continue;
      const int lx = t % (TX + 1), ly = t / (TX + 1);
      const int i = x0 + lx, j = y0 + ly;
      T w0 = 0, w1 = 0, w2 = 0;
      if (i >= 1 && i <= d.X - 2 && j >= 1 && j <= d.Y - 2 && k >= 1 && k <= d.Z - 2) {
        T m0 = 0, m1 = 0, m2 = 0;
        // adjacent cells in ascending cell index: dz, dy, dx
#pragma unroll
        for (int dz = 0; dz < 2; dz++) {
          const T* L = dz ? L1 : L0;
#pragma unroll
          for (int dy = 0; dy < 2; dy++)
#pragma unroll
            for (int dx = 0; dx < 2; dx++) {
              const T v = L[(ly + dy) * (TX + 2) + lx + dx];
              m0 += dx == 0 ? v : -v;
              m1 += dy == 0 ? v : -v;
              m2 += dz == 0 ? v : -v;
            }
        }
        const T mi = mv[h];
        w0 = m0 * mi;
        w1 = m1 * mi;
        w2 = m2 * mi;
      }
      W[3 * t] = w0;
      W[3 * t + 1] = w1;
      W[3 * t + 2] = w2;
    }
  };

  fetch(z0 - 1);
  store(z0 - 1);
  fetch(z0);
  fetchMinv(z0, mvA);
  store(z0);
  __syncthreads();
  nodeLayer(z0, mvA);
  fetch(z0 + 1);
  fetchMinv(z0 + 1, mvB);  // layer z0 + 1 + 2m lives in mvB, z0 + 2 + 2m in mvA
  __syncthreads();
  const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
  const int cx = x0 + lx, cy = y0 + ly;
  for (int cz = z0; cz < zEnd; cz++) {
    // this cell's own input, read before the barrier after which faster
    // threads may already overwrite layer cz's buffer with layer cz + 2
    const T ownIn = xs[cz & 1][(ly + 1) * (TX + 2) + lx + 1];
    store(cz + 1);
    const bool odd = ((cz - z0) & 1) != 0;
    if (cz + 2 <= zEnd) {  // next layer's loads in flight during this layer
      fetch(cz + 2);
      fetchMinv(cz + 2, odd ? mvB : mvA);
    }
    __syncthreads();
    nodeLayer(cz + 1, odd ? mvA : mvB);
    __syncthreads();
    if (cx < d.CX && cy < d.CY) {
      T acc = 0;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int ox = c & 1, oy = (c >> 1) & 1, oz = c >> 2;
        const int ni = cx + ox, nj = cy + oy, nk = cz + oz;
        if (ni < 1 || ni > d.X - 2 || nj < 1 || nj > d.Y - 2 || nk < 1 || nk > d.Z - 2)
          continue;
        const T* W = &ws[(cz + oz) & 1][3 * ((ly + oy) * (TX + 1) + lx + ox)];
        acc += ox ? W[0] : -W[0];
        acc += oy ? W[1] : -W[1];
        acc += oz ? W[2] : -W[2];
      }
      const long c = cx + (long)cy * d.CX + (long)cz * d.cslab;
      if (MODE == M_APPLY) {
        a.out[c] = acc;
      } else if (MODE == M_SETUP_R) {
        const T bv = a.b[c];
        const T rv = (bv - acc) / a.D[c];
        a.r[c] = rv;
        a.p[c] = rv;
        part0 += (double)bv * (double)bv;
      } else if (MODE == M_SETUP_S) {
        a.s[c] = acc;
        a.Ap[c] = acc;
        part0 += (double)ownIn * (double)acc;
        const T q = acc / a.D[c];
        part1 += (double)acc * (double)q;
      } else if (MODE == M_UPDATE) {
        a.sOut[c] = acc;
        // this cell's new direction, formed as on the halo (cache hits)
        const T apn = a.Ap[c] * beta + a.s[c];
        const T dc = a.D[c];
        const T sq = acc / dc;
        part0 += (double)ownIn * (double)acc;
        part1 += (double)acc * (double)sq;
        part2 += (double)apn * (double)sq;
        part3 += (double)apn * (double)(apn / dc);
        // |b - S x| as the recursion carries it: r is D^-1 (b - S x)
        const double rb = (double)dc * (double)ownIn;
        part4 += rb * rb;
      } else if (MODE == M_CHECK) {
        const T t = acc - a.b[c];
        part0 += (double)t * (double)t;
      } else if (MODE == M_IR_RESID) {
        const T bv = a.b[c];
        const T rr = bv - acc;
        a.r32[c] = (float)rr;
        part0 += (double)rr * (double)rr;
        part1 += (double)bv * (double)bv;
      }
    }
  }
  if (MODE == M_APPLY) return;
  if (MODE == M_UPDATE) {
    double mine[5] = {part0, part1, part2, part3, part4};
    blockPartials<5>(mine, a.partials, red);
  } else {
    double mine[2] = {part0, part1};
    blockPartials<2>(mine, a.partials, red);
  }
}

// b = J v0 - c, c (boundary velocities), D = diag(S); also the FP32 copies
// used by the mixed solve.  COUPLED_INTEGRATOR_3D.cpp:286-325 (c),
// :694-709 (D, b).
__global__ void kPressureRhs(Dims d, const double* __restrict__ v0,
                             const double* __restrict__ vel,
                             const double* __restrict__ minv, double* c,
                             double* b, double* D, float* D32) {
  const long r = xcdBlock() * (long)blockDim.x + threadIdx.x;
  if (r >= d.cells) return;
  const int cx = r % d.CX, cy = (r / d.CX) % d.CY, cz = r / d.cslab;
  const long base = cx + (long)cy * d.X + (long)cz * d.slab;
  double tmp = 0, cc = 0, dd = 0;
  for (int k = 0; k < 8; k++) {
    const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
    const long n = base + ox + (long)oy * d.X + (long)oz * d.slab;
    const int ni = cx + ox, nj = cy + oy, nk = cz + oz;
    const bool interior =
        ni >= 1 && ni <= d.X - 2 && nj >= 1 && nj <= d.Y - 2 && nk >= 1 && nk <= d.Z - 2;
    const double sx = ox ? 1.0 : -1.0, sy = oy ? 1.0 : -1.0, sz = oz ? 1.0 : -1.0;
    if (interior) {
      tmp += sx * v0[3 * n];
      tmp += sy * v0[3 * n + 1];
      tmp += sz * v0[3 * n + 2];
      const double mi = minv[n];
      dd += mi * 1.0;
      dd += mi * 1.0;
      dd += mi * 1.0;
    } else {
      cc -= sx * vel[3 * n];
      cc -= sy * vel[3 * n + 1];
      cc -= sz * vel[3 * n + 2];
    }
  }
  c[r] = cc;
  b[r] = tmp - cc;
  D[r] = dd;
  if (D32) D32[r] = (float)dd;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static void tileShape(const Dims& d, int& tx, int& ty, int& tz) {
  tx = cdiv(d.CX, TX);
  ty = cdiv(d.CY, TY);
  tz = cdiv(d.CZ, TZ);
}

template <typename T, int MODE>
esf_status launchSpmv(esf_ctx* ctx, const SpmvArgs<T>& a) {
  int tx, ty, tz;
  tileShape(ctx->d, tx, ty, tz);
  const int nb = tx * ty * tz;
  launchTimed(ctx, MODE + (sizeof(T) == 4 ? 6 : 0), kSpmv<T, MODE>, dim3(nb), dim3(256),
              ctx->d, a, tx, ty, tz);
  // the grid-wide sums of the mode (kSpmv's epilogue partials)
  if constexpr (MODE == M_SETUP_R) launchFinish<F_BNORM, 1>(ctx, nb, a.sc);
  if constexpr (MODE == M_SETUP_S) launchFinish<F_SETUP_S, 2>(ctx, nb, a.sc);
  if constexpr (MODE == M_UPDATE)
    launchFinish<F_CR, 5>(ctx, nb, a.sc, a.iter, a.recursive, nullptr, a.hostOut);
  if constexpr (MODE == M_CHECK)
    launchFinish<F_CHECK, 1>(ctx, nb, a.sc, a.iter, 0, nullptr, a.hostOut);
  if constexpr (MODE == M_IR_RESID) launchFinish<F_IR, 2>(ctx, nb, a.sc, 0, a.first);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("kSpmv launch: ") + hipGetErrorString(e);
    return ESF_ERR_DEVICE;
  }
  return ESF_OK;
}

int spmvBlocks(const Dims& d) {
  int tx, ty, tz;
  tileShape(d, tx, ty, tz);
  return tx * ty * tz;
}

esf_status pressureRhs(esf_ctx* ctx) {
  const Dims& d = ctx->d;
  hipLaunchKernelGGL(kPressureRhs, dim3(cdiv(d.cells, 256)), dim3(256), 0, ctx->stream, d,
                     ctx->v0, ctx->vel, ctx->minv, ctx->c, ctx->b, ctx->D,
                     ctx->p.precision == ESF_MIXED ? ctx->D32 : nullptr);
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

__global__ void kToFloat(long n, const double* a, float* b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}
// streaming copy of n doubles (calibration of the HBM byte counters and
// the box's copy bandwidth): 16-byte vector loads and stores
__global__ void __launch_bounds__(256) kCopy(long n2, const double2* __restrict__ a,
                                             double2* __restrict__ b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n2;
       i += (long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

esf_status benchCopy(esf_ctx* ctx, long bytes, int reps, double* avgMs) {
  const long n2 = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n2 * 16) != hipSuccess) return ESF_ERR_OOM;
  if (hipMalloc(&b, n2 * 16) != hipSuccess) {
    hipFree(a);
    return ESF_ERR_OOM;
  }
  hipMemsetAsync(a, 0, n2 * 16, ctx->stream);
  const int blocks = 256 * 32;
  hipLaunchKernelGGL(kCopy, dim3(blocks), dim3(256), 0, ctx->stream, n2, (const double2*)a, b);
  hipEventRecord(ctx->ev[2], ctx->stream);
  for (int i = 0; i < reps; i++)
    hipLaunchKernelGGL(kCopy, dim3(blocks), dim3(256), 0, ctx->stream, n2, (const double2*)a, b);
  hipEventRecord(ctx->ev[3], ctx->stream);
  hipError_t e = hipEventSynchronize(ctx->ev[3]);
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
  *avgMs = ms / reps;
  hipFree(a);
  hipFree(b);
  return e == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// back-to-back applies between two events on the context stream: the
// average launch duration of the pressure SpMV on the current masses
esf_status benchPressureApply(esf_ctx* ctx, int fp32, int reps, double* avgMs) {
  const Dims& d = ctx->d;
  if (fp32 && !ctx->minv32) {
    ctx->err = "FP32 apply needs a mixed-precision context";
    return ESF_ERR_INVALID;
  }
  esf_status st = ESF_OK;
  if (fp32)
    hipLaunchKernelGGL(kToFloat, dim3(1024), dim3(256), 0, ctx->stream, d.nodes,
                       (const double*)ctx->minv, ctx->minv32);
  hipEventRecord(ctx->ev[2], ctx->stream);
  for (int i = 0; i < reps && !st; i++) {
    if (fp32) {
      SpmvArgs<float> a{};
      a.in = ctx->e32;
      a.minv = ctx->minv32;
      a.out = ctx->s32;
      st = launchSpmv<float, M_APPLY>(ctx, a);
    } else {
      SpmvArgs<double> a{};
      a.in = ctx->lambda;
      a.minv = ctx->minv;
      a.out = ctx->s;
      st = launchSpmv<double, M_APPLY>(ctx, a);
    }
  }
  hipEventRecord(ctx->ev[3], ctx->stream);
  if (hipEventSynchronize(ctx->ev[3]) != hipSuccess) return ESF_ERR_DEVICE;
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
  *avgMs = reps > 0 ? ms / reps : 0;
  return st;
}

esf_status pressureApplyDevice(esf_ctx* ctx, const double* x, double* y) {
  SpmvArgs<double> a{};
  a.in = x;
  a.minv = ctx->minv;
  a.out = y;
  return launchSpmv<double, M_APPLY>(ctx, a);
}

static esf_status initScalars(esf_ctx* ctx, int which, double eps) {
  initScalarsAsync(ctx, &ctx->sc[which], eps);
  return ESF_OK;
}

// synchronous read of scalar block `which` (through report slot 3)
static esf_status readScalars(esf_ctx* ctx, int which, CrScalars* out) {
  copyScalarsAsync(ctx, &ctx->sc[which], &ctx->hostSc[3]);
  ESF_HIP(hipStreamSynchronize(ctx->stream));
  *out = ctx->hostSc[3];
  return ESF_OK;
}

// CONJUGATE_RESIDUAL::solvePCR (matrix-free) on S, FP64 or FP32 vectors.
// x is the start vector.  Scalar block `which`.  *iters receives the
// reference's iteration count; *relres the last checked |Sx-b|/|b|.
// recursive = false: the reference's test, the true residual every 5
// iterations (CONJUGATE_RESIDUAL.inl:166-176).  recursive = true (the FP32
// inner solve of the mixed form, whose result the FP64 outer residual
// checks): the finisher of every iteration tests the residual the
// recursion carries, and no check SpMV runs.
template <typename T>
esf_status crSolve(esf_ctx* ctx, int which, const T* minv, const T* b, const T* D,
                   T* x, T* r, T* r2, T* p, T* s, T* s2, T* Ap, T* Ap2, double eps, int maxIt,
                   int slot, bool recursive) {
  esf_status st = initScalars(ctx, which, eps);
  if (st) return st;
  const Dims& d = ctx->d;
  SpmvArgs<T> a{};
  a.minv = minv;
  a.b = b;
  a.D = D;
  a.x = x;
  a.r = r;
  a.p = p;
  a.s = s;
  a.Ap = Ap;
  a.sc = &ctx->sc[which];
  a.partials = ctx->partials;
  T* rA = r;
  T* rB = r2;
  a.in = x;
  a.r = rA;
  if ((st = launchSpmv<T, M_SETUP_R>(ctx, a))) return st;
  a.in = rA;
  if ((st = launchSpmv<T, M_SETUP_S>(ctx, a))) return st;
  // the FP32 inner solve of the mixed form stops when its residual
  // stagnates; a check is waited for once two iterations are queued behind it
  DonePoller poll(ctx, which, sizeof(T) == 4, 2, 4);
  T *sA = s, *sB = s2, *apA = Ap, *apB = Ap2;
  a.recursive = recursive ? 1 : 0;
  int i = 0;
  for (; i < maxIt; i++) {
    a.iter = i;
    a.in = rA;
    a.r = rB;
    a.s = sA;
    a.sOut = sB;
    a.Ap = apA;
    a.ApOut = apB;
    a.hostOut = recursive ? poll.slot() : nullptr;
    if ((st = launchSpmv<T, M_UPDATE>(ctx, a))) return st;
    std::swap(rA, rB);
    std::swap(sA, sB);
    std::swap(apA, apB);
    if (recursive) {
      poll.mark(i);
    } else if (i > 0 && i % 5 == 0) {
      a.in = x;
      a.hostOut = poll.slot();
      if ((st = launchSpmv<T, M_CHECK>(ctx, a))) return st;
      poll.mark(i);
    }
    if (poll.drain(i)) {
      i++;
      break;
    }
  }
  ctx->crQueued = i;
  // the report is written without waiting; the caller reads report slot
  // `slot` after its next synchronisation
  copyScalarsAsync(ctx, &ctx->sc[which], &ctx->hostSc[slot]);
  return ESF_OK;
}

template esf_status crSolve<double>(esf_ctx*, int, const double*, const double*,
                                    const double*, double*, double*, double*, double*,
                                    double*, double*, double*, double*, double, int, int, bool);
template esf_status crSolve<float>(esf_ctx*, int, const float*, const float*,
                                   const float*, float*, float*, float*, float*, float*,
                                   float*, float*, float*, double, int, int, bool);

__global__ void kZeroF(long n, float* a) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    a[i] = 0.0f;
}

// Mixed-precision pressure solve: FP32 CR on S (FP32 minv, FP32 vectors,
// FP64 inner products) inside FP64 iterative refinement.  The outer
// residual b - S x, the correction x += e and |r| stay in FP64.
esf_status mixedSolve(esf_ctx* ctx, double eps, int maxIt, int* iters, bool* converged,
                      double* relres, int* refinements, int* innerIters) {
  const Dims& d = ctx->d;
  esf_status st = initScalars(ctx, 0, eps);
  if (st) return st;
  hipLaunchKernelGGL(kToFloat, dim3(1024), dim3(256), 0, ctx->stream, d.nodes,
                     (const double*)ctx->minv, ctx->minv32);
  SpmvArgs<double> o{};
  o.minv = ctx->minv;
  o.b = ctx->b;
  o.in = ctx->lambda;
  o.x = ctx->lambda2;
  o.r32 = ctx->rhs32;
  o.e32 = nullptr;
  o.sc = &ctx->sc[0];
  o.partials = ctx->partials;
  o.first = 1;
  if ((st = launchSpmv<double, M_IR_RESID>(ctx, o))) return st;
  CrScalars h;
  if ((st = readScalars(ctx, 0, &h))) return st;
  int outer = 0, inner = 0;
  // refinement rounds: each gains roughly the FP32 solve's attainable
  // reduction (about eps32 * cond on the smooth modes), so an ill-conditioned
  // cold start needs many; the total FP32 iterations stay capped by maxIt
  const int maxOuter = 1000;
  while (!h.done && outer < maxOuter && inner < maxIt) {
    // inner tolerance: ask for the reduction the outer test still needs,
    // bounded by what FP32 can deliver
    double need = h.resnorm > 0 ? 0.3 * eps * h.bnorm / h.resnorm : 1.0;
    double innerEps = std::min(0.5, std::max(need, ctx->p.inner_eps));
    hipLaunchKernelGGL(kZeroF, dim3(1024), dim3(256), 0, ctx->stream, d.cells, ctx->e32);
    if ((st = crSolve<float>(ctx, 1, ctx->minv32, ctx->rhs32, ctx->D32, ctx->e32, ctx->r32,
                             ctx->r32b, ctx->p32, ctx->s32, ctx->s32b, ctx->Ap32, ctx->Ap32b,
                             innerEps, maxIt - inner, 4, true)))
      return st;
    outer++;
    o.e32 = ctx->e32;
    o.first = 0;
    o.in = ctx->lambda;
    o.x = ctx->lambda2;
    if ((st = launchSpmv<double, M_IR_RESID>(ctx, o))) return st;
    std::swap(ctx->lambda, ctx->lambda2);
    if ((st = readScalars(ctx, 0, &h))) return st;  // also completes report slot 4
    int it = 0;
    bool conv = false;
    double rr = 0;
    solveReport(ctx->hostSc[4], maxIt - inner, &it, &conv, &rr);
    // an inner solve that stopped on stagnation ran the iterations it queued
    if (!conv) it = ctx->crQueued;
    inner += it;
  }
  *converged = h.done == 1;
  *iters = inner;
  *relres = h.bnorm > 0 ? h.resnorm / h.bnorm : h.resnorm;
  *refinements = outer;
  *innerIters = inner;
  return ESF_OK;
}

}  // namespace esf
