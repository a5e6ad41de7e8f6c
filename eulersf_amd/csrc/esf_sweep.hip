// esf_sweep.hip -- the solid's fast sweeping on the GPU, bitwise the
// reference's ordered Gauss-Seidel sweeps:
//   computeSDF (DEFOR_SOLID_3D.cpp:216-267): phi = -0.5 on nodes with mass,
//     a large distance elsewhere, the box of the massive nodes widened by the
//     tight border and the extrapolation cells, two rounds of sweepPhi;
//   sweepPhi / solveDistance (:369-454): eight ordered sweeps of the grid
//     Eikonal update over the box;
//   extendVectorField (:566-574): four rounds of sweepVectorField (:521-564),
//     eight sweepOnePass (:457-519) and the box faces copied inwards.
//
// A sweep visits the box in loop order and each node reads the three
// neighbours visited before it along i, j and k (or, for the faces, only
// what the previous loop nest wrote).  A node's new value therefore depends
// only on those three, so the nodes of one plane i' + j' + k' = p (primed:
// positions along the sweep's directions) are independent and the planes
// run in order: one workgroup walks the planes with a barrier between
// them, and every node gets the reference's value by the reference's
// arithmetic.  sweepOnePass also carries three flags across the whole pass
// (solvableX/Y/Z: once cleared they stay cleared).  Whether a node clears
// one depends on phi and the mass only, never on the field being swept, so
// a parallel prepass finds, per flag, the first node in loop order that
// clears it (an integer atomic minimum); a node's flag is then "its loop
// position comes before that one".
//
// These are host-side stages of the reference's finalize that the step does
// not run; the entry points serve a host integrator that keeps them.
#include <climits>
#include <mutex>

#include "esf_internal.h"

namespace esf {

namespace {

constexpr int kWaveThreads = 1024;

struct SweepBox {
  int X, Y, Z;
  int x0, y0, z0, x1, y1, z1;
};

__device__ __host__ inline long nodeAt(const SweepBox& b, int i, int j, int k) {
  return i + (long)b.X * (j + (long)b.Y * k);
}

// one ordered sweep: n* nodes along each axis from (i0, j0, k0) in steps
// (di, dj, dk), loop order k outer, i inner
struct Sweep {
  int i0, j0, k0, di, dj, dk, ni, nj, nk;
};

// phi -> -0.5 / large, and the extent of the massive nodes
template <class GridT>
__device__ __forceinline__ void kSdfInit(const GridT& wg, long n, int X, int Y,
                                         const double* __restrict__ mass, double large, double* phi,
                                         int* ext) {
  for (long c = wg.block() * (long)blockDim.x + threadIdx.x; c < n;
       c += (long)wg.blocks() * blockDim.x) {
    const bool m = mass[c] > 0;
    phi[c] = m ? -0.5 : large;
    if (m) {
      const int x = (int)(c % X), y = (int)((c / X) % Y), z = (int)(c / ((long)X * Y));
      atomicMin(&ext[0], x);
      atomicMin(&ext[1], y);
      atomicMin(&ext[2], z);
      atomicMax(&ext[3], x);
      atomicMax(&ext[4], y);
      atomicMax(&ext[5], z);
    }
  }
}
ESF_BOUNDS(kSdfInit, 1024);

// sweepPhi's sweep s, one workgroup, plane by plane
template <class GridT>
__device__ __forceinline__ void kPhiWave(const GridT& wg, SweepBox b, Sweep s,
                                         const double* __restrict__ mass, double* phi) {
  const int planes = s.ni + s.nj + s.nk - 2;
  for (int p = 0; p < planes; p++) {
    // the (j', k') rectangle that can hold a node of plane p
    const int jLo = max(0, p - (s.ni - 1) - (s.nk - 1)), jHi = min(s.nj - 1, p);
    const int kLo = max(0, p - (s.ni - 1) - (s.nj - 1)), kHi = min(s.nk - 1, p);
    const int wj = jHi - jLo + 1, wk = kHi - kLo + 1;
    const int jk = wj > 0 && wk > 0 ? wj * wk : 0;
    for (int t = threadIdx.x; t < jk; t += blockDim.x) {
      const int jp = jLo + t % wj, kp = kLo + t / wj, ip = p - jp - kp;
      if (ip < 0 || ip >= s.ni) continue;
      const int i = s.i0 + s.di * ip, j = s.j0 + s.dj * jp, k = s.k0 + s.dk * kp;
      const long c = nodeAt(b, i, j, k);
      // the mass and the four phi values in one round of loads
      const double mc = mass[c], pa = phi[nodeAt(b, i - s.di, j, k)],
                   pb = phi[nodeAt(b, i, j - s.dj, k)], pe = phi[nodeAt(b, i, j, k - s.dk)],
                   pc = phi[c];
      if (mc == 0) phi[c] = solveDistance(pa, pb, pe, pc);
    }
    __syncthreads();
  }
}
ESF_BOUNDS(kPhiWave, kWaveThreads);

// sweepOnePass's per-node decisions, from phi and the mass only:
// returns false where the node is not updated; clear* say whether the node
// clears that flag (it does so before any later axis may skip it)
struct PassNode {
  double dp, dq, dr;
  bool clearX, clearY, clearZ;
};
__device__ inline bool passNode(const SweepBox& b, const Sweep& s, const double* mass,
                                const double* phi, int i, int j, int k, PassNode& o) {
  o.clearX = o.clearY = o.clearZ = false;
  const long c = nodeAt(b, i, j, k);
  if (mass[c] != 0) return false;
  const double pc = phi[c];
  o.dp = pc - phi[nodeAt(b, i - s.di, j, k)];
  const double dpn = pc - phi[nodeAt(b, i + s.di, j, k)];
  o.dq = pc - phi[nodeAt(b, i, j - s.dj, k)];
  const double dqn = pc - phi[nodeAt(b, i, j + s.dj, k)];
  o.dr = pc - phi[nodeAt(b, i, j, k - s.dk)];
  const double drn = pc - phi[nodeAt(b, i, j, k + s.dk)];
  if (o.dp < 0) {
    if (dpn >= 0) return false;
    o.clearX = true;
  }
  if (o.dq < 0) {
    if (dqn >= 0) return false;
    o.clearY = true;
  }
  if (o.dr < 0) {
    if (drn >= 0) return false;
    o.clearZ = true;
  }
  return true;
}

// the first node, in loop order, that clears each flag
template <class GridT>
__device__ __forceinline__ void kPassFlags(const GridT& wg, SweepBox b, Sweep s,
                                           const double* __restrict__ mass,
                                           const double* __restrict__ phi, int* first) {
  const long n = (long)s.ni * s.nj * s.nk;
  for (long t = wg.block() * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)wg.blocks() * blockDim.x) {
    const int ip = (int)(t % s.ni), jp = (int)((t / s.ni) % s.nj), kp = (int)(t / ((long)s.ni * s.nj));
    PassNode o;
    passNode(b, s, mass, phi, s.i0 + s.di * ip, s.j0 + s.dj * jp, s.k0 + s.dk * kp, o);
    if (o.clearX) atomicMin(&first[0], (int)t);
    if (o.clearY) atomicMin(&first[1], (int)t);
    if (o.clearZ) atomicMin(&first[2], (int)t);
  }
}
ESF_BOUNDS(kPassFlags, 1024);

// every node's coefficients for the pass, from the flags' first clearing
// nodes: alpha and beta where the node is updated, NaN in alpha where it
// is not (indexed by loop position), so the wave below reads them together
// with the neighbours' field values in one round of loads
template <class GridT>
__device__ __forceinline__ void kPassCoef(const GridT& wg, SweepBox b, Sweep s,
                                          const double* __restrict__ mass,
                                          const double* __restrict__ phi,
                                          const int* __restrict__ first, double2* coef) {
  const long n = (long)s.ni * s.nj * s.nk;
  const int fx = first[0], fy = first[1], fz = first[2];
  for (long t = wg.block() * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)wg.blocks() * blockDim.x) {
    const int ip = (int)(t % s.ni), jp = (int)((t / s.ni) % s.nj), kp = (int)(t / ((long)s.ni * s.nj));
    PassNode o;
    double2 ab = make_double2(NAN, 0.0);
    if (passNode(b, s, mass, phi, s.i0 + s.di * ip, s.j0 + s.dj * jp, s.k0 + s.dk * kp, o)) {
      // a flag is cleared from its first clearing node on, that node included
      const double dp = t >= fx ? 0.0 : o.dp;
      const double dq = t >= fy ? 0.0 : o.dq;
      const double dr = t >= fz ? 0.0 : o.dr;
      if (dp + dq + dr < 1e-3) {
        ab = make_double2(0.333, 0.333);
      } else {
        ab = make_double2(dp / (dp + dq + dr), dq / (dp + dq + dr));
      }
    }
    coef[t] = ab;
  }
}
ESF_BOUNDS(kPassCoef, 1024);

// the pass itself, one workgroup, plane by plane
template <class GridT>
__device__ __forceinline__ void kPassWave(const GridT& wg, SweepBox b, Sweep s,
                                          const double2* __restrict__ coef, double* field) {
  const int planes = s.ni + s.nj + s.nk - 2;
  for (int p = 0; p < planes; p++) {
    const int jLo = max(0, p - (s.ni - 1) - (s.nk - 1)), jHi = min(s.nj - 1, p);
    const int kLo = max(0, p - (s.ni - 1) - (s.nj - 1)), kHi = min(s.nk - 1, p);
    const int wj = jHi - jLo + 1, wk = kHi - kLo + 1;
    const int jk = wj > 0 && wk > 0 ? wj * wk : 0;
    for (int t = threadIdx.x; t < jk; t += blockDim.x) {
      const int jp = jLo + t % wj, kp = kLo + t / wj, ip = p - jp - kp;
      if (ip < 0 || ip >= s.ni) continue;
      const int i = s.i0 + s.di * ip, j = s.j0 + s.dj * jp, k = s.k0 + s.dk * kp;
      const double2 ab = coef[(kp * s.nj + jp) * (long)s.ni + ip];
      const long c = 3 * nodeAt(b, i, j, k), a = 3 * nodeAt(b, i - s.di, j, k),
                 bb = 3 * nodeAt(b, i, j - s.dj, k), e = 3 * nodeAt(b, i, j, k - s.dk);
      double fa[3], fb[3], fe[3];
#pragma unroll
      for (int q = 0; q < 3; q++) {
        fa[q] = field[a + q];
        fb[q] = field[bb + q];
        fe[q] = field[e + q];
      }
      if (isnan(ab.x)) continue;  // not updated in this pass
      const double alpha = ab.x, beta = ab.y, gamma = 1 - alpha - beta;
#pragma unroll
      for (int q = 0; q < 3; q++) field[c + q] = alpha * fa[q] + beta * fb[q] + gamma * fe[q];
    }
    __syncthreads();
  }
}
ESF_BOUNDS(kPassWave, kWaveThreads);

// sweepVectorField's face copies, one axis per launch (z, then y, then x):
// each thread copies both faces of its line in the reference's order
template <class GridT>
__device__ __forceinline__ void kFaceCopy(const GridT& wg, SweepBox b, int axis,
                                          const double* __restrict__ mass, double* field) {
  const int lo[3] = {b.x0, b.y0, b.z0}, hi[3] = {b.x1, b.y1, b.z1};
  const int u = axis == 0 ? 1 : 0, w = axis == 2 ? 1 : 2;  // the two other axes
  const long nu = hi[u] - lo[u], nw = hi[w] - lo[w];
  const long t = wg.block() * (long)blockDim.x + threadIdx.x;
  if (t >= nu * nw) return;
  int pos[3];
  pos[u] = lo[u] + (int)(t / nw);
  pos[w] = lo[w] + (int)(t % nw);
  auto copy = [&](int to, int from) {
    pos[axis] = to;
    const long c = nodeAt(b, pos[0], pos[1], pos[2]);
    pos[axis] = from;
    const long f = nodeAt(b, pos[0], pos[1], pos[2]);
    if (mass[c] == 0)
      for (int q = 0; q < 3; q++) field[3 * c + q] = field[3 * f + q];
  };
  copy(lo[axis], lo[axis] + 1);
  copy(hi[axis] - 1, hi[axis] - 2);
}
ESF_BOUNDS(kFaceCopy, 1024);

template <class GridT>
__device__ __forceinline__ void kFillFirst(const GridT& wg, int* first) {
  if (threadIdx.x < 3) first[threadIdx.x] = INT_MAX;
}
ESF_BOUNDS(kFillFirst, 1024);

// a sweep from a to b exclusive (the reference's sweepOnePass bounds)
Sweep passRange(int i0, int i1, int j0, int j1, int k0, int k1) {
  Sweep s;
  s.i0 = i0;
  s.j0 = j0;
  s.k0 = k0;
  s.di = i0 < i1 ? 1 : -1;
  s.dj = j0 < j1 ? 1 : -1;
  s.dk = k0 < k1 ? 1 : -1;
  s.ni = std::max((i1 - i0) * s.di, 0);
  s.nj = std::max((j1 - j0) * s.dj, 0);
  s.nk = std::max((k1 - k0) * s.dk, 0);
  return s;
}

// the context's sweep integers: three first-clear flags of a pass, or the
// six extent counters of computeSDF
constexpr int kSweepInts = 8;

esf_status checkBox(esf_ctx* ctx, const int box[6]) {
  const Dims& d = ctx->d;
  const int res[3] = {d.X, d.Y, d.Z};
  for (int a = 0; a < 3; a++)
    if (box[a] < 0 || box[3 + a] > res[a] || box[3 + a] - box[a] < 3) {
      ctx->err = "sweep box must lie in the grid and span at least 3 nodes per axis";
      return ESF_ERR_INVALID;
    }
  return ESF_OK;
}

esf_status sweepPhiDevice(esf_ctx* ctx, const SweepBox& b, const double* mass, double* phi) {
  for (int s = 0; s < 8; s++) {
    const int dx = (s & 2) ? -1 : 1, dy = (s & 1) ? -1 : 1, dz = (s & 4) ? -1 : 1;
    // ascending from min + 1 to max - 1, descending from max - 2 to min
    Sweep w;
    w.di = dx;
    w.dj = dy;
    w.dk = dz;
    w.i0 = dx > 0 ? b.x0 + 1 : b.x1 - 2;
    w.j0 = dy > 0 ? b.y0 + 1 : b.y1 - 2;
    w.k0 = dz > 0 ? b.z0 + 1 : b.z1 - 2;
    w.ni = b.x1 - b.x0 - 1;
    w.nj = b.y1 - b.y0 - 1;
    w.nk = b.z1 - b.z0 - 1;
    launch<ESF_K(kPhiWave)>(ctx, dim3(1), dim3(kWaveThreads), b, w, mass, phi);
  }
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

}  // namespace

esf_status sweepPhiRounds(esf_ctx* ctx, const int box[6], const double* mass, double* phi,
                          int rounds) {
  const Dims& d = ctx->d;
  if (!(box[3] > box[0] && box[4] > box[1] && box[5] > box[2])) return ESF_OK;
  const SweepBox b{d.X, d.Y, d.Z, box[0], box[1], box[2], box[3], box[4], box[5]};
  for (int r = 0; r < rounds; r++) {
    esf_status st = sweepPhiDevice(ctx, b, mass, phi);
    if (st) return st;
  }
  return ESF_OK;
}

// extendVectorField (:566-574) on device fields; the passes' flags and
// coefficients in the context's scratch (grown with the box)
esf_status extendDevice(esf_ctx* ctx, const int box[6], const double* m, const double* ph,
                        double* f) {
  esf_status st = checkBox(ctx, box);
  if (st) return st;
  const Dims& d = ctx->d;
  const long boxNodes = (long)(box[3] - box[0]) * (box[4] - box[1]) * (box[5] - box[2]);
  if (!ctx->swFirst && (st = devAlloc(ctx, &ctx->swFirst, kSweepInts))) return st;
  if ((size_t)boxNodes > ctx->swCoefCap) {
    if ((st = devAlloc(ctx, &ctx->swCoef, boxNodes))) return st;
    ctx->swCoefCap = boxNodes;
  }
  const SweepBox b{d.X, d.Y, d.Z, box[0], box[1], box[2], box[3], box[4], box[5]};
  int* first = ctx->swFirst;
  const int x0 = b.x0, y0 = b.y0, z0 = b.z0, x1 = b.x1, y1 = b.y1, z1 = b.z1;
  // sweepVectorField's eight passes (:524-545)
  const Sweep passes[8] = {
      passRange(x0 + 1, x1 - 1, y0 + 1, y1 - 1, z0 + 1, z1 - 1),
      passRange(x0 + 1, x1 - 1, y1 - 2, y0, z0 + 1, z1 - 1),
      passRange(x1 - 2, x0, y0 + 1, y1 - 1, z0 + 1, z1 - 1),
      passRange(x1 - 2, x0, y1 - 2, y0, z0 + 1, z1 - 1),
      passRange(x0 + 1, x1 - 1, y0 + 1, y1 - 1, z1 - 2, z0),
      passRange(x0 + 1, x1 - 1, y1 - 2, y0, z1 - 2, z0),
      passRange(x1 - 2, x0, y0 + 1, y1 - 1, z1 - 2, z0),
      passRange(x1 - 2, x0, y1 - 2, y0, z1 - 2, z0)};
  for (int round = 0; round < 4; round++) {
    for (const Sweep& sw : passes) {
      const long cells = (long)sw.ni * sw.nj * sw.nk;
      if (cells == 0) continue;
      launch<ESF_K(kFillFirst)>(ctx, dim3(1), dim3(64), first);
      launch<ESF_K(kPassFlags)>(ctx, dim3(std::min<long>(cdiv(cells, 256), 2048)), dim3(256), b, sw,
                                m, ph, first);
      launch<ESF_K(kPassCoef)>(ctx, dim3(std::min<long>(cdiv(cells, 256), 2048)), dim3(256), b, sw,
                               m, ph, (const int*)first, ctx->swCoef);
      launch<ESF_K(kPassWave)>(ctx, dim3(1), dim3(kWaveThreads), b, sw, (const double2*)ctx->swCoef,
                               f);
    }
    // the faces: z (:547-552), y (:554-558), x (:560-564)
    const long lines[3] = {(long)(y1 - y0) * (z1 - z0), (long)(x1 - x0) * (z1 - z0),
                           (long)(x1 - x0) * (y1 - y0)};
    for (int axis : {2, 1, 0})
      launch<ESF_K(kFaceCopy)>(ctx, dim3(cdiv(lines[axis], 256)), dim3(256), b, axis, m, f);
  }
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// the host-array entry points' device staging, allocated once per context
esf_status sweepStaging(esf_ctx* ctx) {
  const long n = ctx->d.nodes;
  esf_status st = ESF_OK;
  if (!ctx->swMass && (st = devAlloc(ctx, &ctx->swMass, n))) return st;
  if (!ctx->swPhi && (st = devAlloc(ctx, &ctx->swPhi, n))) return st;
  if (!ctx->swField && (st = devAlloc(ctx, &ctx->swField, 3 * n))) return st;
  if (!ctx->swFirst && (st = devAlloc(ctx, &ctx->swFirst, kSweepInts))) return st;
  return st;
}

}  // namespace esf

using namespace esf;

// The allocation-free computeSDF's device memory (esf_compute_sdf_noalloc):
// one array per GPU, resident while the process lives.  [0, 32) doubles hold
// the extent counters, then the mass and phi of up to kSdfScratchNodes nodes
// (the (128 + 1)^3 nodes of a 128^3-cell grid), 34.3 MB in all.  Host code
// never passes `scratch` itself to a kernel or a copy (on the host the name
// is only a stand-in): it uses the device address hipGetSymbolAddress
// returns, looked up once per device.
constexpr long kSdfScratchNodes = 129L * 129 * 129;
constexpr long kSdfScratchDoubles = 32 + 2 * kSdfScratchNodes;
static __device__ __align__(256) double scratch[kSdfScratchDoubles];

namespace esf {
namespace {

// the extent counters start at (X, Y, Z, 0, 0, 0): written by a kernel,
// so the call works from memory the host cannot address
template <class GridT>
__device__ __forceinline__ void kExtInit(const GridT& wg, int* ext, int X, int Y, int Z) {
  if (threadIdx.x < 6) ext[threadIdx.x] = threadIdx.x == 0 ? X : threadIdx.x == 1 ? Y
                                          : threadIdx.x == 2 ? Z : 0;
}
ESF_BOUNDS(kExtInit, 1024);

// computeSDF on the device buffers dm (the mass, n doubles), dphi (n
// doubles) and dext (6 ints), all the caller's: boxes to the host, phi back
esf_status computeSdfOn(esf_ctx* ctx, const double* mass, double* phi, int box[6], int tight[6],
                        double* dm, double* dphi, int* dext) {
  const Dims& d = ctx->d;
  const long n = d.nodes;
  ESF_HIP(hipMemcpyAsync(dm, mass, 8 * n, hipMemcpyHostToDevice, ctx->stream));
  launch<ESF_K(kExtInit)>(ctx, dim3(1), dim3(64), dext, d.X, d.Y, d.Z);
  launch<ESF_K(kSdfInit)>(ctx, dim3(std::min<long>(cdiv(n, 256), 4096)), dim3(256), n, d.X, d.Y,
                          (const double*)dm, (double)(d.X + d.Y + d.Z + 2), dphi, dext);
  int ext[6];
  ESF_HIP(hipMemcpyAsync(ext, dext, sizeof(ext), hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(streamSync(ctx));
  const int res[3] = {d.X, d.Y, d.Z};
  const int tb = ctx->p.tight_border, ex = ctx->p.extrapolation_cells;
  for (int a = 0; a < 3; a++) {
    tight[a] = std::max(ext[a] - tb, 0);
    tight[3 + a] = std::min(ext[3 + a] + tb, res[a]);
    box[a] = std::max(ext[a] - ex, 0);
    box[3 + a] = std::min(ext[3 + a] + ex, res[a]);
  }
  // no node with mass: the box is empty and the reference's loops run empty
  const bool sweep = box[3] > box[0] && box[4] > box[1] && box[5] > box[2];
  if (sweep) {
    const SweepBox b{d.X, d.Y, d.Z, box[0], box[1], box[2], box[3], box[4], box[5]};
    for (int r = 0; r < 2; r++) {
      esf_status st = sweepPhiDevice(ctx, b, dm, dphi);
      if (st) return st;
    }
  }
  ESF_HIP(hipMemcpyAsync(phi, dphi, 8 * n, hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

constexpr int kScratchDevices = 64;
std::mutex gScratchLock;
double* gScratch[kScratchDevices] = {};

// the scratch array's address on device `dev`, looked up with that device
// current on the first call there
esf_status scratchOn(esf_ctx* ctx, int dev, double** out) {
  if (dev < 0 || dev >= kScratchDevices) {
    ctx->err = "device ordinal out of range for the scratch table";
    return ESF_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(gScratchLock);
  if (!gScratch[dev]) {
    int cur = 0;
    ESF_HIP(hipGetDevice(&cur));
    if (cur != dev) ESF_HIP(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    const hipError_t back = cur != dev ? hipSetDevice(cur) : hipSuccess;
    if (back != hipSuccess) {
      ctx->err = std::string("hipSetDevice back to the caller's device: ") +
                 hipGetErrorString(back);
      return ESF_ERR_DEVICE;
    }
    if (e != hipSuccess || !p) {
      ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return ESF_ERR_DEVICE;
    }
    gScratch[dev] = (double*)p;
  }
  *out = gScratch[dev];
  return ESF_OK;
}

}  // namespace
}  // namespace esf

esf_status esf_compute_sdf(esf_ctx* ctx, const double* mass, double* phi, int box[6],
                           int tight[6]) {
  if (!ctx || !mass || !phi || !box || !tight) return ESF_ERR_INVALID;
  // staging held by the context (allocated on the first call)
  esf_status st = sweepStaging(ctx);
  if (st) return st;
  return computeSdfOn(ctx, mass, phi, box, tight, ctx->swMass, ctx->swPhi, ctx->swFirst);
}

esf_status esf_compute_sdf_noalloc(esf_ctx* ctx, const double* mass, double* phi, int box[6],
                                   int tight[6]) {
  if (!ctx || !mass || !phi || !box || !tight) return ESF_ERR_INVALID;
  if (ctx->d.nodes > kSdfScratchNodes) {
    ctx->err = "compute_sdf_noalloc: the grid has " + std::to_string(ctx->d.nodes) +
               " nodes, more than the " + std::to_string(kSdfScratchNodes) +
               " its scratch holds";
    return ESF_ERR_INVALID;
  }
  double* base = nullptr;
  esf_status st = scratchOn(ctx, ctx->device, &base);
  if (st) return st;
  return computeSdfOn(ctx, mass, phi, box, tight, base + 32, base + 32 + kSdfScratchNodes,
                      (int*)base);
}

esf_status esf_extend_vector_field(esf_ctx* ctx, const int box[6], const double* mass,
                                   const double* phi, double* field) {
  if (!ctx || !box || !mass || !phi || !field) return ESF_ERR_INVALID;
  esf_status st = checkBox(ctx, box);
  if (st) return st;
  // staging held by the context (allocated on the first call)
  if ((st = sweepStaging(ctx))) return st;
  const long n = ctx->d.nodes;
  ESF_HIP(hipMemcpyAsync(ctx->swMass, mass, 8 * n, hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(hipMemcpyAsync(ctx->swPhi, phi, 8 * n, hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(hipMemcpyAsync(ctx->swField, field, 24 * n, hipMemcpyHostToDevice, ctx->stream));
  if ((st = extendDevice(ctx, box, ctx->swMass, ctx->swPhi, ctx->swField))) return st;
  ESF_HIP(hipMemcpyAsync(field, ctx->swField, 24 * n, hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(streamSync(ctx));
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}
