// esf_flip.hip -- the FLIP material particles of a solid on the GPU
// (DEFOR_SOLID_3D.cpp:333-367 initMaterialParticles, :891-1034
// gridToParticle, :1036-1107 particleToGrid).
//
//  - initMaterialParticles: one thread per particle.  The reference draws
//    the offsets from a MERSENNETWISTER seeded with 123456 at every call
//    (MERSENNETWISTER.h:259-263), so the k-th particle of a step always
//    takes draws 3k, 3k + 1, 3k + 2 (x, y, z: the offset triple's arguments
//    evaluated left to right, as clang does): the host keeps a device
//    table of those draws (std::mt19937(123456), the same generator) and
//    grows it with the particle count.
//  - gridToParticle: one thread per particle, the 4^3 cubic B-spline
//    stencil over the solid's nodes with mass, then the walls.  Contact
//    between solids (:969-995) reads the other solid's rest-space SDF,
//    which this library does not keep, so the round trip takes scenes of
//    one solid.
//  - particleToGrid: the reference scatters every particle into its 4^3
//    nodes.  Here a node gathers instead: particles are binned by cell
//    (counting sort, each bin ordered by particle index), and every node
//    visits the 64 bins whose particles' stencils can cover it, so each
//    node's sums are formed in a fixed order (bins, then particles) and the
//    result does not depend on the schedule.  The order differs from the
//    reference's particle order, so sums agree to rounding.
#include <random>

#include "esf_internal.h"
#include "esf_interp.h"

namespace esf {

namespace {

// the cubic B-spline weight of one axis (:917-925)
__device__ inline double bspline(double d) {
  const double a = fabs(d);
  if (a < 1) return 0.5 * fabs(pow(d, 3.0)) - d * d + 2.0 / 3;
  if (a < 2) return -1.0 / 6 * fabs(pow(d, 3.0)) + d * d - 2 * a + 4.0 / 3;
  return 0;
}

template <class GridT>
__device__ __forceinline__ void kInitParticles(const GridT& wg, Dims d, int nValid, int ppc,
                                               const int* __restrict__ cellList,
                                               const double* __restrict__ rnd,
                                               const double* __restrict__ vel, double* pos,
                                               double* pv) {
  const long k = wg.block() * (long)blockDim.x + threadIdx.x;
  if (k >= (long)nValid * ppc) return;
  const long node = cellList[k / ppc];
  int x, y, z;
  nodeXYZ(d, node, x, y, z);
  const double grid[3] = {d.o0 + x * d.dh, d.o1 + y * d.dh, d.o2 + z * d.dh};
  double p[3];
  for (int c = 0; c < 3; c++) p[c] = grid[c] + rnd[3 * k + c] * d.dh;
  double v[3];
  interpV(vel, gridGeom(d), p, v);
  for (int c = 0; c < 3; c++) {
    pos[3 * k + c] = p[c];
    pv[3 * k + c] = v[c];
  }
}
ESF_BOUNDS(kInitParticles, 1024);

struct G2PArgs {
  long n;
  double* pos;
  double* pv;
  const double* mass;   // the solid's _mass
  const double* vel;    // the solid's _velocity (the grid's, just copied)
  const double* velws;  // the solid's _vecWorkspace (its last step's velocity)
  double alpha, dt, thick, ks, friction, pmass;
};

template <class GridT>
__device__ __forceinline__ void kG2P(const GridT& wg, Dims d, G2PArgs a) {
  const long c = wg.block() * (long)blockDim.x + threadIdx.x;
  if (c >= a.n) return;
  double pp[3], pv[3];
  for (int q = 0; q < 3; q++) {
    pp[q] = a.pos[3 * c + q];
    pv[q] = a.pv[3 * c + q];
  }
  const double o[3] = {d.o0, d.o1, d.o2};
  const int res[3] = {d.X, d.Y, d.Z};
  int base[3];
  for (int q = 0; q < 3; q++) base[q] = (int)floor((pp[q] - o[q]) / d.dh) - 1;
  double pic[3] = {0, 0, 0}, flip[3] = {0, 0, 0};
  for (int r = base[2]; r < base[2] + 4; r++)
    for (int t = base[1]; t < base[1] + 4; t++)
      for (int s = base[0]; s < base[0] + 4; s++) {
        if (s < 0 || t < 0 || r < 0 || s > d.X - 1 || t > d.Y - 1 || r > d.Z - 1) continue;
        const long n = r * d.slab + (long)t * d.X + s;
        if (a.mass[n] == 0) continue;
        const double df0 = d.dhInv * (pp[0] - o[0]) - s;
        const double df1 = d.dhInv * (pp[1] - o[1]) - t;
        const double df2 = d.dhInv * (pp[2] - o[2]) - r;
        const double w = (bspline(df0) * bspline(df1)) * bspline(df2);
        for (int q = 0; q < 3; q++) {
          pic[q] += a.vel[3 * n + q] * w;
          flip[q] += (a.vel[3 * n + q] - a.velws[3 * n + q]) * w;
        }
      }
  for (int q = 0; q < 3; q++) {
    flip[q] += pv[q];
    pv[q] = (1 - a.alpha) * pic[q] + a.alpha * flip[q];
  }
  // the walls (:996-1031): the first wall within `thick` that the particle
  // approaches gets a spring/friction impulse against the solid's velocity
  // at the wall node beside it
  int cell[3];
  for (int q = 0; q < 3; q++)
    cell[q] = min(max((int)floor((pp[q] - o[q]) / d.dh), 0), res[q] - 2);
  for (int b = 0; b < 6; b++) {
    const int ax = b % 3;
    const double bound = b < 3 ? o[ax] : o[ax] + d.dh * (res[ax] - 1);
    const double dist = fabs(pp[ax] - bound);
    if (!(dist < a.thick)) continue;
    const double overlap = a.thick - dist;
    int idx[3] = {cell[0], cell[1], cell[2]};
    idx[ax] = b < 3 ? 0 : res[ax] - 1;
    const long n = idx[2] * d.slab + (long)idx[1] * d.X + idx[0];
    double nb[3] = {0, 0, 0};
    nb[ax] = b < 3 ? 1.0 : -1.0;
    double vrel[3];
    for (int q = 0; q < 3; q++) vrel[q] = pv[q] - a.vel[3 * n + q];
    const double vn = vrel[0] * nb[0] + (vrel[1] * nb[1] + vrel[2] * nb[2]);
    if (vn <= 0) {
      const double impulse = -fmin(a.dt * a.ks * overlap, a.pmass * (0.1 * overlap / a.dt - vn));
      const double dvn = -impulse / a.pmass;
      double vt[3];
      for (int q = 0; q < 3; q++) vt[q] = vrel[q] - vn * nb[q];
      const double vtn = sqrt(vt[0] * vt[0] + (vt[1] * vt[1] + vt[2] * vt[2]));
      const double sc = fmax(0.1 - a.friction * dvn / vtn, 0.0);
      for (int q = 0; q < 3; q++) pv[q] = ((vn + dvn) * nb[q] + sc * vt[q]) + a.vel[3 * n + q];
      break;
    }
  }
  for (int q = 0; q < 3; q++) {
    a.pos[3 * c + q] = pp[q] + pv[q] * a.dt;
    a.pv[3 * c + q] = pv[q];
  }
}
ESF_BOUNDS(kG2P, 256);

// bins: the cells of the grid widened by two on every side, a particle's
// bin from its floor((pos - o) / dh) clamped there
struct Bins {
  int nx, ny, nz;
  __host__ __device__ long count() const { return (long)nx * ny * nz; }
};

__device__ inline int binOf(const Dims& d, const Bins& b, const double* p) {
  const double o[3] = {d.o0, d.o1, d.o2};
  const int lim[3] = {b.nx, b.ny, b.nz};
  int q[3];
  for (int k = 0; k < 3; k++) {
    const double f = floor((p[k] - o[k]) / d.dh);
    q[k] = (int)fmin(fmax(f + 2.0, 0.0), (double)(lim[k] - 1));
  }
  return q[0] + b.nx * (q[1] + b.ny * q[2]);
}

template <class GridT>
__device__ __forceinline__ void kZeroInts(const GridT& wg, long n, int* a) {
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)wg.blocks() * blockDim.x)
    a[i] = 0;
}
ESF_BOUNDS(kZeroInts, 1024);

template <class GridT>
__device__ __forceinline__ void kBinCount(const GridT& wg, Dims d, Bins b, long n,
                                          const double* __restrict__ pos, int* pBin, int* count) {
  const long c = wg.block() * (long)blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int bin = binOf(d, b, &pos[3 * c]);
  pBin[c] = bin;
  atomicAdd(&count[bin], 1);
}
ESF_BOUNDS(kBinCount, 1024);

// exclusive scan of counts[0, n) into starts, in tiles of 1024 (block
// totals into blocks[], scanned by one block, then added back)
template <class GridT>
__device__ __forceinline__ void kTileSums(const GridT& wg, long n, const int* __restrict__ count,
                                          int* blocks) {
  __shared__ int red[256];
  const long base = (long)wg.block() * 1024;
  int s = 0;
  for (int k = 0; k < 4; k++) {
    const long i = base + threadIdx.x * 4 + k;
    if (i < n) s += count[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) blocks[wg.block()] = red[0];
}
ESF_BOUNDS(kTileSums, 1024);

template <class GridT>
__device__ __forceinline__ void kTileScan(const GridT& wg, int nb, int* blocks) {
  if (threadIdx.x != 0) return;
  int acc = 0;
  for (int i = 0; i < nb; i++) {
    const int v = blocks[i];
    blocks[i] = acc;
    acc += v;
  }
}
ESF_BOUNDS(kTileScan, 1024);

template <class GridT>
__device__ __forceinline__ void kTileEmit(const GridT& wg, long n, const int* __restrict__ count,
                                          const int* __restrict__ blocks, int* starts, int* cursor) {
  __shared__ int pre[256];
  const long base = (long)wg.block() * 1024;
  int v[4], s = 0;
  for (int k = 0; k < 4; k++) {
    const long i = base + threadIdx.x * 4 + k;
    v[k] = i < n ? count[i] : 0;
    s += v[k];
  }
  pre[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int t = 0; t < 256; t++) {
      const int x = pre[t];
      pre[t] = acc;
      acc += x;
    }
  }
  __syncthreads();
  int acc = blocks[wg.block()] + pre[threadIdx.x];
  for (int k = 0; k < 4; k++) {
    const long i = base + threadIdx.x * 4 + k;
    if (i < n) {
      starts[i] = acc;
      cursor[i] = 0;
    }
    acc += v[k];
  }
  if (base + 1024 >= n && threadIdx.x == 255) starts[n] = acc;
}
ESF_BOUNDS(kTileEmit, 1024);

template <class GridT>
__device__ __forceinline__ void kBinFill(const GridT& wg, long n, const int* __restrict__ pBin,
                                         const int* __restrict__ starts, int* cursor, int* items) {
  const long c = wg.block() * (long)blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int b = pBin[c];
  items[starts[b] + atomicAdd(&cursor[b], 1)] = (int)c;
}
ESF_BOUNDS(kBinFill, 1024);

// each bin's particles in index order (insertion sort; bins hold a few)
template <class GridT>
__device__ __forceinline__ void kBinSort(const GridT& wg, long nbins, const int* __restrict__ starts,
                                         int* items) {
  const long b = wg.block() * (long)blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  const int lo = starts[b], hi = starts[b + 1];
  for (int i = lo + 1; i < hi; i++) {
    const int v = items[i];
    int j = i - 1;
    while (j >= lo && items[j] > v) {
      items[j + 1] = items[j];
      j--;
    }
    items[j + 1] = v;
  }
}
ESF_BOUNDS(kBinSort, 1024);

struct P2GArgs {
  Bins b;
  const int* starts;
  const int* items;
  const double* pos;
  const double* pv;
  double cutoff;
  double* vel;   // the solid's _velocity: the weighted average where the weight reaches the cutoff
  double* mass;  // the solid's _mass: 1 there, 0 elsewhere
};

template <class GridT>
__device__ __forceinline__ void kP2G(const GridT& wg, Dims d, P2GArgs a) {
  const long n = wg.block() * (long)blockDim.x + threadIdx.x;
  if (n >= d.nodes) return;
  int x, y, z;
  nodeXYZ(d, n, x, y, z);
  const double o[3] = {d.o0, d.o1, d.o2};
  double vs[3] = {0, 0, 0}, ws = 0;
  bool touched = false;
  // cells floor(pos) in [node - 2, node + 1]: bins (node - 2 .. node + 1) + 2
  for (int bz = z; bz < z + 4; bz++)
    for (int by = y; by < y + 4; by++)
      for (int bx = x; bx < x + 4; bx++) {
        const long bin = bx + a.b.nx * ((long)by + a.b.ny * bz);
        const int lo = a.starts[bin], hi = a.starts[bin + 1];
        for (int i = lo; i < hi; i++) {
          const int c = a.items[i];
          const double* pp = &a.pos[3 * (long)c];
          // the particle's own stencil must cover the node (a clamped bin
          // holds particles beyond the grid)
          bool cover = true;
          const int node[3] = {x, y, z};
          for (int k = 0; k < 3; k++) {
            const int f = (int)floor((pp[k] - o[k]) / d.dh);
            cover = cover && node[k] >= f - 1 && node[k] <= f + 2;
          }
          if (!cover) continue;
          const double df0 = d.dhInv * (pp[0] - o[0]) - x;
          const double df1 = d.dhInv * (pp[1] - o[1]) - y;
          const double df2 = d.dhInv * (pp[2] - o[2]) - z;
          const double w = (bspline(df0) * bspline(df1)) * bspline(df2);
          const double* v = &a.pv[3 * (long)c];
          if (!touched) {
            touched = true;
            for (int q = 0; q < 3; q++) vs[q] = w * v[q];
            ws = w;
          } else {
            for (int q = 0; q < 3; q++) vs[q] += w * v[q];
            ws += w;
          }
        }
      }
  if (touched && !(ws < a.cutoff)) {
    for (int q = 0; q < 3; q++) a.vel[3 * n + q] = vs[q] / ws;
    a.mass[n] = 1;
  } else {
    a.mass[n] = 0;
  }
}
ESF_BOUNDS(kP2G, 256);

Bins binsOf(const Dims& d) { return Bins{d.X + 4, d.Y + 4, d.Z + 4}; }

// the particle arrays of a solid, for n particles
esf_status growParticles(esf_ctx* ctx, DevSolid& so, long n) {
  if (n <= so.capPart && so.pPos) return ESF_OK;
  const long cap = std::max<long>(n + n / 4 + 1024, 4096);
  esf_status st;
  if ((st = devAlloc(ctx, &so.pPos, 3 * cap))) return st;
  if ((st = devAlloc(ctx, &so.pVel, 3 * cap))) return st;
  if ((st = devAlloc(ctx, &so.pBin, cap))) return st;
  if ((st = devAlloc(ctx, &so.binItems, cap))) return st;
  so.capPart = cap;
  return ESF_OK;
}

// the first `draws` rand() values of MT19937 seeded with 123456, on the device
esf_status growRandTable(esf_ctx* ctx, long draws) {
  if (draws <= ctx->randCap && ctx->randTable) return ESF_OK;
  const long cap = std::max<long>(draws + draws / 4 + 3072, 3 << 12);
  std::vector<double> h(cap);
  std::mt19937 twister(123456);
  for (long i = 0; i < cap; i++) h[i] = double(twister()) * (1.0 / 4294967295.0);
  esf_status st = devAlloc(ctx, &ctx->randTable, cap);
  if (st) return st;
  ESF_HIP(hipMemcpyAsync(ctx->randTable, h.data(), 8 * cap, hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(streamSync(ctx));
  ctx->randCap = cap;
  return ESF_OK;
}

}  // namespace

esf_status growParticlesFor(esf_ctx* ctx, int solid, long n) {
  return growParticles(ctx, ctx->solids[solid], n);
}

esf_status initParticles(esf_ctx* ctx, int solid) {
  DevSolid& so = ctx->solids[solid];
  const int ppc = ctx->p.particles_per_cell;
  if (ppc < 1) {
    ctx->err = "number of particles per material cell must be at least 1";
    return ESF_ERR_INVALID;
  }
  const long n = (long)so.nValid * ppc;
  esf_status st;
  if ((st = growParticles(ctx, so, n))) return st;
  if ((st = growRandTable(ctx, 3 * n))) return st;
  so.nPart = n;
  if (n)
    launch<ESF_K(kInitParticles)>(ctx, dim3(cdiv(n, 256)), dim3(256), ctx->d, so.nValid, ppc,
                                  (const int*)so.cellList, (const double*)ctx->randTable,
                                  (const double*)so.vel, so.pPos, so.pVel);
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status gridToParticle(esf_ctx* ctx, int solid) {
  DevSolid& so = ctx->solids[solid];
  if (so.nPart == 0) return ESF_OK;
  const Dims& d = ctx->d;
  G2PArgs a{};
  a.n = so.nPart;
  a.pos = so.pPos;
  a.pv = so.pVel;
  a.mass = so.mass;
  a.vel = so.vel;
  a.velws = so.velws;
  a.alpha = ctx->p.flip_alpha;
  a.dt = ctx->dt;
  a.thick = ctx->p.wall_thickness;
  a.ks = ctx->p.spring_stiffness;
  a.friction = ctx->p.friction;
  // the particle mass: density * volume, function-local statics of the
  // first solid that created particles (DEFOR_SOLID_3D.cpp:338-340)
  const double dh3 = d.dh * d.dh * d.dh;
  a.pmass = (ctx->solids[0].qcm / (dh3 * 0.125)) * (dh3 / ctx->p.particles_per_cell);
  launch<ESF_K(kG2P)>(ctx, dim3(cdiv(a.n, 256)), dim3(256), d, a);
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status particleToGrid(esf_ctx* ctx, int solid) {
  DevSolid& so = ctx->solids[solid];
  const Dims& d = ctx->d;
  const Bins b = binsOf(d);
  const long nb = b.count();
  esf_status st;
  if (!ctx->binCount) {
    if ((st = devAlloc(ctx, &ctx->binCount, nb))) return st;
    if ((st = devAlloc(ctx, &ctx->binStart, nb + 1))) return st;
    if ((st = devAlloc(ctx, &ctx->binCursor, nb))) return st;
    if ((st = devAlloc(ctx, &ctx->binBlocks, cdiv(nb, 1024) + 1))) return st;
  }
  const long n = so.nPart;
  if (n && (st = growParticles(ctx, so, n))) return st;
  launch<ESF_K(kZeroInts)>(ctx, dim3(std::min<long>(cdiv(nb, 256), 4096)), dim3(256), nb,
                           ctx->binCount);
  if (n)
    launch<ESF_K(kBinCount)>(ctx, dim3(cdiv(n, 256)), dim3(256), d, b, n, (const double*)so.pPos,
                             so.pBin, ctx->binCount);
  const int tiles = cdiv(nb, 1024);
  launch<ESF_K(kTileSums)>(ctx, dim3(tiles), dim3(256), nb, (const int*)ctx->binCount,
                           ctx->binBlocks);
  launch<ESF_K(kTileScan)>(ctx, dim3(1), dim3(64), tiles, ctx->binBlocks);
  launch<ESF_K(kTileEmit)>(ctx, dim3(tiles), dim3(256), nb, (const int*)ctx->binCount,
                           (const int*)ctx->binBlocks, ctx->binStart, ctx->binCursor);
  if (n) {
    launch<ESF_K(kBinFill)>(ctx, dim3(cdiv(n, 256)), dim3(256), n, (const int*)so.pBin,
                            (const int*)ctx->binStart, ctx->binCursor, so.binItems);
    launch<ESF_K(kBinSort)>(ctx, dim3(cdiv(nb, 256)), dim3(256), nb, (const int*)ctx->binStart,
                            so.binItems);
  }
  P2GArgs a{};
  a.b = b;
  a.starts = ctx->binStart;
  a.items = so.binItems;
  a.pos = so.pPos;
  a.pv = so.pVel;
  a.cutoff = ctx->p.flip_cutoff;
  a.vel = so.vel;
  a.mass = so.mass;
  launch<ESF_K(kP2G)>(ctx, dim3(cdiv(d.nodes, 256)), dim3(256), d, a);
  // _mass now holds the particles' 0/1 indicator over the whole grid: the
  // next mass pass clears and rebuilds all of it
  so.listDirty = true;
  so.dirtyAll = true;
  so.wsState = 2;  // the solid's velocity changed beyond the copied box
  ctx->massGen++;
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

}  // namespace esf
