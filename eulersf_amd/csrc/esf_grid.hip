// esf_grid.hip -- node-grid stages of the coupled step: solid mass gathering
// and solid DOF numbering, the solid mass resampling used at scene setup,
// CFL, vorticity and body forces, the scatter/correct steps of the Schur
// solve, semi-Lagrangian advection and the boundary conditions.
//
// Every node-parallel kernel is a gather: each thread writes only its own
// node, and sums that the reference builds by scattering (mass gathering,
// COUPLED_INTEGRATOR_3D.cpp:346-359; solid mass, DEFOR_SOLID_3D.cpp:655-667)
// are re-ordered into per-node sums in the reference's addition order.
#include <algorithm>

#include "esf_internal.h"
#include "esf_interp.h"

namespace esf {

__device__ __forceinline__ bool isInterior(const Dims& d, int x, int y, int z) {
  return x >= 1 && x <= d.X - 2 && y >= 1 && y <= d.Y - 2 && z >= 1 && z <= d.Z - 2;
}

// ---------------------------------------------------------------------------
// stream compaction (exclusive scan of flags) in three launches
// ---------------------------------------------------------------------------
constexpr int kScanTile = 1024;  // 256 threads x 4

// inclusive prefix sum over the 64 lanes of a wave
__device__ inline int waveInclusiveScan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// exclusive prefix sum of v over the block (a whole number of waves, at
// most 16) and, in *total, the block's sum; wsum is 16 ints of LDS
__device__ inline int blockExclusiveScan(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int inc = waveInclusiveScan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int pre = 0, all = 0;
  for (int k = 0; k < nw; k++) {
    const int t = wsum[k];
    pre += k < w ? t : 0;
    all += t;
  }
  *total = all;
  return pre + inc - v;
}

// The compactions walk a box of nodes or cells (BoxWalk, defined below):
// each thread takes 4 consecutive elements of the box, x fastest, decoding
// the first one's coordinates and stepping to the others.
template <class Walk, class Flag, class GridT>
__device__ __forceinline__ void kScanCount(const GridT& wg, Walk w, long n, Flag flag,
                                           int* blockCounts) {
  __shared__ int wsum[16];
  const long t0 = (long)wg.block() * kScanTile + threadIdx.x * 4;
  int c = 0;
  if (t0 < n) {
    int x, y, z;
    w.start(t0, x, y, z);
    for (int k = 0; k < 4 && t0 + k < n; k++, w.next(x, y, z))
      if (flag(x, y, z, w.index(x, y, z))) c++;
  }
  int total;
  blockExclusiveScan(c, wsum, &total);
  if (threadIdx.x == 0) blockCounts[wg.block()] = total;
}
ESF_BOUNDS(kScanCount, 1024);

// exclusive scan of nb block counts by one block; total into blockCounts[nb]
__device__ inline int scanBlockCounts(int nb, int* blockCounts, int* wsum) {
  const int per = (nb + 1023) / 1024;
  const int b0 = threadIdx.x * per;
  int s = 0;
  for (int i = 0; i < per; i++)
    if (b0 + i < nb) s += blockCounts[b0 + i];
  int total;
  int acc = blockExclusiveScan(s, wsum, &total);
  for (int i = 0; i < per; i++)
    if (b0 + i < nb) {
      const int v = blockCounts[b0 + i];
      blockCounts[b0 + i] = acc;
      acc += v;
    }
  if (threadIdx.x == 0) blockCounts[nb] = total;
  return total;
}
template <class GridT>
__device__ __forceinline__ void kScanBlocks(const GridT& wg, int nb, int* blockCounts) {
  __shared__ int wsum[16];
  (void)scanBlockCounts(nb, blockCounts, wsum);
}
ESF_BOUNDS(kScanBlocks, 1024);


template <class Walk, class Flag, class Emit, class GridT>
__device__ __forceinline__ void kScanEmit(const GridT& wg, Walk w, long n, Flag flag,
                                          const int* blockCounts, Emit emit) {
  __shared__ int wsum[16];
  const long t0 = (long)wg.block() * kScanTile + threadIdx.x * 4;
  int f[4] = {0, 0, 0, 0}, c = 0;
  int x = 0, y = 0, z = 0;
  if (t0 < n) {
    w.start(t0, x, y, z);
    int xs = x, ys = y, zs = z;
    for (int k = 0; k < 4 && t0 + k < n; k++, w.next(xs, ys, zs)) {
      f[k] = flag(xs, ys, zs, w.index(xs, ys, zs)) ? 1 : 0;
      c += f[k];
    }
  }
  int total;
  int idx = blockCounts[wg.block()] + blockExclusiveScan(c, wsum, &total);
  if (t0 >= n) return;
  for (int k = 0; k < 4 && t0 + k < n; k++, w.next(x, y, z)) {
    emit(x, y, z, w.index(x, y, z), f[k] ? idx : -1);
    idx += f[k];
  }
}
ESF_BOUNDS(kScanEmit, 1024);

// Two compactions in one set of launches (a step's DOF numbering and a
// solid's valid-cell list): blocks [0, nbA) take walk A, the next nbB walk
// B; A's block counts sit at [0, nbA) with its total behind them, B's at
// [nbA + 1, nbA + 1 + nbB) with its total behind, and the scan block writes
// both totals to pinned host memory.
template <class WA, class FA, class WB, class FB, class GridT>
__device__ __forceinline__ void kScanCount2(const GridT& wg, WA wa, long na, FA fa, int nbA,
                                            WB wb, long nb, FB fb, int* blockCounts) {
  __shared__ int wsum[16];
  const bool isA = (int)wg.block() < nbA;  // uniform per block
  const int lb = isA ? wg.block() : wg.block() - nbA;
  const long t0 = (long)lb * kScanTile + threadIdx.x * 4;
  int c = 0;
  int x, y, z;
  if (isA) {
    if (t0 < na) {
      wa.start(t0, x, y, z);
      for (int k = 0; k < 4 && t0 + k < na; k++, wa.next(x, y, z))
        if (fa(x, y, z, wa.index(x, y, z))) c++;
    }
  } else if (t0 < nb) {
    wb.start(t0, x, y, z);
    for (int k = 0; k < 4 && t0 + k < nb; k++, wb.next(x, y, z))
      if (fb(x, y, z, wb.index(x, y, z))) c++;
  }
  int total;
  blockExclusiveScan(c, wsum, &total);
  if (threadIdx.x == 0) blockCounts[isA ? lb : nbA + 1 + lb] = total;
}
ESF_BOUNDS(kScanCount2, 1024);

template <class GridT>
__device__ __forceinline__ void kScanBlocks2(const GridT& wg, int nbA, int nbB, int* blockCounts,
                                             int* hostA, int* hostB) {
  __shared__ int wsum[16];
  const int ta = scanBlockCounts(nbA, blockCounts, wsum);
  __syncthreads();
  const int tb = scanBlockCounts(nbB, blockCounts + nbA + 1, wsum);
  if (threadIdx.x == 0) {
    *hostA = ta;
    *hostB = tb;
  }
}
ESF_BOUNDS(kScanBlocks2, 1024);

// one block's emit of one walk (kScanEmit's body)
template <class Walk, class Flag, class Emit>
__device__ inline void emitTile(int lb, Walk w, long n, Flag flag, int base, Emit emit,
                                int* wsum) {
  const long t0 = (long)lb * kScanTile + threadIdx.x * 4;
  int f[4] = {0, 0, 0, 0}, c = 0;
  int x = 0, y = 0, z = 0;
  if (t0 < n) {
    w.start(t0, x, y, z);
    int xs = x, ys = y, zs = z;
    for (int k = 0; k < 4 && t0 + k < n; k++, w.next(xs, ys, zs)) {
      f[k] = flag(xs, ys, zs, w.index(xs, ys, zs)) ? 1 : 0;
      c += f[k];
    }
  }
  int total;
  int idx = base + blockExclusiveScan(c, wsum, &total);
  if (t0 >= n) return;
  for (int k = 0; k < 4 && t0 + k < n; k++, w.next(x, y, z)) {
    emit(x, y, z, w.index(x, y, z), f[k] ? idx : -1);
    idx += f[k];
  }
}

template <class WA, class FA, class EA, class WB, class FB, class EB, class GridT>
__device__ __forceinline__ void kScanEmit2(const GridT& wg, WA wa, long na, FA fa, EA ea, int nbA,
                                           WB wb, long nb, FB fb, EB eb,
                                           const int* blockCounts) {
  __shared__ int wsum[16];
  const int blk = wg.block();
  if (blk < nbA)
    emitTile(blk, wa, na, fa, blockCounts[blk], ea, wsum);
  else
    emitTile(blk - nbA, wb, nb, fb, blockCounts[nbA + 1 + blk - nbA], eb, wsum);
}
ESF_BOUNDS(kScanEmit2, 1024);

template <class WA, class FA, class EA, class WB, class FB, class EB>
void compactEnqueue2(esf_ctx* ctx, WA wa, long na, FA fa, EA ea, int* hostA, WB wb, long nb,
                     FB fb, EB eb, int* hostB) {
  const int nbA = std::max(cdiv(na, kScanTile), 1), nbB = std::max(cdiv(nb, kScanTile), 1);
  launch<ESF_KT(kScanCount2, WA, FA, WB, FB)>(ctx, dim3(nbA + nbB), dim3(256), wa, na, fa, nbA,
                                              wb, nb, fb, ctx->scanBlocks);
  launch<ESF_K(kScanBlocks2)>(ctx, dim3(1), dim3(1024), nbA, nbB, ctx->scanBlocks, hostA, hostB);
  launch<ESF_KT(kScanEmit2, WA, FA, EA, WB, FB, EB)>(ctx, dim3(nbA + nbB), dim3(256), wa, na,
                                                     fa, ea, nbA, wb, nb, fb, eb,
                                                     (const int*)ctx->scanBlocks);
}

// count, scan and emit over the n elements of a walk, queued; the count is
// copied to the pinned int *hostTotal, read after a later synchronisation
template <class Walk, class Flag, class Emit>
void compactEnqueue(esf_ctx* ctx, Walk w, long n, Flag flag, Emit emit, int* hostTotal) {
  const int nb = std::max(cdiv(n, kScanTile), 1);
  launch<ESF_KT(kScanCount, Walk, Flag)>(ctx, dim3(nb), dim3(256), w, n, flag, ctx->scanBlocks);
  launch<ESF_K(kScanBlocks)>(ctx, dim3(1), dim3(1024), nb, ctx->scanBlocks);
  launch<ESF_KT(kScanEmit, Walk, Flag, Emit)>(ctx, dim3(nb), dim3(256), w, n, flag,
                                              (const int*)ctx->scanBlocks, emit);
  toHostAsync(ctx, hostTotal, ctx->scanBlocks + nb, sizeof(int));
}

// a box of nodes or cells, [lo, hi) per axis, walked x fastest
struct BoxRange {
  int lo[3], hi[3];
  __host__ __device__ long count() const {
    return (long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
  }
  __device__ void at(long t, int& x, int& y, int& z) const {
    const int nx = hi[0] - lo[0], ny = hi[1] - lo[1];
    if (t <= 0x7fffffffL) {  // 32-bit division where the index fits
      const int ti = (int)t, q = ti / nx;
      x = lo[0] + (ti - q * nx);
      y = lo[1] + q % ny;
      z = lo[2] + q / ny;
      return;
    }
    x = lo[0] + (int)(t % nx);
    y = lo[1] + (int)((t / nx) % ny);
    z = lo[2] + (int)(t / ((long)nx * ny));
  }
  __device__ long node(const Dims& d, long t) const {
    int x, y, z;
    at(t, x, y, z);
    return x + (long)y * d.X + (long)z * d.slab;
  }
};

// the box spanning a and b, clipped to [0, lim); an empty box where both are
inline BoxRange spanBoxes(const int a[6], const int b[6], const int lim[3]) {
  BoxRange r;
  const bool ea = a[0] >= a[3] || a[1] >= a[4] || a[2] >= a[5];
  const bool eb = b[0] >= b[3] || b[1] >= b[4] || b[2] >= b[5];
  for (int k = 0; k < 3; k++) {
    int lo = ea ? b[k] : eb ? a[k] : std::min(a[k], b[k]);
    int hi = ea ? b[3 + k] : eb ? a[3 + k] : std::max(a[3 + k], b[3 + k]);
    r.lo[k] = std::max(lo, 0);
    r.hi[k] = std::min(hi, lim[k]);
    if (r.hi[k] < r.lo[k]) r.hi[k] = r.lo[k];
  }
  return r;
}

// the elements of a box, walked x fastest: element t is grid element
// x + y sy + z sz, so a compaction over the box emits in the whole grid's
// ascending order (box counts stay below 2^31: the grid's do)
struct BoxWalk {
  BoxRange r;
  long sy, sz;
  __device__ void start(long t, int& x, int& y, int& z) const {
    const int nx = r.hi[0] - r.lo[0], ny = r.hi[1] - r.lo[1];
    const int ti = (int)t, q = ti / nx;
    x = r.lo[0] + (ti - q * nx);
    y = r.lo[1] + q % ny;
    z = r.lo[2] + q / ny;
  }
  __device__ void next(int& x, int& y, int& z) const {
    if (++x < r.hi[0]) return;
    x = r.lo[0];
    if (++y < r.hi[1]) return;
    y = r.lo[1];
    ++z;
  }
  __device__ long index(int x, int y, int z) const { return x + y * sy + z * sz; }
};

// ---------------------------------------------------------------------------
// gatherSolidMasses  COUPLED_INTEGRATOR_3D.cpp:327-375
// ---------------------------------------------------------------------------
struct SolidPtrs {
  int n;
  const double* mass[kMaxSolids];
  const double* qfrac[kMaxSolids];
  // per solid, the cells where its quarter fractions may be nonzero ([lo,
  // hi) per axis; its masses reach one node beyond): elsewhere both are 0,
  // and adding 0 leaves a sum unchanged, so the gather skips the load
  int cells[kMaxSolids][6];
};

__device__ inline bool inCells(const int (&r)[6], int x, int y, int z) {
  return x >= r[0] && x < r[3] && y >= r[1] && y < r[4] && z >= r[2] && z < r[5];
}
// a node is a corner of a cell of the range when each coordinate is in [lo, hi]
__device__ inline bool atCellCorner(const int (&r)[6], int x, int y, int z) {
  return x >= r[0] && x <= r[3] && y >= r[1] && y <= r[4] && z >= r[2] && z <= r[5];
}

// over the nodes of box g ([lo, hi) per axis); outside it both fields are 0
template <class GridT>
__device__ __forceinline__ void kGatherMass(const GridT& wg, Dims d, SolidPtrs sp, BoxRange g,
                                            double* solidMass, double* nodeFrac) {
  const long t = wg.block() * (long)blockDim.x + threadIdx.x;
  if (t >= g.count()) return;
  const long n = g.node(d, t);
  // (loops over the solids unrolled: constant indices into the argument
  // arrays, so the arguments stay in registers)
  int x, y, z;
  nodeXYZ(d, n, x, y, z);
  double m = 0;
#pragma unroll
  for (int b = 0; b < kMaxSolids; b++)
    if (b < sp.n && atCellCorner(sp.cells[b], x, y, z))
      m += sp.mass[b][n];
  solidMass[n] = m;
  double f = 0;
  // corner i of the cell whose min corner is n - off[i], for cells with
  // min corner in [1, X-2]^3 (the reference's loop range, :350-356)
  for (int i = 0; i < 8; i++) {
    const int ox = i & 1, oy = (i >> 1) & 1, oz = i >> 2;
    const int cx = x - ox, cy = y - oy, cz = z - oz;
    if (!isInterior(d, cx, cy, cz)) continue;
    const long c = cx + (long)cy * d.X + (long)cz * d.slab;
    double q = 0;
#pragma unroll
    for (int b = 0; b < kMaxSolids; b++)
      if (b < sp.n && inCells(sp.cells[b], cx, cy, cz)) q += sp.qfrac[b][8 * c + i];
    f += q;
  }
  nodeFrac[n] = f * 0.125;
}
ESF_BOUNDS(kGatherMass, 1024);

struct SolidFlag {
  Dims d;
  const double* solidMass;
  __device__ bool operator()(int x, int y, int z, long n) const {
    return isInterior(d, x, y, z) && solidMass[n] > 0;
  }
};
// numbers past `cap` (solidNodes' length) are counted, not stored: the
// count read later shows the list must grow and the scan run again
struct SolidEmit {
  int* sidx;
  int* solidNodes;
  int cap;
  __device__ void operator()(int, int, int, long n, int idx) const {
    sidx[n] = idx;
    if (idx >= 0 && idx < cap) solidNodes[idx] = (int)n;
  }
};

// valid-cell list of one solid: cells of the tight box with _valid == 1
struct ValidFlag {
  const Box* box;  // the solid's boxes in device memory
  const int8_t* valid;
  __device__ bool operator()(int x, int y, int z, long c) const {
    const Box b = *box;
    return x >= b.tMin[0] && x < b.tMax[0] - 1 && y >= b.tMin[1] && y < b.tMax[1] - 1 &&
           z >= b.tMin[2] && z < b.tMax[2] - 1 && valid[c] == 1;
  }
};
struct ValidEmit {
  Dims d;
  int* cellSlot;
  int* cellList;
  int cap;  // as SolidEmit's
  __device__ void operator()(int x, int y, int z, long c, int idx) const {
    cellSlot[c] = idx;
    if (idx >= 0 && idx < cap) cellList[idx] = x + y * d.X + z * (int)d.slab;
  }
};

esf_status growSolidCells(esf_ctx* ctx, DevSolid& so, int need);
esf_status growSolidDofs(esf_ctx* ctx, int need);

static BoxRange rangeOf(const int r[6]) {
  BoxRange b;
  for (int k = 0; k < 3; k++) {
    b.lo[k] = r[k];
    b.hi[k] = r[3 + k];
  }
  return b;
}

// Compactions are queued and their counts read at a later synchronisation
// (gatherFinish): a step reads them at its end, so the stages in between
// need not wait for them.
static void queueSolidDofs(esf_ctx* ctx) {
  const Dims& d = ctx->d;
  const BoxRange r = rangeOf(ctx->dofScan);
  compactEnqueue(ctx, BoxWalk{r, d.X, d.slab}, r.count(), SolidFlag{d, ctx->solidMass},
                 SolidEmit{ctx->sidx, ctx->solidNodes, ctx->capSolid},
                 ctx->hostFlags + kHostDofCount);
  ctx->dofPending = true;
}

static void queueValidScan(esf_ctx* ctx, int b) {
  DevSolid& so = ctx->solids[b];
  const Dims& d = ctx->d;
  const BoxRange r = rangeOf(so.slotScan);
  compactEnqueue(ctx, BoxWalk{r, d.CX, d.cslab}, r.count(), ValidFlag{ctx->dBoxes + b, so.valid},
                 ValidEmit{d, so.cellSlot, so.cellList, so.capValid},
                 ctx->hostFlags + kHostValidCount + b);
  so.listPending = true;
}

// the DOF numbering and solid 0's valid-cell list, queued in one set of
// launches (their scan ranges set up already)
static void queueDofsAndValid(esf_ctx* ctx) {
  DevSolid& so = ctx->solids[0];
  const Dims& d = ctx->d;
  const BoxRange ra = rangeOf(ctx->dofScan), rb = rangeOf(so.slotScan);
  compactEnqueue2(ctx, BoxWalk{ra, d.X, d.slab}, ra.count(), SolidFlag{d, ctx->solidMass},
                  SolidEmit{ctx->sidx, ctx->solidNodes, ctx->capSolid},
                  ctx->hostFlags + kHostDofCount, BoxWalk{rb, d.CX, d.cslab}, rb.count(),
                  ValidFlag{ctx->dBoxes, so.valid},
                  ValidEmit{d, so.cellSlot, so.cellList, so.capValid},
                  ctx->hostFlags + kHostValidCount);
  ctx->dofPending = true;
  so.listPending = true;
}

// the valid-cell list of solid b: its scan range, then (queue) the scan.
// The scan walks the cells of the tight box (ValidFlag's range; a box that
// contains it while the solid's boxes are still on their way to the host)
// and of the last build's, where cellSlot may be >= 0; every other cell's
// slot is -1 already and stays so.
static esf_status queueValidList(esf_ctx* ctx, int b, bool queue = true) {
  DevSolid& so = ctx->solids[b];
  const Dims& d = ctx->d;
  esf_status st = growSolidCells(ctx, so, 0);
  if (st) return st;
  const Box& bx = so.boxPending ? so.boxBound : so.box;
  const int cur[6] = {bx.tMin[0], bx.tMin[1], bx.tMin[2],
                      bx.tMax[0] - 1, bx.tMax[1] - 1, bx.tMax[2] - 1};
  const int lim[3] = {d.CX, d.CY, d.CZ};
  const BoxRange r = spanBoxes(so.slotSpan, cur, lim);
  for (int k = 0; k < 3; k++) {
    so.slotScan[k] = r.lo[k];
    so.slotScan[3 + k] = r.hi[k];
  }
  for (int k = 0; k < 6; k++) so.slotSpan[k] = cur[k];
  if (queue) queueValidScan(ctx, b);
  return ESF_OK;
}

// after a synchronisation that the queued compactions' counts precede: the
// counts; a list that outgrew its capacity is grown and scanned again
esf_status gatherFinish(esf_ctx* ctx) {
  esf_status st;
  if (ctx->dofPending) {
    int total = ctx->hostFlags[kHostDofCount];
    if (total > ctx->capSolid || (ctx->tightLists && total != ctx->capSolid)) {
      if ((st = growSolidDofs(ctx, total))) return st;
      queueSolidDofs(ctx);
      ESF_HIP(streamSync(ctx));
      total = ctx->hostFlags[kHostDofCount];
    }
    ctx->nSolid = total;
    ctx->dofPending = false;
  }
  for (int b = 0; b < (int)ctx->solids.size(); b++) {
    DevSolid& so = ctx->solids[b];
    if (!so.listPending) continue;
    int total = ctx->hostFlags[kHostValidCount + b];
    if (total > so.capValid || (ctx->tightLists && total != so.capValid)) {
      if ((st = growSolidCells(ctx, so, total))) return st;
      queueValidScan(ctx, b);
      ESF_HIP(streamSync(ctx));
      total = ctx->hostFlags[kHostValidCount + b];
    }
    so.nValid = total;
    so.listPending = false;
    so.listDirty = false;
  }
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// gatherSolidMasses  COUPLED_INTEGRATOR_3D.cpp:327-375.  `defer`: leave the
// compactions' counts for gatherFinish after the caller's next
// synchronisation (the DOF count and the valid-cell counts are not read
// before it); otherwise synchronise and finish here.
esf_status gatherSolidMasses(esf_ctx* ctx, bool defer) {
  const Dims& d = ctx->d;
  if (ctx->solids.size() > (size_t)kMaxSolids) {
    ctx->err = "too many solids per context";
    return ESF_ERR_INVALID;
  }
  esf_status st;
  bool dofs = false;  // DOF numbering to queue
  // the gathered fields depend on the solids' masses and quarter fractions
  // only: when neither moved since the last gather (the usual step: they
  // are computed once per scene), the fields and the DOF numbering stand
  if (ctx->gatheredGen != ctx->massGen || ctx->gatheredSolids != (int)ctx->solids.size()) {
    SolidPtrs sp{};
    sp.n = (int)ctx->solids.size();
    for (int b = 0; b < sp.n; b++) {
      const DevSolid& so = ctx->solids[b];
      sp.mass[b] = so.mass;
      sp.qfrac[b] = so.qfrac;
      // (FLIP scenes' particles set masses of their own: no skipping)
      const bool all = so.dirtyAll || ctx->p.flip;
      for (int k = 0; k < 3; k++) {
        sp.cells[b][k] = all ? 0 : so.dirty[k];
        sp.cells[b][3 + k] = all ? (1 << 30) : so.dirty[3 + k];
      }
    }
    // the gathered fields are nonzero only at the nodes of the solids' last
    // mass passes (their dirty cells and one node beyond); recompute those
    // and the region the last gather left nonzero
    const int lim[3] = {d.X, d.Y, d.Z};
    const int full[6] = {0, 0, 0, d.X, d.Y, d.Z};
    int now[6] = {0, 0, 0, 0, 0, 0};
    for (auto& so : ctx->solids) {
      int nb[6];
      for (int k = 0; k < 3; k++) {
        nb[k] = so.dirty[k];
        nb[3 + k] = so.dirty[3 + k] + 1;
      }
      const BoxRange u = spanBoxes(now, so.dirtyAll ? full : nb, lim);
      for (int k = 0; k < 3; k++) {
        now[k] = u.lo[k];
        now[3 + k] = u.hi[k];
      }
    }
    const BoxRange gb = spanBoxes(ctx->gatherBox, now, lim);
    for (int k = 0; k < 6; k++) ctx->gatherBox[k] = now[k];
    if (gb.count() > 0)
      launch<ESF_K(kGatherMass)>(ctx, dim3(cdiv(gb.count(), 256)), dim3(256), d, sp, gb,
                                 ctx->solidMass, ctx->nodeFrac);
    // number the solid DOF nodes.  Outside gb the last gather left no solid
    // node (sidx = -1) and nothing changed, so the scan walks gb only
    if ((st = growSolidDofs(ctx, 0))) return st;
    for (int k = 0; k < 3; k++) {
      ctx->dofScan[k] = gb.lo[k];
      ctx->dofScan[3 + k] = gb.hi[k];
    }
    dofs = true;
    ctx->gatheredGen = ctx->massGen;
    ctx->gatheredSolids = (int)ctx->solids.size();
  }
  // the DOF numbering and the first solid's valid cells (the usual step
  // needs both) share one set of launches
  int b0 = 0;
  if (dofs && !ctx->solids.empty() && ctx->solids[0].listDirty) {
    if ((st = queueValidList(ctx, 0, false))) return st;
    queueDofsAndValid(ctx);
    b0 = 1;
  } else if (dofs) {
    queueSolidDofs(ctx);
  }
  for (int b = b0; b < (int)ctx->solids.size(); b++)
    if (ctx->solids[b].listDirty && (st = queueValidList(ctx, b)))
      return st;
  if (defer) return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
  ESF_HIP(streamSync(ctx));
  return gatherFinish(ctx);
}

// ---------------------------------------------------------------------------
// computeSolidMass(usesdf = false)  DEFOR_SOLID_3D.cpp:576-669
// ---------------------------------------------------------------------------
template <class GridT>
__device__ __forceinline__ void kZero(const GridT& wg, long n, double* a) {
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)wg.blocks() * blockDim.x)
    a[i] = 0.0;
}
ESF_BOUNDS(kZero, 1024);
template <class GridT>
__device__ __forceinline__ void kZero8(const GridT& wg, long n, int8_t* a) {
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)wg.blocks() * blockDim.x)
    a[i] = 0;
}
ESF_BOUNDS(kZero8, 1024);
template <class GridT>
__device__ __forceinline__ void kFillInt(const GridT& wg, long n, int* a, int v) {
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)wg.blocks() * blockDim.x)
    a[i] = v;
}
ESF_BOUNDS(kFillInt, 1024);

__device__ inline double det3(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[3] * (m[1] * m[8] - m[2] * m[7]) +
         m[6] * (m[1] * m[5] - m[2] * m[4]);
}

// computeF's determinant: det F = 1 / det(F^-1) with the |det| <= 1e-4 rule
__device__ inline double detF(const Dims& d, const double* u, long index) {
  const long off[8] = {0, 1, d.X, d.X + 1, d.slab, d.slab + 1, d.slab + d.X, d.slab + d.X + 1};
  const double g = 0.25 * d.dhInv;
  double Finv[9];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      double acc = 0;
      for (int i = 0; i < 8; i++) {
        const int bit = c == 0 ? (i & 1) : (c == 1 ? ((i >> 1) & 1) : (i >> 2));
        const double t = u[3 * (index + off[i]) + r] * ((bit ? 1 : -1) * g);
        acc = (i == 0) ? t : acc + t;
      }
      Finv[3 * r + c] = ((r == c) ? 1.0 : 0.0) - acc;
    }
  double J = det3(Finv);
  if (!(fabs(J) > 1e-4)) J = 1;
  return 1.0 / J;
}

struct MassArgs {
  Box box;
  const double* u;
  const double* X;
  const double* rest;
  FieldGeom rg;
  double sdh, subV, fraction, qcm;
  int msr, half;
  double* detFInv;
  double* qfrac;
  int8_t* valid;
  double* mass;
  const double* phi;  // computeSolidMass(true): cells whose min corner has phi > 2 are skipped
};

__device__ inline bool inTight(const Box& b, int x, int y, int z) {
  return x >= b.tMin[0] && x < b.tMax[0] - 1 && y >= b.tMin[1] && y < b.tMax[1] - 1 &&
         z >= b.tMin[2] && z < b.tMax[2] - 1;
}

// clear q and _valid over the cells of box u (the reference clears the
// whole fields, :578-581; outside u they are already 0), and |det F^-1|:
// computeF's determinant at the cells of the tight box t (:613-619), 0 at
// the others of u
template <class GridT>
__device__ __forceinline__ void kMassClear(const GridT& wg, Dims d, BoxRange u, BoxRange t,
                                           const double* uField, double* qfrac, int8_t* valid,
                                           double* detFInv) {
  const long k = wg.block() * (long)blockDim.x + threadIdx.x;
  if (k >= u.count()) return;
  int x, y, z;
  u.at(k, x, y, z);
  const long index = x + (long)y * d.X + (long)z * d.slab;
  for (int i = 0; i < 8; i++) qfrac[8 * index + i] = 0.0;
  const bool inT = x >= t.lo[0] && x < t.hi[0] && y >= t.lo[1] && y < t.hi[1] && z >= t.lo[2] &&
                   z < t.hi[2];
  detFInv[index] = inT ? fabs(1.0 / detF(d, uField, index)) : 0.0;
  valid[x + (long)y * d.CX + (long)z * d.cslab] = 0;
}
ESF_BOUNDS(kMassClear, 1024);

// one wave per tight-box cell: the cell's msr^3 subsamples spread over the
// lanes (sample s = (p msr + q) msr + r in the reference's loop order, lane
// l taking s = l, l + 64, ...); per octant the count of samples that carry
// mass comes from wave ballots.  The reference adds `fraction` to the
// octant's q once per such sample, so q is that many additions of one
// constant, which depends on the count only: lane 0 performs them, then the
// sum, the 0.5 threshold and the flags.  The cell's phi and the first
// samples' interpolation of X are loaded together, so a cell that phi
// skips costs one round of loads.
template <class GridT>
__device__ __forceinline__ void kMassSamples(const GridT& wg, Dims d, MassArgs a, BoxRange t) {
  const long wave = __builtin_amdgcn_readfirstlane(
      (int)((wg.block() * (long)blockDim.x + threadIdx.x) >> 6));
  const int lane = threadIdx.x & 63;
  if (wave >= t.count()) return;  // uniform per wave
  int x, y, z;
  t.at(wave, x, y, z);
  const long index = x + (long)y * d.X + (long)z * d.slab;
  const long c = x + (long)y * d.CX + (long)z * d.cslab;
  const FieldGeom g = gridGeom(d);
  const double node[3] = {x * d.dh + d.o0, y * d.dh + d.o1, z * d.dh + d.o2};
  const int m2 = a.msr * a.msr, m3 = m2 * a.msr;
  // every sample lies inside node cell (x, y, z) (its offset from the cell's
  // low corner is between sdh / 2 and dh - sdh / 2 per axis), so X's
  // interpolation reads the same 8 corners for all of them: loaded once
  double cv[3][8];
  {
    const long X = d.X, S = d.slab;
    const long off[8] = {0, 1, X, X + 1, S, S + 1, S + X, S + X + 1};
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) cv[c][k] = a.X[3 * (index + off[k]) + c];
  }
  // sample s's position and its rest position X(pos)
  auto restPos = [&](int s, double (&rPos)[3], int& o) {
    const int p = s / m2, qq = (s / a.msr) % a.msr, r = s % a.msr;
    const double pos[3] = {node[0] + (p + 0.5) * a.sdh, node[1] + (qq + 0.5) * a.sdh,
                           node[2] + (r + 0.5) * a.sdh};
    interpVCell(cv, g, pos, rPos);
    o = p < a.half ? 0 : 1;
    o += qq < a.half ? 0 : 2;
    o += r < a.half ? 0 : 4;
  };
  const int s0 = lane < m3 ? lane : m3 - 1;
  double rPos0[3];
  int o0;
  const double ph = a.phi ? a.phi[index] : 0.0;
  restPos(s0, rPos0, o0);
  // both rounds of loads above are issued before the branch on phi
  asm volatile("" ::: "memory");
  // usesdf (DEFOR_SOLID_3D.cpp:620): no samples, q and _valid stay cleared
  if (ph > 2) return;
  int count[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int base = 0; base < m3; base += 64) {
    const int s = base + lane;
    bool hit = false;
    int o = 0;
    if (s < m3) {
      double rPos[3];
      if (base == 0) {
        for (int k = 0; k < 3; k++) rPos[k] = rPos0[k];
        o = o0;
      } else {
        restPos(s, rPos, o);
      }
      const double dd = interpS(a.rest, a.rg, rPos) * a.subV;
      hit = !(dd < 1e-6);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) count[k] += __popcll(__ballot(hit && o == k));
  }
  if (lane != 0) return;
  double q[8];
  for (int k = 0; k < 8; k++) {
    double acc = 0;
    for (int j = 0; j < count[k]; j++) acc += a.fraction;
    q[k] = acc;
  }
  double sum = 0;
  for (int i = 0; i < 8; i++) sum += q[i];
  int8_t v = 0;
  double qv = 0;
  if (sum > 0.5) {
    v = 1;
    qv = 1;
  } else if (sum > 0) {
    v = 2;
  }
  for (int i = 0; i < 8; i++) a.qfrac[8 * index + i] = qv;
  a.valid[c] = v;
}
ESF_BOUNDS(kMassSamples, 256);

// mass per node of box u: corners i = 0..7 in the reference's outer-loop
// order
template <class GridT>
__device__ __forceinline__ void kMassGather(const GridT& wg, Dims d, MassArgs a, BoxRange u) {
  const long t = wg.block() * (long)blockDim.x + threadIdx.x;
  if (t >= u.count()) return;
  int x, y, z;
  u.at(t, x, y, z);
  const long n = x + (long)y * d.X + (long)z * d.slab;
  double m = 0;
  for (int i = 0; i < 8; i++) {
    const int cx = x - (i & 1), cy = y - ((i >> 1) & 1), cz = z - (i >> 2);
    if (cx < 0 || cy < 0 || cz < 0 || cx >= d.CX || cy >= d.CY || cz >= d.CZ) continue;
    if (!inTight(a.box, cx, cy, cz)) continue;
    const long cell = cx + (long)cy * d.CX + (long)cz * d.cslab;
    if (!a.valid[cell]) continue;
    const long index = cx + (long)cy * d.X + (long)cz * d.slab;
    m += a.qfrac[8 * index + i] * a.detFInv[index] * a.qcm;
  }
  a.mass[n] = m;
}
ESF_BOUNDS(kMassGather, 1024);

esf_status computeSolidMass(esf_ctx* ctx, int solid, bool usesdf) {
  const Dims& d = ctx->d;
  DevSolid& so = ctx->solids[solid];
  // T: the tight box's cells; U: T and the cells the last pass (or an
  // upload) may have left nonzero
  const int climit[3] = {d.CX, d.CY, d.CZ};
  const int fullCells[6] = {0, 0, 0, d.CX, d.CY, d.CZ};
  const int tc[6] = {so.box.tMin[0], so.box.tMin[1], so.box.tMin[2], so.box.tMax[0] - 1,
                     so.box.tMax[1] - 1, so.box.tMax[2] - 1};
  const int none[6] = {0, 0, 0, 0, 0, 0};
  const BoxRange T = spanBoxes(tc, none, climit);
  const BoxRange U = spanBoxes(so.dirtyAll ? fullCells : so.dirty, tc, climit);
  // the samples read _X at the corners of T's cells
  if (so.xLazy && T.count() > 0) {
    bool inside = true;
    for (int k = 0; k < 3; k++)
      inside = inside && so.xBox[k] <= T.lo[k] && T.hi[k] + 1 <= so.xBox[3 + k];
    if (!inside) {
      esf_status st = materialiseX(ctx, so);
      if (st) return st;
    }
  }
  if (U.count() > 0)
    launch<ESF_K(kMassClear)>(ctx, dim3(cdiv(U.count(), 256)), dim3(256), d, U, T,
                              (const double*)so.u, so.qfrac, so.valid, so.detFInv);
  MassArgs a{};
  a.box = so.box;
  a.u = so.u;
  a.X = so.X;
  a.rest = so.rest;
  a.rg = fieldGeom(so.rX, so.rY, so.rZ, so.rdh, so.ro[0], so.ro[1], so.ro[2]);
  a.sdh = ctx->p.mass_sample_dh;
  a.subV = a.sdh * a.sdh * a.sdh;
  a.msr = (int)(d.dh / a.sdh);
  a.half = a.msr >> 1;
  a.fraction = 1.0 / (a.msr * a.msr * a.msr);
  a.qcm = so.qcm;
  a.detFInv = so.detFInv;
  a.qfrac = so.qfrac;
  a.valid = so.valid;
  a.mass = so.mass;
  a.phi = usesdf ? so.phi : nullptr;
  if (T.count() > 0)
    launch<ESF_K(kMassSamples)>(ctx, dim3(cdiv(64 * T.count(), 256)), dim3(256), d, a, T);
  // the masses of U's nodes (a cell's corners reach one node past it); the
  // other nodes keep 0 (or a mass that was uploaded: then U is the grid)
  const int nlimit[3] = {d.X, d.Y, d.Z};
  const int un[6] = {U.lo[0], U.lo[1], U.lo[2], U.hi[0] + 1, U.hi[1] + 1, U.hi[2] + 1};
  const BoxRange UN = spanBoxes(so.dirtyAll ? fullCells : un, none, nlimit);
  const BoxRange M = so.dirtyAll ? BoxRange{{0, 0, 0}, {d.X, d.Y, d.Z}} : UN;
  if (M.count() > 0)
    launch<ESF_K(kMassGather)>(ctx, dim3(cdiv(M.count(), 256)), dim3(256), d, a, M);
  for (int k = 0; k < 3; k++) {
    so.dirty[k] = T.lo[k];
    so.dirty[3 + k] = T.hi[k];
  }
  so.dirtyAll = false;
  so.listDirty = true;
  ctx->massGen++;
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// bounding box of nodes with mass > 0 (computeSDF :227-241) over the node
// range r (where the mass can be nonzero): per-block min/max of x, y, z
template <class GridT>
__device__ __forceinline__ void kMassBox(const GridT& wg, Dims d, BoxRange r, const double* mass,
                                         int* part) {
  __shared__ int s[6][256];
  int mn[3] = {1 << 30, 1 << 30, 1 << 30}, mx[3] = {-1, -1, -1};
  const int nx = r.hi[0] - r.lo[0], ny = r.hi[1] - r.lo[1];
  const int cnt = (int)r.count();
  for (int t = wg.block() * blockDim.x + threadIdx.x; t < cnt; t += wg.blocks() * blockDim.x) {
    const int q = t / nx;
    const int c[3] = {r.lo[0] + (t - q * nx), r.lo[1] + q % ny, r.lo[2] + q / ny};
    if (mass[c[0] + (long)c[1] * d.X + (long)c[2] * d.slab] > 0)
      for (int k = 0; k < 3; k++) {
        mn[k] = min(mn[k], c[k]);
        mx[k] = max(mx[k], c[k]);
      }
  }
  for (int k = 0; k < 3; k++) {
    s[k][threadIdx.x] = mn[k];
    s[3 + k][threadIdx.x] = mx[k];
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = min(s[k][threadIdx.x], s[k][threadIdx.x + st]);
        s[3 + k][threadIdx.x] = max(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k++) part[6 * wg.block() + k] = s[k][0];
}
ESF_BOUNDS(kMassBox, 1024);

// the boxes of computeSDF from the node bounds mn, mx: the tight box widened
// by the tight border, the extended box by the extrapolation cells, both
// clipped to the grid
__host__ __device__ inline Box boxesFrom(const int mn[3], const int mx[3], const int res[3],
                                         int tb, int ex) {
  Box b;
  for (int k = 0; k < 3; k++) {
    b.tMin[k] = mn[k] - tb > 0 ? mn[k] - tb : 0;
    b.tMax[k] = mx[k] + tb < res[k] ? mx[k] + tb : res[k];
    b.bMin[k] = mn[k] - ex > 0 ? mn[k] - ex : 0;
    b.bMax[k] = mx[k] + ex < res[k] ? mx[k] + ex : res[k];
  }
  return b;
}

// the per-block boxes of kMassBox into one (blocks with no massive node,
// max -1, left out; none at all: min = resolution, max = 0), as the host
// would: the node bounds into `out` (6 ints, pinned host memory) and the
// solid's boxes into `box` (device memory, read by the later stages of the
// step without a round trip to the host)
template <class GridT>
__device__ __forceinline__ void kMassBoxFinish(const GridT& wg, const int* part, int nb, int X,
                                               int Y, int Z, int tb, int ex, Box* box, int* out) {
  __shared__ int s[6][256];
  int mn[3] = {X, Y, Z}, mx[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int k = 0; k < 3; k++) {
      if (part[6 * b + 3 + k] < 0) continue;
      mn[k] = min(mn[k], part[6 * b + k]);
      mx[k] = max(mx[k], part[6 * b + 3 + k]);
    }
  for (int k = 0; k < 3; k++) {
    s[k][threadIdx.x] = mn[k];
    s[3 + k][threadIdx.x] = mx[k];
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = min(s[k][threadIdx.x], s[k][threadIdx.x + st]);
        s[3 + k][threadIdx.x] = max(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
  if (threadIdx.x == 0) {
    const int bmn[3] = {s[0][0], s[1][0], s[2][0]}, bmx[3] = {s[3][0], s[4][0], s[5][0]};
    const int res[3] = {X, Y, Z};
    *box = boxesFrom(bmn, bmx, res, tb, ex);
  }
}
ESF_BOUNDS(kMassBoxFinish, 256);

// computeSDF's boxes (DEFOR_SOLID_3D.cpp:227-241): the bounding box of the
// nodes with mass, widened by the tight border and the extrapolation cells.
// Enqueue scans the nodes where the mass can be nonzero (the corners of the
// last mass pass's cells, or the whole grid after an upload or in FLIP
// scenes, whose particles set masses), writes the solid's boxes to device
// memory and the node bounds into pinned hostFlags[kHostBoxes + 6 b ...],
// and keeps in so.boxBound boxes that contain whatever it finds; finish
// reads the bounds after a synchronisation.
esf_status updateBoxesEnqueue(esf_ctx* ctx, int solid) {
  DevSolid& so = ctx->solids[solid];
  const Dims& d = ctx->d;
  const int nb = 256;
  const int res[3] = {d.X, d.Y, d.Z};
  const int full[6] = {0, 0, 0, d.X, d.Y, d.Z};
  const int none[6] = {0, 0, 0, 0, 0, 0};
  const int corners[6] = {so.dirty[0], so.dirty[1], so.dirty[2],
                          so.dirty[3] + 1, so.dirty[4] + 1, so.dirty[5] + 1};
  const BoxRange r = spanBoxes(so.dirtyAll || ctx->p.flip ? full : corners, none, res);
  double* part = ctx->partials + 6L * nb * solid;  // one partials region per solid
  launch<ESF_K(kMassBox)>(ctx, dim3(nb), dim3(256), d, r, so.mass, (int*)part);
  launch<ESF_K(kMassBoxFinish)>(ctx, dim3(1), dim3(256), (const int*)part, nb, d.X, d.Y, d.Z,
                                ctx->p.tight_border, ctx->p.extrapolation_cells,
                                ctx->dBoxes + solid, ctx->hostFlags + kHostBoxes + 6 * solid);
  const int mx[3] = {r.hi[0] - 1, r.hi[1] - 1, r.hi[2] - 1};
  so.boxBound = r.count() > 0 ? boxesFrom(r.lo, mx, res, ctx->p.tight_border,
                                          ctx->p.extrapolation_cells)
                              : Box{};
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// markList: the valid-cell list follows the new tight box (a caller that
// built it from the device copy already passes false)
void updateBoxesFinish(esf_ctx* ctx, int solid, bool markList) {
  DevSolid& so = ctx->solids[solid];
  const Dims& d = ctx->d;
  const int* h = ctx->hostFlags + kHostBoxes + 6 * solid;
  const int mn[3] = {h[0], h[1], h[2]}, mx[3] = {h[3], h[4], h[5]};
  const int res[3] = {d.X, d.Y, d.Z};
  so.box = boxesFrom(mn, mx, res, ctx->p.tight_border, ctx->p.extrapolation_cells);
  if (markList) so.listDirty = true;
}

// the host's copy of a solid's boxes into device memory (after a change
// that did not come from kMassBoxFinish)
esf_status uploadBox(esf_ctx* ctx, int solid) {
  ESF_HIP(hipMemcpyAsync(ctx->dBoxes + solid, &ctx->solids[solid].box, sizeof(Box),
                         hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status updateBoxes(esf_ctx* ctx, int solid) {
  esf_status st = updateBoxesEnqueue(ctx, solid);
  if (st) return st;
  ESF_HIP(streamSync(ctx));
  updateBoxesFinish(ctx, solid, true);
  return ESF_OK;
}

// ---------------------------------------------------------------------------
// phi of computeSDF (DEFOR_SOLID_3D.cpp:216-267) for the step.  The
// reference runs two rounds of eight ordered sweeps of solveDistance over
// the extended box; the one reader of phi inside the step is
// computeSolidMass(true) (:620), which asks only whether phi > 2 at a tight
// box node.  Around the massive nodes the sweeps reach the fixed point of
// the upwind update phi = min(phi, solveDistance(min of the two x
// neighbours, min of the two y, min of the two z)) with the neighbours
// outside the box absent, and Jacobi relaxation of the same update, from
// the same start, converges to that fixed point from above, settling one
// ring of neighbours per relaxation.  Each ring adds at least 1/sqrt(3) to
// phi (three equal upwind neighbours), so a value <= 2 lies within 5 rings
// of a massive node; kPhiRelaxations = 8 leaves a margin for the slower
// settling of two-neighbour updates.  On ball scenes, deformed by steps,
// the set phi > 2 over the tight box equals the ordered sweeps' from 5
// relaxations on and the values below 3 agree to 1e-4 at 8 (host test
// against the oracle's sweeps).  Further out phi is an upper bound of the
// swept value.
// ---------------------------------------------------------------------------
constexpr int kPhiRelaxations = 8;

// phi = -0.5 on nodes with mass inside the box, large everywhere else (the
// box read from device memory), over the nodes of r: outside r phi is
// large already
template <class GridT>
__device__ __forceinline__ void kPhiInit(const GridT& wg, Dims d, const Box* box, BoxRange r,
                                         const double* __restrict__ mass, double large, double* phi,
                                         double* phi2) {
  const long t = wg.block() * (long)blockDim.x + threadIdx.x;
  if (t >= r.count()) return;
  const Box bx = *box;
  int x, y, z;
  r.at(t, x, y, z);
  const long n = x + (long)y * d.X + (long)z * d.slab;
  const bool in = x >= bx.bMin[0] && x < bx.bMax[0] && y >= bx.bMin[1] && y < bx.bMax[1] &&
                  z >= bx.bMin[2] && z < bx.bMax[2];
  const double v = (in && mass[n] > 0) ? -0.5 : large;
  phi[n] = v;
  phi2[n] = v;
}
ESF_BOUNDS(kPhiInit, 1024);

// one relaxation over the box: nodes without mass take the upwind update
// of their neighbours' previous values.  The launch covers the extended
// box of `r` (a box that contains the solid's); the box itself comes from
// device memory.
template <class GridT>
__device__ __forceinline__ void kPhiRelax(const GridT& wg, Dims d, const Box* box, BoxRange r,
                                          const double* __restrict__ mass,
                                          const double* __restrict__ in, double* out, double large,
                                          int margin) {
  const int nx = r.hi[0] - r.lo[0], ny = r.hi[1] - r.lo[1];
  const int t = wg.block() * blockDim.x + threadIdx.x;
  if (t >= (int)r.count()) return;
  const int q = t / nx;
  const int x = r.lo[0] + (t - q * nx), y = r.lo[1] + q % ny, z = r.lo[2] + q / ny;
  const Box bx = *box;
  auto inBox = [&](int i, int j, int k) {
    return i >= bx.bMin[0] && i < bx.bMax[0] && j >= bx.bMin[1] && j < bx.bMax[1] &&
           k >= bx.bMin[2] && k < bx.bMax[2];
  };
  if (!inBox(x, y, z)) return;
  // margin >= 0: only the nodes within `margin` of the tight box (a node's
  // value after the remaining relaxations reads no node further out)
  if (margin >= 0 && (x < bx.tMin[0] - margin || x >= bx.tMax[0] + margin ||
                      y < bx.tMin[1] - margin || y >= bx.tMax[1] + margin ||
                      z < bx.tMin[2] - margin || z >= bx.tMax[2] + margin))
    return;
  const long n = x + (long)y * d.X + (long)z * d.slab;
  if (mass[n] != 0) return;
  auto at = [&](int i, int j, int k) {
    return inBox(i, j, k) ? in[i + (long)j * d.X + (long)k * d.slab] : large;
  };
  const double p = fmin(at(x - 1, y, z), at(x + 1, y, z));
  const double q2 = fmin(at(x, y - 1, z), at(x, y + 1, z));
  const double r2 = fmin(at(x, y, z - 1), at(x, y, z + 1));
  out[n] = solveDistance(p, q2, r2, in[n]);
}
ESF_BOUNDS(kPhiRelax, 1024);

// `launchBox`: the solid's boxes or boxes that contain them (the kernels
// read the solid's own from device memory)
// tightOnly (the step's pass, outside FLIP scenes): relaxation k runs only
// within kPhiRelaxations - 1 - k nodes of the tight box, so phi is the
// whole-box relaxation's to the bit inside the tight box, where the next
// mass pass reads it, and phiPartial is set
esf_status solidPhi(esf_ctx* ctx, int solid, const Box& launchBox, bool tightOnly) {
  const Dims& d = ctx->d;
  DevSolid& so = ctx->solids[solid];
  const double large = (double)(d.X + d.Y + d.Z + 2);
  const Box* dbox = ctx->dBoxes + solid;
  // phi differs from `large` only inside the extended box: (re)initialise
  // that box and the one the last initialisation left
  {
    const int lim[3] = {d.X, d.Y, d.Z};
    const int ext[6] = {launchBox.bMin[0], launchBox.bMin[1], launchBox.bMin[2],
                        launchBox.bMax[0], launchBox.bMax[1], launchBox.bMax[2]};
    const BoxRange r = spanBoxes(so.phiSpan, ext, lim);
    if (r.count() > 0)
      launch<ESF_K(kPhiInit)>(ctx, dim3(cdiv(r.count(), 256)), dim3(256), d, dbox, r,
                              (const double*)so.mass, large, so.phi, so.phi2);
    for (int k = 0; k < 6; k++) so.phiSpan[k] = ext[k];
  }
  if (ctx->p.flip) {
    // the FLIP round trip extends the velocity over the whole box by phi:
    // the reference's two rounds of ordered sweeps, bitwise
    so.phiPartial = false;
    const Box& bx = so.box;
    const int box[6] = {bx.bMin[0], bx.bMin[1], bx.bMin[2], bx.bMax[0], bx.bMax[1], bx.bMax[2]};
    return sweepPhiRounds(ctx, box, so.mass, so.phi, 2);
  }
  static_assert(kPhiRelaxations % 2 == 0, "the relaxation ends in phi");
  for (int k = 0; k < kPhiRelaxations; k++) {
    const int margin = tightOnly ? kPhiRelaxations - 1 - k : -1;
    BoxRange r;
    for (int a = 0; a < 3; a++) {
      r.lo[a] = launchBox.bMin[a];
      r.hi[a] = std::max(launchBox.bMax[a], launchBox.bMin[a]);
      if (margin >= 0) {
        r.lo[a] = std::max(r.lo[a], launchBox.tMin[a] - margin);
        r.hi[a] = std::max(r.lo[a], std::min(r.hi[a], launchBox.tMax[a] + margin));
      }
    }
    const long nb = r.count();
    if (nb <= 0) continue;
    const double* src = (k & 1) ? so.phi2 : so.phi;
    double* dst = (k & 1) ? so.phi : so.phi2;
    launch<ESF_K(kPhiRelax)>(ctx, dim3(cdiv(nb, 256)), dim3(256), d, dbox, r,
                             (const double*)so.mass, src, dst, large, margin);
  }
  so.phiPartial = tightOnly;
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status phiWhole(esf_ctx* ctx, int solid) {
  DevSolid& so = ctx->solids[solid];
  if (!so.phiPartial) return ESF_OK;
  return solidPhi(ctx, solid, so.box, false);
}

// ---------------------------------------------------------------------------
// CFL  COUPLED_INTEGRATOR_3D.cpp:88-102 (solid DOF nodes are exactly the
// interior nodes with solid mass).  The reference takes dt = min over the
// nodes of alpha / vmax, starting from 1e9.  Division by a positive vmax
// and its rounding are monotone, so that minimum is alpha / (max vmax),
// rounded once: the kernel reduces vmax, and the host divides once alpha
// is known (the broad phase that picks alpha runs beside it).  Both skip
// NaN as the reference's std::min does.
// ---------------------------------------------------------------------------
template <class GridT>
__device__ __forceinline__ void kCfl(const GridT& wg, int nSolid, const int* solidNodes,
                                     const double* vel, double dhInv, double* part) {
  __shared__ double s[256];
  double vm = 0;
  for (int i = wg.block() * blockDim.x + threadIdx.x; i < nSolid; i += wg.blocks() * blockDim.x) {
    const long n = solidNodes[i];
    const double m = fmax(fmax(fabs(vel[3 * n]), fabs(vel[3 * n + 1])), fabs(vel[3 * n + 2]));
    vm = fmax(vm, m * dhInv);
  }
  s[threadIdx.x] = vm;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[wg.block()] = s[0];
}
ESF_BOUNDS(kCfl, 1024);

// the CFL kernel and the copy of its block maxima to pinned memory, queued;
// cflWait waits for that copy only, so work queued in between (the stress
// stage) runs while the host waits.  `mark` also marks the stream when
// there is no solid node (a caller that queued other copies before it).
static int cflBlocks(const esf_ctx* ctx) { return std::min(cdiv(ctx->nSolid, 256), kCflBlocks); }

esf_status cflEnqueue(esf_ctx* ctx, bool mark) {
  ctx->cflMark = nullptr;
  if (ctx->nSolid > 0) {
    const int nb = cflBlocks(ctx);
    launch<ESF_K(kCfl)>(ctx, dim3(nb), dim3(256), ctx->nSolid, ctx->solidNodes, ctx->vel,
                        ctx->d.dhInv, ctx->partials);
    toHostAsync(ctx, ctx->hostCfl, ctx->partials, nb * sizeof(double));
  }
  if (ctx->nSolid > 0 || mark) ctx->cflMark = markEvent(ctx, ctx->evCfl);
  return ESF_OK;
}

esf_status cflWait(esf_ctx* ctx) {
  if (ctx->cflMark) ESF_HIP(eventSync(ctx, ctx->cflMark));
  return ESF_OK;
}

// after cflWait
esf_status cflFinish(esf_ctx* ctx, double alpha, double* dtOut) {
  double dt = 1e9;
  if (ctx->nSolid > 0) {
    const int nb = cflBlocks(ctx);
    double vm = 0;
    for (int i = 0; i < nb; i++) vm = std::fmax(vm, ctx->hostCfl[i]);
    dt = std::min(dt, alpha / vm);
  }
  dt = std::min(dt, ctx->defaultDt);
  ctx->dt = dt;
  *dtOut = dt;
  return ESF_OK;
}

esf_status cfl(esf_ctx* ctx, double alpha, double* dtOut) {
  esf_status st = cflEnqueue(ctx, false);
  if (!st) st = cflWait(ctx);
  return st ? st : cflFinish(ctx, alpha, dtOut);
}

// ---------------------------------------------------------------------------
// vorticity  COUPLED_INTEGRATOR_3D.cpp:450-497; _workspace cleared first
// (:443) and written with |omega| where the node is not fully solid
// ---------------------------------------------------------------------------
template <class GridT>
__device__ __forceinline__ void kVorticity(const GridT& wg, Dims d, const double* __restrict__ v,
                                           const double* __restrict__ nf, double* vort, double* ws,
                                           int enabled) {
  const long index = wg.block() * (long)blockDim.x + threadIdx.x;
  if (index >= d.nodes) return;
  int x, y, z;
  nodeXYZ(d, index, x, y, z);
  double out = 0.0;
  if (enabled && isInterior(d, x, y, z) && !(nf[index] >= 1)) {
    const long X = d.X, S = d.slab;
    const double gridSize = 0.5 / d.dh, dhInv = d.dhInv;
    const long up = nf[index + X] >= 1 || y == d.Y - 2 ? index : index + X;
    const long down = nf[index - X] >= 1 || y == 1 ? index : index - X;
    const double dy = (up == index || down == index) ? dhInv : gridSize;
    const long out_ = nf[index + S] >= 1 || z == d.Z - 2 ? index : index + S;
    const long in = nf[index - S] >= 1 || z == 1 ? index : index - S;
    const double dz = (out_ == index || in == index) ? dhInv : gridSize;
    const long right = nf[index + 1] >= 1 || x == d.X - 2 ? index : index + 1;
    const long left = nf[index - 1] >= 1 || x == 1 ? index : index - 1;
    const double dx = (right == index || left == index) ? dhInv : gridSize;
    const double w0 = ((v[3 * up + 2] - v[3 * down + 2]) + (-v[3 * out_ + 1] + v[3 * in + 1])) * dz;
    const double w1 = ((v[3 * out_] - v[3 * in]) + (-v[3 * right + 2] + v[3 * left + 2])) * dx;
    const double w2 = ((v[3 * right + 1] - v[3 * left + 1]) + (-v[3 * up] + v[3 * down])) * dy;
    vort[3 * index] = w0;
    vort[3 * index + 1] = w1;
    vort[3 * index + 2] = w2;
    out = sqrt(w0 * w0 + (w1 * w1 + w2 * w2));
  }
  ws[index] = out;
}
ESF_BOUNDS(kVorticity, 1024);

// body forces and mass vectors  COUPLED_INTEGRATOR_3D.cpp:505-606, with the
// fluid part of pcrImplicitSolve folded in: v0 = (dt f + m_f v) / m
// (:669-670, :686-688).  Solid nodes add to the solid force / momentum /
// damped mass (after the stress gather): here unless skipSolid is set, in
// which case kSolidBodyForces does it (the first solid solve needs only
// that part, so the rest can run beside it).
struct BodyArgs {
  const double* vel;
  const double* nf;
  const double* solidMass;
  const double* intensity;
  const double* ws;
  const double* vort;
  const int* sidx;
  double* v0;
  double* minv;
  double* sForce;
  double* sMv;
  double* sDamped;
  double fluidDensity, gravity, buoyancy, dampCoeff, vortEps, dt;
  double* energyPart;
  int skipSolid;
};

// a solid node's share of the body-force pass (the solidIdx >= 0 branch of
// kBodyForces, the same expressions in the same order)
__device__ inline void solidBodyForces(const Dims& d, const BodyArgs& a, long index,
                                       int solidIdx) {
  const double dh = d.dh;
  const double fluidMass = fmax(1.0 - a.nf[index], 0.0) * a.fluidDensity * dh * dh * dh;
  const double mass = fluidMass + a.solidMass[index];
  const double* v = &a.vel[3 * index];
  const long k = 3L * solidIdx;
  const double dm = a.solidMass[index] * a.dampCoeff + fluidMass;
  a.sDamped[k] = dm;
  a.sDamped[k + 1] = dm;
  a.sDamped[k + 2] = dm;
  for (int c = 0; c < 3; c++) a.sMv[k + c] += fluidMass * v[c];
  const double mg = mass * a.gravity;
  a.sForce[k] += mg * 0;
  a.sForce[k + 1] += mg * -1;
  a.sForce[k + 2] += mg * 0;
  a.sForce[k + 1] += a.buoyancy * fluidMass * a.intensity[index];
}

template <class GridT>
__device__ __forceinline__ void kBodyForces(const GridT& wg, Dims d, BodyArgs a) {
  __shared__ double red[16];
  const long index = wg.block() * (long)blockDim.x + threadIdx.x;
  double en = 0;
  if (index < d.nodes) {
    int x, y, z;
    nodeXYZ(d, index, x, y, z);
    if (isInterior(d, x, y, z)) {
      const double* nf = a.nf;
      const double dh = d.dh;
      const double fluidMass = fmax(1.0 - nf[index], 0.0) * a.fluidDensity * dh * dh * dh;
      const double mass = fluidMass + a.solidMass[index];
      const double h = dh * y + d.o1;
      en += mass * a.gravity * h;
      const int solidIdx = a.sidx[index];
      const double* v = &a.vel[3 * index];
      if (solidIdx < 0) {
        const double mg = mass * a.gravity;
        double f0 = 0.0 + mg * 0, f1 = 0.0 + mg * -1, f2 = 0.0 + mg * 0;
        f1 += a.buoyancy * fluidMass * a.intensity[index];
        const double mv0 = fluidMass * v[0], mv1 = fluidMass * v[1], mv2 = fluidMass * v[2];
        if (a.vortEps > 0) {
          const long X = d.X, S = d.slab;
          if (nf[index + X] == 0 && nf[index - X] == 0 && nf[index + 1] == 0 &&
              nf[index - 1] == 0 && nf[index + S] == 0 && nf[index - S] == 0) {
            const double gridSize = 0.5 / dh, dhInv = d.dhInv;
            const long up = y == d.Y - 2 ? index : index + X;
            const long down = nf[index - X] >= 1 || y == 1 ? index : index - X;
            const double dy = (up == index || down == index) ? dhInv : gridSize;
            const long out = nf[index + S] >= 1 || z == d.Z - 2 ? index : index + S;
            const long in = nf[index - S] >= 1 || z == 1 ? index : index - S;
            const double dz = (out == index || in == index) ? dhInv : gridSize;
            const long right = nf[index + 1] >= 1 || x == d.X - 2 ? index : index + 1;
            const long left = nf[index - 1] >= 1 || x == 1 ? index : index - 1;
            const double dx = (right == index || left == index) ? dhInv : gridSize;
            double N0 = (a.ws[right] - a.ws[left]) * dx;
            double N1 = (a.ws[up] - a.ws[down]) * dy;
            double N2 = (a.ws[out] - a.ws[in]) * dz;
            double mag = sqrt(N0 * N0 + N1 * N1 + N2 * N2);
            if (mag > 0.0) {
              mag = 1.0 / mag;
              N0 *= mag;
              N1 *= mag;
              N2 *= mag;
              const double* w = &a.vort[3 * index];
              f0 += (N1 * w[2] - N2 * w[1]) * dh * a.vortEps * fluidMass;
              f1 -= (N0 * w[2] - N2 * w[0]) * dh * a.vortEps * fluidMass;
              f2 += (N0 * w[1] - N1 * w[0]) * dh * a.vortEps * fluidMass;
            }
          }
        }
        const double mi = 1.0 / mass;
        a.v0[3 * index] = (f0 * a.dt + mv0) * mi;
        a.v0[3 * index + 1] = (f1 * a.dt + mv1) * mi;
        a.v0[3 * index + 2] = (f2 * a.dt + mv2) * mi;
      } else if (!a.skipSolid) {
        solidBodyForces(d, a, index, solidIdx);
      }
      a.minv[index] = 1.0 / mass;
      en += 0.5 * fluidMass * (v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]));
    }
  }
  const double t = blockSum(en, red);
  if (threadIdx.x == 0) a.energyPart[wg.block()] = t;
}
ESF_BOUNDS(kBodyForces, 256);

// the solid nodes' part of kBodyForces, one thread per solid DOF node
template <class GridT>
__device__ __forceinline__ void kSolidBodyForces(const GridT& wg, Dims d, BodyArgs a, int nSolid,
                                                 const int* solidNodes) {
  const int s = wg.block() * blockDim.x + threadIdx.x;
  if (s >= nSolid) return;
  const long index = solidNodes[s];
  int x, y, z;
  nodeXYZ(d, index, x, y, z);
  if (!isInterior(d, x, y, z)) return;
  const int solidIdx = a.sidx[index];
  if (solidIdx >= 0) solidBodyForces(d, a, index, solidIdx);
}
ESF_BOUNDS(kSolidBodyForces, 1024);

// solid right-hand side dt F + M v  (COUPLED_INTEGRATOR_3D.cpp:664-665)
template <class GridT>
__device__ __forceinline__ void kSolidRhs(const GridT& wg, long n, const double* f, const double* mv,
                                          double dt, double* rhs) {
  const long i = wg.block() * (long)blockDim.x + threadIdx.x;
  if (i < n) rhs[i] = f[i] * dt + mv[i];
}
ESF_BOUNDS(kSolidRhs, 1024);

// v0 at solid nodes from the first solid solve (:690-691)
template <class GridT>
__device__ __forceinline__ void kScatterSolid(const GridT& wg, int nSolid, const int* solidNodes,
                                              const double* sv, double* v0) {
  const int s = wg.block() * blockDim.x + threadIdx.x;
  if (s >= nSolid) return;
  const long n = solidNodes[s];
  for (int c = 0; c < 3; c++) v0[3 * n + c] = sv[3 * (long)s + c];
}
ESF_BOUNDS(kScatterSolid, 1024);

// g = -J^T lambda per interior node (cells in ascending order, the column
// order of Eigen's transposed product, :719); fluid nodes get
// v1 = v0 + g / m (:745-747, :753-763); solid nodes hand g to the second
// solid solve (:724-729)
template <class GridT>
__device__ __forceinline__ void kCorrect(const GridT& wg, Dims d, const double* __restrict__ lam,
                                         const double* __restrict__ v0,
                                         const double* __restrict__ minv,
                                         const int* __restrict__ sidx, double* vel, double* sB,
                                         double* dvPrev, const Box* boxes, int nBoxes) {
  const long n = wg.block() * (long)blockDim.x + threadIdx.x;
  if (n >= d.nodes) return;
  int i, j, k;
  nodeXYZ(d, n, i, j, k);
  if (!isInterior(d, i, j, k)) return;
  double a0 = 0, a1 = 0, a2 = 0;
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        const double l = lam[(i - 1 + dx) + (long)(j - 1 + dy) * d.CX + (long)(k - 1 + dz) * d.cslab];
        a0 += dx == 0 ? -l : l;
        a1 += dy == 0 ? -l : l;
        a2 += dz == 0 ? -l : l;
      }
  const int s = sidx[n];
  if (s >= 0) {
    sB[3L * s] = a0;
    sB[3L * s + 1] = a1;
    sB[3L * s + 2] = a2;
  } else {
    const double mi = minv[n];
    vel[3 * n] = v0[3 * n] + a0 * mi;
    vel[3 * n + 1] = v0[3 * n + 1] + a1 * mi;
    vel[3 * n + 2] = v0[3 * n + 2] + a2 * mi;
    // dvPrev is read (the next step's second solid solve starts from it) at
    // the next step's solid nodes, which the mass pass of this step finds
    // inside this step's boxes: written there only (everywhere: nBoxes < 0)
    bool keep = dvPrev && nBoxes < 0;
    for (int b = 0; dvPrev && b < nBoxes && !keep; b++) {
      const Box bx = boxes[b];
      keep = i >= min(bx.tMin[0], bx.bMin[0]) && i < max(bx.tMax[0], bx.bMax[0]) &&
             j >= min(bx.tMin[1], bx.bMin[1]) && j < max(bx.tMax[1], bx.bMax[1]) &&
             k >= min(bx.tMin[2], bx.bMin[2]) && k < max(bx.tMax[2], bx.bMax[2]);
    }
    if (keep) {
      dvPrev[3 * n] = a0 * mi;
      dvPrev[3 * n + 1] = a1 * mi;
      dvPrev[3 * n + 2] = a2 * mi;
    }
  }
}
ESF_BOUNDS(kCorrect, 1024);

// v1 at solid nodes = v0 + delta_s  (:748-753)
template <class GridT>
__device__ __forceinline__ void kSolidVelocity(const GridT& wg, int nSolid, const int* solidNodes,
                                               const double* v0, const double* dv, double* vel,
                                               double* dvPrev) {
  const int s = wg.block() * blockDim.x + threadIdx.x;
  if (s >= nSolid) return;
  const long n = solidNodes[s];
  for (int c = 0; c < 3; c++) vel[3 * n + c] = v0[3 * n + c] + dv[3 * (long)s + c];
  if (dvPrev)
    for (int c = 0; c < 3; c++) dvPrev[3 * n + c] = dv[3 * (long)s + c];
}
ESF_BOUNDS(kSolidVelocity, 1024);

// ---------------------------------------------------------------------------
// advection
// ---------------------------------------------------------------------------
// Tiling of the solid-displacement advection: a block owns 32 x 8 nodes of
// each layer and marches kSLZ layers in z; the interpolated field on the
// tile plus a one-node halo sits in LDS for three consecutive layers (loaded
// row by row, each row of the halo box one contiguous run), so the 8-corner
// interpolations read LDS.  A point whose corners leave that window reads
// device memory instead, with the same arithmetic.  (The fluid advection
// kernels stay one thread per node: tiled the same way they measured
// slower, 137 vs 86 us and 66 vs 32 us at 128^3, their departure points
// leaving the window more often.)
constexpr int kSLX = 32, kSLY = 8, kSLZ = 8;
constexpr int kSLHX = kSLX + 2, kSLHY = kSLY + 2;
constexpr int kSLLayer = kSLHX * kSLHY;  // halo nodes per layer
static_assert(kSLX * kSLY == 256, "one node per thread and layer");

__device__ inline void tileOf(int b, int tilesX, int tilesY, int& x0, int& y0, int& z0) {
  const int tx = b % tilesX, ty = (b / tilesX) % tilesY, tz = b / (tilesX * tilesY);
  x0 = tx * kSLX;
  y0 = ty * kSLY;
  z0 = tz * kSLZ;
}

// advectFluidVelocityField  COUPLED_INTEGRATOR_3D.cpp:797-826
template <class GridT>
__device__ __forceinline__ void kAdvectVelocity(const GridT& wg, Dims d,
                                                const double* __restrict__ vel,
                                                const double* __restrict__ solidMass, double dt,
                                                double* velws) {
  const long n = wg.block() * (long)blockDim.x + threadIdx.x;
  if (n >= d.nodes) return;
  int x, y, z;
  nodeXYZ(d, n, x, y, z);
  if (!isInterior(d, x, y, z) || solidMass[n] > 0) return;
  // divisions only: with both the division and the power-of-two path this
  // kernel needs 81 registers (5 waves per SIMD, 73.6 against 70.2 us at
  // 128^3)
  FieldGeom g = gridGeom(d);
  g.pow2 = 0;
  const double node[3] = {x * d.dh + d.o0, y * d.dh + d.o1, z * d.dh + d.o2};
  double mp[3], v[3], np[3];
  for (int c = 0; c < 3; c++) mp[c] = node[c] - 0.5 * vel[3 * n + c] * dt;
  interpV(vel, g, mp, v);
  for (int c = 0; c < 3; c++) np[c] = node[c] - v[c] * dt;
  double mass = interpS(solidMass, g, np);
  double newdt = dt / 2;
  while (mass > 0 && newdt > dt / 16) {
    for (int c = 0; c < 3; c++) np[c] += v[c] * newdt;
    newdt /= 2;
    mass = interpS(solidMass, g, np);
  }
  interpV(vel, g, np, &velws[3 * n]);
}
ESF_BOUNDS(kAdvectVelocity, 256);

// advectField(_intensity)  COUPLED_INTEGRATOR_3D.cpp:766-796 (workspace
// cleared first, so skipped nodes receive 0)
template <class GridT>
__device__ __forceinline__ void kAdvectScalar(const GridT& wg, Dims d, const double* __restrict__ vel,
                                              const double* __restrict__ solidMass,
                                              const double* __restrict__ field, double dt,
                                              double* out) {
  const long n = wg.block() * (long)blockDim.x + threadIdx.x;
  if (n >= d.nodes) return;
  if (solidMass[n] > 0) {
    out[n] = 0.0;
    return;
  }
  int x, y, z;
  nodeXYZ(d, n, x, y, z);
  const FieldGeom g = gridGeom(d);
  const double node[3] = {x * d.dh + d.o0, y * d.dh + d.o1, z * d.dh + d.o2};
  const double v[3] = {vel[3 * n], vel[3 * n + 1], vel[3 * n + 2]};
  double np[3];
  for (int c = 0; c < 3; c++) np[c] = node[c] - v[c] * dt;
  double mass = interpS(solidMass, g, np);
  double newdt = dt / 2;
  while (mass > 0 && newdt > dt / 16) {
    for (int c = 0; c < 3; c++) np[c] += v[c] * newdt;
    newdt /= 2;
    mass = interpS(solidMass, g, np);
  }
  out[n] = interpS(field, g, np);
}
ESF_BOUNDS(kAdvectScalar, 256);

// DEFOR_SOLID_3D::finalizeSolution()  DEFOR_SOLID_3D.cpp:1198-1233.
// u += v dt over all nodes and the advection of u (:1108-1123) fuse: w =
// u + v dt is formed where the interpolation reads it, rounding as the
// separate pass would.
// Tiled as described above; the window holds w.
template <class GridT>
__device__ __forceinline__ void kSolidAdvectU(const GridT& wg, Dims d, const double* __restrict__ u,
                                              const double* __restrict__ sv, double dt,
                                              double* __restrict__ unew, int tilesX, int tilesY) {
  __shared__ double w[3][3 * kSLLayer];
  int x0, y0, z0;
  tileOf(wg.block(), tilesX, tilesY, x0, y0, z0);
  const int zEnd = min(z0 + kSLZ, d.Z);
  // layer z of w into ring slot (z + 3) mod 3 (zeros outside the grid:
  // never read, the interpolation corners stay inside)
  auto loadLayer = [&](int z) {
    double* L = w[(z + 3) % 3];
    for (int e = threadIdx.x; e < 3 * kSLLayer; e += 256) {
      const int hn = e / 3, c = e - 3 * hn;
      const int gx = x0 - 1 + hn % kSLHX, gy = y0 - 1 + hn / kSLHX;
      double val = 0;
      if (z >= 0 && z < d.Z && gx >= 0 && gx < d.X && gy >= 0 && gy < d.Y) {
        const long i = 3 * (gx + (long)gy * d.X + (long)z * d.slab) + c;
        val = u[i] + sv[i] * dt;
      }
      L[e] = val;
    }
  };
  loadLayer(z0 - 1);
  loadLayer(z0);
  const int lx = threadIdx.x % kSLX, ly = threadIdx.x / kSLX;
  const int x = x0 + lx, y = y0 + ly;
  const FieldGeom g = gridGeom(d);
  const long X = d.X, S = d.slab;
  for (int z = z0; z < zEnd; z++) {
    loadLayer(z + 1);
    __syncthreads();
    if (x < d.X && y < d.Y) {
      const long n = x + (long)y * X + (long)z * S;
      const double node[3] = {x * d.dh + d.o0, y * d.dh + d.o1, z * d.dh + d.o2};
      double np[3];
      for (int c = 0; c < 3; c++) np[c] = node[c] - sv[3 * n + c] * dt;
      double fx, fy, fz;
      int bi, bj, bk;
      const long b = locate(g, np, fx, fy, fz, bi, bj, bk);
      const int hx = bi - (x0 - 1), hy = bj - (y0 - 1);
      const bool inWin = hx >= 0 && hx + 1 < kSLHX && hy >= 0 && hy + 1 < kSLHY &&
                         bk >= z - 1 && bk <= z;
      const long cn[8] = {b, b + 1, b + X, b + X + 1, b + S, b + S + 1, b + S + X, b + S + X + 1};
      for (int c = 0; c < 3; c++) {
        double q[8];
        if (inWin) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
            q[k] = w[(bk + oz + 3) % 3][3 * ((hx + ox) + kSLHX * (hy + oy)) + c];
          }
        } else {
#pragma unroll
          for (int k = 0; k < 8; k++) q[k] = u[3 * cn[k] + c] + sv[3 * cn[k] + c] * dt;
        }
        const double z0v = lerp(lerp(q[0], q[1], fx), lerp(q[2], q[3], fx), fy);
        const double z1v = lerp(lerp(q[4], q[5], fx), lerp(q[6], q[7], fx), fy);
        unew[3 * n + c] = lerp(z0v, z1v, fz);
      }
    }
    __syncthreads();
  }
}
ESF_BOUNDS(kSolidAdvectU, 256);

// face extrapolation of u on the extended box (:1205-1220); axis 2 (z
// faces) first, then y, then x, each face pair per thread in the
// reference's statement order
template <class GridT>
__device__ __forceinline__ void kSolidExtrap(const GridT& wg, Dims d, Box bx, int axis, double* u) {
  const int a0 = axis == 2 ? 0 : 0;  // first free axis
  int fa = 0, fb = 1;
  if (axis == 2) {
    fa = 0;
    fb = 1;
  } else if (axis == 1) {
    fa = 0;
    fb = 2;
  } else {
    fa = 1;
    fb = 2;
  }
  (void)a0;
  const int na = bx.bMax[fa] - bx.bMin[fa], nbb = bx.bMax[fb] - bx.bMin[fb];
  const long t = wg.block() * (long)blockDim.x + threadIdx.x;
  if (t >= (long)na * nbb) return;
  int c[3];
  c[fa] = bx.bMin[fa] + (int)(t % na);
  c[fb] = bx.bMin[fb] + (int)(t / na);
  auto at = [&](int v) -> double* {
    c[axis] = v;
    return &u[3 * (c[0] + (long)c[1] * d.X + (long)c[2] * d.slab)];
  };
  const int lo = bx.bMin[axis], hi = bx.bMax[axis];
  for (int k = 0; k < 3; k++) {
    const double a1 = at(lo + 1)[k], a2 = at(lo + 2)[k];
    at(lo)[k] = 2 * a1 - a2;
  }
  for (int k = 0; k < 3; k++) {
    const double a1 = at(hi - 2)[k], a2 = at(hi - 3)[k];
    at(hi - 1)[k] = 2 * a1 - a2;
  }
}
ESF_BOUNDS(kSolidExtrap, 1024);

// X = x - u (:1222-1230) and _vecWorkspace = _velocity (:1232), over the
// 3 N doubles of the AoS fields as flat pairs (16-byte accesses, every
// lane of a load on consecutive bytes); element i is component i mod 3 of
// node i / 3
__device__ inline double recoverXAt(const Dims& d, int x, int y, int z, int c, double u) {
  const double pos = (c == 0 ? x : (c == 1 ? y : z)) * d.dh;
  const double o = c == 0 ? d.o0 : (c == 1 ? d.o1 : d.o2);
  return pos + (o - u);
}
__device__ inline double recoverX(const Dims& d, long i, double u) {
  const int n = (int)i / 3;
  const int c = (int)i - 3 * n;
  int x, y, z;
  nodeXYZ(d, n, x, y, z);
  return recoverXAt(d, x, y, z, c, u);
}

// The face extrapolation of the extended box (:1205-1220: z faces, then
// y, then x, each face value 2 a1 - a2 of the two nodes inward as the
// passes before left them), evaluated per node from the field before it.
// For a box at least kFaceFuseMin nodes wide on every axis the inward
// nodes a face reads are faces of no axis (lo + 1, lo + 2, hi - 2, hi - 3),
// so the chain reads only nodes the extrapolation leaves alone, and a
// kernel may write each face node in place while others read.
struct FaceBox {
  int on;  // 0: no extrapolation here (done by kSolidExtrap)
  int lo[3], hi[3];
};
constexpr int kFaceFuseMin = 5;

__device__ inline bool onFace(const FaceBox& b, int x, int y, int z) {
  const bool in = x >= b.lo[0] && x < b.hi[0] && y >= b.lo[1] && y < b.hi[1] && z >= b.lo[2] &&
                  z < b.hi[2];
  return b.on && in &&
         (x == b.lo[0] || x == b.hi[0] - 1 || y == b.lo[1] || y == b.hi[1] - 1 ||
          z == b.lo[2] || z == b.hi[2] - 1);
}

__device__ inline double extrapolated(const Dims& d, const double* u, const FaceBox& b, int x,
                                      int y, int z, int c) {
  auto at = [&](int i, int j, int k) { return u[3 * (i + (long)j * d.X + (long)k * d.slab) + c]; };
  auto uz = [&](int i, int j, int k) {
    if (k == b.lo[2]) return 2 * at(i, j, k + 1) - at(i, j, k + 2);
    if (k == b.hi[2] - 1) return 2 * at(i, j, k - 1) - at(i, j, k - 2);
    return at(i, j, k);
  };
  auto uy = [&](int i, int j, int k) {
    if (j == b.lo[1]) return 2 * uz(i, j + 1, k) - uz(i, j + 2, k);
    if (j == b.hi[1] - 1) return 2 * uz(i, j - 1, k) - uz(i, j - 2, k);
    return uz(i, j, k);
  };
  if (x == b.lo[0]) return 2 * uy(x + 1, y, z) - uy(x + 2, y, z);
  if (x == b.hi[0] - 1) return 2 * uy(x - 1, y, z) - uy(x - 2, y, z);
  return uy(x, y, z);
}

// X = x - u (:1222-1230) and _vecWorkspace = _velocity (:1232), with the
// face extrapolation of f folded in when f.on (face nodes' u written here)
template <class GridT>
__device__ __forceinline__ void kSolidRecoverX(const GridT& wg, Dims d, double* u,
                                               const double* __restrict__ sv,
                                               double* __restrict__ X, double* __restrict__ svws,
                                               FaceBox f) {
  const long n3 = 3 * d.nodes;
  const long i = 2 * (wg.block() * (long)blockDim.x + threadIdx.x);
  if (i >= n3) return;
  // the pair's elements (the last element may be alone)
  const bool pair = i + 1 < n3;
  double uu[2], vv[2] = {0, 0};
  if (pair) {
    const double2 a = *reinterpret_cast<const double2*>(u + i);
    uu[0] = a.x;
    uu[1] = a.y;
    if (svws) {
      const double2 b = *reinterpret_cast<const double2*>(sv + i);
      vv[0] = b.x;
      vv[1] = b.y;
    }
  } else {
    uu[0] = u[i];
    if (svws) vv[0] = sv[i];
  }
  const int m = pair ? 2 : 1;
  double xx[2];
  for (int e = 0; e < m; e++) {
    const long k = i + e;
    const int n = (int)k / 3;
    const int c = (int)k - 3 * n;
    int x, y, z;
    nodeXYZ(d, n, x, y, z);
    if (onFace(f, x, y, z)) {
      uu[e] = extrapolated(d, u, f, x, y, z, c);
      u[k] = uu[e];
    }
    xx[e] = recoverXAt(d, x, y, z, c, uu[e]);
  }
  if (pair) {
    *reinterpret_cast<double2*>(X + i) = make_double2(xx[0], xx[1]);
    if (svws) *reinterpret_cast<double2*>(svws + i) = make_double2(vv[0], vv[1]);
  } else {
    X[i] = xx[0];
    if (svws) svws[i] = vv[0];
  }
}
ESF_BOUNDS(kSolidRecoverX, 256);

// The step's form of the end of finalizeSolution that stays inside the
// extended box (node range r): the face extrapolation of u at the face
// nodes of f (f.on: every box dimension at least kFaceFuseMin, so each face
// value reads only nodes off every face, which no thread writes), _X = x - u
// inside r (the mass pass reads no other node), and, when svws is given,
// _vecWorkspace = _velocity inside r, the only nodes where the two can
// differ.  A thread per node and component (a face value reads only its
// own component).
template <class GridT>
__device__ __forceinline__ void kSolidBoxFinish(const GridT& wg, Dims d, BoxRange r, FaceBox f,
                                                double* u, const double* __restrict__ sv,
                                                double* __restrict__ svws,
                                                double* __restrict__ X) {
  const long t = wg.block() * (long)blockDim.x + threadIdx.x;
  if (t >= 3 * r.count()) return;
  const long m = t / 3;
  const int c = (int)(t - 3 * m);
  int x, y, z;
  r.at(m, x, y, z);
  const long k = 3 * (x + (long)y * d.X + (long)z * d.slab) + c;
  if (svws) svws[k] = sv[k];
  double v;
  if (f.on && onFace(f, x, y, z)) {
    v = extrapolated(d, u, f, x, y, z, c);
    u[k] = v;
  } else {
    v = u[k];
  }
  X[k] = recoverXAt(d, x, y, z, c, v);
}
ESF_BOUNDS(kSolidBoxFinish, 256);

// DEFOR_SOLID_3D::finalizeSolution(velocity, dt)  (:1182-1197): copy the
// grid velocity into the solid's extended box (nodes outside keep their
// old value: VEC3F_FIELD_3D::clear() is a no-op, VEC3F_FIELD_3D.h:177)
template <class GridT>
__device__ __forceinline__ void kSolidCopy(const GridT& wg, Dims d, BoxRange r, const double* vel,
                                           double* sv) {
  const long t = wg.block() * (long)blockDim.x + threadIdx.x;
  if (t >= r.count()) return;
  const long n = r.node(d, t);
  for (int c = 0; c < 3; c++) sv[3 * n + c] = vel[3 * n + c];
}
ESF_BOUNDS(kSolidCopy, 1024);

// recoverVelocityField for every solid (DEFOR_SOLID_3D.cpp:1235-1241) then
// the Dirichlet faces (COUPLED_INTEGRATOR_3D.cpp:900-916)
struct RecoverArgs {
  int n;
  const Box* box;  // the solids' boxes in device memory
  const double* mass[kMaxSolids];
  const double* vel[kMaxSolids];
};
template <class GridT>
__device__ __forceinline__ void kRecoverBC(const GridT& wg, Dims d, RecoverArgs a, double* vel,
                                           double* intensity, int bc) {
  const long n = wg.block() * (long)blockDim.x + threadIdx.x;
  if (n >= d.nodes) return;
  int x, y, z;
  nodeXYZ(d, n, x, y, z);
#pragma unroll
  for (int b = 0; b < kMaxSolids; b++) {
    if (b >= a.n) break;
    const Box bx = a.box[b];
    if (x >= bx.bMin[0] && x < bx.bMax[0] - 1 && y >= bx.bMin[1] && y < bx.bMax[1] - 1 &&
        z >= bx.bMin[2] && z < bx.bMax[2] - 1 && a.mass[b][n] > 0)
      for (int c = 0; c < 3; c++) vel[3 * n + c] = a.vel[b][3 * n + c];
  }
  if (!bc) return;
  if (x == 0 || x == d.X - 1 || y == 0 || y == d.Y - 1 || z == 0 || z == d.Z - 1)
    for (int c = 0; c < 3; c++) vel[3 * n + c] = 0;
  if (y == 0) intensity[n] = 0;
}
ESF_BOUNDS(kRecoverBC, 1024);

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
#define LAUNCH_NODES(kernel, ...) \
  launch<ESF_K(kernel)>(ctx, dim3(cdiv(ctx->d.nodes, 256)), dim3(256), __VA_ARGS__)

static BodyArgs bodyArgs(esf_ctx* ctx) {
  BodyArgs a{};
  a.vel = ctx->vel;
  a.nf = ctx->nodeFrac;
  a.solidMass = ctx->solidMass;
  a.intensity = ctx->intensity;
  a.ws = ctx->ws;
  a.vort = ctx->vort;
  a.sidx = ctx->sidx;
  a.v0 = ctx->v0;
  a.minv = ctx->minv;
  a.sForce = ctx->sForce;
  a.sMv = ctx->sMv;
  a.sDamped = ctx->sDamped;
  a.fluidDensity = ctx->p.fluid_density;
  a.gravity = ctx->p.gravity;
  a.buoyancy = ctx->p.buoyancy;
  a.dampCoeff = 1 + ctx->dt * ctx->p.rayleigh_alpha;
  a.vortEps = ctx->vortEps;
  a.dt = ctx->dt;
  a.energyPart = ctx->energyPart;
  a.skipSolid = 0;
  return a;
}

esf_status vorticityAndBodyForces(esf_ctx* ctx, double* energy) {
  const Dims& d = ctx->d;
  LAUNCH_NODES(kVorticity, d, ctx->vel, ctx->nodeFrac, ctx->vort, ctx->ws,
               ctx->vortEps > 0 ? 1 : 0);
  const BodyArgs a = bodyArgs(ctx);
  const int nb = cdiv(d.nodes, 256);
  LAUNCH_NODES(kBodyForces, d, a);
  if (energy) {
    std::vector<double> h(nb);
    ESF_HIP(hipMemcpyAsync(h.data(), ctx->energyPart, nb * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    ESF_HIP(streamSync(ctx));
    double e = 0;
    for (double v : h) e += v;
    *energy += e;
  }
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// The same pass split for the step: the solid nodes' part on the context's
// stream (the first solid solve reads it), vorticity and the fluid part on
// the side stream, which the stream joins before the pressure stage
// (joinBodyForces).  The side kernels write vort, ws, v0 at fluid nodes,
// minv and the energy sums; the first solid solve touches none of them.
esf_status forkBodyForces(esf_ctx* ctx) {
  const Dims& d = ctx->d;
  // the side stream is created on first use, not with the context: HIP hands
  // the process's few hardware queues to streams in creation order, so
  // creating it beside every context's own stream would put all of those
  // on every other queue when many contexts are created together
  if (!ctx->side && hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) {
    ctx->side = nullptr;
    return vorticityAndBodyForces(ctx, nullptr);
  }
  BodyArgs a = bodyArgs(ctx);
  if (ctx->nSolid)
    launch<ESF_K(kSolidBodyForces)>(ctx, dim3(cdiv(ctx->nSolid, 256)), dim3(256), d, a, ctx->nSolid,
                                    (const int*)ctx->solidNodes);
  ESF_HIP(hipEventRecord(ctx->evFork, ctx->stream));
  ESF_HIP(hipStreamWaitEvent(ctx->side, ctx->evFork, 0));
  const dim3 grid(cdiv(d.nodes, 256));
  launchOn<ESF_K(kVorticity)>(ctx, ctx->side, grid, dim3(256), d, (const double*)ctx->vel,
                              (const double*)ctx->nodeFrac, ctx->vort, ctx->ws,
                              ctx->vortEps > 0 ? 1 : 0);
  a.skipSolid = 1;
  launchOn<ESF_K(kBodyForces)>(ctx, ctx->side, grid, dim3(256), d, a);
  ESF_HIP(hipEventRecord(ctx->evJoin, ctx->side));
  ctx->joinPending = true;
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status joinBodyForces(esf_ctx* ctx) {
  if (!ctx->joinPending) return ESF_OK;
  ctx->joinPending = false;
  ESF_HIP(hipStreamWaitEvent(ctx->stream, ctx->evJoin, 0));
  return ESF_OK;
}

esf_status solidRhs(esf_ctx* ctx) {
  const long n = 3L * ctx->nSolid;
  if (n)
    launch<ESF_K(kSolidRhs)>(ctx, dim3(cdiv(n, 256)), dim3(256), n, ctx->sForce, ctx->sMv, ctx->dt,
                             ctx->sRhs);
  return ESF_OK;
}

esf_status scatterSolidV0(esf_ctx* ctx) {
  if (ctx->nSolid)
    launch<ESF_K(kScatterSolid)>(ctx, dim3(cdiv(ctx->nSolid, 256)), dim3(256), ctx->nSolid,
                                 ctx->solidNodes, ctx->sV0, ctx->v0);
  return ESF_OK;
}

esf_status correctVelocity(esf_ctx* ctx) {
  // FLIP scenes' particles set masses anywhere: dvPrev kept at every node
  LAUNCH_NODES(kCorrect, ctx->d, ctx->lambda, ctx->v0, ctx->minv, ctx->sidx, ctx->vel, ctx->sB,
               ctx->dvPrev, (const Box*)ctx->dBoxes,
               ctx->p.flip ? -1 : (int)ctx->solids.size());
  return ESF_OK;
}

esf_status solidVelocity(esf_ctx* ctx) {
  if (ctx->nSolid)
    launch<ESF_K(kSolidVelocity)>(ctx, dim3(cdiv(ctx->nSolid, 256)), dim3(256), ctx->nSolid,
                                  ctx->solidNodes, ctx->v0, ctx->sDv, ctx->vel, ctx->dvPrev);
  return ESF_OK;
}

esf_status advectVelocity(esf_ctx* ctx) {
  LAUNCH_NODES(kAdvectVelocity, ctx->d, ctx->vel, ctx->solidMass, ctx->dt, ctx->velws);
  std::swap(ctx->vel, ctx->velws);
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status advectIntensity(esf_ctx* ctx) {
  LAUNCH_NODES(kAdvectScalar, ctx->d, ctx->vel, ctx->solidMass, ctx->intensity, ctx->dt,
               ctx->intws);
  std::swap(ctx->intensity, ctx->intws);
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status solidCopyVelocity(esf_ctx* ctx) {
  // over the extended box's nodes (the solid's velocity elsewhere stays as
  // it was, as the reference's box loop leaves it)
  for (auto& so : ctx->solids) {
    BoxRange r;
    for (int k = 0; k < 3; k++) {
      r.lo[k] = so.box.bMin[k];
      r.hi[k] = std::max(so.box.bMax[k], so.box.bMin[k]);
    }
    if (r.count() > 0) {
      launch<ESF_K(kSolidCopy)>(ctx, dim3(cdiv(r.count(), 256)), dim3(256), ctx->d, r,
                                (const double*)ctx->vel, so.vel);
      // velws now differs from vel at most inside r
      const int rb[6] = {r.lo[0], r.lo[1], r.lo[2], r.hi[0], r.hi[1], r.hi[2]};
      if (so.wsState == 0) {
        so.wsState = 1;
        std::copy(rb, rb + 6, so.wsBox);
      } else if (so.wsState == 1 && !std::equal(rb, rb + 6, so.wsBox)) {
        so.wsState = 2;
      }
    }
  }
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status solidFinalize(esf_ctx* ctx, double dt) {
  for (auto& so : ctx->solids) {
    const int tX = cdiv(ctx->d.X, kSLX), tY = cdiv(ctx->d.Y, kSLY), tZ = cdiv(ctx->d.Z, kSLZ);
    launch<ESF_K(kSolidAdvectU)>(ctx, dim3(tX * tY * tZ), dim3(256), ctx->d, (const double*)so.u,
                                 (const double*)so.vel, dt, so.unew, tX, tY);
    std::swap(so.u, so.unew);
    const Box& b = so.box;
    // the face extrapolation rides with X = x - u where the box is wide
    // enough (every axis at least kFaceFuseMin nodes), else it runs first,
    // axis by axis
    FaceBox f{};
    f.on = 1;
    for (int k = 0; k < 3; k++) {
      f.lo[k] = b.bMin[k];
      f.hi[k] = b.bMax[k];
      f.on = f.on && b.bMax[k] - b.bMin[k] >= kFaceFuseMin;
    }
    if (!f.on) {
      const int axes[3] = {2, 1, 0};
      for (int a : axes) {
        int fa = a == 0 ? 1 : 0, fb = a == 2 ? 1 : 2;
        const long cnt = (long)(b.bMax[fa] - b.bMin[fa]) * (b.bMax[fb] - b.bMin[fb]);
        if (cnt > 0 && b.bMax[a] - b.bMin[a] >= 3)
          launch<ESF_K(kSolidExtrap)>(ctx, dim3(cdiv(cnt, 256)), dim3(256), ctx->d, b, a, so.u);
      }
    }
    // Outside FLIP scenes, with a wide box and velws known to differ from
    // vel at most inside the box the velocity copy wrote, the work stays
    // inside the boxes: the faces, _X and velws there (_X elsewhere waits
    // for materialiseX).  Otherwise one pass over every node: the faces
    // (f.on), _X and velws.
    if (f.on && !ctx->p.flip && so.wsState != 2) {
      BoxRange fr;
      for (int k = 0; k < 3; k++) {
        fr.lo[k] = f.lo[k];
        fr.hi[k] = f.hi[k];
      }
      const int fb[6] = {f.lo[0], f.lo[1], f.lo[2], f.hi[0], f.hi[1], f.hi[2]};
      const bool wsHere = so.wsState == 1 && std::equal(fb, fb + 6, so.wsBox);
      launch<ESF_K(kSolidBoxFinish)>(ctx, dim3(cdiv(3 * fr.count(), 256)), dim3(256), ctx->d,
                                     fr, f, so.u, wsHere ? (const double*)so.vel : nullptr,
                                     wsHere ? so.velws : nullptr, so.X);
      if (so.wsState == 1 && !wsHere) {
        BoxRange wr;
        for (int k = 0; k < 3; k++) {
          wr.lo[k] = so.wsBox[k];
          wr.hi[k] = so.wsBox[3 + k];
        }
        if (wr.count() > 0)
          launch<ESF_K(kSolidCopy)>(ctx, dim3(cdiv(wr.count(), 256)), dim3(256), ctx->d, wr,
                                    (const double*)so.vel, so.velws);
      }
      so.xLazy = true;
      std::copy(fb, fb + 6, so.xBox);
    } else {
      launch<ESF_K(kSolidRecoverX)>(ctx, dim3(cdiv((3 * ctx->d.nodes + 1) / 2, 256)),
                                    dim3(256), ctx->d, so.u, (const double*)so.vel, so.X,
                                    so.velws, f);
      so.xLazy = false;
    }
    so.wsState = 0;
  }
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// _X = x - u at every node, into the buffer X (after a step left it to the
// mass pass)
esf_status materialiseX(esf_ctx* ctx, DevSolid& so) {
  const FaceBox off{};
  launch<ESF_K(kSolidRecoverX)>(ctx, dim3(cdiv((3 * ctx->d.nodes + 1) / 2, 256)), dim3(256),
                                ctx->d, so.u, (const double*)nullptr, so.X, (double*)nullptr,
                                off);
  so.xLazy = false;
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status recoverAndBC(esf_ctx* ctx, int bc) {
  if (ctx->solids.size() > (size_t)kMaxSolids) {
    ctx->err = "too many solids per context";
    return ESF_ERR_INVALID;
  }
  RecoverArgs a{};
  a.n = (int)ctx->solids.size();
  a.box = ctx->dBoxes;
  for (int b = 0; b < a.n; b++) {
    a.mass[b] = ctx->solids[b].mass;
    a.vel[b] = ctx->solids[b].vel;
  }
  LAUNCH_NODES(kRecoverBC, ctx->d, a, ctx->vel, ctx->intensity, bc);
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status fillInt(esf_ctx* ctx, long n, int* p, int v) {
  launch<ESF_K(kFillInt)>(ctx, dim3(1024), dim3(256), n, p, v);
  return ESF_OK;
}
esf_status zeroD(esf_ctx* ctx, long n, double* p) {
  launch<ESF_K(kZero)>(ctx, dim3(1024), dim3(256), n, p);
  return ESF_OK;
}

}  // namespace esf
