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
  const T* Dinv;     // its reciprocals (M_UPDATE)
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

constexpr int TX = 32, TY = 8;
constexpr int HALO = (TX + 2) * (TY + 2);  // cells of one layer incl. halo
constexpr int NODES = (TX + 1) * (TY + 1); // nodes of one layer
constexpr int NRING = HALO - TX * TY;      // halo cells around the tile's own
constexpr int HPT = 2;                     // halo cells per thread: its own, one of the ring
constexpr int NEDGE = NODES - TX * TY;     // nodes on the tile's low x and low y edges
static_assert(TX * TY == 256 && NRING <= 256, "one 256-thread block per tile");

// One block owns a TX x TY tile of cells and `layers` cell layers of it,
// marching up in z.  Iteration k of the march forms node layer k and cell
// layer k - 2, so one barrier per layer separates producers from
// consumers:
//   - node layer k: w = minv * (J^T x) from cell layers k - 1 and k.  Each
//     thread owns the node at the upper corner of its cell and carries that
//     node's sums over its four cells of the lower layer from the previous
//     iteration, so it reads only the four of the upper layer; the NEDGE
//     nodes on the tile's low x and low y edges are formed by the first
//     threads from LDS alone (three x buffers: layers k - 1, k, k + 1);
//   - cell layer k - 2: y = J w, its sum over the corner nodes of layer
//     k - 2 carried from the previous iteration, those of layer k - 1 read
//     from LDS (two w buffers);
//   - x of layer k + 1 goes to LDS; layer k + 2 and the pointwise operands
//     of the next output layer are fetched, minv of layer k + 1 at the top
//     of the iteration.
// The additions keep the reference's sparse row/column order exactly.  A
// corner node off the interior carries w = 0 and is added like the others:
// the running sum starts at +0 and adding a zero never changes it.
// Out-of-grid halo cells are read at a clamped address and replaced by 0,
// so the loop has no per-cell branches.
template <typename T, int MODE, int DEPTH, class GridT>
__device__ __forceinline__ void kSpmv(const GridT& wg, Dims d, SpmvArgs<T> a, int tilesX, int tilesY,
                                      int tilesZ, int layers) {
  if (MODE != M_APPLY && a.sc && a.sc->done) return;
  __shared__ T xs[3 * HALO];
  __shared__ T ws[2 * NODES * 3];
  __shared__ double red[5 * 4];

  // XCD-aware tile order: each XCD gets a contiguous run of tiles, so
  // neighbouring tiles' shared halos stay in one L2
  const int tile = wg.block();
  const int tx = tile % tilesX, ty = (tile / tilesX) % tilesY,
            tz = tile / (tilesX * tilesY);
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * layers;
  const int zEnd = min(z0 + layers, d.CZ);

  T alpha = 0, beta = 0;
  if (MODE == M_UPDATE) {
    alpha = (T)a.sc->alpha;
    beta = (T)a.sc->beta;
  }
  double part0 = 0, part1 = 0, part2 = 0, part3 = 0, part4 = 0;

  const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
  const int cx = x0 + lx, cy = y0 + ly;
  const bool cellXY = cx < d.CX && cy < d.CY;
  // in-layer offset of this thread's cell (clamped into the grid)
  const int cOff = min(cx, d.CX - 1) + min(cy, d.CY - 1) * d.CX;

  // halo slots of this thread: slot 0 is its own cell, slot 1 (threads
  // below NRING) a cell of the ring around the tile -- the low row, the
  // high row, then the low and the high column between them.  Per slot:
  // in-layer offset (clamped), validity and LDS position.  Owning its cell's
  // halo values lets the CR update keep them in registers until the cell's
  // output layer instead of reading them again.
  int hOff[HPT], hIdx[HPT];
  bool hVal[HPT];
  {
    const int r = threadIdx.x < NRING ? threadIdx.x : 0;
    const int q = r - 2 * (TX + 2);
    const int rx = r < TX + 2 ? r : (r < 2 * (TX + 2) ? r - (TX + 2) : (q < TY ? 0 : TX + 1));
    const int ry = r < TX + 2 ? 0 : (r < 2 * (TX + 2) ? TY + 1 : 1 + q % TY);
    const int hx[HPT] = {lx + 1, rx}, hy[HPT] = {ly + 1, ry};
#pragma unroll
    for (int h = 0; h < HPT; h++) {
      const int gx = x0 - 1 + hx[h], gy = y0 - 1 + hy[h];
      hIdx[h] = hy[h] * (TX + 2) + hx[h];
      hVal[h] = gx >= 0 && gx < d.CX && gy >= 0 && gy < d.CY;
      hOff[h] = min(max(gx, 0), d.CX - 1) + min(max(gy, 0), d.CY - 1) * d.CX;
    }
    // a thread without a ring cell re-reads its own (no extra lines)
    if (threadIdx.x >= NRING) hOff[1] = hOff[0];
  }
  // this thread's node: tile node (lx + 1, ly + 1), the upper corner of its
  // cell; its edge node (threads below NEDGE): (0, t) for t <= TY, else
  // (t - TY, 0).  In-layer offsets clamped to the interior.
  const bool hasEdge = threadIdx.x < NEDGE;
  const int enx = threadIdx.x <= TY ? 0 : threadIdx.x - TY;
  const int eny = threadIdx.x <= TY ? threadIdx.x : 0;
  const int mI = x0 + lx + 1, mJ = y0 + ly + 1, eI = x0 + enx, eJ = y0 + eny;
  const bool mainIJ = mI >= 1 && mI <= d.X - 2 && mJ >= 1 && mJ <= d.Y - 2;
  const bool edgeIJ = hasEdge && eI >= 1 && eI <= d.X - 2 && eJ >= 1 && eJ <= d.Y - 2;
  const int mOff = min(max(mI, 1), d.X - 2) + min(max(mJ, 1), d.Y - 2) * d.X;
  // a thread without an edge node re-reads its main node's minv
  const int eOff = hasEdge ? min(max(eI, 1), d.X - 2) + min(max(eJ, 1), d.Y - 2) * d.X : mOff;
  // LDS positions
  const int mHalo = (ly + 1) * (TX + 2) + lx + 1;  // lowest of the main node's cells
  const int eHalo = eny * (TX + 2) + enx;          // lowest of the edge node's cells
  const int mW = 3 * ((ly + 1) * (TX + 1) + lx + 1);
  const int eW = 3 * (eny * (TX + 1) + enx);
  const int cW = 3 * (ly * (TX + 1) + lx);         // this cell's lowest corner node

  // Every global load of the loop is issued unconditionally at the end of
  // an iteration for the iteration DEPTH later (addresses clamped into the
  // grid and to the block's own last layers, so the loads issued past the
  // end of the march re-read lines just fetched; their values are not
  // used), in the order that iteration
  // consumes them -- minv (node stage), the output cell's operands (cell
  // stage), x (LDS store) -- so each consumer waits only for its own
  // loads.  With DEPTH 2 two such sets are in flight.
  struct Pre {
    T mvMain, mvEdge;                 // minv of the main and the edge node
    T in[HPT], ap[HPT], s[HPT], dd[HPT];  // raw values of the halo cells
    float e[HPT];
    T eB, eD, eX, eP;                 // operands of the output cell
  };
  auto fetchX = [&](Pre& P, int cz) {
    const long zb = (long)min(max(cz, 0), min(zEnd, d.CZ - 1)) * d.cslab;
#pragma unroll
    for (int h = 0; h < HPT; h++) {
      const long c = zb + hOff[h];
      P.in[h] = a.in[c];
      if (MODE == M_UPDATE) {
        P.ap[h] = a.Ap[c];
        P.s[h] = a.s[c];
        P.dd[h] = a.Dinv[c];
      } else if (MODE == M_IR_RESID) {
        P.e[h] = a.e32 ? a.e32[c] : 0.0f;
      }
    }
  };
  // x of layer cz into LDS: the halo values (out-of-grid cells 0)
  auto store = [&](const Pre& P, T* L, int cz) {
    const bool zok = cz >= 0 && cz < d.CZ;
#pragma unroll
    for (int h = 0; h < HPT; h++) {
      if (h == 0 || threadIdx.x < NRING) {
        T v = P.in[h];
        if (MODE == M_UPDATE) {
          // r_new from r_old, Ap_old and s_old, on the halo as well (the
          // owner writes it with the cell's other updates)
          const T apn = P.ap[h] * beta + P.s[h];
          const T qq = apn * P.dd[h];
          v = P.in[h] - alpha * qq;
        } else if (MODE == M_IR_RESID) {
          // x_old (+ e); the owner writes x_new to the other buffer
          if (a.e32) v = v + (T)P.e[h];
        }
        L[hIdx[h]] = zok && hVal[h] ? v : (T)0;
      }
    }
  };
  // M_UPDATE: the own cell's Ap_new = beta Ap + s, 1/D and r_old, taken from
  // its halo slot as each layer enters LDS and kept for three layers, until
  // the cell stage of that layer (the slot after every store, so the third
  // set is always the output layer's)
  T oAp0 = 0, oAp1 = 0, oAp2 = 0, oD0 = 0, oD1 = 0, oD2 = 0, oR0 = 0, oR1 = 0, oR2 = 0;
  auto keepOwn = [&](const Pre& P) {
    if (MODE != M_UPDATE) return;
    oAp2 = oAp1;
    oAp1 = oAp0;
    oAp0 = P.ap[0] * beta + P.s[0];
    oD2 = oD1;
    oD1 = oD0;
    oD0 = P.dd[0];
    oR2 = oR1;
    oR1 = oR0;
    oR0 = P.in[0];
  };
  auto fetchMinv = [&](Pre& P, int k) {
    const long kb = (long)max(min(max(k, 1), min(zEnd, d.Z - 2)), 0) * d.slab;
    P.mvMain = a.minv[kb + mOff];
    P.mvEdge = a.minv[kb + eOff];
  };
  auto fetchOut = [&](Pre& P, int cz) {
    const long c = (long)min(max(cz, 0), zEnd - 1) * d.cslab + cOff;
    if (MODE == M_SETUP_R || MODE == M_CHECK || MODE == M_IR_RESID) P.eB = a.b[c];
    if (MODE == M_SETUP_R || MODE == M_SETUP_S) P.eD = a.D[c];
    if (MODE == M_UPDATE) {
      P.eX = a.x[c];
      P.eP = a.p[c];
    }
  };
  // what iteration k consumes: minv of node layer k, the operands of output
  // cell layer k - 2, x of layer k + 1
  auto issue = [&](Pre& P, int k) {
    fetchMinv(P, k);
    if (MODE != M_APPLY) fetchOut(P, k - 2);
    fetchX(P, k + 1);
  };

  // a node's additions over its adjacent cells, in ascending cell index
  // (dz, dy, dx): lowerSums adds the four of the lower layer to zero,
  // nodeW the four of the upper layer, then scales by minv
  auto lowerSums = [&](const T (&lo)[4], T& m0, T& m1, T& m2) {
    m0 = 0;
    m1 = 0;
    m2 = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const T v = lo[q];
      m0 += (q & 1) == 0 ? v : -v;
      m1 += (q >> 1) == 0 ? v : -v;
      m2 += v;
    }
  };
  auto nodeW = [&](T m0, T m1, T m2, const T (&hi)[4], T mi, bool interior, T* W) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const T v = hi[q];
      m0 += (q & 1) == 0 ? v : -v;
      m1 += (q >> 1) == 0 ? v : -v;
      m2 += -v;
    }
    W[0] = interior ? m0 * mi : (T)0;
    W[1] = interior ? m1 * mi : (T)0;
    W[2] = interior ? m2 * mi : (T)0;
  };
  // a cell's additions over four corner nodes of one layer (oz), in
  // ascending node index (oy, ox)
  auto cornerSums = [&](T acc, const T (&wc)[12], int oz) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const T* W = &wc[3 * q];
      acc += (q & 1) ? W[0] : -W[0];
      acc += (q >> 1) ? W[1] : -W[1];
      acc += oz ? W[2] : -W[2];
    }
    return acc;
  };

  T ml0, ml1, ml2;   // main node: sums over its lower-layer cells
  T accLo = 0;       // output cell: sum over its lower-layer corners
  T own0 = 0, own1 = 0, own2 = 0;  // own input of cell layers k, k - 1, k - 2
  // LDS buffer offsets: x of layers k - 1, k, k + 1; w of layers k, k - 1
  int bLo = 0, bHi = HALO, bNx = 2 * HALO;
  int wCur = 0, wPrev = 3 * NODES;

  // one iteration of the march (k: node layer), consuming the loads in P
  // and refilling it for iteration k + DEPTH
  auto step = [&](int k, Pre& P) {
    T up0 = 0;
    if (k <= zEnd) {
      const bool kin = k >= 1 && k <= d.Z - 2;
      const T* Lhi = xs + bHi;
      T hi[4];
      hi[0] = Lhi[mHalo];
      hi[1] = Lhi[mHalo + 1];
      hi[2] = Lhi[mHalo + TX + 2];
      hi[3] = Lhi[mHalo + TX + 3];
      nodeW(ml0, ml1, ml2, hi, P.mvMain, mainIJ && kin, ws + wCur + mW);
      lowerSums(hi, ml0, ml1, ml2);
      up0 = hi[0];
      if (hasEdge) {
        const T* Llo = xs + bLo;
        T elo[4], ehi[4];
        elo[0] = Llo[eHalo];
        elo[1] = Llo[eHalo + 1];
        elo[2] = Llo[eHalo + TX + 2];
        elo[3] = Llo[eHalo + TX + 3];
        ehi[0] = Lhi[eHalo];
        ehi[1] = Lhi[eHalo + 1];
        ehi[2] = Lhi[eHalo + TX + 2];
        ehi[3] = Lhi[eHalo + TX + 3];
        T e0, e1, e2;
        lowerSums(elo, e0, e1, e2);
        nodeW(e0, e1, e2, ehi, P.mvEdge, edgeIJ && kin, ws + wCur + eW);
      }
    }
    own2 = own1;
    own1 = own0;
    own0 = up0;
    if (k >= z0 + 1) {
      // node layer k - 1: the upper corners of cell layer k - 2, the lower
      // corners of cell layer k - 1
      const T* W = ws + wPrev + cW;
      T wu[12];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        wu[c] = W[c];
        wu[3 + c] = W[3 + c];
        wu[6 + c] = W[3 * (TX + 1) + c];
        wu[9 + c] = W[3 * (TX + 1) + 3 + c];
      }
      const int cz = k - 2;
      if (cz >= z0 && cz < zEnd) {
        const T acc = cornerSums(accLo, wu, 1);
        // the cell's own input; the CR update forms r_new again from the kept
        // operands, by the expression that stored it in LDS
        const T ownIn = MODE == M_UPDATE ? oR2 - alpha * (oAp2 * oD2) : own2;
        const long c = (long)cz * d.cslab + cOff;
        if (cellXY) {
          if (MODE == M_APPLY) {
            a.out[c] = acc;
          } else if (MODE == M_SETUP_R) {
            const T rv = (P.eB - acc) / P.eD;
            a.r[c] = rv;
            a.p[c] = rv;
            part0 += (double)P.eB * (double)P.eB;
          } else if (MODE == M_SETUP_S) {
            a.s[c] = acc;
            a.Ap[c] = acc;
            part0 += (double)ownIn * (double)acc;
            const T q = acc / P.eD;
            part1 += (double)acc * (double)q;
          } else if (MODE == M_UPDATE) {
            // the cell's deferred direction update (Ap, p, x) and its
            // r_new, written to the other buffers, and s = S r_new
            // (Ap_new, 1/D and r_old kept from the halo slot)
            const T apn = oAp2;
            const T pn = P.eP * beta + oR2;
            a.r[c] = ownIn;
            a.ApOut[c] = apn;
            a.p[c] = pn;
            a.x[c] = P.eX + alpha * pn;
            a.sOut[c] = acc;
            const T sq = acc * oD2;
            part0 += (double)ownIn * (double)acc;
            part1 += (double)acc * (double)sq;
            part2 += (double)apn * (double)sq;
            part3 += (double)apn * (double)(apn * oD2);
            // |b - S x| as the recursion carries it (r is D^-1 (b - S x));
            // only the FP32 inner solve tests it
            if (a.recursive) {
              const double rb = (double)ownIn / (double)oD2;
              part4 += rb * rb;
            }
          } else if (MODE == M_CHECK) {
            const T t = acc - P.eB;
            part0 += (double)t * (double)t;
          } else if (MODE == M_IR_RESID) {
            if (a.e32) a.x[c] = ownIn;
            const T rr = P.eB - acc;
            a.r32[c] = (float)rr;
            part0 += (double)rr * (double)rr;
            part1 += (double)P.eB * (double)P.eB;
          }
        }
      }
      accLo = cornerSums((T)0, wu, 0);
    }
    if (k + 1 <= zEnd) store(P, xs + bNx, k + 1);
    keepOwn(P);
    issue(P, k + DEPTH);
    // rotate the LDS buffers
    const int b = bLo;
    bLo = bHi;
    bHi = bNx;
    bNx = b;
    const int w = wCur;
    wCur = wPrev;
    wPrev = w;
    __syncthreads();
  };

  Pre P[DEPTH];
  fetchX(P[0], z0 - 1);
  store(P[0], xs + bLo, z0 - 1);
  keepOwn(P[0]);
  fetchX(P[0], z0);
  store(P[0], xs + bHi, z0);
  keepOwn(P[0]);
#pragma unroll
  for (int j = 0; j < DEPTH; j++) issue(P[j], z0 + j);
  __syncthreads();
  {
    T lo[4];
    lo[0] = xs[bLo + mHalo];
    lo[1] = xs[bLo + mHalo + 1];
    lo[2] = xs[bLo + mHalo + TX + 2];
    lo[3] = xs[bLo + mHalo + TX + 3];
    lowerSums(lo, ml0, ml1, ml2);
  }
  // layers z0 - 1 .. zEnd pass through: node layers z0 .. zEnd, cell
  // layers z0 .. zEnd - 1 (two iterations behind).  The march runs in
  // groups of DEPTH iterations, one load set each (trailing iterations past
  // zEnd + 1 do nothing but load).
  if (DEPTH == 1) {
    for (int k = z0; k <= zEnd + 1; k++) step(k, P[0]);
  } else {
    for (int k = z0; k <= zEnd + 1; k += DEPTH) {
#pragma unroll
      for (int j = 0; j < DEPTH; j++) step(k + j, P[j]);
    }
  }
  if (MODE == M_APPLY) return;
  if (MODE == M_UPDATE) {
    double mine[5] = {part0, part1, part2, part3, part4};
    blockPartialsAt<5>(mine, a.partials, red, wg.blocks(), wg.block());
  } else {
    double mine[2] = {part0, part1};
    blockPartialsAt<2>(mine, a.partials, red, wg.blocks(), wg.block());
  }
}
ESF_BOUNDS(kSpmv, 256);

// b = J v0 - c, c (boundary velocities), D = diag(S); also the FP32 copies
// used by the mixed solve.  COUPLED_INTEGRATOR_3D.cpp:286-325 (c),
// :694-709 (D, b).  D == nullptr: b and c only (the mixed step forms the
// diagonal with kPressureDiag when its inner solve first needs it).
template <class GridT>
__device__ __forceinline__ void kPressureRhs(const GridT& wg, Dims d, const double* __restrict__ v0,
                                             const double* __restrict__ vel,
                                             const double* __restrict__ minv, double* c, double* b,
                                             double* D, double* Dinv, float* D32, float* Dinv32,
                                             CrScalars* sc, double eps) {
  const int r = wg.block() * blockDim.x + threadIdx.x;
  if (sc && r == 0) initScalarBlock(sc, eps);
  if (r >= d.cells) return;
  const int q = r / d.CX, cx = r - q * d.CX, cy = q % d.CY, cz = q / d.CY;
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
      if (D) {
        const double mi = minv[n];
        dd += mi * 1.0;
        dd += mi * 1.0;
        dd += mi * 1.0;
      }
    } else {
      cc -= sx * vel[3 * n];
      cc -= sy * vel[3 * n + 1];
      cc -= sz * vel[3 * n + 2];
    }
  }
  c[r] = cc;
  b[r] = tmp - cc;
  if (D) {
    D[r] = dd;
    Dinv[r] = 1.0 / dd;
    if (D32) {
      D32[r] = (float)dd;
      Dinv32[r] = (float)(1.0 / dd);
    }
  }
}
ESF_BOUNDS(kPressureRhs, 1024);

// D = diag(S), its reciprocals and their FP32 copies, and minv in FP32, for
// the mixed solve's inner CR: the sums of kPressureRhs in the same order
template <class GridT>
__device__ __forceinline__ void kPressureDiag(const GridT& wg, Dims d,
                                              const double* __restrict__ minv, double* D,
                                              double* Dinv, float* D32, float* Dinv32,
                                              float* minv32) {
  const int t = wg.block() * blockDim.x + threadIdx.x;
  if (t < d.nodes) minv32[t] = (float)minv[t];
  if (t >= d.cells) return;
  const int q = t / d.CX, cx = t - q * d.CX, cy = q % d.CY, cz = q / d.CY;
  const long base = cx + (long)cy * d.X + (long)cz * d.slab;
  double dd = 0;
  for (int k = 0; k < 8; k++) {
    const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
    const int ni = cx + ox, nj = cy + oy, nk = cz + oz;
    if (ni >= 1 && ni <= d.X - 2 && nj >= 1 && nj <= d.Y - 2 && nk >= 1 && nk <= d.Z - 2) {
      const double mi = minv[base + ox + (long)oy * d.X + (long)oz * d.slab];
      dd += mi * 1.0;
      dd += mi * 1.0;
      dd += mi * 1.0;
    }
  }
  D[t] = dd;
  Dinv[t] = 1.0 / dd;
  D32[t] = (float)dd;
  Dinv32[t] = (float)(1.0 / dd);
}
ESF_BOUNDS(kPressureDiag, 1024);

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// The z-march's launch shape, fixed per context when it is created:
//  - cell layers per block: each block loads two layers of x and forms one
//    node layer beyond the ones it outputs, so deeper blocks waste less,
//    but fewer blocks hide less latency.  The FP64 CR update (the most
//    registers) takes 16 layers, shallower only so that at least 512
//    blocks run; the other launches (the FP32 kernels, the FP64 residual of
//    the refinement, the applies and checks) take at least 1024 blocks.
//    On MI355X at 128^3 that is 8 layers: FP32 update 20.5 vs 22.8 us and
//    refinement residual 23.3 vs 27.9 us against 16, while the FP64 update
//    measured 30.5 at 16 against 32.5 at 8.  ESF_SPMV_LAYERS (4..64)
//    overrides both choices;
//  - how many iterations ahead the loads run (1..3): two sets of loads in
//    flight hide more latency but hold more registers, so the CR update
//    (more operands) keeps one.  ESF_SPMV_DEPTH (1..3) overrides both.
// The overrides are for tuning and tests; they are read at context creation.
static int envInt(const char* name, int lo, int hi) {
  const char* e = getenv(name);
  const int v = e ? atoi(e) : 0;
  return v >= lo && v <= hi ? v : 0;
}

void spmvConfigure(esf_ctx* ctx) {
  const Dims& d = ctx->d;
  const long tilesXY = (long)cdiv(d.CX, TX) * cdiv(d.CY, TY);
  auto layersFor = [&](long minBlocks) {
    int layers = 16;
    while (layers > 4 && tilesXY * cdiv(d.CZ, layers) < minBlocks) layers /= 2;
    return layers;
  };
  const int forcedLayers = envInt("ESF_SPMV_LAYERS", 4, 64);
  ctx->spmvLayers = forcedLayers ? forcedLayers : layersFor(512);
  ctx->spmvLayersWide = forcedLayers ? forcedLayers : layersFor(1024);
  const int forcedDepth = envInt("ESF_SPMV_DEPTH", 1, 3);
  ctx->spmvDepthUpdate = forcedDepth ? forcedDepth : 1;
  ctx->spmvDepthOther = forcedDepth ? forcedDepth : 2;
}

// `deep`: the FP64 CR update's shape
static void tileShape(const esf_ctx* ctx, bool deep, int& tx, int& ty, int& tz, int& layers) {
  const Dims& d = ctx->d;
  tx = cdiv(d.CX, TX);
  ty = cdiv(d.CY, TY);
  layers = deep ? ctx->spmvLayers : ctx->spmvLayersWide;
  tz = cdiv(d.CZ, layers);
}

template <typename T, int MODE>
esf_status launchSpmv(esf_ctx* ctx, const SpmvArgs<T>& a) {
  int tx, ty, tz, layers;
  tileShape(ctx, sizeof(T) == 8 && MODE == M_UPDATE, tx, ty, tz, layers);
  const int nb = tx * ty * tz;
  const int depth = MODE == M_UPDATE ? ctx->spmvDepthUpdate : ctx->spmvDepthOther;
  if (depth == 3)
    launchTimed<ESF_KT(kSpmv, T, MODE, 3)>(ctx, MODE + (sizeof(T) == 4 ? 6 : 0), dim3(nb),
                                           dim3(256), ctx->d, a, tx, ty, tz, layers);
  else if (depth == 2)
    launchTimed<ESF_KT(kSpmv, T, MODE, 2)>(ctx, MODE + (sizeof(T) == 4 ? 6 : 0), dim3(nb),
                                           dim3(256), ctx->d, a, tx, ty, tz, layers);
  else
    launchTimed<ESF_KT(kSpmv, T, MODE, 1)>(ctx, MODE + (sizeof(T) == 4 ? 6 : 0), dim3(nb),
                                           dim3(256), ctx->d, a, tx, ty, tz, layers);
  // the grid-wide sums of the mode (kSpmv's epilogue partials)
  if constexpr (MODE == M_SETUP_R) launchFinish<F_BNORM, 1>(ctx, nb, a.sc);
  if constexpr (MODE == M_SETUP_S) launchFinish<F_SETUP_S, 2>(ctx, nb, a.sc);
  if constexpr (MODE == M_UPDATE)
    launchFinish<F_CR, 5>(ctx, nb, a.sc, a.iter, a.recursive, nullptr, a.hostOut);
  if constexpr (MODE == M_CHECK)
    launchFinish<F_CHECK, 1>(ctx, nb, a.sc, a.iter, 0, nullptr, a.hostOut);
  if constexpr (MODE == M_IR_RESID)
    launchFinish<F_IR, 2>(ctx, nb, a.sc, 0, a.first, nullptr, a.hostOut);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("kSpmv launch: ") + hipGetErrorString(e);
    return ESF_ERR_DEVICE;
  }
  return ESF_OK;
}

// the most blocks a launch takes (the partials buffer's size)
int spmvBlocks(const esf_ctx* ctx) {
  int nb = 0;
  for (bool deep : {true, false}) {
    int tx, ty, tz, layers;
    tileShape(ctx, deep, tx, ty, tz, layers);
    nb = std::max(nb, tx * ty * tz);
  }
  return nb;
}

// sc (may be null): a scalar block initialised with tolerance eps for the
// solve that follows
esf_status pressureRhs(esf_ctx* ctx, CrScalars* sc, double eps) {
  const Dims& d = ctx->d;
  const bool mixed = ctx->p.precision == ESF_MIXED;
  launch<ESF_K(kPressureRhs)>(ctx, dim3(cdiv(d.cells, 256)), dim3(256), d, ctx->v0, ctx->vel,
                              ctx->minv, ctx->c, ctx->b, ctx->D, ctx->Dinv,
                              mixed ? ctx->D32 : nullptr, ctx->Dinv32, sc, eps);
  ctx->diagPending = false;
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// the mixed step's right-hand side: b and c only.  The diagonal and minv in
// FP32 follow (pressureDiagFlush) only where an inner solve runs, or where
// a caller reads the diagonal.
// It also initialises scalar block `sc` with tolerance eps for the solve
// that follows (null: none).
esf_status pressureRhsDeferDiag(esf_ctx* ctx, CrScalars* sc, double eps) {
  const Dims& d = ctx->d;
  launch<ESF_K(kPressureRhs)>(ctx, dim3(cdiv(d.cells, 256)), dim3(256), d, ctx->v0, ctx->vel,
                              ctx->minv, ctx->c, ctx->b, (double*)nullptr, (double*)nullptr,
                              (float*)nullptr, (float*)nullptr, sc, eps);
  ctx->diagPending = true;
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status pressureDiagFlush(esf_ctx* ctx) {
  if (!ctx->diagPending) return ESF_OK;
  const Dims& d = ctx->d;
  launch<ESF_K(kPressureDiag)>(ctx, dim3(cdiv(std::max(d.nodes, d.cells), 256)), dim3(256), d,
                               (const double*)ctx->minv, ctx->D, ctx->Dinv, ctx->D32,
                               ctx->Dinv32, ctx->minv32);
  ctx->diagPending = false;
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

template <class GridT>
__device__ __forceinline__ void kToFloat(const GridT& wg, long n, const double* a, float* b) {
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)wg.blocks() * blockDim.x)
    b[i] = (float)a[i];
}
ESF_BOUNDS(kToFloat, 1024);
// streaming copy of n doubles (calibration of the HBM byte counters and
// the box's copy bandwidth): 16-byte vector loads and stores
template <class GridT>
__device__ __forceinline__ void kCopy(const GridT& wg, long n2, const double2* __restrict__ a,
                                      double2* __restrict__ b) {
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n2;
       i += (long)wg.blocks() * blockDim.x)
    b[i] = a[i];
}
ESF_BOUNDS(kCopy, 256);

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
  launch<ESF_K(kCopy)>(ctx, dim3(blocks), dim3(256), n2, (const double2*)a, b);
  hipEventRecord(ctx->ev[2], ctx->stream);
  for (int i = 0; i < reps; i++)
    launch<ESF_K(kCopy)>(ctx, dim3(blocks), dim3(256), n2, (const double2*)a, b);
  hipEventRecord(ctx->ev[3], ctx->stream);
  hipError_t e = eventSync(ctx, ctx->ev[3]);
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
    launch<ESF_K(kToFloat)>(ctx, dim3(1024), dim3(256), d.nodes, (const double*)ctx->minv,
                            ctx->minv32);
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
  if (eventSync(ctx, ctx->ev[3]) != hipSuccess) return ESF_ERR_DEVICE;
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

// the scalar block a finisher queued into report slot 3, after a
// synchronisation
static esf_status waitSlot3(esf_ctx* ctx, CrScalars* out) {
  ESF_HIP(streamSync(ctx));
  *out = ctx->hostSc[3];
  return ESF_OK;
}

// CONJUGATE_RESIDUAL::solvePCR (matrix-free) on S, FP64 or FP32 vectors.
// The update multiplies by Dinv = 1 / D where the reference divides by D
// (JACOBI_PRECONDITIONER.h:35-37): the iterates differ from the
// reference's in rounding, as alpha's evaluation already makes them.
// x is the start vector.  Scalar block `which`.  *iters receives the
// reference's iteration count; *relres the last checked |Sx-b|/|b|.
// recursive = false: the reference's test, the true residual every 5
// iterations (CONJUGATE_RESIDUAL.inl:166-176).  recursive = true (the FP32
// inner solve of the mixed form, whose result the FP64 outer residual
// checks): the finisher of every iteration tests the residual the
// recursion carries, and no check SpMV runs.
template <typename T>
esf_status crSolve(esf_ctx* ctx, int which, const T* minv, const T* b, const T* D,
                   const T* Dinv, T* x, T* r, T* r2, T* p, T* s, T* s2, T* Ap, T* Ap2,
                   double eps, int maxIt, int slot, bool recursive, bool scalarsReady) {
  // scalarsReady: the kernel before the solve initialised scalar block `which`
  esf_status st = scalarsReady ? ESF_OK : initScalars(ctx, which, eps);
  if (st) return st;
  const Dims& d = ctx->d;
  SpmvArgs<T> a{};
  a.minv = minv;
  a.b = b;
  a.D = D;
  a.Dinv = Dinv;
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
  // the recursive solve's finishers test convergence on the device at every
  // iteration (later launches exit at once); the host reads every third of
  // them, since each read leaves the stream idle for some microseconds
  // (measured at 128^3: mixed steps/s 893 reading every one, 914 every
  // third, 909 every fourth, 902 every sixth)
  constexpr int kPollEvery = 3;
  T *sA = s, *sB = s2, *apA = Ap, *apB = Ap2;
  a.recursive = recursive ? 1 : 0;
  // iteration `it`: the fused update and S r, then the buffers swap
  auto update = [&](int it, CrScalars* hostOut) {
    a.iter = it;
    a.in = rA;
    a.r = rB;
    a.s = sA;
    a.sOut = sB;
    a.Ap = apA;
    a.ApOut = apB;
    a.hostOut = hostOut;
    const esf_status e = launchSpmv<T, M_UPDATE>(ctx, a);
    std::swap(rA, rB);
    std::swap(sA, sB);
    std::swap(apA, apB);
    return e;
  };
  int i = 0;
  for (; i < maxIt; i++) {
    const bool read = recursive && i % kPollEvery == kPollEvery - 1;
    if ((st = update(i, read ? poll.slot() : nullptr))) return st;
    if (recursive) {
      if (read) poll.mark(i);
    } else if (i > 0 && i % 5 == 0) {
      a.in = x;
      a.hostOut = poll.slot();
      if ((st = launchSpmv<T, M_CHECK>(ctx, a))) return st;
      poll.mark(i);
    }
    if (poll.drain(i)) {
      // a solve that stagnated runs on to `lag` iterations past the check
      // that showed it, however early the host saw that check, so the
      // iterate it leaves does not depend on the timing of the polls (a
      // converged solve's further launches exit at once)
      while (poll.stagnated && i < poll.stopIt && i + 1 < maxIt)
        if ((st = update(++i, nullptr))) return st;
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
                                    const double*, const double*, double*, double*, double*,
                                    double*, double*, double*, double*, double*, double, int,
                                    int, bool, bool);
template esf_status crSolve<float>(esf_ctx*, int, const float*, const float*, const float*,
                                   const float*, float*, float*, float*, float*, float*,
                                   float*, float*, float*, double, int, int, bool, bool);

// the inner solve's correction zeroed, and its scalar block initialised
template <class GridT>
__device__ __forceinline__ void kZeroF(const GridT& wg, long n, float* a, CrScalars* sc,
                                       double eps) {
  if (wg.block() == 0 && threadIdx.x == 0) initScalarBlock(sc, eps);
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)wg.blocks() * blockDim.x)
    a[i] = 0.0f;
}
ESF_BOUNDS(kZeroF, 1024);

// Mixed-precision pressure solve: FP32 CR on S (FP32 minv, FP32 vectors,
// FP64 inner products) inside FP64 iterative refinement.  The outer
// residual b - S x, the correction x += e and |r| stay in FP64.
esf_status mixedSolve(esf_ctx* ctx, double eps, int maxIt, int* iters, bool* converged,
                      double* relres, int* refinements, int* innerIters, bool scalarsReady) {
  const Dims& d = ctx->d;
  // scalarsReady: the right-hand side's kernel initialised scalar block 0
  esf_status st = scalarsReady ? ESF_OK : initScalars(ctx, 0, eps);
  if (st) return st;
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
  o.hostOut = &ctx->hostSc[3];
  if ((st = launchSpmv<double, M_IR_RESID>(ctx, o))) return st;
  CrScalars h;
  if ((st = waitSlot3(ctx, &h))) return st;
  int outer = 0, inner = 0;
  // refinement rounds: each gains roughly the FP32 solve's attainable
  // reduction (about eps32 * cond on the smooth modes), so an ill-conditioned
  // cold start needs many; the total FP32 iterations stay capped by maxIt
  const int maxOuter = 1000;
  while (!h.done && outer < maxOuter && inner < maxIt) {
    // inner tolerance: ask for the reduction the outer test still needs,
    // with a margin, bounded by what FP32 can deliver.  The residual a
    // converged step leaves sits in the slowly converging modes, and the
    // inner CR's iterations grow with the reduction asked of them, so the
    // margin is shallow: on the 128^3 ball falling from rest, margins
    // 0.9 / 0.7 / 0.5 measured 7.7 / 6.7 / 8.7 FP32 iterations per step over
    // 30 steps (777 / 792 / 772 steps/s), and on the ball thrown at
    // (0.5, -1.4, 0.3) 13209 / 13879 / 15214 (2.42 / 2.30 / 2.10 steps/s);
    // earlier, slower builds measured 0.3 / 0.1 / 0.03 well below 0.5.
    double need = h.resnorm > 0 ? 0.7 * eps * h.bnorm / h.resnorm : 1.0;
    double innerEps = std::min(0.5, std::max(need, ctx->p.inner_eps));
    if (outer == 0) {
      // the inner solve's operands: minv in FP32, and the diagonal where
      // the step left it for here
      if (ctx->diagPending) {
        if ((st = pressureDiagFlush(ctx))) return st;
      } else {
        launch<ESF_K(kToFloat)>(ctx, dim3(1024), dim3(256), d.nodes, (const double*)ctx->minv,
                                ctx->minv32);
      }
    }
    launch<ESF_K(kZeroF)>(ctx, dim3(1024), dim3(256), d.cells, ctx->e32, &ctx->sc[1], innerEps);
    if ((st = crSolve<float>(ctx, 1, ctx->minv32, ctx->rhs32, ctx->D32, ctx->Dinv32, ctx->e32,
                             ctx->r32,
                             ctx->r32b, ctx->p32, ctx->s32, ctx->s32b, ctx->Ap32, ctx->Ap32b,
                             innerEps, maxIt - inner, 4, true, true)))
      return st;
    outer++;
    o.e32 = ctx->e32;
    o.first = 0;
    o.in = ctx->lambda;
    o.x = ctx->lambda2;
    if ((st = launchSpmv<double, M_IR_RESID>(ctx, o))) return st;
    std::swap(ctx->lambda, ctx->lambda2);
    if ((st = waitSlot3(ctx, &h))) return st;  // also completes report slot 4
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
