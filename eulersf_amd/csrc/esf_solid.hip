// esf_solid.hip -- stress stage and the solid system A = (dt^2 + beta dt) K
// + diag(M~) of the reference, applied matrix-free.
//
// Stress (DEFOR_SOLID_3D::computeStiffnessMatrixAndMaterialForce,
// DEFOR_SOLID_3D.cpp:697-871): one thread per valid cell computes F from the
// 8 corner displacements, the polar decomposition, P, the Cauchy stress, the
// 24 nodal forces, the 24 diagonal stiffness entries, and keeps (R, L, F,
// J^-1) for the matrix-free K.  Nodes then gather their cells' force entries
// in ascending cell order -- the order in which the reference's serial loop
// adds them -- so the force vector needs no atomics.
//
// K x: the reference assembles K_e column by column from 24 calls of dP
// (:805-827) and multiplies the CSR.  dP is linear in dF, so per cell
//   K_e x_e = J^-1 dh^3 dP(sum_i x_i (GF)_i^T) (GF)^T
// needs one dP instead of 24; masked corners (_mass == 0 or no solid index,
// :807, :836-842) contribute nothing, as in the assembled matrix.
#include "esf_corot.h"
#include "esf_internal.h"
#include "esf_krylov.h"

#include <algorithm>
#include <cstdlib>

namespace esf {

__device__ __forceinline__ int sgn(int bit) { return bit ? 1 : -1; }

// kStress: per valid cell slot.  out 24 forces, 24 diag entries, 28 doubles
// of (R, L, F, Jinv).  energy partial per block.
// Per-cell arrays are structure-of-arrays with stride `cap` (the slot
// capacity): component k of slot e lives at [k * cap + e], so a wave reads
// 64 consecutive slots of one component in one coalesced access.
template <class GridT>
__device__ __forceinline__ void kStress(const GridT& wg, Dims d, int nValid, int cap,
                                        const int* __restrict__ cellList,
                                        const double* __restrict__ u,
                                        const double* __restrict__ mass,
                                        const int* __restrict__ sidx, double lambda, double mu,
                                        double* cellMat, double* cellOut, double* cellDiag,
                                        int* cornerDof, double* energyPart,
                                        const double* __restrict__ Fset) {
  __shared__ double red[16];
  const int e = wg.block() * blockDim.x + threadIdx.x;
  double en = 0;
  if (e < nValid) {
    const long index = cellList[e];
    const long off[8] = {0, 1, d.X, d.X + 1, d.slab, d.slab + 1, d.slab + d.X,
                         d.slab + d.X + 1};
    const double g = 0.25 * d.dhInv;
    // computeF  DEFOR_SOLID_3D.cpp:671-682
    double Finv[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const int bit = c == 0 ? (i & 1) : (c == 1 ? ((i >> 1) & 1) : (i >> 2));
          const double G = sgn(bit) * g;
          const double t = u[3 * (index + off[i]) + r] * G;
          acc = (i == 0) ? t : acc + t;
        }
        Finv[3 * r + c] = ((r == c) ? 1.0 : 0.0) - acc;
      }
    double Jd = det(Finv);
    double F[9];
    if (fabs(Jd) > 1e-4) {
      inverse(Finv, F);
    } else {
      for (int k = 0; k < 9; k++) F[k] = (k % 4 == 0) ? 1.0 : 0.0;
      Jd = 1;
    }
    double Jinv = 1.0 / (1.0 / Jd);  // Jinv = 1.0 / computeF(...)  (:760)
    if (Fset) {
      // a deformation gradient given per valid cell (the isolated system's
      // solids): F as given, det F^-1 = 1 / det F
      for (int k = 0; k < 9; k++) F[k] = Fset[9L * e + k];
      Jinv = 1.0 / det(F);
    }
    double R[9], S[9], L[9], P[9];
    corotDecompose(F, lambda, mu, R, S, L);
    corotP(R, S, lambda, mu, P);
    double JP[9], cs[9];
    for (int k = 0; k < 9; k++) JP[k] = Jinv * P[k];
    mulT(JP, F, cs);  // cauchyStress = Jinv * P * F^T  (:763)
    double GF[24];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const double G0 = sgn(i & 1) * g, G1 = sgn((i >> 1) & 1) * g, G2 = sgn(i >> 2) * g;
#pragma unroll
      for (int c = 0; c < 3; c++)
        GF[3 * i + c] = (G0 * F[c] + G1 * F[3 + c]) + G2 * F[6 + c];
    }
    double* out = cellOut + e;
    const double dh = d.dh;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const double G0 = sgn(i & 1) * g, G1 = sgn((i >> 1) & 1) * g, G2 = sgn(i >> 2) * g;
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const double t = (cs[3 * r] * G0 + cs[3 * r + 1] * G1) + cs[3 * r + 2] * G2;
        out[(long)(3 * i + r) * cap] = 0.0 - ((t * dh) * dh) * dh;  // nodalForcesFEM (:801)
      }
    }
    en = ((corotEnergy(S, lambda, mu) * dh) * dh) * dh;
    // diagonal of K_e over unmasked DOFs (:805-827, :834-846)
    double* dg = cellDiag + e;
    for (int i = 0; i < 8; i++) {
      const long n = index + off[i];
      const int si = sidx[n];
      const bool keep = mass[n] != 0 && si >= 0;
      cornerDof[(long)i * cap + e] = keep ? 3 * si : -1;
      for (int j = 0; j < 3; j++) {
        double val = 0;
        if (keep) {
          double dF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
          for (int c = 0; c < 3; c++) dF[3 * j + c] = GF[3 * i + c];
          double dP[9];
          corotDP(R, L, lambda, mu, dF, dP);
          double t = (dP[3 * j] * GF[3 * i] + dP[3 * j + 1] * GF[3 * i + 1]) +
                     dP[3 * j + 2] * GF[3 * i + 2];
          val = (((t * Jinv) * dh) * dh) * dh;
        }
        dg[(long)(3 * i + j) * cap] = val;
      }
    }
    double* m = cellMat + e;
    for (int k = 0; k < 9; k++) {
      m[(long)k * cap] = R[k];
      m[(long)(9 + k) * cap] = L[k];
      m[(long)(18 + k) * cap] = F[k];
    }
    m[27L * cap] = Jinv;
  }
  const double tot = blockSum(en, red);
  if (threadIdx.x == 0) energyPart[wg.block()] = tot;
}
ESF_BOUNDS(kStress, 256);

// Per solid DOF node: gather the stress force and the stiffness diagonal
// from adjacent valid cells (ascending cell order), and computeMV
// (DEFOR_SOLID_3D.cpp:873-889).  `first` initialises (solid 0), later
// solids accumulate, matching `force += localForce` per solid (:865).
template <class GridT>
__device__ __forceinline__ void kSolidGather(const GridT& wg, Dims d, int nSolid, int cap,
                                             const int* __restrict__ solidNodes,
                                             const int* __restrict__ cellSlot,
                                             const double* __restrict__ cellOut,
                                             const double* __restrict__ cellDiag,
                                             const double* __restrict__ smass,
                                             const double* __restrict__ svel, Box box, int first,
                                             double* sForce, double* sKdiag, double* sMv,
                                             uint8_t* nodeMaskByte) {
  const int s = wg.block() * blockDim.x + threadIdx.x;
  if (s >= nSolid) return;
  const long n = solidNodes[s];
  int x, y, z;
  nodeXYZ(d, n, x, y, z);
  double f[3] = {0, 0, 0}, kd[3] = {0, 0, 0};
  const bool rowKeep = smass[n] != 0;
  int nk = 0;
  unsigned mask = 0;
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        const long cell = (x - 1 + dx) + (long)(y - 1 + dy) * d.CX + (long)(z - 1 + dz) * d.cslab;
        const int slot = cellSlot[cell];
        // cells whose stiffness rows reach this node in the Krylov loop:
        // valid cells, and none at all where the node has no mass in this
        // solid (DEFOR_SOLID_3D.cpp:836-842)
        if (slot >= 0 && rowKeep) mask |= 1u << (dx + 2 * dy + 4 * dz);
        if (slot < 0) continue;
        const int corner = (1 - dx) + 2 * (1 - dy) + 4 * (1 - dz);
        for (int r = 0; r < 3; r++) f[r] += cellOut[(long)(3 * corner + r) * cap + slot];
        if (rowKeep) {
          for (int r = 0; r < 3; r++) {
            const double v = cellDiag[(long)(3 * corner + r) * cap + slot];
            kd[r] = nk == 0 ? v : kd[r] + v;
          }
          nk++;
        }
      }
  nodeMaskByte[8L * s] = (uint8_t)mask;
  double mv[3] = {0, 0, 0};
  const bool inBox = x >= box.bMin[0] && x < box.bMax[0] - 1 && y >= box.bMin[1] &&
                     y < box.bMax[1] - 1 && z >= box.bMin[2] && z < box.bMax[2] - 1;
  for (int r = 0; r < 3; r++) {
    if (inBox) mv[r] = smass[n] * svel[3 * n + r];
    if (first) {
      sForce[3 * s + r] = 0.0 + f[r];
      sKdiag[3 * s + r] = kd[r];
      sMv[3 * s + r] = inBox ? 0.0 + mv[r] : 0.0;
    } else {
      sForce[3 * s + r] += f[r];
      if (nk) sKdiag[3 * s + r] += kd[r];
      if (inBox) sMv[3 * s + r] += mv[r];
    }
  }
}
ESF_BOUNDS(kSolidGather, 1024);

// alpha and beta of solid CR iteration i, from the partial sums the node
// pass of iteration i - 1 left in `cur` (delta_i, s.D^-1 s, Ap.D^-1 s,
// Ap.D^-1 Ap; the setup pass for i = 0) and delta_{i-1} from the pass
// before (`prev`).  beta = delta_i / delta_{i-1} (0 for i = 0), alpha =
// delta_i / ((s + beta Ap).D^-1 (s + beta Ap)) with the square expanded
// (CONJUGATE_RESIDUAL.inl:141-160; alpha stays the denominator when that
// is zero, as there).
//
// Canonical order, whatever the block size: the nb block partials form
// groups of 64; a group's sum is a wave sum (waveSum: the same in every
// lane), and the groups are added in ascending order.  Every block of both
// loop kernels therefore holds bitwise the same alpha and beta.
// Split in two so the loads go out with the kernel's first round of loads:
//   scalarLoads  : each wave loads its groups (wave w: groups w, w + nw, ...)
//   scalarFinish : butterflies, group sums through LDS, alpha and beta
constexpr int kScalGroupsPerWave = 2;  // groups a wave loads in one round
constexpr int kScalMaxGroups = 128;    // nb <= 8192 node-pass blocks

struct ScalarLoads {
  double v[kScalGroupsPerWave][5];
};

__device__ inline void scalarLoads(const double* cur, const double* prev, int nb, int g0,
                                   ScalarLoads& L) {
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
#pragma unroll
  for (int u = 0; u < kScalGroupsPerWave; u++) {
    const int i = 64 * (g0 + u * nw) + lane;
    const int ic = i < nb ? i : nb - 1;  // in bounds; masked in scalarFinish
    L.v[u][0] = cur[ic];
    L.v[u][1] = cur[nb + ic];
    L.v[u][2] = cur[2 * nb + ic];
    L.v[u][3] = cur[3 * nb + ic];
    L.v[u][4] = prev[ic];
  }
}

// returns sc->done in every thread (sc->done is read with the partials)
__device__ inline int scalarFinish(const double* cur, const double* prev, int nb, int first,
                                   const CrScalars* sc, ScalarLoads& L, double* grp,
                                   double& alpha, double& beta) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int groups = (nb + 63) / 64;
  const int done = sc->done;
  for (int g0 = w; g0 < groups; g0 += kScalGroupsPerWave * nw) {
    if (g0 != w) scalarLoads(cur, prev, nb, g0, L);  // more groups than one round covers
#pragma unroll
    for (int u = 0; u < kScalGroupsPerWave; u++) {
      const int g = g0 + u * nw;
      const bool ok = 64 * g + lane < nb;
      double t[5];
#pragma unroll
      for (int q = 0; q < 5; q++) t[q] = waveSum(ok ? L.v[u][q] : 0.0);
      if (lane == 0 && g < groups)
#pragma unroll
        for (int q = 0; q < 5; q++) grp[5 * g + q] = t[q];
    }
  }
  __syncthreads();
  double t0 = 0, t1 = 0, t2 = 0, t3 = 0, d0 = 0;
  for (int g = 0; g < groups; g++) {
    t0 += grp[5 * g];
    t1 += grp[5 * g + 1];
    t2 += grp[5 * g + 2];
    t3 += grp[5 * g + 3];
    d0 += grp[5 * g + 4];
  }
  beta = first ? 0.0 : t0 / d0;
  const double apq = (t1 + (2.0 * beta) * t2) + (beta * beta) * t3;
  alpha = (fabs(apq) > 0.0) ? t0 / apq : apq;
  return done;
}

// where the partial sums of the solid CR loop live: three rotating sets of
// 4 x nb (node pass i writes set i mod 3, the setup pass set 2), and the
// residual-check set
struct SolidPartials {
  const double* cur;  // sums of the previous node pass
  const double* prev; // sums of the pass before it
  int nb;             // node-pass blocks
  int first;          // iteration 0: no previous direction, beta = 0
};

// K_e v_e per valid cell, LPC lanes per cell (cell e = thread / LPC); lane l
// of the group takes corners l, l + LPC, ... (8 / LPC of them).  Each lane
// forms its corners' shares v_c (GF)_c^T of dF, a DPP sum over the group
// adds the shares (sumGroup: every lane of the group holds the bitwise-same
// dF), every lane evaluates dP(dF) and writes its corners' three rows of
// K_e v straight into the inbox of the corner's node (24 doubles per solid
// node: 3 rows for each of its 8 cells), so the node pass reads them
// without an index table.  Masked corners write nothing.
//   CELL_IN : v = in                      (apply, residual check)
//   CELL_CR : v = r - alpha (D^-1 Ap beta + D^-1 s), the CR residual update
//             of CONJUGATE_RESIDUAL.inl:223-230 (r -= alpha D^-1 Ap) with the
//             direction update of :257-262 (Ap = Ap beta + s) folded in; the
//             node pass keeps D^-1 Ap and D^-1 s beside Ap and s, so no lane
//             divides, and forms the same values with the same expression
// FMA contraction is on inside the two loop kernels (the products here are
// matrix-free and need no bitwise match with the reference's CSR order).
//   CELL_CRX: CELL_CR and, into a second inbox, K_e x: the true-residual
//             check of the previous iteration's x rides along (the node pass
//             S_CRX finishes it)
enum CellMode { CELL_IN = 0, CELL_CR = 1, CELL_CRX = 2 };
constexpr int kCellLanes = 2;  // lanes per cell of the solid cell pass

template <int LPC>
__device__ inline double sumGroup(double v) {
  if constexpr (LPC == 8) return sum8(v);
  v = v + dppMov<0xB1>(v);
  if constexpr (LPC == 4) v = v + dppMov<0x4E>(v);
  return v;
}

constexpr int kCellThreads = 256;  // cell-pass block (512 measured slower)

template <int MODE, int LPC, class GridT>
__device__ __forceinline__ void kSolidCell(const GridT& wg, Dims d, int nValid, int cap,
                                           const int* __restrict__ cornerDof,
                                           const double* __restrict__ cellMat,
                                           const double* __restrict__ in,
                                           const double* __restrict__ sd,
                                           const double* __restrict__ apd,
                                           const double* __restrict__ xin, const CrScalars* sc,
                                           SolidPartials sp, double lambda, double mu,
                                           double* __restrict__ inbox, double* __restrict__ inboxX) {
#pragma clang fp contract(on)
  static_assert(LPC == 2 || LPC == 4 || LPC == 8, "lanes per cell");
  constexpr bool kCR = MODE != CELL_IN;   // the CR residual update at the corners
  constexpr bool kX = MODE == CELL_CRX;   // also K_e x
  constexpr int CPL = 8 / LPC;  // corners per lane
  __shared__ double grp[5 * kScalMaxGroups];
  const int t = wg.block() * blockDim.x + threadIdx.x;
  const int e = t / LPC, l = t % LPC;
  const bool live = e < nValid;
  // first round of loads, all unconditional (indices clamped in bounds,
  // results masked afterwards) so nothing waits for a branch: the CR
  // scalars' partials, the corners' DOF indices and the cell's (R, L, F, 1/J)
  ScalarLoads SL;
  if (kCR) scalarLoads(sp.cur, sp.prev, sp.nb, threadIdx.x >> 6, SL);
  const int ec = live ? e : nValid - 1;
  int kraw[CPL];
#pragma unroll
  for (int j = 0; j < CPL; j++) kraw[j] = cornerDof[(long)(l + LPC * j) * cap + ec];
  const double* m = cellMat + ec;
  double F[9], R[9], L[9];
#pragma unroll
  for (int q = 0; q < 9; q++) {
    F[q] = m[(long)(18 + q) * cap];
    R[q] = m[(long)q * cap];
    L[q] = m[(long)(9 + q) * cap];
  }
  const double Jinv = m[27L * cap];
  // second round: the corners' r (or v), D^-1 s, D^-1 Ap (and x)
  double rr[CPL][3], ss[CPL][3], aa[CPL][3], xx[CPL][3];
#pragma unroll
  for (int j = 0; j < CPL; j++) {
    const int k = kraw[j] >= 0 ? kraw[j] : 0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      rr[j][q] = in[k + q];
      ss[j][q] = 0;
      aa[j][q] = 0;
      xx[j][q] = 0;
      if (kCR) {
        ss[j][q] = sd[k + q];
        aa[j][q] = apd[k + q];
      }
      if (kX) xx[j][q] = xin[k + q];
    }
  }
  // every load of the pass is issued above; keep the compiler from sinking
  // any of them below the reductions (emits no instruction)
  asm volatile("" ::: "memory");
  double v[CPL][3];
  if (kCR) {
    double alpha, beta;
    if (scalarFinish(sp.cur, sp.prev, sp.nb, sp.first, sc, SL, grp, alpha, beta)) return;
#pragma unroll
    for (int j = 0; j < CPL; j++)
#pragma unroll
      for (int q = 0; q < 3; q++)
        v[j][q] = rr[j][q] - alpha * (aa[j][q] * beta + ss[j][q]);
  } else {
#pragma unroll
    for (int j = 0; j < CPL; j++)
#pragma unroll
      for (int q = 0; q < 3; q++) v[j][q] = rr[j][q];
  }
  // whole groups leave together, so the group sum below only ever reads
  // lanes of a live group
  if (!live) return;
  const double g = 0.25 * d.dhInv;
  const double dh = d.dh;
  double gf[CPL][3];
#pragma unroll
  for (int j = 0; j < CPL; j++) {
    const int c = l + LPC * j;
    const double G0 = sgn(c & 1) * g, G1 = sgn((c >> 1) & 1) * g, G2 = sgn(c >> 2) * g;
#pragma unroll
    for (int q = 0; q < 3; q++) gf[j][q] = (G0 * F[q] + G1 * F[3 + q]) + G2 * F[6 + q];
  }
  // K_e w for the corner values w: dF summed over the group, one dP, the
  // corners' rows into their nodes' inbox (node s = kraw / 3 keeps 8 rows of
  // 3, row 7 - c for the cell at offset 7 - c from it)
  auto product = [&](const double (&w)[CPL][3], double* box) {
    double dF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < CPL; j++)
      if (kraw[j] >= 0)
#pragma unroll
        for (int q = 0; q < 3; q++) {
          dF[q] += w[j][0] * gf[j][q];
          dF[3 + q] += w[j][1] * gf[j][q];
          dF[6 + q] += w[j][2] * gf[j][q];
        }
#pragma unroll
    for (int q = 0; q < 9; q++) dF[q] = sumGroup<LPC>(dF[q]);
    double dP[9];
    corotDP(R, L, lambda, mu, dF, dP);
#pragma unroll
    for (int j = 0; j < CPL; j++) {
      if (kraw[j] < 0) continue;
      const int c = l + LPC * j;
      double* out = box + 8L * kraw[j] + 3 * (7 - c);
#pragma unroll
      for (int r = 0; r < 3; r++) {
        double y = (dP[3 * r] * gf[j][0] + dP[3 * r + 1] * gf[j][1]) + dP[3 * r + 2] * gf[j][2];
        out[r] = (((y * Jinv) * dh) * dh) * dh;
      }
    }
  };
  product(v, inbox);
  if (kX) product(xx, inboxX);
}
ESF_BOUNDS(kSolidCell, kCellThreads);

enum SolidMode { S_APPLY = 0, S_SETUP = 1, S_CR = 2, S_CHECK = 3, S_CRX = 4 };

struct SolidNodeArgs {
  int nSolid;
  int nSolidBodies;       // solids gathered in one launch (<= kMaxSolids)
  const double* inbox;    // solid b: inbox + 24 capS b, 24 doubles per node (kSolidCell)
  const uint64_t* nodeMask; // per node, byte b: bit c = solid b's cell at offset c has rows here
  int capS;
  double scale;           // dt^2 + beta dt
  const double* damped;   // diag mass
  const double* in;       // S_APPLY / S_SETUP / S_CHECK: the vector A acts on
  double* out;            // S_APPLY
  const double* b;        // S_CHECK
  const double* diag;
  double* x;              // S_CR
  double* r;              // S_CR
  double* p;              // S_CR
  double* s;              // S_SETUP, S_CR
  double* Ap;             // S_CR
  double* sd;             // S_SETUP, S_CR: D^-1 s, read by the cell pass
  double* apd;            // S_CR: D^-1 Ap
  const double* inboxX;   // S_CRX: rows of K x (laid out as inbox)
  double* xb;             // S_CRX: x before this iteration's update
  const CrScalars* sc;
  SolidPartials sp;       // S_CR: where alpha and beta come from
  double* partials;       // this pass's block partials
};

// y = scale * (K v) + M~ v per solid DOF (one thread per DOF k = 3 s + q),
// K v summed from the node's inbox rows in ascending cell order, with CR
// epilogues (block partials, summed by the consumers):
//   S_SETUP : s = A r;  r.s, s.D^-1 s
//   S_CR    : Ap = Ap beta + s, p = p beta + r, x += alpha p,
//             r -= alpha (D^-1 Ap beta + D^-1 s) (as the cell pass), s = A r;
//             D^-1 Ap and D^-1 s kept;  r.s, s.D^-1 s, Ap.D^-1 s,
//             Ap.D^-1 Ap
//   S_CHECK : |A x - b|^2  (every 5 iterations, CONJUGATE_RESIDUAL.inl:246-255)
//   S_CRX   : S_CR, and the check of the x it starts from (rows of K x from
//             CELL_CRX): |A x - b|^2 as a fifth sum, x kept in xb
// All loads are in the first round (indices clamped in bounds, masked
// afterwards), so a pass costs one memory round trip.
constexpr int kNodeThreads = 512;  // node-pass block: half the partials of 256 (1024 measured slower)

template <int MODE, class GridT>
__device__ __forceinline__ void kSolidNode(const GridT& wg, Dims d, SolidNodeArgs a) {
#pragma clang fp contract(on)
  constexpr bool kCR = MODE == S_CR || MODE == S_CRX;
  constexpr bool kX = MODE == S_CRX;
  __shared__ double red[5 * (kNodeThreads / 64)];
  __shared__ double grp[5 * kScalMaxGroups];
  const long nDof = 3L * a.nSolid;
  if (nDof == 0) return;
  const int blk = wg.block();
  const long k0 = blk * (long)blockDim.x + threadIdx.x;
  const bool live = k0 < nDof;
  const long k = live ? k0 : nDof - 1;
  const long sn = k / 3;
  const int q = (int)(k - 3 * sn);
  ScalarLoads SL;
  if (kCR) scalarLoads(a.sp.cur, a.sp.prev, a.sp.nb, threadIdx.x >> 6, SL);
  const uint64_t mask = a.nodeMask[sn];
  double rows[8], rowsX[8];
  const double* in0 = a.inbox + 24 * sn + q;
  const double* inX = a.inboxX + 24 * sn + q;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    rows[c] = in0[3 * c];
    rowsX[c] = kX ? inX[3 * c] : 0.0;
  }
  const double dm = a.damped[k];
  double vin = 0, dg = 1, ap0 = 0, s0 = 0, p0 = 0, r0 = 0, x0 = 0, bb = 0, apd0 = 0, sd0 = 0;
  if (kX) bb = a.b[k];
  if (kCR) {
    dg = a.diag[k];
    ap0 = a.Ap[k];
    s0 = a.s[k];
    p0 = a.p[k];
    r0 = a.r[k];
    x0 = a.x[k];
    apd0 = a.apd[k];
    sd0 = a.sd[k];
  } else {
    vin = a.in[k];
    if (MODE == S_SETUP) dg = a.diag[k];
    if (MODE == S_CHECK) bb = a.b[k];
  }
  // every load of the pass is issued above; keep the compiler from sinking
  // any of them below the reductions (emits no instruction)
  asm volatile("" ::: "memory");
  double kx = 0, kxx = 0;
#pragma unroll
  for (int c = 0; c < 8; c++)
    if (mask >> c & 1u) {
      kx += rows[c];
      if (kX) kxx += rowsX[c];
    }
  // further solids (scenes with several bodies): their rows, in solid order
  for (int b = 1; b < a.nSolidBodies; b++) {
    const double* in = in0 + 24L * a.capS * b;
    const double* ix = inX + 24L * a.capS * b;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      rows[c] = in[3 * c];
      if (kX) rowsX[c] = ix[3 * c];
    }
#pragma unroll
    for (int c = 0; c < 8; c++)
      if (mask >> (8 * b + c) & 1u) {
        kx += rows[c];
        if (kX) kxx += rowsX[c];
      }
  }
  double alpha = 0, beta = 0;
  if (kCR) {
    if (scalarFinish(a.sp.cur, a.sp.prev, a.sp.nb, a.sp.first, a.sc, SL, grp, alpha, beta))
      return;
  }
  double part[5] = {0, 0, 0, 0, 0};
  if (kX && live) {
    // the check of x as this iteration found it (as S_CHECK forms it)
    const double y = kxx * a.scale + dm * x0;
    const double t = y - bb;
    part[4] = t * t;
    a.xb[k] = x0;
  }
  if (kCR) {
    const double ap = ap0 * beta + s0;
    const double pp = p0 * beta + r0;
    const double rn = r0 - alpha * (apd0 * beta + sd0);
    const double y = kx * a.scale + dm * rn;
    if (live) {
      const double yq = y / dg, aq = ap / dg;
      a.x[k] = x0 + alpha * pp;
      a.Ap[k] = ap;
      a.p[k] = pp;
      a.r[k] = rn;
      a.s[k] = y;
      a.sd[k] = yq;
      a.apd[k] = aq;
      part[0] = rn * y;
      part[1] = y * yq;
      part[2] = ap * yq;
      part[3] = ap * aq;
    }
  } else {
    const double y = kx * a.scale + dm * vin;
    if (live) {
      if (MODE == S_APPLY) {
        a.out[k] = y;
      } else if (MODE == S_SETUP) {
        const double yq = y / dg;
        a.s[k] = y;
        a.sd[k] = yq;
        part[0] = vin * y;
        part[1] = y * yq;
      } else {
        const double t = y - bb;
        part[0] = t * t;
      }
    }
  }
  if (MODE == S_APPLY) return;
  if (kX) {
    blockPartialsAt<5>(part, a.partials, red, wg.blocks(), blk);
  } else {
    double p4[4] = {part[0], part[1], part[2], part[3]};
    blockPartialsAt<4>(p4, a.partials, red, wg.blocks(), blk);
  }
}
ESF_BOUNDS(kSolidNode, kNodeThreads);

// the end of a solid solve whose convergence a riding check found: x back
// to the iterate that check accepted
template <class GridT>
__device__ __forceinline__ void kSolidFixX(const GridT& wg, long n, const CrScalars* sc,
                                           const double* xb, double* x, CrScalars* report) {
  // the solve's report into its pinned slot (read after a later
  // synchronisation)
  if (wg.block() == 0 && threadIdx.x == 0) *report = *sc;
  if (!(sc->done && sc->aux != 0.0)) return;
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)wg.blocks() * blockDim.x)
    x[i] = xb[i];
}
ESF_BOUNDS(kSolidFixX, 1024);

// diag(A) = scale * diag(K) + M~   (COUPLED_INTEGRATOR_3D.cpp:613-617)
template <class GridT>
__device__ __forceinline__ void kSolidDiag(const GridT& wg, long n, const double* kdiag,
                                           const double* damped, double scale, double* adiag) {
  const long i = wg.block() * (long)blockDim.x + threadIdx.x;
  if (i < n) adiag[i] = kdiag[i] * scale + damped[i];
}
ESF_BOUNDS(kSolidDiag, 1024);


// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
esf_status stressStage(esf_ctx* ctx, double* energy) {
  const Dims& d = ctx->d;
  if (ctx->solids.size() > (size_t)kMaxSolids) {
    ctx->err = "too many solids per context";
    return ESF_ERR_INVALID;
  }
  double en = 0;
  for (size_t b = 0; b < ctx->solids.size(); b++) {
    DevSolid& so = ctx->solids[b];
    const int nb = cdiv(std::max(so.nValid, 1), 256);
    if (so.nValid > 0)
      launch<ESF_K(kStress)>(ctx, dim3(nb), dim3(256), d, so.nValid, so.capValid, so.cellList, so.u,
                             so.mass, ctx->sidx, so.lambda, so.mu, so.cellMat, so.cellOut,
                             so.cellDiag, so.cornerDof, ctx->partials,
                             so.nFset == so.nValid ? (const double*)so.Fset : nullptr);
    if (energy && so.nValid > 0) {
      std::vector<double> h(nb);
      ESF_HIP(hipMemcpyAsync(h.data(), ctx->partials, nb * sizeof(double),
                             hipMemcpyDeviceToHost, ctx->stream));
      ESF_HIP(streamSync(ctx));
      for (double v : h) en += v;
    }
    if (ctx->nSolid > 0)
      launch<ESF_K(kSolidGather)>(ctx, dim3(cdiv(ctx->nSolid, 256)), dim3(256), d, ctx->nSolid,
                                  so.capValid, ctx->solidNodes, so.cellSlot, so.cellOut,
                                  so.cellDiag, so.mass, so.vel, so.box, b == 0 ? 1 : 0, ctx->sForce,
                                  ctx->sKdiag, ctx->sMv, (uint8_t*)ctx->nodeMask + b);
  }
  if (energy) *energy = en;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("stress stage: ") + hipGetErrorString(e);
    return ESF_ERR_DEVICE;
  }
  return ESF_OK;
}

static void fillNodeArgs(esf_ctx* ctx, SolidNodeArgs& a) {
  a.nSolid = ctx->nSolid;
  a.nSolidBodies = (int)ctx->solids.size();
  a.inbox = ctx->inbox;
  a.nodeMask = ctx->nodeMask;
  a.capS = ctx->capSolid;
  a.scale = ctx->dt * ctx->dt + ctx->p.rayleigh_beta * ctx->dt;
  a.damped = ctx->sDamped;
  a.diag = ctx->sAdiag;
}

// the CR-update launches are kernel classes 14 (cells) and 15 (nodes) of
// esf_kernel_timing
// lanes per cell of the cell pass: kCellLanes, or ESF_SOLID_LANES (2, 4
// or 8), for tuning
static int cellLanes() {
  static const int lanes = [] {
    const char* e = getenv("ESF_SOLID_LANES");
    const int v = e ? atoi(e) : 0;
    return v == 2 || v == 4 || v == 8 ? v : kCellLanes;
  }();
  return lanes;
}

template <int MODE, int LPC>
static void cellPassLanes(esf_ctx* ctx, const double* in, const CrScalars* sc,
                          const SolidPartials& sp, const double* x) {
  const int cls = MODE == CELL_CR ? 14 : (MODE == CELL_CRX ? 16 : -1);
  for (size_t b = 0; b < ctx->solids.size(); b++) {
    DevSolid& so = ctx->solids[b];
    if (so.nValid > 0) {
      launchTimed<ESF_KT(kSolidCell, MODE, LPC)>(ctx, cls,
                                                 dim3(cdiv((long)LPC * so.nValid, kCellThreads)),
                                                 dim3(kCellThreads), ctx->d, so.nValid, so.capValid,
                                                 (const int*)so.cornerDof,
                                                 (const double*)so.cellMat, in,
                                                 (const double*)ctx->sSd, (const double*)ctx->sApd,
                                                 x, sc, sp, so.lambda, so.mu,
                                                 ctx->inbox + 24L * ctx->capSolid * b,
                                                 ctx->inboxX + 24L * ctx->capSolid * b);
    }
  }
}

template <int MODE>
static void cellPass(esf_ctx* ctx, const double* in, const CrScalars* sc,
                     const SolidPartials& sp, const double* x = nullptr) {
  const int lanes = cellLanes();
  if (lanes == 8)
    cellPassLanes<MODE, 8>(ctx, in, sc, sp, x);
  else if (lanes == 4)
    cellPassLanes<MODE, 4>(ctx, in, sc, sp, x);
  else
    cellPassLanes<MODE, 2>(ctx, in, sc, sp, x);
}

static int nodeBlocks(const esf_ctx* ctx) {
  return cdiv(std::max(3L * ctx->nSolid, 1L), kNodeThreads);
}

template <int MODE>
static void nodePass(esf_ctx* ctx, SolidNodeArgs& a) {
  launchTimed<ESF_KT(kSolidNode, MODE)>(ctx, MODE == S_CR ? 15 : (MODE == S_CRX ? 17 : -1),
                                        dim3(nodeBlocks(ctx)), dim3(kNodeThreads), ctx->d, a);
}

esf_status solidAdiag(esf_ctx* ctx) {
  const long n = 3L * ctx->nSolid;
  if (n == 0) return ESF_OK;
  const double scale = ctx->dt * ctx->dt + ctx->p.rayleigh_beta * ctx->dt;
  launch<ESF_K(kSolidDiag)>(ctx, dim3(cdiv(n, 256)), dim3(256), n, ctx->sKdiag, ctx->sDamped, scale,
                            ctx->sAdiag);
  return ESF_OK;
}

esf_status solidApplyDevice(esf_ctx* ctx, const double* x, double* y) {
  if (ctx->solids.size() > (size_t)kMaxSolids) {
    ctx->err = "too many solids per context";
    return ESF_ERR_INVALID;
  }
  cellPass<CELL_IN>(ctx, x, nullptr, SolidPartials{});
  SolidNodeArgs a{};
  fillNodeArgs(ctx, a);
  a.in = x;
  a.out = y;
  nodePass<S_APPLY>(ctx, a);
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// CONJUGATE_RESIDUAL::solvePCR(SpMat&, b, x)  (CONJUGATE_RESIDUAL.inl:194-276)
// with the matrix-free A; x starts at zero.  Two launches per iteration:
//   cells : K_e r_new with r_new = r - alpha D^-1 (Ap beta + s) formed at
//           the corners
//   nodes : the direction update p = p beta + r, Ap = Ap beta + s of the
//           previous iteration (.inl:257-262), x += alpha p, r = r_new,
//           s = A r_new, and the reductions for delta, beta and alpha
// plus the true-residual check every 5 iterations (.inl:246-255).  alpha =
// delta / (Ap.D^-1 Ap) needs Ap after beta is known; expanding
// (s + beta Ap).D^-1 (s + beta Ap) = s.D^-1 s + 2 beta Ap.D^-1 s +
// beta^2 Ap.D^-1 Ap lets the node kernel's one reduction provide it, so
// no separate direction pass (and no third grid-wide reduction) is needed.
// Only the rounding of alpha differs from the reference's evaluation.
template <class GridT>
__device__ __forceinline__ void kSolidSetup(const GridT& wg, long n, const double* b,
                                            const double* diag, double* x, double* r, double* p,
                                            double* Ap, double* apd, double* partials,
                                            CrScalars* sc, double eps) {
  __shared__ double red[16];
  if (wg.block() == 0 && threadIdx.x == 0) initScalarBlock(sc, eps);
  double part = 0;
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)wg.blocks() * blockDim.x) {
    x[i] = 0.0;
    r[i] = b[i] / diag[i];
    p[i] = 0.0;   // with beta = 0 the first direction is p = r, Ap = s
    Ap[i] = 0.0;
    apd[i] = 0.0;
    part += b[i] * b[i];
  }
  double mine[1] = {part};
  blockPartialsAt<1>(mine, partials, red, wg.blocks(), wg.block());
}
ESF_BOUNDS(kSolidSetup, 1024);

// a warm start: x at each solid DOF from a node field (3 doubles per node)
template <class GridT>
__device__ __forceinline__ void kWarmGather(const GridT& wg, int nSolid, const int* solidNodes,
                                            const double* guess, double* x, CrScalars* sc,
                                            double eps) {
  const int s = wg.block() * blockDim.x + threadIdx.x;
  if (s == 0) initScalarBlock(sc, eps);
  if (s >= nSolid) return;
  const long n = solidNodes[s];
  for (int c = 0; c < 3; c++) x[3L * s + c] = guess[3 * n + c];
}
ESF_BOUNDS(kWarmGather, 1024);

// the CR set-up from a warm start x0, given y = A x0: r = D^-1 (b - y), no
// direction yet; partial sums |b|^2 and |b - A x0|^2
template <class GridT>
__device__ __forceinline__ void kSolidSetupWarm(const GridT& wg, long n, const double* b,
                                                const double* diag, const double* y, double* r,
                                                double* p, double* Ap, double* apd,
                                                double* partials) {
  __shared__ double red[2 * 4];
  double part[2] = {0, 0};
  for (long i = wg.block() * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)wg.blocks() * blockDim.x) {
    const double res = b[i] - y[i];
    r[i] = res / diag[i];
    p[i] = 0.0;
    Ap[i] = 0.0;
    apd[i] = 0.0;
    part[0] += b[i] * b[i];
    part[1] += res * res;
  }
  blockPartialsAt<2>(part, partials, red, wg.blocks(), wg.block());
}
ESF_BOUNDS(kSolidSetupWarm, 1024);

esf_status solidSolve(esf_ctx* ctx, int which, const double* b, double* x, double eps,
                      int maxIt, int slot, const double* guess) {
  if (ctx->solids.size() > (size_t)kMaxSolids) {
    ctx->err = "too many solids per context";
    return ESF_ERR_INVALID;
  }
  const long n = 3L * ctx->nSolid;
  if (n == 0) {
    // nothing queued for this slot: report "converged after 0 iterations"
    ESF_HIP(streamSync(ctx));
    ctx->hostSc[slot] = CrScalars{};
    ctx->hostSc[slot].done = 1;
    ctx->hostSc[slot].iter = 0;
    return ESF_OK;
  }
  if (nodeBlocks(ctx) > 64 * kScalMaxGroups) {
    ctx->err = "too many solid nodes for one context (limit " +
               std::to_string((long)kNodeThreads * 64 * kScalMaxGroups / 3) + ")";
    return ESF_ERR_INVALID;
  }
  CrScalars* sc = &ctx->sc[which];
  const int eb = std::max(1, std::min(cdiv(n, 1024), 256));
  // a check is waited for once two iterations are queued behind it
  DonePoller poll(ctx, which, false, 2);
  if (guess) {
    // warm start x0 from the node field: r = D^-1 (b - A x0), and done at
    // once if x0 already meets the tolerance
    // (the scalar block is initialised by the first thread)
    launch<ESF_K(kWarmGather)>(ctx, dim3(cdiv(ctx->nSolid, 256)), dim3(256), ctx->nSolid,
                               (const int*)ctx->solidNodes, guess, x, sc, eps);
    esf_status st = solidApplyDevice(ctx, x, ctx->stmp);
    if (st) return st;
    launch<ESF_K(kSolidSetupWarm)>(ctx, dim3(eb), dim3(256), n, b, (const double*)ctx->sAdiag,
                                   (const double*)ctx->stmp, ctx->sr, ctx->sp, ctx->sAp, ctx->sApd,
                                   ctx->partials);
    launchFinish<F_WARM, 2>(ctx, eb, sc, 0, 0, nullptr, poll.slot());
    poll.mark(-1);
  } else {
    launch<ESF_K(kSolidSetup)>(ctx, dim3(eb), dim3(256), n, b, (const double*)ctx->sAdiag, x,
                               ctx->sr, ctx->sp, ctx->sAp, ctx->sApd, ctx->partials, sc, eps);
    launchFinish<F_BNORM, 1>(ctx, eb, sc);
  }
  // partial sets of 5 x nb: node pass i writes set i mod 3 (the setup pass
  // set 2; row 4 is the check that rides along), a lone check its own set
  const int nb = nodeBlocks(ctx);
  double* set[3] = {ctx->sPartials, ctx->sPartials + 5L * nb, ctx->sPartials + 10L * nb};
  double* checkSet = ctx->sPartials + 15L * nb;
  SolidNodeArgs a{};
  fillNodeArgs(ctx, a);
  a.b = b;
  a.x = x;
  a.r = ctx->sr;
  a.p = ctx->sp;
  a.s = ctx->ss;
  a.Ap = ctx->sAp;
  a.sd = ctx->sSd;
  a.apd = ctx->sApd;
  a.sc = sc;
  cellPass<CELL_IN>(ctx, ctx->sr, sc, SolidPartials{});
  a.in = ctx->sr;
  a.partials = set[2];
  nodePass<S_SETUP>(ctx, a);
  a.inboxX = ctx->inboxX;
  a.xb = ctx->sXb;
  // The true residual of the x that iteration i leaves (every 5 iterations,
  // CONJUGATE_RESIDUAL.inl:246-255) is checked by the passes of iteration
  // i + 1: its cell pass forms K_e x beside K_e r_new, its node pass the
  // sum |A x - b|^2 and a copy of x before its own update, and the finisher
  // decides for iteration i.  If that check converges, kSolidFixX restores
  // the accepted x at the end, so the solve returns the reference's iterate
  // and iteration count.  A check after the last iteration runs alone.
  for (int i = 0; i < maxIt; i++) {
    SolidPartials sp{set[(i + 2) % 3], set[(i + 1) % 3], nb, i == 0 ? 1 : 0};
    a.sp = sp;
    a.partials = set[i % 3];
    if (i > 1 && (i - 1) % 5 == 0) {
      cellPass<CELL_CRX>(ctx, ctx->sr, sc, sp, x);
      nodePass<S_CRX>(ctx, a);
      launchFinish<F_CHECK, 1>(ctx, nb, sc, i - 1, 1, set[i % 3] + 4L * nb, poll.slot());
      poll.mark(i);
    } else {
      cellPass<CELL_CR>(ctx, ctx->sr, sc, sp);
      nodePass<S_CR>(ctx, a);
    }
    if (i > 0 && i % 5 == 0 && i + 1 == maxIt) {
      cellPass<CELL_IN>(ctx, x, sc, SolidPartials{});
      a.in = x;
      a.partials = checkSet;
      nodePass<S_CHECK>(ctx, a);
      launchFinish<F_CHECK, 1>(ctx, nb, sc, i, 0, checkSet, poll.slot());
      poll.mark(i);
    }
    if (poll.drain(i)) break;
  }
  // the accepted x, and the report written without waiting (the caller
  // reads report slot `slot` after its next synchronisation)
  launch<ESF_K(kSolidFixX)>(ctx, dim3(std::max(1, std::min(cdiv(n, 256), 1024))), dim3(256), n,
                            (const CrScalars*)sc, (const double*)ctx->sXb, x,
                            &ctx->hostSc[slot]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("solid solve: ") + hipGetErrorString(e);
    return ESF_ERR_DEVICE;
  }
  return ESF_OK;
}

// batch corotation for unit tests: one thread per F
__global__ void kCorotBatch(int n, double lambda, double mu, const double* F, const double* dF,
                            double* out, double* energy, double* dP) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[9], R[9], S[9], L[9], P[9];
  for (int k = 0; k < 9; k++) f[k] = F[9 * (long)i + k];
  corotDecompose(f, lambda, mu, R, S, L);
  corotP(R, S, lambda, mu, P);
  for (int k = 0; k < 9; k++) {
    out[36 * (long)i + k] = R[k];
    out[36 * (long)i + 9 + k] = S[k];
    out[36 * (long)i + 18 + k] = L[k];
    out[36 * (long)i + 27 + k] = P[k];
  }
  energy[i] = corotEnergy(S, lambda, mu);
  double g[9], q[9];
  for (int k = 0; k < 9; k++) g[k] = dF[9 * (long)i + k];
  corotDP(R, L, lambda, mu, g, q);
  for (int k = 0; k < 9; k++) dP[9 * (long)i + k] = q[k];
}

}  // namespace esf
