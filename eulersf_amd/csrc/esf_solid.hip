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

namespace esf {

__device__ __forceinline__ int sgn(int bit) { return bit ? 1 : -1; }

// kStress: per valid cell slot.  out 24 forces, 24 diag entries, 28 doubles
// of (R, L, F, Jinv).  energy partial per block.
// Per-cell arrays are structure-of-arrays with stride `cap` (the slot
// capacity): component k of slot e lives at [k * cap + e], so a wave reads
// 64 consecutive slots of one component in one coalesced access.
__global__ void __launch_bounds__(256) kStress(Dims d, int nValid, int cap,
                                               const int* __restrict__ cellList,
                                               const double* __restrict__ u,
                                               const double* __restrict__ mass,
                                               const int* __restrict__ sidx, double lambda,
                                               double mu, double* cellMat, double* cellOut,
                                               double* cellDiag, int* cornerDof,
                                               double* energyPart) {
  __shared__ double red[8];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
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
    const double Jinv = 1.0 / (1.0 / Jd);  // Jinv = 1.0 / computeF(...)  (:760)
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
  if (threadIdx.x == 0) energyPart[blockIdx.x] = tot;
}

// Per solid DOF node: gather the stress force and the stiffness diagonal
// from adjacent valid cells (ascending cell order), and computeMV
// (DEFOR_SOLID_3D.cpp:873-889).  `first` initialises (solid 0), later
// solids accumulate, matching `force += localForce` per solid (:865).
__global__ void kSolidGather(Dims d, int nSolid, int cap, const int* __restrict__ solidNodes,
                             const int* __restrict__ cellSlot, const double* __restrict__ cellOut,
                             const double* __restrict__ cellDiag,
                             const double* __restrict__ smass, const double* __restrict__ svel,
                             Box box, int first, double* sForce, double* sKdiag, double* sMv,
                             int* nodeSlot, int capS) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nSolid) return;
  const long n = solidNodes[s];
  const int x = n % d.X, y = (n / d.X) % d.Y, z = n / d.slab;
  double f[3] = {0, 0, 0}, kd[3] = {0, 0, 0};
  const bool rowKeep = smass[n] != 0;
  int nk = 0;
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        const long cell = (x - 1 + dx) + (long)(y - 1 + dy) * d.CX + (long)(z - 1 + dz) * d.cslab;
        const int slot = cellSlot[cell];
        // stiffness gather table of the Krylov loop: rows whose node has
        // no mass in this solid take no K contribution (DEFOR_SOLID_3D.cpp:836-842)
        nodeSlot[(long)(dx + 2 * dy + 4 * dz) * capS + s] = rowKeep ? slot : -1;
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

// K_e v_e per valid cell, eight lanes per cell (lane c = corner c of cell
// e = thread / 8).  Each lane forms its corner's share v_c (GF)_c^T of dF,
// a 3-step xor butterfly sums the shares (commutative at every step, so
// all eight lanes hold the bitwise-same dF), every lane evaluates dP(dF)
// and writes its corner's three rows.  Output is per cell, 24 contiguous
// doubles [24 e + 3 c + r], so a wave's stores are one contiguous run.
//   CELL_IN : v = in                      (apply, residual check)
//   CELL_CR : v = r - alpha (Ap beta + s) / D, the CR residual update of
//             CONJUGATE_RESIDUAL.inl:223-230 with the direction update of
//             :257-262 folded in; the node kernel forms the same values
enum CellMode { CELL_IN = 0, CELL_CR = 1 };

template <int MODE>
__global__ void __launch_bounds__(256) kSolidCell8(Dims d, int nValid, int cap,
                                                   const int* __restrict__ cornerDof,
                                                   const double* __restrict__ cellMat,
                                                   const double* __restrict__ in,
                                                   const double* __restrict__ s,
                                                   const double* __restrict__ Ap,
                                                   const double* __restrict__ diag,
                                                   const CrScalars* sc, double lambda, double mu,
                                                   double* __restrict__ cellKx) {
  if (MODE == CELL_CR && sc->done) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = t >> 3, c = t & 7;
  // whole groups of eight leave together, so the butterfly below only
  // ever reads lanes of a live group
  if (e >= nValid) return;
  const int k = cornerDof[(long)c * cap + e];
  double v0 = 0, v1 = 0, v2 = 0;
  if (k >= 0) {
    if (MODE == CELL_IN) {
      v0 = in[k];
      v1 = in[k + 1];
      v2 = in[k + 2];
    } else {
      const double alpha = sc->alpha, beta = sc->beta;
      v0 = in[k] - alpha * ((Ap[k] * beta + s[k]) / diag[k]);
      v1 = in[k + 1] - alpha * ((Ap[k + 1] * beta + s[k + 1]) / diag[k + 1]);
      v2 = in[k + 2] - alpha * ((Ap[k + 2] * beta + s[k + 2]) / diag[k + 2]);
    }
  }
  const double* m = cellMat + e;
  double F[9];
#pragma unroll
  for (int q = 0; q < 9; q++) F[q] = m[(long)(18 + q) * cap];
  const double g = 0.25 * d.dhInv;
  const double G0 = sgn(c & 1) * g, G1 = sgn((c >> 1) & 1) * g, G2 = sgn(c >> 2) * g;
  double gf[3];
#pragma unroll
  for (int q = 0; q < 3; q++) gf[q] = (G0 * F[q] + G1 * F[3 + q]) + G2 * F[6 + q];
  double dF[9];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    dF[q] = v0 * gf[q];
    dF[3 + q] = v1 * gf[q];
    dF[6 + q] = v2 * gf[q];
  }
#pragma unroll
  for (int off = 1; off < 8; off <<= 1)
#pragma unroll
    for (int q = 0; q < 9; q++) dF[q] += __shfl_xor(dF[q], off, 8);
  double R[9], L[9];
#pragma unroll
  for (int q = 0; q < 9; q++) {
    R[q] = m[(long)q * cap];
    L[q] = m[(long)(9 + q) * cap];
  }
  const double Jinv = m[27L * cap];
  double dP[9];
  corotDP(R, L, lambda, mu, dF, dP);
  const double dh = d.dh;
  double* out = cellKx + 24L * e + 3 * c;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    double y = 0;
    if (k >= 0) {
      y = (dP[3 * r] * gf[0] + dP[3 * r + 1] * gf[1]) + dP[3 * r + 2] * gf[2];
      y = (((y * Jinv) * dh) * dh) * dh;
    }
    out[r] = y;
  }
}

enum SolidMode { S_APPLY = 0, S_SETUP = 1, S_CR = 2, S_CHECK = 3 };

struct SolidNodeArgs {
  int nSolid;
  int nSolidBodies;       // up to 4 solids gathered in one launch
  const double* cellKx[4];
  const int* nodeSlot;    // [(8 b + c) capS + s]: slot of the cell at offset c, or -1
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
  CrScalars* sc;
  double* partials;
  int iter;
};

// y = scale * (K v) + M~ v per solid DOF node, K v gathered from the
// adjacent cells' outputs in ascending cell order, with CR epilogues:
//   S_SETUP : s = A r;  r.s and s.D^-1 s  -> delta, alpha (beta = 0)
//   S_CR    : Ap = Ap beta + s, p = p beta + r, x += alpha p,
//             r -= alpha Ap / D, s = A r;  r.s, s.D^-1 s, Ap.D^-1 s,
//             Ap.D^-1 Ap -> delta, beta and the next alpha from
//             (s + beta Ap).D^-1 (s + beta Ap) expanded
//   S_CHECK : |A x - b|  (every 5 iterations, CONJUGATE_RESIDUAL.inl:246-255)
template <int MODE>
__global__ void __launch_bounds__(128) kSolidNode(Dims d, SolidNodeArgs a) {
  if (MODE != S_APPLY && a.sc->done) return;
  __shared__ double red[8];
  const int sn = blockIdx.x * blockDim.x + threadIdx.x;
  double p0 = 0, p1 = 0, p2 = 0, p3 = 0;
  if (sn < a.nSolid) {
    double kx[3] = {0, 0, 0};
    for (int b = 0; b < a.nSolidBodies; b++) {
      int slot[8];
#pragma unroll
      for (int c = 0; c < 8; c++) slot[c] = a.nodeSlot[(long)(8 * b + c) * a.capS + sn];
      const double* out = a.cellKx[b];
#pragma unroll
      for (int c = 0; c < 8; c++) {
        if (slot[c] < 0) continue;
        // the node is corner 7 - c of the cell at offset c
        const double* o = out + 24L * slot[c] + 3 * (7 - c);
        kx[0] += o[0];
        kx[1] += o[1];
        kx[2] += o[2];
      }
    }
    double alpha = 0, beta = 0;
    if (MODE == S_CR) {
      alpha = a.sc->alpha;
      beta = a.sc->beta;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const long k = 3L * sn + q;
      if (MODE == S_CR) {
        const double dg = a.diag[k];
        const double ap = a.Ap[k] * beta + a.s[k];
        const double pp = a.p[k] * beta + a.r[k];
        a.x[k] = a.x[k] + alpha * pp;
        const double rn = a.r[k] - alpha * (ap / dg);
        const double y = kx[q] * a.scale + a.damped[k] * rn;
        a.Ap[k] = ap;
        a.p[k] = pp;
        a.r[k] = rn;
        a.s[k] = y;
        const double yq = y / dg;
        p0 += rn * y;
        p1 += y * yq;
        p2 += ap * yq;
        p3 += ap * (ap / dg);
      } else {
        const double v = a.in[k];
        const double y = kx[q] * a.scale + a.damped[k] * v;
        if (MODE == S_APPLY) {
          a.out[k] = y;
        } else if (MODE == S_SETUP) {
          a.s[k] = y;
          p0 += v * y;
          p1 += y * (y / a.diag[k]);
        } else {
          const double t = y - a.b[k];
          p0 += t * t;
        }
      }
    }
  }
  if (MODE == S_APPLY) return;
  double mine[4] = {p0, p1, p2, p3};
  blockPartials<4>(mine, a.partials, red);
}

// diag(A) = scale * diag(K) + M~   (COUPLED_INTEGRATOR_3D.cpp:613-617)
__global__ void kSolidDiag(long n, const double* kdiag, const double* damped, double scale,
                           double* adiag) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) adiag[i] = kdiag[i] * scale + damped[i];
}


// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
esf_status stressStage(esf_ctx* ctx, double* energy) {
  const Dims& d = ctx->d;
  if (ctx->solids.size() > 4) {
    ctx->err = "at most 4 solids per context";
    return ESF_ERR_INVALID;
  }
  double en = 0;
  for (size_t b = 0; b < ctx->solids.size(); b++) {
    DevSolid& so = ctx->solids[b];
    const int nb = cdiv(std::max(so.nValid, 1), 256);
    if (so.nValid > 0)
      hipLaunchKernelGGL(kStress, dim3(nb), dim3(256), 0, ctx->stream, d, so.nValid, so.capValid,
                         so.cellList, so.u, so.mass, ctx->sidx, so.lambda, so.mu, so.cellMat,
                         so.cellOut, so.cellDiag, so.cornerDof, ctx->partials);
    if (energy && so.nValid > 0) {
      std::vector<double> h(nb);
      ESF_HIP(hipMemcpyAsync(h.data(), ctx->partials, nb * sizeof(double),
                             hipMemcpyDeviceToHost, ctx->stream));
      ESF_HIP(hipStreamSynchronize(ctx->stream));
      for (double v : h) en += v;
    }
    if (ctx->nSolid > 0)
      hipLaunchKernelGGL(kSolidGather, dim3(cdiv(ctx->nSolid, 256)), dim3(256), 0,
                         ctx->stream, d, ctx->nSolid, so.capValid, ctx->solidNodes, so.cellSlot,
                         so.cellOut,
                         so.cellDiag, so.mass, so.vel, so.box, b == 0 ? 1 : 0, ctx->sForce,
                         ctx->sKdiag, ctx->sMv, ctx->nodeSlot + 8L * b * ctx->capSolid,
                         ctx->capSolid);
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
  a.nSolidBodies = (int)std::min<size_t>(ctx->solids.size(), 4);
  for (int b = 0; b < a.nSolidBodies; b++) a.cellKx[b] = ctx->solids[b].cellKx;
  a.nodeSlot = ctx->nodeSlot;
  a.capS = ctx->capSolid;
  a.scale = ctx->dt * ctx->dt + ctx->p.rayleigh_beta * ctx->dt;
  a.damped = ctx->sDamped;
  a.diag = ctx->sAdiag;
}

// the CR-update launches are kernel classes 14 (cells) and 15 (nodes) of
// esf_kernel_timing
template <int MODE>
static void cellPass(esf_ctx* ctx, const double* in, const CrScalars* sc) {
  for (auto& so : ctx->solids)
    if (so.nValid > 0) {
      auto* tl = MODE == CELL_CR ? timedBegin(ctx, 14) : nullptr;
      hipLaunchKernelGGL((kSolidCell8<MODE>), dim3(cdiv(8L * so.nValid, 256)), dim3(256), 0,
                         ctx->stream, ctx->d, so.nValid, so.capValid, (const int*)so.cornerDof,
                         (const double*)so.cellMat, in, (const double*)ctx->ss,
                         (const double*)ctx->sAp, (const double*)ctx->sAdiag, sc, so.lambda,
                         so.mu, so.cellKx);
      timedEnd(ctx, tl);
    }
}

template <int MODE>
static void nodePass(esf_ctx* ctx, SolidNodeArgs& a) {
  auto* tl = MODE == S_CR ? timedBegin(ctx, 15) : nullptr;
  const int nb = cdiv(std::max(ctx->nSolid, 1), 128);
  hipLaunchKernelGGL((kSolidNode<MODE>), dim3(nb), dim3(128), 0, ctx->stream, ctx->d, a);
  timedEnd(ctx, tl);
  if constexpr (MODE == S_SETUP) launchFinish<F_SETUP_S, 2>(ctx, nb, a.sc);
  if constexpr (MODE == S_CR) launchFinish<F_SOLID_CR, 4>(ctx, nb, a.sc);
  if constexpr (MODE == S_CHECK) launchFinish<F_CHECK, 1>(ctx, nb, a.sc, a.iter);
}

esf_status solidAdiag(esf_ctx* ctx) {
  const long n = 3L * ctx->nSolid;
  if (n == 0) return ESF_OK;
  const double scale = ctx->dt * ctx->dt + ctx->p.rayleigh_beta * ctx->dt;
  hipLaunchKernelGGL(kSolidDiag, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, n, ctx->sKdiag,
                     ctx->sDamped, scale, ctx->sAdiag);
  return ESF_OK;
}

esf_status solidApplyDevice(esf_ctx* ctx, const double* x, double* y) {
  if (ctx->solids.size() > 4) {
    ctx->err = "at most 4 solids per context";
    return ESF_ERR_INVALID;
  }
  cellPass<CELL_IN>(ctx, x, nullptr);
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
__global__ void kSolidSetup(long n, const double* b, const double* diag, double* x, double* r,
                            double* p, double* Ap, double* partials) {
  __shared__ double red[8];
  double part = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    x[i] = 0.0;
    r[i] = b[i] / diag[i];
    p[i] = 0.0;   // with beta = 0 the first direction is p = r, Ap = s
    Ap[i] = 0.0;
    part += b[i] * b[i];
  }
  double mine[1] = {part};
  blockPartials<1>(mine, partials, red);
}

esf_status solidSolve(esf_ctx* ctx, int which, const double* b, double* x, double eps,
                      int maxIt, int* iters, bool* converged, double* relres) {
  if (ctx->solids.size() > 4) {
    ctx->err = "at most 4 solids per context";
    return ESF_ERR_INVALID;
  }
  const long n = 3L * ctx->nSolid;
  if (n == 0) {
    *iters = 0;
    *converged = true;
    *relres = 0;
    return ESF_OK;
  }
  CrScalars h0{};
  h0.eps = eps;
  h0.iter = -1;
  ESF_HIP(hipMemcpyAsync(&ctx->sc[which], &h0, sizeof(h0), hipMemcpyHostToDevice, ctx->stream));
  CrScalars* sc = &ctx->sc[which];
  const int eb = std::max(1, std::min(cdiv(n, 1024), 256));
  hipLaunchKernelGGL(kSolidSetup, dim3(eb), dim3(256), 0, ctx->stream, n, b,
                     (const double*)ctx->sAdiag, x, ctx->sr, ctx->sp, ctx->sAp, ctx->partials);
  launchFinish<F_BNORM, 1>(ctx, eb, sc);
  SolidNodeArgs a{};
  fillNodeArgs(ctx, a);
  a.b = b;
  a.x = x;
  a.r = ctx->sr;
  a.p = ctx->sp;
  a.s = ctx->ss;
  a.Ap = ctx->sAp;
  a.sc = sc;
  a.partials = ctx->partials;
  cellPass<CELL_IN>(ctx, ctx->sr, sc);
  a.in = ctx->sr;
  nodePass<S_SETUP>(ctx, a);
  DonePoller poll(ctx, which);
  for (int i = 0; i < maxIt; i++) {
    cellPass<CELL_CR>(ctx, ctx->sr, sc);
    nodePass<S_CR>(ctx, a);
    if (i > 0 && i % 5 == 0) {
      cellPass<CELL_IN>(ctx, x, sc);
      a.in = x;
      a.iter = i;
      nodePass<S_CHECK>(ctx, a);
      if (poll.mark()) break;
    }
  }
  ESF_HIP(hipMemcpyAsync(ctx->hostSc, sc, sizeof(CrScalars), hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(hipStreamSynchronize(ctx->stream));
  const CrScalars& h = *ctx->hostSc;
  *converged = h.done == 1;
  *iters = h.done ? h.iter : maxIt;
  *relres = h.bnorm > 0 ? h.resnorm / h.bnorm : h.resnorm;
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
