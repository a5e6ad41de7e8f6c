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
                                               double* cellDiag, double* energyPart) {
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
      const bool keep = mass[n] != 0 && sidx[n] >= 0;
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
                             Box box, int first, double* sForce, double* sKdiag, double* sMv) {
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

// K_e x_e per valid cell (see file comment)
__global__ void __launch_bounds__(256) kSolidCellApply(Dims d, int nValid, int cap,
                                                       const int* __restrict__ cellList,
                                                       const double* __restrict__ cellMat,
                                                       const double* __restrict__ smass,
                                                       const int* __restrict__ sidx,
                                                       const double* __restrict__ xs, double lambda,
                                                       double mu, double* cellOut) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nValid) return;
  const long index = cellList[e];
  const long off[8] = {0, 1, d.X, d.X + 1, d.slab, d.slab + 1, d.slab + d.X, d.slab + d.X + 1};
  const double* m = cellMat + e;
  double R[9], L[9], F[9];
  for (int k = 0; k < 9; k++) {
    R[k] = m[(long)k * cap];
    L[k] = m[(long)(9 + k) * cap];
    F[k] = m[(long)(18 + k) * cap];
  }
  const double Jinv = m[27L * cap];
  const double g = 0.25 * d.dhInv;
  double GF[24];
  for (int i = 0; i < 8; i++) {
    const double G0 = sgn(i & 1) * g, G1 = sgn((i >> 1) & 1) * g, G2 = sgn(i >> 2) * g;
    for (int c = 0; c < 3; c++) GF[3 * i + c] = (G0 * F[c] + G1 * F[3 + c]) + G2 * F[6 + c];
  }
  double dF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool keep[8];
  for (int i = 0; i < 8; i++) {
    const long n = index + off[i];
    const int si = sidx[n];
    keep[i] = smass[n] != 0 && si >= 0;
    if (!keep[i]) continue;
    for (int j = 0; j < 3; j++) {
      const double xv = xs[3 * (long)si + j];
      for (int c = 0; c < 3; c++) dF[3 * j + c] += xv * GF[3 * i + c];
    }
  }
  double dP[9];
  corotDP(R, L, lambda, mu, dF, dP);
  const double dh = d.dh;
  double* out = cellOut + e;
  for (int i = 0; i < 8; i++)
    for (int r = 0; r < 3; r++) {
      double t = 0;
      if (keep[i]) {
        t = (dP[3 * r] * GF[3 * i] + dP[3 * r + 1] * GF[3 * i + 1]) + dP[3 * r + 2] * GF[3 * i + 2];
        t = (((t * Jinv) * dh) * dh) * dh;
      }
      out[(long)(3 * i + r) * cap] = t;
    }
}

enum SolidMode { S_APPLY = 0, S_SETUP_S = 1, S_CHECK = 3 };

struct SolidNodeArgs {
  int nSolid;
  const int* solidNodes;
  int nSolidBodies;
  // up to 4 solids gathered in one launch
  const int* cellSlot[4];
  const double* cellOut[4];
  int cap[4];
  const double* smass[4];
  double scale;          // dt^2 + beta dt
  const double* damped;  // diag mass
  const double* in;      // x the apply acts on
  double* out;           // S_APPLY
  const double* b;
  const double* diag;
  double* s;
  double* Ap;
  CrScalars* sc;
  double* partials;
  unsigned* counter;
  int iter;
};

// y = scale * (K x) + M~ x per solid DOF node, with CR epilogues
template <int MODE>
__global__ void __launch_bounds__(256) kSolidNodeApply(Dims d, SolidNodeArgs a) {
  if (MODE != S_APPLY && a.sc->done) return;
  __shared__ double red[8];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double p0 = 0, p1 = 0;
  if (s < a.nSolid) {
    const long n = a.solidNodes[s];
    const int x = n % d.X, y = (n / d.X) % d.Y, z = n / d.slab;
    double kx[3] = {0, 0, 0};
    for (int b = 0; b < a.nSolidBodies; b++) {
      if (a.smass[b][n] == 0) continue;
      for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
          for (int dx = 0; dx < 2; dx++) {
            const long cell =
                (x - 1 + dx) + (long)(y - 1 + dy) * d.CX + (long)(z - 1 + dz) * d.cslab;
            const int slot = a.cellSlot[b][cell];
            if (slot < 0) continue;
            const int corner = (1 - dx) + 2 * (1 - dy) + 4 * (1 - dz);
            for (int r = 0; r < 3; r++)
              kx[r] += a.cellOut[b][(long)(3 * corner + r) * a.cap[b] + slot];
          }
    }
    for (int r = 0; r < 3; r++) {
      const long k = 3 * (long)s + r;
      const double xv = a.in[k];
      const double yv = kx[r] * a.scale + a.damped[k] * xv;
      if (MODE == S_APPLY) {
        a.out[k] = yv;
      } else if (MODE == S_SETUP_S) {
        a.s[k] = yv;
        a.Ap[k] = yv;
        p0 += xv * yv;
        const double q = yv / a.diag[k];
        p1 += yv * q;
      } else if (MODE == S_CHECK) {
        const double t = yv - a.b[k];
        p0 += t * t;
      }
    }
  }
  if (MODE == S_APPLY) return;
  double mine[2] = {p0, p1}, tot[2];
  if (lastBlockReduce<2>(mine, a.partials, a.counter, red, tot) && threadIdx.x == 0) {
    CrScalars* sc = a.sc;
    if (MODE == S_SETUP_S) {
      sc->delta = tot[0];
      sc->alpha = (fabs(tot[1]) > 0.0) ? tot[0] / tot[1] : tot[1];
    } else if (MODE == S_CHECK) {
      const double rn = sqrt(tot[0]);
      sc->resnorm = rn;
      if (rn < sc->eps * sc->bnorm) {
        sc->done = 1;
        sc->iter = a.iter;
      }
    }
  }
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
  double en = 0;
  for (size_t b = 0; b < ctx->solids.size(); b++) {
    DevSolid& so = ctx->solids[b];
    const int nb = cdiv(std::max(so.nValid, 1), 256);
    if (so.nValid > 0)
      hipLaunchKernelGGL(kStress, dim3(nb), dim3(256), 0, ctx->stream, d, so.nValid, so.capValid,
                         so.cellList, so.u, so.mass, ctx->sidx, so.lambda, so.mu, so.cellMat,
                         so.cellOut, so.cellDiag, ctx->partials);
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
                         ctx->sKdiag, ctx->sMv);
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
  a.solidNodes = ctx->solidNodes;
  a.nSolidBodies = (int)std::min<size_t>(ctx->solids.size(), 4);
  for (int b = 0; b < a.nSolidBodies; b++) {
    a.cellSlot[b] = ctx->solids[b].cellSlot;
    a.cellOut[b] = ctx->solids[b].cellOut;
    a.cap[b] = ctx->solids[b].capValid;
    a.smass[b] = ctx->solids[b].mass;
  }
  a.scale = ctx->dt * ctx->dt + ctx->p.rayleigh_beta * ctx->dt;
  a.damped = ctx->sDamped;
  a.diag = ctx->sAdiag;
}

static void cellApply(esf_ctx* ctx, const double* x) {
  for (auto& so : ctx->solids)
    if (so.nValid > 0)
      hipLaunchKernelGGL(kSolidCellApply, dim3(cdiv(so.nValid, 256)), dim3(256), 0,
                         ctx->stream, ctx->d, so.nValid, so.capValid, so.cellList, so.cellMat,
                         so.mass,
                         ctx->sidx, x, so.lambda, so.mu, so.cellOut);
}

template <int MODE>
static void nodeApply(esf_ctx* ctx, SolidNodeArgs& a) {
  hipLaunchKernelGGL(kSolidNodeApply<MODE>, dim3(cdiv(std::max(ctx->nSolid, 1), 256)),
                     dim3(256), 0, ctx->stream, ctx->d, a);
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
  cellApply(ctx, x);
  SolidNodeArgs a{};
  fillNodeArgs(ctx, a);
  a.in = x;
  a.out = y;
  nodeApply<S_APPLY>(ctx, a);
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// CONJUGATE_RESIDUAL::solvePCR(SpMat&, b, x)  (CONJUGATE_RESIDUAL.inl:194-276)
// with the matrix-free A; x starts at zero.  Three launches per iteration:
//   cells : K_e r_new per valid cell, r_new = r - alpha D^-1 Ap formed at the
//           corners (the update of .inl:223-230 folded into the apply)
//   nodes : x += alpha p, r = r_new, s = A r_new, partial r.s -> beta
//   dofs  : p = beta p + r, Ap = beta Ap + s, partial Ap.D^-1 Ap -> alpha
// plus the true-residual check every 5 iterations (.inl:246-255).
__global__ void kSolidSetup(long n, const double* b, const double* diag, double* x, double* r,
                            double* p, CrScalars* sc, double* partials, unsigned* counter) {
  __shared__ double red[8];
  double part = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const double rv = b[i] / diag[i];
    x[i] = 0.0;
    r[i] = rv;
    p[i] = rv;
    part += b[i] * b[i];
  }
  double mine[1] = {part}, tot[1];
  if (lastBlockReduce<1>(mine, partials, counter, red, tot) && threadIdx.x == 0)
    sc->bnorm = sqrt(tot[0]);
}

// cell phase of the CR update: v = r - alpha Ap / diag at unmasked corners
__global__ void __launch_bounds__(256) kSolidCellUpdate(Dims d, int nValid, int cap,
                                                        const int* __restrict__ cellList,
                                                        const double* __restrict__ cellMat,
                                                        const double* __restrict__ smass,
                                                        const int* __restrict__ sidx,
                                                        const double* __restrict__ r,
                                                        const double* __restrict__ Ap,
                                                        const double* __restrict__ diag,
                                                        const CrScalars* sc, double lambda,
                                                        double mu, double* cellOut) {
  if (sc->done) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nValid) return;
  const double alpha = sc->alpha;
  const long index = cellList[e];
  const long off[8] = {0, 1, d.X, d.X + 1, d.slab, d.slab + 1, d.slab + d.X, d.slab + d.X + 1};
  const double* m = cellMat + e;
  double R[9], L[9], F[9];
  for (int k = 0; k < 9; k++) {
    R[k] = m[(long)k * cap];
    L[k] = m[(long)(9 + k) * cap];
    F[k] = m[(long)(18 + k) * cap];
  }
  const double Jinv = m[27L * cap];
  const double g = 0.25 * d.dhInv, dh = d.dh;
  double GF[24];
  for (int i = 0; i < 8; i++) {
    const double G0 = sgn(i & 1) * g, G1 = sgn((i >> 1) & 1) * g, G2 = sgn(i >> 2) * g;
    for (int c = 0; c < 3; c++) GF[3 * i + c] = (G0 * F[c] + G1 * F[3 + c]) + G2 * F[6 + c];
  }
  double dF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  unsigned keep = 0;
  for (int i = 0; i < 8; i++) {
    const long n = index + off[i];
    const int si = sidx[n];
    if (!(smass[n] != 0 && si >= 0)) continue;
    keep |= 1u << i;
    for (int j = 0; j < 3; j++) {
      const long k = 3L * si + j;
      const double q = Ap[k] / diag[k];
      const double xv = r[k] - alpha * q;
      for (int c = 0; c < 3; c++) dF[3 * j + c] += xv * GF[3 * i + c];
    }
  }
  double dP[9];
  corotDP(R, L, lambda, mu, dF, dP);
  double* out = cellOut + e;
  for (int i = 0; i < 8; i++)
    for (int rr = 0; rr < 3; rr++) {
      double t = 0;
      if (keep & (1u << i)) {
        t = (dP[3 * rr] * GF[3 * i] + dP[3 * rr + 1] * GF[3 * i + 1]) + dP[3 * rr + 2] * GF[3 * i + 2];
        t = (((t * Jinv) * dh) * dh) * dh;
      }
      out[(long)(3 * i + rr) * cap] = t;
    }
}

// node phase of the CR update: x += alpha p, r = r_new, s = A r_new, r.s
__global__ void __launch_bounds__(256) kSolidNodeUpdate(Dims d, SolidNodeArgs a, double* x,
                                                        double* r, const double* p) {
  if (a.sc->done) return;
  __shared__ double red[8];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const double alpha = a.sc->alpha;
  double prs = 0;
  if (s < a.nSolid) {
    const long n = a.solidNodes[s];
    const int x0 = n % d.X, y0 = (n / d.X) % d.Y, z0 = n / d.slab;
    double kx[3] = {0, 0, 0};
    for (int b = 0; b < a.nSolidBodies; b++) {
      if (a.smass[b][n] == 0) continue;
      for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
          for (int dx = 0; dx < 2; dx++) {
            const long cell =
                (x0 - 1 + dx) + (long)(y0 - 1 + dy) * d.CX + (long)(z0 - 1 + dz) * d.cslab;
            const int slot = a.cellSlot[b][cell];
            if (slot < 0) continue;
            const int corner = (1 - dx) + 2 * (1 - dy) + 4 * (1 - dz);
            for (int c = 0; c < 3; c++)
              kx[c] += a.cellOut[b][(long)(3 * corner + c) * a.cap[b] + slot];
          }
    }
    for (int c = 0; c < 3; c++) {
      const long k = 3L * s + c;
      const double q = a.Ap[k] / a.diag[k];
      const double rn = r[k] - alpha * q;
      x[k] = x[k] + alpha * p[k];
      const double y = kx[c] * a.scale + a.damped[k] * rn;
      r[k] = rn;
      a.s[k] = y;
      prs += rn * y;
    }
  }
  double mine[1] = {prs}, tot[1];
  if (lastBlockReduce<1>(mine, a.partials, a.counter, red, tot) && threadIdx.x == 0) {
    CrScalars* sc = a.sc;
    const double dOld = sc->delta;
    sc->delta = tot[0];
    sc->beta = tot[0] / dOld;
  }
}

static void cellUpdate(esf_ctx* ctx, const CrScalars* sc) {
  for (auto& so : ctx->solids)
    if (so.nValid > 0)
      hipLaunchKernelGGL(kSolidCellUpdate, dim3(cdiv(so.nValid, 256)), dim3(256), 0, ctx->stream,
                         ctx->d, so.nValid, so.capValid, so.cellList, so.cellMat, so.mass,
                         ctx->sidx, (const double*)ctx->sr, (const double*)ctx->sAp,
                         (const double*)ctx->sAdiag, sc, so.lambda, so.mu, so.cellOut);
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
  unsigned* cnt = &ctx->counter[which];
  // few blocks for the DOF-parallel reductions: ~4 elements per thread
  const int eb = std::max(1, std::min(cdiv(n, 1024), 256));
  const int nb = cdiv(ctx->nSolid, 256);
  hipLaunchKernelGGL(kSolidSetup, dim3(eb), dim3(256), 0, ctx->stream, n, b,
                     (const double*)ctx->sAdiag, x, ctx->sr, ctx->sp, sc, ctx->partials, cnt);
  SolidNodeArgs a{};
  fillNodeArgs(ctx, a);
  a.b = b;
  a.s = ctx->ss;
  a.Ap = ctx->sAp;
  a.sc = sc;
  a.partials = ctx->partials;
  a.counter = cnt;
  cellApply(ctx, ctx->sr);
  a.in = ctx->sr;
  nodeApply<S_SETUP_S>(ctx, a);
  DonePoller poll(ctx, which);
  for (int i = 0; i < maxIt; i++) {
    cellUpdate(ctx, sc);
    hipLaunchKernelGGL(kSolidNodeUpdate, dim3(nb), dim3(256), 0, ctx->stream, ctx->d, a, x,
                       ctx->sr, (const double*)ctx->sp);
    hipLaunchKernelGGL(kCrDirection<double>, dim3(eb), dim3(256), 0, ctx->stream, n, ctx->sp,
                       ctx->sAp, (const double*)ctx->sr, (const double*)ctx->ss,
                       (const double*)ctx->sAdiag, sc, ctx->partials, cnt);
    if (i > 0 && i % 5 == 0) {
      cellApply(ctx, x);
      a.in = x;
      a.iter = i;
      nodeApply<S_CHECK>(ctx, a);
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
