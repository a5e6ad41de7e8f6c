// esf_corot.h -- corotational linear elasticity for one deformation gradient
// per GPU thread (src/material/COROTATION.cpp).  FP64, row-major 3x3.
//
// The polar factors come from the eigen-decomposition of F^T F, as in
// COROTATION::svd (COROTATION.cpp:235-280), which takes Eigen::JacobiSVD
// of F^T F: singular values of that SPD matrix are its eigenvalue
// magnitudes in decreasing order, V its eigenvectors.  Here the symmetric
// 3x3 is diagonalised by cyclic Jacobi rotations in registers.  R = U V^T,
// S = V Fhat V^T and L = V Ld V^T do not depend on the eigenvector signs,
// which is why a different eigensolver reproduces the reference's values.
#pragma once
#include <hip/hip_runtime.h>

namespace esf {

struct Mat3 {
  double a[9];
};

__device__ __forceinline__ double s3(double x, double y, double z) {
  // Eigen fixed-size 3-term reductions (trace, dot, squaredNorm) add
  // x + (y + z): redux_novec_unroller splits the range in halves.
  return x + (y + z);
}

// a * b, inner index summed in order (Eigen lazy product, fixed 3x3)
__device__ __forceinline__ void mul(const double* a, const double* b, double* r) {
  double t[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      t[3 * i + j] = (a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j]) + a[3 * i + 2] * b[6 + j];
#pragma unroll
  for (int k = 0; k < 9; k++) r[k] = t[k];
}
__device__ __forceinline__ void mulT(const double* a, const double* b, double* r) {
  double t[9];  // a * b^T
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      t[3 * i + j] =
          (a[3 * i] * b[3 * j] + a[3 * i + 1] * b[3 * j + 1]) + a[3 * i + 2] * b[3 * j + 2];
#pragma unroll
  for (int k = 0; k < 9; k++) r[k] = t[k];
}
__device__ __forceinline__ void Tmul(const double* a, const double* b, double* r) {
  double t[9];  // a^T * b
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      t[3 * i + j] = (a[i] * b[j] + a[3 + i] * b[3 + j]) + a[6 + i] * b[6 + j];
#pragma unroll
  for (int k = 0; k < 9; k++) r[k] = t[k];
}

// Eigen determinant_impl<3>
__device__ __forceinline__ double det(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[3] * (m[1] * m[8] - m[2] * m[7]) +
         m[6] * (m[1] * m[5] - m[2] * m[4]);
}
__device__ __forceinline__ double cofac(const double* m, int i, int j) {
  const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m[3 * i1 + j1] * m[3 * i2 + j2] - m[3 * i1 + j2] * m[3 * i2 + j1];
}
// Eigen compute_inverse<3>: cofactors, determinant from column 0
__device__ __forceinline__ void inverse(const double* m, double* r) {
  const double c0 = cofac(m, 0, 0), c1 = cofac(m, 1, 0), c2 = cofac(m, 2, 0);
  const double dt = (c0 * m[0] + c1 * m[3]) + c2 * m[6];
  const double inv = 1.0 / dt;
  double t[9];
  t[0] = c0 * inv;
  t[1] = c1 * inv;
  t[2] = c2 * inv;
  t[3] = cofac(m, 0, 1) * inv;
  t[4] = cofac(m, 1, 1) * inv;
  t[5] = cofac(m, 2, 1) * inv;
  t[6] = cofac(m, 0, 2) * inv;
  t[7] = cofac(m, 1, 2) * inv;
  t[8] = cofac(m, 2, 2) * inv;
#pragma unroll
  for (int k = 0; k < 9; k++) r[k] = t[k];
}

// symmetric 3x3 eigen-decomposition, cyclic Jacobi; values are magnitudes
// sorted decreasing, vectors in the columns of V
__device__ inline void symEig(const double* A, double ev[3], double* V) {
  double a00 = A[0], a01 = A[1], a02 = A[2], a11 = A[4], a12 = A[5], a22 = A[8];
  double v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 60; sweep++) {
    if (a01 == 0.0 && a02 == 0.0 && a12 == 0.0) break;
#pragma unroll
    for (int pq = 0; pq < 3; pq++) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      double* apq = pq == 0 ? &a01 : (pq == 1 ? &a02 : &a12);
      double* app = p == 0 ? &a00 : &a11;
      double* aqq = q == 1 ? &a11 : &a22;
      if (*apq == 0.0) continue;
      if (sweep > 3 && fabs(*app) + 100.0 * fabs(*apq) == fabs(*app) &&
          fabs(*aqq) + 100.0 * fabs(*apq) == fabs(*aqq)) {
        *apq = 0.0;
        continue;
      }
      const double theta = (*aqq - *app) / (2.0 * *apq);
      const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      // rotate the full symmetric matrix in the (p,q) plane
      double m[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
      for (int k = 0; k < 3; k++) {
        const double mkp = m[k][p], mkq = m[k][q];
        m[k][p] = c * mkp - s * mkq;
        m[k][q] = s * mkp + c * mkq;
      }
      for (int k = 0; k < 3; k++) {
        const double mpk = m[p][k], mqk = m[q][k];
        m[p][k] = c * mpk - s * mqk;
        m[q][k] = s * mpk + c * mqk;
      }
      m[p][q] = m[q][p] = 0.0;
      a00 = m[0][0];
      a11 = m[1][1];
      a22 = m[2][2];
      a01 = m[0][1];
      a02 = m[0][2];
      a12 = m[1][2];
      for (int k = 0; k < 3; k++) {
        const double vkp = v[3 * k + p], vkq = v[3 * k + q];
        v[3 * k + p] = c * vkp - s * vkq;
        v[3 * k + q] = s * vkp + c * vkq;
      }
    }
  }
  double e[3] = {fabs(a00), fabs(a11), fabs(a22)};
  int o[3] = {0, 1, 2};
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2 - i; j++)
      if (e[o[j]] < e[o[j + 1]]) {
        int tmp = o[j];
        o[j] = o[j + 1];
        o[j + 1] = tmp;
      }
  for (int cI = 0; cI < 3; cI++) {
    ev[cI] = e[o[cI]];
    for (int r = 0; r < 3; r++) V[3 * r + cI] = v[3 * r + o[cI]];
  }
}

__device__ __forceinline__ void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// COROTATION::orthogonalizeU  COROTATION.cpp:127-230
__device__ inline void orthogonalizeU(double* U, const double* Fh) {
  const double signU = (det(U) >= 0.0) ? 1.0 : -1.0;
  bool bad[3];
  int nbad = 0;
  for (int x = 0; x < 3; x++) {
    bad[x] = Fh[x] < 1e-4;
    nbad += bad[x];
  }
  if (nbad == 0) return;
  if (nbad == 1) {
    const int w = bad[1] ? 1 : (bad[2] ? 2 : 0);
    double c0[3], c1[3], c2[3];
    for (int r = 0; r < 3; r++) {
      c0[r] = U[3 * r + (w + 1) % 3];
      c1[r] = U[3 * r + (w + 2) % 3];
    }
    cross(c0, c1, c2);
    const double n = sqrt(s3(c2[0] * c2[0], c2[1] * c2[1], c2[2] * c2[2]));
    for (int r = 0; r < 3; r++) c2[r] /= n;
    for (int r = 0; r < 3; r++) U[3 * r + w] = c2[r];
    if (signU * det(U) < 0.0)
      for (int r = 0; r < 3; r++) U[3 * r + w] = -c2[r];
    return;
  }
  if (nbad == 2) {
    const int wv = !bad[1] ? 1 : (!bad[2] ? 2 : 0);
    const double vv[3] = {U[wv], U[3 + wv], U[6 + wv]};
    int sm = (fabs(vv[0]) < fabs(vv[1])) ? 0 : 1;
    sm = (fabs(vv[2]) < fabs(vv[sm])) ? 2 : sm;
    double sec[3], thr[3];
    int nx = (sm + 1) % 3, nn = (sm + 2) % 3;
    sec[sm] = 0.0;
    sec[nx] = -vv[nn];
    sec[nn] = vv[nx];
    cross(vv, sec, thr);
    const double n = sqrt(s3(thr[0] * thr[0], thr[1] * thr[1], thr[2] * thr[2]));
    for (int r = 0; r < 3; r++) thr[r] /= n;
    nx = (wv + 1) % 3;
    nn = (wv + 2) % 3;
    for (int r = 0; r < 3; r++) {
      U[3 * r + nx] = sec[r];
      U[3 * r + nn] = thr[r];
    }
    if (signU * det(U) < 0.0) {
      if (Fh[nx] < Fh[nn])
        for (int r = 0; r < 3; r++) U[3 * r + nx] = -sec[r];
      else
        for (int r = 0; r < 3; r++) U[3 * r + nn] = -thr[r];
    }
    return;
  }
  for (int k = 0; k < 9; k++) U[k] = (k % 4 == 0) ? 1.0 : 0.0;
}

// COROTATION::init -> decomposeF/svd (COROTATION.cpp:37-39, 91-107, 235-280)
// outputs R, S, L
__device__ inline void corotDecompose(const double* F, double lambda, double mu, double* R,
                                      double* S, double* L) {
  double FtF[9], V[9], ev[3], Fh[3];
  Tmul(F, F, FtF);
  symEig(FtF, ev, V);
  for (int i = 0; i < 3; i++) {
    Fh[i] = sqrt(ev[i]);
    if (isnan(Fh[i])) Fh[i] = 0.0;
  }
  if (det(V) < 0.0)
    for (int r = 0; r < 3; r++) V[3 * r] *= -1.0;
  double Dg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 3; i++) Dg[4 * i] = (Fh[i] > 0.0) ? 1.0 / Fh[i] : 0.0;
  double FV[9], U[9];
  mul(F, V, FV);
  mul(FV, Dg, U);
  orthogonalizeU(U, Fh);
  if (det(F) < 0.0) {  // removeUReflection  COROTATION.cpp:112-122
    int sm = (Fh[0] < Fh[1]) ? 0 : 1;
    sm = (Fh[sm] < Fh[2]) ? sm : 2;
    Fh[sm] *= -1.0;
    for (int r = 0; r < 3; r++) U[3 * r + sm] *= -1.0;
  }
  mulT(U, V, R);
  double Fd[9] = {Fh[0], 0, 0, 0, Fh[1], 0, 0, 0, Fh[2]}, VF[9];
  mul(V, Fd, VF);
  mulT(VF, V, S);
  const double trF = s3(Fh[0], Fh[1], Fh[2]);
  const double trE = s3(Fh[0] - 1.0, Fh[1] - 1.0, Fh[2] - 1.0);
  const double sc = lambda * trE - 2 * mu;
  const double Mt[9] = {trF - Fh[0], 0, 0, 0, trF - Fh[1], 0, 0, 0, trF - Fh[2]};
  double Mi[9], Ld[9], VL[9];
  inverse(Mt, Mi);
  for (int k = 0; k < 9; k++) Ld[k] = mu * ((k % 4 == 0) ? 1.0 : 0.0) + sc * Mi[k];
  for (int i = 0; i < 3; i++) Ld[4 * i] = Ld[4 * i] > 0 ? Ld[4 * i] : 0;
  mul(V, Ld, VL);
  mulT(VL, V, L);
}

// COROTATION::firstPiolaKirchhoff  COROTATION.cpp:53-62
__device__ inline void corotP(const double* R, const double* S, double lambda, double mu,
                              double* P) {
  double E[9], t[9];
  for (int k = 0; k < 9; k++) E[k] = S[k] - ((k % 4 == 0) ? 1.0 : 0.0);
  for (int k = 0; k < 9; k++) t[k] = 2.0 * mu * E[k];
  const double tr = lambda * s3(E[0], E[4], E[8]);
  t[0] += tr;
  t[4] += tr;
  t[8] += tr;
  mul(R, t, P);
}

// COROTATION::strainEnergy  COROTATION.cpp:42-49
__device__ inline double corotEnergy(const double* S, double lambda, double mu) {
  double E[9], EtE[9];
  for (int k = 0; k < 9; k++) E[k] = S[k] - ((k % 4 == 0) ? 1.0 : 0.0);
  Tmul(E, E, EtE);
  const double tr = s3(E[0], E[4], E[8]);
  return mu * s3(EtE[0], EtE[4], EtE[8]) + 0.5 * lambda * (tr * tr);
}

// COROTATION::firstPKDifferential  COROTATION.cpp:64-84; the lambda trace
// term enters twice (:69-70 and :75-78) exactly as in the reference.
__device__ inline void corotDP(const double* R, const double* L, double lambda, double mu,
                               const double* dF, double* dP) {
  double h[9];
  Tmul(R, dF, h);
  double sym[9], skw[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      sym[3 * i + j] = 0.5 * (h[3 * i + j] + h[3 * j + i]);
      skw[3 * i + j] = 0.5 * (h[3 * i + j] - h[3 * j + i]);
    }
  const double trs = s3(sym[0], sym[4], sym[8]);
  double d[9];
#pragma unroll
  for (int k = 0; k < 9; k++)
    d[k] = (2 * mu * sym[k]) + (lambda * trs * ((k % 4 == 0) ? 1.0 : 0.0));
  const double f0 = -skw[5] + skw[7], f1 = -skw[6] + skw[2], f2 = -skw[1] + skw[3];
  const double tr = lambda * s3(h[0], h[4], h[8]);
  d[0] += tr;
  d[4] += tr;
  d[8] += tr;
  double Lf[3];
#pragma unroll
  for (int r = 0; r < 3; r++) Lf[r] = (L[3 * r] * f0 + L[3 * r + 1] * f1) + L[3 * r + 2] * f2;
  // + MATRIX_UTIL::cross(L f)  MATRIX_UTIL.h:26-41
  d[0] += 0.0;
  d[1] += -Lf[2];
  d[2] += Lf[1];
  d[3] += Lf[2];
  d[4] += 0.0;
  d[5] += -Lf[0];
  d[6] += -Lf[1];
  d[7] += Lf[0];
  d[8] += 0.0;
  mul(R, d, dP);
}

}  // namespace esf
