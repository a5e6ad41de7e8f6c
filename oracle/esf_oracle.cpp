/*
 * esf_oracle.cpp -- CPU restatement of the reference's coupled substep.
 *
 * TEST INFRASTRUCTURE ONLY (see esf_oracle.h).  Every function names the
 * reference lines it restates.  Arithmetic follows the reference's
 * evaluation order wherever that order is fixed by the source (sparse
 * row/column order, left-to-right scalar chains); vector inner products
 * are plain sums, so the Krylov scalars agree with the reference only to
 * rounding.  Compiled with -ffp-contract=off so no FMA is formed.
 *
 * Functions whose parity is pinned by the compiled reference (oracle/_ref):
 *   corotation (polar / P / dP / energy), CR solvers, trilinear
 *   interpolation, cfg parsing.
 * Functions pinned by closed forms and invariants in tests/:
 *   J, J^T, S = J M^-1 J^T, its diagonal, the divergence RHS, body forces
 *   and vorticity confinement (linear and rotational fields), advection
 *   (linear fields, clamp rule), mass gathering (hand-built fractions).
 */
#include "esf_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <random>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using std::vector;

// ---------------------------------------------------------------------------
// Trilinear interpolation: FIELD_3D.h:198-216, get_barycentric :368-382,
// VEC3F_FIELD_3D.h:179-197, trilerp MATH_UTILS.h:326-342.
// ---------------------------------------------------------------------------
struct Geom {
  int X, Y, Z;
  double dh;
  double o[3];
};

inline void bary(double x, double dh, int hi, int& i, double& f) {
  x /= dh;
  double s = std::floor(x);
  i = int(s);
  if (i < 0) {
    i = 0;
    f = 0.5;
  } else if (i > hi) {
    i = hi;
    f = 0.5;
  } else {
    f = x - s;
  }
}
inline double lerp(double a, double b, double f) { return (1 - f) * a + f * b; }

inline double interpS(const double* d, const Geom& g, const double p[3]) {
  int i, j, k;
  double fx, fy, fz;
  bary(p[0] - g.o[0], g.dh, g.X - 2, i, fx);
  bary(p[1] - g.o[1], g.dh, g.Y - 2, j, fy);
  bary(p[2] - g.o[2], g.dh, g.Z - 2, k, fz);
  const long X = g.X, S = (long)g.X * g.Y;
  const long b = i + j * X + k * S;
  double v000 = d[b], v100 = d[b + 1], v010 = d[b + X], v110 = d[b + X + 1];
  double v001 = d[b + S], v101 = d[b + S + 1], v011 = d[b + S + X],
         v111 = d[b + S + X + 1];
  double z0 = lerp(lerp(v000, v100, fx), lerp(v010, v110, fx), fy);
  double z1 = lerp(lerp(v001, v101, fx), lerp(v011, v111, fx), fy);
  return lerp(z0, z1, fz);
}

inline void interpV(const double* d, const Geom& g, const double p[3],
                    double out[3]) {
  int i, j, k;
  double fx, fy, fz;
  bary(p[0] - g.o[0], g.dh, g.X - 2, i, fx);
  bary(p[1] - g.o[1], g.dh, g.Y - 2, j, fy);
  bary(p[2] - g.o[2], g.dh, g.Z - 2, k, fz);
  const long X = g.X, S = (long)g.X * g.Y;
  const long b = i + j * X + k * S;
  const long n[8] = {b, b + 1, b + X, b + X + 1, b + S, b + S + 1, b + S + X,
                     b + S + X + 1};
  for (int c = 0; c < 3; c++) {
    double v000 = d[3 * n[0] + c], v100 = d[3 * n[1] + c];
    double v010 = d[3 * n[2] + c], v110 = d[3 * n[3] + c];
    double v001 = d[3 * n[4] + c], v101 = d[3 * n[5] + c];
    double v011 = d[3 * n[6] + c], v111 = d[3 * n[7] + c];
    double z0 = lerp(lerp(v000, v100, fx), lerp(v010, v110, fx), fy);
    double z1 = lerp(lerp(v001, v101, fx), lerp(v011, v111, fx), fy);
    out[c] = lerp(z0, z1, fz);
  }
}

// ---------------------------------------------------------------------------
// 3x3 helpers (row-major m[3*r+c]).  Products accumulate k = 0,1,2 in order.
// ---------------------------------------------------------------------------
inline void mm(const double* a, const double* b, double* r) {
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[3 * i + j] = (a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j]) +
                     a[3 * i + 2] * b[6 + j];
  std::memcpy(r, t, sizeof(t));
}
inline void mmT(const double* a, const double* b, double* r) {  // a * b^T
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[3 * i + j] = (a[3 * i] * b[3 * j] + a[3 * i + 1] * b[3 * j + 1]) +
                     a[3 * i + 2] * b[3 * j + 2];
  std::memcpy(r, t, sizeof(t));
}
inline void mTm(const double* a, const double* b, double* r) {  // a^T * b
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[3 * i + j] = (a[i] * b[j] + a[3 + i] * b[3 + j]) + a[6 + i] * b[6 + j];
  std::memcpy(r, t, sizeof(t));
}
// Eigen determinant_impl<3> (bruteforce_det3_helper order)
inline double det3(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) -
         m[3] * (m[1] * m[8] - m[2] * m[7]) +
         m[6] * (m[1] * m[5] - m[2] * m[4]);
}
// Eigen compute_inverse<3>: cofactors, det from column 0, result = cof*invdet
inline double cof(const double* m, int i, int j) {
  int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m[3 * i1 + j1] * m[3 * i2 + j2] - m[3 * i1 + j2] * m[3 * i2 + j1];
}
inline void inv3(const double* m, double* r) {
  double c0 = cof(m, 0, 0), c1 = cof(m, 1, 0), c2 = cof(m, 2, 0);
  double det = (c0 * m[0] + c1 * m[3]) + c2 * m[6];
  double invdet = 1.0 / det;
  double t[9];
  t[0] = c0 * invdet;
  t[1] = c1 * invdet;
  t[2] = c2 * invdet;
  t[3] = cof(m, 0, 1) * invdet;
  t[4] = cof(m, 1, 1) * invdet;
  t[5] = cof(m, 2, 1) * invdet;
  t[6] = cof(m, 0, 2) * invdet;
  t[7] = cof(m, 1, 2) * invdet;
  t[8] = cof(m, 2, 2) * invdet;
  std::memcpy(r, t, sizeof(t));
}

// ---------------------------------------------------------------------------
// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations.  The
// reference obtains the same quantities from Eigen::JacobiSVD of F^T F
// (COROTATION.cpp:237-241): singular values of an SPD matrix are its
// eigenvalues (absolute values), sorted in decreasing order, and V holds
// the eigenvectors as columns.
// ---------------------------------------------------------------------------
void symEig3(const double* Ain, double ev[3], double V[9]) {
  double a[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = Ain[3 * i + j];
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
    if (off == 0.0) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        double apq = a[p][q];
        if (apq == 0.0) continue;
        double app = a[p][p], aqq = a[q][q];
        // annihilate entries that are below rounding of both diagonals
        if (sweep > 3 && std::fabs(app) + 100.0 * std::fabs(apq) ==
                             std::fabs(app) &&
            std::fabs(aqq) + 100.0 * std::fabs(apq) == std::fabs(aqq)) {
          a[p][q] = a[q][p] = 0.0;
          continue;
        }
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) /
                   (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; k++) {  // A <- A J
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; k++) {  // A <- J^T A
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        a[p][q] = a[q][p] = 0.0;
        for (int k = 0; k < 3; k++) {
          double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
  }
  double e[3] = {std::fabs(a[0][0]), std::fabs(a[1][1]), std::fabs(a[2][2])};
  int order[3] = {0, 1, 2};
  // stable descending sort of three values
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2 - i; j++)
      if (e[order[j]] < e[order[j + 1]]) std::swap(order[j], order[j + 1]);
  for (int c = 0; c < 3; c++) {
    ev[c] = e[order[c]];
    for (int r = 0; r < 3; r++) V[3 * r + c] = v[r][order[c]];
  }
}

// COROTATION state for one deformation gradient (COROTATION.h:54)
struct Corot {
  double lambda, mu;
  double U[9], Fhat[3], V[9], R[9], S[9], L[9];
};

// COROTATION::removeUReflection  COROTATION.cpp:112-122
void removeUReflection(double* U, double* Fh) {
  int smallest = (Fh[0] < Fh[1]) ? 0 : 1;
  smallest = (Fh[smallest] < Fh[2]) ? smallest : 2;
  Fh[smallest] *= -1.0;
  for (int r = 0; r < 3; r++) U[3 * r + smallest] *= -1.0;
}

inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
// Eigen fixed-size 3-vector reductions (dot, squaredNorm, trace) add
// e0 + (e1 + e2): redux_novec_unroller splits the range in halves.
inline double sum3(double a, double b, double c) { return a + (b + c); }
// VEC3F::normalize: *this /= norm() with no zero test
inline void normalize3(double* a) {
  double n = std::sqrt(sum3(a[0] * a[0], a[1] * a[1], a[2] * a[2]));
  a[0] /= n;
  a[1] /= n;
  a[2] /= n;
}

// COROTATION::orthogonalizeU  COROTATION.cpp:127-230
void orthogonalizeU(double* U, const double* Fh) {
  double signU = (det3(U) >= 0.0) ? 1.0 : -1.0;
  const double smallEigenvalue = 1e-4;
  bool invalid[3] = {false, false, false};
  int totalInvalid = 0;
  for (int x = 0; x < 3; x++)
    if (Fh[x] < smallEigenvalue) {
      invalid[x] = true;
      totalInvalid++;
    }
  if (totalInvalid == 0) return;
  if (totalInvalid == 1) {
    int w = 0;
    if (invalid[1]) w = 1;
    if (invalid[2]) w = 2;
    double c0[3], c1[3], c2[3];
    for (int r = 0; r < 3; r++) {
      c0[r] = U[3 * r + (w + 1) % 3];
      c1[r] = U[3 * r + (w + 2) % 3];
    }
    cross3(c0, c1, c2);
    normalize3(c2);
    for (int r = 0; r < 3; r++) U[3 * r + w] = c2[r];
    if (signU * det3(U) < 0.0)
      for (int r = 0; r < 3; r++) U[3 * r + w] = -c2[r];
    return;
  } else if (totalInvalid == 2) {
    int wv = 0;
    if (!invalid[1]) wv = 1;
    if (!invalid[2]) wv = 2;
    double vv[3] = {U[wv], U[3 + wv], U[6 + wv]};
    int smallest = (std::fabs(vv[0]) < std::fabs(vv[1])) ? 0 : 1;
    smallest = (std::fabs(vv[2]) < std::fabs(vv[smallest])) ? 2 : smallest;
    double second[3], third[3];
    int next = (smallest + 1) % 3, nextnext = (smallest + 2) % 3;
    second[smallest] = 0.0;
    second[next] = -vv[nextnext];
    second[nextnext] = vv[next];
    cross3(vv, second, third);
    normalize3(third);
    next = (wv + 1) % 3;
    nextnext = (wv + 2) % 3;
    for (int r = 0; r < 3; r++) {
      U[3 * r + next] = second[r];
      U[3 * r + nextnext] = third[r];
    }
    if (signU * det3(U) < 0.0) {
      if (Fh[next] < Fh[nextnext])
        for (int r = 0; r < 3; r++) U[3 * r + next] = -second[r];
      else
        for (int r = 0; r < 3; r++) U[3 * r + nextnext] = -third[r];
    }
    return;
  }
  for (int i = 0; i < 9; i++) U[i] = (i % 4 == 0) ? 1.0 : 0.0;
}

// COROTATION::svd  COROTATION.cpp:235-280
void corotSvd(const double* F, double* U, double* Fh, double* V) {
  double FtF[9];
  mTm(F, F, FtF);
  double ev[3];
  symEig3(FtF, ev, V);
  for (int i = 0; i < 3; i++) {
    Fh[i] = std::sqrt(ev[i]);
    if (std::isnan(Fh[i])) Fh[i] = 0.0;
  }
  if (det3(V) < 0.0)
    for (int r = 0; r < 3; r++) V[3 * r] *= -1.0;
  double D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 3; i++) D[4 * i] = (Fh[i] > 0.0) ? 1.0 / Fh[i] : 0.0;
  double FV[9];
  mm(F, V, FV);
  mm(FV, D, U);
  orthogonalizeU(U, Fh);
  if (det3(F) < 0.0) removeUReflection(U, Fh);
}

// COROTATION::init -> decomposeF  COROTATION.cpp:37-39, 91-107
void corotInit(Corot& m, const double* F) {
  corotSvd(F, m.U, m.Fhat, m.V);
  mmT(m.U, m.V, m.R);
  double FhV[9] = {0};
  double Dg[9] = {m.Fhat[0], 0, 0, 0, m.Fhat[1], 0, 0, 0, m.Fhat[2]};
  mm(m.V, Dg, FhV);
  mmT(FhV, m.V, m.S);
  // Ld = mu I + (lambda tr(Fhat - I) - 2 mu) (tr(Fhat) I - Fhat)^-1
  double trF = sum3(m.Fhat[0], m.Fhat[1], m.Fhat[2]);
  double trE = sum3(m.Fhat[0] - 1.0, m.Fhat[1] - 1.0, m.Fhat[2] - 1.0);
  double sc = m.lambda * trE - 2 * m.mu;
  double Mt[9] = {trF - m.Fhat[0], 0, 0, 0, trF - m.Fhat[1], 0,
                  0,               0, trF - m.Fhat[2]};
  double Mi[9];
  inv3(Mt, Mi);
  double Ld[9];
  for (int i = 0; i < 9; i++) Ld[i] = m.mu * ((i % 4 == 0) ? 1.0 : 0.0) + sc * Mi[i];
  for (int i = 0; i < 3; i++) Ld[4 * i] = Ld[4 * i] > 0 ? Ld[4 * i] : 0;
  double VL[9];
  mm(m.V, Ld, VL);
  mmT(VL, m.V, m.L);
}

// COROTATION::firstPiolaKirchhoff  COROTATION.cpp:53-62
void corotP(const Corot& m, double* P) {
  double E[9], tmp[9];
  for (int i = 0; i < 9; i++) E[i] = m.S[i] - ((i % 4 == 0) ? 1.0 : 0.0);
  for (int i = 0; i < 9; i++) tmp[i] = 2.0 * m.mu * E[i];
  double tr = m.lambda * sum3(E[0], E[4], E[8]);
  tmp[0] += tr;
  tmp[4] += tr;
  tmp[8] += tr;
  mm(m.R, tmp, P);
}

// COROTATION::strainEnergy  COROTATION.cpp:42-49
double corotEnergy(const Corot& m) {
  double E[9], EtE[9];
  for (int i = 0; i < 9; i++) E[i] = m.S[i] - ((i % 4 == 0) ? 1.0 : 0.0);
  mTm(E, E, EtE);
  double trE = sum3(E[0], E[4], E[8]);
  return m.mu * sum3(EtE[0], EtE[4], EtE[8]) + 0.5 * m.lambda * (trE * trE);
}

// COROTATION::firstPKDifferential  COROTATION.cpp:64-84 (lambda trace term
// applied twice, :69-70 and :75-78)
void corotDP(const Corot& m, const double* dF, double* dP) {
  double dFh[9];
  mTm(m.R, dF, dFh);
  double sym[9], skew[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      sym[3 * i + j] = 0.5 * (dFh[3 * i + j] + dFh[3 * j + i]);
      skew[3 * i + j] = 0.5 * (dFh[3 * i + j] - dFh[3 * j + i]);
    }
  double trs = sum3(sym[0], sym[4], sym[8]);
  double dPs[9];
  for (int i = 0; i < 9; i++)
    dPs[i] = (2 * m.mu * sym[i]) + (m.lambda * trs * ((i % 4 == 0) ? 1.0 : 0.0));
  double f[3] = {-skew[5] + skew[7], -skew[6] + skew[2], -skew[1] + skew[3]};
  double tr = m.lambda * sum3(dFh[0], dFh[4], dFh[8]);
  dPs[0] += tr;
  dPs[4] += tr;
  dPs[8] += tr;
  double Lf[3];
  for (int r = 0; r < 3; r++)
    Lf[r] = (m.L[3 * r] * f[0] + m.L[3 * r + 1] * f[1]) + m.L[3 * r + 2] * f[2];
  // MATRIX_UTIL::cross  MATRIX_UTIL.h:26-41
  double cx[9] = {0.0, -Lf[2], Lf[1], Lf[2], 0.0, -Lf[0], -Lf[1], Lf[0], 0.0};
  double d[9];
  for (int i = 0; i < 9; i++) d[i] = dPs[i] + cx[i];
  mm(m.R, d, dP);
}

// ---------------------------------------------------------------------------
// Scene state
// ---------------------------------------------------------------------------
struct CSR {
  int n = 0;
  vector<int> rowptr, col;
  vector<double> val;
};

struct Solid {
  double lambda, mu, qcm;  // qcm = _quarterCellSolidMass (DEFOR_SOLID_3D.h:64-66)
  Geom rg;
  vector<double> rest;
  vector<double> X, u, vel, velws;  // 3*nodes
  vector<double> mass, detFInv;     // nodes
  vector<double> phi;               // nodes (_phi, computeSDF)
  // FLIP material particles (PARTICLE, DEFOR_SOLID_3D.h): position,
  // velocity (3 each) and mass
  vector<double> ppos, pvel, pmass;
  // a deformation gradient per valid cell (valid-cell order) replacing the
  // displacement's in the stress stage (the isolated system's solids)
  vector<double> Fset;
  vector<int8_t> valid;             // cells
  vector<double> qfrac;             // 8*nodes
  int tMin[3], tMax[3], bMin[3], bMax[3];
  double dt;
};

struct Scene {
  int X, Y, Z;
  long slab, total, totalActive, nCells;
  double dh, dhInv, o[3];
  orc_params p;
  double vortEps, defaultDt, dt;
  long off[8];
  vector<double> vel, velws, vort;  // 3*nodes
  vector<double> intensity, ws, solidMass, nodeFrac;
  vector<double> qfracSum;
  vector<int> sol, sidx;
  int nSolid = 0;
  vector<double> unitForce, mv, nzMass, minv;  // 3*active
  vector<double> sForce, sMv, sDamped;         // 3*nSolid
  vector<double> sForceRhs;                    // dt*F + Mv
  vector<double> unitRhs;                      // dt*f + mv
  CSR A;
  vector<double> Adiag;
  vector<double> c, b, lambda, D;  // cells
  bool lambdaInit = false;
  vector<double> v0, v1, g;        // 3*active
  vector<double> sV0, sDv, sB;     // 3*nSolid
  vector<Solid> solids;
  double stageSec[16] = {0};
  Geom geom() const { return Geom{X, Y, Z, dh, {o[0], o[1], o[2]}}; }
};

inline int sgnX(int i) { return i % 2 == 0 ? -1 : 1; }
inline int sgnY(int i) { return i % 4 <= 1 ? -1 : 1; }
inline int sgnZ(int i) { return i < 4 ? -1 : 1; }

double nowSec() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// ---------------------------------------------------------------------------
// parallel vector helpers (per-thread partial sums combined in thread order)
// ---------------------------------------------------------------------------
double vdot(const vector<double>& a, const vector<double>& b) {
  long n = (long)a.size();
  double total = 0;
#ifdef _OPENMP
  int nt = omp_get_max_threads();
  vector<double> part(nt, 0.0);
#pragma omp parallel
  {
    int t = omp_get_thread_num();
    double s = 0;
#pragma omp for schedule(static)
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    part[t] = s;
  }
  for (int t = 0; t < nt; t++) total += part[t];
#else
  for (long i = 0; i < n; i++) total += a[i] * b[i];
#endif
  return total;
}
double vnorm(const vector<double>& a) { return std::sqrt(vdot(a, a)); }

// ---------------------------------------------------------------------------
// Interpolation of scene fields
// ---------------------------------------------------------------------------

// DEFOR_SOLID_3D::computeF  DEFOR_SOLID_3D.cpp:671-682 with G from :304-310.
// Returns det F; F receives inv(F^-1) or I when |det F^-1| <= 1e-4.
double computeF(const Scene& s, const Solid& so, long index, double* F) {
  const double g = 0.25 * s.dhInv;
  double Finv[9];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      double acc = 0;
      for (int i = 0; i < 8; i++) {
        int sg = c == 0 ? sgnX(i) : (c == 1 ? sgnY(i) : sgnZ(i));
        double G = sg * g;
        double t = so.u[3 * (index + s.off[i]) + r] * G;
        acc = (i == 0) ? t : acc + t;
      }
      Finv[3 * r + c] = ((r == c) ? 1.0 : 0.0) - acc;
    }
  double Jinv = det3(Finv);
  if (std::fabs(Jinv) > 1e-4) {
    inv3(Finv, F);
  } else {
    for (int i = 0; i < 9; i++) F[i] = (i % 4 == 0) ? 1.0 : 0.0;
    Jinv = 1;
  }
  return 1.0 / Jinv;
}

// DEFOR_SOLID_3D::computeSolidMass(usesdf)  DEFOR_SOLID_3D.cpp:576-669
void computeSolidMass(Scene& s, Solid& so, bool usesdf) {
  std::fill(so.mass.begin(), so.mass.end(), 0.0);
  std::fill(so.valid.begin(), so.valid.end(), 0);
  std::fill(so.detFInv.begin(), so.detFInv.end(), 0.0);
  std::fill(so.qfrac.begin(), so.qfrac.end(), 0.0);
  const double sdh = s.p.mass_sample_dh;
  const double subV = sdh * sdh * sdh;
  const int msr = (int)(s.dh / sdh);
  const int half = msr >> 1;
  const double fraction = 1.0 / (msr * msr * msr);
  const Geom g = s.geom();
  const long cx = s.X - 1, cxy = (long)(s.X - 1) * (s.Y - 1);
#pragma omp parallel for schedule(static)
  for (int z = so.tMin[2]; z < so.tMax[2] - 1; z++)
    for (int y = so.tMin[1]; y < so.tMax[1] - 1; y++)
      for (int x = so.tMin[0]; x < so.tMax[0] - 1; x++) {
        long index = z * s.slab + (long)y * s.X + x;
        double F[9];
        so.detFInv[index] = std::fabs(1.0 / computeF(s, so, index, F));
      }
#pragma omp parallel for schedule(static)
  for (int z = so.tMin[2]; z < so.tMax[2] - 1; z++)
    for (int y = so.tMin[1]; y < so.tMax[1] - 1; y++)
      for (int x = so.tMin[0]; x < so.tMax[0] - 1; x++) {
        long index = z * s.slab + (long)y * s.X + x;
        double node[3] = {x * s.dh + s.o[0], y * s.dh + s.o[1],
                          z * s.dh + s.o[2]};
        if (usesdf && so.phi[index] > 2) continue;  // :620
        double* q = &so.qfrac[8 * index];
        for (int pp = 0; pp < msr; pp++)
          for (int qq = 0; qq < msr; qq++)
            for (int rr = 0; rr < msr; rr++) {
              double pos[3] = {node[0] + (pp + 0.5) * sdh,
                               node[1] + (qq + 0.5) * sdh,
                               node[2] + (rr + 0.5) * sdh};
              double rPos[3];
              interpV(so.X.data(), g, pos, rPos);
              double d = interpS(so.rest.data(), so.rg, rPos) * subV;
              if (d < 1e-6) continue;
              int offset = pp < half ? 0 : 1;
              offset += qq < half ? 0 : 2;
              offset += rr < half ? 0 : 4;
              q[offset] += fraction;
            }
        double sum = 0;
        for (int i = 0; i < 8; i++) sum += q[i];
        long ci = x + y * cx + z * cxy;
        if (sum > 0.5) {
          for (int i = 0; i < 8; i++) q[i] = 1;
          so.valid[ci] = 1;
        } else {
          if (sum > 0) so.valid[ci] = 2;
          for (int i = 0; i < 8; i++) q[i] = 0;
        }
      }
  for (int i = 0; i < 8; i++) {
#pragma omp parallel for schedule(static)
    for (int z = so.tMin[2]; z < so.tMax[2] - 1; z++)
      for (int y = so.tMin[1]; y < so.tMax[1] - 1; y++)
        for (int x = so.tMin[0]; x < so.tMax[0] - 1; x++) {
          long ci = x + y * cx + z * cxy;
          if (!so.valid[ci]) continue;
          long index = z * s.slab + (long)y * s.X + x;
          so.mass[index + s.off[i]] +=
              so.qfrac[index * 8 + i] * so.detFInv[index] * so.qcm;
        }
  }
}

// Box part of DEFOR_SOLID_3D::computeSDF  DEFOR_SOLID_3D.cpp:216-266
void updateBoxes(Scene& s, Solid& so) {
  int xmax = 0, ymax = 0, zmax = 0, xmin = s.X, ymin = s.Y, zmin = s.Z;
  long index = 0;
  for (int z = 0; z < s.Z; z++)
    for (int y = 0; y < s.Y; y++)
      for (int x = 0; x < s.X; x++, index++)
        if (so.mass[index] > 0) {
          xmin = std::min(xmin, x);
          xmax = std::max(xmax, x);
          ymin = std::min(ymin, y);
          ymax = std::max(ymax, y);
          zmin = std::min(zmin, z);
          zmax = std::max(zmax, z);
        }
  const int tb = s.p.tight_border, ex = s.p.extrapolation_cells;
  so.tMin[0] = std::max(xmin - tb, 0);
  so.tMin[1] = std::max(ymin - tb, 0);
  so.tMin[2] = std::max(zmin - tb, 0);
  so.tMax[0] = std::min(xmax + tb, s.X);
  so.tMax[1] = std::min(ymax + tb, s.Y);
  so.tMax[2] = std::min(zmax + tb, s.Z);
  so.bMin[0] = std::max(xmin - ex, 0);
  so.bMin[1] = std::max(ymin - ex, 0);
  so.bMin[2] = std::max(zmin - ex, 0);
  so.bMax[0] = std::min(xmax + ex, s.X);
  so.bMax[1] = std::min(ymax + ex, s.Y);
  so.bMax[2] = std::min(zmax + ex, s.Z);
}

extern "C" void orc_sweep_phi(int X, int Y, int Z, const int box[6], const double* mass,
                              double* phi);

// DEFOR_SOLID_3D::computeSDF  DEFOR_SOLID_3D.cpp:216-267: the boxes, then
// phi = large, -0.5 on massive nodes, two rounds of sweepPhi over the
// extended box
void computeSDF(Scene& s, Solid& so) {
  updateBoxes(s, so);
  const double large = (double)(s.X + s.Y + s.Z + 2);
  for (long i = 0; i < s.total; i++) so.phi[i] = so.mass[i] > 0 ? -0.5 : large;
  const int box[6] = {so.bMin[0], so.bMin[1], so.bMin[2], so.bMax[0], so.bMax[1], so.bMax[2]};
  for (int r = 0; r < 2; r++) orc_sweep_phi(s.X, s.Y, s.Z, box, so.mass.data(), so.phi.data());
}

// ---------------------------------------------------------------------------
// FLIP material particles  DEFOR_SOLID_3D.cpp:333-367, 891-1107
// ---------------------------------------------------------------------------
// DEFOR_SOLID_3D::initMaterialParticles  :333-367.  A fresh MERSENNETWISTER
// seeded with 123456 every call (MERSENNETWISTER.h:259-263, standard
// MT19937 as std::mt19937); rand(n) = randInt() * (1/4294967295) * n
// (:146-150); the offset triple VEC3F(rand(dh), rand(dh), rand(dh)) takes
// its three draws left to right (clang's argument order, the reference's
// toolchain).  The particle density and volume are function-local statics
// of the first solid that ran it (:338-340): `density` is passed in.
void initParticles(Scene& s, Solid& so, double density) {
  const int ppc = s.p.particles_per_cell;
  const double volume = s.dh * s.dh * s.dh / ppc;
  std::mt19937 twister(123456);
  auto rnd = [&](double n) { return double(twister()) * (1.0 / 4294967295.0) * n; };
  so.ppos.clear();
  so.pvel.clear();
  so.pmass.clear();
  const Geom g = s.geom();
  const long cx = s.X - 1, cxy = (long)(s.X - 1) * (s.Y - 1);
  for (int z = so.tMin[2]; z < so.tMax[2] - 1; z++)
    for (int y = so.tMin[1]; y < so.tMax[1] - 1; y++)
      for (int x = so.tMin[0]; x < so.tMax[0] - 1; x++) {
        if (so.valid[x + y * cx + z * cxy] != 1) continue;
        const double grid[3] = {s.o[0] + x * s.dh, s.o[1] + y * s.dh, s.o[2] + z * s.dh};
        for (int i = 0; i < ppc; i++) {
          const double ox = rnd(s.dh);
          const double oy = rnd(s.dh);
          const double oz = rnd(s.dh);
          const double pos[3] = {grid[0] + ox, grid[1] + oy, grid[2] + oz};
          double v[3];
          interpV(so.vel.data(), g, pos, v);
          for (int c = 0; c < 3; c++) {
            so.ppos.push_back(pos[c]);
            so.pvel.push_back(v[c]);
          }
          so.pmass.push_back(density * volume);
        }
      }
}

// the cubic B-spline weight of one axis (:917-925, :1068-1076)
inline double bspline(double d) {
  if (std::fabs(d) < 1) return 0.5 * std::fabs(std::pow(d, 3)) - d * d + 2.0 / 3;
  if (std::fabs(d) < 2) return -1.0 / 6 * std::fabs(std::pow(d, 3)) + d * d - 2 * std::fabs(d) + 4.0 / 3;
  return 0;
}

// DEFOR_SOLID_3D::gridToParticle  :891-1034, without the contact with other
// solids (:969-995, which reads their rest-space SDF; the caller runs it
// for scenes of one solid), then the walls (:996-1031) and the move
void gridToParticle(Scene& s, Solid& so) {
  const double alpha = s.p.flip_alpha;
  const double bnd[6] = {s.o[0], s.o[1], s.o[2], s.o[0] + s.dh * (s.X - 1),
                         s.o[1] + s.dh * (s.Y - 1), s.o[2] + s.dh * (s.Z - 1)};
  static const double wallN[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                     {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  const long np = (long)so.pmass.size();
  auto V = [&](const vector<double>& f, long n, int c) { return f[3 * n + c]; };
#pragma omp parallel for schedule(static)
  for (long c = 0; c < np; c++) {
    double* pp = &so.ppos[3 * c];
    double* pv = &so.pvel[3 * c];
    const int i = (int)std::floor((pp[0] - s.o[0]) / s.dh) - 1;
    const int j = (int)std::floor((pp[1] - s.o[1]) / s.dh) - 1;
    const int k = (int)std::floor((pp[2] - s.o[2]) / s.dh) - 1;
    double pic[3] = {0, 0, 0}, flip[3] = {0, 0, 0};
    for (int r = k; r < k + 4; r++)
      for (int t = j; t < j + 4; t++)
        for (int q = i; q < i + 4; q++) {
          if (q < 0 || t < 0 || r < 0 || q > s.X - 1 || t > s.Y - 1 || r > s.Z - 1) continue;
          const long n = r * s.slab + (long)t * s.X + q;
          if (so.mass[n] == 0) continue;
          const double diff[3] = {s.dhInv * (pp[0] - s.o[0]) - q, s.dhInv * (pp[1] - s.o[1]) - t,
                                  s.dhInv * (pp[2] - s.o[2]) - r};
          const double w = (bspline(diff[0]) * bspline(diff[1])) * bspline(diff[2]);
          for (int d = 0; d < 3; d++) {
            pic[d] += V(so.vel, n, d) * w;
            flip[d] += (V(so.vel, n, d) - V(so.velws, n, d)) * w;
          }
        }
    for (int d = 0; d < 3; d++) {
      flip[d] += pv[d];
      pv[d] = (1 - alpha) * pic[d] + alpha * flip[d];
    }
  }
  const double thick = s.p.wall_thickness, ks = s.p.spring_stiffness, fr = s.p.friction;
  const double dt = so.dt;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < np; c++) {
    double* pp = &so.ppos[3 * c];
    double* pv = &so.pvel[3 * c];
    int i = (int)std::floor((pp[0] - s.o[0]) / s.dh);
    int j = (int)std::floor((pp[1] - s.o[1]) / s.dh);
    int k = (int)std::floor((pp[2] - s.o[2]) / s.dh);
    i = std::min(std::max(i, 0), s.X - 2);
    j = std::min(std::max(j, 0), s.Y - 2);
    k = std::min(std::max(k, 0), s.Z - 2);
    for (int b = 0; b < 6; b++) {
      const double dist = std::fabs(pp[b % 3] - bnd[b]);
      if (!(dist < thick)) continue;
      const double overlap = thick - dist;
      const int xi = b == 0 ? 0 : (b == 3 ? s.X - 1 : i);
      const int yi = b == 1 ? 0 : (b == 4 ? s.Y - 1 : j);
      const int zi = b == 2 ? 0 : (b == 5 ? s.Z - 1 : k);
      const long n = zi * s.slab + (long)yi * s.X + xi;
      const double* nb = wallN[b];
      double vrel[3];
      for (int d = 0; d < 3; d++) vrel[d] = pv[d] - so.vel[3 * n + d];
      const double vn = sum3(vrel[0] * nb[0], vrel[1] * nb[1], vrel[2] * nb[2]);
      if (vn <= 0) {
        const double impulse = -std::min(dt * ks * overlap, so.pmass[c] * (0.1 * overlap / dt - vn));
        const double dvn = -impulse / so.pmass[c];
        double vt[3];
        for (int d = 0; d < 3; d++) vt[d] = vrel[d] - vn * nb[d];
        const double vtn = std::sqrt(sum3(vt[0] * vt[0], vt[1] * vt[1], vt[2] * vt[2]));
        const double sc = std::max(0.1 - fr * dvn / vtn, 0.0);
        for (int d = 0; d < 3; d++) pv[d] = ((vn + dvn) * nb[d] + sc * vt[d]) + so.vel[3 * n + d];
        break;
      }
    }
    for (int d = 0; d < 3; d++) pp[d] += pv[d] * dt;
  }
}

// DEFOR_SOLID_3D::particleToGrid  :1036-1107: per node the weighted velocity
// and weight sums over the particles whose 4^3 stencil covers it, in
// particle order (the single-threaded build); _mass becomes 1 where the
// weight reaches the cutoff, where the velocity becomes the average, and 0
// elsewhere
void particleToGrid(Scene& s, Solid& so) {
  const long np = (long)so.pmass.size();
  vector<double> vs(3 * s.total, 0.0), ws(s.total, 0.0);
  vector<char> touched(s.total, 0);
  for (long c = 0; c < np; c++) {
    const double* pp = &so.ppos[3 * c];
    const double* pv = &so.pvel[3 * c];
    const int i = (int)std::floor((pp[0] - s.o[0]) / s.dh) - 1;
    const int j = (int)std::floor((pp[1] - s.o[1]) / s.dh) - 1;
    const int k = (int)std::floor((pp[2] - s.o[2]) / s.dh) - 1;
    for (int r = k; r < k + 4; r++)
      for (int t = j; t < j + 4; t++)
        for (int q = i; q < i + 4; q++) {
          if (q < 0 || t < 0 || r < 0 || q > s.X - 1 || t > s.Y - 1 || r > s.Z - 1) continue;
          const double diff[3] = {s.dhInv * (pp[0] - s.o[0]) - q, s.dhInv * (pp[1] - s.o[1]) - t,
                                  s.dhInv * (pp[2] - s.o[2]) - r};
          const double w = (bspline(diff[0]) * bspline(diff[1])) * bspline(diff[2]);
          const long n = r * s.slab + (long)t * s.X + q;
          if (!touched[n]) {
            touched[n] = 1;
            for (int d = 0; d < 3; d++) vs[3 * n + d] = w * pv[d];
            ws[n] = w;
          } else {
            for (int d = 0; d < 3; d++) vs[3 * n + d] += w * pv[d];
            ws[n] += w;
          }
        }
  }
  std::fill(so.mass.begin(), so.mass.end(), 0.0);
  const double cutoff = s.p.flip_cutoff;
  for (long n = 0; n < s.total; n++) {
    if (!touched[n] || ws[n] < cutoff) continue;
    for (int d = 0; d < 3; d++) so.vel[3 * n + d] = vs[3 * n + d] / ws[n];
    so.mass[n] = 1;
  }
}

extern "C" void orc_extend_vector_field(int X, int Y, int Z, const int box[6], const double* mass,
                                        const double* phi, double* field);

// DEFOR_SOLID_3D::extendVectorField(_velocity)  :566-574 over the extended box
void extendSolidVelocity(Scene& s, Solid& so) {
  const int box[6] = {so.bMin[0], so.bMin[1], so.bMin[2], so.bMax[0], so.bMax[1], so.bMax[2]};
  orc_extend_vector_field(s.X, s.Y, s.Z, box, so.mass.data(), so.phi.data(), so.vel.data());
}

// COUPLED_INTEGRATOR_3D::gatherSolidMasses  COUPLED_INTEGRATOR_3D.cpp:327-375
void gather(Scene& s) {
  std::fill(s.solidMass.begin(), s.solidMass.end(), 0.0);
  std::fill(s.qfracSum.begin(), s.qfracSum.end(), 0.0);
  for (Solid& so : s.solids) {
    for (long i = 0; i < s.total; i++) s.solidMass[i] += so.mass[i];
    for (long i = 0; i < 8 * s.total; i++) s.qfracSum[i] += so.qfrac[i];
  }
  std::fill(s.nodeFrac.begin(), s.nodeFrac.end(), 0.0);
  for (int i = 0; i < 8; i++)
    for (int z = 1; z < s.Z - 1; z++)
      for (int y = 1; y < s.Y - 1; y++)
        for (int x = 1; x < s.X - 1; x++) {
          long index = z * s.slab + (long)y * s.X + x;
          s.nodeFrac[index + s.off[i]] += s.qfracSum[index * 8 + i];
        }
  for (long i = 0; i < s.total; i++) s.nodeFrac[i] *= 0.125;
  std::fill(s.sidx.begin(), s.sidx.end(), -1);
  int idx = 0;
  for (int z = 1; z < s.Z - 1; z++)
    for (int y = 1; y < s.Y - 1; y++)
      for (int x = 1; x < s.X - 1; x++) {
        long index = z * s.slab + (long)y * s.X + x;
        s.sidx[index] = s.solidMass[index] > 0 ? idx++ : -1;
      }
  s.nSolid = idx;
}

// DEFOR_SOLID_3D::recoverVelocityField  DEFOR_SOLID_3D.cpp:1235-1241
void recoverVelocity(Scene& s, Solid& so) {
  for (int z = so.bMin[2]; z < so.bMax[2] - 1; z++)
    for (int y = so.bMin[1]; y < so.bMax[1] - 1; y++)
      for (int x = so.bMin[0]; x < so.bMax[0] - 1; x++) {
        long n = z * s.slab + (long)y * s.X + x;
        if (so.mass[n] > 0)
          for (int c = 0; c < 3; c++) s.vel[3 * n + c] = so.vel[3 * n + c];
      }
}

// COUPLED_INTEGRATOR_3D::CFL  COUPLED_INTEGRATOR_3D.cpp:88-102
double cfl(Scene& s, double alpha) {
  double dt = 1e9;
  for (int z = 1; z < s.Z - 1; z++)
    for (int y = 1; y < s.Y - 1; y++)
      for (int x = 1; x < s.X - 1; x++) {
        long n = z * s.slab + (long)y * s.X + x;
        if (s.solidMass[n] > 0) {
          double m = std::max(std::max(std::fabs(s.vel[3 * n]),
                                       std::fabs(s.vel[3 * n + 1])),
                              std::fabs(s.vel[3 * n + 2]));
          double vmax = m * s.dhInv;
          dt = std::min(dt, alpha / vmax);
        }
      }
  dt = std::min(dt, s.defaultDt);
  return dt;
}

// ---------------------------------------------------------------------------
// Element stiffness and force: DEFOR_SOLID_3D::
// computeStiffnessMatrixAndMaterialForce  DEFOR_SOLID_3D.cpp:697-871
// (plastic branch :765-795 is inactive for the shipped cfgs).  Per-cell
// results are produced in parallel and merged serially in cell order, which
// is the order of the reference's single-threaded build (USING_OPENMP 0,
// SETTINGS.h:40).
// ---------------------------------------------------------------------------
struct Trip {
  int r, c;
  double v;
};

double stiffnessAndForce(Scene& s, Solid& so, vector<double>& force,
                         vector<Trip>& trips) {
  const double g = 0.25 * s.dhInv;
  const double dh3 = s.dh;
  vector<long> cells;
  const long cx = s.X - 1, cxy = (long)(s.X - 1) * (s.Y - 1);
  for (int z = so.tMin[2]; z < so.tMax[2] - 1; z++)
    for (int y = so.tMin[1]; y < so.tMax[1] - 1; y++)
      for (int x = so.tMin[0]; x < so.tMax[0] - 1; x++)
        if (so.valid[x + y * cx + z * cxy] == 1)
          cells.push_back(z * s.slab + (long)y * s.X + x);
  const long nc = (long)cells.size();
  vector<double> nodal(nc * 24), stiff(nc * 576), energyCell(nc);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < nc; e++) {
    long index = cells[e];
    double F[9];
    double Jinv = 1.0 / computeF(s, so, index, F);
    if ((long)so.Fset.size() == 9 * nc) {
      for (int k = 0; k < 9; k++) F[k] = so.Fset[9 * e + k];
      Jinv = 1.0 / det3(F);
    }
    Corot m;
    m.lambda = so.lambda;
    m.mu = so.mu;
    corotInit(m, F);
    double P[9];
    corotP(m, P);
    double JP[9], cs[9];
    for (int i = 0; i < 9; i++) JP[i] = Jinv * P[i];
    mmT(JP, F, cs);  // cauchyStress = Jinv * P * F^T  (:763)
    // GF = G * F (8x3)
    double GF[24];
    for (int i = 0; i < 8; i++) {
      double Gi[3] = {sgnX(i) * g, sgnY(i) * g, sgnZ(i) * g};
      for (int c = 0; c < 3; c++)
        GF[3 * i + c] = (Gi[0] * F[c] + Gi[1] * F[3 + c]) + Gi[2] * F[6 + c];
    }
    // nodalForcesFEM -= cauchyStress * G^T * dh^3  (:801)
    for (int i = 0; i < 8; i++) {
      double Gi[3] = {sgnX(i) * g, sgnY(i) * g, sgnZ(i) * g};
      for (int r = 0; r < 3; r++) {
        double t = (cs[3 * r] * Gi[0] + cs[3 * r + 1] * Gi[1]) + cs[3 * r + 2] * Gi[2];
        nodal[e * 24 + 3 * i + r] = 0.0 - ((t * dh3) * dh3) * dh3;
      }
    }
    energyCell[e] = ((corotEnergy(m) * dh3) * dh3) * dh3;
    double* K = &stiff[e * 576];
    for (int k = 0; k < 576; k++) K[k] = 0;
    for (int k = 0; k < 24; k++) {
      long col = index + s.off[k / 3];
      if (so.mass[col] == 0) continue;
      // dF = e_j (GF_i)^T, j = k%3, i = k/3 (computePFPxhat :684-695)
      double dF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      int i = k / 3, j = k % 3;
      for (int c = 0; c < 3; c++) dF[3 * j + c] = GF[3 * i + c];
      double dP[9];
      corotDP(m, dF, dP);
      // deltaForce = deltaP * GF^T * Jinv * dh^3 (:820-821)
      for (int mm_ = 0; mm_ < 8; mm_++)
        for (int r = 0; r < 3; r++) {
          double t = (dP[3 * r] * GF[3 * mm_] + dP[3 * r + 1] * GF[3 * mm_ + 1]) +
                     dP[3 * r + 2] * GF[3 * mm_ + 2];
          t = (((t * Jinv) * dh3) * dh3) * dh3;
          K[(3 * mm_ + r) * 24 + k] += t;
        }
    }
  }
  // serial merge in cell order (:829-846)
  vector<double> local(force.size(), 0.0);
  double energy = 0;
  for (long e = 0; e < nc; e++) {
    long index = cells[e];
    energy += energyCell[e];
    for (int i = 0; i < 8; i++) {
      int idx = s.sidx[index + s.off[i]];
      if (idx >= 0)
        for (int r = 0; r < 3; r++) local[idx * 3 + r] += nodal[e * 24 + 3 * i + r];
    }
    const double* K = &stiff[e * 576];
    for (int i = 0; i < 24; i++) {
      long col = index + s.off[i / 3];
      if (so.mass[col] == 0) continue;
      if (s.sidx[col] < 0) continue;
      int newCol = s.sidx[col] * 3 + i % 3;
      for (int k = 0; k < 24; k++) {
        long row = index + s.off[k / 3];
        if (so.mass[row] == 0) continue;
        if (s.sidx[row] < 0) continue;
        int newRow = s.sidx[row] * 3 + k % 3;
        trips.push_back(Trip{newRow, newCol, K[k * 24 + i]});
      }
    }
  }
  for (size_t i = 0; i < force.size(); i++) force[i] += local[i];
  return energy;
}

// DEFOR_SOLID_3D::computeMV  DEFOR_SOLID_3D.cpp:873-889
double computeMV(Scene& s, Solid& so, vector<double>& mv) {
  double energy = 0;
  for (int z = so.bMin[2]; z < so.bMax[2] - 1; z++)
    for (int y = so.bMin[1]; y < so.bMax[1] - 1; y++)
      for (int x = so.bMin[0]; x < so.bMax[0] - 1; x++) {
        long n = z * s.slab + (long)y * s.X + x;
        int idx = s.sidx[n];
        if (idx >= 0) {
          const double* v = &so.vel[3 * n];
          for (int c = 0; c < 3; c++) mv[idx * 3 + c] += so.mass[n] * v[c];
          energy += 0.5 * so.mass[n] * sum3(v[0] * v[0], v[1] * v[1], v[2] * v[2]);
        }
      }
  return energy;
}

// stepImplicit :413-607 -- stress, computeMV, vorticity, body forces
double forces(Scene& s, vector<Trip>& trips) {
  const long na = s.totalActive;
  s.mv.assign(3 * na, 0.0);
  s.unitForce.assign(3 * na, 0.0);
  s.nzMass.assign(3 * na, 0.0);
  s.minv.assign(3 * na, 0.0);
  s.sForce.assign(3 * (size_t)s.nSolid, 0.0);
  s.sMv.assign(3 * (size_t)s.nSolid, 0.0);
  s.sDamped.assign(3 * (size_t)s.nSolid, 0.0);
  double t0 = nowSec();
  double energy = 0;
  for (Solid& so : s.solids) {
    energy += stiffnessAndForce(s, so, s.sForce, trips);
    energy += computeMV(s, so, s.sMv);
  }
  double t1 = nowSec();
  s.stageSec[1] += t1 - t0;

  const double gridSize = 0.5 / s.dh;
  std::fill(s.ws.begin(), s.ws.end(), 0.0);
  const double solidMassDampCoeff = 1 + s.dt * s.p.rayleigh_alpha;
  const long X = s.X, S = s.slab;
  const double dhInv = s.dhInv;
  const vector<double>& nf = s.nodeFrac;
  const vector<double>& v = s.vel;
  if (s.vortEps > 0) {
#pragma omp parallel for schedule(static)
    for (int z = 1; z < s.Z - 1; z++)
      for (int y = 1; y < s.Y - 1; y++)
        for (int x = 1; x < s.X - 1; x++) {
          long index = z * S + (long)y * X + x;
          if (nf[index] >= 1) continue;
          long up = nf[index + X] >= 1 || y == s.Y - 2 ? index : index + X;
          long down = nf[index - X] >= 1 || y == 1 ? index : index - X;
          double dy = (up == index || down == index) ? dhInv : gridSize;
          long out = nf[index + S] >= 1 || z == s.Z - 2 ? index : index + S;
          long in = nf[index - S] >= 1 || z == 1 ? index : index - S;
          double dz = (out == index || in == index) ? dhInv : gridSize;
          long right = nf[index + 1] >= 1 || x == s.X - 2 ? index : index + 1;
          long left = nf[index - 1] >= 1 || x == 1 ? index : index - 1;
          double dx = (right == index || left == index) ? dhInv : gridSize;
          double w0 = ((v[3 * up + 2] - v[3 * down + 2]) +
                       (-v[3 * out + 1] + v[3 * in + 1])) * dz;
          double w1 = ((v[3 * out] - v[3 * in]) +
                       (-v[3 * right + 2] + v[3 * left + 2])) * dx;
          double w2 = ((v[3 * right + 1] - v[3 * left + 1]) +
                       (-v[3 * up] + v[3 * down])) * dy;
          s.vort[3 * index] = w0;
          s.vort[3 * index + 1] = w1;
          s.vort[3 * index + 2] = w2;
          s.ws[index] = std::sqrt(sum3(w0 * w0, w1 * w1, w2 * w2));
        }
  }
  const double gravity = s.p.gravity, buoy = s.p.buoyancy;
  const double dh3 = s.dh;
  double en = 0;
#pragma omp parallel for schedule(static) reduction(+ : en)
  for (int z = 1; z < s.Z - 1; z++)
    for (int y = 1; y < s.Y - 1; y++)
      for (int x = 1; x < s.X - 1; x++) {
        long index = z * S + (long)y * X + x;
        long idx = s.sol[index];
        int solidIdx = s.sidx[index];
        double fluidMass =
            std::max(1.0 - nf[index], 0.0) * s.p.fluid_density * dh3 * dh3 * dh3;
        double mass = fluidMass + s.solidMass[index];
        double h = s.dh * y + s.o[1];
        en += mass * gravity * h;
        if (solidIdx < 0) {
          double mg = mass * gravity;
          s.unitForce[idx * 3] += mg * 0;
          s.unitForce[idx * 3 + 1] += mg * -1;
          s.unitForce[idx * 3 + 2] += mg * 0;
          s.unitForce[idx * 3 + 1] += buoy * fluidMass * s.intensity[index];
          s.nzMass[idx * 3] = s.nzMass[idx * 3 + 1] = s.nzMass[idx * 3 + 2] = mass;
          for (int c = 0; c < 3; c++) s.mv[idx * 3 + c] = fluidMass * v[3 * index + c];
          if (s.vortEps > 0) {
            if (nf[index + X] == 0 && nf[index - X] == 0 && nf[index + 1] == 0 &&
                nf[index - 1] == 0 && nf[index + S] == 0 && nf[index - S] == 0) {
              long up = y == s.Y - 2 ? index : index + X;
              long down = nf[index - X] >= 1 || y == 1 ? index : index - X;
              double dy = (up == index || down == index) ? dhInv : gridSize;
              long out = nf[index + S] >= 1 || z == s.Z - 2 ? index : index + S;
              long in = nf[index - S] >= 1 || z == 1 ? index : index - S;
              double dz = (out == index || in == index) ? dhInv : gridSize;
              long right = nf[index + 1] >= 1 || x == s.X - 2 ? index : index + 1;
              long left = nf[index - 1] >= 1 || x == 1 ? index : index - 1;
              double dx = (right == index || left == index) ? dhInv : gridSize;
              double N[3];
              N[0] = (s.ws[right] - s.ws[left]) * dx;
              N[1] = (s.ws[up] - s.ws[down]) * dy;
              N[2] = (s.ws[out] - s.ws[in]) * dz;
              double mag = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
              if (mag > 0.0) {
                mag = 1.0 / mag;
                N[0] *= mag;
                N[1] *= mag;
                N[2] *= mag;
                const double* w = &s.vort[3 * index];
                s.unitForce[idx * 3] +=
                    (N[1] * w[2] - N[2] * w[1]) * s.dh * s.vortEps * fluidMass;
                s.unitForce[idx * 3 + 1] -=
                    (N[0] * w[2] - N[2] * w[0]) * s.dh * s.vortEps * fluidMass;
                s.unitForce[idx * 3 + 2] +=
                    (N[0] * w[1] - N[1] * w[0]) * s.dh * s.vortEps * fluidMass;
              }
            }
          }
        } else {
          s.sDamped[solidIdx * 3] = s.sDamped[solidIdx * 3 + 1] =
              s.sDamped[solidIdx * 3 + 2] =
                  s.solidMass[index] * solidMassDampCoeff + fluidMass;
          for (int c = 0; c < 3; c++) s.sMv[solidIdx * 3 + c] += fluidMass * v[3 * index + c];
          double mg = mass * gravity;
          s.sForce[solidIdx * 3] += mg * 0;
          s.sForce[solidIdx * 3 + 1] += mg * -1;
          s.sForce[solidIdx * 3 + 2] += mg * 0;
          s.sForce[solidIdx * 3 + 1] += buoy * fluidMass * s.intensity[index];
        }
        s.minv[idx * 3] = s.minv[idx * 3 + 1] = s.minv[idx * 3 + 2] = 1.0 / mass;
        const double* vv = &v[3 * index];
        en += 0.5 * fluidMass * sum3(vv[0] * vv[0], vv[1] * vv[1], vv[2] * vv[2]);
      }
  s.stageSec[2] += nowSec() - t1;
  return energy + en;
}

// stepImplicit :609-620 (Eigen setFromTriplets sums duplicates in triplet
// order, rows keep ascending columns) and computeDivergenceConstraintRHS
// :286-325
void assemble(Scene& s, vector<Trip>& trips) {
  double t0 = nowSec();
  const int n = 3 * s.nSolid;
  vector<vector<std::pair<int, double>>> rows(n);
  for (const Trip& t : trips) rows[t.r].push_back({t.c, t.v});
  const double scale = s.dt * s.dt + s.p.rayleigh_beta * s.dt;
  CSR& A = s.A;
  A.n = n;
  A.rowptr.assign(n + 1, 0);
  A.col.clear();
  A.val.clear();
  s.Adiag.assign(n, 0.0);
  for (int r = 0; r < n; r++) {
    auto& row = rows[r];
    std::stable_sort(row.begin(), row.end(),
                     [](const std::pair<int, double>& a,
                        const std::pair<int, double>& b) { return a.first < b.first; });
    bool hasDiag = false;
    for (size_t k = 0; k < row.size();) {
      int c = row[k].first;
      double v = row[k].second;
      size_t k2 = k + 1;
      while (k2 < row.size() && row[k2].first == c) v += row[k2++].second;
      k = k2;
      if (!hasDiag && c > r) {  // insert an explicit diagonal before c
        A.col.push_back(r);
        A.val.push_back(0.0 + s.sDamped[r]);
        hasDiag = true;
      }
      double val = v * scale;
      if (c == r) {
        val += s.sDamped[r];
        hasDiag = true;
      }
      A.col.push_back(c);
      A.val.push_back(val);
    }
    if (!hasDiag) {
      A.col.push_back(r);
      A.val.push_back(0.0 + s.sDamped[r]);
    }
    A.rowptr[r + 1] = (int)A.col.size();
  }
  for (int r = 0; r < n; r++)
    for (int k = A.rowptr[r]; k < A.rowptr[r + 1]; k++)
      if (A.col[k] == r) s.Adiag[r] = A.val[k];
  // divergence RHS
  const long cx = s.X - 1, cxy = (long)(s.X - 1) * (s.Y - 1);
  s.c.assign(s.nCells, 0.0);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < s.Z - 1; z++)
    for (int y = 0; y < s.Y - 1; y++)
      for (int x = 0; x < s.X - 1; x++) {
        long cnt = x + y * cx + z * cxy;
        long index = z * s.slab + (long)y * s.X + x;
        double acc = 0;
        for (int i = 0; i < 8; i++) {
          long nd = index + s.off[i];
          if (s.sol[nd] < 0) {
            acc -= sgnX(i) * s.vel[3 * nd];
            acc -= sgnY(i) * s.vel[3 * nd + 1];
            acc -= sgnZ(i) * s.vel[3 * nd + 2];
          }
        }
        s.c[cnt] = acc;
      }
  s.stageSec[3] += nowSec() - t0;
}

// CONJUGATE_RESIDUAL::parallelMatVecMult  CONJUGATE_RESIDUAL.inl:31-44
void csrMul(const CSR& A, const vector<double>& b, vector<double>& x) {
  x.assign(A.n, 0.0);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < A.n; k++) {
    double acc = 0;
    for (int j = A.rowptr[k]; j < A.rowptr[k + 1]; j++) acc += A.val[j] * b[A.col[j]];
    x[k] = acc;
  }
}

// getMatVecMult  COUPLED_INTEGRATOR_3D.cpp:638-660: y = J (minv .* (J^T x)),
// accumulated in the reference's sparse row / column order.
void pressureApply(const Scene& s, const double* x, double* y) {
  const long X = s.X, S = s.slab;
  const long cx = s.X - 1, cxy = (long)(s.X - 1) * (s.Y - 1);
  vector<double> mid(3 * s.total, 0.0);
#pragma omp parallel for schedule(static)
  for (int k = 1; k < s.Z - 1; k++)
    for (int j = 1; j < s.Y - 1; j++)
      for (int i = 1; i < s.X - 1; i++) {
        long n = k * S + (long)j * X + i;
        double m0 = 0, m1 = 0, m2 = 0;
        // adjacent cells in ascending cell index (JT row order)
        for (int dz = 0; dz < 2; dz++)
          for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) {
              int ccx = i - 1 + dx, ccy = j - 1 + dy, ccz = k - 1 + dz;
              double xv = x[ccx + ccy * cx + ccz * cxy];
              // node is corner (1-dx, 1-dy, 1-dz) of this cell
              m0 += (dx == 0 ? 1.0 : -1.0) * xv;
              m1 += (dy == 0 ? 1.0 : -1.0) * xv;
              m2 += (dz == 0 ? 1.0 : -1.0) * xv;
            }
        mid[3 * n] = m0;
        mid[3 * n + 1] = m1;
        mid[3 * n + 2] = m2;
      }
#pragma omp parallel for schedule(static)
  for (int z = 0; z < s.Z - 1; z++)
    for (int yy = 0; yy < s.Y - 1; yy++)
      for (int xx = 0; xx < s.X - 1; xx++) {
        long r = xx + yy * cx + z * cxy;
        long index = z * S + (long)yy * X + xx;
        double acc = 0;
        for (int c = 0; c < 8; c++) {
          long nd = index + s.off[c];
          long sol = s.sol[nd];
          if (sol < 0) continue;
          double mi = s.minv[3 * sol];
          acc += (sgnX(c) * mid[3 * nd]) * mi;
          acc += (sgnY(c) * mid[3 * nd + 1]) * mi;
          acc += (sgnZ(c) * mid[3 * nd + 2]) * mi;
        }
        y[r] = acc;
      }
}

// JMInvJT diagonal  COUPLED_INTEGRATOR_3D.cpp:694-705
void pressureDiag(const Scene& s, double* d) {
  const long cx = s.X - 1, cxy = (long)(s.X - 1) * (s.Y - 1);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < s.Z - 1; z++)
    for (int yy = 0; yy < s.Y - 1; yy++)
      for (int xx = 0; xx < s.X - 1; xx++) {
        long index = z * s.slab + (long)yy * s.X + xx;
        double sum = 0;
        for (int c = 0; c < 8; c++) {
          long sol = s.sol[index + s.off[c]];
          if (sol < 0) continue;
          for (int dd = 0; dd < 3; dd++) sum += s.minv[3 * sol + dd] * 1.0;
        }
        d[xx + yy * cx + z * cxy] = sum;
      }
}

// J v over solution-ordered DOFs (Eigen RowMajor sparse * dense)
void applyJ(const Scene& s, const vector<double>& v, vector<double>& out) {
  const long cx = s.X - 1, cxy = (long)(s.X - 1) * (s.Y - 1);
  out.assign(s.nCells, 0.0);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < s.Z - 1; z++)
    for (int yy = 0; yy < s.Y - 1; yy++)
      for (int xx = 0; xx < s.X - 1; xx++) {
        long index = z * s.slab + (long)yy * s.X + xx;
        double acc = 0;
        for (int c = 0; c < 8; c++) {
          long sol = s.sol[index + s.off[c]];
          if (sol < 0) continue;
          acc += sgnX(c) * v[3 * sol];
          acc += sgnY(c) * v[3 * sol + 1];
          acc += sgnZ(c) * v[3 * sol + 2];
        }
        out[xx + yy * cx + z * cxy] = acc;
      }
}

// -J^T lam (Eigen ColMajor transpose product: cells in ascending order)
void applyNegJT(const Scene& s, const vector<double>& lam, vector<double>& out) {
  const long cx = s.X - 1, cxy = (long)(s.X - 1) * (s.Y - 1);
  out.assign(3 * s.totalActive, 0.0);
#pragma omp parallel for schedule(static)
  for (int k = 1; k < s.Z - 1; k++)
    for (int j = 1; j < s.Y - 1; j++)
      for (int i = 1; i < s.X - 1; i++) {
        long n = k * s.slab + (long)j * s.X + i;
        long sol = s.sol[n];
        double a0 = 0, a1 = 0, a2 = 0;
        for (int dz = 0; dz < 2; dz++)
          for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) {
              double l = lam[(i - 1 + dx) + (j - 1 + dy) * cx + (k - 1 + dz) * cxy];
              a0 += -(dx == 0 ? 1.0 : -1.0) * l;
              a1 += -(dy == 0 ? 1.0 : -1.0) * l;
              a2 += -(dz == 0 ? 1.0 : -1.0) * l;
            }
        out[3 * sol] = a0;
        out[3 * sol + 1] = a1;
        out[3 * sol + 2] = a2;
      }
}

// CONJUGATE_RESIDUAL::solvePCR, both variants  CONJUGATE_RESIDUAL.inl:109-276
template <class MatVec>
bool pcr(MatVec apply, const vector<double>& diag, const vector<double>& b,
         vector<double>& x, double eps, int maxIt, int* iters, double* relRes) {
  const long n = (long)b.size();
  vector<double> q(n), r(n), p(n), s(n), Ap(n), tmp(n);
  apply(x, q);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) r[i] = (b[i] - q[i]) / diag[i];
  p = r;
  apply(r, s);
  Ap = s;
  double deltaNew = vdot(r, s);
  double bNorm = vnorm(b);
  bool converged = false;
  int it = 0;
  double lastRel = -1;
  for (; it < maxIt && !converged; it++) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) q[i] = Ap[i] / diag[i];
    double alpha = vdot(Ap, q);
    if (std::fabs(alpha) > 0.0) alpha = deltaNew / alpha;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    apply(r, s);
    double deltaOld = deltaNew;
    deltaNew = vdot(r, s);
    double beta = deltaNew / deltaOld;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
      p[i] = p[i] * beta + r[i];
      Ap[i] = Ap[i] * beta + s[i];
    }
    if (it > 0 && it % 5 == 0) {
      apply(x, tmp);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; i++) tmp[i] -= b[i];
      double tn = vnorm(tmp);
      lastRel = bNorm > 0 ? tn / bNorm : tn;
      if (tn < eps * bNorm) {
        converged = true;
        break;
      }
    }
  }
  if (iters) *iters = it;
  if (relRes) {
    apply(x, tmp);
    for (long i = 0; i < n; i++) tmp[i] -= b[i];
    double tn = vnorm(tmp);
    *relRes = bNorm > 0 ? tn / bNorm : tn;
    (void)lastRel;
  }
  return converged;
}

// COUPLED_INTEGRATOR_3D::pcrImplicitSolve  COUPLED_INTEGRATOR_3D.cpp:662-764
int solve(Scene& s, orc_stats* st) {
  double t0 = nowSec();
  const long na = s.totalActive;
  const int ns = 3 * s.nSolid;
  s.sForceRhs.resize(ns);
  for (int i = 0; i < ns; i++) s.sForceRhs[i] = s.sForce[i] * s.dt + s.sMv[i];
  s.unitRhs.resize(3 * na);
  for (long i = 0; i < 3 * na; i++) s.unitRhs[i] = s.unitForce[i] * s.dt + s.mv[i];
  auto Aapply = [&](const vector<double>& in, vector<double>& out) { csrMul(s.A, in, out); };
  s.sV0.assign(ns, 0.0);
  int it0 = 0;
  double rr0 = 0;
  bool c0 = pcr(Aapply, s.Adiag, s.sForceRhs, s.sV0, s.p.implicit_eps,
                s.p.cr_max_iterations, &it0, &rr0);
  double t1 = nowSec();
  s.v0.assign(3 * na, 0.0);
  for (long index = 0; index < s.total; index++) {
    int idx = s.sol[index];
    if (idx < 0) continue;
    int sIdx = s.sidx[index];
    for (int c = 0; c < 3; c++)
      s.v0[3 * idx + c] = sIdx < 0 ? s.unitRhs[3 * idx + c] * s.minv[3 * idx]
                                   : s.sV0[3 * sIdx + c];
  }
  s.D.assign(s.nCells, 0.0);
  pressureDiag(s, s.D.data());
  vector<double> Jv;
  applyJ(s, s.v0, Jv);
  s.b.resize(s.nCells);
  for (long r = 0; r < s.nCells; r++) s.b[r] = Jv[r] - s.c[r];
  if (!s.lambdaInit || (long)s.lambda.size() != s.nCells) {
    s.lambda.assign(s.nCells, 0.0);
    s.lambdaInit = true;
  }
  auto Sapply = [&](const vector<double>& in, vector<double>& out) {
    out.resize(s.nCells);
    pressureApply(s, in.data(), out.data());
  };
  int it1 = 0;
  double rr1 = 0;
  bool c1 = pcr(Sapply, s.D, s.b, s.lambda, s.p.divfree_eps,
                s.p.cr_max_iterations, &it1, &rr1);
  double t2 = nowSec();
  applyNegJT(s, s.lambda, s.g);
  s.sB.assign(ns, 0.0);
  for (long x = 0; x < s.total; x++) {
    int idx = s.sol[x], sIdx = s.sidx[x];
    if (sIdx >= 0)
      for (int c = 0; c < 3; c++) s.sB[3 * sIdx + c] = s.g[3 * idx + c];
  }
  s.sDv.assign(ns, 0.0);
  int it2 = 0;
  double rr2 = 0;
  bool c2 = pcr(Aapply, s.Adiag, s.sB, s.sDv, s.p.implicit_eps,
                s.p.cr_max_iterations, &it2, &rr2);
  double t3 = nowSec();
  s.v1.assign(3 * na, 0.0);
  for (long index = 0; index < s.total; index++) {
    int idx = s.sol[index];
    if (idx < 0) continue;
    int sIdx = s.sidx[index];
    for (int c = 0; c < 3; c++) {
      double corr = sIdx < 0 ? s.g[3 * idx + c] * s.minv[3 * idx] : s.sDv[3 * sIdx + c];
      s.v1[3 * idx + c] = s.v0[3 * idx + c] + corr;
    }
  }
  for (long i = 0; i < s.total; i++) {
    int idx = s.sol[i];
    if (idx >= 0)
      for (int c = 0; c < 3; c++) s.vel[3 * i + c] = s.v1[3 * idx + c];
  }
  if (st) {
    st->iterations[0] = it0;
    st->iterations[1] = it1;
    st->iterations[2] = it2;
    st->converged[0] = c0;
    st->converged[1] = c1;
    st->converged[2] = c2;
    st->rel_residual[0] = rr0;
    st->rel_residual[1] = rr1;
    st->rel_residual[2] = rr2;
    st->dt = s.dt;
  }
  s.stageSec[4] += t1 - t0;
  s.stageSec[5] += t2 - t1;
  s.stageSec[6] += nowSec() - t2;
  (void)t3;
  return (c0 && c1 && c2) ? 0 : 1;
}

// DEFOR_SOLID_3D::finalizeSolution(velocity, dt)  DEFOR_SOLID_3D.cpp:1182-1197
void solidCopyVelocity(Scene& s) {
  for (Solid& so : s.solids) {
    for (int z = so.bMin[2]; z < so.bMax[2]; z++)
      for (int y = so.bMin[1]; y < so.bMax[1]; y++)
        for (int x = so.bMin[0]; x < so.bMax[0]; x++) {
          long n = z * s.slab + (long)y * s.X + x;
          for (int c = 0; c < 3; c++) so.vel[3 * n + c] = s.vel[3 * n + c];
        }
    so.dt = s.dt;
  }
}

// DEFOR_SOLID_3D::finalizeSolution()  DEFOR_SOLID_3D.cpp:1198-1233 with
// DEFOR_SOLID_3D::advectField :1108-1123
void solidFinalize(Scene& s) {
  double t0 = nowSec();
  const Geom g = s.geom();
  for (Solid& so : s.solids) {
    const double dt = so.dt;
    for (long i = 0; i < 3 * s.total; i++) so.u[i] += so.vel[i] * dt;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < s.Z; z++)
      for (int y = 0; y < s.Y; y++)
        for (int x = 0; x < s.X; x++) {
          long n = z * s.slab + (long)y * s.X + x;
          double node[3] = {x * s.dh + s.o[0], y * s.dh + s.o[1], z * s.dh + s.o[2]};
          double np[3];
          for (int c = 0; c < 3; c++) np[c] = node[c] - so.vel[3 * n + c] * dt;
          interpV(so.u.data(), g, np, &so.velws[3 * n]);
        }
    std::swap(so.u, so.velws);
    auto U = [&](int x, int y, int z) { return &so.u[3 * (z * s.slab + (long)y * s.X + x)]; };
    const int* mn = so.bMin;
    const int* mx = so.bMax;
    for (int x = mn[0]; x < mx[0]; x++)
      for (int y = mn[1]; y < mx[1]; y++)
        for (int c = 0; c < 3; c++) {
          U(x, y, mn[2])[c] = 2 * U(x, y, mn[2] + 1)[c] - U(x, y, mn[2] + 2)[c];
          U(x, y, mx[2] - 1)[c] = 2 * U(x, y, mx[2] - 2)[c] - U(x, y, mx[2] - 3)[c];
        }
    for (int x = mn[0]; x < mx[0]; x++)
      for (int z = mn[2]; z < mx[2]; z++)
        for (int c = 0; c < 3; c++) {
          U(x, mn[1], z)[c] = 2 * U(x, mn[1] + 1, z)[c] - U(x, mn[1] + 2, z)[c];
          U(x, mx[1] - 1, z)[c] = 2 * U(x, mx[1] - 2, z)[c] - U(x, mx[1] - 3, z)[c];
        }
    for (int y = mn[1]; y < mx[1]; y++)
      for (int z = mn[2]; z < mx[2]; z++)
        for (int c = 0; c < 3; c++) {
          U(mn[0], y, z)[c] = 2 * U(mn[0] + 1, y, z)[c] - U(mn[0] + 2, y, z)[c];
          U(mx[0] - 1, y, z)[c] = 2 * U(mx[0] - 2, y, z)[c] - U(mx[0] - 3, y, z)[c];
        }
    long index = 0;
    for (int z = 0; z < s.Z; z++)
      for (int y = 0; y < s.Y; y++)
        for (int x = 0; x < s.X; x++, index++) {
          double pos[3] = {x * s.dh, y * s.dh, z * s.dh};
          for (int c = 0; c < 3; c++) so.X[3 * index + c] = pos[c] + (s.o[c] - so.u[3 * index + c]);
        }
    so.velws = so.vel;
  }
  s.stageSec[7] += nowSec() - t0;
}

// COUPLED_INTEGRATOR_3D::advectFluidVelocityField  :797-826
void advectVelocity(Scene& s) {
  double t0 = nowSec();
  const Geom g = s.geom();
  const double dt = s.dt;
#pragma omp parallel for schedule(static)
  for (int z = 1; z < s.Z - 1; z++)
    for (int y = 1; y < s.Y - 1; y++)
      for (int x = 1; x < s.X - 1; x++) {
        long n = z * s.slab + (long)y * s.X + x;
        if (s.solidMass[n] > 0) continue;
        double node[3] = {x * s.dh + s.o[0], y * s.dh + s.o[1], z * s.dh + s.o[2]};
        double mp[3], v[3], np[3];
        for (int c = 0; c < 3; c++) mp[c] = node[c] - 0.5 * s.vel[3 * n + c] * dt;
        interpV(s.vel.data(), g, mp, v);
        for (int c = 0; c < 3; c++) np[c] = node[c] - v[c] * dt;
        double mass = interpS(s.solidMass.data(), g, np);
        double newdt = dt / 2;
        while (mass > 0 && newdt > dt / 16) {
          for (int c = 0; c < 3; c++) np[c] += v[c] * newdt;
          newdt /= 2;
          mass = interpS(s.solidMass.data(), g, np);
        }
        interpV(s.vel.data(), g, np, &s.velws[3 * n]);
      }
  std::swap(s.vel, s.velws);
  s.stageSec[8] += nowSec() - t0;
}

// COUPLED_INTEGRATOR_3D::advectField (intensity)  :766-796
void advectIntensity(Scene& s) {
  double t0 = nowSec();
  const Geom g = s.geom();
  const double dt = s.dt;
  std::fill(s.ws.begin(), s.ws.end(), 0.0);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < s.Z; z++)
    for (int y = 0; y < s.Y; y++)
      for (int x = 0; x < s.X; x++) {
        long n = z * s.slab + (long)y * s.X + x;
        if (s.solidMass[n] > 0) continue;
        double node[3] = {x * s.dh + s.o[0], y * s.dh + s.o[1], z * s.dh + s.o[2]};
        double v[3] = {s.vel[3 * n], s.vel[3 * n + 1], s.vel[3 * n + 2]};
        double np[3];
        for (int c = 0; c < 3; c++) np[c] = node[c] - v[c] * dt;
        double mass = interpS(s.solidMass.data(), g, np);
        double newdt = s.dt / 2;
        while (mass > 0 && newdt > dt / 16) {
          for (int c = 0; c < 3; c++) np[c] += v[c] * newdt;
          newdt /= 2;
          mass = interpS(s.solidMass.data(), g, np);
        }
        s.ws[n] = interpS(s.intensity.data(), g, np);
      }
  std::swap(s.intensity, s.ws);
  s.stageSec[9] += nowSec() - t0;
}

// finalizeSolution tail  COUPLED_INTEGRATOR_3D.cpp:894-916
void recoverAndBC(Scene& s) {
  for (Solid& so : s.solids) recoverVelocity(s, so);
  auto V = [&](int x, int y, int z) { return &s.vel[3 * (z * s.slab + (long)y * s.X + x)]; };
  for (int x = 0; x < s.X; x++) {
    for (int y = 0; y < s.Y; y++) {
      for (int c = 0; c < 3; c++) V(x, y, 0)[c] = 0;
      for (int c = 0; c < 3; c++) V(x, y, s.Z - 1)[c] = 0;
    }
    for (int z = 0; z < s.Z; z++) {
      for (int c = 0; c < 3; c++) V(x, 0, z)[c] = 0;
      s.intensity[z * s.slab + x] = 0;
      for (int c = 0; c < 3; c++) V(x, s.Y - 1, z)[c] = 0;
    }
  }
  for (int y = 0; y < s.Y; y++)
    for (int z = 0; z < s.Z; z++) {
      for (int c = 0; c < 3; c++) V(0, y, z)[c] = 0;
      for (int c = 0; c < 3; c++) V(s.X - 1, y, z)[c] = 0;
    }
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

void* orc_create(int X, int Y, int Z, double dh, const double origin[3],
                 const orc_params* p) {
  Scene* s = new Scene();
  s->X = X;
  s->Y = Y;
  s->Z = Z;
  s->slab = (long)X * Y;
  s->total = s->slab * Z;
  s->totalActive = (long)(X - 2) * (Y - 2) * (Z - 2);
  s->nCells = (long)(X - 1) * (Y - 1) * (Z - 1);
  s->dh = dh;
  s->dhInv = 1.0 / dh;
  for (int i = 0; i < 3; i++) s->o[i] = origin[i];
  s->p = *p;
  s->defaultDt = p->dt;
  s->dt = p->dt;
  // vorticity eps scaling  COUPLED_INTEGRATOR_3D.cpp:67-71 (float arithmetic)
  int maxRes = std::max(std::max(X, Y), Z);
  double scaling = 64.0f / maxRes;
  scaling = (scaling < 1.0f) ? 1.0f : scaling;
  s->vortEps = p->vorticity_eps / scaling;
  const long off[8] = {0, 1, X, X + 1, s->slab, s->slab + 1, s->slab + X, s->slab + X + 1};
  for (int i = 0; i < 8; i++) s->off[i] = off[i];
  s->vel.assign(3 * s->total, 0.0);
  s->velws.assign(3 * s->total, 0.0);
  s->vort.assign(3 * s->total, 0.0);
  s->intensity.assign(s->total, 0.0);
  s->ws.assign(s->total, 0.0);
  s->solidMass.assign(s->total, 0.0);
  s->nodeFrac.assign(s->total, 0.0);
  s->qfracSum.assign(8 * s->total, 0.0);
  s->sidx.assign(s->total, -1);
  // _gridIdxToSolutionIdx  COUPLED_INTEGRATOR_3D.cpp:123-131
  s->sol.assign(s->total, -1);
  int idx = 0;
  for (int z = 1; z < Z - 1; z++)
    for (int y = 1; y < Y - 1; y++)
      for (int x = 1; x < X - 1; x++) s->sol[z * s->slab + (long)y * X + x] = idx++;
  return s;
}

void orc_destroy(void* h) { delete static_cast<Scene*>(h); }

int orc_add_solid(void* h, double lambda, double mu, double density, int rX,
                  int rY, int rZ, double rdh, const double rorigin[3],
                  const double* rest_mass, const double rot[9],
                  const double trans[3], const double scale[3],
                  const double v0[3]) {
  Scene& s = *static_cast<Scene*>(h);
  Solid so;
  so.lambda = lambda;
  so.mu = mu;
  so.qcm = density * s.dh * s.dh * s.dh * 0.125;
  so.rg = Geom{rX, rY, rZ, rdh, {rorigin[0], rorigin[1], rorigin[2]}};
  so.rest.assign(rest_mass, rest_mass + (long)rX * rY * rZ);
  so.X.assign(3 * s.total, 0.0);
  so.u.assign(3 * s.total, 0.0);
  so.vel.assign(3 * s.total, 0.0);
  so.velws.assign(3 * s.total, 0.0);
  so.mass.assign(s.total, 0.0);
  so.detFInv.assign(s.total, 0.0);
  so.phi.assign(s.total, 0.0);
  so.valid.assign(s.nCells, 0);
  so.qfrac.assign(8 * s.total, 0.0);
  so.dt = s.p.dt;
  // setInitialState  DEFOR_SOLID_3D.cpp:126-152
  for (int z = 0; z < s.Z; z++)
    for (int y = 0; y < s.Y; y++)
      for (int x = 0; x < s.X; x++) {
        long n = z * s.slab + (long)y * s.X + x;
        double pt[3] = {(double)x, (double)y, (double)z};
        for (int c = 0; c < 3; c++) {
          pt[c] *= s.dh;
          pt[c] += s.o[c];
          pt[c] -= trans[c];
        }
        double q[3];
        for (int r = 0; r < 3; r++)  // rotation^T * point
          q[r] = (rot[r] * pt[0] + rot[3 + r] * pt[1]) + rot[6 + r] * pt[2];
        for (int c = 0; c < 3; c++) so.X[3 * n + c] = q[c] / scale[c];
      }
  for (int z = 0; z < s.Z; z++)
    for (int y = 0; y < s.Y; y++)
      for (int x = 0; x < s.X; x++) {
        long n = z * s.slab + (long)y * s.X + x;
        double pos[3] = {x * s.dh, y * s.dh, z * s.dh};
        for (int c = 0; c < 3; c++) {
          pos[c] += s.o[c];
          so.u[3 * n + c] = pos[c] - so.X[3 * n + c];
        }
      }
  // setInitialVelocity  DEFOR_SOLID_3D.cpp:153-155
  for (long i = 0; i < s.total; i++)
    for (int c = 0; c < 3; c++) so.vel[3 * i + c] = v0[c];
  s.solids.push_back(std::move(so));
  return (int)s.solids.size() - 1;
}

void orc_init(void* h, double carpet_thickness) {
  Scene& s = *static_cast<Scene*>(h);
  // DEFOR_SOLID_3D::init  DEFOR_SOLID_3D.cpp:269-331 (particles and mesh
  // are outside the drop-in path)
  for (Solid& so : s.solids) {
    for (int c = 0; c < 3; c++) {
      so.bMin[c] = 0;
      so.tMin[c] = 0;
    }
    so.bMax[0] = so.tMax[0] = s.X;
    so.bMax[1] = so.tMax[1] = s.Y;
    so.bMax[2] = so.tMax[2] = s.Z;
    computeSolidMass(s, so, false);
    computeSDF(s, so);
    if (s.p.flip) initParticles(s, so, s.solids[0].qcm / (s.dh * s.dh * s.dh * 0.125));
  }
  gather(s);
  for (Solid& so : s.solids) recoverVelocity(s, so);
  // smoke carpet  rubberBall/main.cpp:80-90
  int layers = (int)(carpet_thickness / s.dh);
  for (int z = 1; z < s.Z - 1; z++)
    for (int y = 1; y < layers; y++)
      for (int x = 1; x < s.X - 1; x++) s.intensity[z * s.slab + (long)y * s.X + x] = 0.8f;
}

void* orc_field(void* h, int id, int solid, size_t* n) {
  Scene& s = *static_cast<Scene*>(h);
  auto ret = [&](vector<double>& v) -> void* {
    *n = v.size();
    return v.data();
  };
  switch (id) {
    case ORC_VELOCITY: return ret(s.vel);
    case ORC_INTENSITY: return ret(s.intensity);
    case ORC_SOLID_MASS: return ret(s.solidMass);
    case ORC_NODE_FRACTION: return ret(s.nodeFrac);
    case ORC_LAMBDA:
      if (!s.lambdaInit) {
        s.lambda.assign(s.nCells, 0.0);
        s.lambdaInit = true;
      }
      return ret(s.lambda);
    case ORC_VEL_WORKSPACE: return ret(s.velws);
    case ORC_SOLID_INDEX: *n = s.sidx.size(); return s.sidx.data();
    case ORC_V0: return ret(s.v0);
    case ORC_V1: return ret(s.v1);
    case ORC_MINV: return ret(s.minv);
    case ORC_PRESSURE_RHS: return ret(s.b);
    case ORC_DIV_RHS: return ret(s.c);
    case ORC_JT_LAMBDA: return ret(s.g);
    case ORC_SOLID_FORCE: return ret(s.sForce);
    case ORC_SOLID_MV: return ret(s.sMv);
    case ORC_DAMPED_MASS: return ret(s.sDamped);
    case ORC_UNIT_FORCE: return ret(s.unitForce);
    case ORC_MV: return ret(s.mv);
    case ORC_SOLID_V0: return ret(s.sV0);
    case ORC_SOLID_DV: return ret(s.sDv);
    case ORC_SYS_DIAG: return ret(s.Adiag);
    case ORC_PRESSURE_DIAG: return ret(s.D);
    case ORC_VORTICITY: return ret(s.vort);
    case ORC_WORKSPACE: return ret(s.ws);
    default: break;
  }
  if (solid < 0 || solid >= (int)s.solids.size()) {
    *n = 0;
    return nullptr;
  }
  Solid& so = s.solids[solid];
  switch (id) {
    case ORC_S_U: return ret(so.u);
    case ORC_S_X: return ret(so.X);
    case ORC_S_VELOCITY: return ret(so.vel);
    case ORC_S_MASS: return ret(so.mass);
    case ORC_S_VALID: *n = so.valid.size(); return so.valid.data();
    case ORC_S_QFRAC: return ret(so.qfrac);
    case ORC_S_VEL_WORKSPACE: return ret(so.velws);
    case ORC_S_DETFINV: return ret(so.detFInv);
    case ORC_S_PHI: return ret(so.phi);
    default: break;
  }
  *n = 0;
  return nullptr;
}

int orc_num_solid_dofs(void* h) { return static_cast<Scene*>(h)->nSolid; }

void orc_get_boxes(void* h, int solid, int out[12]) {
  Solid& so = static_cast<Scene*>(h)->solids[solid];
  for (int c = 0; c < 3; c++) {
    out[c] = so.tMin[c];
    out[3 + c] = so.tMax[c];
    out[6 + c] = so.bMin[c];
    out[9 + c] = so.bMax[c];
  }
}
void orc_set_boxes(void* h, int solid, const int in[12]) {
  Solid& so = static_cast<Scene*>(h)->solids[solid];
  for (int c = 0; c < 3; c++) {
    so.tMin[c] = in[c];
    so.tMax[c] = in[3 + c];
    so.bMin[c] = in[6 + c];
    so.bMax[c] = in[9 + c];
  }
}

double orc_cfl(void* h, double alpha) {
  Scene& s = *static_cast<Scene*>(h);
  s.dt = cfl(s, alpha);
  return s.dt;
}
void orc_set_dt(void* h, double dt) { static_cast<Scene*>(h)->dt = dt; }
void orc_set_max_dt(void* h, double dt) { static_cast<Scene*>(h)->defaultDt = dt; }

static vector<Trip>& tripStore(void* h) {
  static thread_local std::vector<std::pair<void*, vector<Trip>>> store;
  for (auto& e : store)
    if (e.first == h) return e.second;
  store.push_back({h, {}});
  return store.back().second;
}

double orc_forces(void* h) {
  Scene& s = *static_cast<Scene*>(h);
  vector<Trip>& t = tripStore(h);
  t.clear();
  return forces(s, t);
}
void orc_assemble(void* h) {
  Scene& s = *static_cast<Scene*>(h);
  assemble(s, tripStore(h));
}
int orc_solve(void* h, orc_stats* st) { return solve(*static_cast<Scene*>(h), st); }
void orc_solid_copy_velocity(void* h) { solidCopyVelocity(*static_cast<Scene*>(h)); }
void orc_solid_finalize(void* h) { solidFinalize(*static_cast<Scene*>(h)); }
void orc_compute_solid_mass(void* h, int solid) {
  Scene& s = *static_cast<Scene*>(h);
  computeSolidMass(s, s.solids[solid], false);
}
long orc_solid_num_cells(void* h, int solid) {
  Scene& s = *static_cast<Scene*>(h);
  Solid& so = s.solids[solid];
  const long cx = s.X - 1, cxy = (long)(s.X - 1) * (s.Y - 1);
  long n = 0;
  for (int z = so.tMin[2]; z < so.tMax[2] - 1; z++)
    for (int y = so.tMin[1]; y < so.tMax[1] - 1; y++)
      for (int x = so.tMin[0]; x < so.tMax[0] - 1; x++)
        n += so.valid[x + y * cx + z * cxy] == 1;
  return n;
}
void orc_solid_set_deformation(void* h, int solid, long n, const double* F) {
  Solid& so = static_cast<Scene*>(h)->solids[solid];
  so.Fset.assign(F, F + 9 * n);
}
// solvePCR on the assembled solid system (CONJUGATE_RESIDUAL.inl:194-276)
// from x = 0, after orc_forces and orc_assemble
int orc_solid_cr(void* h, const double* b, double* x, double eps, int max_it, int* iters) {
  Scene& s = *static_cast<Scene*>(h);
  vector<double> bb(b, b + s.A.n), xx(s.A.n, 0.0);
  auto apply = [&](const vector<double>& in, vector<double>& out) { csrMul(s.A, in, out); };
  bool c = pcr(apply, s.Adiag, bb, xx, eps, max_it, iters, nullptr);
  std::memcpy(x, xx.data(), sizeof(double) * xx.size());
  return c ? 0 : 1;
}
void orc_init_particles(void* h, int solid) {
  Scene& s = *static_cast<Scene*>(h);
  initParticles(s, s.solids[solid], s.solids[0].qcm / (s.dh * s.dh * s.dh * 0.125));
}
void orc_grid_to_particle(void* h, int solid) {
  Scene& s = *static_cast<Scene*>(h);
  gridToParticle(s, s.solids[solid]);
}
void orc_particle_to_grid(void* h, int solid) {
  Scene& s = *static_cast<Scene*>(h);
  particleToGrid(s, s.solids[solid]);
}
long orc_num_particles(void* h, int solid) {
  return (long)static_cast<Scene*>(h)->solids[solid].pmass.size();
}
void orc_get_particles(void* h, int solid, double* pos, double* vel) {
  Solid& so = static_cast<Scene*>(h)->solids[solid];
  if (pos) std::copy(so.ppos.begin(), so.ppos.end(), pos);
  if (vel) std::copy(so.pvel.begin(), so.pvel.end(), vel);
}
void orc_set_particles(void* h, int solid, long n, const double* pos, const double* vel) {
  Scene& s = *static_cast<Scene*>(h);
  Solid& so = s.solids[solid];
  so.ppos.assign(pos, pos + 3 * n);
  so.pvel.assign(vel, vel + 3 * n);
  const double density = s.solids[0].qcm / (s.dh * s.dh * s.dh * 0.125);
  so.pmass.assign(n, density * (s.dh * s.dh * s.dh / s.p.particles_per_cell));
}
void orc_extend_solid_velocity(void* h, int solid) {
  Scene& s = *static_cast<Scene*>(h);
  extendSolidVelocity(s, s.solids[solid]);
}
double orc_compute_F(void* h, int solid, long index, double* F) {
  Scene& s = *static_cast<Scene*>(h);
  return computeF(s, s.solids[solid], index, F);
}
void orc_compute_solid_mass_sdf(void* h, int solid) {
  Scene& s = *static_cast<Scene*>(h);
  computeSolidMass(s, s.solids[solid], true);
}
void orc_solid_sdf(void* h, int solid) {
  Scene& s = *static_cast<Scene*>(h);
  computeSDF(s, s.solids[solid]);
}
void orc_update_boxes(void* h, int solid) {
  Scene& s = *static_cast<Scene*>(h);
  updateBoxes(s, s.solids[solid]);
}
void orc_gather(void* h) { gather(*static_cast<Scene*>(h)); }
void orc_advect_velocity(void* h) { advectVelocity(*static_cast<Scene*>(h)); }
void orc_advect_intensity(void* h) { advectIntensity(*static_cast<Scene*>(h)); }
void orc_recover_and_bc(void* h) { recoverAndBC(*static_cast<Scene*>(h)); }

int orc_step(void* h, double alpha, orc_stats* st) {
  Scene& s = *static_cast<Scene*>(h);
  double t0 = nowSec();
  s.dt = cfl(s, alpha < 0 ? 1.8 : alpha);
  s.stageSec[0] += nowSec() - t0;
  double e = orc_forces(h);
  orc_assemble(h);
  int rc = solve(s, st);
  solidCopyVelocity(s);
  if (s.p.flip) {
    // finalizeSolution :833-854: G2P, P2G and the velocity extension
    for (Solid& so : s.solids) gridToParticle(s, so);
    for (Solid& so : s.solids) particleToGrid(s, so);
    for (Solid& so : s.solids) extendSolidVelocity(s, so);
  }
  solidFinalize(s);
  double t1 = nowSec();
  // finalizeSolution :857-879: computeSolidMass(true) of every solid, then
  // computeSDF (and the new particles) of every solid (surface meshes are
  // not kept here)
  for (Solid& so : s.solids) computeSolidMass(s, so, true);
  for (Solid& so : s.solids) {
    computeSDF(s, so);
    if (s.p.flip) initParticles(s, so, s.solids[0].qcm / (s.dh * s.dh * s.dh * 0.125));
  }
  gather(s);
  s.stageSec[10] += nowSec() - t1;
  advectVelocity(s);
  advectIntensity(s);
  recoverAndBC(s);
  if (st) {
    st->dt = s.dt;
    st->energy = e;
  }
  return rc;
}

void orc_pressure_apply(void* h, const double* x, double* y) {
  pressureApply(*static_cast<Scene*>(h), x, y);
}
void orc_pressure_diag(void* h, double* d) { pressureDiag(*static_cast<Scene*>(h), d); }
void orc_solid_apply(void* h, const double* x, double* y) {
  Scene& s = *static_cast<Scene*>(h);
  vector<double> in(x, x + s.A.n), out;
  csrMul(s.A, in, out);
  std::memcpy(y, out.data(), sizeof(double) * out.size());
}
int orc_pressure_cr(void* h, const double* b, double* x, double eps, int max_it,
                    int* iters) {
  Scene& s = *static_cast<Scene*>(h);
  vector<double> bb(b, b + s.nCells), xx(x, x + s.nCells), D(s.nCells);
  pressureDiag(s, D.data());
  auto Sapply = [&](const vector<double>& in, vector<double>& out) {
    out.resize(s.nCells);
    pressureApply(s, in.data(), out.data());
  };
  bool c = pcr(Sapply, D, bb, xx, eps, max_it, iters, nullptr);
  std::memcpy(x, xx.data(), sizeof(double) * s.nCells);
  return c ? 0 : 1;
}

void orc_apply_J(void* h, const double* v, double* out) {
  Scene& s = *static_cast<Scene*>(h);
  vector<double> vv(v, v + 3 * s.totalActive), o;
  applyJ(s, vv, o);
  std::memcpy(out, o.data(), sizeof(double) * o.size());
}
void orc_apply_negJT(void* h, const double* lam, double* out) {
  Scene& s = *static_cast<Scene*>(h);
  vector<double> l(lam, lam + s.nCells), o;
  applyNegJT(s, l, o);
  std::memcpy(out, o.data(), sizeof(double) * o.size());
}

int orc_pcr_dense(int n, const double* A, const double* diag, const double* b,
                  double* x, double eps, int max_it, int* iters) {
  vector<double> d(diag, diag + n), bb(b, b + n), xx(x, x + n);
  auto apply = [&](const vector<double>& in, vector<double>& out) {
    out.assign(n, 0.0);
    for (int i = 0; i < n; i++) {
      double acc = 0;
      for (int j = 0; j < n; j++) acc += A[(long)i * n + j] * in[j];
      out[i] = acc;
    }
  };
  bool c = pcr(apply, d, bb, xx, eps, max_it, iters, nullptr);
  std::memcpy(x, xx.data(), sizeof(double) * n);
  return c ? 1 : 0;
}

void orc_corotation(double lambda, double mu, const double F[9], const double dF[9],
                    double out[36], double* energy, double dP_out[9]) {
  Corot m;
  m.lambda = lambda;
  m.mu = mu;
  corotInit(m, F);
  std::memcpy(out, m.R, sizeof(m.R));
  std::memcpy(out + 9, m.S, sizeof(m.S));
  std::memcpy(out + 18, m.L, sizeof(m.L));
  corotP(m, out + 27);
  if (energy) *energy = corotEnergy(m);
  if (dF && dP_out) corotDP(m, dF, dP_out);
}

double orc_interp_scalar(const double* f, int X, int Y, int Z, double dh,
                         const double o[3], const double p[3]) {
  Geom g{X, Y, Z, dh, {o[0], o[1], o[2]}};
  return interpS(f, g, p);
}
void orc_interp_vec3(const double* f, int X, int Y, int Z, double dh,
                     const double o[3], const double p[3], double out[3]) {
  Geom g{X, Y, Z, dh, {o[0], o[1], o[2]}};
  interpV(f, g, p, out);
}
// DEFOR_SOLID_3D::advectSurfaceMesh (DEFOR_SOLID_3D.cpp:1125-1180) on one
// solid's fields (velocity and displacement, AoS over the X*Y*Z grid):
// every vertex moves with the interpolated velocity for dt, then takes up to
// 10 Newton steps vertex -= F (vertex - u(vertex) - rest) with F from the
// computeF of the cell holding it (:671-682), then is clamped into the box.
void orc_advect_surface_mesh(int X, int Y, int Z, double dh, const double o[3],
                             const double* vel, const double* u, double dt, int nv,
                             const double* rest, double* vert) {
  Geom g{X, Y, Z, dh, {o[0], o[1], o[2]}};
  const long S = (long)X * Y;
  const long off[8] = {0, 1, X, X + 1, S, S + 1, S + X, S + X + 1};
  const double gg = 0.25 * (1.0 / dh);
  for (int x = 0; x < nv; x++) {
    double p[3] = {vert[3 * x], vert[3 * x + 1], vert[3 * x + 2]};
    double v[3];
    interpV(vel, g, p, v);
    for (int c = 0; c < 3; c++) p[c] += v[c] * dt;
    double residual = 1;
    int count = 0;
    while (residual > 1e-4 && count++ < 10) {
      double uu[3], prest[3], diff[3];
      interpV(u, g, p, uu);
      for (int c = 0; c < 3; c++) prest[c] = p[c] - uu[c];
      for (int c = 0; c < 3; c++) diff[c] = prest[c] - rest[3 * x + c];
      residual = std::sqrt(diff[0] * diff[0] + (diff[1] * diff[1] + diff[2] * diff[2]));
      int i = (int)((p[0] - o[0]) / dh);
      if (i < 0) i = 0;
      if (i > X - 2) i = X - 2;
      int j = (int)((p[1] - o[1]) / dh);
      if (j < 0) j = 0;
      if (j > Y - 2) j = Y - 2;
      int k = (int)((p[2] - o[2]) / dh);
      if (k < 0) k = 0;
      if (k > Z - 2) k = Z - 2;
      const long index = (long)k * S + (long)j * X + i;
      double Finv[9], F[9];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          double acc = 0;
          for (int q = 0; q < 8; q++) {
            int sg = c == 0 ? sgnX(q) : (c == 1 ? sgnY(q) : sgnZ(q));
            double t = u[3 * (index + off[q]) + r] * (sg * gg);
            acc = (q == 0) ? t : acc + t;
          }
          Finv[3 * r + c] = ((r == c) ? 1.0 : 0.0) - acc;
        }
      if (std::fabs(det3(Finv)) > 1e-4) {
        inv3(Finv, F);
      } else {
        for (int q = 0; q < 9; q++) F[q] = (q % 4 == 0) ? 1.0 : 0.0;
      }
      double Fd[3];
      for (int r = 0; r < 3; r++) Fd[r] = (F[3 * r] * diff[0] + F[3 * r + 1] * diff[1]) + F[3 * r + 2] * diff[2];
      for (int r = 0; r < 3; r++) p[r] -= Fd[r];
    }
    const int res[3] = {X, Y, Z};
    for (int c = 0; c < 3; c++) {
      p[c] = std::max(o[c], p[c]);
      p[c] = std::min(o[c] + dh * (res[c] - 1), p[c]);
      vert[3 * x + c] = p[c];
    }
  }
}

// TRIANGLE_BVH (src/geometry/TRIANGLE_BVH.cpp): a bounding-volume tree over
// triangle boxes (leaf: the triangle's vertex bounds, inner: the union of
// its children, :163-177), split at the median of the triangle centres
// along the widest axis of the node, and the reference's recursive collide
// (:179-192) with aabb::overlaps (aabb.cpp:121-133).  The split rule is this
// checker's own (the reference splits at an axis-aligned plane, aap.cpp):
// collide does not depend on it, which is what the GPU's all-pairs test
// relies on.
struct BvhNode {
  double box[6];
  int left = -1, right = -1;
};
static bool boxOverlap(const double* a, const double* b) {
  if (a[0] > b[3] || a[1] > b[4] || a[2] > b[5]) return false;
  if (a[3] < b[0] || a[4] < b[1] || a[5] < b[2]) return false;
  return true;
}
static int bvhBuild(std::vector<BvhNode>& nodes, const std::vector<std::array<double, 6>>& tb,
                    std::vector<int>& ids, int lo, int hi) {
  BvhNode n;
  for (int c = 0; c < 3; c++) {
    n.box[c] = 1e300;
    n.box[3 + c] = -1e300;
  }
  for (int k = lo; k < hi; k++)
    for (int c = 0; c < 3; c++) {
      n.box[c] = std::min(n.box[c], tb[ids[k]][c]);
      n.box[3 + c] = std::max(n.box[3 + c], tb[ids[k]][3 + c]);
    }
  const int me = (int)nodes.size();
  nodes.push_back(n);
  if (hi - lo > 1) {
    int axis = 0;
    for (int c = 1; c < 3; c++)
      if (n.box[3 + c] - n.box[c] > n.box[3 + axis] - n.box[axis]) axis = c;
    const int mid = (lo + hi) / 2;
    std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi, [&](int a, int b) {
      return tb[a][axis] + tb[a][3 + axis] < tb[b][axis] + tb[b][3 + axis];
    });
    const int l = bvhBuild(nodes, tb, ids, lo, mid);
    const int r = bvhBuild(nodes, tb, ids, mid, hi);
    nodes[me].left = l;
    nodes[me].right = r;
  }
  return me;
}
static std::vector<BvhNode> bvhOf(int nt, const int* tri, const double* v) {
  std::vector<std::array<double, 6>> tb(nt);
  for (int t = 0; t < nt; t++)
    for (int c = 0; c < 3; c++) {
      tb[t][c] = tb[t][3 + c] = v[3 * tri[3 * t] + c];
      for (int y = 1; y < 3; y++) {
        tb[t][c] = std::min(tb[t][c], v[3 * tri[3 * t + y] + c]);
        tb[t][3 + c] = std::max(tb[t][3 + c], v[3 * tri[3 * t + y] + c]);
      }
    }
  std::vector<int> ids(nt);
  for (int t = 0; t < nt; t++) ids[t] = t;
  std::vector<BvhNode> nodes;
  if (nt > 0) bvhBuild(nodes, tb, ids, 0, nt);
  return nodes;
}
static bool bvhCollide(const std::vector<BvhNode>& A, int a, const std::vector<BvhNode>& B, int b) {
  const BvhNode& x = A[a];
  const BvhNode& y = B[b];
  if (!boxOverlap(x.box, y.box)) return false;
  if (x.left < 0 && y.left < 0) return true;
  if (x.left < 0) return bvhCollide(A, a, B, y.left) || bvhCollide(A, a, B, y.right);
  if (y.left >= 0)
    return bvhCollide(A, x.left, B, y.left) || bvhCollide(A, x.left, B, y.right) ||
           bvhCollide(A, x.right, B, y.left) || bvhCollide(A, x.right, B, y.right);
  return bvhCollide(A, x.left, B, b) || bvhCollide(A, x.right, B, b);
}

// the broad phase of stepImplicit (COUPLED_INTEGRATOR_3D.cpp:384-411) on n
// solid meshes (mesh s: nv[s] vertices at v + vo[s] (3 doubles each), nt[s]
// triangles at tri + to[s]): 0.6 when a root box leaves the wall-safe box
// [o + 2dh, o + (res - 3)dh] (DEFOR_SOLID_3D.cpp:64-66) or two trees
// collide, else 1.8
double orc_mesh_broad_phase(int n, const int* nt, const long* vo, const long* to,
                            const double* v, const int* tri, int X, int Y, int Z, double dh,
                            const double o[3]) {
  std::vector<std::vector<BvhNode>> trees;
  for (int s = 0; s < n; s++) trees.push_back(bvhOf(nt[s], tri + to[s], v + vo[s]));
  const int res[3] = {X, Y, Z};
  for (int s = 0; s < n; s++) {
    const double* b = trees[s][0].box;
    for (int k = 0; k < 3; k++) {
      const double lo = o[k] + 2.0 * dh, hi = o[k] + (res[k] - 3) * dh;
      if (b[k] < lo || b[k] > hi || b[3 + k] < lo || b[3 + k] > hi) return 0.6;
    }
  }
  for (int a = 0; a + 1 < n; a++)
    for (int b = a + 1; b < n; b++)
      if (bvhCollide(trees[a], 0, trees[b], 0)) return 0.6;
  return 1.8;
}

double orc_last_stage_seconds(void* h, int stage) {
  Scene& s = *static_cast<Scene*>(h);
  if (stage < 0 || stage >= 16) return 0;
  return s.stageSec[stage];
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_sphere_rest_dims(double radius, double mdh, int dims[3], double origin[3]) {
  double bmin = -radius - mdh * 10, bmax = radius + mdh * 10;
  for (int i = 0; i < 3; i++) {
    dims[i] = (int)((bmax - bmin) / mdh);
    origin[i] = bmin;
  }
}
void orc_sphere_rest_fill(double radius, double mdh, double density, double* out) {
  int d[3];
  double o[3];
  orc_sphere_rest_dims(radius, mdh, d, o);
  long idx = 0;
  for (int z = 0; z < d[2]; z++)
    for (int y = 0; y < d[1]; y++)
      for (int x = 0; x < d[0]; x++, idx++) {
        double p[3] = {x * mdh + o[0], y * mdh + o[1], z * mdh + o[2]};
        double sd = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - radius;
        out[idx] = (sd <= mdh / 2) ? density : 0.0;
      }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// fast sweeping (TEST INFRASTRUCTURE): computeSDF, sweepPhi, solveDistance,
// sweepOnePass, sweepVectorField, extendVectorField, restated in the
// reference's loop order (DEFOR_SOLID_3D.cpp:216-267, 369-574)
// ---------------------------------------------------------------------------
// solveDistance (:369-395): the grid Eikonal update from the three upwind
// neighbours, sorted; lowers v only
double orc_solve_distance(double p, double q, double r, double v) {
  double d[3];
  d[0] = std::min(p, std::min(q, r));
  d[2] = std::max(p, std::max(q, r));
  d[1] = p + q + r - d[0] - d[2];
  double b[3], c[3];
  b[0] = d[0];
  c[0] = d[0] * d[0] - 1.0;
  for (int x = 1; x < 3; x++) {
    b[x] = d[x] + b[x - 1];
    c[x] = d[x] * d[x] + c[x - 1];
  }
  int i = 2;
  double discrim = b[i] * b[i] - (double)(i + 1) * c[i];
  double nd = (b[i] + std::sqrt(discrim)) / (double)(i + 1);
  // the reference steps i down while the candidate is invalid; for finite
  // inputs i = 0 always yields a valid one (discrim there is d0^2 - c0 = 1
  // up to rounding), so i never reaches -1
  while ((discrim < 0.0 || nd < d[i]) && i > 0) {
    i--;
    discrim = b[i] * b[i] - (double)(i + 1) * c[i];
    if (discrim > 0.0) nd = (b[i] + std::sqrt(discrim)) / (double)(i + 1);
  }
  return nd < v ? nd : v;
}

namespace {
struct SweepGrid {
  int X, Y, Z;
  long at(int i, int j, int k) const { return i + (long)X * (j + (long)Y * k); }
};
}  // namespace

// sweepPhi (:397-454): eight ordered sweeps over the box, each updating the
// cells without mass from the neighbours already visited
void orc_sweep_phi(int X, int Y, int Z, const int box[6], const double* mass, double* phi) {
  const SweepGrid g{X, Y, Z};
  const int x0 = box[0], y0 = box[1], z0 = box[2], x1 = box[3], y1 = box[4], z1 = box[5];
  for (int s = 0; s < 8; s++) {
    const int dx = (s & 2) ? -1 : 1, dy = (s & 1) ? -1 : 1, dz = (s & 4) ? -1 : 1;
    // the reference's bounds: ascending from min + 1 to max - 1, descending
    // from max - 2 down to min (inclusive)
    const int ia = dx > 0 ? x0 + 1 : x1 - 2, ib = dx > 0 ? x1 : x0 - 1;
    const int ja = dy > 0 ? y0 + 1 : y1 - 2, jb = dy > 0 ? y1 : y0 - 1;
    const int ka = dz > 0 ? z0 + 1 : z1 - 2, kb = dz > 0 ? z1 : z0 - 1;
    for (int k = ka; dz > 0 ? k < kb : k > kb; k += dz)
      for (int j = ja; dy > 0 ? j < jb : j > jb; j += dy)
        for (int i = ia; dx > 0 ? i < ib : i > ib; i += dx) {
          const long c = g.at(i, j, k);
          if (mass[c] == 0)
            phi[c] = orc_solve_distance(phi[g.at(i - dx, j, k)], phi[g.at(i, j - dy, k)],
                                        phi[g.at(i, j, k - dz)], phi[c]);
        }
  }
}

// computeSDF (:216-267): -0.5 on cells with mass, a large distance elsewhere,
// the box of the massive cells widened by the tight border and by the
// extrapolation cells, then two rounds of sweepPhi
void orc_compute_sdf(int X, int Y, int Z, const double* mass, int tight_border, int extrap,
                     double* phi, int box[6], int tight[6]) {
  const double large = (double)(X + Y + Z + 2);
  int mn[3] = {X, Y, Z}, mx[3] = {0, 0, 0};
  long index = 0;
  for (int z = 0; z < Z; z++)
    for (int y = 0; y < Y; y++)
      for (int x = 0; x < X; x++, index++) {
        phi[index] = large;
        if (mass[index] > 0) {
          phi[index] = -0.5;
          const int p[3] = {x, y, z};
          for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], p[a]);
            mx[a] = std::max(mx[a], p[a]);
          }
        }
      }
  const int res[3] = {X, Y, Z};
  for (int a = 0; a < 3; a++) {
    tight[a] = std::max(mn[a] - tight_border, 0);
    tight[3 + a] = std::min(mx[a] + tight_border, res[a]);
    box[a] = std::max(mn[a] - extrap, 0);
    box[3 + a] = std::min(mx[a] + extrap, res[a]);
  }
  for (int r = 0; r < 2; r++) orc_sweep_phi(X, Y, Z, box, mass, phi);
}

// sweepOnePass (:457-519) over field (3 per node): i0 -> i1 exclusive with
// the direction given by their order; solvableX/Y/Z start true and, once
// cleared, stay cleared for the rest of the pass
static void sweepOnePass(const SweepGrid& g, const double* mass, const double* phi,
                         double* field, int i0, int i1, int j0, int j1, int k0, int k1) {
  const int di = i0 < i1 ? 1 : -1, dj = j0 < j1 ? 1 : -1, dk = k0 < k1 ? 1 : -1;
  bool sx = true, sy = true, sz = true;
  for (int k = k0; k != k1; k += dk)
    for (int j = j0; j != j1; j += dj)
      for (int i = i0; i != i1; i += di) {
        const long c = g.at(i, j, k);
        if (mass[c] != 0) continue;
        double dp = phi[c] - phi[g.at(i - di, j, k)];
        const double dpn = phi[c] - phi[g.at(i + di, j, k)];
        if (dp < 0) {
          if (dpn >= 0) continue;
          sx = false;
        }
        double dq = phi[c] - phi[g.at(i, j - dj, k)];
        const double dqn = phi[c] - phi[g.at(i, j + dj, k)];
        if (dq < 0) {
          if (dqn >= 0) continue;
          sy = false;
        }
        double dr = phi[c] - phi[g.at(i, j, k - dk)];
        const double drn = phi[c] - phi[g.at(i, j, k + dk)];
        if (dr < 0) {
          if (drn >= 0) continue;
          sz = false;
        }
        if (!sx) dp = 0;
        if (!sy) dq = 0;
        if (!sz) dr = 0;
        double alpha, beta;
        if (dp + dq + dr < 1e-3) {
          alpha = 0.333;
          beta = 0.333;
        } else {
          alpha = dp / (dp + dq + dr);
          beta = dq / (dp + dq + dr);
        }
        const long a = 3 * g.at(i - di, j, k), b = 3 * g.at(i, j - dj, k),
                   e = 3 * g.at(i, j, k - dk);
        for (int q = 0; q < 3; q++)
          field[3 * c + q] = alpha * field[a + q] + beta * field[b + q] +
                             (1 - alpha - beta) * field[e + q];
      }
}

// sweepVectorField (:521-564): eight passes, then the box faces copied from
// the layer inside (z faces, then y, then x) where the node has no mass
static void sweepVectorField(const SweepGrid& g, const int box[6], const double* mass,
                             const double* phi, double* field) {
  const int x0 = box[0], y0 = box[1], z0 = box[2], x1 = box[3], y1 = box[4], z1 = box[5];
  sweepOnePass(g, mass, phi, field, x0 + 1, x1 - 1, y0 + 1, y1 - 1, z0 + 1, z1 - 1);
  sweepOnePass(g, mass, phi, field, x0 + 1, x1 - 1, y1 - 2, y0, z0 + 1, z1 - 1);
  sweepOnePass(g, mass, phi, field, x1 - 2, x0, y0 + 1, y1 - 1, z0 + 1, z1 - 1);
  sweepOnePass(g, mass, phi, field, x1 - 2, x0, y1 - 2, y0, z0 + 1, z1 - 1);
  sweepOnePass(g, mass, phi, field, x0 + 1, x1 - 1, y0 + 1, y1 - 1, z1 - 2, z0);
  sweepOnePass(g, mass, phi, field, x0 + 1, x1 - 1, y1 - 2, y0, z1 - 2, z0);
  sweepOnePass(g, mass, phi, field, x1 - 2, x0, y0 + 1, y1 - 1, z1 - 2, z0);
  sweepOnePass(g, mass, phi, field, x1 - 2, x0, y1 - 2, y0, z1 - 2, z0);
  auto copy = [&](long to, long from) {
    if (mass[to] == 0)
      for (int q = 0; q < 3; q++) field[3 * to + q] = field[3 * from + q];
  };
  for (int i = x0; i < x1; i++)
    for (int j = y0; j < y1; j++) {
      copy(g.at(i, j, z0), g.at(i, j, z0 + 1));
      copy(g.at(i, j, z1 - 1), g.at(i, j, z1 - 2));
    }
  for (int i = x0; i < x1; i++)
    for (int k = z0; k < z1; k++) {
      copy(g.at(i, y0, k), g.at(i, y0 + 1, k));
      copy(g.at(i, y1 - 1, k), g.at(i, y1 - 2, k));
    }
  for (int j = y0; j < y1; j++)
    for (int k = z0; k < z1; k++) {
      copy(g.at(x0, j, k), g.at(x0 + 1, j, k));
      copy(g.at(x1 - 1, j, k), g.at(x1 - 2, j, k));
    }
}

void orc_extend_vector_field(int X, int Y, int Z, const int box[6], const double* mass,
                             const double* phi, double* field) {
  const SweepGrid g{X, Y, Z};
  for (int r = 0; r < 4; r++) sweepVectorField(g, box, mass, phi, field);
}
