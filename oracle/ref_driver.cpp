/*
 * ref_driver.cpp -- C entry points into the reference's own code, compiled
 * from the reference tree (see Makefile) into oracle/_ref/libesfref.so.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle against the
 * reference.  Nothing here re-implements reference behaviour; every entry
 * point forwards to a reference class:
 *   COROTATION            src/material/COROTATION.cpp
 *   CONJUGATE_RESIDUAL    include/linearalgebra/CONJUGATE_RESIDUAL.inl
 *   JACOBI_PRECONDITIONER include/linearalgebra/JACOBI_PRECONDITIONER.h
 *   FIELD_3D, VEC3F_FIELD_3D include/geometry/*.h (interpolate_value,
 *                         writeGz / readGz)
 *   SIMPLE_PARSER         src/util/SIMPLE_PARSER.cpp
 */
#include <SETTINGS.h>

#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include <geometry/FIELD_3D.h>
#include <geometry/VEC3F_FIELD_3D.h>
#include <linearalgebra/JACOBI_PRECONDITIONER.h>
#include <linearalgebra/MATRIX_UTIL.h>
#include <util/SIMPLE_PARSER.h>
// COROTATION keeps R, S, L private; expose them to read the decomposition.
#define private public
#include <material/COROTATION.h>
#undef private
#include <linearalgebra/CONJUGATE_RESIDUAL.h>

namespace {
struct DenseOp {
  int n;
  const double* A;  // row-major n x n
};
void denseMatVec(void* app, const VECTOR& in, VECTOR& out) {
  const DenseOp* op = static_cast<const DenseOp*>(app);
  out.resize(op->n);
  for (int i = 0; i < op->n; i++) {
    double acc = 0;
    for (int j = 0; j < op->n; j++) acc += op->A[(long)i * op->n + j] * in[j];
    out[i] = acc;
  }
}
}  // namespace

extern "C" {

// COROTATION::init, firstPiolaKirchhoff, strainEnergy, firstPKDifferential.
// out: R, S, L, P (row-major, 36 doubles)
void ref_corotation(double lambda, double mu, const double F[9], const double dF[9],
                    double out[36], double* energy, double dP_out[9]) {
  COROTATION m(lambda, mu);
  MATRIX3 Fm, dFm;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      Fm(r, c) = F[3 * r + c];
      dFm(r, c) = dF ? dF[3 * r + c] : 0.0;
    }
  m.init(Fm);
  MATRIX3 P = m.firstPiolaKirchhoff();
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      out[3 * r + c] = m._R(r, c);
      out[9 + 3 * r + c] = m._S(r, c);
      out[18 + 3 * r + c] = m._L(r, c);
      out[27 + 3 * r + c] = P(r, c);
    }
  if (energy) *energy = m.strainEnergy();
  if (dF && dP_out) {
    MATRIX3 dP = m.firstPKDifferential(dFm);
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) dP_out[3 * r + c] = dP(r, c);
  }
}

// FIELD_3D<double>::interpolate_value over npts points
void ref_interp_scalar(const double* f, int X, int Y, int Z, double dh,
                       const double o[3], int npts, const double* pts, double* out) {
  FIELD_3Df field(X, Y, Z, dh, VEC3F(o[0], o[1], o[2]));
  std::memcpy(field.data().data(), f, sizeof(double) * (long)X * Y * Z);
  for (int i = 0; i < npts; i++)
    out[i] = field.interpolate_value(VEC3F(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
}

// VEC3F_FIELD_3D::interpolate_value over npts points (AoS field)
void ref_interp_vec3(const double* f, int X, int Y, int Z, double dh,
                     const double o[3], int npts, const double* pts, double* out) {
  VEC3F_FIELD_3D field(X, Y, Z, dh, VEC3F(o[0], o[1], o[2]));
  long n = (long)X * Y * Z;
  for (long i = 0; i < n; i++) field[i] = VEC3F(f[3 * i], f[3 * i + 1], f[3 * i + 2]);
  for (int i = 0; i < npts; i++) {
    VEC3F v = field.interpolate_value(VEC3F(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
    for (int c = 0; c < 3; c++) out[3 * i + c] = v[c];
  }
}

// CONJUGATE_RESIDUAL::solvePCR(SpMat&, b, x) on a triplet-built matrix
// (Eigen setFromTriplets sums duplicates).  Returns converged; *iters is
// the reference's iteration count, read back through numberOfIterations().
int ref_pcr_csr(int n, int nnz, const int* rows, const int* cols, const double* vals,
                const double* b, double* x, double eps, int max_it, int* iters) {
  std::vector<TRIPLET> t;
  t.reserve(nnz);
  for (int k = 0; k < nnz; k++) t.push_back(TRIPLET(rows[k], cols[k], vals[k]));
  SpMat A(n, n);
  A.setFromTriplets(t.begin(), t.end());
  A.makeCompressed();
  VECTOR diag = A.diagonal();
  JACOBI_PRECONDITIONER pre(diag);
  CONJUGATE_RESIDUAL<JACOBI_PRECONDITIONER> cr(&pre);
  cr.eps() = eps;
  cr.maxIteration() = max_it;
  VECTOR bb(n), xx(n);
  for (int i = 0; i < n; i++) bb[i] = b[i];
  std::streambuf* old = std::cout.rdbuf(nullptr);
  bool ok = cr.solvePCR(A, bb, xx);
  std::cout.rdbuf(old);
  for (int i = 0; i < n; i++) x[i] = xx[i];
  if (iters) *iters = ok ? cr.numberOfIterations().back() : max_it;
  return ok ? 1 : 0;
}

// CONJUGATE_RESIDUAL::solvePCR(b, x, matvec, app) with a dense operator and
// an explicit Jacobi diagonal; x is a warm start when sizes match.
int ref_pcr_matfree_dense(int n, const double* A, const double* diag, const double* b,
                          double* x, int warm, double eps, int max_it, int* iters) {
  VECTOR d(n), bb(n), xx;
  for (int i = 0; i < n; i++) {
    d[i] = diag[i];
    bb[i] = b[i];
  }
  if (warm) {
    xx.resize(n);
    for (int i = 0; i < n; i++) xx[i] = x[i];
  }
  JACOBI_PRECONDITIONER pre(d);
  CONJUGATE_RESIDUAL<JACOBI_PRECONDITIONER> cr(&pre);
  cr.eps() = eps;
  cr.maxIteration() = max_it;
  DenseOp op{n, A};
  std::streambuf* old = std::cout.rdbuf(nullptr);
  bool ok = cr.solvePCR(bb, xx, &denseMatVec, &op);
  std::cout.rdbuf(old);
  for (int i = 0; i < n; i++) x[i] = xx[i];
  if (iters) *iters = ok ? cr.numberOfIterations().back() : max_it;
  return ok ? 1 : 0;
}

// SIMPLE_PARSER
int ref_parse(const char* path) {
  std::streambuf* old = std::cout.rdbuf(nullptr);
  bool ok = SIMPLE_PARSER::parse(path);
  std::cout.rdbuf(old);
  return ok ? 1 : 0;
}
double ref_get_float(const char* key, double def) { return SIMPLE_PARSER::getFloat(key, def); }
int ref_get_int(const char* key, int def) { return SIMPLE_PARSER::getInt(key, def); }
int ref_get_string(const char* key, const char* def, char* out, int cap) {
  std::string s = SIMPLE_PARSER::getString(key, def);
  std::strncpy(out, s.c_str(), cap - 1);
  out[cap - 1] = 0;
  return (int)s.size();
}
void ref_get_vec3(const char* key, const double def[3], double out[3]) {
  VEC3F v = SIMPLE_PARSER::getVEC3F(key, VEC3F(def[0], def[1], def[2]));
  for (int i = 0; i < 3; i++) out[i] = v[i];
}

// FIELD_3D<double>::writeGz (components 1) or VEC3F_FIELD_3D::writeGz
// (components 3) of an x-fastest field
void ref_field_write_gz(const char* path, int X, int Y, int Z, double dh, const double o[3],
                        int components, const double* data) {
  const long n = (long)X * Y * Z;
  if (components == 1) {
    FIELD_3Df field(X, Y, Z, dh, VEC3F(o[0], o[1], o[2]));
    std::memcpy(field.data().data(), data, sizeof(double) * n);
    field.writeGz(std::string(path));
  } else {
    VEC3F_FIELD_3D field(X, Y, Z, dh, VEC3F(o[0], o[1], o[2]));
    for (long i = 0; i < n; i++) field[i] = VEC3F(data[3 * i], data[3 * i + 1], data[3 * i + 2]);
    field.writeGz(std::string(path));
  }
}

// FIELD_3D<double>::readGz / VEC3F_FIELD_3D::readGz: dims (3 ints), header
// (dh, origin), and up to cap doubles of the payload; returns the node count
long ref_field_read_gz(const char* path, int components, int dims[3], double hdr[4],
                       double* data, long cap) {
  if (components == 1) {
    FIELD_3Df field;
    field.readGz(std::string(path));
    dims[0] = field.xRes();
    dims[1] = field.yRes();
    dims[2] = field.zRes();
    hdr[0] = field.dh();
    for (int c = 0; c < 3; c++) hdr[1 + c] = field.origin()[c];
    const long n = (long)dims[0] * dims[1] * dims[2];
    for (long i = 0; i < n && i < cap; i++) data[i] = field[i];
    return n;
  }
  VEC3F_FIELD_3D field;
  field.readGz(std::string(path));
  dims[0] = field.xRes();
  dims[1] = field.yRes();
  dims[2] = field.zRes();
  hdr[0] = field.dh();
  for (int c = 0; c < 3; c++) hdr[1 + c] = field.origin()[c];
  const long n = (long)dims[0] * dims[1] * dims[2];
  for (long i = 0; i < n && 3 * i + 2 < cap; i++)
    for (int c = 0; c < 3; c++) data[3 * i + c] = field[i][c];
  return n;
}

}  // extern "C"
