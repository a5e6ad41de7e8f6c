// esf_interp.h -- trilinear interpolation of node fields with the
// reference's clamp rule, shared by the grid stages and the surface mesh.
#pragma once
#include "esf_internal.h"

namespace esf {

// ---------------------------------------------------------------------------
// trilinear interpolation with the reference's clamp rule
// (FIELD_3D.h:198-216, :368-382; MATH_UTILS.h:326-342)
// ---------------------------------------------------------------------------
struct FieldGeom {
  int X, Y, Z;
  double dh, o0, o1, o2;
  // 1 / dh, and whether dh is a power of two: then x * inv is x / dh to the
  // bit (the reciprocal is exact, and both round the same real quotient),
  // one multiplication in place of a division
  double inv;
  int pow2;
};
__host__ __device__ inline FieldGeom fieldGeom(int X, int Y, int Z, double dh, double o0,
                                               double o1, double o2) {
  int e = 0;
  const int pow2 = frexp(dh, &e) == 0.5 ? 1 : 0;
  return FieldGeom{X, Y, Z, dh, o0, o1, o2, 1.0 / dh, pow2};
}

__device__ __forceinline__ void bary(double x, const FieldGeom& g, int hi, int& i, double& f) {
  x = g.pow2 ? x * g.inv : x / g.dh;
  const double s = floor(x);
  i = (int)s;
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
__device__ __forceinline__ double lerp(double a, double b, double f) { return (1 - f) * a + f * b; }

// the cell (i, j, k) holding p and p's offsets in it; returns its node index
__device__ __forceinline__ long locate(const FieldGeom& g, const double* p, double& fx, double& fy,
                                       double& fz, int& i, int& j, int& k) {
  bary(p[0] - g.o0, g, g.X - 2, i, fx);
  bary(p[1] - g.o1, g, g.Y - 2, j, fy);
  bary(p[2] - g.o2, g, g.Z - 2, k, fz);
  return i + (long)j * g.X + (long)k * g.X * g.Y;
}
__device__ __forceinline__ long locate(const FieldGeom& g, const double* p, double& fx, double& fy,
                                       double& fz) {
  int i, j, k;
  return locate(g, p, fx, fy, fz, i, j, k);
}

__device__ __forceinline__ double interpS(const double* d, const FieldGeom& g, const double* p) {
  double fx, fy, fz;
  const long b = locate(g, p, fx, fy, fz);
  const long X = g.X, S = (long)g.X * g.Y;
  const double z0 = lerp(lerp(d[b], d[b + 1], fx), lerp(d[b + X], d[b + X + 1], fx), fy);
  const double z1 =
      lerp(lerp(d[b + S], d[b + S + 1], fx), lerp(d[b + S + X], d[b + S + X + 1], fx), fy);
  return lerp(z0, z1, fz);
}

__device__ __forceinline__ void interpV(const double* d, const FieldGeom& g, const double* p,
                                        double* out) {
  double fx, fy, fz;
  const long b = locate(g, p, fx, fy, fz);
  const long X = g.X, S = (long)g.X * g.Y;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double* q = d + c;
    const double z0 = lerp(lerp(q[3 * b], q[3 * (b + 1)], fx),
                           lerp(q[3 * (b + X)], q[3 * (b + X + 1)], fx), fy);
    const double z1 = lerp(lerp(q[3 * (b + S)], q[3 * (b + S + 1)], fx),
                           lerp(q[3 * (b + S + X)], q[3 * (b + S + X + 1)], fx), fy);
    out[c] = lerp(z0, z1, fz);
  }
}

// interpV at a point p inside node cell b (the cell locate(p) finds), with
// the cell's 24 corner values given: cv[c][k] = d[3 (b + off_k) + c] for the
// corners k in interpV's order (b, b+1, b+X, b+X+1, then the same + S).
// The same arithmetic as interpV, without its loads.
__device__ __forceinline__ void interpVCell(const double (&cv)[3][8], const FieldGeom& g,
                                            const double* p, double* out) {
  double fx, fy, fz;
  (void)locate(g, p, fx, fy, fz);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double z0 = lerp(lerp(cv[c][0], cv[c][1], fx), lerp(cv[c][2], cv[c][3], fx), fy);
    const double z1 = lerp(lerp(cv[c][4], cv[c][5], fx), lerp(cv[c][6], cv[c][7], fx), fy);
    out[c] = lerp(z0, z1, fz);
  }
}

__device__ __forceinline__ FieldGeom gridGeom(const Dims& d) {
  return fieldGeom(d.X, d.Y, d.Z, d.dh, d.o0, d.o1, d.o2);
}

}  // namespace esf
