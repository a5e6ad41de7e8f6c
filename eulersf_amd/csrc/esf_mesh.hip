// esf_mesh.hip -- the solids' surface meshes on the GPU: the per-vertex
// material-space projection of DEFOR_SOLID_3D::advectSurfaceMesh
// (DEFOR_SOLID_3D.cpp:1125-1180) and the broad phase of stepImplicit
// (COUPLED_INTEGRATOR_3D.cpp:384-411) over the meshes' triangle boxes.
//
// Broad phase.  The reference refits a TRIANGLE_BVH per mesh
// (TRIANGLE_BVH.cpp:163-177: a leaf's box is its triangle's vertex bounds,
// an inner box the union of its children) and asks (a) whether the root box
// lies inside the wall-safe box [origin + 2dh, origin + (res - 3)dh]
// (DEFOR_SOLID_3D.h:85, DEFOR_SOLID_3D.cpp:64-66) and (b) whether two trees
// collide (TRIANGLE_BVH.cpp:179-192).  Every inner box contains its leaves'
// boxes, so the recursion of (b) reaches a pair of overlapping leaves
// exactly when one exists: (b) is "some pair of triangle boxes overlaps"
// (closed intervals, aabb.cpp:121-133), whatever the tree's split planes.
// Here one thread per triangle of the first mesh tests it against every
// triangle box of the second.
#include <algorithm>

#include "esf_corot.h"
#include "esf_internal.h"
#include "esf_interp.h"

namespace esf {

// DEFOR_SOLID_3D::computeF (:671-682) on the cell whose lowest corner is
// node `index`: F^-1 = I - U_e G with U_e the displacements at the 8
// corners and G = +-1/(4 dh) (:304-310), the sums in ascending corner
// order; F = (F^-1)^-1, or I when |det F^-1| <= 1e-4
__device__ inline void cellF(const Dims& d, const double* u, long index, double* F) {
  const long off[8] = {0, 1, d.X, d.X + 1, d.slab, d.slab + 1, d.slab + d.X, d.slab + d.X + 1};
  const double g = 0.25 * d.dhInv;
  double Finv[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double acc = 0;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int bit = c == 0 ? (i & 1) : (c == 1 ? ((i >> 1) & 1) : (i >> 2));
        const double G = (bit ? 1 : -1) * g;
        const double t = u[3 * (index + off[i]) + r] * G;
        acc = (i == 0) ? t : acc + t;
      }
      Finv[3 * r + c] = ((r == c) ? 1.0 : 0.0) - acc;
    }
  if (fabs(det(Finv)) > 1e-4) {
    inverse(Finv, F);
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) F[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
}

// advectSurfaceMesh, one thread per vertex: move with the solid velocity
// for dt, then up to 10 Newton steps towards x - u(x) = X_rest (the step
// uses F of the cell holding x), then clamp into the domain
template <class GridT>
__device__ __forceinline__ void kAdvectMesh(const GridT& wg, Dims d, int nv,
                                            const double* __restrict__ rest, double* vert,
                                            const double* __restrict__ vel,
                                            const double* __restrict__ u, double dt) {
  const int x = wg.block() * blockDim.x + threadIdx.x;
  if (x >= nv) return;
  const FieldGeom g = gridGeom(d);
  double p[3] = {vert[3 * x], vert[3 * x + 1], vert[3 * x + 2]};
  double v[3];
  interpV(vel, g, p, v);
#pragma unroll
  for (int c = 0; c < 3; c++) p[c] = p[c] + v[c] * dt;
  const double r0 = rest[3 * x], r1 = rest[3 * x + 1], r2 = rest[3 * x + 2];
  double residual = 1;
  int count = 0;
  while (residual > 1e-4 && count++ < 10) {
    double uu[3];
    interpV(u, g, p, uu);
    const double d0 = (p[0] - uu[0]) - r0;
    const double d1 = (p[1] - uu[1]) - r1;
    const double d2 = (p[2] - uu[2]) - r2;
    // VEC3F::norm adds the squares as d0^2 + (d1^2 + d2^2)
    residual = sqrt(d0 * d0 + (d1 * d1 + d2 * d2));
    // cell index by truncation toward zero, then clamped (:1155-1163)
    const int i = min(max((int)((p[0] - d.o0) / d.dh), 0), d.X - 2);
    const int j = min(max((int)((p[1] - d.o1) / d.dh), 0), d.Y - 2);
    const int k = min(max((int)((p[2] - d.o2) / d.dh), 0), d.Z - 2);
    double F[9];
    cellF(d, u, (long)k * d.slab + (long)j * d.X + i, F);
#pragma unroll
    for (int r = 0; r < 3; r++) p[r] = p[r] - ((F[3 * r] * d0 + F[3 * r + 1] * d1) + F[3 * r + 2] * d2);
  }
  // std::max(origin, x), then std::min(origin + dh (res - 1), x)
  const double lo[3] = {d.o0, d.o1, d.o2};
  const int res[3] = {d.X, d.Y, d.Z};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    p[c] = lo[c] < p[c] ? p[c] : lo[c];
    const double hi = lo[c] + d.dh * (res[c] - 1);
    p[c] = p[c] < hi ? p[c] : hi;
    vert[3 * x + c] = p[c];
  }
}
ESF_BOUNDS(kAdvectMesh, 256);

// triangle boxes (min xyz, max xyz per triangle), TRIANGLE_BVH_NODE::refit
template <class GridT>
__device__ __forceinline__ void kTriBoxes(const GridT& wg, int nt, const int* __restrict__ tri,
                                          const double* __restrict__ vert, double* box) {
  const int t = wg.block() * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double mn[3], mx[3];
#pragma unroll
  for (int c = 0; c < 3; c++) mn[c] = mx[c] = vert[3 * tri[3 * t] + c];
#pragma unroll
  for (int y = 1; y < 3; y++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double q = vert[3 * tri[3 * t + y] + c];
      mn[c] = fmin(mn[c], q);
      mx[c] = fmax(mx[c], q);
    }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    box[6 * t + c] = mn[c];
    box[6 * t + 3 + c] = mx[c];
  }
}
ESF_BOUNDS(kTriBoxes, 1024);

// the union of a mesh's triangle boxes (the BVH root), one block
template <class GridT>
__device__ __forceinline__ void kMeshRoot(const GridT& wg, int nt, const double* __restrict__ box,
                                          double* root) {
  __shared__ double s[6][256];
  double m[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int t = threadIdx.x; t < nt; t += blockDim.x)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      m[c] = fmin(m[c], box[6 * t + c]);
      m[3 + c] = fmax(m[3 + c], box[6 * t + 3 + c]);
    }
#pragma unroll
  for (int c = 0; c < 6; c++) s[c][threadIdx.x] = m[c];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
#pragma unroll
      for (int c = 0; c < 6; c++)
        s[c][threadIdx.x] = c < 3 ? fmin(s[c][threadIdx.x], s[c][threadIdx.x + st])
                                  : fmax(s[c][threadIdx.x], s[c][threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x < 6) root[threadIdx.x] = s[threadIdx.x][0];
}
ESF_BOUNDS(kMeshRoot, 256);

// does some triangle box of mesh a overlap some triangle box of mesh b?
// (aabb::overlaps, closed intervals); a hit sets *flag.  Every triangle box
// of b lies inside b's root box, so a triangle of a that misses the root
// misses them all and skips the loop (as a BVH descent prunes at the root).
template <class GridT>
__device__ __forceinline__ void kMeshPairs(const GridT& wg, int na, const double* __restrict__ boxA,
                                           int nb, const double* __restrict__ boxB,
                                           const double* __restrict__ rootB, int* flag) {
  const int t = wg.block() * blockDim.x + threadIdx.x;
  if (t >= na) return;
  double a[6];
#pragma unroll
  for (int c = 0; c < 6; c++) a[c] = boxA[6 * t + c];
  auto overlaps = [&](const double* b) {
    return !(a[0] > b[3] || a[1] > b[4] || a[2] > b[5] || a[3] < b[0] || a[4] < b[1] ||
             a[5] < b[2]);
  };
  if (!overlaps(rootB)) return;
  bool hit = false;
  for (int q = 0; q < nb && !hit; q++) hit = overlaps(boxB + 6 * q);
  if (hit) atomicOr(flag, 1);
}
ESF_BOUNDS(kMeshPairs, 256);

template <class GridT>
__device__ __forceinline__ void kZeroInts(const GridT& wg, int n, int* a) {
  const int i = wg.block() * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}
ESF_BOUNDS(kZeroInts, 1024);

esf_status advectMeshes(esf_ctx* ctx) {
  for (const DevSolid& so : ctx->solids) {
    if (so.meshV == 0) continue;
    launch<ESF_K(kAdvectMesh)>(ctx, dim3(cdiv(so.meshV, 256)), dim3(256), ctx->d, so.meshV,
                               (const double*)so.meshRest, so.meshVert, (const double*)so.vel,
                               (const double*)so.u, ctx->dt);
  }
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

// the mesh broad phase: root boxes of every solid's mesh (6 doubles each)
// and, for every pair i < j in order, whether their meshes collide.  The
// kernels and the copy of their results to pinned memory are queued here;
// meshBroadPhaseRead takes the results once the copy has landed.
esf_status meshBroadPhaseEnqueue(esf_ctx* ctx) {
  const int n = (int)ctx->solids.size();
  const int np = n * (n - 1) / 2;
  double* root = ctx->meshOut;
  int* flags = reinterpret_cast<int*>(ctx->meshOut + 6 * kMaxSolids);
  if (np > 0)
    launch<ESF_K(kZeroInts)>(ctx, dim3(1), dim3(64), np, flags);
  for (int s = 0; s < n; s++) {
    const DevSolid& so = ctx->solids[s];
    launch<ESF_K(kTriBoxes)>(ctx, dim3(cdiv(so.meshT, 256)), dim3(256), so.meshT,
                             (const int*)so.meshTri, (const double*)so.meshVert, so.meshBox);
    launch<ESF_K(kMeshRoot)>(ctx, dim3(1), dim3(256), so.meshT, (const double*)so.meshBox,
                             root + 6 * s);
  }
  int q = 0;
  for (int a = 0; a < n; a++)
    for (int b = a + 1; b < n; b++, q++) {
      const DevSolid& A = ctx->solids[a];
      const DevSolid& B = ctx->solids[b];
      launch<ESF_K(kMeshPairs)>(ctx, dim3(cdiv(A.meshT, 256)), dim3(256), A.meshT,
                                (const double*)A.meshBox, B.meshT, (const double*)B.meshBox,
                                (const double*)(root + 6 * b), flags + q);
    }
  // root boxes and flags, laid out as in meshOut, into pinned memory
  toHostAsync(ctx, ctx->hostMesh, root, (6 * kMaxSolids) * sizeof(double) + np * sizeof(int));
  return hipGetLastError() == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

void meshBroadPhaseRead(esf_ctx* ctx, std::vector<double>& roots, std::vector<int>& pairs) {
  const int n = (int)ctx->solids.size();
  const int np = n * (n - 1) / 2;
  const int* hflags = reinterpret_cast<const int*>(ctx->hostMesh + 6 * kMaxSolids);
  roots.assign(ctx->hostMesh, ctx->hostMesh + 6 * n);
  pairs.assign(hflags, hflags + np);
}

}  // namespace esf
