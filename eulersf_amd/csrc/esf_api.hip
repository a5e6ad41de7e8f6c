// esf_api.hip -- the extern "C" entry points declared in include/esf.h.
// Host code: allocation, transfers in the reference's layouts, and the
// stage sequence of COUPLED_INTEGRATOR_3D::stepImplicit on one HIP stream.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "esf_internal.h"

namespace esf {
// esf_pressure.hip
esf_status pressureRhs(esf_ctx* ctx, CrScalars* sc = nullptr, double eps = 0);
esf_status pressureRhsDeferDiag(esf_ctx* ctx, CrScalars* sc, double eps);
esf_status pressureDiagFlush(esf_ctx* ctx);
esf_status pressureApplyDevice(esf_ctx* ctx, const double* x, double* y);
template <typename T>
esf_status crSolve(esf_ctx* ctx, int which, const T* minv, const T* b, const T* D,
                   const T* Dinv, T* x, T* r,
                   T* r2, T* p, T* s, T* s2, T* Ap, T* Ap2, double eps, int maxIt, int slot,
                   bool recursive = false, bool scalarsReady = false);
esf_status mixedSolve(esf_ctx* ctx, double eps, int maxIt, int* iters, bool* converged,
                      double* relres, int* refinements, int* innerIters,
                      bool scalarsReady = false);
int spmvBlocks(const esf_ctx* ctx);
void spmvConfigure(esf_ctx* ctx);
esf_status benchPressureApply(esf_ctx* ctx, int fp32, int reps, double* avgMs);
esf_status benchCopy(esf_ctx* ctx, long bytes, int reps, double* avgMs);
// esf_solid.hip
esf_status stressStage(esf_ctx* ctx, double* energy);
esf_status solidAdiag(esf_ctx* ctx);
esf_status solidApplyDevice(esf_ctx* ctx, const double* x, double* y);
esf_status solidSolve(esf_ctx* ctx, int which, const double* b, double* x, double eps, int maxIt,
                      int slot, const double* guess);
__global__ void kCorotBatch(int n, double lambda, double mu, const double* F, const double* dF,
                            double* out, double* energy, double* dP);
// esf_grid.hip
esf_status gatherSolidMasses(esf_ctx* ctx, bool defer);
esf_status gatherFinish(esf_ctx* ctx);
esf_status computeSolidMass(esf_ctx* ctx, int solid, bool usesdf);
esf_status updateBoxes(esf_ctx* ctx, int solid);
esf_status updateBoxesEnqueue(esf_ctx* ctx, int solid);
void updateBoxesFinish(esf_ctx* ctx, int solid, bool markList);
esf_status uploadBox(esf_ctx* ctx, int solid);
esf_status solidPhi(esf_ctx* ctx, int solid, const Box& launchBox, bool tightOnly);
esf_status cfl(esf_ctx* ctx, double alpha, double* dtOut);
esf_status vorticityAndBodyForces(esf_ctx* ctx, double* energy);
esf_status forkBodyForces(esf_ctx* ctx);
esf_status cflEnqueue(esf_ctx* ctx, bool mark);
esf_status cflWait(esf_ctx* ctx);
esf_status cflFinish(esf_ctx* ctx, double alpha, double* dtOut);
esf_status joinBodyForces(esf_ctx* ctx);
esf_status solidRhs(esf_ctx* ctx);
esf_status scatterSolidV0(esf_ctx* ctx);
esf_status correctVelocity(esf_ctx* ctx);
esf_status solidVelocity(esf_ctx* ctx);
esf_status advectVelocity(esf_ctx* ctx);
esf_status advectIntensity(esf_ctx* ctx);
esf_status solidCopyVelocity(esf_ctx* ctx);
esf_status solidFinalize(esf_ctx* ctx, double dt);
esf_status recoverAndBC(esf_ctx* ctx, int bc);
esf_status fillInt(esf_ctx* ctx, long n, int* p, int v);
esf_status zeroD(esf_ctx* ctx, long n, double* p);

template <typename T>
static esf_status dalloc(esf_ctx* ctx, T** p, size_t count) {
  if (*p) {
    auto it = ctx->allocs.find((void*)*p);
    if (it != ctx->allocs.end()) {
      ctx->deviceBytes -= it->second;
      ctx->allocs.erase(it);
    }
    hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    ctx->err = std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
               " bytes: " + hipGetErrorString(e);
    *p = nullptr;
    return ESF_ERR_OOM;
  }
  ctx->allocs[(void*)*p] = count * sizeof(T);
  ctx->deviceBytes += count * sizeof(T);
  e = hipMemsetAsync(*p, 0, count * sizeof(T), ctx->stream);
  return e == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status devAllocBytes(esf_ctx* ctx, void** p, size_t bytes) {
  return dalloc(ctx, (char**)p, bytes);
}

// Lists grow with a quarter of headroom; tight lists (tests) take exactly
// a known count (need > 0), shrinking too.
static size_t listCapacity(const esf_ctx* ctx, int need, int cap) {
  if (ctx->tightLists) return std::max(need > 0 ? need : cap, 1);
  need = std::max(need, cap);
  return std::max<size_t>(need + need / 4 + 64, 256);
}
static bool listFits(const esf_ctx* ctx, int need, int cap) {
  return need <= cap && !(ctx->tightLists && need > 0 && need != cap);
}

esf_status growSolidDofs(esf_ctx* ctx, int need) {
  if (listFits(ctx, need, ctx->capSolid) && ctx->sForce &&
      ctx->inboxSolids >= (int)ctx->solids.size())
    return ESF_OK;
  const size_t cap = listCapacity(ctx, need, ctx->capSolid);
  double** vecs[] = {&ctx->sForce, &ctx->sMv,  &ctx->sDamped, &ctx->sKdiag, &ctx->sAdiag,
                     &ctx->sRhs,   &ctx->sV0,  &ctx->sDv,     &ctx->sB,     &ctx->sr,
                     &ctx->sp,     &ctx->ss,   &ctx->sAp,     &ctx->stmp,   &ctx->sSd,
                     &ctx->sApd,   &ctx->sXb};
  for (double** v : vecs) {
    esf_status st = dalloc(ctx, v, 3 * cap);
    if (st) return st;
  }
  esf_status st = dalloc(ctx, &ctx->solidNodes, cap);
  if (st) return st;
  if ((st = dalloc(ctx, &ctx->nodeMask, cap))) return st;
  ctx->inboxSolids = std::max<int>(1, (int)ctx->solids.size());
  if ((st = dalloc(ctx, &ctx->inbox, 24 * cap * ctx->inboxSolids))) return st;
  if ((st = dalloc(ctx, &ctx->inboxX, 24 * cap * ctx->inboxSolids))) return st;
  // 3 sets x 5 sums and a set of 4 x node-pass blocks (3 cap DOFs / 256)
  if ((st = dalloc(ctx, &ctx->sPartials, 20 * (3 * cap / 256 + 1)))) return st;
  ctx->capSolid = (int)cap;
  return ESF_OK;
}

esf_status growSolidCells(esf_ctx* ctx, DevSolid& so, int need) {
  if (listFits(ctx, need, so.capValid) && so.cellList) return ESF_OK;
  const size_t cap = listCapacity(ctx, need, so.capValid);
  esf_status st;
  if ((st = dalloc(ctx, &so.cellList, cap))) return st;
  if ((st = dalloc(ctx, &so.cellMat, 28 * cap))) return st;
  if ((st = dalloc(ctx, &so.cellOut, 24 * cap))) return st;
  if ((st = dalloc(ctx, &so.cellDiag, 24 * cap))) return st;
  if ((st = dalloc(ctx, &so.cornerDof, 8 * cap))) return st;
  so.capValid = (int)cap;
  return ESF_OK;
}

static void freeAll(esf_ctx* ctx) {
  // every device buffer of the context went through dalloc
  for (auto& kv : ctx->allocs) hipFree(kv.first);
  ctx->allocs.clear();
  ctx->deviceBytes = 0;
  if (ctx->hostFlags) hipHostFree(ctx->hostFlags);
  if (ctx->hostCfl) hipHostFree(ctx->hostCfl);
  if (ctx->hostMesh) hipHostFree(ctx->hostMesh);
  if (ctx->hostRing) hipHostFree(ctx->hostRing);
  if (ctx->hostSc) hipHostFree(ctx->hostSc);
}

struct StageTimer {
  esf_ctx* ctx;
  int stage;
  // stage timing is opt-in: it synchronises the stream at every stage end
  StageTimer(esf_ctx* c, int s) : ctx(c), stage(s) {
    if (ctx->stageTiming) hipEventRecord(ctx->ev[0], ctx->stream);
  }
  ~StageTimer() {
    if (!ctx->stageTiming) return;
    hipEventRecord(ctx->ev[1], ctx->stream);
    eventSync(ctx, ctx->ev[1]);
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->stageMs[stage] += ms;
  }
};

}  // namespace esf

using namespace esf;

static esf_status checkCtx(esf_ctx* ctx) { return ctx ? ESF_OK : ESF_ERR_INVALID; }

static void applyParams(esf_ctx* ctx, const esf_params* p) {
  ctx->p = *p;
  if (ctx->p.inner_eps <= 0) ctx->p.inner_eps = 1e-4;
  const int maxRes = std::max(std::max(ctx->d.X, ctx->d.Y), ctx->d.Z);
  // COUPLED_INTEGRATOR_3D.cpp:67-71, float arithmetic as written there
  double scaling = 64.0f / maxRes;
  scaling = (scaling < 1.0f) ? 1.0f : scaling;
  ctx->vortEps = p->vorticity_eps / scaling;
}

extern "C" {

esf_status esf_ctx_create(esf_ctx** out, const esf_grid* g, const esf_params* p) {
  if (!out || !g || !p) return ESF_ERR_INVALID;
  *out = nullptr;
  if (g->xres < 4 || g->yres < 4 || g->zres < 4 || !(g->dh > 0)) return ESF_ERR_INVALID;
  if ((long)g->xres * g->yres * g->zres > (1L << 31) / 8) return ESF_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ESF_ERR_DEVICE;
  esf_ctx* ctx = new esf_ctx();
  hipGetDevice(&ctx->device);
  // ESF_TIGHT_LISTS=1 (tests): the solid DOF and valid-cell lists are
  // sized to exactly their last counts, so a step whose count rises takes
  // the path that grows a list after its scan and scans again
  const char* tight = getenv("ESF_TIGHT_LISTS");
  ctx->tightLists = tight && tight[0] == '1';
  // ESF_PHI_AS_LEFT=1 (tests): a download of phi returns what the step left
  // (relaxed only near the tight box) instead of the whole-box relaxation
  const char* asLeft = getenv("ESF_PHI_AS_LEFT");
  ctx->phiAsLeft = asLeft && asLeft[0] == '1';
  Dims& d = ctx->d;
  d.X = g->xres;
  d.Y = g->yres;
  d.Z = g->zres;
  d.CX = d.X - 1;
  d.CY = d.Y - 1;
  d.CZ = d.Z - 1;
  d.slab = (long)d.X * d.Y;
  d.nodes = d.slab * d.Z;
  d.cslab = (long)d.CX * d.CY;
  d.cells = d.cslab * d.CZ;
  d.dh = g->dh;
  d.dhInv = 1.0 / g->dh;
  d.o0 = g->origin[0];
  d.o1 = g->origin[1];
  d.o2 = g->origin[2];
  applyParams(ctx, p);
  ctx->defaultDt = p->dt;
  ctx->dt = p->dt;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return ESF_ERR_DEVICE;
  }
  for (int i = 0; i < 24; i++) hipEventCreate(&ctx->ev[i]);
  hipEventCreateWithFlags(&ctx->evFork, hipEventDisableTiming);
  hipEventCreateWithFlags(&ctx->evJoin, hipEventDisableTiming);
  hipEventCreateWithFlags(&ctx->evCfl, hipEventDisableTiming);
  hipEventCreateWithFlags(&ctx->evSettle, hipEventDisableTiming);
  esf_status st = ESF_OK;
  const long N = d.nodes, C = d.cells;
  const bool mixed = p->precision == ESF_MIXED;
  struct {
    double** p;
    long n;
  } dd[] = {{&ctx->vel, 3 * N},  {&ctx->velws, 3 * N}, {&ctx->vort, 3 * N},
            {&ctx->intensity, N}, {&ctx->intws, N},    {&ctx->ws, N},
            {&ctx->solidMass, N}, {&ctx->nodeFrac, N}, {&ctx->v0, 3 * N},
            {&ctx->minv, N},     {&ctx->c, C},         {&ctx->b, C},
            {&ctx->D, C},        {&ctx->lambda, C},    {&ctx->lambda2, C},
            {&ctx->r, C},        {&ctx->r2, C},        {&ctx->pdir, C},
            {&ctx->s, C},        {&ctx->Ap, C},        {&ctx->s2, C},
            {&ctx->Ap2, C},      {&ctx->Dinv, C}};
  for (auto& e : dd)
    if (!st) st = dalloc(ctx, e.p, e.n);
  if (!st && mixed) {
    float** fs[] = {&ctx->D32, &ctx->e32, &ctx->r32,  &ctx->r32b, &ctx->p32,
                    &ctx->s32, &ctx->s32b, &ctx->Ap32, &ctx->Ap32b, &ctx->rhs32,
                    &ctx->Dinv32};
    for (float** f : fs)
      if (!st) st = dalloc(ctx, f, C);
    if (!st) st = dalloc(ctx, &ctx->minv32, N);
    if (!st) st = dalloc(ctx, &ctx->dvPrev, 3 * N);
  }
  if (!st) st = dalloc(ctx, &ctx->sidx, N);
  if (!st) st = dalloc(ctx, &ctx->dBoxes, kMaxSolids);
  // block counts of two compactions at once (nodes and cells) and totals
  const long scanBlocks = N / 1024 + C / 1024 + 4;
  if (!st) st = dalloc(ctx, &ctx->scanBlocks, scanBlocks);
  // up to 5 partials per block of the widest reducing launch (the pressure
  // CR update over every SpMV tile, or 128-thread blocks over every node)
  spmvConfigure(ctx);
  ctx->nPartials = std::max<int>(5 * std::max(spmvBlocks(ctx), cdiv(N, 128)) + 64, 1 << 16);
  if (!st) st = dalloc(ctx, &ctx->partials, ctx->nPartials);
  if (!st) st = dalloc(ctx, &ctx->energyPart, cdiv(N, 256));
  if (!st) st = dalloc(ctx, &ctx->sc, 8);
  if (!st) st = dalloc(ctx, &ctx->dmisc, 16);
  if (!st) st = dalloc(ctx, &ctx->meshOut, 6 * kMaxSolids + 32);
  if (!st) st = growSolidDofs(ctx, 0);
  if (!st && hipHostMalloc((void**)&ctx->hostRing, 8 * sizeof(CrScalars)) != hipSuccess)
    st = ESF_ERR_OOM;
  if (!st && hipHostMalloc((void**)&ctx->hostFlags, kHostFlagInts * sizeof(int)) != hipSuccess)
    st = ESF_ERR_OOM;
  if (!st && hipHostMalloc((void**)&ctx->hostCfl, kCflBlocks * sizeof(double)) != hipSuccess)
    st = ESF_ERR_OOM;
  if (!st && hipHostMalloc((void**)&ctx->hostMesh, (6 * kMaxSolids + 32) * sizeof(double)) !=
                 hipSuccess)
    st = ESF_ERR_OOM;
  if (!st && hipHostMalloc((void**)&ctx->hostSc, esf_ctx::kReportSlots * sizeof(CrScalars)) !=
                 hipSuccess)
    st = ESF_ERR_OOM;
  if (!st) fillInt(ctx, N, ctx->sidx, -1);
  if (!st && streamSync(ctx) != hipSuccess) st = ESF_ERR_DEVICE;
  if (st) {
    esf_ctx_destroy(ctx);
    return st;
  }
  *out = ctx;
  return ESF_OK;
}

void esf_ctx_destroy(esf_ctx* ctx) {
  if (!ctx) return;
  if (ctx->side) hipStreamSynchronize(ctx->side);
  streamSync(ctx);
  freeAll(ctx);
  for (int i = 0; i < 24; i++) hipEventDestroy(ctx->ev[i]);
  hipEventDestroy(ctx->evFork);
  hipEventDestroy(ctx->evJoin);
  hipEventDestroy(ctx->evCfl);
  hipEventDestroy(ctx->evSettle);
  for (auto& t : ctx->timed) {
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  if (ctx->side) hipStreamDestroy(ctx->side);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* esf_last_error(const esf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

esf_status esf_set_params(esf_ctx* ctx, const esf_params* p) {
  if (!ctx || !p) return ESF_ERR_INVALID;
  if (p->precision != ctx->p.precision) {
    ctx->err = "precision is fixed at context creation";
    return ESF_ERR_INVALID;
  }
  applyParams(ctx, p);
  return ESF_OK;
}
esf_status esf_set_max_dt(esf_ctx* ctx, double v) {
  if (!ctx) return ESF_ERR_INVALID;
  ctx->defaultDt = v;
  return ESF_OK;
}
esf_status esf_set_dt(esf_ctx* ctx, double v) {
  if (!ctx) return ESF_ERR_INVALID;
  ctx->dt = v;
  return ESF_OK;
}
esf_status esf_get_dt(esf_ctx* ctx, double* v) {
  if (!ctx || !v) return ESF_ERR_INVALID;
  *v = ctx->dt;
  return ESF_OK;
}
esf_status esf_num_solid_nodes(esf_ctx* ctx, int* n) {
  if (!ctx || !n) return ESF_ERR_INVALID;
  *n = ctx->nSolid;
  return ESF_OK;
}

esf_status esf_solid_add(esf_ctx* ctx, double lambda, double mu, double density, int rx, int ry,
                         int rz, double rdh, const double ro[3], const double* rest, int* id) {
  if (!ctx || !rest || !ro || rx < 2 || ry < 2 || rz < 2 || !(rdh > 0)) return ESF_ERR_INVALID;
  if (ctx->solids.size() >= (size_t)kMaxSolids) {
    ctx->err = "at most " + std::to_string(kMaxSolids) + " solids per context";
    return ESF_ERR_INVALID;
  }
  ctx->solids.emplace_back();
  DevSolid& so = ctx->solids.back();
  const Dims& d = ctx->d;
  so.lambda = lambda;
  so.mu = mu;
  so.qcm = density * d.dh * d.dh * d.dh * 0.125;  // setMassDensity
  so.rX = rx;
  so.rY = ry;
  so.rZ = rz;
  so.rdh = rdh;
  for (int k = 0; k < 3; k++) so.ro[k] = ro[k];
  for (int k = 0; k < 3; k++) {
    so.box.tMin[k] = so.box.bMin[k] = 0;
  }
  so.box.tMax[0] = so.box.bMax[0] = d.X;
  so.box.tMax[1] = so.box.bMax[1] = d.Y;
  so.box.tMax[2] = so.box.bMax[2] = d.Z;
  const long N = d.nodes;
  esf_status st = ESF_OK;
  double** v3[] = {&so.u, &so.unew, &so.X, &so.vel, &so.velws};
  for (double** p : v3)
    if (!st) st = dalloc(ctx, p, 3 * N);
  if (!st) st = dalloc(ctx, &so.mass, N);
  if (!st) st = dalloc(ctx, &so.detFInv, N);
  if (!st) st = dalloc(ctx, &so.phi, N);
  if (!st) st = dalloc(ctx, &so.phi2, N);
  if (!st) st = dalloc(ctx, &so.qfrac, 8 * N);
  if (!st) st = dalloc(ctx, &so.valid, d.cells);
  if (!st) st = dalloc(ctx, &so.cellSlot, d.cells);
  if (!st) st = dalloc(ctx, &so.rest, (long)rx * ry * rz);
  if (!st) st = growSolidCells(ctx, so, 0);
  if (st) {
    ctx->solids.pop_back();
    return st;
  }
  ESF_HIP(hipMemcpyAsync(so.rest, rest, sizeof(double) * (long)rx * ry * rz,
                         hipMemcpyHostToDevice, ctx->stream));
  fillInt(ctx, d.cells, so.cellSlot, -1);
  ESF_HIP(streamSync(ctx));
  if ((st = uploadBox(ctx, (int)ctx->solids.size() - 1))) return st;
  if (id) *id = (int)ctx->solids.size() - 1;
  return ESF_OK;
}

esf_status esf_solid_set_mesh(esf_ctx* ctx, int solid, int nv, const double* rest,
                             const double* vertices, int nt, const int* tris) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size() || nv < 0 || nt < 0)
    return ESF_ERR_INVALID;
  if (nv > 0 && (!rest || !vertices || nt <= 0 || !tris)) return ESF_ERR_INVALID;
  for (long k = 0; k < 3L * nt; k++)
    if (tris[k] < 0 || tris[k] >= nv) {
      ctx->err = "triangle vertex index out of range";
      return ESF_ERR_INVALID;
    }
  DevSolid& so = ctx->solids[solid];
  esf_status st = ESF_OK;
  if (nv == 0) {
    so.meshV = so.meshT = 0;
    so.meshTriHost.clear();
    return ESF_OK;
  }
  if (!st) st = dalloc(ctx, &so.meshRest, 3L * nv);
  if (!st) st = dalloc(ctx, &so.meshVert, 3L * nv);
  if (!st) st = dalloc(ctx, &so.meshTri, 3L * nt);
  if (!st) st = dalloc(ctx, &so.meshBox, 6L * nt);
  if (st) return st;
  ESF_HIP(hipMemcpyAsync(so.meshRest, rest, 24L * nv, hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(hipMemcpyAsync(so.meshVert, vertices, 24L * nv, hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(hipMemcpyAsync(so.meshTri, tris, 12L * nt, hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(streamSync(ctx));
  so.meshTriHost.assign(tris, tris + 3L * nt);
  so.meshV = nv;
  so.meshT = nt;
  return ESF_OK;
}

esf_status esf_solid_mesh_size(esf_ctx* ctx, int solid, int* nv, int* nt) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size() || !nv || !nt)
    return ESF_ERR_INVALID;
  *nv = ctx->solids[solid].meshV;
  *nt = ctx->solids[solid].meshT;
  return ESF_OK;
}

esf_status esf_solid_get_mesh(esf_ctx* ctx, int solid, double* vertices, double* rest,
                              int* tris) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  const DevSolid& so = ctx->solids[solid];
  if (so.meshV == 0) {
    ctx->err = "solid has no surface mesh";
    return ESF_ERR_STATE;
  }
  if (vertices)
    ESF_HIP(hipMemcpyAsync(vertices, so.meshVert, 24L * so.meshV, hipMemcpyDeviceToHost,
                           ctx->stream));
  if (rest)
    ESF_HIP(hipMemcpyAsync(rest, so.meshRest, 24L * so.meshV, hipMemcpyDeviceToHost,
                           ctx->stream));
  ESF_HIP(streamSync(ctx));
  if (tris) std::copy(so.meshTriHost.begin(), so.meshTriHost.end(), tris);
  return ESF_OK;
}

esf_status esf_advect_surface_mesh(esf_ctx* ctx) {
  if (!ctx) return ESF_ERR_INVALID;
  esf_status st = advectMeshes(ctx);
  if (st) return st;
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_solid_set_boxes(esf_ctx* ctx, int solid, const int b[12]) {
  if (!ctx || !b || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  Box& bx = ctx->solids[solid].box;
  const int res[3] = {ctx->d.X, ctx->d.Y, ctx->d.Z};
  for (int k = 0; k < 3; k++) {
    if (b[k] < 0 || b[3 + k] > res[k] || b[6 + k] < 0 || b[9 + k] > res[k]) {
      ctx->err = "box outside the grid";
      return ESF_ERR_INVALID;
    }
  }
  esf_status st = phiWhole(ctx, solid);  // relaxed on the boxes it was computed for
  if (st) return st;
  for (int k = 0; k < 3; k++) {
    bx.tMin[k] = b[k];
    bx.tMax[k] = b[3 + k];
    bx.bMin[k] = b[6 + k];
    bx.bMax[k] = b[9 + k];
  }
  ctx->solids[solid].listDirty = true;
  return uploadBox(ctx, solid);
}

esf_status esf_solid_get_boxes(esf_ctx* ctx, int solid, int b[12]) {
  if (!ctx || !b || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  const Box& bx = ctx->solids[solid].box;
  for (int k = 0; k < 3; k++) {
    b[k] = bx.tMin[k];
    b[3 + k] = bx.tMax[k];
    b[6 + k] = bx.bMin[k];
    b[9 + k] = bx.bMax[k];
  }
  return ESF_OK;
}

static esf_status fieldPtr(esf_ctx* ctx, int field, int solid, void** ptr, size_t* bytes,
                           bool forWrite) {
  const Dims& d = ctx->d;
  const size_t N = d.nodes, C = d.cells, NS = 3 * (size_t)ctx->nSolid;
  *ptr = nullptr;
  switch (field) {
    case ESF_F_VELOCITY: *ptr = ctx->vel; *bytes = 24 * N; return ESF_OK;
    case ESF_F_INTENSITY: *ptr = ctx->intensity; *bytes = 8 * N; return ESF_OK;
    case ESF_F_SOLID_MASS: *ptr = ctx->solidMass; *bytes = 8 * N; break;
    case ESF_F_NODE_FRACTION: *ptr = ctx->nodeFrac; *bytes = 8 * N; break;
    case ESF_F_LAMBDA:
      *ptr = ctx->lambda;
      *bytes = 8 * C;
      if (forWrite) ctx->lambdaInit = true;
      return ESF_OK;
    case ESF_F_SOLID_INDEX: *ptr = ctx->sidx; *bytes = 4 * N; break;
    case ESF_F_MINV: *ptr = ctx->minv; *bytes = 8 * N; break;
    case ESF_F_PRESSURE_RHS:  // writable: the right-hand side of esf_pressure_solve_resident
      *ptr = ctx->b;
      *bytes = 8 * C;
      return ESF_OK;
    case ESF_F_DIV_RHS: *ptr = ctx->c; *bytes = 8 * C; break;
    case ESF_F_V0: *ptr = ctx->v0; *bytes = 24 * N; break;
    case ESF_F_SOLID_FORCE: *ptr = ctx->sForce; *bytes = 8 * NS; break;
    case ESF_F_SOLID_MV: *ptr = ctx->sMv; *bytes = 8 * NS; break;
    case ESF_F_DAMPED_MASS: *ptr = ctx->sDamped; *bytes = 8 * NS; break;
    case ESF_F_SYS_DIAG: *ptr = ctx->sAdiag; *bytes = 8 * NS; break;
    case ESF_F_PRESSURE_DIAG:
      if (ctx->diagPending) {
        esf_status st = pressureDiagFlush(ctx);
        if (st) return st;
        ESF_HIP(streamSync(ctx));
      }
      *ptr = ctx->D;
      *bytes = 8 * C;
      break;
    case ESF_F_SOLID_V0: *ptr = ctx->sV0; *bytes = 8 * NS; break;
    case ESF_F_SOLID_RHS:  // writable: the right-hand side of esf_solid_solve_resident
      *ptr = ctx->sRhs;
      *bytes = 8 * NS;
      return ESF_OK;
    case ESF_F_SOLID_DV: *ptr = ctx->sDv; *bytes = 8 * NS; break;
    case ESF_F_WORKSPACE: *ptr = ctx->ws; *bytes = 8 * N; break;
    default: break;
  }
  if (*ptr) {
    if (forWrite) {
      ctx->err = "field is derived and read-only";
      return ESF_ERR_INVALID;
    }
    return ESF_OK;
  }
  if (solid < 0 || solid >= (int)ctx->solids.size()) {
    ctx->err = "bad field id or solid id";
    return ESF_ERR_INVALID;
  }
  DevSolid& so = ctx->solids[solid];
  switch (field) {
    case ESF_S_U:
      // _X keeps x - u of the u being replaced
      if (forWrite && so.xLazy) {
        esf_status st = materialiseX(ctx, so);
        if (st) return st;
      }
      *ptr = so.u;
      *bytes = 24 * N;
      break;
    case ESF_S_X:
      if (so.xLazy && !forWrite) {
        esf_status st = materialiseX(ctx, so);
        if (st) return st;
      }
      so.xLazy = false;  // an upload replaces the whole field
      *ptr = so.X;
      *bytes = 24 * N;
      break;
    case ESF_S_VELOCITY:
      *ptr = so.vel;
      *bytes = 24 * N;
      if (forWrite) so.wsState = 2;
      break;
    case ESF_S_MASS:
      *ptr = so.mass;
      *bytes = 8 * N;
      if (forWrite) {
        esf_status st = phiWhole(ctx, solid);  // from the mass being replaced
        if (st) return st;
        so.listDirty = true;
        so.dirtyAll = true;
        ctx->massGen++;
      }
      break;
    case ESF_S_VALID:
      *ptr = so.valid;
      *bytes = C;
      if (forWrite) {
        so.listDirty = true;
        so.dirtyAll = true;
      }
      break;
    case ESF_S_QFRAC:
      *ptr = so.qfrac;
      *bytes = 64 * N;
      if (forWrite) {
        so.dirtyAll = true;
        ctx->massGen++;
      }
      break;
    case ESF_S_VEL_WORKSPACE:
      *ptr = so.velws;
      *bytes = 24 * N;
      if (forWrite) so.wsState = 2;
      break;
    case ESF_S_PHI:
      if (!forWrite && ctx->phiAsLeft) {
        *ptr = so.phi;
        *bytes = 8 * N;
        return ESF_OK;
      }
      if (!forWrite) {
        esf_status st = phiWhole(ctx, solid);
        if (st) return st;
      }
      so.phiPartial = false;  // an upload replaces the whole field
      *ptr = so.phi;
      *bytes = 8 * N;
      if (forWrite)
        for (int k = 0; k < 3; k++) {
          so.phiSpan[k] = 0;
          so.phiSpan[3 + k] = 1 << 30;
        }
      break;
    default:
      ctx->err = "bad field id";
      return ESF_ERR_INVALID;
  }
  return ESF_OK;
}

esf_status esf_upload(esf_ctx* ctx, int field, int solid, const void* src, size_t bytes) {
  if (!ctx || !src) return ESF_ERR_INVALID;
  void* p;
  size_t want;
  esf_status st = fieldPtr(ctx, field, solid, &p, &want, true);
  if (st) return st;
  if (bytes != want) {
    ctx->err = "upload size " + std::to_string(bytes) + " != " + std::to_string(want);
    return ESF_ERR_INVALID;
  }
  ESF_HIP(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_download(esf_ctx* ctx, int field, int solid, void* dst, size_t bytes) {
  if (!ctx || !dst) return ESF_ERR_INVALID;
  void* p;
  size_t want;
  esf_status st = fieldPtr(ctx, field, solid, &p, &want, false);
  if (st) return st;
  if (bytes != want) {
    ctx->err = "download size " + std::to_string(bytes) + " != " + std::to_string(want);
    return ESF_ERR_INVALID;
  }
  ESF_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_compute_solid_mass(esf_ctx* ctx, int solid) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  if (!(ctx->p.mass_sample_dh > 0) || (int)(ctx->d.dh / ctx->p.mass_sample_dh) < 1) {
    ctx->err = "mass sample dh must be positive and at most dh";
    return ESF_ERR_INVALID;
  }
  esf_status st = phiWhole(ctx, solid);  // from the mass it was computed for
  if (!st) st = computeSolidMass(ctx, solid, false);
  if (st) return st;
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_compute_solid_mass_sdf(esf_ctx* ctx, int solid) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  if (!(ctx->p.mass_sample_dh > 0) || (int)(ctx->d.dh / ctx->p.mass_sample_dh) < 1) {
    ctx->err = "mass sample dh must be positive and at most dh";
    return ESF_ERR_INVALID;
  }
  esf_status st = phiWhole(ctx, solid);
  if (!st) st = computeSolidMass(ctx, solid, true);
  if (st) return st;
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_update_boxes(esf_ctx* ctx, int solid) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  esf_status st = phiWhole(ctx, solid);
  return st ? st : updateBoxes(ctx, solid);
}

static esf_status flipCall(esf_ctx* ctx, int solid, esf_status (*fn)(esf_ctx*, int)) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  esf_status st = fn(ctx, solid);
  if (st) return st;
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_init_particles(esf_ctx* ctx, int solid) {
  if (ctx && solid >= 0 && solid < (int)ctx->solids.size() && ctx->solids[solid].listDirty) {
    ctx->err = "solid masses changed: call esf_gather_solid_masses first";
    return ESF_ERR_STATE;
  }
  return flipCall(ctx, solid, initParticles);
}
esf_status esf_grid_to_particle(esf_ctx* ctx, int solid) {
  return flipCall(ctx, solid, gridToParticle);
}
esf_status esf_particle_to_grid(esf_ctx* ctx, int solid) {
  return flipCall(ctx, solid, particleToGrid);
}
esf_status esf_extend_solid_velocity(esf_ctx* ctx, int solid) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  DevSolid& so = ctx->solids[solid];
  const int box[6] = {so.box.bMin[0], so.box.bMin[1], so.box.bMin[2],
                      so.box.bMax[0], so.box.bMax[1], so.box.bMax[2]};
  so.wsState = 2;
  esf_status st = phiWhole(ctx, solid);
  if (!st) st = extendDevice(ctx, box, so.mass, so.phi, so.vel);
  if (st) return st;
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}
esf_status esf_solid_num_cells(esf_ctx* ctx, int solid, int* n) {
  if (!ctx || !n || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  if (ctx->solids[solid].listDirty) {
    ctx->err = "solid masses changed: call esf_gather_solid_masses first";
    return ESF_ERR_STATE;
  }
  *n = ctx->solids[solid].nValid;
  return ESF_OK;
}

esf_status esf_solid_set_deformation(esf_ctx* ctx, int solid, long n, const double* F) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size() || n < 0 || (n && !F))
    return ESF_ERR_INVALID;
  DevSolid& so = ctx->solids[solid];
  if (n == 0) {
    so.nFset = -1;
    return ESF_OK;
  }
  if (so.listDirty || n != so.nValid) {
    ctx->err = "one deformation gradient per valid cell (esf_solid_num_cells) is needed";
    return ESF_ERR_INVALID;
  }
  esf_status st = dalloc(ctx, &so.Fset, 9 * n);
  if (st) return st;
  ESF_HIP(hipMemcpyAsync(so.Fset, F, 72 * n, hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(streamSync(ctx));
  so.nFset = n;
  return ESF_OK;
}

esf_status esf_solid_solve_resident(esf_ctx* ctx, double eps, int maxIt, int* iters,
                                    double* relres, double* deviceMs) {
  if (!ctx || !(eps > 0) || maxIt < 1) return ESF_ERR_INVALID;
  esf_status st = solidAdiag(ctx);
  if (st) return st;
  ESF_HIP(hipEventRecord(ctx->ev[4], ctx->stream));
  if ((st = solidSolve(ctx, 2, ctx->sRhs, ctx->sV0, eps, maxIt, 3, nullptr))) return st;
  ESF_HIP(hipEventRecord(ctx->ev[5], ctx->stream));
  ESF_HIP(eventSync(ctx, ctx->ev[5]));
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]);
  int it = 0;
  bool conv = false;
  double rr = 0;
  solveReport(ctx->hostSc[3], maxIt, &it, &conv, &rr);
  if (iters) *iters = it;
  if (relres) *relres = rr;
  if (deviceMs) *deviceMs = ms;
  return conv ? ESF_OK : ESF_NOT_CONVERGED;
}

esf_status esf_num_particles(esf_ctx* ctx, int solid, long* n) {
  if (!ctx || !n || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  *n = ctx->solids[solid].nPart;
  return ESF_OK;
}
esf_status esf_get_particles(esf_ctx* ctx, int solid, double* pos, double* vel) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  const DevSolid& so = ctx->solids[solid];
  if (pos && so.nPart)
    ESF_HIP(hipMemcpyAsync(pos, so.pPos, 24 * so.nPart, hipMemcpyDeviceToHost, ctx->stream));
  if (vel && so.nPart)
    ESF_HIP(hipMemcpyAsync(vel, so.pVel, 24 * so.nPart, hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}
esf_status esf_set_particles(esf_ctx* ctx, int solid, long n, const double* pos,
                             const double* vel) {
  if (!ctx || n < 0 || (n && (!pos || !vel)) || solid < 0 || solid >= (int)ctx->solids.size())
    return ESF_ERR_INVALID;
  esf_status st = growParticlesFor(ctx, solid, n);
  if (st) return st;
  DevSolid& so = ctx->solids[solid];
  if (n) {
    ESF_HIP(hipMemcpyAsync(so.pPos, pos, 24 * n, hipMemcpyHostToDevice, ctx->stream));
    ESF_HIP(hipMemcpyAsync(so.pVel, vel, 24 * n, hipMemcpyHostToDevice, ctx->stream));
  }
  ESF_HIP(streamSync(ctx));
  so.nPart = n;
  return ESF_OK;
}

esf_status esf_solid_sdf(esf_ctx* ctx, int solid) {
  if (!ctx || solid < 0 || solid >= (int)ctx->solids.size()) return ESF_ERR_INVALID;
  esf_status st = updateBoxes(ctx, solid);
  if (!st) st = solidPhi(ctx, solid, ctx->solids[solid].box, false);
  if (st) return st;
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_gather_solid_masses(esf_ctx* ctx) {
  if (!ctx) return ESF_ERR_INVALID;
  StageTimer t(ctx, 5);
  return gatherSolidMasses(ctx, false);
}

esf_status esf_recover_velocity(esf_ctx* ctx) {
  if (!ctx) return ESF_ERR_INVALID;
  esf_status st = recoverAndBC(ctx, 0);
  if (st) return st;
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_cfl(esf_ctx* ctx, double alpha, double* dt) {
  if (!ctx) return ESF_ERR_INVALID;
  double v;
  StageTimer t(ctx, 0);
  esf_status st = cfl(ctx, alpha, &v);
  if (dt) *dt = v;
  return st;
}

esf_status esf_forces(esf_ctx* ctx, double* energy) {
  if (!ctx) return ESF_ERR_INVALID;
  StageTimer t(ctx, 1);
  for (auto& so : ctx->solids)
    if (so.listDirty) {
      ctx->err = "solid masses changed: call esf_gather_solid_masses first";
      return ESF_ERR_STATE;
    }
  double e = 0;
  esf_status st = stressStage(ctx, energy ? &e : nullptr);
  if (st) return st;
  st = vorticityAndBodyForces(ctx, energy ? &e : nullptr);
  if (st) return st;
  if (energy) *energy = e;
  return ESF_OK;
}

// stepImplicit :609-632 + pcrImplicitSolve (:662-764), queued on the
// stream.  The FP64 solves leave their reports in the pinned slots 0-2
// (read by coupledSolveReport after the next synchronisation); the mixed
// pressure solve synchronises per refinement and fills *s itself.
static esf_status coupledSolveEnqueue(esf_ctx* ctx, esf_solve_stats* s) {
  esf_status st;
  const int maxIt = ctx->p.cr_max_iterations;
  *s = esf_solve_stats{};
  batchBarrier(ctx);
  {
    StageTimer t(ctx, 2);
    if ((st = solidAdiag(ctx))) return st;
    if ((st = solidRhs(ctx))) return st;
    // the mixed form starts the solid solves from the previous step's
    // solutions at the same nodes (v0, and the correction kept in dvPrev);
    // the FP64 form starts from 0 as the reference does
    // (CONJUGATE_RESIDUAL.inl:197-198)
    const bool warm = ctx->p.precision == ESF_MIXED;
    if ((st = solidSolve(ctx, 2, ctx->sRhs, ctx->sV0, ctx->p.implicit_eps, maxIt, 0,
                         warm ? ctx->v0 : nullptr)))
      return st;
    if ((st = scatterSolidV0(ctx))) return st;
    // the fluid's body forces, if they ran beside the solve (esf_step)
    if ((st = joinBodyForces(ctx))) return st;
  }
  batchBarrier(ctx);
  {
    StageTimer t(ctx, 3);
    // the mixed form forms the diagonal only where its inner solve runs;
    // both forms' right-hand side kernels initialise the solve's scalars
    CrScalars* sc0 = &ctx->sc[0];
    const double eps = ctx->p.divfree_eps;
    if ((st = ctx->p.precision == ESF_MIXED ? pressureRhsDeferDiag(ctx, sc0, eps)
                                            : pressureRhs(ctx, sc0, eps)))
      return st;
    if (!ctx->lambdaInit) {
      if ((st = zeroD(ctx, ctx->d.cells, ctx->lambda))) return st;
      ctx->lambdaInit = true;
    }
    if (ctx->p.precision == ESF_MIXED) {
      bool conv = false;
      st = mixedSolve(ctx, eps, maxIt, &s->iterations[1], &conv, &s->rel_residual[1],
                      &s->refinements, &s->inner_iterations, true);
      s->converged[1] = conv;
    } else {
      st = crSolve<double>(ctx, 0, ctx->minv, ctx->b, ctx->D, ctx->Dinv, ctx->lambda, ctx->r, ctx->r2,
                           ctx->pdir, ctx->s, ctx->s2, ctx->Ap, ctx->Ap2, eps, maxIt, 1, false, true);
    }
    if (st) return st;
    if ((st = correctVelocity(ctx))) return st;
  }
  batchBarrier(ctx);
  {
    StageTimer t(ctx, 2);
    if ((st = solidSolve(ctx, 2, ctx->sB, ctx->sDv, ctx->p.implicit_eps, maxIt, 2,
                         ctx->p.precision == ESF_MIXED ? ctx->dvPrev : nullptr)))
      return st;
    if ((st = solidVelocity(ctx))) return st;
  }
  return ESF_OK;
}

// after a synchronisation of the stream: the reports of the queued solves,
// and the status (NaN residual, not converged)
static esf_status coupledSolveReport(esf_ctx* ctx, esf_solve_stats* s) {
  const int maxIt = ctx->p.cr_max_iterations;
  for (int k = 0; k < 3; k++) {
    if (k == 1 && ctx->p.precision == ESF_MIXED) continue;
    bool conv = false;
    solveReport(ctx->hostSc[k], maxIt, &s->iterations[k], &conv, &s->rel_residual[k]);
    s->converged[k] = conv;
  }
  s->dt = ctx->dt;
  s->n_solid_nodes = ctx->nSolid;
  for (auto& so : ctx->solids) s->n_solid_cells += so.nValid;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("coupled solve: ") + hipGetErrorString(e);
    return ESF_ERR_DEVICE;
  }
  for (double r : s->rel_residual)
    if (std::isnan(r)) {
      ctx->err = "NaN residual in the coupled solve";
      return ESF_ERR_NAN;
    }
  return (s->converged[0] && s->converged[1] && s->converged[2]) ? ESF_OK : ESF_NOT_CONVERGED;
}

esf_status esf_coupled_solve(esf_ctx* ctx, esf_solve_stats* stats) {
  if (!ctx) return ESF_ERR_INVALID;
  esf_solve_stats s{};
  esf_status st = coupledSolveEnqueue(ctx, &s);
  if (st) return st;
  ESF_HIP(streamSync(ctx));
  st = coupledSolveReport(ctx, &s);
  if (stats) *stats = s;
  return st;
}

esf_status esf_solid_copy_velocity(esf_ctx* ctx) {
  if (!ctx) return ESF_ERR_INVALID;
  StageTimer t(ctx, 4);
  return solidCopyVelocity(ctx);
}
esf_status esf_solid_finalize(esf_ctx* ctx) {
  if (!ctx) return ESF_ERR_INVALID;
  StageTimer t(ctx, 4);
  return solidFinalize(ctx, ctx->dt);
}
esf_status esf_advect_velocity(esf_ctx* ctx) {
  if (!ctx) return ESF_ERR_INVALID;
  StageTimer t(ctx, 6);
  return advectVelocity(ctx);
}
esf_status esf_advect_intensity(esf_ctx* ctx) {
  if (!ctx) return ESF_ERR_INVALID;
  StageTimer t(ctx, 7);
  return advectIntensity(ctx);
}
esf_status esf_recover_and_bc(esf_ctx* ctx) {
  if (!ctx) return ESF_ERR_INVALID;
  StageTimer t(ctx, 8);
  return recoverAndBC(ctx, 1);
}

// Steps in flight in this process.  A step forks its fluid forces onto its
// context's second stream only when it runs alone: the process has a few
// hardware queues (4 by default) that all contexts' streams share, and a
// cross-stream wait on a shared queue can hold up the other scenes queued
// behind it.  (64 concurrent scenes measured 1.49 x their sequential rate
// with every step forking and a second stream created with every context,
// 2.15 x with neither.)
static std::atomic<int> gStepsInFlight{0};

namespace {
struct StepCount {
  bool alone;
  StepCount() : alone(gStepsInFlight.fetch_add(1) == 0) {}
  ~StepCount() { gStepsInFlight.fetch_sub(1); }
};
}  // namespace

// a step that ended early (an error) may leave box reads and list counts
// queued: settle them before anything reads the host copies
// `ready`: an event after which everything pending has landed (null: the
// whole stream)
static esf_status settlePending(esf_ctx* ctx, hipEvent_t ready = nullptr) {
  bool any = ctx->dofPending;
  for (auto& so : ctx->solids) any = any || so.boxPending || so.listPending;
  if (!any) return ESF_OK;
  ESF_HIP(ready ? eventSync(ctx, ready) : streamSync(ctx));
  for (int b = 0; b < (int)ctx->solids.size(); b++)
    if (ctx->solids[b].boxPending) {
      updateBoxesFinish(ctx, b, false);
      ctx->solids[b].boxPending = false;
    }
  return gatherFinish(ctx);
}

esf_status esf_step(esf_ctx* ctx, double alpha, esf_solve_stats* stats) {
  if (!ctx) return ESF_ERR_INVALID;
  const StepCount inFlight;
  double dt;
  esf_status st, solveSt;
  if ((st = settlePending(ctx))) return st;
  for (auto& so : ctx->solids)
    if (so.listDirty) {
      ctx->err = "solid masses changed: call esf_gather_solid_masses first";
      return ESF_ERR_STATE;
    }
  if (ctx->p.flip && ctx->solids.size() > 1) {
    ctx->err = "the FLIP round trip takes scenes of one solid (contact between solids reads "
               "rest-space SDFs this library does not keep)";
    return ESF_ERR_INVALID;
  }
  // alpha < 0: the broad phase of stepImplicit decides it
  // (COUPLED_INTEGRATOR_3D.cpp:384-411), as in esf_frame.  Its mesh kernels
  // and their read-back are queued ahead of the CFL kernel, the stress stage
  // behind it (it does not use dt), and the host waits once for both
  // read-backs.  A step running alone sends the fluid's vorticity and body
  // forces to the side stream, beside the first solid solve, which needs
  // only the solid nodes' part (same results either way).
  batchBarrier(ctx);
  {
    StageTimer t(ctx, 1);
    const bool broad = alpha < 0;
    bool meshes = false;
    if (broad && (st = broadPhaseEnqueue(ctx, &meshes))) return st;
    if ((st = cflEnqueue(ctx, broad && meshes))) return st;
    if ((st = stressStage(ctx, nullptr))) return st;
    if ((st = cflWait(ctx))) return st;
    if (broad && (st = broadPhaseFinish(ctx, meshes, &alpha))) return st;
    if ((st = cflFinish(ctx, alpha, &dt))) return st;
    st = inFlight.alone && !ctx->fiber ? forkBodyForces(ctx)
                                       : vorticityAndBodyForces(ctx, nullptr);
    if (st) return st;
  }
  // the solve's reports are read at the end of the step: no stall between
  // the solve and the stages after it
  esf_solve_stats s{};
  if ((st = coupledSolveEnqueue(ctx, &s))) {
    joinBodyForces(ctx);
    return st;
  }
  batchBarrier(ctx);
  if ((st = esf_solid_copy_velocity(ctx))) return st;
  if (ctx->p.flip) {
    // the FLIP round trip: G2P, P2G, then the solid velocity extended over
    // its box by phi (COUPLED_INTEGRATOR_3D.cpp:833-854)
    StageTimer t(ctx, 10);
    for (int b = 0; b < (int)ctx->solids.size(); b++)
      if ((st = gridToParticle(ctx, b))) return st;
    for (int b = 0; b < (int)ctx->solids.size(); b++)
      if ((st = particleToGrid(ctx, b))) return st;
    for (int b = 0; b < (int)ctx->solids.size(); b++) {
      const Box& bx = ctx->solids[b].box;
      const int box[6] = {bx.bMin[0], bx.bMin[1], bx.bMin[2], bx.bMax[0], bx.bMax[1], bx.bMax[2]};
      if ((st = extendDevice(ctx, box, ctx->solids[b].mass, ctx->solids[b].phi,
                             ctx->solids[b].vel)))
        return st;
    }
  }
  if ((st = esf_solid_finalize(ctx))) return st;
  batchBarrier(ctx);
  {
    // computeSolidMass(true) of every solid, the surface meshes, then
    // computeSDF's boxes (all solids' read back at one synchronisation) and
    // phi (COUPLED_INTEGRATOR_3D.cpp:857-879)
    StageTimer t(ctx, 9);
    for (int b = 0; b < (int)ctx->solids.size(); b++)
      if ((st = computeSolidMass(ctx, b, true))) return st;
    if ((st = advectMeshes(ctx))) return st;
    for (int b = 0; b < (int)ctx->solids.size(); b++)
      if ((st = updateBoxesEnqueue(ctx, b))) return st;
    if (ctx->p.flip) {
      // the FLIP stages of this step and the next read the boxes on the host
      ESF_HIP(streamSync(ctx));
      for (int b = 0; b < (int)ctx->solids.size(); b++) {
        updateBoxesFinish(ctx, b, true);
        if ((st = solidPhi(ctx, b, ctx->solids[b].box, false))) return st;
      }
    } else {
      // the stages up to the step's end read the boxes from device memory
      // and launch over boxes that contain them; the host copy is read at
      // the end
      for (int b = 0; b < (int)ctx->solids.size(); b++) {
        DevSolid& so = ctx->solids[b];
        so.boxPending = true;
        so.listDirty = true;
        if ((st = solidPhi(ctx, b, so.boxBound, true))) return st;
      }
    }
  }
  batchBarrier(ctx);
  {
    // the DOF and valid-cell counts are read after the step's last
    // synchronisation (nothing before it launches by them), except where
    // the FLIP particles are seeded on the new valid cells
    StageTimer t(ctx, 5);
    if ((st = gatherSolidMasses(ctx, !ctx->p.flip))) return st;
  }
  // initMaterialParticles on the new valid cells (:870-877)
  if (ctx->p.flip)
    for (int b = 0; b < (int)ctx->solids.size(); b++)
      if ((st = initParticles(ctx, b))) return st;
  // A step alone returns once what the host reads at its end has landed:
  // the advection and recovery behind it run on while the caller comes back
  // for the next step (queued behind them on the same stream; every read
  // through this interface synchronises first).  Batched scenes wait for
  // their whole step.
  const hipEvent_t ready = ctx->fiber ? nullptr : markEvent(ctx, ctx->evSettle);
  batchBarrier(ctx);
  if ((st = esf_advect_velocity(ctx))) return st;
  if ((st = esf_advect_intensity(ctx))) return st;
  if ((st = esf_recover_and_bc(ctx))) return st;
  if ((st = settlePending(ctx, ready))) return st;
  solveSt = coupledSolveReport(ctx, &s);
  if (stats) *stats = s;
  return solveSt;
}

esf_status esf_pressure_apply(esf_ctx* ctx, const double* x, double* y) {
  if (!ctx || !x || !y) return ESF_ERR_INVALID;
  const size_t bytes = 8 * ctx->d.cells;
  ESF_HIP(hipMemcpyAsync(ctx->r, x, bytes, hipMemcpyHostToDevice, ctx->stream));
  esf_status st = pressureApplyDevice(ctx, ctx->r, ctx->s);
  if (st) return st;
  ESF_HIP(hipMemcpyAsync(y, ctx->s, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_pressure_diag(esf_ctx* ctx, double* dOut) {
  if (!ctx || !dOut) return ESF_ERR_INVALID;
  esf_status st = pressureRhs(ctx);
  if (st) return st;
  ESF_HIP(hipMemcpyAsync(dOut, ctx->D, 8 * ctx->d.cells, hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_solid_apply(esf_ctx* ctx, const double* x, double* y) {
  if (!ctx || !x || !y) return ESF_ERR_INVALID;
  const size_t bytes = 24 * (size_t)ctx->nSolid;
  if (!bytes) return ESF_OK;
  ESF_HIP(hipMemcpyAsync(ctx->stmp, x, bytes, hipMemcpyHostToDevice, ctx->stream));
  esf_status st = solidApplyDevice(ctx, ctx->stmp, ctx->sr);
  if (st) return st;
  ESF_HIP(hipMemcpyAsync(y, ctx->sr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_pressure_solve(esf_ctx* ctx, const double* b, double* x, double eps, int maxIt,
                              int* iters, double* relres) {
  if (!ctx || !b || !x) return ESF_ERR_INVALID;
  const size_t bytes = 8 * ctx->d.cells;
  ESF_HIP(hipMemcpyAsync(ctx->b, b, bytes, hipMemcpyHostToDevice, ctx->stream));
  ESF_HIP(hipMemcpyAsync(ctx->lambda, x, bytes, hipMemcpyHostToDevice, ctx->stream));
  int it = 0;
  bool conv = false;
  double rr = 0;
  esf_status st = pressureDiagFlush(ctx);
  if (st) return st;
  st = crSolve<double>(ctx, 0, ctx->minv, ctx->b, ctx->D, ctx->Dinv, ctx->lambda, ctx->r,
                                  ctx->r2, ctx->pdir, ctx->s, ctx->s2, ctx->Ap, ctx->Ap2, eps, maxIt, 3);
  if (st) return st;
  ESF_HIP(hipMemcpyAsync(x, ctx->lambda, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ESF_HIP(streamSync(ctx));
  solveReport(ctx->hostSc[3], maxIt, &it, &conv, &rr);
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return conv ? ESF_OK : ESF_NOT_CONVERGED;
}

esf_status esf_pressure_solve_resident(esf_ctx* ctx, double eps, int maxIt, int* iters,
                                       int* refinements, double* relres, double* deviceMs) {
  if (!ctx || !(eps > 0) || maxIt < 1) return ESF_ERR_INVALID;
  int it = 0, outer = 0, inner = 0;
  bool conv = false;
  double rr = 0;
  ESF_HIP(hipEventRecord(ctx->ev[4], ctx->stream));
  esf_status st;
  if (ctx->p.precision == ESF_MIXED) {
    st = mixedSolve(ctx, eps, maxIt, &it, &conv, &rr, &outer, &inner);
  } else {
    st = crSolve<double>(ctx, 0, ctx->minv, ctx->b, ctx->D, ctx->Dinv, ctx->lambda, ctx->r, ctx->r2,
                         ctx->pdir, ctx->s, ctx->s2, ctx->Ap, ctx->Ap2, eps, maxIt, 3);
  }
  if (st) return st;
  ESF_HIP(hipEventRecord(ctx->ev[5], ctx->stream));
  ESF_HIP(eventSync(ctx, ctx->ev[5]));
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]);
  if (ctx->p.precision != ESF_MIXED) solveReport(ctx->hostSc[3], maxIt, &it, &conv, &rr);
  ctx->lambdaInit = true;
  if (iters) *iters = it;
  if (refinements) *refinements = outer;
  if (relres) *relres = rr;
  if (deviceMs) *deviceMs = ms;
  return conv ? ESF_OK : ESF_NOT_CONVERGED;
}

esf_status esf_corotation_batch(int n, double lambda, double mu, const double* F,
                                const double* dF, double* out, double* energy, double* dP) {
  if (n <= 0 || !F || !dF || !out || !energy || !dP) return ESF_ERR_INVALID;
  double *dFm = nullptr, *ddF = nullptr, *dout = nullptr, *den = nullptr, *ddP = nullptr;
  auto release = [&] {
    for (double* q : {dFm, ddF, dout, den, ddP})
      if (q) hipFree(q);
  };
  if (hipMalloc(&dFm, 72L * n) != hipSuccess || hipMalloc(&ddF, 72L * n) != hipSuccess ||
      hipMalloc(&dout, 288L * n) != hipSuccess || hipMalloc(&den, 8L * n) != hipSuccess ||
      hipMalloc(&ddP, 72L * n) != hipSuccess) {
    release();
    return ESF_ERR_OOM;
  }
  hipError_t e = hipMemcpy(dFm, F, 72L * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ddF, dF, 72L * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(kCorotBatch, dim3(cdiv(n, 128)), dim3(128), 0, 0, n, lambda, mu, dFm,
                       ddF, dout, den, ddP);
    e = hipDeviceSynchronize();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dout, 288L * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(energy, den, 8L * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(dP, ddP, 72L * n, hipMemcpyDeviceToHost);
  release();
  return e == hipSuccess ? ESF_OK : ESF_ERR_DEVICE;
}

esf_status esf_stage_times(esf_ctx* ctx, double* ms, int n, int reset) {
  if (!ctx || !ms) return ESF_ERR_INVALID;
  for (int i = 0; i < n && i < 16; i++) ms[i] = ctx->stageMs[i];
  if (reset)
    for (double& v : ctx->stageMs) v = 0;
  return ESF_OK;
}

esf_status esf_ctx_device_bytes(const esf_ctx* ctx, size_t* bytes) {
  if (!ctx || !bytes) return ESF_ERR_INVALID;
  *bytes = ctx->deviceBytes;
  return ESF_OK;
}

esf_status esf_synchronize(esf_ctx* ctx) {
  if (!ctx) return ESF_ERR_INVALID;
  ESF_HIP(streamSync(ctx));
  return ESF_OK;
}

esf_status esf_bench_pressure_apply(esf_ctx* ctx, int fp32, int reps, double* avg_ms) {
  if (!ctx || !avg_ms || reps < 1) return ESF_ERR_INVALID;
  return benchPressureApply(ctx, fp32, reps, avg_ms);
}

esf_status esf_bench_copy(esf_ctx* ctx, long bytes, int reps, double* avg_ms) {
  if (!ctx || !avg_ms || reps < 1 || bytes < 16) return ESF_ERR_INVALID;
  return benchCopy(ctx, bytes, reps, avg_ms);
}

esf_status esf_set_stage_timing(esf_ctx* ctx, int on) {
  if (!ctx) return ESF_ERR_INVALID;
  ctx->stageTiming = on != 0;
  return ESF_OK;
}

esf_status esf_set_kernel_timing(esf_ctx* ctx, int on) {
  if (!ctx) return ESF_ERR_INVALID;
  ctx->kernelTiming = on != 0;
  return ESF_OK;
}

esf_status esf_kernel_timing(esf_ctx* ctx, double* total_ms, long* counts, int reset) {
  if (!ctx) return ESF_ERR_INVALID;
  ESF_HIP(streamSync(ctx));
  for (size_t i = 0; i < ctx->timedUsed; i++) {
    float ms = 0;
    const auto& t = ctx->timed[i];
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      ctx->kernelMs[t.cls] += ms;
      ctx->kernelCount[t.cls]++;
    }
  }
  ctx->timedUsed = 0;
  for (int c = 0; c < ESF_KERNEL_CLASSES; c++) {
    if (total_ms) total_ms[c] = ctx->kernelMs[c];
    if (counts) counts[c] = ctx->kernelCount[c];
    if (reset) {
      ctx->kernelMs[c] = 0;
      ctx->kernelCount[c] = 0;
    }
  }
  return ESF_OK;
}

}  // extern "C"
