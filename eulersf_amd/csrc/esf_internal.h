// esf_internal.h -- device context and helpers shared by the libesf.so
// translation units.  gfx950 only: wave64, FP64 arithmetic compiled with
// -ffp-contract=off so every stage that mirrors a reference expression
// rounds exactly as the reference's left-to-right evaluation does.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/esf.h"

namespace esf {

// ---------------------------------------------------------------------------
// Grid dimensions passed by value to every kernel
// ---------------------------------------------------------------------------
struct Dims {
  int X, Y, Z;       // nodes
  int CX, CY, CZ;    // cells = nodes - 1
  long slab, nodes;  // X*Y, X*Y*Z
  long cslab, cells; // CX*CY, CX*CY*CZ
  double dh, dhInv;
  double o0, o1, o2;
};

// solids per context (scene files of the reference use one; the isolated
// 256^3 system uses eight)
constexpr int kMaxSolids = 8;
constexpr int kCflBlocks = 256;  // most blocks of the CFL minimum (grid-stride)
// pinned ints of a context (hostFlags): the solid DOF count, the solids'
// boxes (6 each from kHostBoxes) and their valid-cell counts
constexpr int kHostDofCount = 0;
constexpr int kHostBoxes = 8;
constexpr int kHostValidCount = kHostBoxes + 6 * kMaxSolids;
constexpr int kHostFlagInts = kHostValidCount + kMaxSolids;

// Per-solid box, in node coordinates (DEFOR_SOLID_3D.h:138-140)
struct Box {
  int tMin[3], tMax[3];  // tight box  (_tightMin/_tightMax)
  int bMin[3], bMax[3];  // extended box (_xmin.._xmax)
};

// Scalars of one Krylov solve, resident in device memory.  The finisher
// after each reducing kernel writes them; the next kernel reads them.
struct CrScalars {
  double delta;    // r . s
  double beta;     // delta_new / delta_old
  double alpha;    // delta / (Ap . D^-1 Ap)
  double bnorm;    // |b|
  double resnorm;  // last true residual |Ax - b|
  double eps;      // relative tolerance
  double aux;      // solid CR: 1 when a check riding along with the next
                   // iteration converged (x is then in the backup)
  int done;        // 1 = converged
  int iter;        // iteration count at convergence
  int pad[2];
};

// node index -> (x, y, z).  A context holds at most 2^28 nodes (checked when
// it is created), so 32-bit division suffices.
__device__ inline void nodeXYZ(const Dims& d, long n, int& x, int& y, int& z) {
  const int ni = (int)n, q = ni / d.X;
  x = ni - q * d.X;
  y = q % d.Y;
  z = q / d.Y;
}

// a solve's scalar block at its start: tolerance eps, nothing done yet
__device__ inline void initScalarBlock(CrScalars* sc, double eps) {
  CrScalars h{};
  h.eps = eps;
  h.iter = -1;
  *sc = h;
}

// per-solid device state: DEFOR_SOLID_3D members kept on the GPU
struct DevSolid {
  double lambda = 0, mu = 0, qcm = 0;  // qcm = _quarterCellSolidMass
  int rX = 0, rY = 0, rZ = 0;
  double rdh = 0, ro[3] = {0, 0, 0};
  double* rest = nullptr;              // rest mass FIELD_3Df
  double *u = nullptr, *unew = nullptr, *X = nullptr;
  double *vel = nullptr, *velws = nullptr;  // 3*nodes
  // xLazy: the buffer X holds _X = x - u only inside the node range xBox
  // ([lo, hi) per axis, where the step wrote it); materialiseX writes the
  // rest before X is handed to a caller, u is overwritten, or a mass pass
  // reads outside xBox
  bool xLazy = false;
  int xBox[6] = {0, 0, 0, 0, 0, 0};
  // velws against vel: 0 equal at every node, 1 equal outside the node
  // range wsBox ([lo, hi) per axis: the copy that changed vel since), 2 not
  // known (an upload or a FLIP transfer changed either)
  int wsState = 0;
  int wsBox[6] = {0, 0, 0, 0, 0, 0};
  double *mass = nullptr, *detFInv = nullptr;
  // _phi: the solid's distance field in cells (computeSDF); phi2 is the
  // second buffer of its relaxation
  double *phi = nullptr, *phi2 = nullptr;
  double* qfrac = nullptr;                  // 8*nodes
  int8_t* valid = nullptr;                  // cells
  Box box;
  // cells where qfrac, valid and detFInv may be nonzero (the tight box of
  // the last mass pass; [lo, hi) per axis), or anywhere after an upload of
  // the mass fields: the next mass pass clears only that region
  int dirty[6] = {0, 0, 0, 0, 0, 0};
  bool dirtyAll = false;
  // compact list of cells with _valid == 1 inside the tight box
  int nValid = 0, capValid = 0;
  int* cellList = nullptr;   // node index of each valid cell's min corner
  int* cellSlot = nullptr;   // cells -> slot or -1
  // cells where cellSlot may be >= 0 (the tight box's cell range at the
  // last list build; [lo, hi) per axis, the whole grid until the first)
  int slotSpan[6] = {0, 0, 0, 1 << 30, 1 << 30, 1 << 30};
  int slotScan[6] = {0, 0, 0, 0, 0, 0};  // the cells the queued list scan walks
  bool listPending = false;  // list scan queued, its count not read yet
  // boxes that contain the ones the last box scan finds (its scan range
  // widened as computeSDF widens), and whether the host copy `box` still
  // waits for that scan's result (the device copy is current)
  Box boxBound{};
  bool boxPending = false;
  // nodes where phi (and its relaxation buffer) may differ from `large`:
  // the extended box of the last phi pass, the whole grid after an upload
  int phiSpan[6] = {0, 0, 0, 1 << 30, 1 << 30, 1 << 30};
  // phi holds its relaxed value only where the next mass pass reads it (the
  // tight box; the step's own pass); phiWhole relaxes the whole box before
  // phi is handed out or read elsewhere, or the mass or the boxes change
  bool phiPartial = false;
  double* cellMat = nullptr; // per slot: R(9) L(9) F(9) Jinv(1)
  double* cellOut = nullptr; // per slot: 24 values (force, K x, diag K)
  double* cellDiag = nullptr;// per slot: 24 diagonal stiffness entries
  int* cornerDof = nullptr;  // per slot and corner: 3 * solid index, or -1 if masked
  bool listDirty = true;
  // surface mesh (DEFOR_SOLID_3D::_surfaceMesh): rest pose and current
  // vertices (3 doubles each), triangles (3 vertex indices), triangle boxes
  int meshV = 0, meshT = 0;
  double* meshRest = nullptr;
  double* meshVert = nullptr;
  int* meshTri = nullptr;
  double* meshBox = nullptr;
  std::vector<int> meshTriHost;
  // FLIP material particles (DEFOR_SOLID_3D::_materialParticles): position
  // and velocity, 3 doubles each; their bin (cell) and the bins' contents
  // a deformation gradient per valid cell set by esf_solid_set_deformation
  // (9 doubles each, valid-cell order), used by the stress stage while
  // nFset equals the valid-cell count
  double* Fset = nullptr;
  long nFset = -1;
  long nPart = 0, capPart = 0;
  double *pPos = nullptr, *pVel = nullptr;
  int *pBin = nullptr, *binItems = nullptr;
};

struct SceneFiber;  // esf_batch.hip: a scene's step inside esf_step_batch

}  // namespace esf

struct esf_ctx {
  esf::Dims d;
  esf_params p;
  double vortEps = 0, defaultDt = 0, dt = 0;
  hipStream_t stream = nullptr;
  // the step runs the fluid's vorticity and body forces on `side` while the
  // first solid solve runs on `stream` (evFork / evJoin order the two)
  hipStream_t side = nullptr;
  hipEvent_t evFork = nullptr, evJoin = nullptr;
  bool joinPending = false;
  int device = 0;
  std::string err;
  // grid fields
  double *vel = nullptr, *velws = nullptr, *vort = nullptr;  // 3*nodes
  double *intensity = nullptr, *intws = nullptr, *ws = nullptr;
  double *solidMass = nullptr, *nodeFrac = nullptr, *qfracSum = nullptr;
  int* sidx = nullptr;        // nodes -> solid DOF node index or -1
  int* solidNodes = nullptr;  // solid DOF node -> node index
  int nSolid = 0, capSolid = 0;
  // generation of the solids' masses and quarter fractions (bumped by every
  // write of them); the gathered fields above are rebuilt only when it moved
  long massGen = 1, gatheredGen = 0;
  int gatheredSolids = -1;
  // nodes where the gathered mass and fraction may be nonzero ([lo, hi) per
  // axis; the whole grid until the first gather)
  int gatherBox[6] = {0, 0, 0, 1 << 30, 1 << 30, 1 << 30};
  int dofScan[6] = {0, 0, 0, 0, 0, 0};  // the nodes the queued DOF numbering walks
  bool dofPending = false;  // DOF numbering queued, its count not read yet
  bool tightLists = false;  // lists grow to exactly their counts (ESF_TIGHT_LISTS)
  bool phiAsLeft = false;   // downloads of phi as a step left it (ESF_PHI_AS_LEFT)
  bool diagPending = false; // the pressure diagonal of the current masses not formed yet
  esf::Box* dBoxes = nullptr;  // device copies of the solids' boxes (kMaxSolids)
  uint64_t* nodeMask = nullptr; // per solid node, byte b: solid b's cells with K rows there
  double* inbox = nullptr;      // solid b at 24 capSolid b: 24 doubles per node
  int inboxSolids = 0;          // solids the inbox was sized for
  int* scanBlocks = nullptr;  // scan scratch
  // solve: node-based
  double *v0 = nullptr, *minv = nullptr;  // 3*nodes, nodes
  // mixed form: the velocity correction of the last solve at every interior
  // node (solid DOFs and fluid nodes), the next solid correction solve's
  // warm start
  double* dvPrev = nullptr;
  // fast sweeping and particles: staging of the host-array entry points,
  // the sweep passes' flags and coefficients, the particle bins (counts,
  // starts, fill cursors: one set shared by the solids, used in turn), the
  // weighted sums of particleToGrid and the MT19937 draws of
  // initMaterialParticles (rand() values, the same every step)
  double *swMass = nullptr, *swPhi = nullptr, *swField = nullptr;
  int* swFirst = nullptr;
  double2* swCoef = nullptr;
  size_t swCoefCap = 0;
  int *binCount = nullptr, *binStart = nullptr, *binCursor = nullptr, *binBlocks = nullptr;
  double* randTable = nullptr;
  long randCap = 0;
  float* minv32 = nullptr;
  // solve: cell-based (pressure)
  double *c = nullptr, *b = nullptr, *D = nullptr, *lambda = nullptr,
         *lambda2 = nullptr;
  double* Dinv = nullptr;  // 1 / D (the CR update multiplies by it)
  float* Dinv32 = nullptr;
  double *r = nullptr, *r2 = nullptr, *pdir = nullptr, *s = nullptr, *Ap = nullptr;
  double *s2 = nullptr, *Ap2 = nullptr;  // second buffers of s and Ap (CR loop)
  float *D32 = nullptr, *e32 = nullptr, *r32 = nullptr, *r32b = nullptr,
        *p32 = nullptr, *s32 = nullptr, *Ap32 = nullptr, *rhs32 = nullptr;
  float *s32b = nullptr, *Ap32b = nullptr;
  bool lambdaInit = false;
  // solid DOF vectors (3*capSolid)
  double *sForce = nullptr, *sMv = nullptr, *sDamped = nullptr,
         *sKdiag = nullptr, *sAdiag = nullptr, *sRhs = nullptr,
         *sV0 = nullptr, *sDv = nullptr, *sB = nullptr;
  double *sr = nullptr, *sp = nullptr, *ss = nullptr, *sAp = nullptr,
         *stmp = nullptr, *sSd = nullptr, *sApd = nullptr;  // sSd = D^-1 s, sApd = D^-1 Ap
  double* sXb = nullptr;        // solid CR: x before the iteration a check rides along with
  double* inboxX = nullptr;     // rows of K x for that check, laid out as inbox
  // reductions
  double* partials = nullptr;  // per-block partial sums
  double* energyPart = nullptr;  // per-block energy sums of the body-force pass
  double* sPartials = nullptr; // solid CR: 4 sets of 4 x node-pass blocks
  esf::CrScalars* sc = nullptr;   // device scalars (several solves)
  double* dmisc = nullptr;        // device misc scalars (dt, cfl min, energy)
  double* meshOut = nullptr;      // mesh broad phase: root boxes, then pair flags
  int* hostFlags = nullptr;       // pinned: counts read back by the host
  double* hostCfl = nullptr;      // pinned: the CFL kernel's block minima
  hipEvent_t evCfl = nullptr;     // their copy has landed
  hipEvent_t cflMark = nullptr;   // the event cflFinish waits for (evCfl alone)
  // after a step's gather stage: what the host reads at the step's end is
  // ready there (the advection and recovery behind it may still run)
  hipEvent_t evSettle = nullptr;
  double* hostMesh = nullptr;     // pinned: root boxes and pair flags of the broad phase
  esf::CrScalars* hostRing = nullptr;  // pinned ring of polled scalar blocks
  int crQueued = 0;  // iterations the last pressure CR solve queued
  // pressure SpMV launch shape (fixed at creation): cell layers per block,
  // load sets in flight for the CR update and for the other modes
  int spmvLayers = 16, spmvLayersWide = 8, spmvDepthUpdate = 1, spmvDepthOther = 2;
  // device allocations held by the context (admission control of served scenes)
  std::unordered_map<void*, size_t> allocs;
  size_t deviceBytes = 0;
  // pinned copies of the Krylov scalars, one slot per solve of a step (0:
  // solid 1, 1: pressure, 2: solid 2, 3: scratch, 4: inner FP32); a solve
  // queues its copy and the host reads it after the step's next sync
  static constexpr int kReportSlots = 8;
  esf::CrScalars* hostSc = nullptr;
  int nPartials = 0;
  std::vector<esf::DevSolid> solids;
  // timing
  hipEvent_t ev[24];
  double stageMs[16] = {0};
  bool stageTiming = false;
  // kernel timing: event pairs recorded around Krylov kernel launches
  // without synchronising, read back by esf_kernel_timing()
  bool kernelTiming = false;
  struct TimedLaunch {
    hipEvent_t a, b;
    int cls;
  };
  std::vector<TimedLaunch> timed;
  size_t timedUsed = 0;
  double kernelMs[ESF_KERNEL_CLASSES] = {0};
  long kernelCount[ESF_KERNEL_CLASSES] = {0};
  // set while the context's step runs inside esf_step_batch: launches go to
  // the batch scheduler, waits let the other scenes go on
  esf::SceneFiber* fiber = nullptr;
};

namespace esf {

// error helpers
#define ESF_HIP(call)                                                   \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) {                                             \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);      \
      return ESF_ERR_DEVICE;                                            \
    }                                                                   \
  } while (0)

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;

// Cross-lane moves of doubles with DPP (two 32-bit v_mov_dpp, VALU speed,
// no LDS traffic).  Controls used (gfx9 DPP encoding): quad_perm [1,0,3,2]
// = 0xB1 (lane ^ 1), quad_perm [2,3,0,1] = 0x4E (lane ^ 2), row_half_mirror
// = 0x141 (lane l <-> 7 - l within each 8), row_mirror = 0x140 (l <-> 15 - l
// within each 16).  Every lane of the pattern has a source, so no lane
// reads a disabled one.
template <int CTRL>
__device__ inline double dppMov(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ inline double readlaneD(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Sum over each aligned group of 8 lanes, bitwise the same in all 8 (each
// step adds a pair and its partner adds the same pair commuted).  The
// pairing is that of an xor butterfly over lane bits 0, 1, 2.
__device__ inline double sum8(double v) {
  v = v + dppMov<0xB1>(v);
  v = v + dppMov<0x4E>(v);
  v = v + dppMov<0x141>(v);
  return v;
}

// Sum over the 64 lanes of the wave, uniform (bitwise the same in every
// lane): 16-lane sums by DPP, then the four rows added in a fixed order.
// All 64 lanes must be active.
__device__ inline double waveSum(double v) {
  v = sum8(v);
  v = v + dppMov<0x140>(v);
  return (readlaneD(v, 0) + readlaneD(v, 16)) + (readlaneD(v, 32) + readlaneD(v, 48));
}

// wave sums, then a fixed-order sum across the block's waves; the result
// is valid in thread 0.  Deterministic for a fixed launch shape.  All
// threads of the block must call it.
__device__ inline double blockSum(double v, double* lds) {
  v = waveSum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += lds[w];
  return r;
}

// Grid-wide sums.  Each block of a reducing kernel writes its K block sums
// with plain stores, block-major ([k * nb + block]); a one-block finisher
// kernel launched next on the same stream adds them in block order (fixed
// order: deterministic for a fixed launch shape).  The kernel boundary
// orders the partials for the finisher.  A last-block reduction inside the
// producer would need an agent-scope release fence per block, which MI355X
// serialises at ~20 ns per block (1024 blocks: ~22 us per reduction).
// All K values in one pass: wave sums, one LDS exchange, thread k adds
// quantity k over the waves in ascending order.  lds: K * (waves per
// block) doubles.  All threads of the block must call it.
// Work-group renumbering for the 8 XCDs: the dispatcher deals the
// work-groups of a launch to the XCDs in turn (group b to XCD b mod 8), and
// each XCD has its own L2.  xcdBlock() gives each XCD one contiguous run of
// logical indices instead, so blocks that read overlapping data (stencil
// neighbours, shared corner nodes) find it in the same L2.  A bijection on
// [0, gridDim.x) for every grid size; 1-D grids only.
__device__ __forceinline__ int xcdBlock() {
  const int b = blockIdx.x, nb = gridDim.x, x = b & 7;
  const int q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + (b >> 3);
}

// the same with the block's slot given (a kernel numbering its blocks by
// xcdBlock() keeps its partials in that order)
template <int K>
__device__ inline void blockPartialsAt(const double (&mine)[K], double* partials, double* lds,
                                       unsigned nb, unsigned bid) {
  double w[K];
#pragma unroll
  for (int k = 0; k < K; k++) w[k] = waveSum(mine[k]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; k++) lds[K * wid + k] = w[k];
  __syncthreads();
  if (threadIdx.x < K) {
    double r = 0;
    for (int v = 0; v < nw; v++) r += lds[K * v + threadIdx.x];
    partials[(size_t)threadIdx.x * nb + bid] = r;
  }
}

// in the finisher (one block, up to 1024 threads): the K totals of the
// producer's nb blocks, valid in thread 0.  Every load of a round of 8
// strides is issued for all K sums before any is added, then one LDS
// exchange serves all K; per sum the order of additions is that of
// blockSum over per-thread sums taken in stride order.  lds: 16 * K doubles.
template <int K>
__device__ inline void finishSums(int nb, const double* partials, double* lds,
                                  double (&tot)[K]) {
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0;
  for (int base = threadIdx.x; base < nb; base += 8 * blockDim.x) {
    double c[K][8];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int i = base + u * blockDim.x;
        c[k][u] = i < nb ? partials[(size_t)k * nb + i] : 0.0;
      }
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int u = 0; u < 8; u++) acc[k] += c[k][u];
  }
  double w[K];
#pragma unroll
  for (int k = 0; k < K; k++) w[k] = waveSum(acc[k]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; k++) lds[16 * k + wid] = w[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; k++) {
      double r = 0;
      for (int v = 0; v < nw; v++) r += lds[16 * k + v];
      tot[k] = r;
    }
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// solveDistance (DEFOR_SOLID_3D.cpp:369-395): the grid Eikonal update from
// the three upwind neighbours p, q, r (x, y, z), sorted; lowers v only.  For
// finite inputs the candidate at i = 0 is always valid, so the reference's
// loop never steps below 0.
__device__ inline double solveDistance(double p, double q, double r, double v) {
  double d[3];
  d[0] = fmin(p, fmin(q, r));
  d[2] = fmax(p, fmax(q, r));
  d[1] = p + q + r - d[0] - d[2];
  double b[3], c[3];
  b[0] = d[0];
  c[0] = d[0] * d[0] - 1.0;
#pragma unroll
  for (int x = 1; x < 3; x++) {
    b[x] = d[x] + b[x - 1];
    c[x] = d[x] * d[x] + c[x - 1];
  }
  int i = 2;
  double discrim = b[2] * b[2] - 3.0 * c[2];
  double nd = (b[2] + sqrt(discrim)) / 3.0;
  while ((discrim < 0.0 || nd < d[i]) && i > 0) {
    i--;
    discrim = b[i] * b[i] - (double)(i + 1) * c[i];
    if (discrim > 0.0) nd = (b[i] + sqrt(discrim)) / (double)(i + 1);
  }
  return nd < v ? nd : v;
}

// what a finished Krylov solve reports, from its scalar block
inline void solveReport(const CrScalars& h, int maxIt, int* iters, bool* converged,
                        double* relres) {
  *converged = h.done == 1;
  *iters = h.done ? h.iter : maxIt;
  *relres = h.bnorm > 0 ? h.resnorm / h.bnorm : h.resnorm;
}

// kernel classes for esf_kernel_timing: SpMV modes 0-5 (FP64), 6-11 (FP32),
// 12/13 unused, solid cell pass 14, solid node pass 15, the same two
// carrying the true-residual check 16 / 17
//
// esf_api.hip: a device buffer of `count` elements held by the context
// (replacing the one at *p, which is freed); zero-filled
esf_status devAllocBytes(esf_ctx* ctx, void** p, size_t bytes);
template <typename T>
inline esf_status devAlloc(esf_ctx* ctx, T** p, size_t count) {
  return devAllocBytes(ctx, (void**)p, count * sizeof(T));
}

// esf_sweep.hip: rounds of sweepPhi over box (xmin, ymin, zmin, xmax,
// ymax, zmax) on device fields; extendVectorField (four rounds of
// sweepVectorField) of a device node field over box
esf_status sweepPhiRounds(esf_ctx* ctx, const int box[6], const double* mass, double* phi,
                          int rounds);
esf_status extendDevice(esf_ctx* ctx, const int box[6], const double* mass, const double* phi,
                        double* field);

// esf_grid.hip: _X = x - u at every node into the solid's X buffer; phi
// relaxed over the whole extended box where a step left it partial
esf_status materialiseX(esf_ctx* ctx, DevSolid& so);
esf_status phiWhole(esf_ctx* ctx, int solid);

// esf_flip.hip: the FLIP round trip of one solid
esf_status initParticles(esf_ctx* ctx, int solid);
esf_status gridToParticle(esf_ctx* ctx, int solid);
esf_status particleToGrid(esf_ctx* ctx, int solid);
esf_status growParticlesFor(esf_ctx* ctx, int solid, long n);

// esf_mesh.hip: surface meshes (advectSurfaceMesh for every solid with a
// mesh; the mesh broad phase: root boxes, 6 doubles per solid, and one
// collide flag per solid pair i < j)
esf_status advectMeshes(esf_ctx* ctx);
esf_status meshBroadPhaseEnqueue(esf_ctx* ctx);
void meshBroadPhaseRead(esf_ctx* ctx, std::vector<double>& roots, std::vector<int>& pairs);
// the broad phase of stepImplicit in two halves (esf_scene.cpp): queue
// (mesh kernels and their read-back, when every solid has a mesh), then,
// after the stream has passed the read-back, the CFL alpha
esf_status broadPhaseEnqueue(esf_ctx* ctx, bool* meshes);
esf_status broadPhaseFinish(esf_ctx* ctx, bool meshes, double* alpha);

}  // namespace esf

#include "esf_launch.h"
