// esf_scene.cpp -- host C++ around the C ABI: the reference's cfg format and
// scene setup (projects/rubberBall/main.cpp:32-160), and its frame loop
// (:174-205), driving the device stages of libesf.so.
#include <zlib.h>

#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/esf.h"
#include "esf_cfg.h"
#include "../csrc/esf_internal.h"

namespace esf {

// ---------------------------------------------------------------------------
// Cfg -- SIMPLE_PARSER semantics (src/util/SIMPLE_PARSER.cpp:64-236): lines
// "key = value", '#' comments in column 0, keys and values lowercased,
// numbers through atof, integers truncated from atof, vectors split on ','.
// ---------------------------------------------------------------------------
static std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < s.size() && std::isspace((unsigned char)s[b])) b++;
  while (e > b && std::isspace((unsigned char)s[e - 1])) e--;
  return s.substr(b, e - b);
}
static std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

bool Cfg::parse(const std::string& path, std::string* err) {
  std::ifstream f(path);
  if (!f.good()) {
    if (err) *err = "cannot open " + path;
    return false;
  }
  std::string line;
  int lineNo = 1;
  while (std::getline(f, line)) {
    if (line.size() > 511) line.resize(511);  // the reference reads 512-byte lines
    if (line.empty()) continue;
    if (line[0] == '#') continue;
    size_t pos = line.find('=');
    if (pos != std::string::npos) {
      vals[lower(trim(line.substr(0, pos)))] = lower(trim(line.substr(pos + 1)));
    } else if (!trim(line).empty()) {
      if (err) *err = "unable to parse line " + std::to_string(lineNo) + ": '" + line + "'";
      return false;
    }
    lineNo++;
  }
  return true;
}

bool Cfg::has(const std::string& k) const { return vals.count(lower(k)) != 0; }
double Cfg::getFloat(const std::string& k, double def) const {
  auto it = vals.find(lower(k));
  return it == vals.end() ? def : std::atof(it->second.c_str());
}
int Cfg::getInt(const std::string& k, int def) const {
  auto it = vals.find(lower(k));
  return it == vals.end() ? def : (int)std::atof(it->second.c_str());
}
std::string Cfg::getString(const std::string& k, const std::string& def) const {
  auto it = vals.find(lower(k));
  return it == vals.end() ? def : it->second;
}
void Cfg::getVec3(const std::string& k, const double def[3], double out[3]) const {
  auto it = vals.find(lower(k));
  if (it == vals.end()) {
    for (int i = 0; i < 3; i++) out[i] = def[i];
    return;
  }
  std::stringstream ss(it->second);
  std::string tok;
  for (int i = 0; i < 3; i++) {
    out[i] = 0;
    if (std::getline(ss, tok, ',')) out[i] = std::atof(tok.c_str());
  }
}

void paramsFromCfg(const Cfg& c, double dh, esf_params* p) {
  p->dt = c.getFloat("dt", 0.01);                          // COUPLED_INTEGRATOR_3D.cpp:43
  p->vorticity_eps = c.getFloat("vorticity eps", 2.0);     // :67
  p->gravity = c.getFloat("gravity", 0);                   // :499
  p->buoyancy = c.getFloat("buoyancy", 0.1);               // :501
  p->rayleigh_alpha = c.getFloat("rayleigh alpha", 0);     // :445
  p->rayleigh_beta = c.getFloat("rayleigh beta", 0);       // :446
  p->fluid_density = c.getFloat("fluid density", 1);       // main.cpp:73
  // CONJUGATE_RESIDUAL ctor reads "cr eps", the integrator overrides it
  p->implicit_eps = c.getFloat("implicit solver eps", 1e-8);  // :73
  p->divfree_eps = c.getFloat("div free solver eps", 1e-8);   // :74
  p->cr_max_iterations = c.getInt("cr max iterations", 10);   // CONJUGATE_RESIDUAL.inl:26
  p->mass_sample_dh = c.getFloat("mass sample dh", dh / 10);  // DEFOR_SOLID_3D.cpp:62
  p->tight_border = c.getInt("tight border", 3);              // :243
  p->extrapolation_cells = c.getInt("extrapolation cells", 20);  // :244
  p->inner_eps = c.getFloat("mixed inner eps", 1e-4);
  p->particles_per_cell = c.getInt("number of particles per material cell", 8);  // :336
  p->flip_alpha = c.getFloat("flip alpha", 0.95);                                 // :901
  p->flip_cutoff = c.getInt("flip cutoff", 3);                                    // :1101
  p->wall_thickness = c.getFloat("wall thickness", dh * 1.5);                     // :949
  p->spring_stiffness = c.getFloat("contact spring stiffness", 10000);            // :950
  p->friction = c.getFloat("friction", 0.5);                                      // :952
  p->flip = c.getInt("flip particles", 0);
}

// ---------------------------------------------------------------------------
// rest mass: FIELD_3D::readGz (FIELD_3D.h:335-365) or an analytic sphere
// rasterized as projects/rasterizeSolid/main.cpp:42-72 does from an SDF
// ---------------------------------------------------------------------------
bool readMassGz(const std::string& path, RestField* f) {
  gzFile g = gzopen(path.c_str(), "rb");
  if (!g) return false;
  int dims[3];
  double hdr[4];
  bool ok = gzread(g, dims, sizeof(dims)) == (int)sizeof(dims) &&
            gzread(g, hdr, sizeof(hdr)) == (int)sizeof(hdr);
  if (ok && dims[0] > 1 && dims[1] > 1 && dims[2] > 1) {
    f->X = dims[0];
    f->Y = dims[1];
    f->Z = dims[2];
    f->dh = hdr[0];
    for (int i = 0; i < 3; i++) f->origin[i] = hdr[1 + i];
    const size_t n = (size_t)f->X * f->Y * f->Z;
    f->data.resize(n);
    ok = gzread(g, f->data.data(), (unsigned)(n * sizeof(double))) == (int)(n * sizeof(double));
  } else {
    ok = false;
  }
  gzclose(g);
  return ok;
}

void rasterizeSphere(double radius, double massDh, double density, RestField* f) {
  // mesh bounding box of a sphere centred at the rest-space origin,
  // expanded by 10 mass cells (rasterizeSolid/main.cpp:44-51)
  double bmin[3], bmax[3];
  for (int i = 0; i < 3; i++) {
    bmin[i] = -radius - massDh * 10;
    bmax[i] = radius + massDh * 10;
  }
  f->X = (int)((bmax[0] - bmin[0]) / massDh);
  f->Y = (int)((bmax[1] - bmin[1]) / massDh);
  f->Z = (int)((bmax[2] - bmin[2]) / massDh);
  f->dh = massDh;
  for (int i = 0; i < 3; i++) f->origin[i] = bmin[i];
  f->data.assign((size_t)f->X * f->Y * f->Z, 0.0);
  size_t idx = 0;
  for (int z = 0; z < f->Z; z++)
    for (int y = 0; y < f->Y; y++)
      for (int x = 0; x < f->X; x++, idx++) {
        const double p[3] = {x * massDh + bmin[0], y * massDh + bmin[1], z * massDh + bmin[2]};
        const double sd = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - radius;
        if (sd <= massDh / 2) f->data[idx] = density;
      }
}

// Eigen AngleAxisd(a pi, X) * AngleAxisd(b pi, Y) * AngleAxisd(c pi, Z)
// (rubberBall/main.cpp:133-136): the product is taken as quaternions and
// converted to a matrix.
void rotationFromCfg(const double r[3], double R[9]) {
  double q[4] = {1, 0, 0, 0};  // w x y z
  for (int axis = 0; axis < 3; axis++) {
    const double half = 0.5 * (r[axis] * M_PI);
    double a[4] = {std::cos(half), 0, 0, 0};
    a[1 + axis] = std::sin(half);
    const double w = q[0] * a[0] - q[1] * a[1] - q[2] * a[2] - q[3] * a[3];
    const double x = q[0] * a[1] + q[1] * a[0] + q[2] * a[3] - q[3] * a[2];
    const double y = q[0] * a[2] + q[2] * a[0] + q[3] * a[1] - q[1] * a[3];
    const double z = q[0] * a[3] + q[3] * a[0] + q[1] * a[2] - q[2] * a[1];
    q[0] = w;
    q[1] = x;
    q[2] = y;
    q[3] = z;
  }
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
  const double twx = tx * w, twy = ty * w, twz = tz * w;
  const double txx = tx * x, txy = ty * x, txz = tz * x;
  const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
  R[0] = 1 - (tyy + tzz);
  R[1] = txy - twz;
  R[2] = txz + twy;
  R[3] = txy + twz;
  R[4] = 1 - (txx + tzz);
  R[5] = tyz - twx;
  R[6] = txz - twy;
  R[7] = tyz + twx;
  R[8] = 1 - (txx + tyy);
}

// A triangulated sphere of radius r about the rest-space origin, the
// surface of an analytic solid sphere (the reference loads its sphere's
// surface from an OBJ file): the icosahedron with each triangle split in
// four `level` times, the new vertices pushed out onto the sphere.
static void icosphere(double r, int level, std::vector<double>& v, std::vector<int>& t) {
  const double a = (1.0 + std::sqrt(5.0)) / 2.0;
  const double base[12][3] = {{-1, a, 0}, {1, a, 0}, {-1, -a, 0}, {1, -a, 0},
                              {0, -1, a}, {0, 1, a}, {0, -1, -a}, {0, 1, -a},
                              {a, 0, -1}, {a, 0, 1}, {-a, 0, -1}, {-a, 0, 1}};
  const int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                            {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                            {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                            {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
  std::vector<std::array<double, 3>> p;
  auto push = [&](double x, double y, double z) {
    const double n = std::sqrt(x * x + y * y + z * z);
    p.push_back({x / n, y / n, z / n});
    return (int)p.size() - 1;
  };
  for (const auto& q : base) push(q[0], q[1], q[2]);
  std::vector<std::array<int, 3>> tri;
  for (const auto& f : faces) tri.push_back({f[0], f[1], f[2]});
  for (int l = 0; l < level; l++) {
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int i, int j) {
      const std::pair<int, int> key(std::min(i, j), std::max(i, j));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      const int k = push(p[i][0] + p[j][0], p[i][1] + p[j][1], p[i][2] + p[j][2]);
      mid[key] = k;
      return k;
    };
    std::vector<std::array<int, 3>> next;
    for (const auto& f : tri) {
      const int ab = midpoint(f[0], f[1]), bc = midpoint(f[1], f[2]), ca = midpoint(f[2], f[0]);
      next.push_back({f[0], ab, ca});
      next.push_back({f[1], bc, ab});
      next.push_back({f[2], ca, bc});
      next.push_back({ab, bc, ca});
    }
    tri.swap(next);
  }
  v.clear();
  for (const auto& q : p)
    for (int c = 0; c < 3; c++) v.push_back(r * q[c]);
  t.clear();
  for (const auto& f : tri)
    for (int c = 0; c < 3; c++) t.push_back(f[c]);
}

// DEFOR_SOLID_3D::setInitialState  DEFOR_SOLID_3D.cpp:126-152
void initialState(int X, int Y, int Z, double dh, const double o[3], const double R[9],
                  const double t[3], const double s[3], std::vector<double>& Xm,
                  std::vector<double>& u) {
  const size_t N = (size_t)X * Y * Z;
  Xm.resize(3 * N);
  u.resize(3 * N);
  size_t n = 0;
  for (int z = 0; z < Z; z++)
    for (int y = 0; y < Y; y++)
      for (int x = 0; x < X; x++, n++) {
        double p[3] = {(double)x, (double)y, (double)z};
        for (int c = 0; c < 3; c++) {
          p[c] *= dh;
          p[c] += o[c];
          p[c] -= t[c];
        }
        for (int r = 0; r < 3; r++) {
          const double q = (R[r] * p[0] + R[3 + r] * p[1]) + R[6 + r] * p[2];
          Xm[3 * n + r] = q / s[r];
        }
        const double pos[3] = {x * dh + o[0], y * dh + o[1], z * dh + o[2]};
        for (int c = 0; c < 3; c++) u[3 * n + c] = pos[c] - Xm[3 * n + c];
      }
}

}  // namespace esf

using namespace esf;

extern "C" esf_status esf_cfg_read(const char* path, esf_grid* g, esf_params* p,
                                   esf_scene_info* info) {
  if (!path || !g || !p) return ESF_ERR_INVALID;
  Cfg cfg;
  std::string err;
  if (!cfg.parse(path, &err)) return ESF_ERR_IO;
  g->xres = cfg.getInt("xRes", 100);
  g->yres = cfg.getInt("yRes", 100);
  g->zres = cfg.getInt("zRes", 100);
  g->dh = cfg.getFloat("dh", 0.01);
  const double zero[3] = {0, 0, 0};
  cfg.getVec3("origin", zero, g->origin);
  paramsFromCfg(cfg, g->dh, p);
  p->precision = ESF_PLAIN_FP64;
  if (info) {
    info->n_solids = cfg.getInt("number of solids", 1);
    info->frame_start = cfg.getInt("start frame", 0);
    info->frame_end = cfg.getInt("end frame", 100);
    info->frame_dt = cfg.getFloat("dt", 0.005);
    info->carpet_thickness = cfg.getFloat("smoke carpet thickness", 0.06);
  }
  return ESF_OK;
}

extern "C" esf_status esf_scene_load_cfg(const char* path, int precision, esf_ctx** out,
                                         esf_scene_info* info) {
  if (!path || !out) return ESF_ERR_INVALID;
  Cfg cfg;
  std::string err;
  if (!cfg.parse(path, &err)) {
    std::fprintf(stderr, "esf: %s\n", err.c_str());
    return ESF_ERR_IO;
  }
  // APPLICATION::init  rubberBall/main.cpp:63-97
  esf_grid g{};
  g.xres = cfg.getInt("xRes", 100);
  g.yres = cfg.getInt("yRes", 100);
  g.zres = cfg.getInt("zRes", 100);
  g.dh = cfg.getFloat("dh", 0.01);
  const double zero[3] = {0, 0, 0}, one[3] = {1, 1, 1};
  cfg.getVec3("origin", zero, g.origin);
  esf_params p{};
  paramsFromCfg(cfg, g.dh, &p);
  p.precision = precision;
  esf_ctx* ctx = nullptr;
  esf_status st = esf_ctx_create(&ctx, &g, &p);
  if (st) return st;
  const int nSolids = cfg.getInt("number of solids", 1);
  for (int i = 0; i < nSolids; i++) {
    const std::string idx = std::to_string(i);
    const std::string base =
        cfg.getString("solid path " + idx, "") + cfg.getString("solid name " + idx, "");
    const double density = cfg.getFloat("solid mass density " + idx, 1);
    RestField rest;
    if (!readMassGz(base + ".mass", &rest)) {
      if (!cfg.has("solid sphere radius " + idx)) {
        std::fprintf(stderr, "esf: no rest mass %s.mass and no 'solid sphere radius %d'\n",
                     base.c_str(), i);
        esf_ctx_destroy(ctx);
        return ESF_ERR_IO;
      }
      rasterizeSphere(cfg.getFloat("solid sphere radius " + idx, 0.1),
                      cfg.getFloat("sphere mass dh " + idx, p.mass_sample_dh), density, &rest);
    }
    if (cfg.getFloat("solid plastic yield " + idx, -1) > 0) {
      std::fprintf(stderr, "esf: plastic solids are not supported by the drop-in\n");
      esf_ctx_destroy(ctx);
      return ESF_ERR_INVALID;
    }
    int id = -1;
    st = esf_solid_add(ctx, cfg.getFloat("material lambda " + idx, 100),
                       cfg.getFloat("material mu " + idx, 100), density, rest.X, rest.Y, rest.Z,
                       rest.dh, rest.origin, rest.data.data(), &id);
    if (st) {
      esf_ctx_destroy(ctx);
      return st;
    }
    double t[3], s[3], r[3], R[9], v0[3];
    cfg.getVec3("solid translation " + idx, zero, t);
    cfg.getVec3("solid scaling " + idx, one, s);
    cfg.getVec3("solid rotation " + idx, zero, r);
    cfg.getVec3("solid initial velocity " + idx, zero, v0);
    rotationFromCfg(r, R);
    std::vector<double> Xm, u;
    initialState(g.xres, g.yres, g.zres, g.dh, g.origin, R, t, s, Xm, u);
    const size_t N = (size_t)g.xres * g.yres * g.zres;
    std::vector<double> vel(3 * N);
    for (size_t n = 0; n < N; n++)
      for (int c = 0; c < 3; c++) vel[3 * n + c] = v0[c];
    if ((st = esf_upload(ctx, ESF_S_X, id, Xm.data(), Xm.size() * 8)) ||
        (st = esf_upload(ctx, ESF_S_U, id, u.data(), u.size() * 8)) ||
        (st = esf_upload(ctx, ESF_S_VELOCITY, id, vel.data(), vel.size() * 8))) {
      esf_ctx_destroy(ctx);
      return st;
    }
    // surface mesh (main.cpp:140-152): the OBJ's vertices are the rest pose
    // (OBJ.cpp:255); the current vertices are scaled, rotated, translated.
    // Without "<name>.obj" an analytic sphere gets a triangulated sphere
    // ("solid sphere mesh level i" subdivisions, default 2; negative: none).
    std::vector<double> mrest;
    std::vector<int> mtri;
    int nv = 0, nt = 0;
    if (esf_obj_read((base + ".obj").c_str(), &nv, &nt, nullptr, 0, nullptr, 0) == ESF_OK &&
        nv > 0 && nt > 0) {
      mrest.resize(3L * nv);
      mtri.resize(3L * nt);
      esf_obj_read((base + ".obj").c_str(), &nv, &nt, mrest.data(), nv, mtri.data(), nt);
    } else if (cfg.has("solid sphere radius " + idx)) {
      const int level = cfg.getInt("solid sphere mesh level " + idx, 2);
      if (level >= 0)
        icosphere(cfg.getFloat("solid sphere radius " + idx, 0.1), std::min(level, 6), mrest,
                  mtri);
    }
    if (!mrest.empty()) {
      std::vector<double> mv(mrest.size());
      for (size_t k = 0; k < mrest.size() / 3; k++) {
        double q[3];
        for (int c = 0; c < 3; c++) q[c] = mrest[3 * k + c] * s[c];  // OBJ::scale
        for (int r2 = 0; r2 < 3; r2++)                                // transformMesh
          mv[3 * k + r2] = ((R[3 * r2] * q[0] + R[3 * r2 + 1] * q[1]) + R[3 * r2 + 2] * q[2]) +
                           t[r2];                                     // translateMesh
      }
      if ((st = esf_solid_set_mesh(ctx, id, (int)(mrest.size() / 3), mrest.data(), mv.data(),
                                   (int)(mtri.size() / 3), mtri.data()))) {
        esf_ctx_destroy(ctx);
        return st;
      }
    }
  }
  // COUPLED_INTEGRATOR_3D::init  COUPLED_INTEGRATOR_3D.cpp:104-135 with
  // DEFOR_SOLID_3D::init (:269-331): full-grid boxes, mass, boxes from mass
  for (int i = 0; i < nSolids; i++) {
    if ((st = esf_compute_solid_mass(ctx, i)) || (st = esf_solid_sdf(ctx, i))) {
      esf_ctx_destroy(ctx);
      return st;
    }
  }
  if ((st = esf_gather_solid_masses(ctx)) || (st = esf_recover_velocity(ctx))) {
    esf_ctx_destroy(ctx);
    return st;
  }
  // initMaterialParticles of DEFOR_SOLID_3D::init (:327) when the step runs
  // the FLIP round trip
  if (p.flip) {
    if (nSolids > 1) {
      esf_ctx_destroy(ctx);
      return ESF_ERR_INVALID;
    }
    for (int i = 0; i < nSolids; i++)
      if ((st = esf_init_particles(ctx, i))) {
        esf_ctx_destroy(ctx);
        return st;
      }
  }
  // smoke carpet  main.cpp:80-90
  const double carpet = cfg.getFloat("smoke carpet thickness", 0.06);
  const int layers = (int)(carpet / g.dh);
  const size_t N = (size_t)g.xres * g.yres * g.zres;
  std::vector<double> inten(N, 0.0);
  for (int z = 1; z < g.zres - 1; z++)
    for (int y = 1; y < layers; y++)
      for (int x = 1; x < g.xres - 1; x++)
        inten[x + (size_t)y * g.xres + (size_t)z * g.xres * g.yres] = 0.8f;
  if ((st = esf_upload(ctx, ESF_F_INTENSITY, -1, inten.data(), N * 8))) {
    esf_ctx_destroy(ctx);
    return st;
  }
  if (info) {
    info->n_solids = nSolids;
    info->frame_start = cfg.getInt("start frame", 0);
    info->frame_end = cfg.getInt("end frame", 100);
    info->frame_dt = cfg.getFloat("dt", 0.005);
    info->carpet_thickness = carpet;
  }
  *out = ctx;
  return ESF_OK;
}

// The broad phase of stepImplicit (COUPLED_INTEGRATOR_3D.cpp:384-411): a
// solid within 2 cells of a wall (its surface mesh's BVH root box outside
// [origin + 2dh, origin + (res - 3)dh], DEFOR_SOLID_3D.h:85 and
// DEFOR_SOLID_3D.cpp:64-66) or two solids whose meshes collide
// (DEFOR_SOLID_3D.h:84) give CFL(0.6), otherwise CFL(1.8).  When every
// solid has a surface mesh the test runs on the meshes (on the GPU,
// esf_mesh.hip); otherwise each solid's mass bounding box (tight box less
// its border) stands in for its mesh.
static double boxBroadPhase(esf_ctx* ctx) {
  const int res[3] = {ctx->d.X, ctx->d.Y, ctx->d.Z};
  const int border = ctx->p.tight_border;
  std::vector<std::array<int, 6>> bb;
  for (const DevSolid& so : ctx->solids) {
    std::array<int, 6> m;
    for (int k = 0; k < 3; k++) {
      m[k] = so.box.tMin[k] + border;
      m[3 + k] = so.box.tMax[k] - border;
      if (so.box.tMin[k] == 0 || m[k] < 2 || so.box.tMax[k] == res[k] || m[3 + k] > res[k] - 3)
        return 0.6;
    }
    bb.push_back(m);
  }
  for (size_t i = 0; i < bb.size(); i++)
    for (size_t j = i + 1; j < bb.size(); j++) {
      bool ov = true;
      for (int k = 0; k < 3; k++) ov = ov && bb[i][k] <= bb[j][3 + k] && bb[j][k] <= bb[i][3 + k];
      if (ov) return 0.6;
    }
  return 1.8;
}

namespace esf {
esf_status broadPhaseEnqueue(esf_ctx* ctx, bool* meshes) {
  bool m = !ctx->solids.empty();
  for (const DevSolid& so : ctx->solids) m = m && so.meshV > 0;
  *meshes = m;
  return m ? meshBroadPhaseEnqueue(ctx) : ESF_OK;
}

esf_status broadPhaseFinish(esf_ctx* ctx, bool meshes, double* alpha) {
  if (!meshes) {
    *alpha = boxBroadPhase(ctx);
    return ESF_OK;
  }
  std::vector<double> roots;
  std::vector<int> pairs;
  meshBroadPhaseRead(ctx, roots, pairs);
  const double o[3] = {ctx->d.o0, ctx->d.o1, ctx->d.o2};
  const int res[3] = {ctx->d.X, ctx->d.Y, ctx->d.Z};
  bool contact = false;
  for (size_t s = 0; s < ctx->solids.size() && !contact; s++) {
    // _safeFromWallBox.inside(root min) && inside(root max) (aabb.cpp:135-142)
    for (int k = 0; k < 3; k++) {
      const double lo = o[k] + 2.0 * ctx->d.dh;
      const double hi = o[k] + (res[k] - 3) * ctx->d.dh;
      const double mn = roots[6 * s + k], mx = roots[6 * s + 3 + k];
      if (mn < lo || mn > hi || mx < lo || mx > hi) contact = true;
    }
  }
  for (int f : pairs) contact = contact || f != 0;
  *alpha = contact ? 0.6 : 1.8;
  return ESF_OK;
}
}  // namespace esf

extern "C" esf_status esf_broad_phase(esf_ctx* ctx, double* alpha) {
  if (!ctx || !alpha) return ESF_ERR_INVALID;
  bool meshes = false;
  esf_status st = broadPhaseEnqueue(ctx, &meshes);
  if (st) return st;
  if (meshes) ESF_HIP(streamSync(ctx));
  return broadPhaseFinish(ctx, meshes, alpha);
}

extern "C" esf_status esf_frame(esf_ctx* ctx, double frameDt, double alpha, int* substeps) {
  if (!ctx || !(frameDt > 0)) return ESF_ERR_INVALID;
  // APPLICATION::step  rubberBall/main.cpp:180-198
  double sim = 0;
  int n = 0;
  esf_status worst = ESF_OK;
  while (sim < frameDt) {
    esf_set_max_dt(ctx, frameDt - sim);
    esf_solve_stats s;
    double a = alpha;
    if (a < 0) {
      esf_status bs = esf_broad_phase(ctx, &a);
      if (bs) return bs;
    }
    esf_status st = esf_step(ctx, a, &s);
    if (st < 0) return st;
    if (st) worst = st;
    double dt = 0;
    esf_get_dt(ctx, &dt);
    sim += dt;
    n++;
    if (!(dt > 0)) return ESF_ERR_STATE;
  }
  if (substeps) *substeps = n;
  return worst;
}
