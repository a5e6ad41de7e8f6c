// esf_frames.cpp -- the reference's on-disk frame format, on the host: gz
// field files (FIELD_3D::writeGz / readGz, FIELD_3D.h:286-365;
// VEC3F_FIELD_3D::writeGz / readGz, VEC3F_FIELD_3D.h:199-267) and the frame
// files the integrator writes and resumes from
// (COUPLED_INTEGRATOR_3D::writeFrame / readFrame,
// COUPLED_INTEGRATOR_3D.cpp:976-1011; DEFOR_SOLID_3D::writeFrame /
// readFrame, DEFOR_SOLID_3D.cpp:1299-1371), and the OBJ surface meshes
// (OBJ::Load / Save, src/geometry/OBJ.cpp:97-375).  Device fields cross PCIe
// through esf_download / esf_upload; nothing here runs on the GPU.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/esf.h"
#include "../csrc/esf_internal.h"

namespace esf {

// a gz field file: header int32 X, Y, Z; double dh, origin[3]; then
// components * X * Y * Z doubles, x fastest, vector components interleaved
static esf_status writeGzField(const std::string& path, const esf_grid& g, int components,
                               const double* data) {
  // level 1, as the reference opens it ("wb1")
  gzFile f = gzopen(path.c_str(), "wb1");
  if (!f) return ESF_ERR_IO;
  const int dims[3] = {g.xres, g.yres, g.zres};
  const double hdr[4] = {g.dh, g.origin[0], g.origin[1], g.origin[2]};
  bool ok = gzwrite(f, dims, sizeof(dims)) == (int)sizeof(dims) &&
            gzwrite(f, hdr, sizeof(hdr)) == (int)sizeof(hdr);
  // the payload in chunks (gzwrite takes an unsigned length)
  const size_t total = (size_t)components * g.xres * g.yres * g.zres * sizeof(double);
  const char* p = (const char*)data;
  for (size_t off = 0; ok && off < total;) {
    const unsigned len = (unsigned)std::min<size_t>(total - off, (size_t)1 << 30);
    ok = gzwrite(f, p + off, len) == (int)len;
    off += len;
  }
  if (gzclose(f) != Z_OK) ok = false;
  return ok ? ESF_OK : ESF_ERR_IO;
}

static esf_status readGzHeader(gzFile f, esf_grid* g) {
  int dims[3];
  double hdr[4];
  if (gzread(f, dims, sizeof(dims)) != (int)sizeof(dims) ||
      gzread(f, hdr, sizeof(hdr)) != (int)sizeof(hdr))
    return ESF_ERR_IO;
  if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return ESF_ERR_IO;
  g->xres = dims[0];
  g->yres = dims[1];
  g->zres = dims[2];
  g->dh = hdr[0];
  g->origin[0] = hdr[1];
  g->origin[1] = hdr[2];
  g->origin[2] = hdr[3];
  return ESF_OK;
}

static esf_status readGzField(const std::string& path, esf_grid* g, int components,
                              double* data, size_t capacity) {
  gzFile f = gzopen(path.c_str(), "rb");
  if (!f) return ESF_ERR_IO;
  esf_status st = readGzHeader(f, g);
  if (!st && data) {
    const size_t n = (size_t)components * g->xres * g->yres * g->zres;
    if (n > capacity) {
      st = ESF_ERR_INVALID;
    } else {
      char* p = (char*)data;
      const size_t total = n * sizeof(double);
      for (size_t off = 0; !st && off < total;) {
        const unsigned len = (unsigned)std::min<size_t>(total - off, (size_t)1 << 30);
        if (gzread(f, p + off, len) != (int)len) st = ESF_ERR_IO;
        off += len;
      }
    }
  }
  gzclose(f);
  return st;
}

// IO::itoPaddedString (include/util/IO.h:27-37): at least four digits
static std::string padded(int frame) {
  std::string s = std::to_string(frame);
  if (frame < 10) s = "0" + s;
  if (frame < 100) s = "0" + s;
  if (frame < 1000) s = "0" + s;
  return s;
}

static esf_grid gridOf(const esf_ctx* ctx) {
  esf_grid g;
  g.xres = ctx->d.X;
  g.yres = ctx->d.Y;
  g.zres = ctx->d.Z;
  g.dh = ctx->d.dh;
  g.origin[0] = ctx->d.o0;
  g.origin[1] = ctx->d.o1;
  g.origin[2] = ctx->d.o2;
  return g;
}

// OBJ::Load (src/geometry/OBJ.cpp:97-295) for what the solver uses: "v"
// lines give the vertices, each "f" line (indices v, v/t, v/t/n or v//n,
// 1-based, negative = relative to the end) a triangle from its first three
// vertices (:256-277).  Other lines are skipped.
static bool readObj(const std::string& path, std::vector<double>& v, std::vector<int>& tri) {
  std::ifstream in(path);
  if (in.fail()) return false;
  v.clear();
  tri.clear();
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream is(line);
    std::string type;
    is >> type;
    if (is.fail()) continue;
    if (type == "v") {
      double x = 0, y = 0, z = 0;
      is >> x >> y >> z;
      if (is.fail()) return false;
      v.push_back(x);
      v.push_back(y);
      v.push_back(z);
    } else if (type == "f") {
      std::vector<int> idx;
      std::string tok;
      while (is >> tok) {
        const long nvert = (long)(v.size() / 3);
        const size_t slash = tok.find('/');
        const std::string first = tok.substr(0, slash);
        char* end = nullptr;
        long k = std::strtol(first.c_str(), &end, 10);
        if (end == first.c_str() || *end != '\0') return false;
        if (k < 0) k = nvert + 1 + k;
        if (k < 1 || k > nvert) return false;
        idx.push_back((int)(k - 1));
      }
      if (idx.size() < 3) return false;
      for (int y = 0; y < 3; y++) tri.push_back(idx[y]);
    }
  }
  return true;
}

static bool writeObj(const std::string& path, int nv, const double* v, int nt, const int* tri) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  // full precision (OBJ::Save, OBJ.cpp:336-375, streams 6 significant
  // digits; the extra digits keep a resumed mesh identical)
  for (int i = 0; i < nv; i++)
    std::fprintf(f, "v %.17g %.17g %.17g\n", v[3 * i], v[3 * i + 1], v[3 * i + 2]);
  for (int t = 0; t < nt; t++)
    std::fprintf(f, "f %d %d %d\n", tri[3 * t] + 1, tri[3 * t + 1] + 1, tri[3 * t + 2] + 1);
  const bool ok = std::ferror(f) == 0;
  return std::fclose(f) == 0 && ok;
}

static bool sameGrid(const esf_grid& a, const esf_grid& b) {
  return a.xres == b.xres && a.yres == b.yres && a.zres == b.zres && a.dh == b.dh &&
         a.origin[0] == b.origin[0] && a.origin[1] == b.origin[1] && a.origin[2] == b.origin[2];
}

}  // namespace esf

using namespace esf;

extern "C" esf_status esf_field_write_gz(const char* path, const esf_grid* grid, int components,
                                         const double* data) {
  if (!path || !grid || !data || (components != 1 && components != 3) || grid->xres <= 0 ||
      grid->yres <= 0 || grid->zres <= 0)
    return ESF_ERR_INVALID;
  return writeGzField(path, *grid, components, data);
}

extern "C" esf_status esf_field_read_gz(const char* path, esf_grid* grid, int components,
                                        double* data, size_t capacity) {
  if (!path || !grid || (components != 1 && components != 3)) return ESF_ERR_INVALID;
  return readGzField(path, grid, components, data, capacity);
}

extern "C" esf_status esf_obj_read(const char* path, int* nv, int* nt, double* vertices,
                                   int cap_v, int* tris, int cap_t) {
  if (!path || !nv || !nt) return ESF_ERR_INVALID;
  std::vector<double> v;
  std::vector<int> t;
  if (!readObj(path, v, t)) return ESF_ERR_IO;
  *nv = (int)(v.size() / 3);
  *nt = (int)(t.size() / 3);
  if (vertices) {
    if (cap_v < *nv) return ESF_ERR_INVALID;
    std::copy(v.begin(), v.end(), vertices);
  }
  if (tris) {
    if (cap_t < *nt) return ESF_ERR_INVALID;
    std::copy(t.begin(), t.end(), tris);
  }
  return ESF_OK;
}

extern "C" esf_status esf_obj_write(const char* path, int nv, const double* vertices, int nt,
                                    const int* tris) {
  if (!path || nv < 0 || nt < 0 || (nv && !vertices) || (nt && !tris)) return ESF_ERR_INVALID;
  for (long k = 0; k < 3L * nt; k++)
    if (tris[k] < 0 || tris[k] >= nv) return ESF_ERR_INVALID;
  return writeObj(path, nv, vertices, nt, tris) ? ESF_OK : ESF_ERR_IO;
}

extern "C" esf_status esf_write_frame(esf_ctx* ctx, const char* dataPath, int frame, int debug) {
  if (!ctx || !dataPath || frame < 0) return ESF_ERR_INVALID;
  const esf_grid g = gridOf(ctx);
  const size_t N = ctx->d.nodes;
  const std::string base = std::string(dataPath) + "frame." + padded(frame);
  const bool full = debug || frame % 10 == 0;
  std::vector<double> buf(3 * N);
  esf_status st = ESF_OK;
  // solids first (COUPLED_INTEGRATOR_3D.cpp:999-1004): displacement and
  // velocity (DEFOR_SOLID_3D.cpp:1362-1366)
  for (size_t i = 0; full && !st && i < ctx->solids.size(); i++) {
    const std::string sb = base + ".solid." + std::to_string(i);
    st = esf_download(ctx, ESF_S_U, (int)i, buf.data(), 24 * N);
    if (!st) st = writeGzField(sb + ".displacement", g, 3, buf.data());
    if (!st) st = esf_download(ctx, ESF_S_VELOCITY, (int)i, buf.data(), 24 * N);
    if (!st) st = writeGzField(sb + ".velocity", g, 3, buf.data());
    if (!st) st = esf_download(ctx, ESF_S_PHI, (int)i, buf.data(), 8 * N);
    if (!st) st = writeGzField(sb + ".phi", g, 1, buf.data());
  }
  // the surface mesh every frame (DEFOR_SOLID_3D.cpp:1368)
  for (size_t i = 0; !st && i < ctx->solids.size(); i++) {
    const DevSolid& so = ctx->solids[i];
    if (so.meshV == 0) continue;
    std::vector<double> mv(3L * so.meshV);
    st = esf_solid_get_mesh(ctx, (int)i, mv.data(), nullptr, nullptr);
    if (!st && !writeObj(base + ".solid." + std::to_string(i) + ".obj", so.meshV, mv.data(),
                         so.meshT, so.meshTriHost.data()))
      st = ESF_ERR_IO;
  }
  if (full && !st) {
    st = esf_download(ctx, ESF_F_VELOCITY, -1, buf.data(), 24 * N);
    if (!st) st = writeGzField(base + ".velocity", g, 3, buf.data());
  }
  if (!st) {
    st = esf_download(ctx, ESF_F_INTENSITY, -1, buf.data(), 8 * N);
    if (!st) st = writeGzField(base + ".intensity", g, 1, buf.data());
  }
  if (st == ESF_ERR_IO) ctx->err = "cannot write frame files under " + base;
  return st;
}

extern "C" esf_status esf_read_frame(esf_ctx* ctx, const char* dataPath, int frame, int debug) {
  if (!ctx || !dataPath || frame < 0) return ESF_ERR_INVALID;
  const esf_grid g = gridOf(ctx);
  const size_t N = ctx->d.nodes;
  const std::string base = std::string(dataPath) + "frame." + padded(frame);
  const bool full = debug || frame % 10 == 0;
  std::vector<double> buf(3 * N), u(3 * N);
  esf_status st = ESF_OK;
  auto readField = [&](const std::string& path, int components, double* dst) {
    esf_grid fg;
    esf_status s = readGzField(path, &fg, components, dst, 3 * N);
    if (s == ESF_ERR_IO) {
      ctx->err = "cannot read " + path;
    } else if (!s && !sameGrid(fg, g)) {
      ctx->err = path + ": grid differs from the context's";
      s = ESF_ERR_INVALID;
    }
    return s;
  };
  // solids (DEFOR_SOLID_3D.cpp:1299-1314): displacement and velocity, the
  // velocity workspace from the velocity, X = x - u
  for (size_t i = 0; !st && i < ctx->solids.size(); i++) {
    const std::string sb = base + ".solid." + std::to_string(i);
    if (full) {
      st = readField(sb + ".displacement", 3, u.data());
      if (!st) st = esf_upload(ctx, ESF_S_U, (int)i, u.data(), 24 * N);
      if (!st) st = readField(sb + ".phi", 1, buf.data());
      if (!st) st = esf_upload(ctx, ESF_S_PHI, (int)i, buf.data(), 8 * N);
      if (!st) st = readField(sb + ".velocity", 3, buf.data());
      if (!st) st = esf_upload(ctx, ESF_S_VELOCITY, (int)i, buf.data(), 24 * N);
    } else {
      st = esf_download(ctx, ESF_S_U, (int)i, u.data(), 24 * N);
      if (!st) st = esf_download(ctx, ESF_S_VELOCITY, (int)i, buf.data(), 24 * N);
    }
    if (!st) st = esf_upload(ctx, ESF_S_VEL_WORKSPACE, (int)i, buf.data(), 24 * N);
    if (st) break;
    size_t n = 0;
    for (int z = 0; z < g.zres; z++)
      for (int y = 0; y < g.yres; y++)
        for (int x = 0; x < g.xres; x++, n++) {
          const double pos[3] = {x * g.dh + g.origin[0], y * g.dh + g.origin[1],
                                 z * g.dh + g.origin[2]};
          for (int c = 0; c < 3; c++) buf[3 * n + c] = pos[c] - u[3 * n + c];
        }
    st = esf_upload(ctx, ESF_S_X, (int)i, buf.data(), 24 * N);
    // the surface mesh's vertices (:1315-1317); the topology must match
    DevSolid& so = ctx->solids[i];
    if (!st && so.meshV > 0) {
      const std::string path = sb + ".obj";
      std::vector<double> mv;
      std::vector<int> mt;
      if (!readObj(path, mv, mt)) {
        ctx->err = "cannot read " + path;
        st = ESF_ERR_IO;
      } else if ((int)(mv.size() / 3) != so.meshV || mt != so.meshTriHost) {
        ctx->err = path + ": mesh differs from the solid's";
        st = ESF_ERR_INVALID;
      } else {
        ESF_HIP(hipMemcpyAsync(so.meshVert, mv.data(), mv.size() * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
        ESF_HIP(hipStreamSynchronize(ctx->stream));
        // the boxes from the mesh's bounding box, its corners truncated to
        // node indices (:1318-1355)
        double lo[3] = {1e9, 1e9, 1e9}, hi[3] = {-1e9, -1e9, -1e9};
        for (size_t v = 0; v < mv.size(); v += 3)
          for (int c = 0; c < 3; c++) {
            lo[c] = std::min(lo[c], mv[v + c]);
            hi[c] = std::max(hi[c], mv[v + c]);
          }
        const double o[3] = {g.origin[0], g.origin[1], g.origin[2]};
        const int res[3] = {g.xres, g.yres, g.zres};
        const int tb = ctx->p.tight_border, ex = ctx->p.extrapolation_cells;
        int boxes[12];
        for (int c = 0; c < 3; c++) {
          const int mn = (int)((lo[c] - o[c]) / g.dh), mx = (int)((hi[c] - o[c]) / g.dh);
          boxes[c] = std::max(mn - tb, 0);
          boxes[3 + c] = std::min(mx + tb, res[c]);
          boxes[6 + c] = std::max(mn - ex, 0);
          boxes[9 + c] = std::min(mx + ex, res[c]);
        }
        st = esf_solid_set_boxes(ctx, (int)i, boxes);
      }
    }
    // computeSolidMass(true) on the read state (:1357)
    if (!st) st = esf_compute_solid_mass_sdf(ctx, (int)i);
  }
  // grid (COUPLED_INTEGRATOR_3D.cpp:983-990)
  if (!st && full) {
    st = readField(base + ".velocity", 3, buf.data());
    if (!st) st = esf_upload(ctx, ESF_F_VELOCITY, -1, buf.data(), 24 * N);
  }
  if (!st) {
    st = readField(base + ".intensity", 1, buf.data());
    if (!st) st = esf_upload(ctx, ESF_F_INTENSITY, -1, buf.data(), 8 * N);
  }
  // :991-995
  if (!st) st = esf_gather_solid_masses(ctx);
  if (!st) st = esf_recover_velocity(ctx);
  return st;
}
