// esf_run.cpp -- headless counterpart of projects/rubberBall/main.cpp:
// loads a reference cfg, builds the scene on the GPU and advances frames
// with APPLICATION::step's substep loop (main.cpp:174-205), printing the
// per-frame solver report instead of opening a GLUT window.
//
//   esf_run <scene.cfg> [frames] [--mixed] [--alpha A]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/esf.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("usage: esf_run <scene.cfg> [frames] [--mixed] [--alpha A]\n");
    return 2;
  }
  int frames = -1;
  int precision = ESF_PLAIN_FP64;
  double alpha = -1;
  for (int i = 2; i < argc; i++) {
    if (!std::strcmp(argv[i], "--mixed"))
      precision = ESF_MIXED;
    else if (!std::strcmp(argv[i], "--alpha") && i + 1 < argc)
      alpha = std::atof(argv[++i]);
    else
      frames = std::atoi(argv[i]);
  }
  esf_ctx* ctx = nullptr;
  esf_scene_info info{};
  esf_status st = esf_scene_load_cfg(argv[1], precision, &ctx, &info);
  if (st) {
    std::fprintf(stderr, "scene setup failed (%d): %s\n", st, ctx ? esf_last_error(ctx) : "");
    return 1;
  }
  if (frames < 0) frames = info.frame_end - info.frame_start;
  int nsolid = 0;
  esf_num_solid_nodes(ctx, &nsolid);
  std::printf("scene %s: %d solid(s), %d solid nodes, frame dt %g\n", argv[1], info.n_solids,
              nsolid, info.frame_dt);
  for (int f = 0; f < frames; f++) {
    auto t0 = std::chrono::steady_clock::now();
    int sub = 0;
    st = esf_frame(ctx, info.frame_dt, alpha, &sub);
    auto t1 = std::chrono::steady_clock::now();
    if (st < 0) {
      std::fprintf(stderr, "frame %d failed (%d): %s\n", f, st, esf_last_error(ctx));
      esf_ctx_destroy(ctx);
      return 1;
    }
    std::printf("frame %4d: %d substep(s), %.3f ms%s\n", info.frame_start + f, sub,
                std::chrono::duration<double, std::milli>(t1 - t0).count(),
                st == ESF_NOT_CONVERGED ? " (a solve hit cr max iterations)" : "");
    std::fflush(stdout);
  }
  esf_ctx_destroy(ctx);
  return 0;
}
