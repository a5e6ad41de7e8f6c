// Three fibers interleave through yields, each with floating-point work,
// a deep call chain on its own stack and a changed rounding mode that must
// not leak into the others; then the fibers are reset and run again.
#include <cfenv>
#include <cmath>
#include <cstdio>

#include "esf_fiber.h"

using esf::Fiber;

struct Task {
  Fiber f;
  int id = 0;
  double acc = 0;
  int trace[8] = {0};
};

static int order[64];
static int nOrder = 0;

static double chain(int d) {
  volatile double pad[32];
  for (int i = 0; i < 32; i++) pad[i] = d * 0.5 + i;
  return d ? pad[7] + chain(d - 1) : 0.0;
}

static void body(void* p) {
  Task* t = (Task*)p;
  if (t->id == 1) std::fesetround(FE_UPWARD);
  for (int k = 0; k < 5; k++) {
    t->acc += std::sqrt(t->id + k + 1.0) + chain(200);
    order[nOrder++] = 10 * t->id + k;
    esf::esf_fiber_yield(&t->f);
  }
}

static int runAll(Task* t, int n) {
  int rounds = 0;
  for (bool live = true; live; rounds++) {
    live = false;
    for (int i = 0; i < n; i++)
      if (!t[i].f.finished) {
        esf::esf_fiber_resume(&t[i].f);
        live = true;
      }
  }
  return rounds;
}

int main() {
  Task t[3];
  for (int i = 0; i < 3; i++) {
    t[i].id = i;
    if (!esf::esf_fiber_create(&t[i].f, 256 * 1024, body, &t[i])) return 2;
  }
  const int rounds = runAll(t, 3);
  std::printf("rounds %d\n", rounds);
  std::printf("order");
  for (int i = 0; i < nOrder; i++) std::printf(" %d", order[i]);
  std::printf("\n");
  // the thread's own rounding mode is untouched by fiber 1's
  std::printf("round %d\n", std::fegetround() == FE_TONEAREST ? 1 : 0);
  for (int i = 0; i < 3; i++) std::printf("acc %d %.17g\n", i, t[i].acc);
  for (int i = 0; i < 3; i++) {
    t[i].acc = 0;
    esf::esf_fiber_reset(&t[i].f, body, &t[i]);
  }
  nOrder = 0;
  runAll(t, 3);
  for (int i = 0; i < 3; i++) std::printf("again %d %.17g\n", i, t[i].acc);
  for (int i = 0; i < 3; i++) esf::esf_fiber_destroy(&t[i].f);
  return 0;
}
