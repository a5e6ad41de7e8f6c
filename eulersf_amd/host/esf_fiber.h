// esf_fiber.h -- cooperative user-level threads on one OS thread.
//
// A fiber runs a function on its own stack and hands control back to the
// thread that resumed it whenever it calls esf_fiber_yield.  The batched
// step (esf_step_batch) runs each scene's step in a fiber, so every
// scene's step keeps its own control flow while one OS thread collects the
// kernel launches the scenes reach together into launches over all of
// them.  x86-64 System V only: the switch saves and restores the
// callee-saved registers, the SSE control word and the x87 control word.
#pragma once
#include <cstddef>

namespace esf {

struct Fiber {
  void* sp = nullptr;       // the fiber's stack pointer while it is switched out
  void* back = nullptr;     // the resumer's stack pointer while the fiber runs
  void* stack = nullptr;    // mapping: guard page + stack
  size_t mapped = 0;
  void (*fn)(void*) = nullptr;
  void* arg = nullptr;
  bool finished = false;
};

// a fiber with a stack of `stackBytes` (rounded up to pages, plus a guard
// page below it) that will run fn(arg) when first resumed; false when the
// stack cannot be mapped
bool esf_fiber_create(Fiber* f, size_t stackBytes, void (*fn)(void*), void* arg);
// a finished fiber (or one never resumed) made ready to run fn(arg) again
// on the same stack
void esf_fiber_reset(Fiber* f, void (*fn)(void*), void* arg);
void esf_fiber_destroy(Fiber* f);
// run the fiber until it yields or its function returns
void esf_fiber_resume(Fiber* f);
// from inside the fiber: back to the thread that resumed it
void esf_fiber_yield(Fiber* f);

}  // namespace esf
