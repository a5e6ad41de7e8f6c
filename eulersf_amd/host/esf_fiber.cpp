// esf_fiber.cpp -- the fiber switch (x86-64 System V) and fiber stacks.
// Host-only: built with g++, not through the offload compiler.
#include "esf_fiber.h"

#include <sys/mman.h>
#include <unistd.h>

#include <cstdint>
#include <cstring>

// esf_fiber_switch(save, load): push the callee-saved registers and the
// MXCSR / x87 control words on the current stack, store the stack pointer
// in *save, switch to the stack `load` and pop the same set from it.  A new
// fiber's stack is laid out so that the pops land in esf_fiber_entry with
// the Fiber in %rbx.
extern "C" void esf_fiber_switch(void** save, void* load);
extern "C" void esf_fiber_entry();

asm(R"(
  .text
  .p2align 4
  .globl esf_fiber_switch
  .hidden esf_fiber_switch
  .type esf_fiber_switch, @function
esf_fiber_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  subq $8, %rsp
  stmxcsr (%rsp)
  fnstcw 4(%rsp)
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  ldmxcsr (%rsp)
  fldcw 4(%rsp)
  addq $8, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
  .size esf_fiber_switch, .-esf_fiber_switch

  .p2align 4
  .globl esf_fiber_entry
  .hidden esf_fiber_entry
  .type esf_fiber_entry, @function
esf_fiber_entry:
  movq %rbx, %rdi
  call esf_fiber_main
  ud2
  .size esf_fiber_entry, .-esf_fiber_entry
)");

namespace esf {

extern "C" __attribute__((used)) void esf_fiber_main(Fiber* f) {
  f->fn(f->arg);
  f->finished = true;
  // never resumed again: the resumer sees `finished`
  esf_fiber_switch(&f->sp, f->back);
  __builtin_trap();
}

static void layout(Fiber* f) {
  uintptr_t top = (uintptr_t)f->stack + f->mapped;
  top &= ~(uintptr_t)15;
  uint64_t* s = (uint64_t*)top;
  // from the top down: return address, rbp, rbx (the Fiber), r12 - r15,
  // then the control words; after the final `ret` the stack pointer is
  // the 16-byte aligned `top`, as a call site needs it
  uint32_t mx;
  uint16_t cw;
  asm volatile("stmxcsr %0" : "=m"(mx));
  asm volatile("fnstcw %0" : "=m"(cw));
  s[-1] = (uint64_t)(uintptr_t)&esf_fiber_entry;
  s[-2] = 0;                        // rbp
  s[-3] = (uint64_t)(uintptr_t)f;   // rbx
  s[-4] = 0;                        // r12
  s[-5] = 0;                        // r13
  s[-6] = 0;                        // r14
  s[-7] = 0;                        // r15
  uint64_t ctl = 0;
  std::memcpy(&ctl, &mx, 4);
  std::memcpy((char*)&ctl + 4, &cw, 2);
  s[-8] = ctl;
  f->sp = &s[-8];
  f->finished = false;
}

bool esf_fiber_create(Fiber* f, size_t stackBytes, void (*fn)(void*), void* arg) {
  const size_t page = (size_t)sysconf(_SC_PAGESIZE);
  const size_t bytes = ((stackBytes + page - 1) / page) * page + page;
  void* m = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
  if (m == MAP_FAILED) return false;
  // the lowest page faults on overflow
  mprotect(m, page, PROT_NONE);
  f->stack = m;
  f->mapped = bytes;
  f->fn = fn;
  f->arg = arg;
  layout(f);
  return true;
}

void esf_fiber_reset(Fiber* f, void (*fn)(void*), void* arg) {
  f->fn = fn;
  f->arg = arg;
  layout(f);
}

void esf_fiber_destroy(Fiber* f) {
  if (f->stack) munmap(f->stack, f->mapped);
  f->stack = nullptr;
  f->mapped = 0;
}

void esf_fiber_resume(Fiber* f) { esf_fiber_switch(&f->back, f->sp); }

void esf_fiber_yield(Fiber* f) { esf_fiber_switch(&f->sp, f->back); }

}  // namespace esf
