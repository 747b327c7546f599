// Test infrastructure: CPU execution of the library's kernels.  All lanes of
// a workgroup run as fibers on one OS thread (an x86-64 stack switch written
// below); __syncthreads switches back to the scheduler, which resumes the
// next lane, so every lane reaches each barrier before any passes it.  A
// barrier that some lanes skip aborts with a message.
#include "hip/hip_runtime.h"
#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define EMU_ASAN 1
#endif
#endif
#if defined(__SANITIZE_ADDRESS__)
#define EMU_ASAN 1
#endif
#ifdef EMU_ASAN
#include <sanitizer/common_interface_defs.h>
#else
#define __sanitizer_start_switch_fiber(a, b, c) ((void)0)
#define __sanitizer_finish_switch_fiber(a, b, c) ((void)0)
#endif
thread_local dim3 threadIdx, blockIdx;
dim3 gridDim, blockDim;
static double g_xbuf[1024];

extern "C" void emu_switch(void **save_sp, void *new_sp);
asm(R"(
.text
.globl emu_switch
.type emu_switch,@function
emu_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
)");

struct Fiber { void *sp; std::vector<char> stack; bool done; void *fake; };
static std::vector<Fiber> g_fib;
static void *g_sched_sp;
static int g_cur = -1;
static std::function<void()> *g_body = nullptr;
static const void *g_main_bottom; static size_t g_main_size;

static void to_sched() {
  Fiber &F = g_fib[g_cur];
  __sanitizer_start_switch_fiber(&F.fake, g_main_bottom, g_main_size);
  emu_switch(&F.sp, g_sched_sp);
  __sanitizer_finish_switch_fiber(F.fake, nullptr, nullptr);
}
void __syncthreads() { to_sched(); }
double emu_shfl_xor(double v, int off) {
  g_xbuf[threadIdx.x] = v;
  __syncthreads();
  double r = g_xbuf[threadIdx.x ^ off];
  __syncthreads();
  return r;
}
void emu_record_or_run(hipStream_t st, std::function<void()> f) {
  if (st && st->capturing) st->rec->push_back(f);
  else f();
}
extern "C" void fiber_entry() {
  __sanitizer_finish_switch_fiber(nullptr, &g_main_bottom, &g_main_size);
  (*g_body)();
  g_fib[g_cur].done = true;
  Fiber &F = g_fib[g_cur];
  __sanitizer_start_switch_fiber(nullptr, g_main_bottom, g_main_size);
  emu_switch(&F.sp, g_sched_sp);
  abort();
}
void emu_run_grid(dim3 grid, dim3 block, std::function<void()> body) {
  gridDim = grid;
  blockDim = block;
  int nt = block.x * block.y * block.z;
  g_fib.resize(nt);
  const size_t SZ = 1 << 20;
  for (auto &f : g_fib) if (f.stack.size() < SZ) f.stack.resize(SZ);
  g_body = &body;
  for (unsigned bz = 0; bz < grid.z; bz++)
    for (unsigned by = 0; by < grid.y; by++)
      for (unsigned bx = 0; bx < grid.x; bx++) {
        blockIdx = dim3(bx, by, bz);
        for (int t = 0; t < nt; t++) {
          Fiber &F = g_fib[t];
          uintptr_t top = ((uintptr_t)(F.stack.data() + SZ)) & ~(uintptr_t)15;
          void **sp = (void **)top;
          *--sp = nullptr;                 // alignment slot: entry sees rsp % 16 == 8
          *--sp = (void *)&fiber_entry;    // return address for emu_switch's ret
          for (int r = 0; r < 6; r++) *--sp = nullptr;   // rbp rbx r12-r15
          F.sp = sp;
          F.done = false;
        }
        int alive = nt;
        while (alive > 0) {
          int ndone = 0, nbar = 0;
          for (int t = 0; t < nt; t++) {
            if (g_fib[t].done) { ndone++; continue; }
            g_cur = t;
            threadIdx = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
            void *fake = nullptr;
            __sanitizer_start_switch_fiber(&fake, g_fib[t].stack.data(), SZ);
            emu_switch(&g_sched_sp, g_fib[t].sp);
            __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
            if (g_fib[t].done) ndone++; else nbar++;
          }
          if (ndone > 0 && nbar > 0) { fprintf(stderr, "EMU: divergent barrier (%d done, %d waiting)\n", ndone, nbar); abort(); }
          alive = nt - ndone;
        }
      }
  g_body = nullptr;
}
