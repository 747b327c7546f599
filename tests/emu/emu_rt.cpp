// Test infrastructure: CPU execution of the library's kernels.  All lanes of
// a workgroup run as fibers on one OS thread (an x86-64 stack switch written
// below); __syncthreads switches back to the scheduler, which resumes the
// next lane, so every lane reaches each barrier before any passes it.
// Checks: a barrier that some lanes skip, or lanes waiting at different
// barriers, abort with a message; EMU_ORDER_SEED runs the lanes of each
// barrier interval in a seeded random order (a result that changes with it
// reads shared data another lane writes in the same interval); the library's
// WX_EMU_CHECK index ranges abort with the failing lane.
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

namespace wx {
void wx_emu_check_failed(const char *what, const char *file, int line) {
  fprintf(stderr, "EMU: index check failed: %s at %s:%d (block %u %u %u, thread %u %u %u)\n", what, file, line,
          blockIdx.x, blockIdx.y, blockIdx.z, threadIdx.x, threadIdx.y, threadIdx.z);
  abort();
}
}  // namespace wx

static void to_sched() {
  Fiber &F = g_fib[g_cur];
  __sanitizer_start_switch_fiber(&F.fake, g_main_bottom, g_main_size);
  emu_switch(&F.sp, g_sched_sp);
  __sanitizer_finish_switch_fiber(F.fake, nullptr, nullptr);
}
// source position (file, line) of each lane's current barrier: every lane
// of a round must wait at the same __syncthreads of the source (a barrier
// that lanes reach from different places of the kernel is divergence the GPU
// does not tolerate; copies of one call that the compiler makes are the
// same barrier)
struct EmuSite { const char *file; int line; };
static std::vector<EmuSite> g_site;
void emu_syncthreads(const char *file, int line) {
  g_site[g_cur] = EmuSite{file, line};
  to_sched();
}
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
  g_site.assign(nt, EmuSite{nullptr, 0});
  // EMU_ORDER_SEED: run the lanes of each barrier interval in a seeded
  // random order instead of lane order (a kernel whose result depends on
  // that order reads shared data another lane writes in the same interval)
  static const char *seed_env = getenv("EMU_ORDER_SEED");
  const unsigned long long seed = seed_env ? strtoull(seed_env, nullptr, 10) : 0;
  std::vector<int> order(nt);
  unsigned long long rng = seed * 0x9E3779B97F4A7C15ull + 1;
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
          for (int t = 0; t < nt; t++) order[t] = t;
          if (seed_env) {
            for (int t = nt - 1; t > 0; t--) {
              rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
              std::swap(order[t], order[rng % (unsigned long long)(t + 1)]);
            }
          }
          for (int n = 0; n < nt; n++) {
            const int t = order[n];
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
          if (nbar > 0) {
            const EmuSite *site = nullptr;
            for (int t = 0; t < nt; t++) {
              if (g_fib[t].done) continue;
              if (!site) site = &g_site[t];
              else if (g_site[t].line != site->line || strcmp(g_site[t].file, site->file) != 0) {
                fprintf(stderr, "EMU: lanes wait at different barriers (%s:%d, %s:%d)\n", site->file, site->line,
                        g_site[t].file, g_site[t].line);
                abort();
              }
            }
          }
          alive = nt - ndone;
        }
      }
  g_body = nullptr;
}
