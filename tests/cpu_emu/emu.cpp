// Fiber scheduler and host API of the CPU emulation (see hip/hip_runtime.h).
#include "hip/hip_runtime.h"

#include <sys/mman.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#ifdef __SANITIZE_ADDRESS__
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#else
#define EMU_ASAN 0
#endif

extern "C" void emu_swap(void** save_sp, void* load_sp);
asm(".text\n.globl emu_swap\n.type emu_swap,@function\nemu_swap:\n"
    "pushq %rbp\npushq %rbx\npushq %r12\npushq %r13\npushq %r14\npushq %r15\n"
    "movq %rsp, (%rdi)\nmovq %rsi, %rsp\n"
    "popq %r15\npopq %r14\npopq %r13\npopq %r12\npopq %rbx\npopq %rbp\nret\n"
    ".size emu_swap, .-emu_swap\n");

namespace emu {

thread_local dim3 t_block, t_grid, t_bdim;

enum { RUN = 0, BAR = 1, WAVE = 2, DONE = 3 };
static const size_t kStack = 512 * 1024;

struct Fiber {
  void* sp = nullptr;
  char* stack = nullptr;
  dim3 tid;
  int state = RUN;
  bool slept = false;
  int kind = 0, arg = 0, width = 64;
  uint64_t val = 0, res = 0;
};

struct Block {
  std::vector<Fiber> f;
  void* sched_sp = nullptr;
  const std::function<void()>* fn = nullptr;
  std::vector<char> dyn;
  void* fake = nullptr;
  const void* sched_bottom = nullptr;
  size_t sched_size = 0;
};
thread_local Block* t_blk = nullptr;
thread_local Fiber* t_cur = nullptr;
std::atomic<long> g_inactive_src{0};

Fiber* cur() { return t_cur; }
const dim3& tid_of(Fiber* f) { return f->tid; }
void* dyn_lds() { return t_blk->dyn.data(); }

static void to_sched(Fiber* f) {
#if EMU_ASAN
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, t_blk->sched_bottom, t_blk->sched_size);
  emu_swap(&f->sp, t_blk->sched_sp);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#else
  emu_swap(&f->sp, t_blk->sched_sp);
#endif
}

[[noreturn]] static void finish_fiber() {
  Fiber* f = t_cur;
  f->state = DONE;
#if EMU_ASAN
  __sanitizer_start_switch_fiber(nullptr, t_blk->sched_bottom, t_blk->sched_size);
#endif
  emu_swap(&f->sp, t_blk->sched_sp);
  std::abort();
}

static void fiber_main() {
#if EMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &t_blk->sched_bottom, &t_blk->sched_size);
#endif
  (*t_blk->fn)();
  finish_fiber();
}

// The kernel's frames on this fiber's stack are abandoned; the stack is
// unpoisoned again when the next workgroup takes it.
void end_program() { finish_fiber(); }

void barrier() {
  Fiber* f = t_cur;
  f->state = BAR;
  to_sched(f);
}
void yield() {
  Fiber* f = t_cur;
  f->slept = true;
  to_sched(f);
}
uint64_t wave_op(int kind, uint64_t v, int arg, int width) {
  Fiber* f = t_cur;
  f->kind = kind;
  f->val = v;
  f->arg = arg;
  f->width = width;
  f->state = WAVE;
  to_sched(f);
  return f->res;
}

// Resolves the wave operations of segment [s0, s0 + w) when no lane of it
// can still arrive: participants are the lanes waiting in WAVE.
static bool resolve_segment(Block& B, uint32_t s0, uint32_t w, uint32_t nthreads) {
  uint32_t waiting = 0, running = 0;
  for (uint32_t l = s0; l < s0 + w && l < nthreads; ++l) {
    if (B.f[l].state == WAVE) ++waiting;
    if (B.f[l].state == RUN) ++running;
  }
  if (!waiting || running) return false;
  const Fiber* first = nullptr;
  for (uint32_t l = s0; l < s0 + w && l < nthreads; ++l)
    if (B.f[l].state == WAVE) {
      first = &B.f[l];
      break;
    }
  for (uint32_t l = s0; l < s0 + w && l < nthreads; ++l) {
    const Fiber& x = B.f[l];
    if (x.state != WAVE) continue;
    if (x.kind != first->kind || x.width != first->width) {
      std::fprintf(stderr, "emu: mismatched wave ops in block %u lanes %u.. (%d/%d vs %d/%d)\n", t_block.x, s0,
                   x.kind, x.width, first->kind, first->width);
      std::abort();
    }
  }
  auto active = [&](uint32_t l) { return l < nthreads && B.f[l].state == WAVE; };
  uint64_t mask = 0;
  for (uint32_t l = s0; l < s0 + w && l < nthreads; ++l)
    if (active(l) && B.f[l].val) mask |= 1ull << ((l - s0) & 63);
  for (uint32_t l = s0; l < s0 + w && l < nthreads; ++l) {
    Fiber& x = B.f[l];
    if (x.state != WAVE) continue;
    const uint32_t lane = l - s0;
    int64_t src = -1;
    switch (x.kind) {
      case EMU_SHFL: src = (int64_t)s0 + (x.arg & (w - 1)); break;
      case EMU_SHFL_UP: src = lane >= (uint32_t)x.arg ? (int64_t)l - x.arg : (int64_t)l; break;
      case EMU_SHFL_DOWN: src = lane + (uint32_t)x.arg < w ? (int64_t)l + x.arg : (int64_t)l; break;
      default: break;
    }
    if (x.kind == EMU_BALLOT) {
      x.res = mask;
    } else if (src >= 0 && active((uint32_t)src)) {
      x.res = B.f[src].val;
    } else {
      x.res = 0;
      g_inactive_src.fetch_add(1);
    }
  }
  for (uint32_t l = s0; l < s0 + w && l < nthreads; ++l)
    if (B.f[l].state == WAVE) B.f[l].state = RUN;
  return true;
}

struct StackPool {
  std::vector<char*> s;
  ~StackPool() {
    for (char* p : s) munmap(p, kStack);
  }
  char* get(size_t i) {
    while (s.size() <= i) {
      void* p = mmap(nullptr, kStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
      if (p == MAP_FAILED) {
        std::perror("mmap");
        std::abort();
      }
      mprotect(p, 4096, PROT_NONE);  // guard page
      s.push_back((char*)p);
    }
    return s[i];
  }
};

static void run_block(uint32_t bx, dim3 grid, dim3 block, size_t shmem, const std::function<void()>& fn) {
  static thread_local StackPool pool;
  Block B;
  t_blk = &B;
  t_block = dim3(bx);
  t_grid = grid;
  t_bdim = block;
  B.fn = &fn;
  B.dyn.assign(shmem ? shmem : 16, (char)0xA5);
  const uint32_t n = block.x;
  B.f.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    Fiber& f = B.f[i];
    f.stack = pool.get(i);
#if EMU_ASAN
    __asan_unpoison_memory_region(f.stack + 4096, kStack - 4096);
#endif
    f.tid = dim3(i);
    uint64_t* s = (uint64_t*)(f.stack + kStack);
    *--s = 0;
    *--s = (uint64_t)(void*)&fiber_main;
    for (int r = 0; r < 6; ++r) *--s = 0;
    f.sp = s;
  }
  uint32_t live = n;
  for (;;) {
    bool progress = false, real = false;
    for (uint32_t i = 0; i < n; ++i) {
      Fiber& f = B.f[i];
      if (f.state != RUN) continue;
      t_cur = &f;
      f.slept = false;
#if EMU_ASAN
      __sanitizer_start_switch_fiber(&B.fake, f.stack + 4096, kStack - 4096);
      emu_swap(&B.sched_sp, f.sp);
      __sanitizer_finish_switch_fiber(B.fake, nullptr, nullptr);
#else
      emu_swap(&B.sched_sp, f.sp);
#endif
      progress = true;
      if (!f.slept) real = true;
      if (f.state == DONE) --live;
    }
    t_cur = nullptr;
    if (live == 0) break;
    uint32_t nbar = 0, nrun = 0;
    for (uint32_t i = 0; i < n; ++i) {
      nbar += B.f[i].state == BAR;
      nrun += B.f[i].state == RUN;
    }
    bool resolved = false;
    // wave operations: segments of the width each waiting lane asked for
    for (uint32_t i = 0; i < n; ++i) {
      if (B.f[i].state != WAVE) continue;
      const uint32_t w = (uint32_t)B.f[i].width;
      const uint32_t s0 = i & ~(w - 1);
      if (resolve_segment(B, s0, w, n)) resolved = true;
    }
    if (resolved) continue;
    if (nbar == live) {
      for (uint32_t i = 0; i < n; ++i)
        if (B.f[i].state == BAR) B.f[i].state = RUN;
      continue;
    }
    if (nrun == 0) {
      std::fprintf(stderr, "emu: deadlock in block %u: %u live, %u at barrier\n", bx, live, nbar);
      std::abort();
    }
    if (progress && !real) std::this_thread::yield();
  }
  t_blk = nullptr;
}

void launch(dim3 grid, dim3 block, size_t shmem, std::function<void()> fn) {
  const uint32_t G = grid.x;
  if (G == 0) return;
  const char* mb = std::getenv("EMU_MAX_BLOCKS");
  const uint32_t max_blocks = mb ? (uint32_t)std::atoi(mb) : 0u;
  if (max_blocks) {  // at most K workgroups resident at a time
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (uint32_t k = 0; k < max_blocks && k < G; ++k)
      th.emplace_back([&] {
        for (;;) {
          const uint32_t b = next.fetch_add(1);
          if (b >= G) break;
          run_block(b, grid, block, shmem, fn);
        }
      });
    for (auto& t : th) t.join();
    return;
  }
  if (G <= 256) {
    std::vector<std::thread> th;
    th.reserve(G);
    for (uint32_t b = 0; b < G; ++b) th.emplace_back([&, b] { run_block(b, grid, block, shmem, fn); });
    for (auto& t : th) t.join();
  } else {
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (int k = 0; k < 8; ++k)
      th.emplace_back([&] {
        for (;;) {
          const uint32_t b = next.fetch_add(1);
          if (b >= G) break;
          run_block(b, grid, block, shmem, fn);
        }
      });
    for (auto& t : th) t.join();
  }
}

}  // namespace emu

// ------------------------------------------------------------- host API
struct emu_stream {
  bool capturing = false;
  std::vector<std::function<void()>> cap;
  emu_stream* redirect = nullptr;  // joined the capture of another stream
};
static emu_stream* g_capture_of = nullptr;
static std::vector<emu_stream*> g_joined;
struct emu_graph {
  std::vector<std::function<void()>> ops;
};
struct emu_event {
  double t = 0.0;
};

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

namespace emu {
void submit(hipStream_t st, std::function<void()> fn) {
  if (st && st->redirect) st = st->redirect;
  if (st && st->capturing) st->cap.push_back(std::move(fn));
  else fn();
}
}  // namespace emu

hipError_t hipMalloc(void** p, size_t n) {
  void* q = std::malloc(n ? n : 1);
  if (!q) return hipErrorOutOfMemory;
  const char* fill = std::getenv("EMU_FILL");
  std::memset(q, fill ? std::atoi(fill) : 0, n);
  *p = q;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st) {
  emu::submit(st, [=] { std::memmove(d, s, n); });
  return hipSuccess;
}
hipError_t hipMemset(void* d, int v, size_t n) {
  std::memset(d, v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st) {
  emu::submit(st, [=] { std::memset(d, v, n); });
  return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t* s) {
  *s = new emu_stream();
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  s->capturing = true;
  s->cap.clear();
  g_capture_of = s;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  *g = new emu_graph{std::move(s->cap)};
  s->cap.clear();
  s->capturing = false;
  g_capture_of = nullptr;
  for (emu_stream* j : g_joined) j->redirect = nullptr;
  g_joined.clear();
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *e = new emu_graph{g->ops};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  emu_graph* g = e;
  emu::submit(s, [g] {
    for (auto& op : g->ops) op();
  });
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new emu_event();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
// Work runs when it is submitted (or replayed), so a cross-stream wait has
// nothing to wait for; inside a capture, the second stream's work is recorded
// into the capturing stream's graph.
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t, unsigned) {
  if (s && g_capture_of && s != g_capture_of && !s->redirect) {
    s->redirect = g_capture_of;
    g_joined.push_back(s);
  }
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  emu::submit(s, [e] { e->t = now_ms(); });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = (float)(b->t - a->t);
  return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emu error"; }
hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
  if (a == hipDeviceAttributeWallClockRate) {
    *v = 100000;  // kHz, the rate of wall_clock64 above
    return hipSuccess;
  }
  const char* cu = std::getenv("EMU_CUS");
  *v = cu ? std::atoi(cu) : 256;
  return hipSuccess;
}
