// CPU emulation runtime for the HIP subset in hip/hip_runtime.h (debug tool).
#include <hip/hip_runtime.h>

#include <sys/mman.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <thread>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define EMU_ASAN 1
#endif
#endif
#ifdef EMU_ASAN
extern "C" void __sanitizer_start_switch_fiber(void **fake_stack_save, const void *bottom, size_t size);
extern "C" void __sanitizer_finish_switch_fiber(void *fake_stack_save, const void **bottom_old,
                                                size_t *size_old);
#endif

extern "C" void emu_swap(void **save_sp, void *load_sp);
asm(R"(
.text
.globl emu_swap
.type emu_swap,@function
emu_swap:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  subq $16, %rsp
  stmxcsr 8(%rsp)
  fnstcw (%rsp)
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  fldcw (%rsp)
  ldmxcsr 8(%rsp)
  addq $16, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
.size emu_swap,.-emu_swap
)");

namespace emu {

ThreadState *g_cur = nullptr;
uint3_ g_bidx{0, 0, 0};
dim3 g_bdim, g_gdim;

namespace {

constexpr size_t kStack = size_t(4) << 20;  // includes one guard page at the bottom
constexpr int kMaxThreads = 1024;
enum State { kRun = 0, kBarrier = 1, kWaveOp = 2, kDone = 3 };

struct Fiber {
  void *sp = nullptr;
  char *stack = nullptr;
  ThreadState ts{};
  int state = kDone;
  int op = 0, arg = 0, width = 64, line = 0;
  const char *file = "";
  uint64_t in = 0, out = 0;
  void *fake = nullptr;
};

std::recursive_mutex g_mu;
Fiber g_fib[kMaxThreads];
void *g_sched_sp = nullptr;
void *g_sched_fake = nullptr;
const void *g_sched_bottom = nullptr;
size_t g_sched_size = 0;
Fiber *g_curf = nullptr;
const std::function<void()> *g_body = nullptr;
const char *g_kname = "";
hipError_t g_last = hipSuccess;
std::set<std::string> g_warned;
unsigned long long g_launches = 0;

std::atomic<const char *> g_watch_name{""};
std::atomic<unsigned> g_watch_bx{0};
std::atomic<unsigned long long> g_watch_launch{0};
std::atomic<unsigned long long> g_watch_blocks{0};

struct Watchdog {
  Watchdog() {
    const char *e = std::getenv("EMU_WATCH");
    if (!e)
      return;
    const int sec = std::atoi(e);
    std::thread([sec] {
      unsigned long long last_launch = 0, last_blocks = 0;
      for (;;) {
        std::this_thread::sleep_for(std::chrono::seconds(sec));
        const unsigned long long l = g_watch_launch.load(), b = g_watch_blocks.load();
        std::fprintf(stderr, "[emu-watch] launches %llu (+%llu) blocks %llu (+%llu) now in %s block %u\n",
                     l, l - last_launch, b, b - last_blocks, g_watch_name.load(), g_watch_bx.load());
        std::fflush(stderr);
        last_launch = l;
        last_blocks = b;
      }
    }).detach();
  }
} g_watchdog;

void report(const char *what, const std::string &detail, bool fatal) {
  std::string key = std::string(what) + "|" + detail;
  if (!fatal && !g_warned.insert(key).second)
    return;
  std::fprintf(stderr, "[emu] %s: %s (kernel %s, block %u,%u,%u)\n", what, detail.c_str(), g_kname,
               g_bidx.x, g_bidx.y, g_bidx.z);
  std::fflush(stderr);
  if (const char *log = std::getenv("EMU_LOG")) {
    if (FILE *fh = std::fopen(log, "a")) {
      std::fprintf(fh, "[emu %d] %s: %s (kernel %s, block %u,%u,%u)\n", static_cast<int>(getpid()), what,
                   detail.c_str(), g_kname, g_bidx.x, g_bidx.y, g_bidx.z);
      std::fclose(fh);
    }
  }
  if (fatal)
    std::abort();
}

void to_sched(bool finishing) {
  Fiber *f = g_curf;
#ifdef EMU_ASAN
  __sanitizer_start_switch_fiber(finishing ? nullptr : &f->fake, g_sched_bottom, g_sched_size);
#endif
  emu_swap(&f->sp, g_sched_sp);
#ifdef EMU_ASAN
  __sanitizer_finish_switch_fiber(f->fake, nullptr, nullptr);
#endif
}

void fiber_entry() {
#ifdef EMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
#endif
  (*g_body)();
  g_curf->state = kDone;
  to_sched(true);
  std::abort();
}

void init_fiber(Fiber &f) {
  if (!f.stack) {
    void *m = mmap(nullptr, kStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                   -1, 0);
    if (m == MAP_FAILED) {
      std::perror("mmap");
      std::abort();
    }
    f.stack = static_cast<char *>(m);
    if (mprotect(m, 4096, PROT_NONE) != 0) {
      std::perror("mprotect");
      std::abort();
    }
  }
  uintptr_t top = (reinterpret_cast<uintptr_t>(f.stack) + kStack) & ~uintptr_t(15);
  uint64_t *s = reinterpret_cast<uint64_t *>(top);
  s[-1] = 0;                                             // return address of fiber_entry
  s[-2] = reinterpret_cast<uint64_t>(&fiber_entry);      // ret target
  for (int k = 3; k <= 8; ++k)                           // rbp rbx r12 r13 r14 r15
    s[-k] = 0;
  uint32_t *ctl = reinterpret_cast<uint32_t *>(top - 80);
  ctl[0] = 0x037F;  // x87 control word
  ctl[1] = 0;
  ctl[2] = 0x1F80;  // mxcsr
  ctl[3] = 0;
  f.sp = reinterpret_cast<void *>(top - 80);
  f.fake = nullptr;
}

void resume(Fiber &f) {
  g_curf = &f;
  g_cur = &f.ts;
#ifdef EMU_ASAN
  __sanitizer_start_switch_fiber(&g_sched_fake, f.stack, kStack);
#endif
  emu_swap(&g_sched_sp, f.sp);
#ifdef EMU_ASAN
  __sanitizer_finish_switch_fiber(g_sched_fake, nullptr, nullptr);
#endif
}

void resolve_wave(int base, int n) {
  Fiber *L = g_fib + base;
  int op = -1, line = -1, width = -1;
  const char *file = "";
  for (int l = 0; l < n; ++l) {
    if (L[l].state == kDone)
      continue;
    if (op < 0) {
      op = L[l].op;
      line = L[l].line;
      width = L[l].width;
      file = L[l].file;
    } else if (L[l].op != op || L[l].line != line || L[l].width != width) {
      report("divergent wave op",
             std::string(file) + ":" + std::to_string(line) + " vs " + L[l].file + ":" +
                 std::to_string(L[l].line) + " (lane " + std::to_string(l) + ")",
             true);
    }
  }
  if (op == kOpBallot) {
    uint64_t m = 0;
    for (int l = 0; l < n; ++l)
      if (L[l].state != kDone && L[l].in)
        m |= uint64_t(1) << l;
    for (int l = 0; l < n; ++l)
      if (L[l].state != kDone) {
        L[l].out = m;
        L[l].state = kRun;
      }
    return;
  }
  for (int l = 0; l < n; ++l) {
    if (L[l].state == kDone)
      continue;
    const int self = base % 64 + l;  // lane index in the wave
    const int w = L[l].width, a = L[l].arg;
    int idx = self;
    switch (op) {
      case kOpShfl: idx = (a + (self & ~(w - 1))) & 63; break;
      case kOpShflUp: idx = self - a; if (idx < (self & ~(w - 1))) idx = self; break;
      case kOpShflDown: idx = self + a; if ((self & (w - 1)) + a >= w) idx = self; break;
      case kOpShflXor: idx = self ^ a; if (idx >= ((self + w) & ~(w - 1))) idx = self; break;
      default: report("unknown wave op", std::to_string(op), true);
    }
    const int src = idx - base % 64;
    if (src < 0 || src >= n || L[src].state == kDone) {
      report("shuffle reads an inactive lane",
             std::string(L[l].file) + ":" + std::to_string(L[l].line) + " lane " +
                 std::to_string(self) + " <- " + std::to_string(idx),
             false);
      L[l].out = 0xDEADBEEFDEADBEEFull;
    } else {
      L[l].out = L[src].in;
    }
  }
  for (int l = 0; l < n; ++l)
    if (L[l].state != kDone)
      L[l].state = kRun;
}

void run_block(const std::function<void()> &fn, int n) {
  g_body = &fn;
  for (int t = 0; t < n; ++t) {
    Fiber &f = g_fib[t];
    init_fiber(f);
    f.state = kRun;
    f.ts.idx = uint3_{static_cast<unsigned>(t % g_bdim.x),
                      static_cast<unsigned>((t / g_bdim.x) % g_bdim.y),
                      static_cast<unsigned>(t / (g_bdim.x * g_bdim.y))};
  }
  for (;;) {
    for (int t = 0; t < n; ++t)
      if (g_fib[t].state == kRun)
        resume(g_fib[t]);
    int live = 0;
    for (int t = 0; t < n; ++t)
      live += g_fib[t].state != kDone;
    if (live == 0)
      break;
    bool progress = false;
    for (int base = 0; base < n; base += 64) {
      const int m = std::min(64, n - base);
      bool any = false, all = true;
      for (int l = 0; l < m; ++l) {
        const int s = g_fib[base + l].state;
        if (s == kDone)
          continue;
        any = true;
        if (s != kWaveOp)
          all = false;
      }
      if (any && all) {
        resolve_wave(base, m);
        progress = true;
      }
    }
    if (progress)
      continue;
    bool all_barrier = true;
    for (int t = 0; t < n; ++t)
      if (g_fib[t].state != kDone && g_fib[t].state != kBarrier)
        all_barrier = false;
    if (all_barrier) {
      int line = -1;
      for (int t = 0; t < n; ++t)
        if (g_fib[t].state == kBarrier) {
          if (line < 0)
            line = g_fib[t].line;
          else if (g_fib[t].line != line)
            report("threads at different barriers",
                   std::to_string(line) + " vs " + std::to_string(g_fib[t].line), false);
          g_fib[t].state = kRun;
        }
      continue;
    }
    std::string s;
    for (int t = 0; t < n; ++t)
      if (g_fib[t].state != kDone)
        s += " t" + std::to_string(t) + ":" + (g_fib[t].state == kBarrier ? "bar@" : "wave@") +
             std::to_string(g_fib[t].line);
    report("deadlock", s.substr(0, 600), true);
  }
  g_body = nullptr;
  g_cur = nullptr;
}

struct Alloc {
  size_t n;
  int kind;  // 0 device, 1 pinned, 2 symbol
};
std::map<uintptr_t, Alloc> g_allocs;

bool inside(const void *p, size_t n, int kinds) {
  if (n == 0)
    return true;
  uintptr_t a = reinterpret_cast<uintptr_t>(p);
  auto it = g_allocs.upper_bound(a);
  if (it == g_allocs.begin())
    return false;
  --it;
  if (!((kinds >> it->second.kind) & 1))
    return false;
  return a >= it->first && a + n <= it->first + it->second.n;
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct Stream {
  bool capturing = false;
  std::vector<std::function<void()>> ops;
  std::deque<std::function<void()>> q;  // deferred work (EMU_DEFER=1)
};
struct Event {
  double t = 0;
  bool done = true;
  Stream *stream = nullptr;
};
struct Graph {
  std::vector<std::function<void()>> ops;
  int pending = 0;  // launches queued and not yet run
};

bool g_defer = [] {
  const char *e = std::getenv("EMU_DEFER");
  return e && std::atoi(e) != 0;
}();
Stream g_null_stream;
std::set<Stream *> g_streams{&g_null_stream};

Stream *sref(hipStream_t s) { return s ? s : &g_null_stream; }

bool run_one(Stream *s) {
  if (s->q.empty())
    return false;
  std::function<void()> op = std::move(s->q.front());
  s->q.pop_front();
  op();
  return true;
}
void drain(Stream *s) {
  while (run_one(s)) {
  }
}
void drain_all() {
  for (bool any = true; any;) {
    any = false;
    for (Stream *s : g_streams)
      while (run_one(s))
        any = true;
  }
}

uint64_t wave_op(int op, uint64_t bits, int arg, int width, int line, const char *file) {
  Fiber *f = g_curf;
  if (!f)
    report("wave op outside a kernel", file, true);
  f->op = op;
  f->in = bits;
  f->arg = arg;
  f->width = width;
  f->line = line;
  f->file = file;
  f->state = kWaveOp;
  to_sched(false);
  return f->out;
}

void barrier(int line, const char *file) {
  Fiber *f = g_curf;
  if (!f)
    report("barrier outside a kernel", file, true);
  f->line = line;
  f->file = file;
  f->state = kBarrier;
  to_sched(false);
}

unsigned lane_id() {
  const unsigned t = g_cur->idx.x + g_bdim.x * (g_cur->idx.y + g_bdim.y * g_cur->idx.z);
  return t % 64;
}

static void run_grid(const char *name, dim3 g, dim3 b, const std::function<void()> &fn) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  const unsigned n = b.x * b.y * b.z;
  if (n == 0 || n > static_cast<unsigned>(kMaxThreads) || g.x == 0 || g.y == 0 || g.z == 0) {
    std::fprintf(stderr, "[emu] invalid launch configuration of %s: grid %u,%u,%u block %u,%u,%u\n",
                 name, g.x, g.y, g.z, b.x, b.y, b.z);
    g_last = hipErrorInvalidConfiguration;
    return;
  }
  const char *prev = g_kname;
  g_kname = name;
  g_bdim = b;
  g_gdim = g;
  ++g_launches;
  g_watch_launch.fetch_add(1);
  for (unsigned z = 0; z < g.z; ++z)
    for (unsigned y = 0; y < g.y; ++y)
      for (unsigned x = 0; x < g.x; ++x) {
        g_bidx = uint3_{x, y, z};
        g_watch_name.store(name);
        g_watch_bx.store(x);
        g_watch_blocks.fetch_add(1);
        run_block(fn, static_cast<int>(n));
      }
  g_kname = prev;
}

void check_arg(const char *kernel, int index, const void *p, size_t n) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  const unsigned char *b = static_cast<const unsigned char *>(p);
  for (size_t o = 0; o + 8 <= n; o += 8) {
    uint64_t w;
    std::memcpy(&w, b + o, 8);
    if (w < 0x500000000000ull || w >= 0x800000000000ull || (w & 3))
      continue;
    auto it = g_allocs.upper_bound(static_cast<uintptr_t>(w));
    bool ok = false;
    if (it != g_allocs.begin()) {
      --it;
      ok = w >= it->first && w <= it->first + it->second.n;
    }
    if (!ok) {
      const char *prev = g_kname;
      g_kname = kernel;
      report("kernel argument holds a host address outside device memory",
             "argument " + std::to_string(index) + " byte " + std::to_string(o), false);
      g_kname = prev;
    }
  }
}

static void enqueue(hipStream_t s, std::function<void()> op) {
  if (s && s->capturing)
    s->ops.push_back(std::move(op));
  else if (g_defer)
    sref(s)->q.push_back(std::move(op));
  else
    op();
}

void submit(hipStream_t s, const char *name, dim3 g, dim3 b, std::function<void()> fn) {
  enqueue(s, [name, g, b, fn]() { run_grid(name, g, b, fn); });
}

}  // namespace emu

using emu::g_allocs;

hipError_t hipMalloc(void **p, size_t n) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  const size_t m = (std::max<size_t>(n, 1) + 255) & ~size_t(255);
  void *q = std::aligned_alloc(256, m);
  if (!q)
    return hipErrorOutOfMemory;
  static const int fill = [] {
    const char *e = std::getenv("EMU_FILL");
    return e ? std::atoi(e) : 0xFF;
  }();
  std::memset(q, fill, m);
  g_allocs[reinterpret_cast<uintptr_t>(q)] = emu::Alloc{n, 0};
  *p = q;
  return hipSuccess;
}
hipError_t hipFree(void *p) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  if (!p)
    return hipSuccess;
  emu::drain_all();  // hipFree synchronizes the device
  auto it = g_allocs.find(reinterpret_cast<uintptr_t>(p));
  if (it == g_allocs.end() || it->second.kind != 0) {
    emu::report("hipFree of a pointer hipMalloc did not return", "", true);
    return hipErrorInvalidValue;
  }
  g_allocs.erase(it);
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  void *q = std::aligned_alloc(256, (std::max<size_t>(n, 1) + 255) & ~size_t(255));
  if (!q)
    return hipErrorOutOfMemory;
  g_allocs[reinterpret_cast<uintptr_t>(q)] = emu::Alloc{n, 1};
  *p = q;
  return hipSuccess;
}
hipError_t hipHostFree(void *p) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  if (!p)
    return hipSuccess;
  emu::drain_all();
  auto it = g_allocs.find(reinterpret_cast<uintptr_t>(p));
  if (it == g_allocs.end() || it->second.kind != 1) {
    emu::report("hipHostFree of a pointer hipHostMalloc did not return", "", true);
    return hipErrorInvalidValue;
  }
  g_allocs.erase(it);
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostGetDevicePointer(void **d, void *h, unsigned) {
  if (!emu::inside(h, 1, 2)) {
    emu::report("hipHostGetDevicePointer of memory hipHostMalloc did not return", "", true);
    return hipErrorInvalidValue;
  }
  *d = h;
  return hipSuccess;
}

static void check_copy(void *d, const void *s, size_t n, hipMemcpyKind k) {
  const int dev = 1 | 2 | 4;  // device, pinned (mapped), symbol
  bool ok = true;
  switch (k) {
    case hipMemcpyHostToDevice: ok = emu::inside(d, n, dev); break;
    case hipMemcpyDeviceToHost: ok = emu::inside(s, n, dev); break;
    case hipMemcpyDeviceToDevice: ok = emu::inside(d, n, dev) && emu::inside(s, n, dev); break;
    default: break;
  }
  if (!ok)
    emu::report("copy touches memory outside a device allocation",
                "kind " + std::to_string(int(k)) + " bytes " + std::to_string(n), true);
}

// host side of a copy in page-locked memory (then the copy is asynchronous)
static bool host_side_pinned(void *d, const void *s, size_t n, hipMemcpyKind k) {
  if (k == hipMemcpyHostToDevice)
    return emu::inside(s, n, 2);
  if (k == hipMemcpyDeviceToHost)
    return emu::inside(d, n, 2);
  return true;
}

hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  check_copy(d, s, n, k);
  // ordered after the null stream only (the library's streams are non-blocking)
  emu::drain(&emu::g_null_stream);
  std::memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  check_copy(d, s, n, k);
  if (emu::g_defer && !(st && st->capturing) && !host_side_pinned(d, s, n, k)) {
    // pageable host memory: the call returns when the copy is done
    emu::drain(emu::sref(st));
    std::memcpy(d, s, n);
    return hipSuccess;
  }
  emu::enqueue(st, [d, s, n]() { std::memcpy(d, s, n); });
  return hipSuccess;
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  if (!emu::inside(d, n, 1 | 2 | 4))
    emu::report("memset outside a device allocation", std::to_string(n), true);
  if (st && st->capturing)
    emu::report("memset node in a captured graph", "", false);
  emu::enqueue(st, [d, v, n]() { std::memset(d, v, n); });
  return hipSuccess;
}
hipError_t hipMemset(void *d, int v, size_t n) {
  hipMemsetAsync(d, v, n, nullptr);
  emu::drain(&emu::g_null_stream);
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  *s = new emu::Stream();
  emu::g_streams.insert(*s);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  emu::drain(s);
  emu::g_streams.erase(s);
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  if (s && s->capturing)
    emu::report("stream synchronize during capture", "", true);
  emu::drain(emu::sref(s));
  return hipSuccess;
}
hipError_t hipStreamQuery(hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  emu::Stream *S = emu::sref(s);
  emu::run_one(S);  // progress for host loops that poll
  return S->q.empty() ? hipSuccess : hipErrorNotReady;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (!s || s->capturing)
    return hipErrorInvalidValue;
  s->capturing = true;
  s->ops.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (!s || !s->capturing)
    return hipErrorInvalidValue;
  s->capturing = false;
  auto *G = new emu::Graph();
  G->ops.swap(s->ops);
  *g = G;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  auto *E = new emu::Graph();
  E->ops = g->ops;
  *e = E;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  if (s && s->capturing) {
    for (auto &op : e->ops)
      s->ops.push_back(op);
    return hipSuccess;
  }
  if (!emu::g_defer) {
    for (auto &op : e->ops)
      op();
    return hipSuccess;
  }
  ++e->pending;
  emu::sref(s)->q.push_back([e]() {
    for (auto &op : e->ops)
      op();
    --e->pending;
  });
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  if (e && e->pending > 0)
    emu::report("graph exec destroyed with launches still queued", std::to_string(e->pending), true);
  delete e;
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new emu::Event();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  if (!(s && s->capturing) && emu::g_defer) {
    e->done = false;
    e->stream = emu::sref(s);
  }
  emu::enqueue(s, [e]() {
    e->t = emu::now_ms();
    e->done = true;
  });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  while (!e->done && e->stream && emu::run_one(e->stream)) {
  }
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  // the work recorded before the event (on its stream) runs first
  if (!(s && s->capturing))
    while (!e->done && e->stream && emu::run_one(e->stream)) {
    }
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  if (!a->done || !b->done)
    return hipErrorNotReady;
  *ms = static_cast<float>(b->t - a->t);
  return hipSuccess;
}
hipError_t hipGetLastError() {
  hipError_t e = emu::g_last;
  emu::g_last = hipSuccess;
  return e;
}
const char *hipGetErrorString(hipError_t e) {
  switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorInvalidValue: return "hipErrorInvalidValue";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorInvalidConfiguration: return "hipErrorInvalidConfiguration";
    case hipErrorLaunchTimeOut: return "hipErrorLaunchTimeOut";
    default: return "hipError";
  }
}
hipError_t hipGetDeviceCount(int *n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  emu::drain_all();
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  std::lock_guard<std::recursive_mutex> lk(emu::g_mu);
  // the symbol's size is unknown here: register a 4 MiB window, the size of
  // the only symbol the library looks up
  g_allocs[reinterpret_cast<uintptr_t>(sym)] = emu::Alloc{size_t(4) << 20, 2};
  *p = const_cast<void *>(sym);
  return hipSuccess;
}

extern "C" unsigned long long emu_launch_count() { return emu::g_launches; }
