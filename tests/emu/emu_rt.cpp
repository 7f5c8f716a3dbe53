#include <hip/hip_runtime.h>
hipError_t emu_last_error = 0;
namespace emu {
dim3 g_grid, g_block, g_bidx;
unsigned char* g_dyn = nullptr;
size_t g_dyn_bytes = 0;
static Block B;
Block& blk() { return B; }
static void tramp() {
  B.body();
  B.f[B.cur].state = 2;
  swapcontext(&B.f[B.cur].ctx, &B.sched);
}
void barrier() {
  Fiber& m = B.f[B.cur];
  m.state = 1;
  swapcontext(&m.ctx, &B.sched);
}
void run_block(std::function<void()> body, unsigned n) {
  B.body = body;
  B.n = n;
  if (B.f.size() < n) B.f.resize(n);
  for (unsigned i = 0; i < n; ++i) {
    Fiber& f = B.f[i];
    if (f.stack.empty()) f.stack.resize(256 * 1024);
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack.data();
    f.ctx.uc_stack.ss_size = f.stack.size();
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, tramp, 0);
    f.tid = i;
    f.state = 0;
  }
  for (;;) {
    bool any_run = false;
    for (unsigned i = 0; i < n; ++i) {
      if (B.f[i].state != 0) continue;
      B.cur = (int)i;
      swapcontext(&B.sched, &B.f[i].ctx);
      any_run = true;
    }
    unsigned done = 0, wait = 0;
    for (unsigned i = 0; i < n; ++i) {
      if (B.f[i].state == 2) ++done;
      else if (B.f[i].state == 1) ++wait;
    }
    if (done == n) break;
    if (wait && !any_run) {
      // everyone alive is at the barrier: release
      if (done) { fprintf(stderr, "emu: barrier with %u exited threads of %u\n", done, n); }
      for (unsigned i = 0; i < n; ++i)
        if (B.f[i].state == 1) B.f[i].state = 0;
    } else if (!any_run) {
      fprintf(stderr, "emu: deadlock\n");
      abort();
    }
  }
  B.cur = -1;
}
}  // namespace emu
