#include "emu_runtime.h"
#include <ucontext.h>
#include <vector>

namespace emu {
int g_lane = 0;
unsigned g_block = 0, g_grid = 1;
long long g_collectives = 0;
namespace {
enum State { READY, WAIT, DONE };
struct Lane {
  ucontext_t ctx;
  std::vector<char> stack;
  State st = READY;
  int kind = -1, line = 0;
  const char* file = "";
  long long gen = 0;
};
Lane lanes[W];
ucontext_t sched;
uint64_t slots[2][W];
const std::function<void()>* g_body = nullptr;
void entry(int lane) {
  (*g_body)();
  lanes[lane].st = DONE;
  swapcontext(&lanes[lane].ctx, &sched);
}
void dump(const char* why) {
  std::fprintf(stderr, "EMU: %s (block %u)\n", why, g_block);
  for (int i = 0; i < W; ++i) {
    const Lane& L = lanes[i];
    std::fprintf(stderr, "  lane %2d: %s kind %d at %s:%d gen %lld\n", i, L.st == DONE ? "done" : (L.st == WAIT ? "wait" : "ready"),
                 L.kind, L.file, L.line, L.gen);
  }
  std::exit(3);
}
}  // namespace

const uint64_t* collective(uint64_t v, int kind, const char* file, int line) {
  Lane& L = lanes[g_lane];
  const int buf = (int)(L.gen & 1);
  slots[buf][g_lane] = v;
  L.st = WAIT;
  L.kind = kind;
  L.file = file;
  L.line = line;
  const int me = g_lane;
  swapcontext(&L.ctx, &sched);
  g_lane = me;
  ++lanes[me].gen;
  return slots[buf];
}

void run_workgroup(const std::function<void()>& body, unsigned block, unsigned grid) {
  g_body = &body;
  g_block = block;
  g_grid = grid;
  for (int i = 0; i < W; ++i) {
    Lane& L = lanes[i];
    L.stack.assign(1 << 20, 0);
    getcontext(&L.ctx);
    L.ctx.uc_stack.ss_sp = L.stack.data();
    L.ctx.uc_stack.ss_size = L.stack.size();
    L.ctx.uc_link = nullptr;
    makecontext(&L.ctx, (void (*)())entry, 1, i);
    L.st = READY;
    L.gen = 0;
    L.kind = -1;
  }
  for (;;) {
    for (int i = 0; i < W; ++i) {
      if (lanes[i].st != READY) continue;
      g_lane = i;
      swapcontext(&sched, &lanes[i].ctx);
    }
    int ndone = 0, nwait = 0;
    for (int i = 0; i < W; ++i) ndone += lanes[i].st == DONE, nwait += lanes[i].st == WAIT;
    if (ndone == W) break;
    if (ndone) dump("some lanes finished while others wait at a collective");
    for (int i = 1; i < W; ++i)
      if (lanes[i].kind != lanes[0].kind || lanes[i].line != lanes[0].line || lanes[i].file != lanes[0].file ||
          lanes[i].gen != lanes[0].gen)
        dump("lanes at different collectives");
    ++g_collectives;
    for (int i = 0; i < W; ++i) lanes[i].st = READY;
  }
}
}  // namespace emu

void emu_report_nonuniform(const char* f, int l, int a, int b) {
  std::fprintf(stderr, "EMU: readlane with a non-uniform or out-of-range lane index at %s:%d (%d vs %d)\n", f, l, a, b);
}
