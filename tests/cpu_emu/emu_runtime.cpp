// Host execution of emulated workgroups (see hip/hip_runtime.h): the threads
// of a workgroup are ucontext fibers on the calling thread, resumed in
// round-robin order; __syncthreads() yields, so every fiber reaches a barrier
// before any leaves it.
#include "hip/hip_runtime.h"

#include <ucontext.h>

#include <memory>

emu_idx threadIdx, blockIdx;
emu_idx blockDim, gridDim;

namespace emu {
double xbuf[1024];
float mfa[1024][8], mfb[1024][8];
bool capturing = false;
std::vector<std::function<void()>> captured;

namespace {
const size_t kStack = 256 * 1024;
struct Fiber {
    ucontext_t ctx;
    std::unique_ptr<char[]> stack;
    bool done = false;
};
ucontext_t sched;
std::vector<Fiber> fibers;
unsigned cur = 0;
const std::function<void()> *cur_body = nullptr;

void fiber_main()
{
    (*cur_body)();
    fibers[cur].done = true;
    swapcontext(&fibers[cur].ctx, &sched);
}
}  // namespace

void sync() { swapcontext(&fibers[cur].ctx, &sched); }

void launch(unsigned grid, unsigned block, const std::function<void()> &body)
{
    blockDim.x = block;
    gridDim.x = grid;
    if (fibers.size() < block) fibers.resize(block);
    cur_body = &body;
    for (unsigned b = 0; b < grid; ++b) {
        blockIdx.x = b;
        for (unsigned t = 0; t < block; ++t) {
            Fiber &f = fibers[t];
            if (!f.stack) f.stack.reset(new char[kStack]);
            getcontext(&f.ctx);
            f.ctx.uc_stack.ss_sp = f.stack.get();
            f.ctx.uc_stack.ss_size = kStack;
            f.ctx.uc_link = nullptr;
            f.done = false;
            makecontext(&f.ctx, fiber_main, 0);
        }
        for (unsigned alive = block; alive > 0;) {
            alive = 0;
            for (unsigned t = 0; t < block; ++t) {
                if (fibers[t].done) continue;
                cur = t;
                threadIdx.x = t;
                swapcontext(&sched, &fibers[t].ctx);
                if (!fibers[t].done) ++alive;
            }
        }
    }
}
}  // namespace emu
