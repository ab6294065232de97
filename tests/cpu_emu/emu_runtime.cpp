// Host execution of emulated workgroups (see hip/hip_runtime.h): the threads
// of a workgroup are ucontext fibers on the calling thread, resumed in
// round-robin order.  sync() is a barrier of the calling thread's wave (the
// lane exchanges and matrix operations), block_sync() one of the whole
// workgroup (__syncthreads): a fiber waiting at one yields until every live
// fiber of its wave (or block) has arrived, so waves may run different
// instruction streams between workgroup barriers, as on the GPU.
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

void release_if_complete(unsigned t);

void fiber_main()
{
    (*cur_body)();
    fibers[cur].done = true;
    release_if_complete(cur);   // barriers the finished thread will not reach
    swapcontext(&fibers[cur].ctx, &sched);
}
}  // namespace

// per wave and for the block: live threads, arrivals at the current barrier,
// and the barrier's generation
unsigned wave_live[16], wave_arrived[16], block_live = 0, block_arrived = 0;
unsigned long wave_gen[16], block_gen = 0;

namespace {
unsigned wave_of(unsigned t) { return t / 64; }

void release_if_complete(unsigned t)
{
    const unsigned w = wave_of(t);
    --wave_live[w];
    --block_live;
    if (wave_arrived[w] && wave_arrived[w] == wave_live[w]) {
        wave_arrived[w] = 0;
        ++wave_gen[w];
    }
    if (block_arrived && block_arrived == block_live) {
        block_arrived = 0;
        ++block_gen;
    }
}
}  // namespace

void sync()
{
    const unsigned t = cur, w = wave_of(t);
    const unsigned long g = wave_gen[w];
    if (++wave_arrived[w] == wave_live[w]) {
        wave_arrived[w] = 0;
        ++wave_gen[w];
    }
    while (wave_gen[w] == g) swapcontext(&fibers[t].ctx, &sched);
}

void block_sync()
{
    const unsigned t = cur;
    const unsigned long g = block_gen;
    if (++block_arrived == block_live) {
        block_arrived = 0;
        ++block_gen;
    }
    while (block_gen == g) swapcontext(&fibers[t].ctx, &sched);
}

void launch(unsigned grid, unsigned block, const std::function<void()> &body)
{
    blockDim.x = block;
    gridDim.x = grid;
    if (fibers.size() < block) fibers.resize(block);
    cur_body = &body;
    for (unsigned b = 0; b < grid; ++b) {
        blockIdx.x = b;
        block_arrived = 0;
        block_live = block;
        for (unsigned w = 0; w < 16; ++w) {
            wave_arrived[w] = 0;
            const unsigned lo = 64 * w;
            wave_live[w] = lo >= block ? 0 : (block - lo < 64 ? block - lo : 64);
        }
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
