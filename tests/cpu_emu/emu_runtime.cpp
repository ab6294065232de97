// Host execution of emulated workgroups (see hip/hip_runtime.h).
#include "hip/hip_runtime.h"

thread_local emu_idx threadIdx, blockIdx;
emu_idx blockDim, gridDim;

namespace emu {
pthread_barrier_t bar;
double xbuf[1024];
bool capturing = false;
std::vector<std::function<void()>> captured;

void launch(unsigned grid, unsigned block, const std::function<void()> &body)
{
    blockDim.x = block;
    gridDim.x = grid;
    for (unsigned b = 0; b < grid; ++b) {
        pthread_barrier_init(&bar, nullptr, block);
        std::vector<std::thread> th;
        th.reserve(block);
        for (unsigned t = 0; t < block; ++t)
            th.emplace_back([&, b, t]() {
                threadIdx.x = t;
                blockIdx.x = b;
                body();
            });
        for (auto &x : th) x.join();
        pthread_barrier_destroy(&bar);
    }
}
}  // namespace emu
