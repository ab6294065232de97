// CPU emulation of the small HIP surface used by deep-green_amd/csrc.
// TEST INFRASTRUCTURE ONLY: lets g++ compile the HIP kernels for the host so
// their logic (indexing, barriers, shuffles, boundary rules) can be checked
// bit for bit against the oracle on machines without a GPU.  Each workgroup
// runs as blockDim cooperative fibers on one host thread (a barrier yields to
// the next fiber; a full round of yields is a full barrier); blocks run in
// sequence.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <functional>
#include <tuple>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define amdgpu_waves_per_eu(...) unused
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx { unsigned x = 0, y = 0, z = 0; };
extern emu_idx threadIdx, blockIdx;
extern emu_idx blockDim, gridDim;

namespace emu {
extern double xbuf[1024];
void sync();         // barrier of the calling thread's wave
void block_sync();   // barrier of the workgroup
template <typename T> T exch(T v, int src_lane_delta, int mode)
{
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid & ~63;
    sync();
    xbuf[tid] = (double)v;
    std::memcpy(&xbuf[tid], &v, sizeof(T));
    sync();
    int src;
    if (mode == 0) src = lane + src_lane_delta < 64 ? lane + src_lane_delta : lane;   // down
    else if (mode == 1) src = lane - src_lane_delta >= 0 ? lane - src_lane_delta : lane; // up
    else src = lane ^ src_lane_delta;                                                    // xor
    T out;
    std::memcpy(&out, &xbuf[wave + src], sizeof(T));
    sync();
    return out;
}
inline int exch_dpp(int old, int v, int dir)
{
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid & ~63;
    sync();
    std::memcpy(&xbuf[tid], &v, sizeof(int));
    sync();
    const int src = lane + dir;
    int out = old;
    if (src >= 0 && src < 64) std::memcpy(&out, &xbuf[wave + src], sizeof(int));
    sync();
    return out;
}
}  // namespace emu

inline void __syncthreads() { emu::block_sync(); }
// wave vote; every thread of the block must reach it (as every call site in
// the kernels does)
inline int __any(int p)
{
    const int tid = (int)threadIdx.x, wave = tid & ~63;
    emu::sync();
    emu::xbuf[tid] = p ? 1.0 : 0.0;
    emu::sync();
    int r = 0;
    for (int l = 0; l < 64 && wave + l < (int)blockDim.x; ++l) r |= emu::xbuf[wave + l] != 0.0;
    emu::sync();
    return r;
}
// only called with wave-uniform values in the kernels
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }
template <typename T> inline T __shfl_down(T v, unsigned d) { return emu::exch<T>(v, (int)d, 0); }
template <typename T> inline T __shfl_up(T v, unsigned d) { return emu::exch<T>(v, (int)d, 1); }
template <typename T> inline T __shfl_xor(T v, int m) { return emu::exch<T>(v, m, 2); }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
inline float __builtin_amdgcn_sqrtf(float x) { return std::sqrt(x); }
inline float __builtin_amdgcn_rsqf(float x) { return 1.0f / std::sqrt(x); }
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline double __builtin_amdgcn_rsq(double x) { return 1.0 / std::sqrt(x); }
inline float __builtin_amdgcn_fmed3f(float a, float b, float c)   // median of three (ordered inputs)
{
    return std::max(std::min(a, b), std::min(std::max(a, b), c));
}
// DPP wave_shl:1 (0x130: lane i <- lane i+1) and wave_shr:1 (0x138: lane i <- lane i-1);
// lanes without a source keep `old`.
inline int __builtin_amdgcn_update_dpp(int old, int v, int ctrl, int, int, bool)
{
    return ctrl == 0x130 ? emu::exch_dpp(old, v, +1) : emu::exch_dpp(old, v, -1);
}

typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
enum { hipSuccess = 0, hipErrorUnknown = 999 };
enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3, hipStreamNonBlocking = 1,
       hipHostMallocDefault = 0, hipStreamCaptureModeThreadLocal = 0,
       hipDeviceAttributeMultiprocessorCount = 16, hipEventDisableTiming = 2 };
typedef int hipMemcpyKind;

namespace emu {
void launch(unsigned grid, unsigned block, const std::function<void()> &body);
extern bool capturing;
extern std::vector<std::function<void()>> captured;
}  // namespace emu

// Arguments are evaluated at launch (as HIP does), so a captured launch
// replays with the values it was recorded with.
#define hipLaunchKernelGGL(K, G, B, SH, S, ...)                                          \
    do {                                                                                 \
        auto emu_args_ = std::make_tuple(__VA_ARGS__);                                   \
        auto emu_body_ = [=]() { std::apply(K, emu_args_); };                            \
        const unsigned g_ = dim3(G).x, b_ = dim3(B).x;                                   \
        if (emu::capturing) emu::captured.push_back([=]() { emu::launch(g_, b_, emu_body_); }); \
        else emu::launch(g_, b_, emu_body_);                                             \
    } while (0)

inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
// a __device__ array is a host array here: its address is the symbol itself
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipMalloc(void **p, size_t n) { *p = std::calloc(1, n ? n : 1); return *p ? hipSuccess : 2; }
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void **p, size_t n, int) { return hipMalloc(p, n); }
template <typename T> inline hipError_t hipHostMalloc(T **p, size_t n, int f) { return hipHostMalloc((void **)p, n, f); }
inline hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, int) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int, hipStream_t)
{
    if (emu::capturing) return hipErrorUnknown;
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemset(void *p, int v, size_t n) { std::memset(p, v, n); return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, int) { *s = (void *)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = (void *)1; return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = (void *)1; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0; return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { emu::capturing = true; emu::captured.clear(); return hipSuccess; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g)
{
    emu::capturing = false;
    *g = new std::vector<std::function<void()>>(emu::captured);
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t)
{
    *e = new std::vector<std::function<void()>>(*static_cast<std::vector<std::function<void()>> *>(g));
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete static_cast<std::vector<std::function<void()>> *>(g); return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t g) { delete static_cast<std::vector<std::function<void()>> *>(g); return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t g, hipStream_t)
{
    for (auto &f : *static_cast<std::vector<std::function<void()>> *>(g)) f();
    return hipSuccess;
}
// A small device: 4 CUs, one resident workgroup per CU (persistent kernels
// then stride over many items per workgroup).
inline hipError_t hipDeviceGetAttribute(int *v, int, int) { *v = 4; return hipSuccess; }
struct hipDeviceProp_t { char gcnArchName[256]; int multiProcessorCount; };
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int)
{
    std::strcpy(p->gcnArchName, "emulated");
    p->multiProcessorCount = 4;
    return hipSuccess;
}
template <typename F> inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, F, int, size_t)
{
    *n = 1;
    return hipSuccess;
}
// Matrix-core operations, computed from the operands of all 64 lanes of the
// wave with the documented lane layouts (see dg_cnn.hip); sums in fp32.
typedef __bf16 emu_bf16x8 __attribute__((ext_vector_type(8)));
typedef float emu_f32x16 __attribute__((ext_vector_type(16)));
typedef float emu_f32x4 __attribute__((ext_vector_type(4)));
namespace emu {
extern float mfa[1024][8], mfb[1024][8];
inline void mf_publish(const emu_bf16x8 &a, const emu_bf16x8 &b)
{
    const int tid = (int)threadIdx.x;
    sync();
    for (int i = 0; i < 8; ++i) { mfa[tid][i] = (float)a[i]; mfb[tid][i] = (float)b[i]; }
    sync();
}
}  // namespace emu
inline emu_f32x16 __builtin_amdgcn_mfma_f32_32x32x16_bf16(emu_bf16x8 a, emu_bf16x8 b, emu_f32x16 c, int, int, int)
{
    emu::mf_publish(a, b);
    const int tid = (int)threadIdx.x, l = tid & 63, w = tid & ~63, j = l % 32;
    emu_f32x16 d = c;
    for (int i = 0; i < 16; ++i) {
        const int r = 8 * (i / 4) + 4 * (l / 32) + (i % 4);
        float s = 0.0f;
        for (int k = 0; k < 16; ++k) s += emu::mfa[w + r + 32 * (k / 8)][k % 8] * emu::mfb[w + j + 32 * (k / 8)][k % 8];
        d[i] = c[i] + s;
    }
    emu::sync();
    return d;
}
inline emu_f32x4 __builtin_amdgcn_mfma_f32_16x16x32_bf16(emu_bf16x8 a, emu_bf16x8 b, emu_f32x4 c, int, int, int)
{
    emu::mf_publish(a, b);
    const int tid = (int)threadIdx.x, l = tid & 63, w = tid & ~63, j = l % 16;
    emu_f32x4 d = c;
    for (int i = 0; i < 4; ++i) {
        const int r = 4 * (l / 16) + i;
        float s = 0.0f;
        for (int k = 0; k < 32; ++k) s += emu::mfa[w + r + 16 * (k / 8)][k % 8] * emu::mfb[w + j + 16 * (k / 8)][k % 8];
        d[i] = c[i] + s;
    }
    emu::sync();
    return d;
}
// Direct global->LDS load (size bytes per lane): lane l's bytes land at
// lds + size * l, visible once the issuing wave waits and the block syncs.
template <typename P> inline void __builtin_amdgcn_global_load_lds(const void *src, P *lds, int size, int, int)
{
    std::memcpy((char *)lds + (threadIdx.x & 63) * size, src, size);
}
inline void __builtin_amdgcn_s_waitcnt(int) {}
inline void __builtin_amdgcn_sched_barrier(int) {}
inline void __builtin_amdgcn_sched_group_barrier(int, int, int) {}
inline int atomicAdd(int *p, int v)
{
    const int old = *p;   // fibers of one host thread: no concurrent access
    *p = old + v;
    return old;
}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
