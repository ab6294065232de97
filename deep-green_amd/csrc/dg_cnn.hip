// dg_cnn.hip -- kernels of the learned conservative flux correction.
//
// The network is a 3D CNN 5 -> 32 -> 32 -> 15 (3x3x3 kernels, GELU after the
// first two layers): from the primitive state of a grid it predicts, for
// each cell c and axis d, a correction to the five flux components on the
// face c + e_d/2.  Convolutions run on MFMA (bf16 operands, fp32
// accumulation) as implicit GEMMs:
//
//   forward / input-gradient  D[cout][cell] = sum_k W[k][cout] * X[cell, k],
//                             k = tap * CIN + cin, tap = (dx*3 + dy)*3 + dz;
//                             A = weights (pre-packed in MFMA fragment order,
//                             held in VGPRs), B = activations (channels-last)
//                             read from a ring of input planes in LDS, 16 B
//                             per lane per k-step (conv_ring_kernel);
//   weight gradient           D[cout][k] = sum_cells dZ[cell][cout] * X[cell + tap][cin],
//                             both operands staged channel-major in LDS per
//                             brick of cells; one partial per workgroup,
//                             reduced in a fixed order.
//
// MFMA lane layouts (gfx950):
//   v_mfma_f32_32x32x16_bf16: A lane l holds row l%32, k = 8*(l/32) + i;
//     B lane l holds column l%32, same k; D lane l holds column l%32,
//     rows 8*(i/4) + 4*(l/32) + i%4 for its 16 values i.
//   v_mfma_f32_16x16x32_bf16: A lane l holds row l%16, k = 8*(l/16) + i;
//     B lane l holds column l%16, same k; D lane l holds column l%16,
//     rows 4*(l/16) + i for its 4 values i.
//
// Geometry: convolutions map an input box of (in0, in1, in2) cells per batch
// item to an output box; "valid" (PAD 0) reads in[o + d], out = in - 2;
// "full" (PAD 2, the input gradient) reads in[o + d - 2] with zeros outside,
// out = in + 2.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dg_cnn.h"

namespace dg {
namespace cnn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------- helpers --
// GELU(z) = z Phi(z) = 0.5 z (1 + erf(z / sqrt 2)), with erf from Abramowitz &
// Stegun 7.1.26 (rational approximation in t = 1 / (1 + p|x|) times exp(-x^2),
// absolute error <= 1.5e-7, far below the bf16 rounding of the activations):
// one reciprocal and one exp2 instead of a library erff.
struct ErfExp {
    float erf, e;   // erf(x), exp(-x^2)
};
__device__ __forceinline__ ErfExp erf_as(float x)
{
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ax, 1.0f));
    const float e = exp2f(-1.44269504088896341f * x * x);
    float p = fmaf(1.061405429f, t, -1.453152027f);
    p = fmaf(p, t, 1.421413741f);
    p = fmaf(p, t, -0.284496736f);
    p = fmaf(p, t, 0.254829592f);
    p *= t;
    return {copysignf(fmaf(-p, e, 1.0f), x), e};
}
__device__ __forceinline__ float gelu(float z)
{
    const float h = 0.5f * z;
    return fmaf(h, erf_as(z * 0.70710678118654752f).erf, h);
}
// GELU'(z) = Phi(z) + z phi(z), phi(z) = exp(-z^2 / 2) / sqrt(2 pi)
__device__ __forceinline__ float gelu_grad(float z)
{
    const ErfExp r = erf_as(z * 0.70710678118654752f);
    return fmaf(0.5f, r.erf, 0.5f) + z * 0.39894228040143268f * r.e;
}

// 32-bit integer hash (xorshift-multiply finalizer)
__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (bf16)0.0f;
    return v;
}

template <int COUT> struct Mfma;
template <> struct Mfma<32> {
    typedef f32x16 acc_t;
    static constexpr int KSTEP = 16, TILE = 32, NACC = 16;
    static __device__ __forceinline__ acc_t op(bf16x8 a, bf16x8 b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    // output channel of accumulator i on lane l
    static __device__ __forceinline__ int row(int i, int lane) { return 8 * (i / 4) + 4 * (lane / 32) + (i % 4); }
};
template <> struct Mfma<16> {
    typedef f32x4 acc_t;
    static constexpr int KSTEP = 32, TILE = 16, NACC = 4;
    static __device__ __forceinline__ acc_t op(bf16x8 a, bf16x8 b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int i, int lane) { return 4 * (lane / 16) + i; }
};

template <int CIN, int COUT> struct ConvCfg {
    static constexpr int KSTEP = Mfma<COUT>::KSTEP;
    static constexpr int KS = (27 * CIN + KSTEP - 1) / KSTEP;   // k-steps
    static constexpr int TILE = Mfma<COUT>::TILE;               // cells per MFMA
    static constexpr int NT = 64 / TILE;                        // cell tiles per wave
    static constexpr int WAVE_CELLS = 64;
};

// ---------------------------------------------------------- conv (MFMA) --
// EPI_TRAIN: Z = acc + b (fp32 out), X = bf16(gelu(Z));  EPI_INFER: X only;
// EPI_OUT: Y = acc + b (fp32, COUT channels);  EPI_BWD: X = bf16(acc * gelu'(Zprev));
// EPI_OUT16: X = bf16(acc + b) (the output layer for the learned step's correction).
enum { EPI_TRAIN = 0, EPI_INFER = 1, EPI_OUT = 2, EPI_BWD = 3, EPI_OUT16 = 4 };

// ------------------------------------------------ conv (MFMA, LDS ring) --
// The same convolutions with the activations staged in LDS.  A workgroup
// (4 waves, one per SIMD) owns a rectangle of RY x RZ output (y, z)
// positions, flattened row-major (at most 256: 64 per wave, NT MFMA tiles of
// TILE positions), and marches it along x over a chunk of output planes.
// Output plane x reads the input planes x - PAD + {0, 1, 2}; they sit in a
// ring of 4 LDS slots, and the plane needed next is copied into the fourth
// slot while the current plane is computed (one barrier per plane; its loads
// are issued before the plane's MFMAs and written to LDS after them).  A slot
// holds (RY + 2) rows of RING_ROWP cells; a cell
// is CIN bf16 channels in 16-B chunks, padded by one chunk for CIN >= 16 so
// that the 128-bit reads of consecutive cells hit distinct banks.  Cells
// outside the input box load zeros (the full convolution's padding).  The
// weights stay in VGPRs for the whole kernel, so LDS carries only the
// activations: one 1-KB read per MFMA.
constexpr int RING_WAVES = 4;
constexpr int RING_POS = 64 * RING_WAVES;   // output positions per workgroup plane
constexpr int RING_ROWP = 48;               // cells per LDS row: RZ + 2 <= RING_ROWP
constexpr int RING_RMAX = 9;                // LDS rows per slot: RY + 2 <= RING_RMAX
constexpr int RING_SLOTS = 4;
constexpr int RING_DEPTH = 3;               // k-steps of B operands loaded ahead

struct RingGeom {
    ConvGeom g;
    int ry, rz;        // output rectangle of a workgroup: ry * rz <= RING_POS
    int ny, nz;        // rectangles along y, z
    int cx, nxc;       // output planes per work item, chunks along x
    long long items;   // B * nxc * ny * nz
};

template <int CIN> struct RingCell {
    static constexpr int NC = CIN / 8;                   // 16-B chunks of channels per cell
    static constexpr int CS = NC == 1 ? 1 : NC + 1;      // chunk stride of a cell
    static constexpr int CHUNKS = RING_RMAX * RING_ROWP * CS;
    static constexpr int BLOCKS = (CHUNKS + 63) / 64;    // 1-KB blocks (one load instruction per wave)
    static constexpr int SLOT = BLOCKS * 1024;           // bytes per slot
};

// zeros for the cells outside the input box (never written)
static __device__ __attribute__((aligned(16))) uint32_t ring_zero[4];

// Filling a slot with input plane xi of batch item b (rows y0 .. y0 + rows - 1,
// cells z0 .. z0 + cells - 1): each thread loads its share of 16-B chunks into
// registers (fill_load, at the top of a plane) and writes them to LDS
// (fill_store, after the plane's MFMAs), so the loads' latency hides under
// the MFMAs and the LDS reads keep exact wait counts.
template <int CIN> struct RingFill {
    static constexpr int NC = CIN / 8;
    static constexpr int PER = (RING_RMAX * RING_ROWP * NC + 64 * RING_WAVES - 1) / (64 * RING_WAVES);
};
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int CIN>
__device__ __forceinline__ void fill_load(u32x4 (&v)[RingFill<CIN>::PER], const bf16 *__restrict__ X,
                                          const ConvGeom &g, int b, int xi, int y0, int z0, int rows, int cells,
                                          int tid)
{
    typedef RingFill<CIN> F;
    const bool xin = xi >= 0 && xi < g.in[0];
    const long long pbase = ((long long)b * g.in[0] + xi) * g.in[1];
#pragma unroll
    for (int k = 0; k < F::PER; ++k) {
        const int idx = k * 64 * RING_WAVES + tid;
        const int cell = idx / F::NC, q = idx % F::NC;
        const int r = cell / RING_ROWP, c = cell % RING_ROWP;
        const int yi = y0 + r, zi = z0 + c;
        const bool in = xin && r < rows && c < cells && yi >= 0 && yi < g.in[1] && zi >= 0 && zi < g.in[2];
        const u32x4 *src = in ? reinterpret_cast<const u32x4 *>(X + ((pbase + yi) * g.in[2] + zi) * CIN + q * 8)
                              : reinterpret_cast<const u32x4 *>(ring_zero);
        v[k] = *src;
    }
}

template <int CIN>
__device__ __forceinline__ void fill_store(char *slot, const u32x4 (&v)[RingFill<CIN>::PER], int rows, int cells,
                                           int tid)
{
    typedef RingFill<CIN> F;
#pragma unroll
    for (int k = 0; k < F::PER; ++k) {
        const int idx = k * 64 * RING_WAVES + tid;
        const int cell = idx / F::NC, q = idx % F::NC;
        const int r = cell / RING_ROWP, c = cell % RING_ROWP;
        if (r < rows && c < cells)
            *reinterpret_cast<u32x4 *>(slot + (cell * RingCell<CIN>::CS + q) * 16) = v[k];
    }
}

template <int CIN, int COUT, int EPI, int PAD>
__global__ void __launch_bounds__(64 * RING_WAVES) __attribute__((amdgpu_waves_per_eu(1, 1)))
conv_ring_kernel(const RingGeom rg, const bf16 *__restrict__ X, const bf16x8 *__restrict__ Wp,
                 const float *__restrict__ bias, float *__restrict__ Zout, bf16 *__restrict__ Xout,
                 float *__restrict__ Yout, const float *__restrict__ Zprev)
{
    typedef ConvCfg<CIN, COUT> C;
    typedef Mfma<COUT> M;
    typedef RingCell<CIN> RC;
    constexpr int NT = C::NT;
    // a k-step lies inside one tap unless CIN < KSTEP (then each lane group has its own tap)
    constexpr bool ALIGNED = CIN % C::KSTEP == 0;
    static_assert(ALIGNED || (CIN == 8 && C::KSTEP == 16), "k-step layout");
    __shared__ __attribute__((aligned(1024))) char ring[RING_SLOTS * RC::SLOT];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane % C::TILE, part = lane / C::TILE;
    const ConvGeom &g = rg.g;

    bf16x8 wa[C::KS];
#pragma unroll
    for (int s = 0; s < C::KS; ++s) wa[s] = Wp[s * 64 + lane];
    // this lane's output channels' biases (no memory wait in the epilogue)
    float bv[M::NACC];
#pragma unroll
    for (int gq = 0; gq < M::NACC / 4; ++gq)
#pragma unroll
        for (int q = 0; q < 4; ++q) bv[4 * gq + q] = EPI == EPI_BWD ? 0.0f : bias[M::row(4 * gq, lane) + q];

    const int RY = rg.ry, RZ = rg.rz, npos = RY * RZ;
    const int o0 = g.out[0], o1 = g.out[1], o2 = g.out[2];
    for (long long item = blockIdx.x; item < rg.items; item += gridDim.x) {
        long long r = item;
        const int iz = (int)(r % rg.nz); r /= rg.nz;
        const int iy = (int)(r % rg.ny); r /= rg.ny;
        const int xc = (int)(r % rg.nxc);
        const int b = (int)(r / rg.nxc);
        const int y0 = iy * RY, z0 = iz * RZ;
        const int xs = xc * rg.cx, xe = min(xs + rg.cx, o0);

        // this lane's positions: LDS byte offset within a slot, output (y, z) index
        int lb[NT], oyz[NT];
        bool ok[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int p = (wave * NT + t) * C::TILE + col;
            const int pp = p < npos ? p : 0;
            const int py = pp / RZ, pz = pp - py * RZ;
            ok[t] = p < npos && y0 + py < o1 && z0 + pz < o2;
            lb[t] = (py * RING_ROWP + pz) * RC::CS * 16 + (ALIGNED ? part * 16 : 0);
            oyz[t] = (y0 + py) * o2 + z0 + pz;
        }

        // prologue: the first three input planes
        u32x4 fv[RingFill<CIN>::PER];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            fill_load<CIN>(fv, X, g, b, xs + j - PAD, y0 - PAD, z0 - PAD, RY + 2, RZ + 2, threadIdx.x);
            fill_store<CIN>(ring + ((xs + j) & 3) * RC::SLOT, fv, RY + 2, RZ + 2, threadIdx.x);
        }
        __syncthreads();

        // The epilogue of plane x runs in slices between the k-steps of plane
        // x + 1 (no store waits at the plane's end); the last plane's runs
        // after the loop.
        constexpr int NEPI = NT * (M::NACC / 4);
        auto epilogue = [&](const typename M::acc_t *A, long long mb, int e) {
            const int t = e / (M::NACC / 4), gq = e % (M::NACC / 4);
            if (!ok[t]) return;
            const long long m = mb + oyz[t];
            const int ch = M::row(4 * gq, lane);
            float v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = A[t][4 * gq + q];
            if (EPI == EPI_BWD) {
                const f32x4 zp = *reinterpret_cast<const f32x4 *>(Zprev + m * COUT + ch);
#pragma unroll
                for (int q = 0; q < 4; ++q) v[q] *= gelu_grad(zp[q]);
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) v[q] += bv[4 * gq + q];
            }
            if (EPI == EPI_TRAIN) {
                f32x4 z;
#pragma unroll
                for (int q = 0; q < 4; ++q) z[q] = v[q];
                *reinterpret_cast<f32x4 *>(Zout + m * COUT + ch) = z;
            }
            if (EPI == EPI_OUT) {
                f32x4 y;
#pragma unroll
                for (int q = 0; q < 4; ++q) y[q] = v[q];
                *reinterpret_cast<f32x4 *>(Yout + m * COUT + ch) = y;
            } else {
                bf16x4 xo;
#pragma unroll
                for (int q = 0; q < 4; ++q) xo[q] = (bf16)(EPI == EPI_BWD || EPI == EPI_OUT16 ? v[q] : gelu(v[q]));
                *reinterpret_cast<bf16x4 *>(Xout + m * COUT + ch) = xo;
            }
        };
        const long long plane_cells = (long long)o1 * o2;
        typename M::acc_t prv[NT];
        long long pmb = -1;   // output index of (plane of prv, y = z = 0); -1: none pending

        for (int x = xs; x < xe; ++x) {
            const bool more = x + 1 < xe;
            if (more) fill_load<CIN>(fv, X, g, b, x + 3 - PAD, y0 - PAD, z0 - PAD, RY + 2, RZ + 2, threadIdx.x);
            const char *sb[3];
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) sb[dx] = ring + ((x + dx) & 3) * RC::SLOT;

            typename M::acc_t acc[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int i = 0; i < M::NACC; ++i) acc[t][i] = 0.0f;

            // B operand of k-step s for tile t (8 channels of one tap of this lane's position)
            auto load_b = [&](int s, int t) -> bf16x8 {
                if (ALIGNED) {
                    const int k = s * C::KSTEP, tap = k / CIN, q = (k % CIN) / 8;
                    const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;
                    const int off = ((dy * RING_ROWP + dz) * RC::CS + q) * 16;
                    return *reinterpret_cast<const bf16x8 *>(sb[dx] + lb[t] + off);
                } else {
                    // CIN 8, KSTEP 16: lanes of part 0 read tap 2s, part 1 tap 2s + 1 (zero past tap 26)
                    const int ta = 2 * s, tb = 2 * s + 1;
                    const int offa = ((ta / 3) % 3 * RING_ROWP + ta % 3) * 16;
                    const int offb = ((tb / 3) % 3 * RING_ROWP + tb % 3) * 16;
                    const char *pa = sb[ta / 9] + lb[t] + offa;
                    if (tb >= 27) return part == 0 ? *reinterpret_cast<const bf16x8 *>(pa) : zero8();
                    const char *pb = sb[tb / 9] + lb[t] + offb;
                    return *reinterpret_cast<const bf16x8 *>(part == 0 ? pa : pb);
                }
            };
            // B operands of RING_DEPTH k-steps in flight (LDS latency under the MFMAs);
            // the scheduling barriers keep the issue order as written
            bf16x8 bq[RING_DEPTH][NT];
#pragma unroll
            for (int d = 0; d < RING_DEPTH; ++d)
                if (d < C::KS)
#pragma unroll
                    for (int t = 0; t < NT; ++t) bq[d][t] = load_b(d, t);
            // NEPI segments of SEG k-steps, each followed by one epilogue slice
            constexpr int SEG = (C::KS + NEPI - 1) / NEPI;
#pragma unroll
            for (int e = 0; e < NEPI; ++e) {
#pragma unroll
                for (int j = 0; j < SEG; ++j) {
                    const int s = e * SEG + j;
                    if (s < C::KS) {
#pragma unroll
                        for (int t = 0; t < NT; ++t) acc[t] = M::op(wa[s], bq[s % RING_DEPTH][t], acc[t]);
                        if (s + RING_DEPTH < C::KS)
#pragma unroll
                            for (int t = 0; t < NT; ++t) bq[s % RING_DEPTH][t] = load_b(s + RING_DEPTH, t);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
                if (pmb >= 0) epilogue(prv, pmb, e);
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) prv[t] = acc[t];
            pmb = ((long long)b * o0 + x) * plane_cells;
            // the slot of plane x - 1 is free since the last barrier
            if (more) fill_store<CIN>(ring + ((x + 3) & 3) * RC::SLOT, fv, RY + 2, RZ + 2, threadIdx.x);
            __syncthreads();
        }
        if (pmb >= 0)
#pragma unroll
            for (int e = 0; e < NEPI; ++e) epilogue(prv, pmb, e);
    }
}

// ----------------------------------------------------- weight gradient --
// dW[k][cout] for k = tap*CIN + cin, over the cells of a valid convolution:
// sum_cells dZ[cell][cout] * X[cell + tap][cin].  Bricks of 2 x 4 x 32 output
// cells; LDS holds the input brick (4 x 6 x 34) and the dZ brick channel-major.
constexpr int DW_BX = 2, DW_BY = 4, DW_BZ = 32;
constexpr int DW_IX = DW_BX + 2, DW_IY = DW_BY + 2, DW_IZ = DW_BZ + 2;

template <int CIN, int COUT>
__global__ void __launch_bounds__(512)
dw_kernel(const ConvGeom g, const bf16 *__restrict__ X, const bf16 *__restrict__ dZ, float *__restrict__ partials)
{
    constexpr int NCOL = 27 * CIN;
    constexpr int NTILES = (NCOL + 31) / 32;
    constexpr int TPW = (NTILES + 7) / 8;
    __shared__ __attribute__((aligned(16))) bf16 xs[CIN][DW_IX][DW_IY][DW_IZ];
    __shared__ __attribute__((aligned(16))) bf16 zs[32][DW_BX][DW_BY][DW_BZ];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, h = lane / 32;
    const int nbx = (g.out[0] + DW_BX - 1) / DW_BX, nby = (g.out[1] + DW_BY - 1) / DW_BY,
              nbz = (g.out[2] + DW_BZ - 1) / DW_BZ;
    const long long nbricks = (long long)g.B * nbx * nby * nbz;

    // columns of this wave's tiles
    int ctap[TPW], ccin[TPW];
    bool cok[TPW];
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
        const int n = (wave + 8 * q) * 32 + (lane % 32);
        cok[q] = (wave + 8 * q) < NTILES && n < NCOL;
        ctap[q] = cok[q] ? n / CIN : 0;
        ccin[q] = cok[q] ? n % CIN : 0;
    }
    f32x16 acc[TPW];
#pragma unroll
    for (int q = 0; q < TPW; ++q)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[q][i] = 0.0f;

    for (long long br = blockIdx.x; br < nbricks; br += gridDim.x) {
        long long r = br;
        const int bz = (int)(r % nbz); r /= nbz;
        const int by = (int)(r % nby); r /= nby;
        const int bx = (int)(r % nbx); r /= nbx;
        const int b = (int)r;
        const int x0 = bx * DW_BX, y0 = by * DW_BY, z0 = bz * DW_BZ;
        // stage X (input brick) and dZ (output brick), channel-major, zeros outside
        constexpr int XCH = CIN / 8;
        for (int e = tid; e < DW_IX * DW_IY * DW_IZ * XCH; e += 512) {
            const int c8 = e % XCH, cell = e / XCH;
            const int iz = cell % DW_IZ, iy = (cell / DW_IZ) % DW_IY, ix = cell / (DW_IZ * DW_IY);
            const int gx = x0 + ix, gy = y0 + iy, gz = z0 + iz;
            bf16x8 v = zero8();
            if (gx < g.in[0] && gy < g.in[1] && gz < g.in[2])
                v = *reinterpret_cast<const bf16x8 *>(
                    X + ((((long long)b * g.in[0] + gx) * g.in[1] + gy) * g.in[2] + gz) * CIN + c8 * 8);
#pragma unroll
            for (int i = 0; i < 8; ++i) xs[c8 * 8 + i][ix][iy][iz] = v[i];
        }
        constexpr int ZCH = COUT / 8;
        for (int e = tid; e < DW_BX * DW_BY * DW_BZ * 4; e += 512) {
            const int c8 = e % 4, cell = e / 4;
            const int iz = cell % DW_BZ, iy = (cell / DW_BZ) % DW_BY, ix = cell / (DW_BZ * DW_BY);
            const int gx = x0 + ix, gy = y0 + iy, gz = z0 + iz;
            bf16x8 v = zero8();
            if (c8 < ZCH && gx < g.out[0] && gy < g.out[1] && gz < g.out[2])
                v = *reinterpret_cast<const bf16x8 *>(
                    dZ + ((((long long)b * g.out[0] + gx) * g.out[1] + gy) * g.out[2] + gz) * COUT + c8 * 8);
#pragma unroll
            for (int i = 0; i < 8; ++i) zs[c8 * 8 + i][ix][iy][iz] = v[i];
        }
        __syncthreads();
        for (int x = 0; x < DW_BX; ++x)
            for (int y = 0; y < DW_BY; ++y)
                for (int zk = 0; zk < DW_BZ; zk += 16) {
                    // A = dZ^T: row cout = lane%32, k = cells z = zk + 8h + i
                    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&zs[lane % 32][x][y][zk + 8 * h]);
#pragma unroll
                    for (int q = 0; q < TPW; ++q) {
                        const int tap = ctap[q], cin = ccin[q];
                        const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;
                        bf16x8 bv;
#pragma unroll
                        for (int i = 0; i < 8; ++i) bv[i] = xs[cin][x + dx][y + dy][zk + 8 * h + i + dz];
                        if (!cok[q]) bv = zero8();
                        acc[q] = Mfma<32>::op(a, bv, acc[q]);
                    }
                }
        __syncthreads();
    }
    // partial of this workgroup: [NTILES*32 columns][32 couts]
    float *P = partials + (size_t)blockIdx.x * NTILES * 32 * 32;
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
        const int T = wave + 8 * q;
        if (T >= NTILES) continue;
        const int n = T * 32 + lane % 32;
#pragma unroll
        for (int i = 0; i < 16; ++i) P[(size_t)n * 32 + Mfma<32>::row(i, lane)] = acc[q][i];
    }
}

// grad[tap][cin_real][cout_real] = sum over workgroups of the partials:
// one block per (tap, cin); 8 groups of 32 lanes (one per output channel)
// each sum every 8th workgroup in order, then the 8 group sums are added in
// order (a fixed order: deterministic).
__global__ void __launch_bounds__(256)
dw_reduce_kernel(const float *__restrict__ partials, int nwg, int cin_pad, int cin_real, int cout_real,
                 float *__restrict__ grad)
{
    __shared__ double red[8][32];
    const int ncol = 27 * cin_pad;
    const int ntiles = (ncol + 31) / 32;
    const int tap = blockIdx.x / cin_real, ci = blockIdx.x % cin_real;
    const int n = tap * cin_pad + ci;
    const int co = threadIdx.x & 31, grp = threadIdx.x >> 5;
    double s = 0.0;
    for (int w = grp; w < nwg; w += 8) s += (double)partials[((size_t)w * ntiles * 32 + n) * 32 + co];
    red[grp][co] = s;
    __syncthreads();
    if (grp == 0 && co < cout_real) {
        double t = 0.0;
        for (int q = 0; q < 8; ++q) t += red[q][co];
        grad[(size_t)blockIdx.x * cout_real + co] = (float)t;
    }
}

// bias gradient: db[c] = sum over cells of dZ[cell][c] (partials per block)
__global__ void __launch_bounds__(256)
db_kernel(const bf16 *__restrict__ dZ, long long ncells, int cpad, float *__restrict__ partials)
{
    __shared__ float red[256];
    const int c = threadIdx.x % 32, r = threadIdx.x / 32;
    float s = 0.0f;
    if (c < cpad)
        for (long long m = (long long)blockIdx.x * 8 + r; m < ncells; m += (long long)gridDim.x * 8)
            s += (float)dZ[m * cpad + c];
    red[threadIdx.x] = s;
    __syncthreads();
    if (r == 0) {
        for (int q = 1; q < 8; ++q) s += red[q * 32 + c];
        partials[blockIdx.x * 32 + c] = s;
    }
}

// grad[c] = sum over blocks of the partials: 8 groups of 32 channels each sum
// every 8th block in order, then the group sums are added in order.
__global__ void __launch_bounds__(256)
db_reduce_kernel(const float *__restrict__ partials, int nblk, int creal, float *__restrict__ grad)
{
    __shared__ double red[8][32];
    const int c = threadIdx.x & 31, grp = threadIdx.x >> 5;
    double s = 0.0;
    for (int b = grp; b < nblk; b += 8) s += (double)partials[b * 32 + c];
    red[grp][c] = s;
    __syncthreads();
    if (grp == 0 && c < creal) {
        double t = 0.0;
        for (int q = 0; q < 8; ++q) t += red[q][c];
        grad[c] = (float)t;
    }
}

// ---------------------------------------------------- weights, Adam --
// Fragment-order bf16 weights of a forward layer (transpose = 0) or of its
// input-gradient convolution (transpose = 1: taps flipped, cin <-> cout).
__global__ void pack_kernel(const float *__restrict__ W, int cin_real, int cout_real, int cin_pad, int cout_pad,
                            int transpose, bf16 *__restrict__ Wp)
{
    // as a convolution: CI input channels (padded CIP), CO output channels (padded COP)
    const int CIP = transpose ? cout_pad : cin_pad, COP = transpose ? cin_pad : cout_pad;
    const int kstep = COP == 32 ? 16 : 32, tile = COP == 32 ? 32 : 16;
    const int ks = (27 * CIP + kstep - 1) / kstep;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ks * 64 * 8) return;
    const int i = e % 8, lane = (e / 8) % 64, s = e / 512;
    const int k = s * kstep + (lane / tile) * 8 + i, n = lane % tile;
    const int tap = k / CIP, c = k % CIP;
    float v = 0.0f;
    if (tap < 27) {
        if (!transpose) {
            if (c < cin_real && n < cout_real) v = W[((size_t)tap * cin_real + c) * cout_real + n];
        } else {
            // input gradient: out channel n = forward cin, in channel c = forward cout, tap flipped
            if (n < cin_real && c < cout_real) v = W[((size_t)(26 - tap) * cin_real + n) * cout_real + c];
        }
    }
    Wp[e] = (bf16)v;
}

__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ gr, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1,
                            float bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    w[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ----------------------------------------------- inputs, loss, update --
// Normalized primitives (bf16, 8 channels, 5 used) of a periodic n0 x n1 x n2
// SoA conserved state, for the box [o - halo, o - halo + dims) (wrapped),
// one box per batch item: boxes[b] = (sample, o0, o1, o2).
__global__ void __launch_bounds__(256)
prep_input_kernel(const PrepArgs a, const float *const *__restrict__ samples, const int *__restrict__ boxes,
                  bf16 *__restrict__ X, int *__restrict__ zero)
{
    if (zero && blockIdx.x == 0 && threadIdx.x == 0) *zero = 0;
    const long long per = (long long)a.dims[0] * a.dims[1] * a.dims[2];
    const long long total = per * a.B;
    const float V = a.ds * a.ds * a.ds, gm1 = a.gamma - 1.0f;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e / per);
        long long r = e % per;
        const int z = (int)(r % a.dims[2]); r /= a.dims[2];
        const int y = (int)(r % a.dims[1]);
        const int x = (int)(r / a.dims[1]);
        const int *bx = boxes + 4 * b;
        const float *const *U = samples + 5 * bx[0];
        const int gx = ((bx[1] + x - a.halo) % a.n[0] + a.n[0]) % a.n[0];
        const int gy = ((bx[2] + y - a.halo) % a.n[1] + a.n[1]) % a.n[1];
        const int gz = ((bx[3] + z - a.halo) % a.n[2] + a.n[2]) % a.n[2];
        const long long c = ((long long)gx * a.n[1] + gy) * a.n[2] + gz;
        const float rho = U[0][c] / V;
        const float vx = U[1][c] / (rho * V), vy = U[2][c] / (rho * V), vz = U[3][c] / (rho * V);
        const float p = (U[4][c] / V - 0.5f * rho * (vx * vx + vy * vy + vz * vz)) * gm1;
        const float q[5] = {rho, vx, vy, vz, p};
        bf16x8 out = zero8();
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            float v = (q[f] - a.shift[f]) * a.inv_scale[f];
            if (a.noise > 0.0f) {
                // counter-based normal variate (Box-Muller on two hashed uniforms)
                const uint32_t k = (uint32_t)(e * 5 + f);
                const uint32_t h1 = mix32(k ^ a.seed), h2 = mix32(k ^ (a.seed * 0x9E3779B9u + 0x7F4A7C15u));
                const float u1 = ((h1 >> 8) + 1u) * (1.0f / 16777217.0f), u2 = (h2 >> 8) * (1.0f / 16777216.0f);
                v += a.noise * sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
            }
            out[f] = (bf16)v;
        }
        *reinterpret_cast<bf16x8 *>(X + e * 8) = out;
    }
}

// Targets of a batch: R / out_scale over the loss box (wrapped), fp32 [B][L^3][5].
__global__ void __launch_bounds__(256)
prep_target_kernel(const PrepArgs a, const float *const *__restrict__ samples, const int *__restrict__ boxes,
                   float *__restrict__ T)
{
    const int L = a.dims[0];
    const long long per = (long long)L * L * L, total = per * a.B;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e / per);
        long long r = e % per;
        const int z = (int)(r % L); r /= L;
        const int y = (int)(r % L);
        const int x = (int)(r / L);
        const int *bx = boxes + 4 * b;
        const float *const *R = samples + 5 * bx[0];
        const int gx = (bx[1] + x) % a.n[0], gy = (bx[2] + y) % a.n[1], gz = (bx[3] + z) % a.n[2];
        const long long c = ((long long)gx * a.n[1] + gy) * a.n[2] + gz;
#pragma unroll
        for (int f = 0; f < 5; ++f) T[e * 5 + f] = R[f][c] * a.inv_out[f];
    }
}

// Divergence loss over the L^3 box: P[c][q] = -sum_d (G[c+1][5d+q] - G[c+1-e_d][5d+q])
// (G on (L+1)^3, index c+1 = face c + e_d/2), loss = mean (P - T)^2 over cells and
// the 5 components; dG = dloss/dG (bf16, 16 channels, channel 15 zero); per-block
// partial loss sums.
__global__ void __launch_bounds__(256)
loss_kernel(int B, int L, const float *__restrict__ G, const float *__restrict__ T, bf16 *__restrict__ dG,
            double *__restrict__ loss_part)
{
    __shared__ double red[256];
    const int Lg = L + 1;
    const long long perg = (long long)Lg * Lg * Lg, total = perg * B;
    const double N = (double)B * L * L * L * 5;
    const float scale = (float)(2.0 / N);
    double lsum = 0.0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e / perg);
        long long r = e % perg;
        const int gz = (int)(r % Lg); r /= Lg;
        const int gy = (int)(r % Lg);
        const int gx = (int)(r / Lg);
        // error of cell c = g - 1 (as the +G[c+1] term) and residual helpers
        auto err = [&](int cx, int cy, int cz, int q) -> float {
            if (cx < 0 || cy < 0 || cz < 0 || cx >= L || cy >= L || cz >= L) return 0.0f;
            const long long gb = (long long)b * perg;
            const long long g0 = gb + (((long long)(cx + 1) * Lg + (cy + 1)) * Lg + (cz + 1));
            const long long gxm = g0 - (long long)Lg * Lg, gym = g0 - Lg, gzm = g0 - 1;
            const float P = -((G[g0 * 16 + q] - G[gxm * 16 + q]) + (G[g0 * 16 + 5 + q] - G[gym * 16 + 5 + q]) +
                              (G[g0 * 16 + 10 + q] - G[gzm * 16 + 10 + q]));
            const long long t = (((long long)b * L + cx) * L + cy) * L + cz;
            return P - T[t * 5 + q];
        };
        bf16x8 lo = zero8(), hi = zero8();
        const int cx = gx - 1, cy = gy - 1, cz = gz - 1;
        for (int q = 0; q < 5; ++q) {
            const float e0 = err(cx, cy, cz, q);
            if (cx >= 0 && cy >= 0 && cz >= 0) lsum += (double)e0 * e0;
            // d/dG[g][5d+q]: -e(c) from cell c = g-1, + e(c + e_d) from cell g-1+e_d
            const float gx_ = scale * (-e0 + err(cx + 1, cy, cz, q));
            const float gy_ = scale * (-e0 + err(cx, cy + 1, cz, q));
            const float gz_ = scale * (-e0 + err(cx, cy, cz + 1, q));
            const float v[3] = {gx_, gy_, gz_};
            for (int d = 0; d < 3; ++d) {
                const int ch = 5 * d + q;
                if (ch < 8) lo[ch] = (bf16)v[d];
                else hi[ch - 8] = (bf16)v[d];
            }
        }
        *reinterpret_cast<bf16x8 *>(dG + e * 16) = lo;
        *reinterpret_cast<bf16x8 *>(dG + e * 16 + 8) = hi;
    }
    red[threadIdx.x] = lsum;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] / N;
}

// Host-array staging: fp32 [cells][creal] -> bf16 [cells][cpad]
__global__ void to_bf16_kernel(const float *__restrict__ x, long long ncells, int creal, int cpad, bf16 *__restrict__ y)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ncells * cpad) return;
    const long long m = e / cpad;
    const int c = (int)(e % cpad);
    y[e] = (bf16)(c < creal ? x[m * creal + c] : 0.0f);
}

// ------------------------------------------------------------ launches --
static ConvGeom geom(int B, const int in[3], int pad)
{
    ConvGeom g;
    g.B = B;
    for (int d = 0; d < 3; ++d) {
        g.in[d] = in[d];
        g.out[d] = pad ? in[d] + 2 : in[d] - 2;
    }
    g.ncells = (long long)B * g.out[0] * g.out[1] * g.out[2];
    return g;
}

static int grid_for(long long n, long long per, int cap)
{
    long long b = (n + per - 1) / per;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

static int cu_count()
{
    static int n = 0;
    if (!n) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n = v;
        else
            n = 256;
    }
    return n;
}

// Rectangles of at most RING_POS positions with RZ + 2 <= RING_ROWP and
// RY + 2 <= RING_RMAX, balanced over the output extents; x chunks so that
// there are about four work items per CU.
static RingGeom ring_geom(const ConvGeom &g)
{
    RingGeom r;
    r.g = g;
    const int o0 = g.out[0], o1 = g.out[1], o2 = g.out[2];
    r.nz = (o2 + RING_ROWP - 3) / (RING_ROWP - 2);
    r.rz = (o2 + r.nz - 1) / r.nz;
    int ry = RING_POS / r.rz;
    if (ry > RING_RMAX - 2) ry = RING_RMAX - 2;
    if (ry > o1) ry = o1;
    if (ry < 1) ry = 1;
    r.ny = (o1 + ry - 1) / ry;
    r.ry = (o1 + r.ny - 1) / r.ny;
    const long long cols = (long long)g.B * r.ny * r.nz;
    long long nxc = (4LL * cu_count() + cols - 1) / cols;
    if (nxc > o0) nxc = o0;
    if (nxc < 1) nxc = 1;
    r.cx = (int)((o0 + nxc - 1) / nxc);
    r.nxc = (o0 + r.cx - 1) / r.cx;
    r.items = cols * r.nxc;
    return r;
}

int conv(int layer, int mode, int B, const int in[3], const void *X, const void *Wp, const float *bias, float *Z,
         void *Xout, float *Y, const float *Zprev, hipStream_t s)
{
    const int pad = mode == CONV_BWD ? 2 : 0;
    const ConvGeom g = geom(B, in, pad);
    for (int d = 0; d < 3; ++d)
        if (g.out[d] < 1) return -1;
    const RingGeom rg = ring_geom(g);
    const int blocks = (int)(rg.items < cu_count() ? rg.items : cu_count());
    const bf16 *x = static_cast<const bf16 *>(X);
    const bf16x8 *w = static_cast<const bf16x8 *>(Wp);
    bf16 *xo = static_cast<bf16 *>(Xout);
#define DG_CONV(CI, CO, E, P)                                                                                 \
    hipLaunchKernelGGL((conv_ring_kernel<CI, CO, E, P>), dim3(blocks), dim3(64 * RING_WAVES), 0, s, rg, x, w, bias, Z, \
                       xo, Y, Zprev)
    if (layer == 1) {
        if (mode == CONV_TRAIN) DG_CONV(8, 32, EPI_TRAIN, 0);
        else if (mode == CONV_INFER) DG_CONV(8, 32, EPI_INFER, 0);
        else return -1;
    } else if (layer == 2) {
        if (mode == CONV_TRAIN) DG_CONV(32, 32, EPI_TRAIN, 0);
        else if (mode == CONV_INFER) DG_CONV(32, 32, EPI_INFER, 0);
        else DG_CONV(32, 32, EPI_BWD, 2);   // input gradient of layer 2 (dZ2 -> dZ1)
    } else if (layer == 3) {
        if (mode == CONV_BWD) DG_CONV(16, 32, EPI_BWD, 2);   // input gradient of layer 3 (dG -> dZ2)
        else if (mode == CONV_OUT16) DG_CONV(32, 16, EPI_OUT16, 0);
        else DG_CONV(32, 16, EPI_OUT, 0);
    } else {
        return -1;
    }
#undef DG_CONV
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int dw_tiles(int layer)
{
    const int cin = layer == 1 ? 8 : 32;
    return (27 * cin + 31) / 32;
}

int dw(int layer, int B, const int in[3], const void *X, const void *dZ, float *partials, int nwg, hipStream_t s)
{
    const ConvGeom g = geom(B, in, 0);
    const bf16 *x = static_cast<const bf16 *>(X);
    const bf16 *z = static_cast<const bf16 *>(dZ);
    if (layer == 1) hipLaunchKernelGGL((dw_kernel<8, 32>), dim3(nwg), dim3(512), 0, s, g, x, z, partials);
    else if (layer == 2) hipLaunchKernelGGL((dw_kernel<32, 32>), dim3(nwg), dim3(512), 0, s, g, x, z, partials);
    else hipLaunchKernelGGL((dw_kernel<32, 16>), dim3(nwg), dim3(512), 0, s, g, x, z, partials);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int dw_reduce(int layer, const float *partials, int nwg, float *grad, hipStream_t s)
{
    const int cin_pad = layer == 1 ? 8 : 32, cin_real = layer == 1 ? 5 : 32, cout_real = layer == 3 ? 15 : 32;
    hipLaunchKernelGGL(dw_reduce_kernel, dim3(27 * cin_real), dim3(256), 0, s, partials, nwg, cin_pad, cin_real,
                       cout_real, grad);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int db(const void *dZ, long long ncells, int cpad, int creal, float *partials, int nblk, float *grad, hipStream_t s)
{
    hipLaunchKernelGGL(db_kernel, dim3(nblk), dim3(256), 0, s, static_cast<const bf16 *>(dZ), ncells, cpad, partials);
    hipLaunchKernelGGL(db_reduce_kernel, dim3(1), dim3(256), 0, s, partials, nblk, creal, grad);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int pack(int layer, int transpose, const float *W, void *Wp, hipStream_t s)
{
    const int cin_real = layer == 1 ? 5 : 32, cout_real = layer == 3 ? 15 : 32;
    const int cin_pad = layer == 1 ? 8 : 32, cout_pad = layer == 3 ? 16 : 32;
    const int n = pack_elems(layer, transpose);
    hipLaunchKernelGGL(pack_kernel, dim3((n + 255) / 256), dim3(256), 0, s, W, cin_real, cout_real, cin_pad, cout_pad,
                       transpose, static_cast<bf16 *>(Wp));
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int pack_elems(int layer, int transpose)
{
    const int cin_pad = layer == 1 ? 8 : 32, cout_pad = layer == 3 ? 16 : 32;
    const int CIP = transpose ? cout_pad : cin_pad, COP = transpose ? cin_pad : cout_pad;
    const int kstep = COP == 32 ? 16 : 32;
    return ((27 * CIP + kstep - 1) / kstep) * 64 * 8;
}

int adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps, long long t,
         hipStream_t s)
{
    const float bc1 = 1.0f - powf(b1, (float)t), bc2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int prep_input(const PrepArgs &a, const float *const *samples, const int *boxes, void *X, hipStream_t s, int *zero)
{
    const long long n = (long long)a.B * a.dims[0] * a.dims[1] * a.dims[2];
    hipLaunchKernelGGL(prep_input_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, a, samples, boxes,
                       static_cast<bf16 *>(X), zero);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int prep_target(const PrepArgs &a, const float *const *samples, const int *boxes, float *T, hipStream_t s)
{
    const long long n = (long long)a.B * a.dims[0] * a.dims[0] * a.dims[0];
    hipLaunchKernelGGL(prep_target_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, a, samples, boxes, T);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int loss(int B, int L, const float *G, const float *T, void *dG, double *loss_part, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL(loss_kernel, dim3(nblk), dim3(256), 0, s, B, L, G, T, static_cast<bf16 *>(dG), loss_part);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int to_bf16(const float *x, long long ncells, int creal, int cpad, void *y, hipStream_t s)
{
    const long long n = ncells * cpad;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, ncells, creal, cpad,
                       static_cast<bf16 *>(y));
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace cnn
}  // namespace dg
