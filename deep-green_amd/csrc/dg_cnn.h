// dg_cnn.h -- launchers of the flux-correction network kernels (dg_cnn.hip),
// used by dg_learned.cpp.  Not part of the public interface.
//
// Layers: 1 = 5 (padded 8) -> 32, 2 = 32 -> 32, 3 = 32 -> 15 (padded 16).
// Activations are bf16 channels-last ([batch][x][y][z][channel]); master
// weights fp32 [tap][cin][cout] with tap = (dx*3 + dy)*3 + dz.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dg {
namespace cnn {

struct ConvGeom {
    int B;
    int in[3], out[3];
    long long ncells;  // B * out cells
};

// Input preparation from periodic SoA conserved fp32 states.
struct PrepArgs {
    int B;
    int n[3];        // grid of the sampled states
    int dims[3];     // box per batch item (prep_target: dims[0] = L)
    int halo;        // box origin offset: cell o - halo
    float ds, gamma;
    float shift[5], inv_scale[5];   // primitive normalization
    float inv_out[5];               // target normalization
    float noise;                    // std of Gaussian noise added to the normalized inputs (training)
    unsigned int seed;              // noise stream
};

enum { CONV_TRAIN = 0, CONV_INFER = 1, CONV_OUT = 2, CONV_BWD = 3, CONV_OUT16 = 4 };

// Forward (valid) convolution of `layer` (CONV_TRAIN: Z fp32 + X = gelu(Z) bf16;
// CONV_INFER: X only; layer 3: Y fp32 16 channels, or with CONV_OUT16 X bf16
// 16 channels), or the input gradient
// (CONV_BWD, full convolution with the flipped weights, times gelu'(Zprev)).
int conv(int layer, int mode, int B, const int in[3], const void *X, const void *Wp, const float *bias, float *Z,
         void *Xout, float *Y, const float *Zprev, hipStream_t s);
// Weight gradient of `layer` over a valid convolution with input box `in`:
// one partial per workgroup (nwg workgroups), then the fixed-order reduction.
int dw(int layer, int B, const int in[3], const void *X, const void *dZ, float *partials, int nwg, hipStream_t s);
int dw_tiles(int layer);
int dw_reduce(int layer, const float *partials, int nwg, float *grad, hipStream_t s);
int db(const void *dZ, long long ncells, int cpad, int creal, float *partials, int nblk, float *grad, hipStream_t s);
int pack(int layer, int transpose, const float *W, void *Wp, hipStream_t s);
int pack_elems(int layer, int transpose);
int adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps, long long t,
         hipStream_t s);
// zero (optional): a counter the kernel also sets to 0 (the learned step's
// fallback count, which saves that step a launch).
int prep_input(const PrepArgs &a, const float *const *samples, const int *boxes, void *X, hipStream_t s,
               int *zero = nullptr);
int prep_target(const PrepArgs &a, const float *const *samples, const int *boxes, float *T, hipStream_t s);
int loss(int B, int L, const float *G, const float *T, void *dG, double *loss_part, int nblk, hipStream_t s);
int to_bf16(const float *x, long long ncells, int creal, int cpad, void *y, hipStream_t s);

}  // namespace cnn
}  // namespace dg
