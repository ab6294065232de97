// dg_learned.cpp -- C ABI of the learned conservative flux correction
// (include/deepgreen.h, dg_cnn_* and dg_learned_*).
//
// A dg_cnn holds the network (fp32 master weights, Adam moments, bf16 weights
// packed for the MFMA kernels), the input/output normalization, a set of
// training samples produced by the solver itself, and the workspaces of
// training and inference.  Samples: the fine grid's state restricted 2x to
// the coarse grid (U^n); the coarse plain step with dt; two fine steps of
// dt/2 restricted again (U_target); the residual R = (U_target - U_plain) /
// (dt ds_c^2) is the flux divergence the coarse update misses.  The network
// G (normalized by the per-field scale of R) is trained so that -div G = R.
//
// A learned step on a periodic FP32 coarse grid: network on U^n -> G; the
// plain step kernel -> U_plain; the masked, telescoping correction (cells
// whose corrected state is not positive keep the plain flux on all faces);
// then dt and clock as for a plain step.  No reference counterpart: the
// reference has no learned component.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/deepgreen.h"
#include "dg_cnn.h"
#include "dg_ctx.h"
#include "dg_internal.h"

namespace {

// parameter layout (fp32): W1 [27][5][32], b1 [32], W2 [27][32][32], b2 [32],
// W3 [27][32][15], b3 [15]
const int kCin[3] = {5, 32, 32}, kCout[3] = {32, 32, 15};
const int64_t kW1 = 0, kB1 = kW1 + 27 * 5 * 32, kW2 = kB1 + 32, kB2 = kW2 + 27 * 32 * 32, kW3 = kB2 + 32,
              kB3 = kW3 + 27 * 32 * 15, kNParams = kB3 + 15;
const int64_t kWoff[3] = {kW1, kW2, kW3}, kBoff[3] = {kB1, kB2, kB3};
const int kDwGroups = 256;     // workgroups of the weight-gradient kernels
const int kDbBlocks = 256;     // blocks of the bias-gradient kernels
const int kLossBlocks = 256;
const int kGraphSteps = 16;

struct Buf {
    void *p = nullptr;
    size_t n = 0;
};

}  // namespace

struct dg_cnn {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    float *bias_pad[3] = {nullptr, nullptr, nullptr};   // 32 floats each
    void *Wp[3] = {nullptr, nullptr, nullptr};          // forward, fragment order
    void *Wt[3] = {nullptr, nullptr, nullptr};          // input gradient (layers 2, 3)
    long long adam_t = 0;
    float norm[15];          // shift[5], scale[5], out_scale[5]
    bool norm_set = false;
    float input_noise = 0.0f;  // training-time input noise (normalized units)
    // samples (device, fp32 SoA interior): U^n and R
    int sn[3] = {0, 0, 0};
    double sds = 0, sgamma = 1.4;
    std::vector<float *> sU, sR;          // 5 per sample
    float **d_sU = nullptr, **d_sR = nullptr;
    size_t d_cap = 0;
    float *tmp = nullptr;                 // restriction scratch (5 coarse fields)
    size_t tmp_cells = 0;
    // workspaces
    Buf ws[16];
    float *dw_part = nullptr, *db_part = nullptr;
    double *loss_part = nullptr;
    // inference tables (per buffer parity of the coarse context)
    const float **d_ptrs = nullptr;       // [2][5]
    int *d_box0 = nullptr;
    unsigned char *flags = nullptr;
    size_t flags_n = 0;
    uint64_t graph_uid = 0;   // context the tables and captured steps belong to
    size_t graph_n3 = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
};

namespace {

int mfail(dg_cnn *m, int code, const std::string &msg)
{
    if (m) m->err = msg;
    return code;
}

#define MHIP(m, expr)                                                                                        \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return mfail((m), DG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define MCHK(m, expr, what)                                                \
    do {                                                                   \
        if ((expr) != 0) return mfail((m), DG_ERR_DEVICE, (what));         \
    } while (0)

void drop_graphs(dg_cnn *m);

// workspace slot k of at least n bytes.  Growing a slot frees its old
// buffer, which captured learned steps may reference: they are dropped (and
// captured again by the next learned run).
void *ws(dg_cnn *m, int k, size_t n)
{
    Buf &b = m->ws[k];
    if (b.n >= n && b.p) return b.p;
    if (b.p) {
        (void)hipStreamSynchronize(m->stream);
        drop_graphs(m);
        (void)hipFree(b.p);
    }
    b.p = nullptr;
    b.n = 0;
    if (hipMalloc(&b.p, n ? n : 1) != hipSuccess) return nullptr;
    b.n = n;
    return b.p;
}

enum { WS_X0, WS_Z1, WS_X1, WS_Z2, WS_X2, WS_G, WS_T, WS_DG, WS_DZ2, WS_DZ1, WS_Y, WS_BOX, WS_IX1, WS_IX2, WS_IG };

int sync_weights(dg_cnn *m)
{
    for (int l = 0; l < 3; ++l) {
        MCHK(m, dg::cnn::pack(l + 1, 0, m->params + kWoff[l], m->Wp[l], m->stream), "pack");
        if (l > 0) MCHK(m, dg::cnn::pack(l + 1, 1, m->params + kWoff[l], m->Wt[l], m->stream), "pack");
        MHIP(m, hipMemcpyAsync(m->bias_pad[l], m->params + kBoff[l], kCout[l] * sizeof(float),
                               hipMemcpyDeviceToDevice, m->stream));
    }
    return DG_OK;
}

inline size_t cube(int a) { return (size_t)a * a * a; }

// forward of the training path on an input box in[3] (bf16 X0 already in WS_X0)
int forward_train(dg_cnn *m, int B, const int in[3])
{
    const int d1[3] = {in[0] - 2, in[1] - 2, in[2] - 2}, d2[3] = {in[0] - 4, in[1] - 4, in[2] - 4},
              d3[3] = {in[0] - 6, in[1] - 6, in[2] - 6};
    const size_t c1 = (size_t)B * d1[0] * d1[1] * d1[2], c2 = (size_t)B * d2[0] * d2[1] * d2[2],
                 c3 = (size_t)B * d3[0] * d3[1] * d3[2];
    float *Z1 = (float *)ws(m, WS_Z1, c1 * 32 * 4);
    void *X1 = ws(m, WS_X1, c1 * 32 * 2);
    float *Z2 = (float *)ws(m, WS_Z2, c2 * 32 * 4);
    void *X2 = ws(m, WS_X2, c2 * 32 * 2);
    float *G = (float *)ws(m, WS_G, c3 * 16 * 4);
    if (!Z1 || !X1 || !Z2 || !X2 || !G) return mfail(m, DG_ERR_ALLOC, "workspace");
    MCHK(m, dg::cnn::conv(1, dg::cnn::CONV_TRAIN, B, in, m->ws[WS_X0].p, m->Wp[0], m->bias_pad[0], Z1, X1, nullptr,
                          nullptr, m->stream), "conv1");
    MCHK(m, dg::cnn::conv(2, dg::cnn::CONV_TRAIN, B, d1, X1, m->Wp[1], m->bias_pad[1], Z2, X2, nullptr, nullptr,
                          m->stream), "conv2");
    MCHK(m, dg::cnn::conv(3, dg::cnn::CONV_OUT, B, d2, X2, m->Wp[2], m->bias_pad[2], nullptr, nullptr, G, nullptr,
                          m->stream), "conv3");
    return DG_OK;
}

// backward from dG (bf16, 16 channels, in WS_DG) -> m->grads
int backward(dg_cnn *m, int B, const int in[3])
{
    const int d1[3] = {in[0] - 2, in[1] - 2, in[2] - 2}, d2[3] = {in[0] - 4, in[1] - 4, in[2] - 4},
              d3[3] = {in[0] - 6, in[1] - 6, in[2] - 6};
    const size_t c1 = (size_t)B * d1[0] * d1[1] * d1[2], c2 = (size_t)B * d2[0] * d2[1] * d2[2],
                 c3 = (size_t)B * d3[0] * d3[1] * d3[2];
    void *dZ2 = ws(m, WS_DZ2, c2 * 32 * 2);
    void *dZ1 = ws(m, WS_DZ1, c1 * 32 * 2);
    if (!dZ2 || !dZ1) return mfail(m, DG_ERR_ALLOC, "workspace");
    void *dG = m->ws[WS_DG].p;
    // layer 3
    MCHK(m, dg::cnn::dw(3, B, d2, m->ws[WS_X2].p, dG, m->dw_part, kDwGroups, m->stream), "dw3");
    MCHK(m, dg::cnn::dw_reduce(3, m->dw_part, kDwGroups, m->grads + kW3, m->stream), "dw3 reduce");
    MCHK(m, dg::cnn::db(dG, (long long)c3, 16, 15, m->db_part, kDbBlocks, m->grads + kB3, m->stream), "db3");
    MCHK(m, dg::cnn::conv(3, dg::cnn::CONV_BWD, B, d3, dG, m->Wt[2], nullptr, nullptr, dZ2, nullptr,
                          (const float *)m->ws[WS_Z2].p, m->stream), "conv3 bwd");
    // layer 2
    MCHK(m, dg::cnn::dw(2, B, d1, m->ws[WS_X1].p, dZ2, m->dw_part, kDwGroups, m->stream), "dw2");
    MCHK(m, dg::cnn::dw_reduce(2, m->dw_part, kDwGroups, m->grads + kW2, m->stream), "dw2 reduce");
    MCHK(m, dg::cnn::db(dZ2, (long long)c2, 32, 32, m->db_part, kDbBlocks, m->grads + kB2, m->stream), "db2");
    MCHK(m, dg::cnn::conv(2, dg::cnn::CONV_BWD, B, d2, dZ2, m->Wt[1], nullptr, nullptr, dZ1, nullptr,
                          (const float *)m->ws[WS_Z1].p, m->stream), "conv2 bwd");
    // layer 1
    MCHK(m, dg::cnn::dw(1, B, in, m->ws[WS_X0].p, dZ1, m->dw_part, kDwGroups, m->stream), "dw1");
    MCHK(m, dg::cnn::dw_reduce(1, m->dw_part, kDwGroups, m->grads + kW1, m->stream), "dw1 reduce");
    MCHK(m, dg::cnn::db(dZ1, (long long)c1, 32, 32, m->db_part, kDbBlocks, m->grads + kB1, m->stream), "db1");
    return DG_OK;
}

int adam_update(dg_cnn *m, float lr, float b1, float b2, float eps)
{
    m->adam_t += 1;
    MCHK(m, dg::cnn::adam(m->params, m->grads, m->adam_m, m->adam_v, (int)kNParams, lr, b1, b2, eps, m->adam_t,
                          m->stream), "adam");
    return sync_weights(m);
}

// losses: per-block partials of one batch, summed on the host later
int loss_and_backward(dg_cnn *m, int B, int L, double *loss_part)
{
    const int in[3] = {L + 7, L + 7, L + 7};
    int rc = forward_train(m, B, in);
    if (rc) return rc;
    void *dG = ws(m, WS_DG, (size_t)B * cube(L + 1) * 16 * 2);
    if (!dG) return mfail(m, DG_ERR_ALLOC, "workspace");
    MCHK(m, dg::cnn::loss(B, L, (const float *)m->ws[WS_G].p, (const float *)m->ws[WS_T].p, dG, loss_part,
                          kLossBlocks, m->stream), "loss");
    return backward(m, B, in);
}

bool periodic_fp32(const dg_ctx *c)
{
    if (c->prec == DG_FP64) return false;
    for (int i = 0; i < 6; ++i)
        if (c->cfg.bc[i] != DG_PERIODIC) return false;
    return true;
}

int upload_tables(dg_cnn *m)
{
    const size_t n = m->sU.size();
    if (n > m->d_cap) {
        if (m->d_sU) (void)hipFree(m->d_sU);
        if (m->d_sR) (void)hipFree(m->d_sR);
        m->d_sU = m->d_sR = nullptr;
        m->d_cap = 0;
        const size_t cap = std::max<size_t>(n * 2, 64);
        MHIP(m, hipMalloc((void **)&m->d_sU, cap * sizeof(float *)));
        MHIP(m, hipMalloc((void **)&m->d_sR, cap * sizeof(float *)));
        m->d_cap = cap;
    }
    MHIP(m, hipMemcpy(m->d_sU, m->sU.data(), n * sizeof(float *), hipMemcpyHostToDevice));
    MHIP(m, hipMemcpy(m->d_sR, m->sR.data(), n * sizeof(float *), hipMemcpyHostToDevice));
    return DG_OK;
}

dg::cnn::PrepArgs prep_args(const dg_cnn *m, int B, const int n[3], int box, int halo, double ds, double gamma)
{
    dg::cnn::PrepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.B = B;
    for (int d = 0; d < 3; ++d) {
        a.n[d] = n[d];
        a.dims[d] = box;
    }
    a.halo = halo;
    a.ds = (float)ds;
    a.gamma = (float)gamma;
    for (int q = 0; q < 5; ++q) {
        a.shift[q] = m->norm[q];
        a.inv_scale[q] = 1.0f / m->norm[5 + q];
        a.inv_out[q] = 1.0f / m->norm[10 + q];
    }
    return a;
}

// Network forward on the current state of a periodic coarse context -> G
// (n^3 x 16, WS_IG): fp32, or bf16 (half) for the learned step's correction.
int infer(dg_cnn *m, dg_ctx *c, bool half)
{
    const int n[3] = {c->args.n[0], c->args.n[1], c->args.n[2]};
    const int in0[3] = {n[0] + 6, n[1] + 6, n[2] + 6}, in1[3] = {n[0] + 4, n[1] + 4, n[2] + 4},
              in2[3] = {n[0] + 2, n[1] + 2, n[2] + 2};
    void *X0 = ws(m, WS_X0, (size_t)in0[0] * in0[1] * in0[2] * 8 * 2);
    void *X1 = ws(m, WS_IX1, (size_t)in1[0] * in1[1] * in1[2] * 32 * 2);
    void *X2 = ws(m, WS_IX2, (size_t)in2[0] * in2[1] * in2[2] * 32 * 2);
    void *G = ws(m, WS_IG, (size_t)n[0] * n[1] * n[2] * 16 * 4);
    if (!X0 || !X1 || !X2 || !G) return mfail(m, DG_ERR_ALLOC, "workspace");
    dg::cnn::PrepArgs a = prep_args(m, 1, n, 0, 3, c->cfg.ds, c->cfg.gamma);
    for (int d = 0; d < 3; ++d) a.dims[d] = in0[d];
    // the learned step's forward also clears the correction's fallback count
    MCHK(m, dg::cnn::prep_input(a, m->d_ptrs + 5 * c->cur, m->d_box0, X0, m->stream,
                                half ? dg::corr_count(m->flags, c->n3) : nullptr), "prep");
    MCHK(m, dg::cnn::conv(1, dg::cnn::CONV_INFER, 1, in0, X0, m->Wp[0], m->bias_pad[0], nullptr, X1, nullptr, nullptr,
                          m->stream), "conv1");
    MCHK(m, dg::cnn::conv(2, dg::cnn::CONV_INFER, 1, in1, X1, m->Wp[1], m->bias_pad[1], nullptr, X2, nullptr, nullptr,
                          m->stream), "conv2");
    if (half)
        MCHK(m, dg::cnn::conv(3, dg::cnn::CONV_OUT16, 1, in2, X2, m->Wp[2], m->bias_pad[2], nullptr, G, nullptr,
                              nullptr, m->stream), "conv3");
    else
        MCHK(m, dg::cnn::conv(3, dg::cnn::CONV_OUT, 1, in2, X2, m->Wp[2], m->bias_pad[2], nullptr, nullptr,
                              static_cast<float *>(G), nullptr, m->stream), "conv3");
    return DG_OK;
}

// Captured learned steps bake in the normalization and the context's buffers.
void drop_graphs(dg_cnn *m)
{
    for (int b = 0; b < 2; ++b)
        if (m->graph[b]) {
            (void)hipGraphExecDestroy(m->graph[b]);
            m->graph[b] = nullptr;
        }
}

int prepare_infer(dg_cnn *m, dg_ctx *c)
{
    if (!periodic_fp32(c)) return mfail(m, DG_ERR_INVALID, "learned steps need an FP32 grid, periodic on all faces");
    if (!c->have_state) return mfail(m, DG_ERR_STATE, "no state set");
    if (c->stream != m->stream && hipStreamSynchronize(c->stream) != hipSuccess)
        return mfail(m, DG_ERR_DEVICE, "synchronize");
    if (!m->d_ptrs) {
        MHIP(m, hipMalloc((void **)&m->d_ptrs, 10 * sizeof(float *)));
        MHIP(m, hipMalloc((void **)&m->d_box0, 4 * sizeof(int)));
        const int zero[4] = {0, 0, 0, 0};
        MHIP(m, hipMemcpy(m->d_box0, zero, sizeof(zero), hipMemcpyHostToDevice));
    }
    if (m->graph_uid != c->uid) {
        drop_graphs(m);
        const float *p[10];
        for (int b = 0; b < 2; ++b)
            for (int q = 0; q < 5; ++q) p[5 * b + q] = static_cast<const float *>(c->buf[b][q]);
        MHIP(m, hipMemcpy(m->d_ptrs, p, sizeof(p), hipMemcpyHostToDevice));
        m->graph_uid = c->uid;
        m->graph_n3 = c->n3;
    }
    if (m->flags_n < c->n3) {
        if (m->flags) (void)hipFree(m->flags);
        m->flags = nullptr;
        m->flags_n = 0;
        MHIP(m, hipMalloc((void **)&m->flags, dg::corr_flags_bytes(c->n3)));
        m->flags_n = c->n3;
    }
    // workspaces of infer(), so that a captured step allocates nothing
    const size_t n0 = (size_t)(c->args.n[0] + 6) * (c->args.n[1] + 6) * (c->args.n[2] + 6),
                 n1 = (size_t)(c->args.n[0] + 4) * (c->args.n[1] + 4) * (c->args.n[2] + 4),
                 n2 = (size_t)(c->args.n[0] + 2) * (c->args.n[1] + 2) * (c->args.n[2] + 2);
    if (!ws(m, WS_X0, n0 * 16) || !ws(m, WS_IX1, n1 * 64) || !ws(m, WS_IX2, n2 * 64) || !ws(m, WS_IG, c->n3 * 64))
        return mfail(m, DG_ERR_ALLOC, "workspace");
    return DG_OK;
}

// Plain step, then the correction in WS_IG (bf16 when half), then dt and clock.
int correct_and_finalize(dg_cnn *m, dg_ctx *c, bool half)
{
    const void *in[5];
    void *out[5];
    for (int q = 0; q < 5; ++q) {
        in[q] = c->buf[c->cur][q];
        out[q] = c->buf[1 - c->cur][q];
    }
    MCHK(m, dg::launch_step(c->prec, c->fast, c->tile_y, c->args, in, out, c->sc, c->partials, m->stream), "step");
    // half: called after infer(half), whose input preparation cleared the fallback count
    MCHK(m, dg::launch_corr(c->fast, c->args, out, m->ws[WS_IG].p, half ? 1 : 0, m->norm + 10, c->sc, m->flags,
                            c->partials, m->stream, half ? 1 : 0), "correction");
    MCHK(m, dg::launch_finalize(c->prec, c->args, c->sc, c->partials, c->nparts_reduce, c->cfg.cfl,
                                c->cfg.dt_nan_fallback, 1, m->stream), "finalize");
    c->cur = 1 - c->cur;
    c->sums_stale = true;
    return DG_OK;
}

// One learned step on the model's stream (the context's buffers and scalars).
int learned_launch(dg_cnn *m, dg_ctx *c)
{
    int rc = infer(m, c, true);
    if (rc) return rc;
    return correct_and_finalize(m, c, true);
}

__global__ void arm_run_kernel(dg::Scalars *sc, double t_end, long long max_steps)
{
    const long long start = sc->steps;
    sc->t_end = t_end;
    sc->max_steps = (max_steps > LLONG_MAX - start) ? LLONG_MAX : start + max_steps;
    sc->active = !((sc->t > sc->t_end) || (sc->steps >= sc->max_steps));
}

// moments of the five primitives (mode 0, from conserved U) or of five raw
// fields (mode 1): per-block partial sums of q and q^2 (double)
__global__ void __launch_bounds__(256)
moments_kernel(long long n3, const float *f0, const float *f1, const float *f2, const float *f3, const float *f4,
               int mode, float V, float gm1, double *part)
{
    __shared__ double red[256][10];
    double s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long long)gridDim.x * blockDim.x) {
        float q[5] = {f0[i], f1[i], f2[i], f3[i], f4[i]};
        if (mode == 0) {
            const float r = q[0] / V;
            const float vx = q[1] / (r * V), vy = q[2] / (r * V), vz = q[3] / (r * V);
            const float p = (q[4] / V - 0.5f * r * (vx * vx + vy * vy + vz * vz)) * gm1;
            q[0] = r; q[1] = vx; q[2] = vy; q[3] = vz; q[4] = p;
        }
        for (int k = 0; k < 5; ++k) {
            s[k] += q[k];
            s[5 + k] += (double)q[k] * q[k];
        }
    }
    for (int k = 0; k < 10; ++k) red[threadIdx.x][k] = s[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int t = 1; t < 256; ++t)
            for (int k = 0; k < 10; ++k) s[k] += red[t][k];
        for (int k = 0; k < 10; ++k) part[blockIdx.x * 10 + k] = s[k];
    }
}

int moments(dg_cnn *m, long long n3, const float *const f[5], int mode, double out[10])
{
    const int nb = 512;
    double *part = (double *)ws(m, WS_BOX, (size_t)nb * 10 * sizeof(double) + 64);
    if (!part) return mfail(m, DG_ERR_ALLOC, "workspace");
    hipLaunchKernelGGL(moments_kernel, dim3(nb), dim3(256), 0, m->stream, n3, f[0], f[1], f[2], f[3], f[4], mode,
                       (float)(m->sds * m->sds * m->sds), (float)(m->sgamma - 1.0), part);
    std::vector<double> h((size_t)nb * 10);
    MHIP(m, hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    for (int k = 0; k < 10; ++k) out[k] = 0;
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 10; ++k) out[k] += h[(size_t)b * 10 + k];
    return DG_OK;
}

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int64_t dg_cnn_param_count(void) { return kNParams; }

int dg_cnn_create(int device, uint64_t seed, dg_cnn **out)
{
    if (!out) return DG_ERR_INVALID;
    *out = nullptr;
    dg_cnn *m = new (std::nothrow) dg_cnn();
    if (!m) return DG_ERR_ALLOC;
    m->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return DG_ERR_DEVICE;
    }
    bool ok = hipMalloc((void **)&m->params, kNParams * 4) == hipSuccess &&
              hipMalloc((void **)&m->grads, kNParams * 4) == hipSuccess &&
              hipMalloc((void **)&m->adam_m, kNParams * 4) == hipSuccess &&
              hipMalloc((void **)&m->adam_v, kNParams * 4) == hipSuccess &&
              hipMalloc((void **)&m->dw_part, (size_t)kDwGroups * 27 * 32 * 32 * 4) == hipSuccess &&
              hipMalloc((void **)&m->db_part, (size_t)kDbBlocks * 32 * 4) == hipSuccess;
    for (int l = 0; l < 3 && ok; ++l) {
        ok = hipMalloc((void **)&m->bias_pad[l], 32 * 4) == hipSuccess &&
             hipMemset(m->bias_pad[l], 0, 32 * 4) == hipSuccess &&
             hipMalloc(&m->Wp[l], dg::cnn::pack_elems(l + 1, 0) * 2) == hipSuccess;
        if (ok && l > 0) ok = hipMalloc(&m->Wt[l], dg::cnn::pack_elems(l + 1, 1) * 2) == hipSuccess;
    }
    if (!ok) {
        dg_cnn_destroy(m);
        return DG_ERR_ALLOC;
    }
    // He-normal weights (std sqrt(2 / fan_in)); the output layer scaled by 0.1
    // so that the initial corrections are small; zero biases.
    std::vector<float> p(kNParams, 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    for (int l = 0; l < 3; ++l) {
        const float sd = std::sqrt(2.0f / (27.0f * kCin[l])) * (l == 2 ? 0.1f : 1.0f);
        for (int64_t i = 0; i < 27 * kCin[l] * kCout[l]; ++i) p[kWoff[l] + i] = sd * nd(rng);
    }
    for (int q = 0; q < 5; ++q) {
        m->norm[q] = 0.0f;
        m->norm[5 + q] = 1.0f;
        m->norm[10 + q] = 1.0f;
    }
    if (hipMemcpy(m->params, p.data(), kNParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->adam_m, 0, kNParams * 4) != hipSuccess || hipMemset(m->adam_v, 0, kNParams * 4) != hipSuccess ||
        hipMemset(m->grads, 0, kNParams * 4) != hipSuccess || sync_weights(m) != DG_OK ||
        hipStreamSynchronize(m->stream) != hipSuccess) {
        dg_cnn_destroy(m);
        return DG_ERR_DEVICE;
    }
    *out = m;
    return DG_OK;
}

void dg_cnn_destroy(dg_cnn *m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (int b = 0; b < 2; ++b)
        if (m->graph[b]) (void)hipGraphExecDestroy(m->graph[b]);
    float *fs[] = {m->params, m->grads, m->adam_m, m->adam_v, m->dw_part, m->db_part, m->tmp,
                   m->bias_pad[0], m->bias_pad[1], m->bias_pad[2]};
    for (float *f : fs)
        if (f) (void)hipFree(f);
    for (int l = 0; l < 3; ++l) {
        if (m->Wp[l]) (void)hipFree(m->Wp[l]);
        if (m->Wt[l]) (void)hipFree(m->Wt[l]);
    }
    for (float *f : m->sU) (void)hipFree(f);
    for (float *f : m->sR) (void)hipFree(f);
    if (m->d_sU) (void)hipFree(m->d_sU);
    if (m->d_sR) (void)hipFree(m->d_sR);
    for (Buf &b : m->ws)
        if (b.p) (void)hipFree(b.p);
    if (m->loss_part) (void)hipFree(m->loss_part);
    if (m->d_ptrs) (void)hipFree(m->d_ptrs);
    if (m->d_box0) (void)hipFree(m->d_box0);
    if (m->flags) (void)hipFree(m->flags);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char *dg_cnn_last_error(const dg_cnn *m) { return m ? m->err.c_str() : "null model"; }

int dg_cnn_get_params(dg_cnn *m, float *params)
{
    if (!m || !params) return DG_ERR_INVALID;
    MHIP(m, hipMemcpyAsync(params, m->params, kNParams * 4, hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return DG_OK;
}

int dg_cnn_set_params(dg_cnn *m, const float *params)
{
    if (!m || !params) return DG_ERR_INVALID;
    MHIP(m, hipMemcpyAsync(m->params, params, kNParams * 4, hipMemcpyHostToDevice, m->stream));
    int rc = sync_weights(m);
    if (rc) return rc;
    MHIP(m, hipStreamSynchronize(m->stream));
    return DG_OK;
}

int dg_cnn_get_normalization(dg_cnn *m, float norm[15])
{
    if (!m || !norm) return DG_ERR_INVALID;
    std::memcpy(norm, m->norm, sizeof(m->norm));
    return DG_OK;
}

int dg_cnn_set_normalization(dg_cnn *m, const float norm[15])
{
    if (!m || !norm) return DG_ERR_INVALID;
    for (int q = 0; q < 5; ++q)
        if (!(norm[5 + q] > 0) || !(norm[10 + q] > 0)) return mfail(m, DG_ERR_INVALID, "scales must be positive");
    std::memcpy(m->norm, norm, sizeof(m->norm));
    m->norm_set = true;
    drop_graphs(m);
    return DG_OK;
}

int dg_cnn_forward(dg_cnn *m, int batch, const int32_t dims[3], const float *x, float *y)
{
    if (!m || batch <= 0 || !dims || !x || !y) return DG_ERR_INVALID;
    for (int d = 0; d < 3; ++d)
        if (dims[d] < 7) return mfail(m, DG_ERR_INVALID, "dims must be at least 7");
    MHIP(m, hipSetDevice(m->device));
    const int in0[3] = {dims[0], dims[1], dims[2]}, in1[3] = {dims[0] - 2, dims[1] - 2, dims[2] - 2},
              in2[3] = {dims[0] - 4, dims[1] - 4, dims[2] - 4};
    const size_t c0 = (size_t)batch * in0[0] * in0[1] * in0[2], c1 = (size_t)batch * in1[0] * in1[1] * in1[2],
                 c2 = (size_t)batch * in2[0] * in2[1] * in2[2],
                 c3 = (size_t)batch * (dims[0] - 6) * (dims[1] - 6) * (dims[2] - 6);
    float *xin = (float *)ws(m, WS_T, c0 * 5 * 4);
    void *X0 = ws(m, WS_X0, c0 * 8 * 2);
    void *X1 = ws(m, WS_IX1, c1 * 32 * 2);
    void *X2 = ws(m, WS_IX2, c2 * 32 * 2);
    float *G = (float *)ws(m, WS_IG, c3 * 16 * 4);
    if (!xin || !X0 || !X1 || !X2 || !G) return mfail(m, DG_ERR_ALLOC, "workspace");
    MHIP(m, hipMemcpyAsync(xin, x, c0 * 5 * 4, hipMemcpyHostToDevice, m->stream));
    MCHK(m, dg::cnn::to_bf16(xin, (long long)c0, 5, 8, X0, m->stream), "to_bf16");
    MCHK(m, dg::cnn::conv(1, dg::cnn::CONV_INFER, batch, in0, X0, m->Wp[0], m->bias_pad[0], nullptr, X1, nullptr,
                          nullptr, m->stream), "conv1");
    MCHK(m, dg::cnn::conv(2, dg::cnn::CONV_INFER, batch, in1, X1, m->Wp[1], m->bias_pad[1], nullptr, X2, nullptr,
                          nullptr, m->stream), "conv2");
    MCHK(m, dg::cnn::conv(3, dg::cnn::CONV_OUT, batch, in2, X2, m->Wp[2], m->bias_pad[2], nullptr, nullptr, G, nullptr,
                          m->stream), "conv3");
    std::vector<float> h(c3 * 16);
    MHIP(m, hipMemcpyAsync(h.data(), G, h.size() * 4, hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    for (size_t i = 0; i < c3; ++i)
        for (int k = 0; k < 15; ++k) y[i * 15 + k] = h[i * 16 + k];
    return DG_OK;
}

int dg_cnn_backward(dg_cnn *m, int batch, const int32_t dims[3], const float *x, const float *dy, float *grads)
{
    if (!m || batch <= 0 || !dims || !x || !dy) return DG_ERR_INVALID;
    for (int d = 0; d < 3; ++d)
        if (dims[d] < 7) return mfail(m, DG_ERR_INVALID, "dims must be at least 7");
    MHIP(m, hipSetDevice(m->device));
    const int in[3] = {dims[0], dims[1], dims[2]};
    const size_t c0 = (size_t)batch * in[0] * in[1] * in[2],
                 c3 = (size_t)batch * (dims[0] - 6) * (dims[1] - 6) * (dims[2] - 6);
    float *xin = (float *)ws(m, WS_T, std::max(c0 * 5, c3 * 15) * 4);
    void *X0 = ws(m, WS_X0, c0 * 8 * 2);
    void *dG = ws(m, WS_DG, c3 * 16 * 2);
    if (!xin || !X0 || !dG) return mfail(m, DG_ERR_ALLOC, "workspace");
    MHIP(m, hipMemcpyAsync(xin, x, c0 * 5 * 4, hipMemcpyHostToDevice, m->stream));
    MCHK(m, dg::cnn::to_bf16(xin, (long long)c0, 5, 8, X0, m->stream), "to_bf16");
    int rc = forward_train(m, batch, in);
    if (rc) return rc;
    MHIP(m, hipStreamSynchronize(m->stream));   // xin is reused for dy
    MHIP(m, hipMemcpyAsync(xin, dy, c3 * 15 * 4, hipMemcpyHostToDevice, m->stream));
    MCHK(m, dg::cnn::to_bf16(xin, (long long)c3, 15, 16, dG, m->stream), "to_bf16");
    if ((rc = backward(m, batch, in))) return rc;
    if (grads) MHIP(m, hipMemcpyAsync(grads, m->grads, kNParams * 4, hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return DG_OK;
}

int dg_cnn_train_batch(dg_cnn *m, int batch, int L, const float *x, const float *target, double *loss,
                       float *grads)
{
    if (!m || batch <= 0 || L < 1 || !x || !target) return DG_ERR_INVALID;
    MHIP(m, hipSetDevice(m->device));
    const size_t c0 = (size_t)batch * cube(L + 7), ct = (size_t)batch * cube(L);
    float *xin = (float *)ws(m, WS_Y, c0 * 5 * 4);
    void *X0 = ws(m, WS_X0, c0 * 8 * 2);
    float *T = (float *)ws(m, WS_T, ct * 5 * 4);
    if (!m->loss_part) MHIP(m, hipMalloc((void **)&m->loss_part, kLossBlocks * sizeof(double)));
    if (!xin || !X0 || !T) return mfail(m, DG_ERR_ALLOC, "workspace");
    MHIP(m, hipMemcpyAsync(xin, x, c0 * 5 * 4, hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(T, target, ct * 5 * 4, hipMemcpyHostToDevice, m->stream));
    MCHK(m, dg::cnn::to_bf16(xin, (long long)c0, 5, 8, X0, m->stream), "to_bf16");
    int rc = loss_and_backward(m, batch, L, m->loss_part);
    if (rc) return rc;
    std::vector<double> lp(kLossBlocks);
    MHIP(m, hipMemcpyAsync(lp.data(), m->loss_part, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    if (grads) MHIP(m, hipMemcpyAsync(grads, m->grads, kNParams * 4, hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    double s = 0;
    for (double v : lp) s += v;
    if (loss) *loss = s;
    return DG_OK;
}

int dg_cnn_adam_step(dg_cnn *m, const float *grads, float lr, float beta1, float beta2, float eps)
{
    if (!m || !(lr >= 0)) return DG_ERR_INVALID;
    MHIP(m, hipSetDevice(m->device));
    if (grads) MHIP(m, hipMemcpyAsync(m->grads, grads, kNParams * 4, hipMemcpyHostToDevice, m->stream));
    int rc = adam_update(m, lr, beta1, beta2, eps);
    if (rc) return rc;
    MHIP(m, hipStreamSynchronize(m->stream));
    return DG_OK;
}

int dg_cnn_add_sample(dg_cnn *m, dg_ctx *fine, dg_ctx *coarse) { return dg_cnn_add_sample_mode(m, fine, coarse, 0, 1); }

int dg_cnn_add_sample_mode(dg_cnn *m, dg_ctx *fine, dg_ctx *coarse, int keep_coarse, int store)
{
    if (!m || !fine || !coarse) return DG_ERR_INVALID;
    if (!periodic_fp32(fine) || !periodic_fp32(coarse)) return mfail(m, DG_ERR_INVALID, "FP32 periodic grids expected");
    for (int d = 0; d < 3; ++d)
        if (fine->args.n[d] != 2 * coarse->args.n[d]) return mfail(m, DG_ERR_INVALID, "fine grid must be 2x coarse");
    if (std::fabs(coarse->cfg.ds - 2 * fine->cfg.ds) > 1e-12 * coarse->cfg.ds)
        return mfail(m, DG_ERR_INVALID, "coarse ds must be 2x fine ds");
    if (!fine->have_state) return mfail(m, DG_ERR_STATE, "fine grid has no state");
    if (keep_coarse && !coarse->have_state) return mfail(m, DG_ERR_STATE, "coarse grid has no state");
    if (!m->sU.empty())
        for (int d = 0; d < 3; ++d)
            if (m->sn[d] != coarse->args.n[d]) return mfail(m, DG_ERR_INVALID, "samples must share the coarse grid");
    MHIP(m, hipSetDevice(m->device));
    // both grids must advance: a finished dg_run (a clock query is a run of
    // zero steps) leaves their scalars inactive, and an inactive step carries
    // the state through unchanged
    if (dg::arm_unbounded(fine) || dg::arm_unbounded(coarse)) return mfail(m, DG_ERR_DEVICE, "arm");
    MHIP(m, hipStreamSynchronize(fine->stream));
    MHIP(m, hipStreamSynchronize(coarse->stream));
    const size_t n3 = coarse->n3;
    if (!keep_coarse) {
        // U^n = R(fine), with its dt
        MCHK(m, dg::launch_restrict(fine->args, fine->buf[fine->cur], coarse->args, coarse->buf[coarse->cur],
                                    coarse->stream), "restrict");
        if (dg::reduce_current(coarse, 0)) return mfail(m, DG_ERR_DEVICE, coarse->err);
        coarse->have_state = true;
    }
    // otherwise U^n is the coarse grid's own state (its trajectory), dt already on the device
    std::vector<float *> u(5, nullptr), r(5, nullptr);
    if (store)
        for (int q = 0; q < 5; ++q) {
            if (hipMalloc((void **)&u[q], n3 * 4) != hipSuccess || hipMalloc((void **)&r[q], n3 * 4) != hipSuccess) {
                for (float *f : u) if (f) (void)hipFree(f);
                for (float *f : r) if (f) (void)hipFree(f);
                return mfail(m, DG_ERR_ALLOC, "sample storage");
            }
            MHIP(m, hipMemcpyAsync(u[q], coarse->buf[coarse->cur][q], n3 * 4, hipMemcpyDeviceToDevice,
                                   coarse->stream));
        }
    dg::Scalars sf, scs;
    MHIP(m, hipMemcpyAsync(&scs, coarse->sc, sizeof(scs), hipMemcpyDeviceToHost, coarse->stream));
    MHIP(m, hipMemcpyAsync(&sf, fine->sc, sizeof(sf), hipMemcpyDeviceToHost, fine->stream));
    MHIP(m, hipStreamSynchronize(coarse->stream));
    MHIP(m, hipStreamSynchronize(fine->stream));
    const double dt = std::min(scs.dt, 2.0 * sf.dt);
    if (!(dt > 0) || !std::isfinite(dt)) return mfail(m, DG_ERR_NONFINITE, "dt is not finite");
    // coarse plain step with dt; two fine steps with dt/2
    MCHK(m, dg::launch_set_dt(coarse->sc, dt, coarse->stream), "set dt");
    if (dg::launch_one_step(coarse)) return mfail(m, DG_ERR_DEVICE, coarse->err);
    for (int k = 0; k < 2; ++k) {
        MCHK(m, dg::launch_set_dt(fine->sc, 0.5 * dt, fine->stream), "set dt");
        if (dg::launch_one_step(fine)) return mfail(m, DG_ERR_DEVICE, fine->err);
    }
    if (m->tmp_cells < n3) {
        if (m->tmp) (void)hipFree(m->tmp);
        m->tmp = nullptr;
        MHIP(m, hipMalloc((void **)&m->tmp, 5 * n3 * 4));
        m->tmp_cells = n3;
    }
    if (!store) {
        // both grids advanced by dt, nothing kept
        MHIP(m, hipStreamSynchronize(fine->stream));
        MHIP(m, hipStreamSynchronize(coarse->stream));
        return DG_OK;
    }
    void *t[5];
    for (int q = 0; q < 5; ++q) t[q] = m->tmp + q * n3;
    MHIP(m, hipStreamSynchronize(fine->stream));
    MCHK(m, dg::launch_restrict(fine->args, fine->buf[fine->cur], coarse->args, t, coarse->stream), "restrict");
    const void *tc[5] = {t[0], t[1], t[2], t[3], t[4]};
    void *rr[5] = {r[0], r[1], r[2], r[3], r[4]};
    MCHK(m, dg::launch_residual((long long)n3, tc, coarse->buf[coarse->cur], dt, coarse->cfg.ds, rr, coarse->stream),
         "residual");
    MHIP(m, hipStreamSynchronize(coarse->stream));
    for (int q = 0; q < 5; ++q) {
        m->sU.push_back(u[q]);
        m->sR.push_back(r[q]);
    }
    for (int d = 0; d < 3; ++d) m->sn[d] = coarse->args.n[d];
    m->sds = coarse->cfg.ds;
    m->sgamma = coarse->cfg.gamma;
    return upload_tables(m);
}

int dg_cnn_sample_count(dg_cnn *m) { return m ? (int)(m->sU.size() / 5) : 0; }

int dg_cnn_set_input_noise(dg_cnn *m, float std)
{
    if (!m || !(std >= 0.0f)) return DG_ERR_INVALID;
    m->input_noise = std;
    return DG_OK;
}

int dg_cnn_fit_normalization(dg_cnn *m)
{
    if (!m) return DG_ERR_INVALID;
    const int ns = dg_cnn_sample_count(m);
    if (ns == 0) return mfail(m, DG_ERR_STATE, "no samples");
    const long long n3 = (long long)m->sn[0] * m->sn[1] * m->sn[2];
    double su[10] = {0}, sr[10] = {0};
    for (int k = 0; k < ns; ++k) {
        double a[10], b[10];
        int rc = moments(m, n3, &m->sU[5 * k], 0, a);
        if (!rc) rc = moments(m, n3, &m->sR[5 * k], 1, b);
        if (rc) return rc;
        for (int i = 0; i < 10; ++i) {
            su[i] += a[i];
            sr[i] += b[i];
        }
    }
    const double N = (double)n3 * ns;
    for (int q = 0; q < 5; ++q) {
        const double mu = su[q] / N, var = std::max(su[5 + q] / N - mu * mu, 0.0);
        m->norm[q] = (float)mu;
        m->norm[5 + q] = (float)std::max(std::sqrt(var), 1e-6);
        const double rms = std::sqrt(std::max(sr[5 + q] / N, 0.0));   // output scale: RMS of R
        m->norm[10 + q] = (float)std::max(rms, 1e-12);
    }
    m->norm_set = true;
    drop_graphs(m);
    return DG_OK;
}

int dg_cnn_train(dg_cnn *m, int steps, int batch, int L, float lr, uint64_t seed, float *losses)
{
    if (!m || steps <= 0 || batch <= 0 || L < 1) return DG_ERR_INVALID;
    const int ns = dg_cnn_sample_count(m);
    if (ns == 0) return mfail(m, DG_ERR_STATE, "no samples");
    MHIP(m, hipSetDevice(m->device));
    if (!m->norm_set) {
        int rc = dg_cnn_fit_normalization(m);
        if (rc) return rc;
    }
    const size_t c0 = (size_t)batch * cube(L + 7), ct = (size_t)batch * cube(L);
    void *X0 = ws(m, WS_X0, c0 * 8 * 2);
    float *T = (float *)ws(m, WS_T, ct * 5 * 4);
    int *boxes = (int *)ws(m, WS_BOX, (size_t)steps * batch * 4 * sizeof(int));
    double *lp = (double *)ws(m, WS_Y, (size_t)steps * kLossBlocks * sizeof(double));
    if (!X0 || !T || !boxes || !lp) return mfail(m, DG_ERR_ALLOC, "workspace");
    std::vector<int> hb((size_t)steps * batch * 4);
    uint64_t st = seed;
    for (int s = 0; s < steps; ++s)
        for (int b = 0; b < batch; ++b) {
            int *e = &hb[((size_t)s * batch + b) * 4];
            e[0] = (int)(splitmix(st) % (uint64_t)ns);
            for (int d = 0; d < 3; ++d) e[1 + d] = (int)(splitmix(st) % (uint64_t)m->sn[d]);
        }
    MHIP(m, hipMemcpyAsync(boxes, hb.data(), hb.size() * sizeof(int), hipMemcpyHostToDevice, m->stream));
    for (int s = 0; s < steps; ++s) {
        const int *bx = boxes + (size_t)s * batch * 4;
        dg::cnn::PrepArgs a = prep_args(m, batch, m->sn, L + 7, 4, m->sds, m->sgamma);
        a.noise = m->input_noise;
        a.seed = (unsigned int)(splitmix(st) & 0xffffffffu);
        MCHK(m, dg::cnn::prep_input(a, m->d_sU, bx, X0, m->stream), "prep input");
        dg::cnn::PrepArgs at = prep_args(m, batch, m->sn, L, 0, m->sds, m->sgamma);
        MCHK(m, dg::cnn::prep_target(at, m->d_sR, bx, T, m->stream), "prep target");
        int rc = loss_and_backward(m, batch, L, lp + (size_t)s * kLossBlocks);
        if (rc) return rc;
        if ((rc = adam_update(m, lr, 0.9f, 0.999f, 1e-8f))) return rc;
    }
    std::vector<double> h((size_t)steps * kLossBlocks);
    MHIP(m, hipMemcpyAsync(h.data(), lp, h.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    if (losses)
        for (int s = 0; s < steps; ++s) {
            double v = 0;
            for (int k = 0; k < kLossBlocks; ++k) v += h[(size_t)s * kLossBlocks + k];
            losses[s] = (float)v;
        }
    return DG_OK;
}

int dg_learned_step(dg_ctx *c, dg_cnn *m, double dt, double *dt_used)
{
    if (!c || !m) return DG_ERR_INVALID;
    MHIP(m, hipSetDevice(m->device));
    int rc = prepare_infer(m, c);
    if (rc) return rc;
    hipLaunchKernelGGL(arm_run_kernel, dim3(1), dim3(1), 0, m->stream, c->sc, (double)INFINITY, (long long)LLONG_MAX);
    if (dt > 0) MCHK(m, dg::launch_set_dt(c->sc, dt, m->stream), "set dt");
    dg::Scalars s;
    MHIP(m, hipMemcpyAsync(&s, c->sc, sizeof(s), hipMemcpyDeviceToHost, m->stream));
    if ((rc = learned_launch(m, c))) return rc;
    MHIP(m, hipStreamSynchronize(m->stream));
    if (dt_used) *dt_used = s.dt;
    return DG_OK;
}

int dg_learned_run(dg_ctx *c, dg_cnn *m, int64_t max_steps, double t_end, double *t_out, int64_t *steps_out)
{
    if (!c || !m || max_steps < 0) return DG_ERR_INVALID;
    MHIP(m, hipSetDevice(m->device));
    int rc = prepare_infer(m, c);
    if (rc) return rc;
    hipLaunchKernelGGL(arm_run_kernel, dim3(1), dim3(1), 0, m->stream, c->sc, t_end, (long long)max_steps);
    const bool timed = std::isfinite(t_end);
    dg::Scalars s;
    int64_t left = max_steps;
    while (left > 0) {
        if (left >= kGraphSteps) {
            if (!m->graph[c->cur]) {
                const int cur0 = c->cur;
                hipGraph_t g = nullptr;
                MHIP(m, hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < kGraphSteps && rc == DG_OK; ++k) rc = learned_launch(m, c);
                hipError_t e = hipStreamEndCapture(m->stream, &g);
                c->cur = cur0;
                if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
                if (e != hipSuccess) return mfail(m, DG_ERR_DEVICE, "hipStreamEndCapture");
                e = hipGraphInstantiate(&m->graph[cur0], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) { m->graph[cur0] = nullptr; return mfail(m, DG_ERR_DEVICE, "hipGraphInstantiate"); }
            }
            MHIP(m, hipGraphLaunch(m->graph[c->cur], m->stream));
            left -= kGraphSteps;
            c->sums_stale = true;
        } else {
            if ((rc = learned_launch(m, c))) return rc;
            left -= 1;
        }
        if (timed) {
            MHIP(m, hipMemcpyAsync(&s, c->sc, sizeof(s), hipMemcpyDeviceToHost, m->stream));
            MHIP(m, hipStreamSynchronize(m->stream));
            if (!s.active) break;
        }
    }
    MHIP(m, hipMemcpyAsync(&s, c->sc, sizeof(s), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    if (t_out) *t_out = s.t;
    if (steps_out) *steps_out = s.steps;
    return DG_OK;
}

int dg_learned_step_given(dg_cnn *m, dg_ctx *c, const float *G, double *dt_used)
{
    if (!c || !m || !G) return DG_ERR_INVALID;
    MHIP(m, hipSetDevice(m->device));
    int rc = prepare_infer(m, c);
    if (rc) return rc;
    MHIP(m, hipMemcpyAsync(m->ws[WS_IG].p, G, c->n3 * 16 * 4, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(arm_run_kernel, dim3(1), dim3(1), 0, m->stream, c->sc, (double)INFINITY, (long long)LLONG_MAX);
    dg::Scalars s;
    MHIP(m, hipMemcpyAsync(&s, c->sc, sizeof(s), hipMemcpyDeviceToHost, m->stream));
    if ((rc = correct_and_finalize(m, c, false))) return rc;
    MHIP(m, hipStreamSynchronize(m->stream));
    if (dt_used) *dt_used = s.dt;
    return DG_OK;
}

int dg_learned_time(dg_cnn *m, dg_ctx *c, int nrep, double ms[5])
{
    if (!m || !c || nrep <= 0 || !ms) return DG_ERR_INVALID;
    MHIP(m, hipSetDevice(m->device));
    int rc = prepare_infer(m, c);
    if (rc) return rc;
    const int n[3] = {c->args.n[0], c->args.n[1], c->args.n[2]};
    const int in0[3] = {n[0] + 6, n[1] + 6, n[2] + 6}, in1[3] = {n[0] + 4, n[1] + 4, n[2] + 4},
              in2[3] = {n[0] + 2, n[1] + 2, n[2] + 2};
    dg::EventSet es;
    MHIP(m, es.create(6));
    std::vector<hipEvent_t> &ev = es.ev;
    double acc[5] = {0, 0, 0, 0, 0};
    dg::cnn::PrepArgs a = prep_args(m, 1, n, 0, 3, c->cfg.ds, c->cfg.gamma);
    for (int d = 0; d < 3; ++d) a.dims[d] = in0[d];
    for (int r = 0; r < nrep && rc == DG_OK; ++r) {
        MHIP(m, hipEventRecord(ev[0], m->stream));
        MCHK(m, dg::cnn::prep_input(a, m->d_ptrs + 5 * c->cur, m->d_box0, m->ws[WS_X0].p, m->stream), "prep");
        MHIP(m, hipEventRecord(ev[1], m->stream));
        MCHK(m, dg::cnn::conv(1, dg::cnn::CONV_INFER, 1, in0, m->ws[WS_X0].p, m->Wp[0], m->bias_pad[0], nullptr,
                              m->ws[WS_IX1].p, nullptr, nullptr, m->stream), "conv1");
        MHIP(m, hipEventRecord(ev[2], m->stream));
        MCHK(m, dg::cnn::conv(2, dg::cnn::CONV_INFER, 1, in1, m->ws[WS_IX1].p, m->Wp[1], m->bias_pad[1], nullptr,
                              m->ws[WS_IX2].p, nullptr, nullptr, m->stream), "conv2");
        MHIP(m, hipEventRecord(ev[3], m->stream));
        MCHK(m, dg::cnn::conv(3, dg::cnn::CONV_OUT16, 1, in2, m->ws[WS_IX2].p, m->Wp[2], m->bias_pad[2], nullptr,
                              m->ws[WS_IG].p, nullptr, nullptr, m->stream), "conv3");   // as the learned step
        MHIP(m, hipEventRecord(ev[4], m->stream));
        MHIP(m, hipEventSynchronize(ev[4]));
        float t;
        for (int k = 0; k < 4; ++k) {
            (void)hipEventElapsedTime(&t, ev[k], ev[k + 1]);
            acc[k] += t;
        }
        (void)hipEventElapsedTime(&t, ev[0], ev[4]);
        acc[4] += t;
    }
    for (int k = 0; k < 5; ++k) ms[k] = acc[k] / nrep;
    return DG_OK;
}

int dg_learned_flagged(dg_cnn *m, int64_t *count)
{
    if (!m || !count || !m->flags || !m->graph_uid) return DG_ERR_INVALID;
    std::vector<unsigned char> h(m->graph_n3);
    MHIP(m, hipMemcpyAsync(h.data(), m->flags, h.size(), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    int64_t n = 0;
    for (unsigned char f : h) n += f;
    *count = n;
    return DG_OK;
}

int dg_learned_correction(dg_cnn *m, dg_ctx *c, float *G)
{
    if (!m || !c || !G) return DG_ERR_INVALID;
    MHIP(m, hipSetDevice(m->device));
    int rc = prepare_infer(m, c);
    if (rc) return rc;
    if ((rc = infer(m, c, false))) return rc;
    MHIP(m, hipMemcpyAsync(G, m->ws[WS_IG].p, c->n3 * 16 * 4, hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return DG_OK;
}

int dg_step_dt(dg_ctx *c, double dt)
{
    if (!c || !(dt > 0)) return DG_ERR_INVALID;
    if (!c->have_state) return DG_ERR_STATE;
    if (hipSetDevice(c->cfg.device) != hipSuccess) return DG_ERR_DEVICE;
    int rc = dg::arm_unbounded(c);   // as dg_step: a finished dg_run must not turn this step into a no-op
    if (rc) return rc;
    if (dg::launch_set_dt(c->sc, dt, c->stream)) return DG_ERR_DEVICE;
    rc = dg::launch_one_step(c);
    if (rc) return rc;
    return hipStreamSynchronize(c->stream) == hipSuccess ? DG_OK : DG_ERR_DEVICE;
}

int dg_restrict(dg_ctx *fine, dg_ctx *coarse)
{
    if (!fine || !coarse || fine->prec == DG_FP64 || coarse->prec == DG_FP64) return DG_ERR_INVALID;
    for (int d = 0; d < 3; ++d)
        if (fine->args.n[d] != 2 * coarse->args.n[d]) return DG_ERR_INVALID;
    if (!fine->have_state) return DG_ERR_STATE;
    if (hipStreamSynchronize(fine->stream) != hipSuccess) return DG_ERR_DEVICE;
    if (dg::launch_restrict(fine->args, fine->buf[fine->cur], coarse->args, coarse->buf[coarse->cur], coarse->stream))
        return DG_ERR_DEVICE;
    int rc = dg::reduce_current(coarse, 0);
    if (rc) return rc;
    coarse->have_state = true;
    return hipStreamSynchronize(coarse->stream) == hipSuccess ? DG_OK : DG_ERR_DEVICE;
}

}  // extern "C"
