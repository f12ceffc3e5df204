// fhip_mlp.hip -- the learned warm start: a 3-128-128-13 multilayer perceptron
// (GELU hidden layers) mapping (temperature, field, pressure) to the
// equilibrium dipole statistics and homogeneous strain that plain MD reaches.
// It is trained on equilibrated states that the MD itself produces and seeds
// new runs near equilibrium; the MD still judges equilibrium.  This part has
// no counterpart in the reference.
//
// Arithmetic: BF16 operands, FP32 accumulation, on gfx950's
// v_mfma_f32_16x16x32_bf16.  One 64-lane workgroup owns 16 samples; every
// layer, forward and backward, is a sequence of 16x16x32 MFMA tiles whose
// operands come from registers (inputs) or LDS (activations, transposes).
// Weight gradients of each workgroup go to its own slot of a partial buffer
// and the Adam kernel folds them in a fixed order, so training is
// deterministic.  The optimiser keeps FP32 master weights and writes the two
// BF16 packings the MFMA tiles read (forward W^T and backward W).
#include "../../include/feram_hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

constexpr int NIN = 3, NH = 128, NOUT = 13;
constexpr int KIN = 32;                 // input width padded to one MFMA K step
constexpr int MS = 16;                  // samples per workgroup (one M tile)
constexpr int WL = 64;                  // lanes per workgroup (one wave)

// FP32 parameter layout (row-major W[out][in]), as fhip_mlp_set_params takes it.
constexpr int OW1 = 0, OB1 = OW1 + NH * NIN, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NOUT * NH, NPARAM = OB3 + NOUT;
constexpr int NPART = NPARAM + 1;       // per-workgroup partial: gradients, then the loss

// BF16 packings.  Forward operand of a layer with W[NO][NI]: B[k][n] = W[n][k]
// (K = inputs, N = outputs); backward: B[k][n] = W[k][n] (K = outputs, N = inputs).
// A packed operand is [N tile][K step][lane][8]: lane l of tile t, step s holds
// W-derived values at n = 16 t + l % 16, k = 32 s + 8 (l / 16) + j.
constexpr int F1_T = NH / 16, F1_S = 1;          // W1 forward: 8 tiles x 1 step
constexpr int F2_T = NH / 16, F2_S = NH / 32;    // W2 forward / backward: 8 x 4
constexpr int F3_T = 1, F3_S = NH / 32;          // W3 forward: 1 tile (13 of 16) x 4 steps
constexpr int B3_T = NH / 16, B3_S = 1;          // W3 backward: 8 tiles x 1 step (13 of 32)
constexpr int PK = 64 * 8;                       // bf16 per (tile, step)
constexpr int PF1 = 0, PF2 = PF1 + F1_T * F1_S * PK, PF3 = PF2 + F2_T * F2_S * PK,
              PB2 = PF3 + F3_T * F3_S * PK, PB3 = PB2 + F2_T * F2_S * PK, NPACK = PB3 + B3_T * B3_S * PK;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even FP32 -> BF16 (finite inputs).
__device__ __forceinline__ __bf16 to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    const unsigned short h = (unsigned short)(u >> 16);
    return __builtin_bit_cast(__bf16, h);
}

__device__ __forceinline__ float bf16_to_f(__bf16 b)
{
    const unsigned short h = __builtin_bit_cast(unsigned short, b);
    return __uint_as_float((unsigned)h << 16);
}

// GELU with the exact erf form and its derivative.
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Packed B operand of tile t, step s for lane l.
__device__ __forceinline__ bf16x8 load_pack(const __bf16* pk, int base, int nsteps, int t, int s, int l)
{
    return *reinterpret_cast<const bf16x8*>(pk + base + ((size_t)(t * nsteps + s) * 64 + l) * 8);
}

// A operand from an LDS activation matrix act[16][ld] (row = sample): row
// l % 16, columns 32 s + 8 (l / 16) .. + 7.
__device__ __forceinline__ bf16x8 load_rows(const __bf16* act, int ld, int s, int l)
{
    return *reinterpret_cast<const bf16x8*>(act + (l % 16) * ld + 32 * s + 8 * (l / 16));
}

// Operand with K = samples (16 valid of 32) from act[16][ld]: lane l takes
// column c0 + l % 16 at samples 8 (l / 16) + j; lanes 32..63 are the zero pad.
__device__ __forceinline__ bf16x8 load_cols(const __bf16* act, int ld, int c0, int l)
{
    bf16x8 v;
    const int r0 = 8 * (l / 16);
#pragma unroll
    for (int j = 0; j < 8; ++j)
        v[j] = (r0 + j < MS) ? act[(r0 + j) * ld + c0 + l % 16] : to_bf16(0.0f);
    return v;
}

constexpr int LD_H = NH + 8;            // LDS row stride of 128-wide activations (bf16)
constexpr int LD_I = KIN + 8;           // of the 32-wide input and output-gradient rows

struct MlpArgs {
    const float* x;          // [n][3] normalised inputs
    const float* y;          // [n][13] normalised targets (training)
    const float* wout;       // [13] loss weight per output
    float* out;              // [n][13] predictions (forward only) or nullptr
    float* part;             // [nwg][NPART] gradients + loss (training) or nullptr
    const float* master;     // FP32 parameters (biases)
    const __bf16* pack;      // BF16 packings
    int n;
    float inv_n;             // 1 / n
};

// One workgroup: 16 samples through the network; with TRAIN the loss, the
// backward pass and this workgroup's weight gradients.
template <int TRAIN>
__global__ __launch_bounds__(WL) void k_mlp(MlpArgs a)
{
    __shared__ __attribute__((aligned(16))) __bf16 X0[MS * LD_I];
    __shared__ __attribute__((aligned(16))) __bf16 A1[MS * LD_H];
    __shared__ __attribute__((aligned(16))) __bf16 A2[MS * LD_H];
    __shared__ __attribute__((aligned(16))) __bf16 G3[MS * LD_I];
    __shared__ __attribute__((aligned(16))) __bf16 G2[MS * LD_H];
    __shared__ __attribute__((aligned(16))) __bf16 G1[MS * LD_H];
    const int l = threadIdx.x;
    const int s0 = blockIdx.x * MS;
    const int col = l % 16, rq = 4 * (l / 16);     // D layout: column, first of 4 rows
    const float* b1 = a.master + OB1;
    const float* b2 = a.master + OB2;
    const float* b3 = a.master + OB3;

    // inputs (zero-padded to 32 columns and to 16 rows)
    for (int e = l; e < MS * LD_I; e += WL) {
        const int r = e / LD_I, c = e % LD_I;
        const int smp = s0 + r;
        X0[e] = to_bf16((c < NIN && smp < a.n) ? a.x[(size_t)smp * NIN + c] : 0.0f);
    }
    __syncthreads();

    // layer 1: Z1 = X0 W1^T + b1, A1 = gelu(Z1)
    f32x4 z1[NH / 16];
    {
        const bf16x8 xa = load_rows(X0, LD_I, 0, l);
#pragma unroll
        for (int t = 0; t < NH / 16; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(xa, load_pack(a.pack, PF1, F1_S, t, 0, l), acc);
            const float bb = b1[16 * t + col];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[j] += bb;
                A1[(rq + j) * LD_H + 16 * t + col] = to_bf16(gelu(acc[j]));
            }
            z1[t] = acc;
        }
    }
    __syncthreads();

    // layer 2
    f32x4 z2[NH / 16];
    {
        bf16x8 ar[NH / 32];
#pragma unroll
        for (int s = 0; s < NH / 32; ++s) ar[s] = load_rows(A1, LD_H, s, l);
#pragma unroll
        for (int t = 0; t < NH / 16; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < NH / 32; ++s) acc = mfma(ar[s], load_pack(a.pack, PF2, F2_S, t, s, l), acc);
            const float bb = b2[16 * t + col];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[j] += bb;
                A2[(rq + j) * LD_H + 16 * t + col] = to_bf16(gelu(acc[j]));
            }
            z2[t] = acc;
        }
    }
    __syncthreads();

    // layer 3 (13 outputs in a 16-wide tile)
    f32x4 z3 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < NH / 32; ++s) z3 = mfma(load_rows(A2, LD_H, s, l), load_pack(a.pack, PF3, F3_S, 0, s, l), z3);
    const float bb3 = col < NOUT ? b3[col] : 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) z3[j] += bb3;

    if (!TRAIN) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int smp = s0 + rq + j;
            if (col < NOUT && smp < a.n) a.out[(size_t)smp * NOUT + col] = z3[j];
        }
        return;
    }

    // ---- loss and backward ----------------------------------------------------
    float* part = a.part + (size_t)blockIdx.x * NPART;
    float g3[4];
    float lsum = 0.0f;
    const float wo = col < NOUT ? a.wout[col] : 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int smp = s0 + rq + j;
        float d = 0.0f;
        if (col < NOUT && smp < a.n) {
            const float r = z3[j] - a.y[(size_t)smp * NOUT + col];
            lsum += wo * r * r;
            d = 2.0f * wo * r * a.inv_n;
        }
        g3[j] = d;
    }
    for (int e = l; e < MS * LD_I; e += WL) G3[e] = to_bf16(0.0f);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) G3[(rq + j) * LD_I + col] = to_bf16(g3[j]);
    // db3 and the loss: column sums over the 16 rows (4 per lane, 4 lane groups)
    {
        float cs = g3[0] + g3[1] + g3[2] + g3[3];
        cs += __shfl_xor(cs, 16);
        cs += __shfl_xor(cs, 32);
        if (l < NOUT) part[OB3 + l] = cs;
        float ls = lsum;
        for (int o = 32; o >= 1; o >>= 1) ls += __shfl_xor(ls, o);
        if (l == 0) part[NPARAM] = ls * a.inv_n;
    }
    __syncthreads();

    // dA2 = G3 W3 (K = outputs padded to 32), G2 = dA2 * gelu'(Z2)
    {
        const bf16x8 ga = load_rows(G3, LD_I, 0, l);
#pragma unroll
        for (int t = 0; t < NH / 16; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(ga, load_pack(a.pack, PB3, B3_S, t, 0, l), acc);
            float cs = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float g = acc[j] * gelu_d(z2[t][j]);
                cs += g;
                G2[(rq + j) * LD_H + 16 * t + col] = to_bf16(g);
            }
            cs += __shfl_xor(cs, 16);
            cs += __shfl_xor(cs, 32);
            if (l < 16) part[OB2 + 16 * t + l] = cs;
        }
    }
    __syncthreads();

    // dA1 = G2 W2, G1 = dA1 * gelu'(Z1)
    {
        bf16x8 gr[NH / 32];
#pragma unroll
        for (int s = 0; s < NH / 32; ++s) gr[s] = load_rows(G2, LD_H, s, l);
#pragma unroll
        for (int t = 0; t < NH / 16; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < NH / 32; ++s) acc = mfma(gr[s], load_pack(a.pack, PB2, F2_S, t, s, l), acc);
            float cs = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float g = acc[j] * gelu_d(z1[t][j]);
                cs += g;
                G1[(rq + j) * LD_H + 16 * t + col] = to_bf16(g);
            }
            cs += __shfl_xor(cs, 16);
            cs += __shfl_xor(cs, 32);
            if (l < 16) part[OB1 + 16 * t + l] = cs;
        }
    }
    __syncthreads();

    // weight gradients, K = the 16 samples: dW[m][n] = sum_i G[i][m] act[i][n]
    // dW3 (m = outputs, 13 of 16; n = hidden)
    {
        const bf16x8 ga = load_cols(G3, LD_I, 0, l);
#pragma unroll
        for (int tn = 0; tn < NH / 16; ++tn) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(ga, load_cols(A2, LD_H, 16 * tn, l), acc);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (rq + j < NOUT) part[OW3 + (rq + j) * NH + 16 * tn + col] = acc[j];
        }
    }
    // dW2 (m = hidden 2, n = hidden 1)
#pragma unroll 1
    for (int tm = 0; tm < NH / 16; ++tm) {
        const bf16x8 ga = load_cols(G2, LD_H, 16 * tm, l);
#pragma unroll
        for (int tn = 0; tn < NH / 16; ++tn) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(ga, load_cols(A1, LD_H, 16 * tn, l), acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) part[OW2 + (16 * tm + rq + j) * NH + 16 * tn + col] = acc[j];
        }
    }
    // dW1 (m = hidden 1, n = inputs, 3 of 16)
#pragma unroll
    for (int tm = 0; tm < NH / 16; ++tm) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(load_cols(G1, LD_H, 16 * tm, l), load_cols(X0, LD_I, 0, l), acc);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (col < NIN) part[OW1 + (16 * tm + rq + j) * NIN + col] = acc[j];
    }
}

// BF16 packings of the FP32 master weights: each parameter goes to its
// forward and (W2, W3) backward positions; the padding stays zero.
__device__ __forceinline__ void pack_param(__bf16* pk, int p, float w)
{
    const __bf16 b = to_bf16(w);
    auto put = [&](int base, int nsteps, int n, int k) {
        const int t = n / 16, s = k / 32, kk = k % 32;
        const int lane = (n % 16) + 16 * (kk / 8), j = kk % 8;
        pk[base + ((size_t)(t * nsteps + s) * 64 + lane) * 8 + j] = b;
    };
    if (p < OB1) {                      // W1[o][i]: forward n = o, k = i
        put(PF1, F1_S, p / NIN, p % NIN);
    } else if (p >= OW2 && p < OB2) {   // W2[o][i]
        const int o = (p - OW2) / NH, i = (p - OW2) % NH;
        put(PF2, F2_S, o, i);
        put(PB2, F2_S, i, o);           // backward: n = i, k = o
    } else if (p >= OW3 && p < OB3) {   // W3[o][i]
        const int o = (p - OW3) / NH, i = (p - OW3) % NH;
        put(PF3, F3_S, o, i);
        put(PB3, B3_S, i, o);
    }
}

__global__ __launch_bounds__(256) void k_mlp_pack(const float* master, __bf16* pk)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < NPARAM) pack_param(pk, p, master[p]);
}

struct AdamArgs {
    float* master;
    float* m;
    float* v;
    __bf16* pack;
    const float* part;
    int nwg;
    float lr, beta1, beta2, eps, c1, c2;   // c1 = 1 - beta1^t, c2 = 1 - beta2^t
    float* loss;                           // loss of the parameters before this update
};

// Fold the per-workgroup gradients in workgroup order, Adam update, repack.
__global__ __launch_bounds__(256) void k_mlp_adam(AdamArgs a)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p > NPARAM) return;
    float g = 0.0f;
    for (int w = 0; w < a.nwg; ++w) g += a.part[(size_t)w * NPART + p];
    if (p == NPARAM) {
        if (a.loss) *a.loss = g;
        return;
    }
    const float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * g;
    const float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * g * g;
    a.m[p] = m;
    a.v[p] = v;
    const float w = a.master[p] - a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
    a.master[p] = w;
    pack_param(a.pack, p, w);
}

// Gradient of the whole batch folded in workgroup order (for the check).
__global__ __launch_bounds__(256) void k_mlp_fold(const float* part, int nwg, float* grad)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p > NPARAM) return;
    float g = 0.0f;
    for (int w = 0; w < nwg; ++w) g += part[(size_t)w * NPART + p];
    grad[p] = g;
}

}  // namespace

struct fhip_mlp {
    hipStream_t stream = nullptr;
    float *master = nullptr, *m = nullptr, *v = nullptr;
    __bf16 *pack = nullptr;
    int step = 0;                       // Adam steps taken
    // data of the last upload
    float *x = nullptr, *y = nullptr, *wout = nullptr, *out = nullptr, *part = nullptr, *loss = nullptr,
          *grad = nullptr;
    int cap = 0;                        // samples the data buffers hold
    std::string err;
};

namespace {

int mfail(fhip_mlp *q, int code, const std::string &msg)
{
    if (q) q->err = msg;
    return code;
}

#define MHIP(expr)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return mfail(q, FHIP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int nwg_of(int n) { return (n + MS - 1) / MS; }

int ensure_data(fhip_mlp *q, int n)
{
    if (n <= q->cap) return FHIP_OK;
    (void)hipFree(q->x);
    (void)hipFree(q->y);
    (void)hipFree(q->out);
    (void)hipFree(q->part);
    q->x = q->y = q->out = q->part = nullptr;
    q->cap = 0;
    MHIP(hipMalloc(&q->x, sizeof(float) * (size_t)n * NIN));
    MHIP(hipMalloc(&q->y, sizeof(float) * (size_t)n * NOUT));
    MHIP(hipMalloc(&q->out, sizeof(float) * (size_t)n * NOUT));
    MHIP(hipMalloc(&q->part, sizeof(float) * (size_t)nwg_of(n) * NPART));
    q->cap = n;
    return FHIP_OK;
}

int repack(fhip_mlp *q)
{
    MHIP(hipMemsetAsync(q->pack, 0, sizeof(__bf16) * NPACK, q->stream));
    k_mlp_pack<<<(NPARAM + 255) / 256, 256, 0, q->stream>>>(q->master, q->pack);
    MHIP(hipGetLastError());
    return FHIP_OK;
}

int upload(fhip_mlp *q, int n, const float *x, const float *y, const float *wout)
{
    if (n < 1 || n > FHIP_MLP_MAX_SAMPLES) return mfail(q, FHIP_ERR_ARG, "sample count out of range");
    int rc = ensure_data(q, n);
    if (rc != FHIP_OK) return rc;
    MHIP(hipMemcpyAsync(q->x, x, sizeof(float) * (size_t)n * NIN, hipMemcpyHostToDevice, q->stream));
    if (y) MHIP(hipMemcpyAsync(q->y, y, sizeof(float) * (size_t)n * NOUT, hipMemcpyHostToDevice, q->stream));
    float w[NOUT];
    for (int o = 0; o < NOUT; ++o) w[o] = wout ? wout[o] : 1.0f;
    MHIP(hipMemcpyAsync(q->wout, w, sizeof(w), hipMemcpyHostToDevice, q->stream));
    MHIP(hipStreamSynchronize(q->stream));
    return FHIP_OK;
}

MlpArgs args(fhip_mlp *q, int n, bool train)
{
    MlpArgs a{};
    a.x = q->x;
    a.y = q->y;
    a.wout = q->wout;
    a.out = train ? nullptr : q->out;
    a.part = train ? q->part : nullptr;
    a.master = q->master;
    a.pack = q->pack;
    a.n = n;
    a.inv_n = 1.0f / (float)n;
    return a;
}

}  // namespace

extern "C" {

int fhip_mlp_nparams(void) { return NPARAM; }

const char *fhip_mlp_last_error(const fhip_mlp *q) { return q ? q->err.c_str() : "null network"; }

void fhip_mlp_destroy(fhip_mlp *q)
{
    if (!q) return;
    if (q->stream) (void)hipStreamSynchronize(q->stream);
    for (void *p : {(void *)q->master, (void *)q->m, (void *)q->v, (void *)q->pack, (void *)q->x, (void *)q->y,
                    (void *)q->wout, (void *)q->out, (void *)q->part, (void *)q->loss, (void *)q->grad})
        if (p) (void)hipFree(p);
    if (q->stream) (void)hipStreamDestroy(q->stream);
    delete q;
}

int fhip_mlp_create(uint64_t seed, fhip_mlp **out)
{
    if (!out) return FHIP_ERR_ARG;
    fhip_mlp *q = new fhip_mlp();
    *out = q;
    MHIP(hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking));
    MHIP(hipMalloc(&q->master, sizeof(float) * NPARAM));
    MHIP(hipMalloc(&q->m, sizeof(float) * NPARAM));
    MHIP(hipMalloc(&q->v, sizeof(float) * NPARAM));
    MHIP(hipMalloc(&q->pack, sizeof(__bf16) * NPACK));
    MHIP(hipMalloc(&q->wout, sizeof(float) * NOUT));
    MHIP(hipMalloc(&q->loss, sizeof(float)));
    MHIP(hipMalloc(&q->grad, sizeof(float) * NPART));
    // Glorot-uniform weights, zero biases, from a seeded host generator
    std::vector<float> h(NPARAM, 0.0f);
    std::mt19937_64 gen(seed);
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> d(-a, a);
        for (int i = 0; i < fan_out * fan_in; ++i) h[off + i] = (float)d(gen);
    };
    fill(OW1, NH, NIN);
    fill(OW2, NH, NH);
    fill(OW3, NOUT, NH);
    return fhip_mlp_set_params(q, h.data());
}

int fhip_mlp_set_params(fhip_mlp *q, const float *params)
{
    if (!q || !params) return mfail(q, FHIP_ERR_ARG, "null argument");
    MHIP(hipMemcpyAsync(q->master, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, q->stream));
    MHIP(hipMemsetAsync(q->m, 0, sizeof(float) * NPARAM, q->stream));
    MHIP(hipMemsetAsync(q->v, 0, sizeof(float) * NPARAM, q->stream));
    q->step = 0;
    int rc = repack(q);
    if (rc != FHIP_OK) return rc;
    MHIP(hipStreamSynchronize(q->stream));
    return FHIP_OK;
}

int fhip_mlp_get_params(fhip_mlp *q, float *params)
{
    if (!q || !params) return mfail(q, FHIP_ERR_ARG, "null argument");
    MHIP(hipMemcpyAsync(params, q->master, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, q->stream));
    MHIP(hipStreamSynchronize(q->stream));
    return FHIP_OK;
}

int fhip_mlp_forward(fhip_mlp *q, int n, const float *x, float *y)
{
    if (!q || !x || !y) return mfail(q, FHIP_ERR_ARG, "null argument");
    int rc = upload(q, n, x, nullptr, nullptr);
    if (rc != FHIP_OK) return rc;
    k_mlp<0><<<nwg_of(n), WL, 0, q->stream>>>(args(q, n, false));
    MHIP(hipGetLastError());
    MHIP(hipMemcpyAsync(y, q->out, sizeof(float) * (size_t)n * NOUT, hipMemcpyDeviceToHost, q->stream));
    MHIP(hipStreamSynchronize(q->stream));
    return FHIP_OK;
}

int fhip_mlp_gradient(fhip_mlp *q, int n, const float *x, const float *y, const float *wout, float *loss,
                      float *grad)
{
    if (!q || !x || !y || !grad) return mfail(q, FHIP_ERR_ARG, "null argument");
    int rc = upload(q, n, x, y, wout);
    if (rc != FHIP_OK) return rc;
    const int nwg = nwg_of(n);
    k_mlp<1><<<nwg, WL, 0, q->stream>>>(args(q, n, true));
    MHIP(hipGetLastError());
    k_mlp_fold<<<(NPART + 255) / 256, 256, 0, q->stream>>>(q->part, nwg, q->grad);
    MHIP(hipGetLastError());
    std::vector<float> h(NPART);
    MHIP(hipMemcpyAsync(h.data(), q->grad, sizeof(float) * NPART, hipMemcpyDeviceToHost, q->stream));
    MHIP(hipStreamSynchronize(q->stream));
    std::memcpy(grad, h.data(), sizeof(float) * NPARAM);
    if (loss) *loss = h[NPARAM];
    return FHIP_OK;
}

int fhip_mlp_train(fhip_mlp *q, int n, const float *x, const float *y, const float *wout, int iters, float lr,
                   float *loss_history)
{
    if (!q || !x || !y) return mfail(q, FHIP_ERR_ARG, "null argument");
    if (iters < 1) return mfail(q, FHIP_ERR_ARG, "iters must be >= 1");
    if (!(lr > 0.0f)) return mfail(q, FHIP_ERR_ARG, "learning rate must be positive");
    int rc = upload(q, n, x, y, wout);
    if (rc != FHIP_OK) return rc;
    const int nwg = nwg_of(n);
    float *hist = nullptr;
    if (loss_history) MHIP(hipMalloc(&hist, sizeof(float) * (size_t)iters));
    const float b1 = 0.9f, b2 = 0.999f;
    for (int it = 0; it < iters; ++it) {
        k_mlp<1><<<nwg, WL, 0, q->stream>>>(args(q, n, true));
        q->step += 1;
        AdamArgs aa{};
        aa.master = q->master;
        aa.m = q->m;
        aa.v = q->v;
        aa.pack = q->pack;
        aa.part = q->part;
        aa.nwg = nwg;
        aa.lr = lr;
        aa.beta1 = b1;
        aa.beta2 = b2;
        aa.eps = 1e-8f;
        aa.c1 = 1.0f - std::pow(b1, (float)q->step);
        aa.c2 = 1.0f - std::pow(b2, (float)q->step);
        aa.loss = hist ? hist + it : nullptr;
        k_mlp_adam<<<(NPART + 255) / 256, 256, 0, q->stream>>>(aa);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            if (hist) (void)hipFree(hist);
            return mfail(q, FHIP_ERR_HIP, std::string("mlp training launch: ") + hipGetErrorString(e));
        }
    }
    hipError_t e = hipSuccess;
    if (hist) e = hipMemcpyAsync(loss_history, hist, sizeof(float) * (size_t)iters, hipMemcpyDeviceToHost, q->stream);
    hipError_t e2 = hipStreamSynchronize(q->stream);
    if (hist) (void)hipFree(hist);
    if (e != hipSuccess || e2 != hipSuccess)
        return mfail(q, FHIP_ERR_HIP, std::string("mlp training: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    return FHIP_OK;
}

}  // extern "C"
