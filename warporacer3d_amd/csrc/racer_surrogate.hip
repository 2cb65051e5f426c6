// A learned dynamics surrogate (no reference counterpart): a residual MLP that
// predicts one environment step's pose change (racer/sim.mojo:184-219's RK4
// integration) from 11 inputs, trained on transitions the simulator emits and
// used for imagined steps with the exact RK4 step as its fallback.
//
// Network: 11 inputs -> 128 -> 128 -> 128 -> 5, tanh on the hidden layers.
// Layer l's weights W_l are [w_(l+1), w_l + 1] with the bias as the last
// column, fed by a constant-1 row under each activation (the layout the
// policy uses, racer/ppo.mojo:40-44), so every product is one GEMM of the
// project's own kernels: bf16 MFMA for the hidden layers in precision 1
// with the branch-free racer_tanh_mlp (training and imagined steps alike),
// float32 with k in order and racer_tanh in precision 0 (bitwise the
// oracle's), the output layer in float32 always. Targets are the displacement in the car's frame
// (forward, left), the heading, speed and steer changes, normalised by their
// dataset mean and deviation; the loss is their masked mean squared error.
#include "racer_internal.h"

#include <cmath>

namespace {

constexpr int S_IN = 11, S_H = 128, S_OUT = 5, S_LAYERS = 4;
constexpr int S_WIDTH[S_LAYERS + 1] = {S_IN, S_H, S_H, S_H, S_OUT};
// stats: in_mean[11], in_inv_std[11], out_mean[5], out_std[5], out_inv_std[5]
constexpr int ST_IN_MEAN = 0, ST_IN_INV = 11, ST_OUT_MEAN = 22, ST_OUT_STD = 27, ST_OUT_INV = 32, ST_LEN = 37;
constexpr int SUM_BLOCKS_S = 512;

// W ~ N(0, 1 / fan_in) with the bias column zero.
__global__ void k_sur_init(float* theta, int64_t P, const int64_t* off, int seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    int l = 0;
    while (l + 1 < S_LAYERS && i >= off[l + 1]) ++l;
    const int cols = S_WIDTH[l] + 1;
    const int64_t k = (i - off[l]) % cols;
    float z[4];
    racer_normal4(racer_random((uint64_t)(int64_t)seed, (uint64_t)i, 7), z);
    theta[i] = k == cols - 1 ? 0.0f : z[0] * sqrtf(1.0f / (float)S_WIDTH[l]);
}

// wt_l[k, j] = W_l[j, k] for k < w_l (the bias column is not propagated back).
__global__ void k_sur_transpose(const float* theta, const int64_t* off, const int64_t* offt, float* wt) {
    const int l = blockIdx.y;
    const int rows = S_WIDTH[l + 1], cols = S_WIDTH[l];
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)rows * cols) return;
    const int k = (int)(idx / rows), j = (int)(idx % rows);
    wt[offt[l] + idx] = theta[off[l] + (int64_t)j * (cols + 1) + k];
}

// Minibatch j of M interleaved ones: sample s = b M + j, normalised.
__global__ void k_sur_gather(const float* __restrict__ feat, const float* __restrict__ tgt,
                             const float* __restrict__ mask, int64_t cap, int64_t M, int64_t j, int64_t first,
                             int64_t B, const float* __restrict__ st, float* __restrict__ x, float* __restrict__ y,
                             float* __restrict__ m) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int64_t s = first + b * M + j;
#pragma unroll
    for (int r = 0; r < S_IN; ++r) x[r * B + b] = (feat[r * cap + s] - st[ST_IN_MEAN + r]) * st[ST_IN_INV + r];
    x[S_IN * B + b] = 1.0f;
#pragma unroll
    for (int o = 0; o < S_OUT; ++o) y[o * B + b] = (tgt[o * cap + s] - st[ST_OUT_MEAN + o]) * st[ST_OUT_INV + o];
    m[b] = mask[s];
}

// d loss / d out for loss = sum_b m_b sum_o (out - y)^2 / (5 B).
__global__ void k_sur_loss_grad(const float* __restrict__ out, const float* __restrict__ y,
                                const float* __restrict__ m, int64_t B, float* __restrict__ d) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float denom = (float)(S_OUT * B);
#pragma unroll
    for (int o = 0; o < S_OUT; ++o) {
        const float e = out[o * B + b] - y[o * B + b];
        d[o * B + b] = ((e + e) * m[b]) / denom;
    }
}

// Per-block float64 partial sums of m_b sum_o (out - y)^2 and of m_b.
__global__ void k_sur_sq_partial(const float* __restrict__ out, const float* __restrict__ y,
                                 const float* __restrict__ m, int64_t B, double* __restrict__ partial) {
    __shared__ double se[256], sm[256];
    double e2 = 0.0, mm = 0.0;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < B; b += (int64_t)gridDim.x * blockDim.x) {
        double row = 0.0;
        for (int o = 0; o < S_OUT; ++o) {
            const double e = (double)out[o * B + b] - (double)y[o * B + b];
            row += e * e;
        }
        e2 += (double)m[b] * row;
        mm += (double)m[b];
    }
    se[threadIdx.x] = e2;
    sm[threadIdx.x] = mm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            se[threadIdx.x] += se[threadIdx.x + w];
            sm[threadIdx.x] += sm[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = se[0];
        partial[2 * blockIdx.x + 1] = sm[0];
    }
}

// acc += the partial sums, in block order (one thread: a fixed order).
__global__ void k_sur_sq_final(const double* __restrict__ partial, int blocks, double* acc) {
    double e2 = acc[0], mm = acc[1];
    for (int k = 0; k < blocks; ++k) {
        e2 += partial[2 * k];
        mm += partial[2 * k + 1];
    }
    acc[0] = e2;
    acc[1] = mm;
}

// Masked float64 moments of `rows` rows of a [rows, ld] array over `count`
// columns: per-block partials of (sum m v, sum m v^2, sum m) per row.
__global__ void k_sur_moments_partial(const float* __restrict__ a, int64_t ld, int rows, int64_t count,
                                      const float* __restrict__ mask, double* __restrict__ partial) {
    const int r = blockIdx.y;
    __shared__ double s1[256], s2[256], s3[256];
    double v1 = 0.0, v2 = 0.0, v3 = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < count;
         c += (int64_t)gridDim.x * blockDim.x) {
        const double w = (double)mask[c];
        const double v = (double)a[r * ld + c];
        v1 += w * v;
        v2 += w * v * v;
        v3 += w;
    }
    s1[threadIdx.x] = v1;
    s2[threadIdx.x] = v2;
    s3[threadIdx.x] = v3;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
            s3[threadIdx.x] += s3[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* p = partial + ((int64_t)r * gridDim.x + blockIdx.x) * 3;
        p[0] = s1[0];
        p[1] = s2[0];
        p[2] = s3[0];
    }
}

// The hidden layers' weights in v_mfma_f32_32x32x16_bf16 A-operand order for
// the fused step: block (layer l, tile t, k-step q) holds lane l's 8 values
// of row 32 t + l % 32, k = 16 q + 8 (l / 32) .. + 7 of W_l (the bias column
// at k = w_l), zeros past it; the same bf16 rounding as the packed copy.
__global__ void k_sur_pack_frags(const float* __restrict__ theta, const int64_t* off, uint16_t* __restrict__ frags) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)RACER_SF_BLOCKS * 512) return;
    const int block = (int)(idx / 512), within = (int)(idx % 512);
    const int l = block >= RACER_SF_BASE2 ? 2 : (block >= RACER_SF_BASE1 ? 1 : 0);
    const int steps = l == 0 ? RACER_SF_STEPS0 : RACER_SF_STEPS;
    const int base = l == 0 ? 0 : (l == 1 ? RACER_SF_BASE1 : RACER_SF_BASE2);
    const int local = block - base, tile = local / steps, step = local % steps;
    const int lane = within / 8, e = within % 8;
    const int row = 32 * tile + lane % 32, k = 16 * step + (lane / 32) * 8 + e;
    const int cols = S_WIDTH[l] + 1;
    frags[idx] = k < cols ? __builtin_bit_cast(uint16_t, (__bf16)theta[off[l] + (int64_t)row * cols + k]) : (uint16_t)0;
}

// Adam (Kingma and Ba) without clipping; c1, c2 are the bias corrections.
__global__ void k_sur_adam(float* theta, int64_t P, const float* grad, float* first, float* second, float lr, float c1,
                           float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float g = grad[i];
    const float m = 0.9f * first[i] + 0.1f * g;
    const float v = 0.999f * second[i] + (0.001f * g) * g;
    first[i] = m;
    second[i] = v;
    theta[i] = theta[i] - (lr * (m / c1)) / (sqrtf(v / c2) + 1e-8f);
}

}  // namespace

struct racer_surrogate {
    racer_device* dev = nullptr;
    int precision = 1, seed = 0, step = 0;
    int64_t P = 0, PT = 0;
    int64_t off[S_LAYERS + 1] = {}, offt[S_LAYERS] = {}, off16[S_LAYERS] = {}, ld16[S_LAYERS] = {};
    int64_t* doff = nullptr;  // device copies of off and offt
    racer_buf *theta = nullptr, *grad = nullptr, *first = nullptr, *second = nullptr, *wt = nullptr,
              *stats = nullptr;
    uint16_t* theta_bf16 = nullptr;
    uint16_t* frags = nullptr;  // hidden layers in MFMA fragment order (the fused step)
    bool stale = true;
    bool fused = true;  // bf16 imagined steps in one launch
    int64_t cols = 0;  // capacity in samples of the working buffers
    int64_t ones_for = -1;  // the sample count whose constant rows are in place
    racer_buf* act[S_LAYERS + 1] = {};
    racer_buf* delta[S_LAYERS + 1] = {};
    racer_buf *y = nullptr, *mask = nullptr;
    double* partial = nullptr;  // 2 SUM_BLOCKS_S doubles for the loss, 3 x 16 x SUM_BLOCKS_S for moments
    double* acc = nullptr;      // loss accumulator (2 doubles)
};

struct racer_dataset {
    racer_device* dev = nullptr;
    int64_t cap = 0, count = 0;
    racer_buf *feat = nullptr, *tgt = nullptr, *mask = nullptr, *prev = nullptr;
};

// ---------------------------------------------------------------- surrogate
extern "C" void racer_surrogate_destroy(racer_surrogate* s) {
    if (!s) return;
    if (s->dev) (void)hipStreamSynchronize(s->dev->stream);
    racer_buf* bufs[] = {s->theta, s->grad, s->first, s->second, s->wt, s->stats, s->y, s->mask};
    for (auto* b : bufs) racer_free(b);
    for (int l = 0; l <= S_LAYERS; ++l) {
        racer_free(s->act[l]);
        racer_free(s->delta[l]);
    }
    if (s->theta_bf16) (void)hipFree(s->theta_bf16);
    if (s->frags) (void)hipFree(s->frags);
    if (s->doff) (void)hipFree(s->doff);
    if (s->partial) (void)hipFree(s->partial);
    if (s->acc) (void)hipFree(s->acc);
    racer_device* dev = s->dev;
    delete s;
    racer_device_release(dev);
}

static int sur_refresh(racer_surrogate* s) {
    if (!s->stale) return 0;
    for (int l = 0; l < S_LAYERS; ++l) {
        const int rows = S_WIDTH[l + 1], cols = S_WIDTH[l];
        if (s->precision == 1 && l < S_LAYERS - 1 &&
            racer_pack_bf16(s->dev, s->theta->ptr + s->off[l], cols + 1, s->theta_bf16 + s->off16[l], s->ld16[l],
                            rows, cols + 1))
            return 1;
    }
    hipLaunchKernelGGL(k_sur_pack_frags, dim3((RACER_SF_BLOCKS * 512 + 255) / 256), dim3(256), 0, s->dev->stream,
                       s->theta->ptr, s->doff, s->frags);
    {
        racer_timed timer(s->dev, "surrogate_transpose");
        hipLaunchKernelGGL(k_sur_transpose, dim3((unsigned)((S_H * S_H + 255) / 256), S_LAYERS), dim3(256), 0,
                           s->dev->stream, s->theta->ptr, s->doff, s->doff + S_LAYERS + 1, s->wt->ptr);
    }
    if (racer_launch_check("surrogate transpose")) return 1;
    s->stale = false;
    return 0;
}

extern "C" int racer_surrogate_create(racer_device* dev, int seed, int precision, int64_t cols, racer_surrogate** out) {
    racer_clear_error();
    *out = nullptr;
    RACER_REQUIRE(dev, "null device");
    RACER_REQUIRE(precision == 0 || precision == 1, "surrogate precision must be 0 (float32) or 1 (bf16)");
    RACER_REQUIRE(cols >= 1, "the surrogate needs room for at least one sample");
    auto* s = new racer_surrogate();
    s->dev = dev;
    racer_device_retain(dev);
    s->precision = precision;
    s->seed = seed;
    s->cols = cols;
    int64_t p = 0, pt = 0, p16 = 0;
    for (int l = 0; l < S_LAYERS; ++l) {
        s->off[l] = p;
        s->offt[l] = pt;
        s->off16[l] = p16;
        s->ld16[l] = racer_pad8(S_WIDTH[l] + 1);
        p += (int64_t)S_WIDTH[l + 1] * (S_WIDTH[l] + 1);
        pt += (int64_t)S_WIDTH[l + 1] * S_WIDTH[l];
        p16 += (int64_t)S_WIDTH[l + 1] * s->ld16[l];
    }
    s->off[S_LAYERS] = p;
    s->P = p;
    s->PT = pt;
    bool bad = racer_alloc(dev, (size_t)p, 0.0f, &s->theta) || racer_alloc(dev, (size_t)p, 0.0f, &s->grad) ||
               racer_alloc(dev, (size_t)p, 0.0f, &s->first) || racer_alloc(dev, (size_t)p, 0.0f, &s->second) ||
               racer_alloc(dev, (size_t)pt, 0.0f, &s->wt) || racer_alloc(dev, ST_LEN, 0.0f, &s->stats) ||
               racer_alloc(dev, (size_t)(S_OUT * cols), 0.0f, &s->y) || racer_alloc(dev, (size_t)cols, 0.0f, &s->mask);
    for (int l = 0; l <= S_LAYERS && !bad; ++l) {
        const int64_t rows = l < S_LAYERS ? S_WIDTH[l] + 1 : S_OUT;
        bad = racer_alloc(dev, (size_t)(rows * cols), 0.0f, &s->act[l]) ||
              racer_alloc(dev, (size_t)(S_WIDTH[l] * cols), 0.0f, &s->delta[l]);
    }
    if (bad || hipMalloc(&s->theta_bf16, (size_t)p16 * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc(&s->frags, (size_t)RACER_SF_BLOCKS * 512 * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc(&s->doff, (size_t)(2 * S_LAYERS + 1) * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&s->partial, (size_t)3 * 16 * SUM_BLOCKS_S * sizeof(double)) != hipSuccess ||
        hipMalloc(&s->acc, 2 * sizeof(double)) != hipSuccess) {
        if (!bad) racer_set_error("surrogate allocation failed");
        racer_surrogate_destroy(s);
        return 1;
    }
    int64_t hoff[2 * S_LAYERS + 1];
    for (int l = 0; l <= S_LAYERS; ++l) hoff[l] = s->off[l];
    for (int l = 0; l < S_LAYERS; ++l) hoff[S_LAYERS + 1 + l] = s->offt[l];
    RACER_TRY(hipMemcpy(s->doff, hoff, sizeof(hoff), hipMemcpyHostToDevice));
    // identity normalisation until fitted
    float st[ST_LEN];
    for (int r = 0; r < ST_LEN; ++r) st[r] = (r >= ST_IN_INV && r < ST_OUT_MEAN) || r >= ST_OUT_STD ? 1.0f : 0.0f;
    if (racer_write(s->stats, st)) return 1;
    hipLaunchKernelGGL(k_sur_init, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, dev->stream, s->theta->ptr, p,
                       s->doff, seed);
    if (racer_launch_check("surrogate init")) return 1;
    RACER_TRY(hipStreamSynchronize(dev->stream));
    *out = s;
    return 0;
}

extern "C" racer_buf* racer_surrogate_theta(racer_surrogate* s) { return s->theta; }
extern "C" racer_buf* racer_surrogate_grad(racer_surrogate* s) { return s->grad; }
extern "C" racer_buf* racer_surrogate_stats(racer_surrogate* s) { return s->stats; }
extern "C" int64_t racer_surrogate_params(const racer_surrogate* s) { return s->P; }
extern "C" int racer_surrogate_mark_updated(racer_surrogate* s) {
    s->stale = true;
    return 0;
}

// The constant-1 rows under the inputs of layers 1 .. 3 for B samples.
static int sur_ones(racer_surrogate* s, int64_t B) {
    if (s->ones_for == B) return 0;
    for (int l = 1; l < S_LAYERS; ++l)
        if (racer_fill(s->dev, s->act[l]->ptr + (int64_t)S_WIDTH[l] * B, (size_t)B, 1.0f)) return 1;
    s->ones_for = B;
    return 0;
}

// act[0] holds B normalised inputs (with the constant row); act[4] gets the outputs.
static int sur_forward(racer_surrogate* s, int64_t B) {
    if (sur_refresh(s) || sur_ones(s, B)) return 1;
    for (int l = 0; l < S_LAYERS; ++l) {
        const int M = S_WIDTH[l + 1], K = S_WIDTH[l] + 1;
        const int epi = l < S_LAYERS - 1 ? RACER_EPI_TANH : RACER_EPI_NONE;
        int rc;
        if (s->precision == 1 && l < S_LAYERS - 1)
            rc = racer_gemm_nn_bf16(s->dev, s->theta_bf16 + s->off16[l], s->ld16[l], s->act[l]->ptr, B,
                                    s->act[l + 1]->ptr, B, M, (int)B, K, RACER_EPI_TANH_MLP);
        else
            rc = racer_gemm_nn_f32(s->dev, s->theta->ptr + s->off[l], K, s->act[l]->ptr, B, s->act[l + 1]->ptr, B, M,
                                   (int)B, K, epi, nullptr, 0);
        if (rc) return 1;
    }
    return 0;
}

// Gradient of the masked loss for the B samples in act[0], y, mask (into grad).
static int sur_backward(racer_surrogate* s, int64_t B) {
    racer_device* dev = s->dev;
    const unsigned blocks = (unsigned)((B + 255) / 256);
    {
        racer_timed timer(dev, "surrogate_loss");
        hipLaunchKernelGGL(k_sur_loss_grad, dim3(blocks), dim3(256), 0, dev->stream, s->act[S_LAYERS]->ptr,
                           s->y->ptr, s->mask->ptr, B, s->delta[S_LAYERS]->ptr);
    }
    if (racer_launch_check("surrogate loss")) return 1;
    for (int l = S_LAYERS - 1; l >= 0; --l) {
        const int M = S_WIDTH[l + 1], N = S_WIDTH[l] + 1;
        int rc;
        if (s->precision == 1 && l < S_LAYERS - 1)
            rc = racer_gemm_nt_bf16(dev, s->delta[l + 1]->ptr, B, s->act[l]->ptr, B, s->grad->ptr + s->off[l], N, M,
                                    N, (int)B);
        else
            rc = racer_gemm_nt_f32(dev, s->delta[l + 1]->ptr, B, s->act[l]->ptr, B, s->grad->ptr + s->off[l], N, M, N,
                                   (int)B);
        if (rc) return 1;
        if (l > 0 && racer_gemm_nn_f32(dev, s->wt->ptr + s->offt[l], M, s->delta[l + 1]->ptr, B, s->delta[l]->ptr, B,
                                       S_WIDTH[l], (int)B, M, RACER_EPI_MASK, s->act[l]->ptr, B))
            return 1;
    }
    return 0;
}

static int sur_adam(racer_surrogate* s, float lr) {
    s->step += 1;
    const float c1 = (float)(1.0 - std::pow(0.9, (double)s->step));
    const float c2 = (float)(1.0 - std::pow(0.999, (double)s->step));
    {
        racer_timed timer(s->dev, "surrogate_adam");
        hipLaunchKernelGGL(k_sur_adam, dim3((unsigned)((s->P + 255) / 256)), dim3(256), 0, s->dev->stream,
                           s->theta->ptr, s->P, s->grad->ptr, s->first->ptr, s->second->ptr, lr, c1, c2);
    }
    s->stale = true;
    return racer_launch_check("surrogate adam");
}

// The loss and gradient of one batch given directly: x [12, B] normalised
// inputs with the constant row, y [5, B] normalised targets, mask [B].
extern "C" int racer_surrogate_batch_gradient(racer_surrogate* s, const racer_view* x, const racer_view* y,
                                              const racer_view* mask) {
    racer_clear_error();
    RACER_REQUIRE(s && x && y && mask, "null argument");
    const int64_t B = x->cols;
    RACER_REQUIRE(B >= 1 && B <= s->cols, "batch larger than the surrogate's working buffers");
    if (racer_check_view(*x, S_IN + 1, B, "surrogate inputs") || racer_check_view(*y, S_OUT, B, "surrogate targets") ||
        racer_check_view(*mask, 1, B, "surrogate mask"))
        return 1;
    RACER_REQUIRE(x->buf->len >= (size_t)((S_IN + 1) * B), "inputs");
    if (racer_copy_raw(s->dev, s->act[0]->ptr, view_ptr(*x), (size_t)((S_IN + 1) * B)) ||
        racer_copy_raw(s->dev, s->y->ptr, view_ptr(*y), (size_t)(S_OUT * B)) ||
        racer_copy_raw(s->dev, s->mask->ptr, view_ptr(*mask), (size_t)B))
        return 1;
    return sur_forward(s, B) || sur_backward(s, B);
}

// Outputs (normalised) of the last forward: [5, B] in the first 5 B floats.
extern "C" racer_buf* racer_surrogate_outputs(racer_surrogate* s) { return s->act[S_LAYERS]; }

// ---------------------------------------------------------------- dataset
extern "C" void racer_dataset_destroy(racer_dataset* d) {
    if (!d) return;
    racer_buf* bufs[] = {d->feat, d->tgt, d->mask, d->prev};
    for (auto* b : bufs) racer_free(b);
    racer_device* dev = d->dev;
    delete d;
    racer_device_release(dev);
}

extern "C" int racer_dataset_create(racer_device* dev, int64_t capacity, racer_dataset** out) {
    racer_clear_error();
    *out = nullptr;
    RACER_REQUIRE(dev && capacity >= 1, "bad dataset size");
    auto* d = new racer_dataset();
    d->dev = dev;
    racer_device_retain(dev);
    d->cap = capacity;
    if (racer_alloc(dev, (size_t)((S_IN + 1) * capacity), 0.0f, &d->feat) ||
        racer_alloc(dev, (size_t)(S_OUT * capacity), 0.0f, &d->tgt) ||
        racer_alloc(dev, (size_t)capacity, 0.0f, &d->mask)) {
        racer_dataset_destroy(d);
        return 1;
    }
    *out = d;
    return 0;
}

extern "C" int64_t racer_dataset_count(const racer_dataset* d) { return d->count; }
extern "C" racer_buf* racer_dataset_features(racer_dataset* d) { return d->feat; }
extern "C" racer_buf* racer_dataset_targets(racer_dataset* d) { return d->tgt; }
extern "C" racer_buf* racer_dataset_mask(racer_dataset* d) { return d->mask; }

// `steps` Runner steps (main.mojo:183-191) of the sim with the policy, each
// recording its cars' transitions: the inputs before the step, the targets
// and mask after it. Step t uses Philox offset t0 + t.
extern "C" int racer_dataset_record(racer_dataset* d, racer_sim* sim, racer_policy* p, racer_acts* acts,
                                    racer_buf* scratch, int steps, int64_t t0) {
    racer_clear_error();
    RACER_REQUIRE(d && sim && p && acts && scratch, "null argument");
    const int64_t n = sim->n;
    RACER_REQUIRE(d->count + n * steps <= d->cap, "the dataset is full");
    RACER_REQUIRE(scratch->len >= (size_t)(2 * n), "scratch must hold 2 x cars floats");
    RACER_REQUIRE(acts->cols == n && acts->hidden == p->hidden, "activations must have one column per car");
    if (!d->prev || d->prev->len < (size_t)(RACER_STATE * n)) {
        racer_free(d->prev);
        d->prev = nullptr;
        if (racer_alloc(d->dev, (size_t)(RACER_STATE * n), 0.0f, &d->prev)) return 1;
    }
    for (int t = 0; t < steps; ++t) {
        if (racer_policy_act(p, sim->obs->ptr, n, acts, (int)n, sim->actions->ptr, scratch->ptr, scratch->ptr + n, 0,
                             t0 + t, sim->seed) ||
            racer_launch_surrogate_features(sim, sim->actions->ptr, n, d->feat->ptr + d->count, d->cap, nullptr,
                                            nullptr) ||
            racer_copy_raw(d->dev, d->prev->ptr, sim->state->ptr, (size_t)(RACER_STATE * n)) ||
            racer_launch_step(sim, sim->actions->ptr, sim->reward->ptr, sim->done->ptr, sim->obs->ptr, n) ||
            racer_launch_surrogate_targets(sim, d->prev->ptr, d->tgt->ptr + d->count, d->cap, d->mask->ptr + d->count))
            return 1;
        d->count += n;
    }
    return 0;
}

// Masked mean and deviation of the inputs and targets over the dataset.
extern "C" int racer_surrogate_fit_normalization(racer_surrogate* s, racer_dataset* d) {
    racer_clear_error();
    RACER_REQUIRE(d->count >= 2, "the dataset is empty");
    racer_device* dev = s->dev;
    double sums[2][16][3] = {};
    const racer_buf* srcs[2] = {d->feat, d->tgt};
    const int rows[2] = {S_IN, S_OUT};
    for (int w = 0; w < 2; ++w) {
        hipLaunchKernelGGL(k_sur_moments_partial, dim3(SUM_BLOCKS_S, rows[w]), dim3(256), 0, dev->stream,
                           srcs[w]->ptr, d->cap, rows[w], d->count, d->mask->ptr, s->partial);
        if (racer_launch_check("surrogate moments")) return 1;
        std::vector<double> part((size_t)3 * rows[w] * SUM_BLOCKS_S);
        RACER_TRY(hipMemcpyAsync(part.data(), s->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 dev->stream));
        RACER_TRY(hipStreamSynchronize(dev->stream));
        for (int r = 0; r < rows[w]; ++r)
            for (int k = 0; k < SUM_BLOCKS_S; ++k)
                for (int q = 0; q < 3; ++q) sums[w][r][q] += part[((size_t)r * SUM_BLOCKS_S + k) * 3 + q];
    }
    RACER_REQUIRE(sums[0][0][2] >= 1.0, "no transition in the dataset is usable (all masked)");
    float st[ST_LEN];
    for (int w = 0; w < 2; ++w)
        for (int r = 0; r < rows[w]; ++r) {
            const double cnt = sums[w][r][2];
            const double mean = sums[w][r][0] / cnt;
            const double var = std::max(sums[w][r][1] / cnt - mean * mean, 1e-12);
            const double sd = std::sqrt(var);
            if (w == 0) {
                st[ST_IN_MEAN + r] = (float)mean;
                st[ST_IN_INV + r] = (float)(1.0 / sd);
            } else {
                st[ST_OUT_MEAN + r] = (float)mean;
                st[ST_OUT_STD + r] = (float)sd;
                st[ST_OUT_INV + r] = (float)(1.0 / sd);
            }
        }
    return racer_write(s->stats, st);
}

// `epochs` passes over the dataset in minibatches of `batch` interleaved
// samples (sample b M + j in minibatch j of M), one Adam step each.
extern "C" int racer_surrogate_train(racer_surrogate* s, racer_dataset* d, int epochs, int64_t batch, float lr) {
    racer_clear_error();
    RACER_REQUIRE(batch >= 1 && batch <= s->cols, "batch larger than the surrogate's working buffers");
    const int64_t M = d->count / batch;
    RACER_REQUIRE(M >= 1, "fewer samples than one batch");
    for (int e = 0; e < epochs; ++e)
        for (int64_t j = 0; j < M; ++j) {
            {
                racer_timed timer(s->dev, "surrogate_gather");
                hipLaunchKernelGGL(k_sur_gather, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s->dev->stream,
                                   d->feat->ptr, d->tgt->ptr, d->mask->ptr, d->cap, M, j, (int64_t)0, batch,
                                   s->stats->ptr, s->act[0]->ptr, s->y->ptr, s->mask->ptr);
            }
            if (racer_launch_check("surrogate gather") || sur_forward(s, batch) || sur_backward(s, batch) ||
                sur_adam(s, lr))
                return 1;
        }
    return 0;
}

// Mean over the dataset's usable samples of the squared error summed over
// the 5 normalised outputs.
extern "C" int racer_surrogate_loss(racer_surrogate* s, racer_dataset* d, double* loss) {
    racer_clear_error();
    racer_device* dev = s->dev;
    RACER_TRY(hipMemsetAsync(s->acc, 0, 2 * sizeof(double), dev->stream));
    for (int64_t first = 0; first < d->count; first += s->cols) {
        const int64_t B = std::min<int64_t>(s->cols, d->count - first);
        hipLaunchKernelGGL(k_sur_gather, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, dev->stream, d->feat->ptr,
                           d->tgt->ptr, d->mask->ptr, d->cap, (int64_t)1, (int64_t)0, first, B, s->stats->ptr,
                           s->act[0]->ptr, s->y->ptr, s->mask->ptr);
        if (racer_launch_check("surrogate gather") || sur_forward(s, B)) return 1;
        const int blocks = (int)std::min<int64_t>(SUM_BLOCKS_S, (B + 255) / 256);
        hipLaunchKernelGGL(k_sur_sq_partial, dim3(blocks), dim3(256), 0, dev->stream, s->act[S_LAYERS]->ptr,
                           s->y->ptr, s->mask->ptr, B, s->partial);
        hipLaunchKernelGGL(k_sur_sq_final, dim3(1), dim3(1), 0, dev->stream, s->partial, blocks, s->acc);
        if (racer_launch_check("surrogate loss sum")) return 1;
    }
    double acc[2];
    RACER_TRY(hipMemcpyAsync(acc, s->acc, sizeof(acc), hipMemcpyDeviceToHost, dev->stream));
    RACER_TRY(hipStreamSynchronize(dev->stream));
    *loss = acc[1] > 0.0 ? acc[0] / acc[1] : 0.0;
    return 0;
}

// ---------------------------------------------------------------- imagined steps
// Sim.step with the surrogate in place of the RK4 integration for the cars
// whose prediction stays at least min_clearance from the walls on the same
// layer (the others take the exact step); the lidar is exact. fallback, when
// given, receives 1 per car that took the exact step.
int racer_surrogate_step_raw(racer_surrogate* s, racer_sim* sim, float min_clearance, float* fallback) {
    const int64_t n = sim->n;
    RACER_REQUIRE(n <= s->cols, "more cars than the surrogate's working buffers");
    RACER_REQUIRE(s->dev == sim->dev, "the surrogate lives on another device");
    const float* st = s->stats->ptr;
    if (s->precision == 1 && s->fused) {
        if (sur_refresh(s)) return 1;
        return racer_launch_surrogate_fused(sim, s->frags, s->theta->ptr + s->off[S_LAYERS - 1], st, min_clearance,
                                            fallback);
    }
    if (racer_launch_surrogate_features(sim, sim->actions->ptr, n, s->act[0]->ptr, n, st + ST_IN_MEAN,
                                        st + ST_IN_INV) ||
        sur_forward(s, n))
        return 1;
    return racer_launch_surrogate_step(sim, s->act[S_LAYERS]->ptr, n, st + ST_OUT_MEAN, st + ST_OUT_STD,
                                       min_clearance, fallback);
}

// Whether bf16 imagined steps run as one fused launch (the default) or as
// the separate inputs, products and step kernels; both give the same result.
extern "C" int racer_surrogate_set_fused(racer_surrogate* s, int fused) {
    racer_clear_error();
    RACER_REQUIRE(s, "null surrogate");
    s->fused = fused != 0;
    return 0;
}

extern "C" int racer_surrogate_step(racer_surrogate* s, racer_sim* sim, float min_clearance, racer_buf* fallback) {
    racer_clear_error();
    RACER_REQUIRE(s && sim, "null argument");
    RACER_REQUIRE(!fallback || fallback->len >= (size_t)sim->n, "fallback must hold one float per car");
    return racer_surrogate_step_raw(s, sim, min_clearance, fallback ? fallback->ptr : nullptr);
}

extern "C" int racer_surrogate_imagine(racer_surrogate* s, racer_sim* sim, const racer_view* actions_p, int steps,
                                       float min_clearance, racer_buf* fallback) {
    racer_clear_error();
    RACER_REQUIRE(s && sim && actions_p, "null argument");
    RACER_REQUIRE(steps >= 1, "imagine at least one step");
    RACER_REQUIRE(s->precision == 1 && s->fused, "open-loop imagined steps run the fused bf16 surrogate step");
    RACER_REQUIRE(sim->n <= s->cols, "more cars than the surrogate's working buffers");
    RACER_REQUIRE(s->dev == sim->dev, "the surrogate lives on another device");
    RACER_REQUIRE(!fallback || fallback->len >= (size_t)sim->n, "fallback must hold one float per car");
    if (racer_check_view(*actions_p, 2 * (int64_t)steps, sim->n, "actions")) return 1;
    if (sur_refresh(s)) return 1;
    return racer_launch_surrogate_imagine(sim, view_ptr(*actions_p), steps, s->frags,
                                          s->theta->ptr + s->off[S_LAYERS - 1], s->stats->ptr, min_clearance,
                                          fallback ? fallback->ptr : nullptr);
}

// Runner.step (main.mojo:183-191) with the surrogate step.
extern "C" int racer_surrogate_runner_step(racer_surrogate* s, racer_sim* sim, racer_policy* p, racer_acts* acts,
                                           racer_buf* scratch, int deterministic, int64_t t, float min_clearance,
                                           racer_buf* fallback) {
    racer_clear_error();
    RACER_REQUIRE(s && sim && p && acts && scratch, "null argument");
    const int64_t n = sim->n;
    RACER_REQUIRE(p->beams == sim->beams && acts->cols == n && acts->hidden == p->hidden, "runner shapes differ");
    RACER_REQUIRE(scratch->len >= (size_t)(2 * n), "scratch must hold 2 x cars floats");
    RACER_REQUIRE(!fallback || fallback->len >= (size_t)n, "fallback must hold one float per car");
    if (racer_policy_act(p, sim->obs->ptr, n, acts, (int)n, sim->actions->ptr, scratch->ptr, scratch->ptr + n,
                         deterministic, t, sim->seed))
        return 1;
    return racer_surrogate_step_raw(s, sim, min_clearance, fallback ? fallback->ptr : nullptr);
}
