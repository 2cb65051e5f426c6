// Policy, activations and the PPO trainer on gfx950 (racer/ppo.mojo), the
// .npy checkpoint format (racer/ppo.mojo:341-355, racer/device.mojo:137-143)
// and the Runner step/benchmark loop (main.mojo:183-235).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include <type_traits>

#include "racer_internal.h"


namespace {

// Policy.__init__'s init kernel (racer/ppo.mojo:74-89), including the
// reference's absolute-index bias-column tests for W1 and W2.
__global__ void k_policy_init(float* theta, int64_t P, int64_t IN, int64_t OBS, int64_t H, int64_t HA, int64_t W1,
                              int64_t W2, int64_t LOGSTD, int seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float scale = 0.0f;
    if (i < W1) {
        scale = (i % IN < OBS) ? sqrtf(2.0f / (float)OBS) : 0.0f;
    } else if (i < W2) {
        scale = (i % HA < H) ? sqrtf(1.0f / (float)H) : 0.0f;
    } else if (i < LOGSTD) {
        scale = (i % HA < H) ? (((i - W2) / HA < 2) ? 0.00125f : 0.125f) : 0.0f;
    } else {
        theta[i] = -0.5f;
        return;
    }
    float z[4];
    racer_normal4(racer_random((uint64_t)(int64_t)seed, (uint64_t)i, 0), z);
    theta[i] = scale * z[0];
}

// Policy.transpose (racer/ppo.mojo:92-103): w1t[k, j] = W1[j, k], w2t[k, j] = W2[j, k].
__global__ void k_transpose(const float* theta, int64_t W1, int64_t W2, int H, int HA, float* w1t, float* w2t) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)HA * H) return;
    int k = (int)(idx / H), j = (int)(idx % H);
    w1t[(int64_t)k * H + j] = theta[W1 + (int64_t)j * HA + k];
    if (j < RACER_OUT) w2t[(int64_t)k * RACER_OUT + j] = theta[W2 + (int64_t)j * HA + k];
}

// GAE (racer/ppo.mojo:218-230), one car per thread, the running advantage in float64.
__global__ void k_gae(int n, int rollout, const float* reward, const float* done, const float* value,
                      const float* last_value, float* advantage) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double next = last_value[i];
    double running = 0.0;
    for (int t = rollout - 1; t >= 0; --t) {
        int64_t k = (int64_t)t * n + i;
        double alive = done[k] == 0.0f ? 1.0 : 0.0;
        double delta = (double)reward[k] + (double)RACER_GAMMA * alive * next - (double)value[k];
        running = delta + (double)RACER_GAMMA * (double)RACER_LAMBDA * alive * running;
        advantage[k] = (float)running;
        next = value[k];
    }
}

// Deterministic two-pass float64 sums of up to four rows (x, and x^2 when
// `square`): pass 1 writes one partial per block, pass 2 sums them in order.
constexpr int SUM_BLOCKS = 1024;
static_assert(4 * SUM_BLOCKS <= RACER_SUM_PARTIALS, "the device's partial-sum buffer holds 4 outputs");

__global__ void __launch_bounds__(256) k_sum_partial(const float* x, int64_t count, int64_t row_stride, int rows,
                                                     int square, double* partial) {
    __shared__ double red[256];
    for (int r = 0; r < rows; ++r) {
        const float* row = x + r * row_stride;
        double acc = 0.0, acc2 = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) {
            double v = row[i];
            acc += v;
            acc2 += v * v;
        }
        for (int q = 0; q < 1 + square; ++q) {
            red[threadIdx.x] = q ? acc2 : acc;
            __syncthreads();
            for (int s = 128; s > 0; s >>= 1) {
                if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                __syncthreads();
            }
            if (threadIdx.x == 0) partial[((int64_t)(r * (1 + square) + q)) * gridDim.x + blockIdx.x] = red[0];
            __syncthreads();
        }
    }
}

// Output o = the sum of its `blocks` partials in a fixed order: thread t adds
// partials t, t + 256, ... in order, then a fixed tree over the 256 threads.
__global__ void __launch_bounds__(256) k_sum_final(const double* partial, int blocks, int outputs, double* out,
                                                   float* out_f32, int64_t f32_slot) {
    __shared__ double red[256];
    for (int o = 0; o < outputs; ++o) {
        double acc = 0.0;
        for (int b = (int)threadIdx.x; b < blocks; b += 256) acc += partial[(int64_t)o * blocks + b];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (out) out[o] = red[0];
            if (out_f32) out_f32[f32_slot + o] = (float)red[0];
        }
        __syncthreads();
    }
}

// minibatch.gather (racer/ppo.mojo:276-279): x[r, b] = obs of sample s = b MB + m.
__global__ void k_gather(const float* obs, int64_t n, int64_t IN, int64_t batch, int minibatches, int m, float* x) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    int64_t s = b * minibatches + m;
    const float* src = obs + (s / n) * IN * n + s % n;
    for (int64_t r = 0; r < IN; ++r) x[r * batch + b] = src[r * n];
}

// The gather of every minibatch at once (racer/ppo.mojo:277-279 for all m):
// xall[m][r][b] = obs[s / n][r][s % n] with s = b MB + m. With n % MB == 0,
// sample s = t n + i has m = i % MB and b = t (n / MB) + i / MB, so a thread
// reads MB consecutive cars of one obs row and writes one column of each
// minibatch: reads and writes are both contiguous across the wavefront.
// X16: xall holds bf16 (round to nearest even, the conversion the bf16
// kernels apply on load), two values a float.
template <int MB, bool X16>
__global__ void k_permute(const float* __restrict__ obs, int64_t n, int64_t IN, int64_t rows, int64_t batch,
                          float* __restrict__ xall) {
    const int64_t per_row = n / MB;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rows * per_row) return;
    const int64_t q = g % per_row, row = g / per_row;  // row = t IN + r
    const int64_t r = row % IN, t = row / IN;
    const float* src = obs + row * n + q * MB;
    float v[MB];
    if constexpr (MB >= 4) {
#pragma unroll
        for (int h = 0; h < MB / 4; ++h) *reinterpret_cast<float4*>(v + 4 * h) = reinterpret_cast<const float4*>(src)[h];
    } else {
#pragma unroll
        for (int e = 0; e < MB; ++e) v[e] = src[e];
    }
    const int64_t b = t * per_row + q;
    if constexpr (X16) {
        uint16_t* x16 = reinterpret_cast<uint16_t*>(xall);
#pragma unroll
        for (int m = 0; m < MB; ++m)
            x16[((int64_t)m * IN + r) * batch + b] = __builtin_bit_cast(uint16_t, (__bf16)v[m]);
    } else {
#pragma unroll
        for (int m = 0; m < MB; ++m) xall[((int64_t)m * IN + r) * batch + b] = v[m];
    }
}

// minibatch.loss (racer/ppo.mojo:281-304): clipped surrogate and Huber value
// derivatives, d log std per sample, and the per-sample KL estimate.
__global__ void k_loss(int64_t batch, int64_t n, int minibatches, int m, const float* actions, const float* o,
                       const float* logp, const float* value, const float* advantage, const float* log_std,
                       float adv_mean, float adv_scale, float* d3) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    int64_t s = b * minibatches + m;
    int64_t t = s / n, i = s % n;
    float new_logp = 0.0f;
    float error[2], invvar[2];
    for (int a = 0; a < 2; ++a) {
        float ls = log_std[a];
        error[a] = actions[(t * 2 + a) * n + i] - o[a * batch + b];
        invvar[a] = racer_exp(-2.0f * ls);
        new_logp = new_logp - ((((0.5f * error[a]) * error[a]) * invvar[a] + ls) + RACER_HALF_LOG_2PI);
    }
    float logratio = new_logp - logp[s];
    float ratio = racer_exp(logratio);
    float a_hat = (advantage[s] - adv_mean) * adv_scale;
    float dlogp = (-a_hat * ratio) / (float)batch;
    if ((a_hat >= 0.0f && ratio > 1.0f + RACER_CLIP) || (a_hat < 0.0f && ratio < 1.0f - RACER_CLIP)) dlogp = 0.0f;
    for (int a = 0; a < 2; ++a) {
        d3[a * batch + b] = (dlogp * error[a]) * invvar[a];
        d3[(3 + a) * batch + b] = dlogp * ((error[a] * error[a]) * invvar[a] - 1.0f);
    }
    d3[2 * batch + b] = (0.5f * racer_clamp((o[2 * batch + b] - value[s]) - advantage[s], -1.0f, 1.0f)) / (float)batch;
    d3[5 * batch + b] = (ratio - 1.0f) - logratio;
}

// Trainer.adam (racer/ppo.mojo:325-336). Bias corrections come from the host.
__global__ void k_adam(float* theta, int64_t P, int64_t logstd_at, const float* grad, float* first, float* second,
                       const float* stats, float c1, float c2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float clip = racer_min(1.0f, RACER_MAX_NORM / (sqrtf(stats[RACER_STAT_NORM]) + 1e-8f));
    float g = grad[i] * clip;
    float m = 0.9f * first[i] + 0.1f * g;
    float v = 0.999f * second[i] + (0.001f * g) * g;
    first[i] = m;
    second[i] = v;
    float m_hat = m / c1;
    float v_hat = v / c2;
    float value = theta[i] - (RACER_LR * m_hat) / (sqrtf(v_hat) + 1e-5f);
    theta[i] = i >= logstd_at ? racer_min(racer_max(value, -2.0f), 0.0f) : value;
}

__global__ void k_store_norm(const double* pair, float* stats) { stats[RACER_STAT_NORM] = (float)pair[1]; }

// Output layer o = W2 h2 (racer/ppo.mojo:111) for HEAD_COLS samples per
// workgroup: the workgroup stages the [HA, HEAD_COLS] tile of h2 in LDS with
// row-contiguous loads, then 3 x HEAD_COLS threads each accumulate one output
// with fmaf over k = 0..HA-1 in order (the order of the float32 products).
// When SAMPLE, Sim.sample's Gaussian action, log probability and value
// (racer/sim.mojo:337-348) follow from those outputs.
constexpr int HEAD_COLS = 32;

template <bool SAMPLE>
__global__ void __launch_bounds__(256) k_head(const float* __restrict__ W2, const float* __restrict__ h2, int HA,
                                              int64_t cols, float* __restrict__ o, const float* log_std,
                                              float* act, int64_t act_cols, float* logp, float* value,
                                              int deterministic, int64_t step, const int64_t* step_base,
                                              int seed) {
    extern __shared__ float tile[];  // [HA][HEAD_COLS], then W2 [3][HA], then out[3][HEAD_COLS]
    float* wl = tile + HA * HEAD_COLS;
    float* out = wl + 3 * HA;
    const int64_t c0 = (int64_t)blockIdx.x * HEAD_COLS;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < HA * HEAD_COLS; idx += blockDim.x) {
        int k = idx / HEAD_COLS, c = idx % HEAD_COLS;
        tile[idx] = c0 + c < cols ? h2[(int64_t)k * cols + c0 + c] : 0.0f;
    }
    for (int idx = tid; idx < 3 * HA; idx += blockDim.x) wl[idx] = W2[idx];
    __syncthreads();
    if (tid < 3 * HEAD_COLS) {
        const int j = tid / HEAD_COLS, c = tid % HEAD_COLS;
        const float* w = wl + j * HA;
        float acc = 0.0f;
        for (int k = 0; k < HA; ++k) acc = fmaf(w[k], tile[k * HEAD_COLS + c], acc);
        out[j * HEAD_COLS + c] = acc;
        if (c0 + c < cols) o[j * cols + c0 + c] = acc;
    }
    if constexpr (SAMPLE) {
        __syncthreads();
        const int64_t col = c0 + tid;
        if (tid < HEAD_COLS && col < cols) {
            if (step_base) step += *step_base;
            float noise[4];
            racer_normal4(racer_random((uint64_t)(int64_t)seed + 1, (uint64_t)col, (uint64_t)step), noise);
            float probability = 0.0f;
            for (int a = 0; a < 2; ++a) {
                float z = deterministic ? 0.0f : noise[a];
                float ls = log_std[a];
                act[a * act_cols + col] = out[a * HEAD_COLS + tid] + racer_exp(ls) * z;
                probability = probability - (((0.5f * z) * z + ls) + RACER_HALF_LOG_2PI);
            }
            logp[col] = probability;
            value[col] = out[2 * HEAD_COLS + tid];
        }
    }
}

int launch_head(racer_policy* p, const float* h2, int64_t cols, float* o, bool sample, float* act, float* logp,
                float* value, int deterministic, int64_t step, int seed) {
    const float* W2 = p->theta->ptr + p->W2;
    const float* log_std = p->theta->ptr + p->LOGSTD;
    unsigned blocks = (unsigned)((cols + HEAD_COLS - 1) / HEAD_COLS);
    size_t shmem = ((size_t)(p->HA + 3) * HEAD_COLS + 3 * (size_t)p->HA) * sizeof(float);
    if (sample) {
        racer_timed timer(p->dev, "head_sample");
        hipLaunchKernelGGL(k_head<true>, dim3(blocks), dim3(128), shmem, p->dev->stream, W2, h2, (int)p->HA, cols, o,
                           log_std, act, cols, logp, value, deterministic, step, p->dev->step_base, seed);
    } else {
        racer_timed timer(p->dev, "head");
        hipLaunchKernelGGL(k_head<false>, dim3(blocks), dim3(128), shmem, p->dev->stream, W2, h2, (int)p->HA, cols,
                           o, log_std, act, cols, logp, value, deterministic, step, p->dev->step_base, seed);
    }
    return racer_launch_check("head");
}

int row_sums(racer_device* dev, const float* x, int64_t count, int64_t row_stride, int rows, int square,
             double* out64, float* out32, int64_t slot) {
    int outputs = rows * (1 + square);
    int blocks = (int)std::min<int64_t>(SUM_BLOCKS, std::max<int64_t>(1, (count + 255) / 256));
    RACER_REQUIRE(outputs * blocks <= RACER_SUM_PARTIALS, "row sums: too many outputs");
    double* partial = dev->sums;  // fixed for the device's lifetime, so captured graphs may hold it
    {
        racer_timed timer(dev, "sum");
        hipLaunchKernelGGL(k_sum_partial, dim3(blocks), dim3(256), 0, dev->stream, x, count, row_stride, rows, square,
                           partial);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, dev->stream, partial, blocks, outputs, out64, out32,
                           slot);
    }
    return racer_launch_check("row sums");
}

float bias_correction(float beta, int step) { return (float)(1.0 - std::pow((double)beta, (double)step)); }

}  // namespace

// ---------------------------------------------------------------- activations
extern "C" int racer_acts_create(racer_device* dev, int cols, const racer_config* cfg, racer_acts** out) {
    racer_clear_error();
    *out = nullptr;
    racer_config def;
    racer_config_default(&def);
    if (!cfg) cfg = &def;
    RACER_REQUIRE(cols > 0, "activations need at least one column");
    auto* a = new racer_acts();
    a->dev = dev;
    racer_device_retain(dev);
    a->cols = cols;
    a->hidden = cfg->hidden;
    size_t HA = (size_t)cfg->hidden + 1;
    if (racer_alloc(dev, HA * cols, 1.0f, &a->h1) || racer_alloc(dev, HA * cols, 1.0f, &a->h2) ||
        racer_alloc(dev, (size_t)RACER_OUT * cols, 0.0f, &a->o)) {
        racer_acts_destroy(a);
        return 1;
    }
    *out = a;
    return 0;
}

extern "C" void racer_acts_destroy(racer_acts* a) {
    if (!a) return;
    racer_free(a->h1);
    racer_free(a->h2);
    racer_free(a->o);
    racer_device* dev = a->dev;
    delete a;
    racer_device_release(dev);
}

extern "C" racer_buf* racer_acts_h1(racer_acts* a) { return a->h1; }
extern "C" racer_buf* racer_acts_h2(racer_acts* a) { return a->h2; }
extern "C" racer_buf* racer_acts_o(racer_acts* a) { return a->o; }

// ---------------------------------------------------------------- policy
extern "C" int racer_policy_create(racer_device* dev, int seed, const racer_config* cfg, racer_policy** out) {
    racer_clear_error();
    *out = nullptr;
    racer_config def;
    racer_config_default(&def);
    if (!cfg) cfg = &def;
    RACER_REQUIRE(cfg->beams >= 2 && cfg->hidden >= 1, "bad network shape");
    RACER_REQUIRE(cfg->hidden <= 460, "hidden width above 460 is not supported (the output layer stages h2 in LDS)");
    RACER_REQUIRE(cfg->precision >= 0 && cfg->precision <= 2, "precision must be 0 (float32), 1 (bf16) or 2 (plain)");
    auto* p = new racer_policy();
    p->dev = dev;
    racer_device_retain(dev);
    p->beams = cfg->beams;
    p->hidden = cfg->hidden;
    p->precision = cfg->precision;
    p->IN = cfg->beams + 5;
    p->HA = cfg->hidden + 1;
    int64_t H = cfg->hidden;
    p->W1 = H * p->IN;
    p->W2 = p->W1 + H * p->HA;
    p->LOGSTD = p->W2 + RACER_OUT * p->HA;
    p->P = p->LOGSTD + 2;
    if (racer_alloc(dev, (size_t)p->P, 0.0f, &p->theta) || racer_alloc(dev, (size_t)(p->HA * H), 0.0f, &p->w1t) ||
        racer_alloc(dev, (size_t)(p->HA * RACER_OUT), 0.0f, &p->w2t)) {
        racer_policy_destroy(p);
        return 1;
    }
    p->in_pad = racer_pad32(p->IN);
    p->ha_pad = racer_pad32(p->HA);
    RACER_TRY(hipMalloc(&p->theta_bf16, (size_t)(H * (p->in_pad + p->ha_pad)) * sizeof(uint16_t)));
    p->h_pad = racer_pad8(H);
    RACER_TRY(hipMalloc(&p->w1t_bf16, (size_t)(p->HA * p->h_pad) * sizeof(uint16_t)));
    if (H <= 256) {
        RACER_TRY(hipMalloc(&p->frag0, racer_fragment_elems((int)p->IN) * sizeof(uint16_t)));
        RACER_TRY(hipMalloc(&p->frag1, racer_fragment_elems((int)p->HA) * sizeof(uint16_t)));
        if (p->precision == 1 && racer_policy_fused_setup()) return 1;
    }
    if (p->precision == 0 && racer_policy_f32_fits((int)H, (int)p->IN)) {
        RACER_TRY(hipMalloc(&p->frag0_f32, racer_f32_fragment_floats((int)p->IN) * sizeof(float)));
        RACER_TRY(hipMalloc(&p->frag1_f32, racer_f32_fragment_floats((int)p->HA) * sizeof(float)));
        if (racer_policy_f32_setup()) return 1;
    }
    hipLaunchKernelGGL(k_policy_init, dim3((unsigned)((p->P + 255) / 256)), dim3(256), 0, dev->stream, p->theta->ptr,
                       p->P, p->IN, p->IN - 1, H, p->HA, p->W1, p->W2, p->LOGSTD, seed);
    if (racer_launch_check("policy init")) return 1;
    if (racer_policy_transpose(p)) return 1;
    RACER_TRY(hipStreamSynchronize(dev->stream));
    *out = p;
    return 0;
}

extern "C" void racer_policy_destroy(racer_policy* p) {
    if (!p) return;
    racer_free(p->theta);
    racer_free(p->w1t);
    racer_free(p->w2t);
    if (p->theta_bf16) (void)hipFree(p->theta_bf16);
    if (p->w1t_bf16) (void)hipFree(p->w1t_bf16);
    if (p->frag0) (void)hipFree(p->frag0);
    if (p->frag1) (void)hipFree(p->frag1);
    if (p->frag0_f32) (void)hipFree(p->frag0_f32);
    if (p->frag1_f32) (void)hipFree(p->frag1_f32);
    racer_device* dev = p->dev;
    delete p;
    racer_device_release(dev);
}

extern "C" racer_buf* racer_policy_theta(racer_policy* p) { return p->theta; }
extern "C" size_t racer_policy_params(const racer_policy* p) { return (size_t)p->P; }

extern "C" int racer_policy_log_std(racer_policy* p, racer_view* v) {
    racer_clear_error();
    RACER_REQUIRE(p && v, "null argument");
    v->buf = p->theta;
    v->offset = (size_t)p->LOGSTD;
    v->rows = 1;
    v->cols = 2;
    return 0;
}

extern "C" int racer_policy_transpose(racer_policy* p) {
    racer_clear_error();
    int64_t total = p->HA * p->hidden;
    {
        racer_timed timer(p->dev, "transpose");
        hipLaunchKernelGGL(k_transpose, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, p->dev->stream,
                           p->theta->ptr, p->W1, p->W2, p->hidden, (int)p->HA, p->w1t->ptr, p->w2t->ptr);
    }
    p->bf16_stale = true;
    p->f32_stale = true;
    return racer_launch_check("transpose");
}

// The bf16 weight copies used by the MFMA products, refreshed after updates.
static int refresh_bf16(racer_policy* p) {
    if (!p->bf16_stale && p->bf16_gen == p->theta->gen) return 0;
    const float* th = p->theta->ptr;
    int H = p->hidden;
    uint16_t* w1 = p->theta_bf16 + (int64_t)H * p->in_pad;
    if (racer_pack_bf16(p->dev, th, p->IN, p->theta_bf16, p->in_pad, H, (int)p->IN) ||
        racer_pack_bf16(p->dev, th + p->W1, p->HA, w1, p->ha_pad, H, (int)p->HA))
        return 1;
    if (p->frag0 && (racer_pack_fragments(p->dev, th, p->IN, H, (int)p->IN, p->frag0) ||
                     racer_pack_fragments(p->dev, th + p->W1, p->HA, H, (int)p->HA, p->frag1)))
        return 1;
    if (racer_pack_bf16(p->dev, p->w1t->ptr, H, p->w1t_bf16, p->h_pad, (int)p->HA, H)) return 1;
    p->bf16_stale = false;
    p->bf16_gen = p->theta->gen;
    return 0;
}

// The float32 operand-order weight copies of the fused float32 forward,
// rebuilt when an update or a write changed theta.
static int refresh_f32(racer_policy* p) {
    if (!p->f32_stale && p->f32_gen == p->theta->gen) return 0;
    const float* th = p->theta->ptr;
    if (racer_pack_f32_fragments(p->dev, th, p->IN, p->hidden, (int)p->IN, p->frag0_f32) ||
        racer_pack_f32_fragments(p->dev, th + p->W1, p->HA, p->hidden, (int)p->HA, p->frag1_f32))
        return 1;
    p->f32_stale = false;
    p->f32_gen = p->theta->gen;
    return 0;
}

// The float32 forward runs as one kernel on K = 1 float32 MFMA (bitwise the
// float32 products) at every column count: at the PPO update's 262144-column
// minibatches it takes 877 us against 1498 us for the two tiled products and
// the head at 113 inputs, and 2355 us against 4107 us at 1086, although each
// 16-column workgroup streams all of W0.
static bool use_f32_fused(const racer_policy* p, int cols) {
    (void)cols;
    return p->precision == 0 && p->frag0_f32;
}

int racer_policy_refresh(racer_policy* p) {
    if (p->precision == 1) return refresh_bf16(p);
    return p->frag0_f32 ? refresh_f32(p) : 0;
}

// Hidden layers of the forward pass (racer/ppo.mojo:109-110) as two products.
static int forward_hidden(racer_policy* p, const float* x, int64_t x_cols, racer_acts* a, int cols) {
    const float* th = p->theta->ptr;
    int H = p->hidden;
    if (p->precision == 1) {
        if (refresh_bf16(p)) return 1;
        uint16_t* w0 = p->theta_bf16;
        uint16_t* w1 = p->theta_bf16 + (int64_t)H * p->in_pad;
        if (racer_gemm_nn_bf16(p->dev, w0, p->in_pad, x, x_cols, a->h1->ptr, cols, H, cols, (int)p->IN,
                               RACER_EPI_TANH_MLP) ||
            racer_gemm_nn_bf16(p->dev, w1, p->ha_pad, a->h1->ptr, cols, a->h2->ptr, cols, H, cols, (int)p->HA,
                               RACER_EPI_TANH_MLP))
            return 1;
    } else {
        if (racer_gemm_nn_f32(p->dev, th, p->IN, x, x_cols, a->h1->ptr, cols, H, cols, (int)p->IN, RACER_EPI_TANH,
                              nullptr, 0) ||
            racer_gemm_nn_f32(p->dev, th + p->W1, p->HA, a->h1->ptr, cols, a->h2->ptr, cols, H, cols, (int)p->HA,
                              RACER_EPI_TANH, nullptr, 0))
            return 1;
    }
    return 0;
}

// The bf16 forward runs as one fused kernel when the hidden width fits its
// workgroup (eight wavefronts of 32 rows).
static bool use_fused(const racer_policy* p) { return p->precision == 1 && p->hidden <= 256; }

int racer_policy_forward_raw(racer_policy* p, const float* x, int64_t x_cols, racer_acts* a, int cols) {
    if (use_fused(p)) {
        if (refresh_bf16(p)) return 1;
        return racer_policy_fused(p, x, x_cols, a, cols, false, nullptr, nullptr, nullptr, 0, 0, 0);
    }
    if (use_f32_fused(p, cols)) {
        if (refresh_f32(p)) return 1;
        return racer_policy_f32(p, x, x_cols, a, cols, false, nullptr, nullptr, nullptr, 0, 0, 0);
    }
    if (forward_hidden(p, x, x_cols, a, cols)) return 1;
    if (p->precision == 2)  // the reference's decomposition: the output layer as a third product
        return racer_gemm_nn_f32(p->dev, p->theta->ptr + p->W2, p->HA, a->h2->ptr, cols, a->o->ptr, cols, RACER_OUT,
                                 cols, (int)p->HA, RACER_EPI_NONE, nullptr, 0);
    return launch_head(p, a->h2->ptr, cols, a->o->ptr, false, nullptr, nullptr, nullptr, 0, 0, 0);
}

// Policy forward plus Sim.sample for one column per car (main.mojo:186-190,
// racer/ppo.mojo:192-202).
int racer_policy_act(racer_policy* p, const float* x, int64_t x_cols, racer_acts* a, int cols, float* act,
                     float* logp, float* value, int deterministic, int64_t step, int seed) {
    const int kernel = p->act_kernel;
    if (use_fused(p)) {
        if (refresh_bf16(p)) return 1;
        if (kernel == RACER_ACT_TILED) {  // the tiled forward, then Sim.sample on its outputs
            if (racer_policy_fused(p, x, x_cols, a, cols, false, nullptr, nullptr, nullptr, 0, 0, 0, kernel))
                return 1;
            return racer_launch_sample_raw(p->dev, cols, a->o->ptr, cols, p->theta->ptr + p->LOGSTD, act, cols, logp,
                                           value, deterministic, step, seed);
        }
        return racer_policy_fused(p, x, x_cols, a, cols, true, act, logp, value, deterministic, step, seed, kernel);
    }
    if (kernel != RACER_ACT_TILED && use_f32_fused(p, cols)) {
        if (refresh_f32(p)) return 1;
        return racer_policy_f32(p, x, x_cols, a, cols, true, act, logp, value, deterministic, step, seed);
    }
    if (p->precision == 2) {  // three products, then a separate sample kernel (racer/sim.mojo:323-348)
        if (racer_policy_forward_raw(p, x, x_cols, a, cols)) return 1;
        return racer_launch_sample_raw(p->dev, cols, a->o->ptr, cols, p->theta->ptr + p->LOGSTD, act, cols, logp,
                                       value, deterministic, step, seed);
    }
    if (forward_hidden(p, x, x_cols, a, cols)) return 1;
    return launch_head(p, a->h2->ptr, cols, a->o->ptr, true, act, logp, value, deterministic, step, seed);
}

extern "C" int racer_policy_forward(racer_policy* p, const racer_view* x_p, racer_acts* a) {
    racer_clear_error();
    RACER_REQUIRE(p && x_p && a, "null argument");
    const racer_view& x = *x_p;
    RACER_REQUIRE(a && a->hidden == p->hidden, "activations do not match the policy's hidden width");
    if (racer_check_view(x, p->IN, a->cols, "policy input")) return 1;
    return racer_policy_forward_raw(p, view_ptr(x), x.cols, a, a->cols);
}

// Workspace floats the three weight gradients of backward_raw need (they run
// one after another on the stream, so they share it).
static size_t backward_partials(const racer_policy* p, int64_t cols) {
    const int H = p->hidden, K = (int)cols;
    size_t w01 = p->precision == 1
                     ? std::max(racer_gemm_nt_bf16_partials(H, p->IN, K), racer_gemm_nt_bf16_partials(H, p->HA, K))
                     : std::max(racer_gemm_nt_f32_partials(H, p->IN, K), racer_gemm_nt_f32_partials(H, p->HA, K));
    return std::max(w01, racer_gemm_nt_f32_partials(RACER_OUT, p->HA, K));
}

// ws: the weight gradients' split-K workspace (ws_len floats), or null for the
// device's grow-only scratch. x16 (bf16 policies, the PPO update's aligned
// minibatches): x already in bf16, read by the W0 gradient in place of x; d2
// and d1 are then stored in bf16 in their buffers (two values a float), as
// each of their consumers rounds them on load, and d1's bias row is left
// unset (no weight gradient reads it): the same weight gradients bit for bit.
static int backward_raw(racer_policy* p, const float* x, racer_acts* a, const float* d3, float* d2, float* d1,
                        float* grad, int64_t cols, float* ws = nullptr, size_t ws_len = 0,
                        const uint16_t* x16 = nullptr) {
    racer_device* dev = p->dev;
    int H = p->hidden;
    int64_t HA = p->HA, IN = p->IN;
    const float* h1 = a->h1->ptr;
    const float* h2 = a->h2->ptr;
    const bool io16 = x16 != nullptr && p->precision == 1;
    // through (racer/ppo.mojo:125-138): d2 = (1 - h2^2) (w2t d3); d1 = (1 - h1^2) (w1t d2[:H]),
    // the second on MFMA in bf16 precision
    if (racer_mask_k3(dev, p->w2t->ptr, d3, cols, d2, io16, cols, (int)HA, (int)cols, h2, cols)) return 1;
    if (io16) {
        if (refresh_bf16(p) ||
            racer_gemm_nn_bf16_mask16(dev, p->w1t_bf16, p->h_pad, reinterpret_cast<const uint16_t*>(d2), cols,
                                      reinterpret_cast<uint16_t*>(d1), cols, H, (int)cols, H, h1, cols))
            return 1;
    } else if (p->precision == 1) {
        // rows 0 .. H-1 on MFMA; row H is the bias row, masked by 1 - h1[H]^2 = 0
        // and read by no weight gradient (they take d1's first H rows), so it
        // is filled with zeros rather than spending a fifth row tile on it
        if (refresh_bf16(p) ||
            racer_gemm_nn_bf16(dev, p->w1t_bf16, p->h_pad, d2, cols, d1, cols, H, (int)cols, H, RACER_EPI_MASK, h1,
                               cols) ||
            racer_fill(dev, d1 + (int64_t)H * cols, (size_t)cols, 0.0f))
            return 1;
    } else if (racer_gemm_nn_f32(dev, p->w1t->ptr, H, d2, cols, d1, cols, (int)HA, (int)cols, H, RACER_EPI_MASK, h1,
                                 cols)) {
        return 1;
    }
    // weight gradients (racer/ppo.mojo:121-123)
    if (io16) {
        if (racer_gemm_nt_bf16_io(dev, d1, true, cols, x16, true, cols, grad, IN, H, (int)IN, (int)cols, ws, ws_len) ||
            racer_gemm_nt_bf16_io(dev, d2, true, cols, h1, false, cols, grad + p->W1, HA, H, (int)HA, (int)cols, ws,
                                  ws_len))
            return 1;
    } else if (p->precision == 1) {
        if (racer_gemm_nt_bf16(dev, d1, cols, x, cols, grad, IN, H, (int)IN, (int)cols, ws, ws_len) ||
            racer_gemm_nt_bf16(dev, d2, cols, h1, cols, grad + p->W1, HA, H, (int)HA, (int)cols, ws, ws_len))
            return 1;
    } else {
        if (racer_gemm_nt_f32(dev, d1, cols, x, cols, grad, IN, H, (int)IN, (int)cols, ws, ws_len) ||
            racer_gemm_nt_f32(dev, d2, cols, h1, cols, grad + p->W1, HA, H, (int)HA, (int)cols, ws, ws_len))
            return 1;
    }
    return racer_gemm_nt_f32(dev, d3, cols, h2, cols, grad + p->W2, HA, RACER_OUT, (int)HA, (int)cols, ws, ws_len);
}

extern "C" int racer_policy_backward(racer_policy* p, const racer_view* x_p, racer_acts* a, const racer_view* d3_p,
                                     racer_buf* d2, racer_buf* d1, racer_buf* grad) {
    racer_clear_error();
    RACER_REQUIRE(p && x_p && a && d3_p, "null argument");
    const racer_view &x = *x_p, &d3 = *d3_p;
    int64_t cols = a->cols;
    if (racer_check_view(x, p->IN, cols, "policy input") || racer_check_view(d3, RACER_OUT, cols, "d3")) return 1;
    RACER_REQUIRE(d2 && d2->len >= (size_t)(p->HA * cols) && d1 && d1->len >= (size_t)(p->HA * cols) && grad &&
                      grad->len >= (size_t)p->P,
                  "backward buffers are too small");
    return backward_raw(p, view_ptr(x), a, view_ptr(d3), d2->ptr, d1->ptr, grad->ptr, cols);
}

// racer_policy_backward with the weight gradients' split-K partials in the
// GPU's fixed scratch array instead of the grow-only device scratch: the same
// kernels in the same order (so bitwise the same results), and no allocation.
// Calls that overlap on one GPU share that array and race.
extern "C" int racer_policy_backward_noalloc(racer_policy* p, const racer_view* x_p, racer_acts* a,
                                             const racer_view* d3_p, racer_buf* d2, racer_buf* d1, racer_buf* grad) {
    racer_clear_error();
    RACER_REQUIRE(p && x_p && a && d3_p, "null argument");
    const racer_view &x = *x_p, &d3 = *d3_p;
    int64_t cols = a->cols;
    if (racer_check_view(x, p->IN, cols, "policy input") || racer_check_view(d3, RACER_OUT, cols, "d3")) return 1;
    RACER_REQUIRE(d2 && d2->len >= (size_t)(p->HA * cols) && d1 && d1->len >= (size_t)(p->HA * cols) && grad &&
                      grad->len >= (size_t)p->P,
                  "backward buffers are too small");
    const size_t need = backward_partials(p, cols);
    if (need > RACER_FIXED_SCRATCH_FLOATS) {
        racer_set_error("the allocation-free backward needs " + std::to_string(need) +
                        " floats of split-K scratch for " + std::to_string(cols) + " columns; its fixed scratch holds " +
                        std::to_string(RACER_FIXED_SCRATCH_FLOATS) +
                        " (16 MiB): use racer_policy_backward, or fewer columns per call");
        return 1;
    }
    RACER_REQUIRE(p->dev->fixed_scratch, "the device has no fixed scratch array");
    return backward_raw(p, view_ptr(x), a, view_ptr(d3), d2->ptr, d1->ptr, grad->ptr, cols, p->dev->fixed_scratch,
                        RACER_FIXED_SCRATCH_FLOATS);
}

// ---------------------------------------------------------------- trainer
// Whether the update's permuted inputs can be bf16 for this trainer.
static bool x16_applies(const racer_trainer* tr) {
    return racer_tiled_x16_applies(tr->precision, tr->hidden, tr->IN, tr->batch, tr->batch);
}

// (Re)allocates the permuted inputs, float32 or bf16 (two values a float).
static int alloc_permuted(racer_trainer* tr, bool x16) {
    const size_t values = (size_t)tr->rollout * tr->IN * tr->n;
    const size_t floats = x16 ? (values + 1) / 2 : values;
    if (tr->xall && tr->xall->len == floats && tr->x16 == x16) return 0;
    racer_free(tr->xall);
    tr->xall = nullptr;
    tr->x16 = x16;
    return racer_alloc(tr->dev, floats, 0.0f, &tr->xall);
}

extern "C" int racer_trainer_create(racer_device* dev, int n, const racer_config* cfg, racer_trainer** out) {
    racer_clear_error();
    *out = nullptr;
    racer_config def;
    racer_config_default(&def);
    if (!cfg) cfg = &def;
    RACER_REQUIRE(n > 0 && cfg->rollout > 0 && cfg->minibatches > 0 && cfg->epochs > 0, "bad trainer shape");
    RACER_REQUIRE(((int64_t)n * cfg->rollout) % cfg->minibatches == 0,
                  "cars x rollout must be a multiple of the minibatch count");
    // The bf16 forward reads its inputs through buffer loads with 32-bit byte
    // offsets: a minibatch's [beams + 5, batch] floats must stay below 4 GiB.
    const int64_t batch = (int64_t)n * cfg->rollout / cfg->minibatches;
    RACER_REQUIRE(cfg->precision != 1 || racer_fused_reach(cfg->beams + 5, batch, batch),
                  "a bf16 minibatch of (beams + 5) x cars x rollout / minibatches floats must stay below 4 GiB "
                  "(32-bit buffer offsets): use more minibatches or fewer cars");
    auto* tr = new racer_trainer();
    tr->dev = dev;
    racer_device_retain(dev);
    tr->n = n;
    tr->beams = cfg->beams;
    tr->hidden = cfg->hidden;
    tr->rollout = cfg->rollout;
    tr->epochs = cfg->epochs;
    tr->minibatches = cfg->minibatches;
    tr->precision = cfg->precision;
    tr->IN = cfg->beams + 5;
    tr->HA = cfg->hidden + 1;
    tr->batch = (int64_t)n * cfg->rollout / cfg->minibatches;
    tr->P = cfg->hidden * tr->IN + cfg->hidden * tr->HA + RACER_OUT * tr->HA + 2;
    size_t R = cfg->rollout, N = n, B = tr->batch;
    if (racer_alloc(dev, (R + 1) * tr->IN * N, 0.0f, &tr->obs) || racer_alloc(dev, R * 2 * N, 0.0f, &tr->actions) ||
        racer_alloc(dev, R * N, 0.0f, &tr->logp) || racer_alloc(dev, R * N, 0.0f, &tr->value) ||
        racer_alloc(dev, R * N, 0.0f, &tr->reward) || racer_alloc(dev, R * N, 0.0f, &tr->done) ||
        racer_alloc(dev, R * N, 0.0f, &tr->advantage) || racer_alloc(dev, tr->IN * B, 0.0f, &tr->x) ||
        racer_alloc(dev, 6 * B, 0.0f, &tr->d3) || racer_alloc(dev, tr->HA * B, 0.0f, &tr->d2) ||
        racer_alloc(dev, tr->HA * B, 0.0f, &tr->d1) || racer_alloc(dev, tr->P, 0.0f, &tr->grad) ||
        racer_alloc(dev, tr->P, 0.0f, &tr->first) || racer_alloc(dev, tr->P, 0.0f, &tr->second) ||
        racer_alloc(dev, RACER_STATS, 0.0f, &tr->stats) || racer_acts_create(dev, (int)B, cfg, &tr->acts)) {
        racer_trainer_destroy(tr);
        return 1;
    }
    RACER_TRY_OR(hipMalloc(&tr->dstats, 8 * sizeof(double)), racer_trainer_destroy(tr));
    const int MB = tr->minibatches;
    if ((MB == 1 || MB == 2 || MB == 4 || MB == 8) && N % MB == 0 && alloc_permuted(tr, x16_applies(tr))) {
        racer_trainer_destroy(tr);
        return 1;
    }
    *out = tr;
    return 0;
}

extern "C" int racer_trainer_set_bf16_inputs(racer_trainer* tr, int on) {
    racer_clear_error();
    RACER_REQUIRE(tr, "null argument");
    if (!tr->xall) return 0;  // no permuted copy: each minibatch is gathered in float32
    return alloc_permuted(tr, on != 0 && x16_applies(tr));
}

extern "C" int racer_trainer_bf16_inputs(const racer_trainer* tr) { return tr->xall && tr->x16 ? 1 : 0; }

extern "C" void racer_trainer_destroy(racer_trainer* tr) {
    if (!tr) return;
    racer_buf* bufs[] = {tr->obs, tr->actions, tr->logp, tr->value, tr->reward, tr->done, tr->advantage, tr->x,
                         tr->d3, tr->d2, tr->d1, tr->grad, tr->first, tr->second, tr->stats, tr->xall};
    for (auto* b : bufs) racer_free(b);
    racer_acts_destroy(tr->acts);
    if (tr->dstats) (void)hipFree(tr->dstats);
    racer_device* dev = tr->dev;
    delete tr;
    racer_device_release(dev);
}

static int check_compatible(racer_trainer* tr, racer_policy* p) {
    RACER_REQUIRE(p->beams == tr->beams && p->hidden == tr->hidden, "trainer and policy shapes differ");
    return 0;
}

// racer/ppo.mojo:190-231: everything up to the reward sum, which stays on the
// device (float64 in dstats[2]); the sample of step t uses Philox offset step0 + t.
int racer_rollout_enqueue(racer_trainer* tr, racer_sim* sim, racer_policy* p, racer_acts* acts, int64_t step0) {
    if (check_compatible(tr, p)) return 1;
    RACER_REQUIRE(sim->n == tr->n && sim->beams == tr->beams, "trainer and sim shapes differ");
    RACER_REQUIRE(acts->cols == tr->n && acts->hidden == tr->hidden, "activations must have one column per car");
    racer_device* dev = tr->dev;
    int64_t n = tr->n, IN = tr->IN, R = tr->rollout;
    if (racer_copy_raw(dev, tr->obs->ptr, sim->obs->ptr, (size_t)(IN * n))) return 1;
    for (int64_t t = 0; t < R; ++t) {
        float* obs_t = tr->obs->ptr + t * IN * n;
        if (racer_policy_act(p, obs_t, n, acts, (int)n, tr->actions->ptr + t * 2 * n, tr->logp->ptr + t * n,
                             tr->value->ptr + t * n, 0, step0 + t, sim->seed))
            return 1;
        if (racer_launch_step(sim, tr->actions->ptr + t * 2 * n, tr->reward->ptr + t * n, tr->done->ptr + t * n,
                              tr->obs->ptr + (t + 1) * IN * n, n))
            return 1;
    }
    if (racer_copy_raw(dev, sim->obs->ptr, tr->obs->ptr + R * IN * n, (size_t)(IN * n))) return 1;
    if (racer_policy_forward_raw(p, tr->obs->ptr + R * IN * n, n, acts, (int)n)) return 1;
    {
        racer_timed timer(dev, "gae");
        hipLaunchKernelGGL(k_gae, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, dev->stream, (int)n, (int)R,
                           tr->reward->ptr, tr->done->ptr, tr->value->ptr, acts->o->ptr + 2 * n, tr->advantage->ptr);
    }
    if (racer_launch_check("gae")) return 1;
    return row_sums(dev, tr->reward->ptr, R * n, 0, 1, 0, tr->dstats + 2, tr->stats->ptr, RACER_STAT_REWARD);
}

extern "C" int racer_trainer_rollout(racer_trainer* tr, racer_sim* sim, racer_policy* p, racer_acts* acts,
                                     int iteration, float* mean_reward) {
    racer_clear_error();
    if (racer_rollout_enqueue(tr, sim, p, acts, (int64_t)iteration * tr->rollout)) return 1;
    return racer_trainer_mean_reward(tr, mean_reward);
}

int racer_trainer_mean_reward(racer_trainer* tr, float* mean_reward) {
    racer_device* dev = tr->dev;
    int64_t n = tr->n, R = tr->rollout;
    double total = 0.0;
    RACER_TRY(hipMemcpyAsync(&total, tr->dstats + 2, sizeof(double), hipMemcpyDeviceToHost, dev->stream));
    RACER_TRY(hipStreamSynchronize(dev->stream));
    *mean_reward = (float)(total / (double)(R * n));
    return 0;
}

static int minibatch_core(racer_trainer* tr, racer_policy* p, int m, float adv_mean, float adv_scale, const float* x,
                          const uint16_t* x16 = nullptr);

extern "C" int racer_trainer_minibatch(racer_trainer* tr, racer_policy* p, int m, float adv_mean, float adv_scale) {
    racer_clear_error();
    if (check_compatible(tr, p)) return 1;
    RACER_REQUIRE(m >= 0 && m < tr->minibatches, "minibatch index out of range");
    racer_device* dev = tr->dev;
    int64_t B = tr->batch, n = tr->n;
    unsigned blocks = (unsigned)((B + 255) / 256);
    {
        racer_timed timer(dev, "gather");
        hipLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, dev->stream, tr->obs->ptr, n, tr->IN, B,
                           tr->minibatches, m, tr->x->ptr);
    }
    if (racer_launch_check("gather")) return 1;
    return minibatch_core(tr, p, m, adv_mean, adv_scale, tr->x->ptr);
}

// Every minibatch's inputs at once, into xall (when the trainer has it).
static int permute_all(racer_trainer* tr) {
    racer_device* dev = tr->dev;
    int64_t rows = (int64_t)tr->rollout * tr->IN, per_row = tr->n / tr->minibatches;
    unsigned blocks = (unsigned)((rows * per_row + 255) / 256);
    racer_timed timer(dev, "permute");
    auto kernel = [&](auto mb) {
        constexpr int MB = decltype(mb)::value;
        return tr->x16 ? k_permute<MB, true> : k_permute<MB, false>;
    };
    auto pick = tr->minibatches == 1   ? kernel(std::integral_constant<int, 1>{})
                : tr->minibatches == 2 ? kernel(std::integral_constant<int, 2>{})
                : tr->minibatches == 4 ? kernel(std::integral_constant<int, 4>{})
                                       : kernel(std::integral_constant<int, 8>{});
    hipLaunchKernelGGL(pick, dim3(blocks), dim3(256), 0, dev->stream, tr->obs->ptr, (int64_t)tr->n, tr->IN, rows,
                       tr->batch, tr->xall->ptr);
    return racer_launch_check("permute");
}

// x16: the minibatch's inputs in bf16 instead of x (bf16 policies on the
// tiled forward), the same bits through every product.
static int minibatch_core(racer_trainer* tr, racer_policy* p, int m, float adv_mean, float adv_scale, const float* x,
                          const uint16_t* x16) {
    racer_device* dev = tr->dev;
    int64_t B = tr->batch, n = tr->n;
    unsigned blocks = (unsigned)((B + 255) / 256);
    if (x16 ? racer_policy_tiled_x16(p, x16, B, tr->acts, (int)B) : racer_policy_forward_raw(p, x, B, tr->acts, (int)B))
        return 1;
    {
        racer_timed timer(dev, "loss");
        hipLaunchKernelGGL(k_loss, dim3(blocks), dim3(256), 0, dev->stream, B, n, tr->minibatches, m,
                           tr->actions->ptr, tr->acts->o->ptr, tr->logp->ptr, tr->value->ptr, tr->advantage->ptr,
                           p->theta->ptr + p->LOGSTD, adv_mean, adv_scale, tr->d3->ptr);
    }
    if (racer_launch_check("loss")) return 1;
    if (backward_raw(p, x, tr->acts, tr->d3->ptr, tr->d2->ptr, tr->d1->ptr, tr->grad->ptr, B, nullptr, 0, x16))
        return 1;
    // log-std gradient: row sums of d3[3..4] (racer/ppo.mojo:313)
    if (row_sums(dev, tr->d3->ptr + 3 * B, B, B, 2, 0, nullptr, tr->grad->ptr, p->LOGSTD)) return 1;
    // squared gradient norm (racer/ppo.mojo:314-315)
    if (row_sums(dev, tr->grad->ptr, tr->P, 0, 1, 1, tr->dstats + 4, nullptr, 0)) return 1;
    // stats[NORM] = float(sum of squares), the second of the (sum, sum of squares) pair
    hipLaunchKernelGGL(k_store_norm, dim3(1), dim3(1), 0, dev->stream, (const double*)(tr->dstats + 4),
                       tr->stats->ptr);
    return racer_launch_check("norm");
}

extern "C" int racer_trainer_adam(racer_trainer* tr, racer_policy* p) {
    racer_clear_error();
    if (check_compatible(tr, p)) return 1;
    racer_device* dev = tr->dev;
    float c1 = bias_correction(0.9f, tr->step), c2 = bias_correction(0.999f, tr->step);
    {
        racer_timed timer(dev, "adam");
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((tr->P + 255) / 256)), dim3(256), 0, dev->stream, p->theta->ptr,
                           tr->P, p->LOGSTD, tr->grad->ptr, tr->first->ptr, tr->second->ptr, tr->stats->ptr, c1, c2);
    }
    if (racer_launch_check("adam")) return 1;
    return racer_policy_transpose(p);
}

extern "C" int racer_trainer_update(racer_trainer* tr, racer_policy* p, float* kl) {
    racer_clear_error();
    if (check_compatible(tr, p)) return 1;
    racer_device* dev = tr->dev;
    int64_t samples = (int64_t)tr->rollout * tr->n;
    // advantage moments in float64 (racer/ppo.mojo:242-250)
    if (row_sums(dev, tr->advantage->ptr, samples, 0, 1, 1, tr->dstats, nullptr, 0)) return 1;
    double mom[2];
    RACER_TRY(hipMemcpyAsync(mom, tr->dstats, sizeof(mom), hipMemcpyDeviceToHost, dev->stream));
    RACER_TRY(hipStreamSynchronize(dev->stream));
    double mean = mom[0] / (double)samples;
    double var = mom[1] / (double)samples - mean * mean;
    float adv_mean = (float)mean;
    float adv_scale = (float)(1.0 / std::sqrt(std::max(var, 1e-8)));
    // the permuted copy serves a policy of the precision it was made for: a
    // bf16 copy only a bf16 policy (a float32 policy then gathers each minibatch)
    const bool permuted = tr->xall && (!tr->x16 || p->precision == 1);
    if (permuted && permute_all(tr)) return 1;
    for (int epoch = 0; epoch < tr->epochs; ++epoch)
        for (int m = 0; m < tr->minibatches; ++m) {
            const int mm = (m + epoch) % tr->minibatches;
            const int64_t at = (int64_t)mm * tr->IN * tr->batch;
            if (permuted && tr->x16) {
                if (minibatch_core(tr, p, mm, adv_mean, adv_scale, nullptr,
                                   reinterpret_cast<const uint16_t*>(tr->xall->ptr) + at))
                    return 1;
            } else if (permuted) {
                if (minibatch_core(tr, p, mm, adv_mean, adv_scale, tr->xall->ptr + at)) return 1;
            } else if (racer_trainer_minibatch(tr, p, mm, adv_mean, adv_scale)) {
                return 1;
            }
            tr->step += 1;
            if (racer_trainer_adam(tr, p)) return 1;
        }
    // KL of the last minibatch (racer/ppo.mojo:256-257)
    if (row_sums(dev, tr->d3->ptr + 5 * tr->batch, tr->batch, 0, 1, 0, tr->dstats + 6, tr->stats->ptr, RACER_STAT_KL))
        return 1;
    double klsum = 0.0;
    RACER_TRY(hipMemcpyAsync(&klsum, tr->dstats + 6, sizeof(double), hipMemcpyDeviceToHost, dev->stream));
    RACER_TRY(hipStreamSynchronize(dev->stream));
    *kl = (float)(klsum / (double)tr->batch);
    return 0;
}

extern "C" int racer_trainer_step_count(const racer_trainer* tr) { return tr->step; }
extern "C" racer_buf* racer_trainer_obs(racer_trainer* tr) { return tr->obs; }
extern "C" racer_buf* racer_trainer_actions(racer_trainer* tr) { return tr->actions; }
extern "C" racer_buf* racer_trainer_logp(racer_trainer* tr) { return tr->logp; }
extern "C" racer_buf* racer_trainer_value(racer_trainer* tr) { return tr->value; }
extern "C" racer_buf* racer_trainer_reward(racer_trainer* tr) { return tr->reward; }
extern "C" racer_buf* racer_trainer_done(racer_trainer* tr) { return tr->done; }
extern "C" racer_buf* racer_trainer_advantage(racer_trainer* tr) { return tr->advantage; }
extern "C" racer_buf* racer_trainer_grad(racer_trainer* tr) { return tr->grad; }
extern "C" racer_buf* racer_trainer_d3(racer_trainer* tr) { return tr->d3; }
extern "C" racer_buf* racer_trainer_stats(racer_trainer* tr) { return tr->stats; }
extern "C" racer_acts* racer_trainer_acts(racer_trainer* tr) { return tr->acts; }

// ---------------------------------------------------------------- checkpoints (.npy, float32, 1-D)
extern "C" int racer_save(racer_policy* p, const char* path, int iteration) {
    racer_clear_error();
    std::vector<float> values((size_t)p->P + 1);
    values[0] = (float)iteration;
    if (racer_read(p->theta, values.data() + 1)) return 1;
    std::string header = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + std::to_string(values.size()) +
                         ",), }";
    size_t total = 10 + header.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    header += std::string(pad, ' ') + "\n";
    std::ofstream f(path, std::ios::binary);
    RACER_REQUIRE(f.good(), std::string("cannot write ") + path);
    const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
    f.write(magic, 8);
    uint16_t hl = (uint16_t)header.size();
    f.write(reinterpret_cast<const char*>(&hl), 2);
    f.write(header.data(), (std::streamsize)header.size());
    f.write(reinterpret_cast<const char*>(values.data()), (std::streamsize)(values.size() * sizeof(float)));
    RACER_REQUIRE(f.good(), std::string("cannot write ") + path);
    return 0;
}

extern "C" int racer_load(racer_policy* p, const char* path, int* iteration) {
    racer_clear_error();
    std::ifstream f(path, std::ios::binary);
    RACER_REQUIRE(f.good(), std::string("cannot read ") + path);
    char magic[8];
    f.read(magic, 8);
    RACER_REQUIRE(f.good() && std::memcmp(magic, "\x93NUMPY", 6) == 0, "not a .npy file");
    uint32_t hl = 0;
    if (magic[6] == 1) {
        uint16_t h16 = 0;
        f.read(reinterpret_cast<char*>(&h16), 2);
        hl = h16;
    } else {
        f.read(reinterpret_cast<char*>(&hl), 4);
    }
    std::string header(hl, ' ');
    f.read(&header[0], hl);
    RACER_REQUIRE(f.good(), "truncated .npy header");
    bool f4 = header.find("'<f4'") != std::string::npos, f8 = header.find("'<f8'") != std::string::npos;
    RACER_REQUIRE(f4 || f8, "checkpoint must hold float32 or float64 values");
    RACER_REQUIRE(header.find("'fortran_order': False") != std::string::npos, "checkpoint must be C-ordered");
    size_t lp = header.find("'shape': ("), rp = header.find(')', lp);
    RACER_REQUIRE(lp != std::string::npos && rp != std::string::npos, "checkpoint has no shape");
    std::string dims = header.substr(lp + 10, rp - lp - 10);
    size_t count = 1;
    std::stringstream ss(dims);
    std::string item;
    while (std::getline(ss, item, ','))
        if (item.find_first_not_of(' ') != std::string::npos) count *= (size_t)std::stoull(item);
    RACER_REQUIRE(count == (size_t)p->P + 1, "checkpoint does not match the network size");
    std::vector<float> values(count);
    if (f4) {
        f.read(reinterpret_cast<char*>(values.data()), (std::streamsize)(count * 4));
    } else {
        std::vector<double> tmp(count);
        f.read(reinterpret_cast<char*>(tmp.data()), (std::streamsize)(count * 8));
        for (size_t i = 0; i < count; ++i) values[i] = (float)tmp[i];
    }
    RACER_REQUIRE(f.good(), "truncated checkpoint");
    if (racer_write(p->theta, values.data() + 1)) return 1;
    if (racer_policy_transpose(p)) return 1;
    if (iteration) *iteration = (int)values[0];
    return 0;
}

// ---------------------------------------------------------------- runner
extern "C" int racer_runner_step(racer_sim* sim, racer_policy* p, racer_acts* acts, racer_buf* scratch,
                                 int deterministic, int64_t t) {
    racer_clear_error();
    int64_t n = sim->n;
    RACER_REQUIRE(p->beams == sim->beams && acts->cols == sim->n && acts->hidden == p->hidden,
                  "runner shapes differ");
    RACER_REQUIRE(scratch && scratch->len >= (size_t)(2 * n), "scratch must hold 2 x cars floats");
    if (racer_policy_act(p, sim->obs->ptr, n, acts, (int)n, sim->actions->ptr, scratch->ptr, scratch->ptr + n,
                         deterministic, t, sim->seed))
        return 1;
    return racer_launch_step(sim, sim->actions->ptr, sim->reward->ptr, sim->done->ptr, sim->obs->ptr, n);
}

extern "C" int racer_benchmark(racer_sim* sim, racer_policy* p, racer_acts* acts, racer_buf* scratch, int steps,
                               double* ms_per_step) {
    racer_clear_error();
    for (int t = 0; t < 10; ++t)
        if (racer_runner_step(sim, p, acts, scratch, 0, t)) return 1;
    RACER_TRY(hipStreamSynchronize(sim->dev->stream));
    auto start = std::chrono::steady_clock::now();
    for (int t = 0; t < steps; ++t)
        if (racer_runner_step(sim, p, acts, scratch, 0, t)) return 1;
    RACER_TRY(hipStreamSynchronize(sim->dev->stream));
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    *ms_per_step = steps > 0 ? ms / steps : 0.0;
    return 0;
}
