// The whole policy forward for the bf16 precision in one kernel
// (racer/ppo.mojo:108-111, then Sim.sample, racer/sim.mojo:337-348):
// h1 = tanh(W0 x), h2 = tanh(W1 h1) on MFMA (bf16 operands, float32
// accumulation), o = W2 h2 in float32, and optionally the Gaussian actions.
//
// One workgroup owns FT = 16 samples (columns) and all H rows. Eight
// wavefronts each own 32 rows (two 16-row MFMA tiles). Weight fragments come
// straight from the row-padded bf16 weight copies as 16-byte loads (8
// consecutive k of one row per lane, the layout v_mfma_f32_16x16x32_bf16
// takes); the input tile is staged in LDS as [column][k] bf16. h1 stays in LDS
// as the next product's input; h2 goes to LDS in float32 for the output layer,
// which accumulates fmaf over k in order exactly as the float32 path does.
#include "racer_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int FT = 16;       // columns per workgroup
constexpr int WAVES = 8;     // 32 rows each: H <= 256
constexpr int KS = 128;      // k per input stage
constexpr int LDX = KS + 8;  // LDS row (column-major tile) stride in bf16
constexpr int W1_STEPS = 9;  // k-steps of the second product held in registers (H + 1 <= 288)

__device__ __forceinline__ uint16_t bf16_of(float f) {
    uint32_t u = racer_f2u(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <bool SAMPLE>
__global__ void __launch_bounds__(64 * WAVES) k_policy_fused(
    const uint16_t* __restrict__ W0, int64_t ld0, const uint16_t* __restrict__ W1, int64_t ld1,
    const float* __restrict__ W2, const float* __restrict__ x, int64_t ldx, int H, int IN, int64_t cols,
    float* __restrict__ h1out, float* __restrict__ h2out, float* __restrict__ o, const float* __restrict__ log_std,
    float* act, int64_t act_cols, float* logp, float* value, int deterministic, int64_t step, int seed) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int HA = H + 1;
    const int ha32 = (HA + 31) & ~31;  // k extent of the second product
    uint16_t* xs = reinterpret_cast<uint16_t*>(smem);                 // [2][FT][LDX]
    uint16_t* h1s = xs + 2 * FT * LDX;                                // [FT][ha32 + 8]
    const int ldh = ha32 + 8;
    float* h2f = reinterpret_cast<float*>(h1s + FT * ldh);            // [HA][FT]
    float* w2s = h2f + HA * FT;                                       // [3][HA]
    float* outs = w2s + 3 * HA;                                       // [3][FT]

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t c0 = (int64_t)blockIdx.x * FT;
    const int r = lane & 15, kg = (lane >> 4) * 8;  // MFMA operand lane layout
    const int row0 = wave * 32;                      // this wavefront's rows: row0 .. row0 + 31

    // ---- first product: h1 = tanh(W0 x) ----
    // input staging: thread t takes k = t / 4 (0..127) and columns 4 (t % 4) .. 4 (t % 4) + 3
    const int sk = tid >> 2, sc = (tid & 3) * 4;
    float xr[4];
    auto load_x = [&](int k0) {
        const int k = k0 + sk;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t c = c0 + sc + e;
            xr[e] = (k < IN && c < cols) ? x[(int64_t)k * ldx + c] : 0.0f;
        }
    };
    auto store_x = [&](int buf) {
        uint16_t* dst = xs + buf * FT * LDX;
#pragma unroll
        for (int e = 0; e < 4; ++e) dst[(sc + e) * LDX + sk] = bf16_of(xr[e]);
    };
    // weight fragments of this wavefront's two row tiles for one stage (4 k-steps)
    constexpr int KSTEPS = KS / 32;
    uint4 wa[2][KSTEPS], wn[2][KSTEPS];
    auto load_w0 = [&](int k0, uint4 (&w)[2][KSTEPS]) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) {
                const int row = row0 + 16 * t + r;
                const int k = k0 + 32 * s + kg;
                w[t][s] = (row < H && k < IN) ? *reinterpret_cast<const uint4*>(W0 + (int64_t)row * ld0 + k)
                                              : make_uint4(0, 0, 0, 0);
            }
    };
    f32x4 acc[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int idx = tid; idx < 3 * HA; idx += blockDim.x) w2s[idx] = W2[idx];

    load_x(0);
    load_w0(0, wa);
    store_x(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < IN; k0 += KS) {
        const bool more = k0 + KS < IN;
        if (more) {
            load_x(k0 + KS);
            load_w0(k0 + KS, wn);
        }
        const uint16_t* xb = xs + buf * FT * LDX;
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) {
            bf16x8 b = *reinterpret_cast<const bf16x8*>(xb + r * LDX + 32 * s + kg);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                bf16x8 a = *reinterpret_cast<const bf16x8*>(&wa[t][s]);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
            }
        }
        if (more) {
            store_x(buf ^ 1);
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int s = 0; s < KSTEPS; ++s) wa[t][s] = wn[t][s];
        }
        __syncthreads();
        buf ^= 1;
    }
    // request the second product's weight fragments now, so their latency
    // overlaps the first product's epilogue
    uint4 w1f[2][W1_STEPS];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int s = 0; s < W1_STEPS; ++s) {
            const int row = row0 + 16 * t + r;
            const int k = 32 * s + kg;
            w1f[t][s] = (row < H && 32 * s < ha32) ? *reinterpret_cast<const uint4*>(W1 + (int64_t)row * ld1 + k)
                                                    : make_uint4(0, 0, 0, 0);
        }
    // D[i][j] of a 16 x 16 tile: column j = lane % 16, rows 4 (lane / 16) + q
    const int jc = lane & 15, ib = (lane >> 4) * 4;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = row0 + 16 * t + ib + q;
            if (row < H) {
                float v = racer_tanh(acc[t][q]);
                h1s[jc * ldh + row] = bf16_of(v);
                if (c0 + jc < cols) h1out[(int64_t)row * cols + c0 + jc] = v;
            }
        }
    // bias row (the constant 1) and zero padding of the second product's k range
    for (int idx = tid; idx < FT * (ha32 - H); idx += blockDim.x) {
        int j = idx / (ha32 - H), k = H + idx % (ha32 - H);
        h1s[j * ldh + k] = k == H ? bf16_of(1.0f) : (uint16_t)0;
    }
    __syncthreads();

    // ---- second product: h2 = tanh(W1 h1) ----
    // W1 fragments were requested before the first product's epilogue (w1f);
    // up to W1_STEPS k-steps are held in registers, any further ones load here.
#pragma unroll
    for (int t = 0; t < 2; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < W1_STEPS; ++s) {
        if (32 * s < ha32) {
            bf16x8 b = *reinterpret_cast<const bf16x8*>(h1s + r * ldh + 32 * s + kg);
#pragma unroll
            for (int t = 0; t < 2; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<bf16x8*>(&w1f[t][s]), b, acc[t], 0,
                                                                 0, 0);
        }
    }
    for (int k0 = 32 * W1_STEPS; k0 < ha32; k0 += 32) {
        bf16x8 b = *reinterpret_cast<const bf16x8*>(h1s + r * ldh + k0 + kg);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int row = row0 + 16 * t + r;
            uint4 w = row < H ? *reinterpret_cast<const uint4*>(W1 + (int64_t)row * ld1 + k0 + kg)
                              : make_uint4(0, 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<bf16x8*>(&w), b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = row0 + 16 * t + ib + q;
            if (row < H) {
                float v = racer_tanh(acc[t][q]);
                h2f[row * FT + jc] = v;
                if (c0 + jc < cols) h2out[(int64_t)row * cols + c0 + jc] = v;
            }
        }
    for (int j = tid; j < FT; j += blockDim.x) h2f[H * FT + j] = 1.0f;
    __syncthreads();

    // ---- output layer o = W2 h2 in float32, k in order ----
    if (tid < 3 * FT) {
        const int j = tid / FT, c = tid % FT;
        const float* w = w2s + j * HA;
        float a = 0.0f;
        for (int k = 0; k < HA; ++k) a = fmaf(w[k], h2f[k * FT + c], a);
        outs[j * FT + c] = a;
        if (c0 + c < cols) o[j * cols + c0 + c] = a;
    }
    if constexpr (SAMPLE) {
        __syncthreads();
        const int64_t col = c0 + tid;
        if (tid < FT && col < cols) {
            float noise[4];
            racer_normal4(racer_random((uint64_t)(int64_t)seed + 1, (uint64_t)col, (uint64_t)step), noise);
            float probability = 0.0f;
            for (int a = 0; a < 2; ++a) {
                float z = deterministic ? 0.0f : noise[a];
                float ls = log_std[a];
                act[a * act_cols + col] = outs[a * FT + tid] + racer_exp(ls) * z;
                probability = probability - (((0.5f * z) * z + ls) + RACER_HALF_LOG_2PI);
            }
            logp[col] = probability;
            value[col] = outs[2 * FT + tid];
        }
    }
}

}  // namespace

size_t racer_policy_fused_smem(int H) {
    int HA = H + 1, ha32 = (HA + 31) & ~31;
    return (size_t)2 * FT * LDX * 2 + (size_t)FT * (ha32 + 8) * 2 + (size_t)HA * FT * 4 + (size_t)3 * HA * 4 +
           (size_t)3 * FT * 4;
}

int racer_policy_fused(racer_policy* p, const float* x, int64_t x_cols, racer_acts* a, int cols, bool sample,
                       float* act, float* logp, float* value, int deterministic, int64_t step, int seed) {
    racer_device* dev = p->dev;
    const uint16_t* w0 = p->theta_bf16;
    const uint16_t* w1 = p->theta_bf16 + (int64_t)p->hidden * p->in_pad;
    const float* W2 = p->theta->ptr + p->W2;
    const float* log_std = p->theta->ptr + p->LOGSTD;
    dim3 grid((unsigned)((cols + FT - 1) / FT)), block(64 * WAVES);
    size_t shmem = racer_policy_fused_smem(p->hidden);
    if (sample) {
        racer_timed timer(dev, "policy_fused_sample");
        hipLaunchKernelGGL(k_policy_fused<true>, grid, block, shmem, dev->stream, w0, p->in_pad, w1, p->ha_pad, W2, x,
                           x_cols, p->hidden, (int)p->IN, (int64_t)cols, a->h1->ptr, a->h2->ptr, a->o->ptr, log_std,
                           act, (int64_t)cols, logp, value, deterministic, step, seed);
    } else {
        racer_timed timer(dev, "policy_fused");
        hipLaunchKernelGGL(k_policy_fused<false>, grid, block, shmem, dev->stream, w0, p->in_pad, w1, p->ha_pad, W2,
                           x, x_cols, p->hidden, (int)p->IN, (int64_t)cols, a->h1->ptr, a->h2->ptr, a->o->ptr,
                           log_std, act, (int64_t)cols, logp, value, deterministic, step, seed);
    }
    return racer_launch_check("policy_fused");
}
