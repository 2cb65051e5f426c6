// The whole policy forward for the bf16 precision in one kernel
// (racer/ppo.mojo:108-111, then Sim.sample, racer/sim.mojo:337-348):
// h1 = tanh(W0 x), h2 = tanh(W1 h1) on MFMA (bf16 operands, float32
// accumulation), o = W2 h2 in float32, and optionally the Gaussian actions.
// The hidden layers' tanh is the branch-free racer_tanh_mlp (a few ulp from
// tanh; the bf16 products' own rounding is far larger), as in the bf16 GEMM
// forward for wider policies, so every bf16 forward computes the same values.
//
// One workgroup owns FT = 16 samples (columns) and all H <= 256 rows, with
// eight wavefronts. The design keeps many loads in flight per wavefront,
// because one workgroup per CU has nothing else to hide latency with:
//  - phase 0: the whole [IN x 16] input slice is loaded at once, rounded to
//    bf16 and stored in LDS as [column][k] (16-byte stores);
//  - phase 1 (h1): wavefront w takes rows 64 (w % 4) .. + 63 (four 16-row MFMA
//    tiles) over half of the k range (w / 4); weight fragments (8 consecutive
//    k of one row per lane, v_mfma_f32_16x16x32_bf16's operand layout) stream
//    from a fragment-ordered bf16 copy of the weights (one contiguous 1 KB per
//    tile and k-step) through a 4-k-step register ring; the two halves are
//    added in a fixed order through LDS;
//  - phase 2 (h2): wavefront w takes rows 32 w .. + 31 over all k, its W1
//    fragments requested before phase 1's reduction;
//  - phase 3: o = W2 h2 in float32 with fmaf over k in order (as the float32
//    path), then the sample.
#include <type_traits>

#include "racer_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int FT = 16;       // columns per workgroup
constexpr int WAVES = 8;
constexpr int RING = 4;      // k-steps of weight fragments in flight per wavefront
constexpr int W1_STEPS = 9;  // k-steps of the second product held in registers (H + 1 <= 288)
constexpr int FRAG_TILES = 16;  // 16-row tiles of a fragment-order weight copy (256 rows, zeros past H)

// Round to nearest even with gfx950's v_cvt_pk_bf16_f32 (equal to the integer
// rounding (u + 0x7fff + lsb) >> 16 for every finite float; a NaN comes out quiet).
__device__ __forceinline__ uint16_t bf16_of(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

struct FusedLayout {
    int in32, kx, ha32, ldh, HA, ldo;  // ldo: row stride (floats) of h2 [column][k] and w2 [3][k]
    size_t xs, red, h1s, h2f, w2s, outs, total;  // byte offsets
};

// The output layer o[j][c] = sum_k w[k] h[k] of one output and column in
// float32, an fmaf chain from 0 with k in order (so the same bits as one
// fmaf per k), reading 4 consecutive k of the weight row and of the
// column's h2 at once (both 16-byte aligned, k contiguous).
__device__ __forceinline__ float output_chain(const float* w, const float* h, int HA) {
    float a = 0.0f;
    const int k4 = HA & ~3;
#pragma unroll 8
    for (int k = 0; k < k4; k += 4) {
        const float4 wv = *reinterpret_cast<const float4*>(w + k);
        const float4 hv = *reinterpret_cast<const float4*>(h + k);
        a = fmaf(wv.x, hv.x, a);
        a = fmaf(wv.y, hv.y, a);
        a = fmaf(wv.z, hv.z, a);
        a = fmaf(wv.w, hv.w, a);
    }
    for (int k = k4; k < HA; ++k) a = fmaf(w[k], h[k], a);
    return a;
}

// wide: the reduction buffer and h2 also reuse the input slice's bytes, which
// the first product's MFMA loop is the last to read (a barrier separates them),
// so three workgroups fit one CU's LDS.
__host__ __device__ inline FusedLayout fused_layout(int H, int IN, bool wide = false) {
    FusedLayout L;
    L.HA = H + 1;
    L.in32 = (IN + 31) & ~31;
    L.kx = L.in32 + 8;
    L.ha32 = (L.HA + 31) & ~31;
    L.ldh = L.ha32 + 8;
    L.ldo = (L.HA + 3) & ~3;
    const size_t xs_bytes = (size_t)FT * L.kx * 2;
    const size_t red_bytes = (size_t)4 * 4 * 64 * 4 * 4;  // [4 row groups][4 tiles][64 lanes][4]
    const size_t h2_bytes = (size_t)FT * L.ldo * 4;
    const size_t rh_bytes = red_bytes > h2_bytes ? red_bytes : h2_bytes;
    L.xs = 0;
    if (wide) {
        L.red = 0;
        L.h1s = xs_bytes > rh_bytes ? xs_bytes : rh_bytes;
        L.w2s = L.h1s + (size_t)FT * L.ldh * 2;
    } else {
        L.h1s = xs_bytes;
        L.red = L.h1s + (size_t)FT * L.ldh * 2;
        L.w2s = L.red + rh_bytes;
    }
    L.h2f = L.red;  // h2 reuses the reduction buffer
    L.outs = L.w2s + (size_t)3 * L.ldo * 4;
    L.total = L.outs + (size_t)3 * FT * 4;
    return L;
}

// One workgroup's 16 columns c0 .. c0 + 15 of the forward (and sample).
// HS > 0: each half of the first product's k range is exactly HS k-steps and
// the fragment ring is unrolled over them (straight-line code, so each k-step
// waits only for its own fragments); HS = 0: any k range, as a loop.
// WIDE, for launches of many workgroups per CU: a two-step ring, the second
// product's fragments loaded as it goes rather than held from phase 1, and
// the aliased layout, so three workgroups share a CU and one's loads overlap
// another's MFMA and tanh work (with one workgroup per CU the deep ring and the
// early W1 requests hide the latency instead). Same arithmetic, same results.
template <bool SAMPLE, int HS, bool VEC, bool WIDE>
__device__ __forceinline__ void policy_tile(
    const uint16_t* __restrict__ F0, const uint16_t* __restrict__ F1,
    const float* __restrict__ W2, const float* __restrict__ x, int64_t ldx, int H, int IN, int64_t cols,
    float* __restrict__ h1out, float* __restrict__ h2out, float* __restrict__ o, const float* __restrict__ log_std,
    float* act, int64_t act_cols, float* logp, float* value, int deterministic, int64_t step,
    const int64_t* step_base, int seed, const int64_t c0) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const FusedLayout Ly = fused_layout(H, IN, WIDE);
    const int HA = Ly.HA;
    constexpr int R = WIDE ? 2 : RING;  // k-steps of weight fragments in flight
    uint16_t* xs = reinterpret_cast<uint16_t*>(smem + Ly.xs);     // [FT][kx]
    uint16_t* h1s = reinterpret_cast<uint16_t*>(smem + Ly.h1s);   // [FT][ldh]
    float* red = reinterpret_cast<float*>(smem + Ly.red);         // [4][4][64][4]
    float* h2f = reinterpret_cast<float*>(smem + Ly.h2f);         // [FT][ldo]
    float* w2s = reinterpret_cast<float*>(smem + Ly.w2s);         // [3][ldo]
    float* outs = reinterpret_cast<float*>(smem + Ly.outs);       // [3][FT]

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, kg = (lane >> 4) * 8;  // MFMA operand lane layout
    // buffer descriptors: out-of-range offsets (rows >= H, k >= IN, columns >= cols) read zeros
    const auto xres = racer_rsrc(x, ((uint64_t)(IN - 1) * ldx + cols) * 4);
    const int steps0 = Ly.in32 / 32, steps1 = Ly.ha32 / 32;
    const auto w0res = racer_rsrc(F0, (uint64_t)FRAG_TILES * steps0 * 1024);
    const auto w1res = racer_rsrc(F1, (uint64_t)FRAG_TILES * steps1 * 1024);

    // ---- phase 0: the input slice, bf16, [column][k] ----
    if constexpr (VEC) {
        // 16-byte loads of 4 adjacent columns (x 16-byte aligned, ldx and cols
        // multiples of 4): a thread takes 8 consecutive k of 4 columns with 8
        // loads and stores 4 columns' 8 k as 16-byte LDS writes
        // two items per thread in flight at once (1086 inputs: 544 items for
        // 512 threads), so the slice costs one HBM round trip, not two
        const int chunks = Ly.in32 / 8, items = chunks * (FT / 4);
        const int nthreads = (int)blockDim.x;
        for (int base = 0; base < items; base += 2 * nthreads) {
            float4 f[2][8];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int idx = base + tid + h * nthreads;
                const int cq = (idx % (FT / 4)) * 4, kc = (idx / (FT / 4)) * 8;
                const int64_t col = c0 + cq;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = kc + e;
                    const uint32_t off = (idx < items && k < IN && col < cols)
                                             ? (uint32_t)(((int64_t)k * ldx + col) * 4) : RACER_OOB;
                    f[h][e] = __builtin_bit_cast(float4, racer_ld_b128(xres, off));
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int idx = base + tid + h * nthreads;
                if (idx < items) {
                    const int cq = (idx % (FT / 4)) * 4, kc = (idx / (FT / 4)) * 8;
                    uint16_t v[4][8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        v[0][e] = bf16_of(f[h][e].x);
                        v[1][e] = bf16_of(f[h][e].y);
                        v[2][e] = bf16_of(f[h][e].z);
                        v[3][e] = bf16_of(f[h][e].w);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        *reinterpret_cast<uint4*>(xs + (cq + q) * Ly.kx + kc) = *reinterpret_cast<uint4*>(v[q]);
                }
            }
        }
        for (int idx = tid; idx < 3 * HA; idx += blockDim.x) w2s[(idx / HA) * Ly.ldo + idx % HA] = W2[idx];
    } else {
        const int chunks = Ly.in32 / 8;  // 8 consecutive k per chunk
        for (int idx = tid; idx < chunks * FT; idx += blockDim.x) {
            const int c = idx % FT, kc = (idx / FT) * 8;
            const int64_t col = c0 + c;
            float f[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = kc + e;
                const uint32_t off = (k < IN && col < cols) ? (uint32_t)(((int64_t)k * ldx + col) * 4) : RACER_OOB;
                f[e] = racer_ld_f32(xres, off);
            }
            uint16_t v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = bf16_of(f[e]);
            *reinterpret_cast<uint4*>(xs + c * Ly.kx + kc) = *reinterpret_cast<uint4*>(v);
        }
        for (int idx = tid; idx < 3 * HA; idx += blockDim.x) w2s[(idx / HA) * Ly.ldo + idx % HA] = W2[idx];
    }
    __syncthreads();

    // ---- phase 1: h1 = tanh(W0 x) ----
    const int mg = wave & 3, kh = wave >> 2;
    const int steps = Ly.in32 / 32;
    const int s0 = kh == 0 ? 0 : (steps + 1) / 2, s1 = kh == 0 ? (steps + 1) / 2 : steps;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    uint4 ring[R][4];
    auto load_step = [&](int s, uint4 (&w)[4]) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int tile = 4 * mg + t;
            const uint32_t off = s < s1 ? (uint32_t)(((tile * steps0 + s) * 64 + lane) * 16) : RACER_OOB;
            w[t] = racer_ld_b128(w0res, off);
        }
    };
#pragma unroll
    for (int u = 0; u < R; ++u) load_step(s0 + u, ring[u]);
    // No branches inside the ring: a k-step past s1 has all-zero weight
    // fragments (out-of-range loads) against finite, clamped input columns, so
    // it adds exactly zero, and the compiler can keep the ring's loads in flight.
    if constexpr (HS > 0) {
        // Unrolled: k-step j consumes ring slot j % R and refills it with
        // step j + R, so R - 1 later steps' fragments stay in flight
        // behind each step's MFMAs. (As a loop, the wait at the loop head
        // covers every outstanding load, and the ring degenerates into
        // load-all, wait-all, compute-all.)
#pragma unroll
        for (int j = 0; j < HS; ++j) {
            const int u = j % R;
            const int kk = 32 * (s0 + j);
            bf16x8 b = *reinterpret_cast<const bf16x8*>(xs + r * Ly.kx + kk + kg);
#pragma unroll
            for (int t = 0; t < 4; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<bf16x8*>(&ring[u][t]), b, acc[t],
                                                                 0, 0, 0);
            // pinned here, so R - 1 k-steps of fragments stay in flight
            // (the scheduler otherwise sinks each refill to just before its use)
            __builtin_amdgcn_sched_barrier(0);
            if (j + R < HS) load_step(s0 + j + R, ring[u]);
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        for (int s = s0; s < s1; s += R) {
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int kk = min(32 * (s + u), Ly.in32 - 32);
                bf16x8 b = *reinterpret_cast<const bf16x8*>(xs + r * Ly.kx + kk + kg);
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<bf16x8*>(&ring[u][t]), b,
                                                                     acc[t], 0, 0, 0);
                load_step(s + u + R, ring[u]);
            }
        }
    }
    // request the second product's weight fragments now (rows 32 w .. + 31)
    constexpr int W1_HELD = WIDE ? 0 : W1_STEPS;
    uint4 w1f[2][W1_HELD > 0 ? W1_HELD : 1];
    if constexpr (W1_HELD > 0) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int s = 0; s < W1_HELD; ++s) {
                const int tile = 2 * wave + t;
                const uint32_t off = s < steps1 ? (uint32_t)(((tile * steps1 + s) * 64 + lane) * 16) : RACER_OOB;
                w1f[t][s] = racer_ld_b128(w1res, off);
            }
    }
    // the reduction buffer overlays the input slice: every wavefront's last
    // read of it (its final k-step) comes first
    if constexpr (WIDE) __syncthreads();
    // The two k halves finish two tiles each: half 0 tiles 0 and 1, half 1
    // tiles 2 and 3, each handing the other its partial sums of the other two
    // through LDS. Every element is (first half) + (second half), whichever
    // wavefront adds them (addition commutes), and the epilogue's tanh work is
    // spread over all eight wavefronts instead of four.
    const bool first = kh == 0;
    const f32x4 give[2] = {first ? acc[2] : acc[0], first ? acc[3] : acc[1]};
    const f32x4 keep[2] = {first ? acc[0] : acc[2], first ? acc[1] : acc[3]};
    const int give_t = first ? 2 : 0, keep_t = first ? 0 : 2;
#pragma unroll
    for (int t = 0; t < 2; ++t)
        *reinterpret_cast<f32x4*>(red + ((mg * 4 + give_t + t) * 64 + lane) * 4) = give[t];
    __syncthreads();
    // D[i][j] of a 16 x 16 tile: column j = lane % 16, rows 4 (lane / 16) + q
    const int jc = lane & 15, ib = (lane >> 4) * 4;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int tile = keep_t + t;
        const f32x4 other = *reinterpret_cast<const f32x4*>(red + ((mg * 4 + tile) * 64 + lane) * 4);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 64 * mg + 16 * tile + ib + q;
            if (row < H) {
                float v = racer_tanh_mlp(keep[t][q] + other[q]);
                h1s[jc * Ly.ldh + row] = bf16_of(v);
                if (c0 + jc < cols) h1out[(int64_t)row * cols + c0 + jc] = v;
            }
        }
    }
    // bias row (the constant 1) and zero padding of the second product's k range
    for (int idx = tid; idx < FT * (Ly.ha32 - H); idx += blockDim.x) {
        int j = idx / (Ly.ha32 - H), k = H + idx % (Ly.ha32 - H);
        h1s[j * Ly.ldh + k] = k == H ? bf16_of(1.0f) : (uint16_t)0;
    }
    __syncthreads();

    // ---- phase 2: h2 = tanh(W1 h1) ----
    f32x4 acc2[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
    if constexpr (W1_HELD > 0) {
#pragma unroll
        for (int s = 0; s < W1_HELD; ++s) {
            if (32 * s < Ly.ha32) {
                bf16x8 b = *reinterpret_cast<const bf16x8*>(h1s + r * Ly.ldh + 32 * s + kg);
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    acc2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<bf16x8*>(&w1f[t][s]), b,
                                                                      acc2[t], 0, 0, 0);
            }
        }
    }
    for (int k0 = 32 * W1_HELD; k0 < Ly.ha32; k0 += 32) {
        bf16x8 b = *reinterpret_cast<const bf16x8*>(h1s + r * Ly.ldh + k0 + kg);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int tile = 2 * wave + t;
            uint4 w = racer_ld_b128(w1res, (uint32_t)(((tile * steps1 + k0 / 32) * 64 + lane) * 16));
            acc2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<bf16x8*>(&w), b, acc2[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 32 * wave + 16 * t + ib + q;
            if (row < H) {
                float v = racer_tanh_mlp(acc2[t][q]);
                h2f[jc * Ly.ldo + row] = v;
                if (c0 + jc < cols) h2out[(int64_t)row * cols + c0 + jc] = v;
            }
        }
    for (int j = tid; j < FT; j += blockDim.x) h2f[j * Ly.ldo + H] = 1.0f;
    __syncthreads();

    // ---- phase 3: output layer o = W2 h2 in float32, k in order ----
    if (tid < 3 * FT) {
        const int j = tid / FT, c = tid % FT;
        const float a = output_chain(w2s + j * Ly.ldo, h2f + c * Ly.ldo, HA);
        outs[j * FT + c] = a;
        if (c0 + c < cols) o[j * cols + c0 + c] = a;
    }
    if constexpr (SAMPLE) {
        __syncthreads();
        const int64_t col = c0 + tid;
        if (tid < FT && col < cols) {
            if (step_base) step += *step_base;
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

template <bool SAMPLE, int HS, bool VEC, bool WIDE>
__global__ void __launch_bounds__(64 * WAVES, WIDE ? 6 : 2) k_policy_fused(
    const uint16_t* __restrict__ F0, const uint16_t* __restrict__ F1,
    const float* __restrict__ W2, const float* __restrict__ x, int64_t ldx, int H, int IN, int64_t cols,
    float* __restrict__ h1out, float* __restrict__ h2out, float* __restrict__ o, const float* __restrict__ log_std,
    float* act, int64_t act_cols, float* logp, float* value, int deterministic, int64_t step,
    const int64_t* step_base, int seed) {
    policy_tile<SAMPLE, HS, VEC, WIDE>(F0, F1, W2, x, ldx, H, IN, cols, h1out, h2out, o, log_std, act, act_cols,
                                       logp, value, deterministic, step, step_base, seed,
                                       (int64_t)xcd_block() * FT);
}

// The forward and sample of several sims' observations in one launch:
// workgroup b serves the instance whose tile range [start[k], start[k + 1])
// holds b, with that instance's buffers and Philox seed.
template <int HS>
__global__ void __launch_bounds__(64 * WAVES) k_policy_fused_batch(
    const uint16_t* __restrict__ F0, const uint16_t* __restrict__ F1, const float* __restrict__ W2, int H, int IN,
    const float* __restrict__ log_std, racer_policy_table tab, int deterministic, int64_t step,
    const int64_t* step_base) {
    const int64_t b = blockIdx.x;
    int k = 0;
    while (k + 1 < tab.count && b >= tab.start[k + 1]) ++k;
    const racer_policy_slot& q = tab.slot[k];
    policy_tile<true, HS, false, false>(F0, F1, W2, q.x, q.ldx, H, IN, q.cols, q.h1, q.h2, q.o, log_std, q.act, q.act_cols, q.logp,
                      q.value, deterministic, step, step_base, q.seed, (b - tab.start[k]) * FT);
}

// Weights in MFMA fragment order: block (tile, step) holds the 64 lanes'
// operands of rows 16 tile .. + 15 and k 32 step .. + 31, lane l taking row
// 16 tile + l % 16 and k 32 step + 8 (l / 16) .. + 7, as 1 KB of bf16. A
// wavefront's operand load for one tile and k-step is then one contiguous 1 KB.
__global__ void k_pack_fragments(const float* __restrict__ src, int64_t ld_src, int rows, int cols, int steps,
                                 uint16_t* __restrict__ dst) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)FRAG_TILES * steps * 512) return;
    const int block = (int)(idx / 512), within = (int)(idx % 512);
    const int tile = block / steps, step = block % steps;
    const int lane = within / 8, e = within % 8;
    const int row = 16 * tile + lane % 16;
    const int k = 32 * step + (lane / 16) * 8 + e;
    dst[idx] = (row < rows && k < cols) ? bf16_of(src[(int64_t)row * ld_src + k]) : (uint16_t)0;
}


// ---- the tiled forward for large launches (the PPO update's minibatches) ----
//
// Every 16-column workgroup above streams all of W0's fragments (557 KB at
// 1086 inputs) for its 16 samples, so at 262144 samples the weight stream,
// not the samples' 4.3 KB each, sets the time. Here a workgroup owns TC = 64
// columns, so each weight fragment serves four 16-column MFMA tiles, and the
// input slice (too large for LDS at 64 columns) is staged through LDS in
// double-buffered blocks of KB k-steps per k half. The arithmetic is the
// fused kernel's: the same MFMA on the same operand layouts, each k half
// accumulated over its k-steps in order and the halves added (half 0 +
// half 1), the same tanh, the second product over all its k-steps in order,
// and the output layer with fmaf over k in order; so the outputs equal the
// fused kernel's bit for bit, and the update sees the values the rollout saw.
constexpr int TC = 64;           // columns per workgroup
constexpr int TCT = TC / 16;     // 16-column MFMA tiles per workgroup
constexpr int TKB = 2;           // k-steps per staged block and k half
constexpr int TXLD = TKB * 32 + 8;  // bf16 per staged column (row pitch, padded)
constexpr int TRING = 2;         // k-steps of weight fragments in flight per wavefront
constexpr int TDX = 2;           // staged blocks of inputs in flight (register sets)

struct TiledLayout {
    int HA, ha32, ldh, ldo;  // ldo: row stride (floats) of h2 [column][k] and w2 [3][k]
    size_t xst, h2f, h1s, w2s, total;
};

__host__ __device__ inline TiledLayout tiled_layout(int H) {
    TiledLayout L;
    L.HA = H + 1;
    L.ha32 = (L.HA + 31) & ~31;
    L.ldh = L.ha32 + 8;
    L.ldo = (L.HA + 3) & ~3;
    const size_t xst_bytes = (size_t)2 * 2 * TC * TXLD * 2;       // [buffer][half][column][k]
    const size_t red_bytes = (size_t)4 * 4 * 2 * 64 * 4 * 4;      // one round: [4][4 row tiles][2 column tiles][64][4]
    const size_t h2_bytes = (size_t)TC * L.ldo * 4;               // [TC][ldo]
    size_t r0 = xst_bytes > red_bytes ? xst_bytes : red_bytes;
    r0 = r0 > h2_bytes ? r0 : h2_bytes;
    L.xst = 0;
    L.h2f = 0;  // the staged inputs, the halves' exchange and h2 share region 0 in turn
    L.h1s = r0;
    L.w2s = L.h1s + (size_t)TC * L.ldh * 2;
    L.total = L.w2s + (size_t)3 * L.ldo * 4;
    return L;
}

// XB16: the inputs are already bf16 (rounded by the same conversion), read as
// 8-byte loads of 4 columns: the same bits on half of the inputs' bytes.
template <int HS, bool XB16>
__global__ void __launch_bounds__(64 * WAVES) k_policy_tiled(
    const uint16_t* __restrict__ F0, const uint16_t* __restrict__ F1, const float* __restrict__ W2,
    const void* __restrict__ x, int64_t ldx, int H, int IN, int64_t cols, float* __restrict__ h1out,
    float* __restrict__ h2out, float* __restrict__ o) {
    static_assert(HS > 0, "the tiled forward needs equal unrolled k halves");
    constexpr int NB = (HS + TKB - 1) / TKB;  // staged blocks per k half
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const TiledLayout Ly = tiled_layout(H);
    const int HA = Ly.HA;
    uint16_t* xst = reinterpret_cast<uint16_t*>(smem + Ly.xst);   // [2][2][TC][TXLD]
    float* red = reinterpret_cast<float*>(smem + Ly.xst);         // [4][4][2][64][4] per round
    float* h2f = reinterpret_cast<float*>(smem + Ly.h2f);         // [HA][TC]
    uint16_t* h1s = reinterpret_cast<uint16_t*>(smem + Ly.h1s);   // [TC][ldh]
    float* w2s = reinterpret_cast<float*>(smem + Ly.w2s);         // [3][ldo]

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, kg = (lane >> 4) * 8;  // MFMA operand lane layout
    const int64_t c0 = (int64_t)xcd_block() * TC;
    const int steps0 = 2 * HS, steps1 = Ly.ha32 / 32;
    constexpr int XSZ = XB16 ? 2 : 4;  // bytes of one input element
    const auto xres = racer_rsrc(x, ((uint64_t)(IN - 1) * ldx + cols) * XSZ);
    const auto w0res = racer_rsrc(F0, (uint64_t)FRAG_TILES * steps0 * 1024);
    const auto w1res = racer_rsrc(F1, (uint64_t)FRAG_TILES * steps1 * 1024);
    for (int idx = tid; idx < 3 * HA; idx += blockDim.x) w2s[(idx / HA) * Ly.ldo + idx % HA] = W2[idx];

    // staging: thread tid covers k-half tid / 256 and 4 adjacent columns
    // (cq); piece m of a block is 4 consecutive k at kq(m), as 4 float4 loads
    // (consecutive lanes along the columns: 256 contiguous bytes per k row;
    // lanes along k instead avoided LDS bank conflicts on the stores but ran
    // slower)
    static_assert(TKB % 2 == 0, "512 threads stage a block as TKB / 2 pieces each");
    constexpr int PIECES = TKB / 2;
    const int sh = tid >> 8, cq = (tid & 15) * 4;
    const int64_t scol = c0 + cq;
    auto kq = [&](int m) { return (((tid & 255) >> 4) + 16 * m) * 4; };
    // blocks b .. b + TDX - 1 are in flight in register set b % TDX (HBM
    // latency is several blocks' worth of MFMA work)
    using XReg = std::conditional_t<XB16, uint2, float4>;
    XReg xr[TDX][PIECES][4];
    auto stage_load = [&](int b) {
#pragma unroll
        for (int m = 0; m < PIECES; ++m)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int kin = b * TKB * 32 + kq(m) + e;  // k within the half
                const int k = sh * HS * 32 + kin;
                const bool ok = kin < HS * 32 && k < IN && scol < cols;
                const uint32_t off = ok ? (uint32_t)(((int64_t)k * ldx + scol) * XSZ) : RACER_OOB;
                if constexpr (XB16)
                    xr[b % TDX][m][e] = __builtin_bit_cast(uint2, racer_ld_f32x2(xres, off));
                else
                    xr[b % TDX][m][e] = __builtin_bit_cast(float4, racer_ld_b128(xres, off));
            }
    };
    auto stage_store = [&](int b) {
#pragma unroll
        for (int m = 0; m < PIECES; ++m) {
            uint16_t* dst = xst + ((size_t)((b & 1) * 2 + sh) * TC + cq) * TXLD + kq(m);
            uint16_t v[4][4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if constexpr (XB16) {
                    const uint2 w = xr[b % TDX][m][e];
                    v[0][e] = (uint16_t)w.x;
                    v[1][e] = (uint16_t)(w.x >> 16);
                    v[2][e] = (uint16_t)w.y;
                    v[3][e] = (uint16_t)(w.y >> 16);
                } else {
                    v[0][e] = bf16_of(xr[b % TDX][m][e].x);
                    v[1][e] = bf16_of(xr[b % TDX][m][e].y);
                    v[2][e] = bf16_of(xr[b % TDX][m][e].z);
                    v[3][e] = bf16_of(xr[b % TDX][m][e].w);
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) *reinterpret_cast<uint2*>(dst + q * TXLD) = *reinterpret_cast<uint2*>(v[q]);
        }
    };

    // ---- first product: wavefront (mg, kh) = rows 64 mg .. + 63 over k half kh, all TC columns ----
    const int mg = wave & 3, kh = wave >> 2;
    f32x4 acc[4][TCT];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int c = 0; c < TCT; ++c) acc[t][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    uint4 ring[TRING][4];
    auto load_step = [&](int j, uint4 (&w)[4]) {  // j: k-step within the half
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int tile = 4 * mg + t;
            const uint32_t off = j < HS ? (uint32_t)(((tile * steps0 + kh * HS + j) * 64 + lane) * 16) : RACER_OOB;
            w[t] = racer_ld_b128(w0res, off);
        }
    };
    // Issue order matters: loads retire in order (one vmcnt counter), so a
    // wait for a k-step's weight fragments (from L2) also waits for every
    // input load (from HBM) issued before them. Block b + TDX's inputs are
    // therefore requested at the end of block b, after the block's weight
    // refills: the inputs a weight wait covers are then those of the block
    // that consumes the weights, needed there anyway.
    stage_load(0);
#pragma unroll
    for (int u = 0; u < TRING; ++u) load_step(u, ring[u]);
#pragma unroll
    for (int b = 1; b < TDX; ++b)
        if (b < NB) stage_load(b);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        stage_store(b);
        __syncthreads();
        const uint16_t* xb = xst + (size_t)((b & 1) * 2 + kh) * TC * TXLD;
#pragma unroll
        for (int i = 0; i < TKB; ++i) {
            const int j = b * TKB + i;
            if (j < HS) {
                const int u = j % TRING;
                bf16x8 bx[TCT];
#pragma unroll
                for (int c = 0; c < TCT; ++c)
                    bx[c] = *reinterpret_cast<const bf16x8*>(xb + (c * 16 + r) * TXLD + i * 32 + kg);
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int c = 0; c < TCT; ++c)
                        acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<bf16x8*>(&ring[u][t]),
                                                                            bx[c], acc[t][c], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (j + TRING < HS) load_step(j + TRING, ring[u]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (b + TDX < NB) stage_load(b + TDX);
        __builtin_amdgcn_sched_barrier(0);
    }
    // request the second product's weight fragments now (rows 32 w .. + 31),
    // so their latency hides behind the halves' exchange and the tanh
    constexpr int TW1 = W1_STEPS;  // k-steps held (H + 1 <= 288)
    uint4 w1f[2][TW1];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int s = 0; s < TW1; ++s) {
            const int tile = 2 * wave + t;
            const uint32_t off = s < steps1 ? (uint32_t)(((tile * steps1 + s) * 64 + lane) * 16) : RACER_OOB;
            w1f[t][s] = racer_ld_b128(w1res, off);
        }
    __syncthreads();  // region 0 now holds the halves' exchange: every read of the staged inputs is done

    // halves: half 0 finishes row tiles 0 and 1, half 1 tiles 2 and 3 (each
    // element half 0 + half 1), in two rounds of two column tiles
    // exchange slots are keyed by the absolute row tile: half 0 writes tiles
    // 2 and 3 and reads 0 and 1 (written by half 1), and the reverse
    const bool first = kh == 0;
    const int give_t = first ? 2 : 0, keep_t = first ? 0 : 2;
    const int jc = lane & 15, ib = (lane >> 4) * 4;
#pragma unroll
    for (int rd = 0; rd < TCT / 2; ++rd) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const f32x4 g = first ? acc[2 + t][2 * rd + c] : acc[t][2 * rd + c];
                *reinterpret_cast<f32x4*>(red + (((mg * 4 + give_t + t) * 2 + c) * 64 + lane) * 4) = g;
            }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const f32x4 keep = first ? acc[t][2 * rd + c] : acc[2 + t][2 * rd + c];
                const f32x4 other =
                    *reinterpret_cast<const f32x4*>(red + (((mg * 4 + keep_t + t) * 2 + c) * 64 + lane) * 4);
                const int col = (2 * rd + c) * 16 + jc;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = 64 * mg + 16 * (keep_t + t) + ib + q;
                    if (row < H) {
                        const float v = racer_tanh_mlp(keep[q] + other[q]);
                        h1s[col * Ly.ldh + row] = bf16_of(v);
                        if (c0 + col < cols) h1out[(int64_t)row * cols + c0 + col] = v;
                    }
                }
            }
        __syncthreads();
    }
    // bias row (the constant 1) and zero padding of the second product's k range
    for (int idx = tid; idx < TC * (Ly.ha32 - H); idx += blockDim.x) {
        const int j = idx / (Ly.ha32 - H), k = H + idx % (Ly.ha32 - H);
        h1s[j * Ly.ldh + k] = k == H ? bf16_of(1.0f) : (uint16_t)0;
    }
    __syncthreads();

    // ---- second product: wavefront w = rows 32 w .. + 31 over all k, all TC columns ----
    f32x4 acc2[2][TCT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < TCT; ++c) acc2[t][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < TW1; ++s) {
        if (s < steps1) {
#pragma unroll
            for (int c = 0; c < TCT; ++c) {
                const bf16x8 bh = *reinterpret_cast<const bf16x8*>(h1s + (c * 16 + r) * Ly.ldh + 32 * s + kg);
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    acc2[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<bf16x8*>(&w1f[t][s]), bh,
                                                                         acc2[t][c], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < TCT; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = 32 * wave + 16 * t + ib + q, col = c * 16 + jc;
                if (row < H) {
                    const float v = racer_tanh_mlp(acc2[t][c][q]);
                    h2f[col * Ly.ldo + row] = v;
                    if (c0 + col < cols) h2out[(int64_t)row * cols + c0 + col] = v;
                }
            }
    for (int j = tid; j < TC; j += blockDim.x) h2f[j * Ly.ldo + H] = 1.0f;
    __syncthreads();

    // ---- output layer o = W2 h2 in float32, k in order ----
    if (tid < 3 * TC) {
        const int j = tid / TC, c = tid % TC;
        const float a = output_chain(w2s + j * Ly.ldo, h2f + c * Ly.ldo, HA);
        if (c0 + c < cols) o[j * cols + c0 + c] = a;
    }
}

}  // namespace

size_t racer_fragment_elems(int cols) { return (size_t)FRAG_TILES * ((cols + 31) / 32) * 512; }

int racer_pack_fragments(racer_device* dev, const float* src, int64_t ld_src, int rows, int cols, uint16_t* dst) {
    const int steps = (cols + 31) / 32;
    const int64_t total = (int64_t)FRAG_TILES * steps * 512;
    {
        racer_timed timer(dev, "pack_fragments");
        hipLaunchKernelGGL(k_pack_fragments, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, dev->stream, src,
                           ld_src, rows, cols, steps, dst);
    }
    return racer_launch_check("pack_fragments");
}

size_t racer_policy_fused_smem(int H, int IN) { return fused_layout(H, IN).total; }

// Allows the fused kernel more than the default 64 KB of dynamic LDS (gfx950
// has 160 KB per CU); called when a policy that uses it is created.
// The unrolled ring's instantiations: half k ranges of 17 k-steps (1086
// inputs: 1081 beams) and 2 (113 inputs: 108 beams); other shapes loop.
template <int HS>
static int fused_setup_one() {
    const void* kernels[] = {
        (const void*)k_policy_fused<true, HS, false, false>, (const void*)k_policy_fused<false, HS, false, false>,
        (const void*)k_policy_fused<true, HS, true, false>,  (const void*)k_policy_fused<false, HS, true, false>,
        (const void*)k_policy_fused<true, HS, true, true>,   (const void*)k_policy_fused<false, HS, true, true>};
    for (const void* k : kernels)
        RACER_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    RACER_TRY(hipFuncSetAttribute((const void*)k_policy_fused_batch<HS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024));
    return 0;
}

int racer_policy_fused_setup() {
    if (fused_setup_one<0>() || fused_setup_one<2>() || fused_setup_one<17>()) return 1;
    for (const void* k : {(const void*)k_policy_tiled<2, false>, (const void*)k_policy_tiled<17, false>,
                          (const void*)k_policy_tiled<2, true>, (const void*)k_policy_tiled<17, true>})
        RACER_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    return 0;
}

// HS for a first product of IN inputs: its k-steps split into two equal halves of HS.
static int unrolled_half(int IN) {
    const int steps = (IN + 31) / 32;
    return steps == 34 ? 17 : (steps == 4 ? 2 : 0);
}

int racer_policy_fused_batch(racer_policy* p, const racer_policy_table& tab, int deterministic, int64_t step) {
    racer_device* dev = p->dev;
    if (tab.count <= 0 || tab.start[tab.count] <= 0) return 0;
    for (int k = 0; k < tab.count; ++k)
        RACER_REQUIRE(racer_fused_reach((int)p->IN, tab.slot[k].ldx, tab.slot[k].cols),
                      "bf16 policy input exceeds the 32-bit offsets of the buffer loads (4 GiB)");
    dim3 grid((unsigned)tab.start[tab.count]), block(64 * WAVES);
    size_t shmem = racer_policy_fused_smem(p->hidden, (int)p->IN);
    {
        racer_timed timer(dev, "policy_fused_sample");
        const int hs = unrolled_half((int)p->IN);
        auto kernel = hs == 17 ? k_policy_fused_batch<17> : (hs == 2 ? k_policy_fused_batch<2> : k_policy_fused_batch<0>);
        hipLaunchKernelGGL(kernel, grid, block, shmem, dev->stream, p->frag0, p->frag1, p->theta->ptr + p->W2,
                           p->hidden, (int)p->IN, p->theta->ptr + p->LOGSTD, tab, deterministic, step,
                           dev->step_base);
    }
    return racer_launch_check("policy_fused_batch");
}

// The kernel for one launch: the unrolled half (HS) of the input width,
// 16-byte input loads when the columns allow them, and the wide variant when
// the grid holds at least WIDE_MIN_TILES workgroups (four per CU).
constexpr unsigned WIDE_MIN_TILES = 1024;
// The forward without the sample over at least this many columns (the PPO
// update's minibatches) takes the tiled kernel.
constexpr int TILED_MIN_COLS = 65536;
template <bool SAMPLE>
static auto fused_kernel(int hs, bool vec, bool wide) {
    if (vec && wide)
        return hs == 17 ? k_policy_fused<SAMPLE, 17, true, true>
                        : (hs == 2 ? k_policy_fused<SAMPLE, 2, true, true> : k_policy_fused<SAMPLE, 0, true, true>);
    if (vec)
        return hs == 17 ? k_policy_fused<SAMPLE, 17, true, false>
                        : (hs == 2 ? k_policy_fused<SAMPLE, 2, true, false> : k_policy_fused<SAMPLE, 0, true, false>);
    return hs == 17 ? k_policy_fused<SAMPLE, 17, false, false>
                    : (hs == 2 ? k_policy_fused<SAMPLE, 2, false, false> : k_policy_fused<SAMPLE, 0, false, false>);
}

bool racer_tiled_x16_applies(int precision, int hidden, int64_t IN, int64_t x_cols, int64_t cols) {
    return precision == 1 && hidden <= 256 && unrolled_half((int)IN) > 0 && x_cols % 4 == 0 && cols % 4 == 0 &&
           cols >= TILED_MIN_COLS && ((uint64_t)(IN - 1) * x_cols + cols) * 2 < 0xFFFFFF00ull;
}

int racer_policy_tiled_x16(racer_policy* p, const uint16_t* x, int64_t x_cols, racer_acts* a, int cols) {
    racer_device* dev = p->dev;
    RACER_REQUIRE(racer_tiled_x16_applies(p->precision, p->hidden, p->IN, x_cols, cols) && (uintptr_t)x % 8 == 0,
                  "the bf16-input tiled forward does not apply to these shapes");
    if (racer_policy_refresh(p)) return 1;
    const int hs = unrolled_half((int)p->IN);
    const dim3 tgrid((unsigned)((cols + TC - 1) / TC)), block(64 * WAVES);
    const size_t tshmem = tiled_layout(p->hidden).total;
    {
        racer_timed timer(dev, "policy_tiled");
        hipLaunchKernelGGL((hs == 17 ? k_policy_tiled<17, true> : k_policy_tiled<2, true>), tgrid, block, tshmem,
                           dev->stream, p->frag0, p->frag1, p->theta->ptr + p->W2, x, x_cols, p->hidden, (int)p->IN,
                           (int64_t)cols, a->h1->ptr, a->h2->ptr, a->o->ptr);
    }
    return racer_launch_check("policy_tiled");
}

int racer_policy_fused(racer_policy* p, const float* x, int64_t x_cols, racer_acts* a, int cols, bool sample,
                       float* act, float* logp, float* value, int deterministic, int64_t step, int seed, int kernel) {
    racer_device* dev = p->dev;
    RACER_REQUIRE(racer_fused_reach((int)p->IN, x_cols, cols),
                  "bf16 policy input exceeds the 32-bit offsets of the buffer loads (4 GiB)");
    const uint16_t* f0 = p->frag0;
    const uint16_t* f1 = p->frag1;
    const float* W2 = p->theta->ptr + p->W2;
    const float* log_std = p->theta->ptr + p->LOGSTD;
    dim3 grid((unsigned)((cols + FT - 1) / FT)), block(64 * WAVES);
    const int hs = unrolled_half((int)p->IN);
    const bool vec = x_cols % 4 == 0 && cols % 4 == 0 && (uintptr_t)x % 16 == 0;
    const bool tiled_kernel = kernel == RACER_ACT_TILED || (kernel == RACER_ACT_AUTO && cols >= TILED_MIN_COLS);
    if (!sample && vec && hs > 0 && tiled_kernel) {
        racer_timed timer(dev, "policy_tiled");
        const dim3 tgrid((unsigned)((cols + TC - 1) / TC));
        const size_t tshmem = tiled_layout(p->hidden).total;
        hipLaunchKernelGGL((hs == 17 ? k_policy_tiled<17, false> : k_policy_tiled<2, false>), tgrid, block, tshmem,
                           dev->stream, f0,
                           f1, W2, x, x_cols, p->hidden, (int)p->IN, (int64_t)cols, a->h1->ptr, a->h2->ptr,
                           a->o->ptr);
        return racer_launch_check("policy_tiled");
    }
    const bool wide = vec && (kernel == RACER_ACT_WIDE || (kernel != RACER_ACT_FUSED && grid.x >= WIDE_MIN_TILES));
    const size_t shmem = fused_layout(p->hidden, (int)p->IN, wide).total;
    auto launch = sample ? fused_kernel<true>(hs, vec, wide) : fused_kernel<false>(hs, vec, wide);
    racer_timed timer(dev, sample ? "policy_fused_sample" : "policy_fused");
    hipLaunchKernelGGL(launch, grid, block, shmem, dev->stream, f0, f1, W2, x, x_cols, p->hidden, (int)p->IN,
                       (int64_t)cols, a->h1->ptr, a->h2->ptr, a->o->ptr, log_std, act, (int64_t)cols, logp, value,
                       deterministic, step, dev->step_base, seed);
    return racer_launch_check("policy_fused");
}
