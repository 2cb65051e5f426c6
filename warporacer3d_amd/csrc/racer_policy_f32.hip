// The whole float32 policy forward in one kernel (racer/ppo.mojo:108-111, then
// Sim.sample, racer/sim.mojo:337-348): h1 = tanh(W0 x), h2 = tanh(W1 h1) and
// o = W2 h2, bitwise the CPU restatement's float32 products (every output one
// fused multiply-add chain over k in increasing order, and the same tanh).
//
// The hidden layers run on gfx950's K = 1 float32 MFMA,
// v_mfma_f32_16x16x1_4b_f32: D = A B + C for one k, 4 blocks of 16 x 16. Each
// output element of one such MFMA is a single correctly rounded fused
// multiply-add (checked against fmaf chains on random, cancelling and
// subnormal-product operands), so a chain of them over k equals the fmaf
// chain bit for bit, at the matrix cores' float32 rate (twice the non-packed
// vector FMA rate).
//
// One workgroup owns FT = 16 columns (samples) and H <= 256 rows, one
// wavefront per 64 rows (4 MFMA blocks of 16 rows on the same 16 columns):
//  - phase 0: the [IN x 16] input slice goes to LDS as [column][k], so one
//    16-byte LDS read gives a lane 4 k-steps of its column. The loads are
//    issued in rounds of 16 k quads per wavefront; for the network shapes in
//    use every round's loads go out at once, only round 0 is stored before
//    the first product starts, and round r is stored (behind a barrier) just
//    before the product's LDS reads reach it, so the staging hides under the
//    MFMAs;
//  - phase 1 (h1): weights stream from a fragment-ordered float32 copy of W0
//    (block (row group, k quad) = the 64 lanes' rows, 4 consecutive k each,
//    1 KB contiguous) through a register ring, the inputs' LDS reads two
//    quads ahead; one MFMA per k; tanh as racer_tanh_select (racer_tanh
//    without branches, the same bits), 16 independent evaluations per lane;
//  - phase 2 (h2): the same over h1 in LDS with W1's fragments;
//  - phase 3: o = W2 h2 with fmaf over k in order (16-byte LDS reads), then
//    the sample.
#include "racer_internal.h"

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int FT = 16;     // columns per workgroup
constexpr int RING = 8;    // k quads of weight fragments in flight per wavefront
constexpr int BRING = 2;   // k quads of inputs read from LDS ahead of their MFMAs
constexpr int MAX_RG = 4;  // 64-row groups (wavefronts): H <= 256

struct F32Layout {
    int HA, KB0, KB1, ldx, ldh, w2p;
    size_t xs, h1s, h2s, w2s, outs, total;  // byte offsets
};

// LDS pitches of 4 (mod 32) floats: the 16 columns' 16-byte reads of one k
// quad fall in distinct bank groups, two phases per wavefront read.
__host__ __device__ inline int pitch(int kq) { return ((4 * kq + 31) & ~31) + 4; }

__host__ __device__ inline F32Layout f32_layout(int H, int IN) {
    F32Layout L;
    L.HA = H + 1;
    L.KB0 = (IN + 3) / 4;
    L.KB1 = (L.HA + 3) / 4;
    L.ldx = pitch(L.KB0);
    L.ldh = pitch(L.KB1);
    L.w2p = 4 * L.KB1;
    L.xs = 0;
    const size_t xs_bytes = (size_t)FT * L.ldx * 4, h_bytes = (size_t)FT * L.ldh * 4;
    L.h1s = xs_bytes;
    L.w2s = L.h1s + h_bytes;
    L.outs = L.w2s + (size_t)3 * L.w2p * 4;
    L.total = L.outs + (size_t)3 * FT * 4;
    // h2 [FT][ldh] over the input slice (phase 1 is the last to read it) when it fits there
    L.h2s = h_bytes <= xs_bytes ? 0 : L.total;
    if (L.h2s) L.total += h_bytes;
    return L;
}

__device__ __forceinline__ f32x16 mfma1(float a, float b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_16x16x1f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float4 ld4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float4, racer_ld_b128(r, off));
}

// One wavefront's chain over k quads LO .. HI - 1 (of FULL whole quads):
// fragment block kb of this wavefront at byte fbase + 1024 kb, its column's
// inputs at brow[4 kb ..]. Unrolled, each quad's refill pinned behind its
// MFMAs, so RING - 1 later quads of weights stay in flight (a loop's head
// would wait for all of them), and the inputs read BRING quads ahead (a
// dependent MFMA cannot issue before its operands arrive; four MFMAs cover
// only part of an LDS round trip). The ring and read-ahead registers carry
// over from one segment to the next.
template <int FULL, int LO, int HI>
__device__ __forceinline__ void chain_seg(f32x16& acc, float4 (&ring)[RING], float4 (&bq)[BRING],
                                          __amdgpu_buffer_rsrc_t fr, uint32_t fbase, const float* brow) {
#pragma unroll
    for (int kb = LO; kb < HI; ++kb) {
        const int u = kb % RING, v = kb % BRING;
        const float4 b = bq[v];
        acc = mfma1(ring[u].x, b.x, acc);
        acc = mfma1(ring[u].y, b.y, acc);
        acc = mfma1(ring[u].z, b.z, acc);
        acc = mfma1(ring[u].w, b.w, acc);
        __builtin_amdgcn_sched_barrier(0);
        if (kb + BRING < FULL) bq[v] = *reinterpret_cast<const float4*>(brow + 4 * (kb + BRING));
        if (kb + RING < FULL) ring[u] = ld4(fr, fbase + 1024u * (kb + RING));
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <int FULL>
__device__ __forceinline__ void ring_start(float4 (&ring)[RING], __amdgpu_buffer_rsrc_t fr, uint32_t fbase) {
#pragma unroll
    for (int u = 0; u < RING; ++u) ring[u] = ld4(fr, u < FULL ? fbase + 1024u * u : RACER_OOB);
}

template <int FULL>
__device__ __forceinline__ void inputs_start(float4 (&bq)[BRING], const float* brow) {
#pragma unroll
    for (int u = 0; u < BRING; ++u)
        if (u < FULL) bq[u] = *reinterpret_cast<const float4*>(brow + 4 * u);
}

__device__ __forceinline__ void chain_loop(f32x16& acc, __amdgpu_buffer_rsrc_t fr, uint32_t fbase, const float* brow,
                                           int full) {
    float4 ring[RING];
#pragma unroll
    for (int u = 0; u < RING; ++u) ring[u] = ld4(fr, u < full ? fbase + 1024u * u : RACER_OOB);
    float4 b = full > 0 ? *reinterpret_cast<const float4*>(brow) : float4{};
    for (int kb0 = 0; kb0 < full; kb0 += RING) {
#pragma unroll
        for (int u = 0; u < RING; ++u) {
            const int kb = kb0 + u;
            if (kb < full) {
                const float4 cur = b;
                if (kb + 1 < full) b = *reinterpret_cast<const float4*>(brow + 4 * (kb + 1));
                acc = mfma1(ring[u].x, cur.x, acc);
                acc = mfma1(ring[u].y, cur.y, acc);
                acc = mfma1(ring[u].z, cur.z, acc);
                acc = mfma1(ring[u].w, cur.w, acc);
            }
            ring[u] = ld4(fr, kb + RING < full ? fbase + 1024u * (kb + RING) : RACER_OOB);
        }
    }
}

// The last (partial) k quad: its first `rem` k only, so the chain has exactly
// K steps as the fmaf chain does.
__device__ __forceinline__ void chain_tail(f32x16& acc, __amdgpu_buffer_rsrc_t fr, uint32_t fbase, const float* brow,
                                           int full, int rem) {
    if (rem == 0) return;
    const float4 a = ld4(fr, fbase + 1024u * full);
    const float4 b = *reinterpret_cast<const float4*>(brow + 4 * full);
    acc = mfma1(a.x, b.x, acc);
    if (rem > 1) acc = mfma1(a.y, b.y, acc);
    if (rem > 2) acc = mfma1(a.z, b.z, acc);
}

// A whole product: unrolled when FULL > 0 (the whole-quad count of K), else a loop.
template <int FULL>
__device__ __forceinline__ void chain(f32x16& acc, __amdgpu_buffer_rsrc_t fr, uint32_t fbase, const float* brow,
                                      int K) {
    const int full = K / 4;
    if constexpr (FULL > 0) {
        float4 ring[RING], bq[BRING];
        ring_start<FULL>(ring, fr, fbase);
        inputs_start<FULL>(bq, brow);
        chain_seg<FULL, 0, FULL>(acc, ring, bq, fr, fbase, brow);
    } else {
        chain_loop(acc, fr, fbase, brow, full);
    }
    chain_tail(acc, fr, fbase, brow, full, K % 4);
}

// NR rounds of the input slice from round r0: round r holds k quads
// [QR r, QR (r + 1)) (QR = blockDim / 4); thread t takes quad QR r + t / 4 of
// columns 4 (t % 4) .. + 3: four 16-byte loads, stored transposed as four
// 16-byte LDS writes into [column][k]. Quads past KB0 are skipped; k past IN
// and columns past cols read 0.
template <int NR>
struct Staged {
    float4 f[NR][4];
    __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t xres, int64_t ldx_g, int IN, int KB0, int64_t c0,
                                         int64_t cols, int tid, int QR, int r0) {
        const int cq = (tid & 3) * 4;
        const int64_t col = c0 + cq;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int kq = QR * (r0 + r) + (tid >> 2);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k = 4 * kq + e;
                const bool in = kq < KB0 && k < IN && col < cols;
                f[r][e] = ld4(xres, in ? (uint32_t)(((int64_t)k * ldx_g + col) * 4) : RACER_OOB);
            }
        }
    }
    template <int R>
    __device__ __forceinline__ void store(float* xs, int ldx, int KB0, int tid, int QR, int r0) const {
        const int cq = (tid & 3) * 4, kq = QR * (r0 + R) + (tid >> 2);
        if (kq >= KB0) return;
        float* d = xs + cq * ldx + 4 * kq;
        *reinterpret_cast<float4*>(d) = make_float4(f[R][0].x, f[R][1].x, f[R][2].x, f[R][3].x);
        *reinterpret_cast<float4*>(d + ldx) = make_float4(f[R][0].y, f[R][1].y, f[R][2].y, f[R][3].y);
        *reinterpret_cast<float4*>(d + 2 * ldx) = make_float4(f[R][0].z, f[R][1].z, f[R][2].z, f[R][3].z);
        *reinterpret_cast<float4*>(d + 3 * ldx) = make_float4(f[R][0].w, f[R][1].w, f[R][2].w, f[R][3].w);
    }
};

// The first product with its inputs staged under it: segment R runs the quads
// up to the opening of round R + 1 (less the read-ahead), then round R + 1 is
// stored and every wavefront's store waited for.
template <int FULL, int R, int NR, int QR>
__device__ __forceinline__ void staged_rounds(f32x16& acc, float4 (&ring)[RING], float4 (&bq)[BRING],
                                              __amdgpu_buffer_rsrc_t fr, uint32_t fbase, const float* brow,
                                              const Staged<NR>& st, float* xs, int ldx, int KB0, int tid) {
    constexpr int lo = R == 0 ? 0 : QR * R - BRING;
    constexpr int hi = R + 1 < NR ? QR * (R + 1) - BRING : FULL;
    static_assert(lo <= hi && hi <= FULL, "staging rounds must open inside the product");
    chain_seg<FULL, lo, hi>(acc, ring, bq, fr, fbase, brow);
    if constexpr (R + 1 < NR) {
        st.template store<R + 1>(xs, ldx, KB0, tid, QR, 0);
        __syncthreads();
        staged_rounds<FULL, R + 1, NR, QR>(acc, ring, bq, fr, fbase, brow, st, xs, ldx, KB0, tid);
    }
}

// The 16 results of a lane's accumulators as tanh, then into the [column][row]
// LDS tile: every evaluation first, in one basic block, so the 16 independent
// chains interleave (a guarded store after each one put a branch between them
// and left one wavefront per SIMD waiting on each chain in turn).
__device__ __forceinline__ void tanh_tile(const f32x16& acc, float* hs, int ldh, int wave, int ib, int jc, int H) {
    float v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = racer_tanh_select(acc[r]);
    const int row0 = 64 * wave + ib;
    if (64 * wave + 64 <= H) {
#pragma unroll
        for (int r = 0; r < 16; ++r) hs[jc * ldh + row0 + 16 * (r >> 2) + (r & 3)] = v[r];
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = row0 + 16 * (r >> 2) + (r & 3);
            if (row < H) hs[jc * ldh + row] = v[r];
        }
    }
}

// The workgroup's h rows 0 .. H - 1 of its FT columns from the LDS tile to
// out [rows][cols] (the activations the backward reads): thread t writes rows
// t, t + blockDim, ..., each as FT consecutive floats (16-byte stores when
// aligned), off the products' path at the end of the kernel.
__device__ __forceinline__ void tile_out(const float* hs, int ldh, float* out, int H, int64_t cols, int64_t c0,
                                         bool vec) {
    const int ncol = cols - c0 < FT ? (int)(cols - c0) : FT;
    for (int row = (int)threadIdx.x; row < H; row += (int)blockDim.x) {
        float* dst = out + (int64_t)row * cols + c0;
        if (vec && ncol == FT) {
#pragma unroll
            for (int q = 0; q < FT / 4; ++q)
                *reinterpret_cast<float4*>(dst + 4 * q) = make_float4(
                    hs[(4 * q) * ldh + row], hs[(4 * q + 1) * ldh + row], hs[(4 * q + 2) * ldh + row],
                    hs[(4 * q + 3) * ldh + row]);
        } else {
            for (int c = 0; c < ncol; ++c) dst[c] = hs[c * ldh + row];
        }
    }
}

// FULL0 / FULL1: whole k quads of the two products when unrolled (0 = loop);
// NW: wavefronts (64-row groups) of the unrolled shapes, NR0 their staging rounds.
template <bool SAMPLE, bool VEC, int FULL0, int FULL1, int NW, int NR0>
__global__ void __launch_bounds__(64 * MAX_RG, 1) k_policy_f32(
    const float* __restrict__ F0, const float* __restrict__ F1, const float* __restrict__ W2,
    const float* __restrict__ x, int64_t ldx_g, int H, int IN, int64_t cols, float* __restrict__ h1out,
    float* __restrict__ h2out, float* __restrict__ o, const float* __restrict__ log_std, float* act, int64_t act_cols,
    float* logp, float* value, int deterministic, int64_t step, const int64_t* step_base, int seed) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const F32Layout Ly = f32_layout(H, IN);
    const int HA = Ly.HA;
    float* xs = reinterpret_cast<float*>(smem + Ly.xs);    // [FT][ldx]
    float* h1s = reinterpret_cast<float*>(smem + Ly.h1s);  // [FT][ldh]
    float* h2s = reinterpret_cast<float*>(smem + Ly.h2s);  // [FT][ldh]
    float* w2s = reinterpret_cast<float*>(smem + Ly.w2s);  // [3][w2p]
    float* outs = reinterpret_cast<float*>(smem + Ly.outs);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t c0 = (int64_t)xcd_block() * FT;
    const auto xres = racer_rsrc(x, ((uint64_t)(IN - 1) * ldx_g + cols) * 4);
    // W2 (3 HA <= 4 blockDim floats for every H <= 256) is loaded now and
    // stored to LDS after the first product, off the start's critical path
    float w2r[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int idx = tid + q * (int)blockDim.x;
        w2r[q] = idx < 3 * HA ? W2[idx] : 0.0f;
    }

    // D of v_mfma_f32_16x16x1_4b_f32: block b = r / 4 holds rows 16 b + 4 (lane / 16) + r % 4,
    // column lane % 16; lane l supplies A of block l / 16, row l % 16, and B of column l % 16.
    const int jc = lane & 15, ib = (lane >> 4) * 4;

    // ---- phases 0 and 1: the input slice, h1 = tanh(W0 x) ----
    f32x16 acc = {};
    {
        const auto fr = racer_rsrc(F0, (uint64_t)MAX_RG * Ly.KB0 * 1024);
        const uint32_t fbase = (uint32_t)(((wave * Ly.KB0) * 64 + lane) * 16);
        const float* brow = xs + jc * Ly.ldx;
        if constexpr (VEC && NW > 0) {
            // blockDim = 64 NW: the weight ring's loads first (loads retire in
            // issue order, so its first waits do not wait for the inputs'),
            // then every staging round's loads, round 0's stores, and the rest
            // of the rounds stored under the product
            constexpr int QR = 16 * NW;
            float4 ring[RING], bq[BRING];
            ring_start<FULL0>(ring, fr, fbase);
            Staged<NR0> st;
            st.load(xres, ldx_g, IN, Ly.KB0, c0, cols, tid, QR, 0);
            st.template store<0>(xs, Ly.ldx, Ly.KB0, tid, QR, 0);
            __syncthreads();
            inputs_start<FULL0>(bq, brow);
            staged_rounds<FULL0, 0, NR0, QR>(acc, ring, bq, fr, fbase, brow, st, xs, Ly.ldx, Ly.KB0, tid);
            chain_tail(acc, fr, fbase, brow, FULL0, IN % 4);
        } else {
            if constexpr (VEC) {
                // four rounds' loads in flight, then their stores
                const int QR = (int)blockDim.x / 4;
                for (int r0 = 0; QR * r0 < Ly.KB0; r0 += 4) {
                    Staged<4> st;
                    st.load(xres, ldx_g, IN, Ly.KB0, c0, cols, tid, QR, r0);
                    st.template store<0>(xs, Ly.ldx, Ly.KB0, tid, QR, r0);
                    st.template store<1>(xs, Ly.ldx, Ly.KB0, tid, QR, r0);
                    st.template store<2>(xs, Ly.ldx, Ly.KB0, tid, QR, r0);
                    st.template store<3>(xs, Ly.ldx, Ly.KB0, tid, QR, r0);
                }
            } else {
                for (int idx = tid; idx < Ly.KB0 * 4 * FT; idx += blockDim.x) {
                    const int c = idx % FT, k = idx / FT;
                    const int64_t col = c0 + c;
                    const uint32_t off =
                        (k < IN && col < cols) ? (uint32_t)(((int64_t)k * ldx_g + col) * 4) : RACER_OOB;
                    xs[c * Ly.ldx + k] = racer_ld_f32(xres, off);
                }
            }
            __syncthreads();
            chain<FULL0>(acc, fr, fbase, brow, IN);
        }
    }
    tanh_tile(acc, h1s, Ly.ldh, wave, ib, jc, H);
    for (int j = tid; j < FT; j += blockDim.x) h1s[j * Ly.ldh + H] = 1.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int idx = tid + q * (int)blockDim.x;
        if (idx < 3 * HA) w2s[(idx / HA) * Ly.w2p + idx % HA] = w2r[q];
    }
    __syncthreads();

    // ---- phase 2: h2 = tanh(W1 h1) ----
    {
        const auto fr = racer_rsrc(F1, (uint64_t)MAX_RG * Ly.KB1 * 1024);
        const uint32_t fbase = (uint32_t)(((wave * Ly.KB1) * 64 + lane) * 16);
        acc = f32x16{};
        chain<FULL1>(acc, fr, fbase, h1s + jc * Ly.ldh, HA);
        tanh_tile(acc, h2s, Ly.ldh, wave, ib, jc, H);
    }
    for (int j = tid; j < FT; j += blockDim.x) h2s[j * Ly.ldh + H] = 1.0f;
    __syncthreads();

    // ---- phase 3: o = W2 h2, fmaf over k in order, then the sample ----
    if (tid < 3 * FT) {
        const int j = tid / FT, c = tid % FT;
        const float* w = w2s + j * Ly.w2p;
        const float* h = h2s + c * Ly.ldh;
        float a = 0.0f;
        int k = 0;
#pragma unroll 4
        for (; k + 4 <= HA; k += 4) {
            const float4 wv = *reinterpret_cast<const float4*>(w + k);
            const float4 hv = *reinterpret_cast<const float4*>(h + k);
            a = fmaf(wv.x, hv.x, a);
            a = fmaf(wv.y, hv.y, a);
            a = fmaf(wv.z, hv.z, a);
            a = fmaf(wv.w, hv.w, a);
        }
        for (; k < HA; ++k) a = fmaf(w[k], h[k], a);
        outs[j * FT + c] = a;
        if (c0 + c < cols) o[j * cols + c0 + c] = a;
    }
    // h1 and h2 stay in LDS to the end (h2 overlays only the input slice)
    const bool ovec = cols % 4 == 0 && (uintptr_t)h1out % 16 == 0 && (uintptr_t)h2out % 16 == 0;
    tile_out(h1s, Ly.ldh, h1out, H, cols, c0, ovec);
    tile_out(h2s, Ly.ldh, h2out, H, cols, c0, ovec);
    if constexpr (SAMPLE) {
        __syncthreads();
        const int64_t col = c0 + tid;
        if (tid < FT && col < cols) {
            if (step_base) step += *step_base;
            float noise[4];
            racer_normal4(racer_random((uint64_t)(int64_t)seed + 1, (uint64_t)col, (uint64_t)step), noise);
            float probability = 0.0f;
            for (int a = 0; a < 2; ++a) {
                const float z = deterministic ? 0.0f : noise[a];
                const float ls = log_std[a];
                act[a * act_cols + col] = outs[a * FT + tid] + racer_exp(ls) * z;
                probability = probability - (((0.5f * z) * z + ls) + RACER_HALF_LOG_2PI);
            }
            logp[col] = probability;
            value[col] = outs[2 * FT + tid];
        }
    }
}

// Float32 weights in the MFMA operand order above: block (row group g, k
// quad q) holds lane l's row 64 g + l and k 4 q .. 4 q + 3, 64 lanes x 16
// bytes contiguous; zeros past the rows and columns of the source.
__global__ void k_pack_f32_fragments(const float* __restrict__ src, int64_t ld_src, int rows, int cols, int KB,
                                     float* __restrict__ dst) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)MAX_RG * KB * 256) return;
    const int e = (int)(idx & 3), l = (int)((idx >> 2) & 63);
    const int64_t blk = idx >> 8;
    const int q = (int)(blk % KB), g = (int)(blk / KB);
    const int row = 64 * g + l, k = 4 * q + e;
    dst[idx] = (row < rows && k < cols) ? src[(int64_t)row * ld_src + k] : 0.0f;
}

// Unrolled instantiations for the network shapes in use (1081 and 108 beams;
// H 256 and 64): FULL0 = IN / 4, FULL1 = (H + 1) / 4, NW = H / 64 wavefronts,
// NR0 = ceil(ceil(IN / 4) / (16 NW)) staging rounds. Anything else takes the loops.
template <bool SAMPLE, bool VEC>
const void* pick_f32(int IN, int H) {
    if (IN == 1086 && H == 256) return (const void*)k_policy_f32<SAMPLE, VEC, 271, 64, 4, 5>;
    if (IN == 113 && H == 64) return (const void*)k_policy_f32<SAMPLE, VEC, 28, 16, 1, 2>;
    if (IN == 113 && H == 256) return (const void*)k_policy_f32<SAMPLE, VEC, 28, 64, 4, 1>;
    return (const void*)k_policy_f32<SAMPLE, VEC, 0, 0, 0, 1>;
}

const void* pick_kernel(bool sample, bool vec, int IN, int H) {
    if (sample) return vec ? pick_f32<true, true>(IN, H) : pick_f32<true, false>(IN, H);
    return vec ? pick_f32<false, true>(IN, H) : pick_f32<false, false>(IN, H);
}

}  // namespace

size_t racer_f32_fragment_floats(int K) { return (size_t)MAX_RG * ((K + 3) / 4) * 256; }

int racer_pack_f32_fragments(racer_device* dev, const float* src, int64_t ld_src, int rows, int cols, float* dst) {
    const int KB = (cols + 3) / 4;
    const int64_t total = (int64_t)MAX_RG * KB * 256;
    {
        racer_timed timer(dev, "pack_f32_fragments");
        hipLaunchKernelGGL(k_pack_f32_fragments, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, dev->stream,
                           src, ld_src, rows, cols, KB, dst);
    }
    return racer_launch_check("pack_f32_fragments");
}

bool racer_policy_f32_fits(int H, int IN) {
    return H >= 1 && H <= 64 * MAX_RG && f32_layout(H, IN).total <= 160 * 1024;
}

int racer_policy_f32_setup() {
    // allow every variant the whole LDS of a CU (the unrolled shapes and the generic loop)
    const int shapes[4][2] = {{1086, 256}, {113, 64}, {113, 256}, {0, 0}};
    for (bool s : {false, true})
        for (bool v : {false, true})
            for (const auto& sh : shapes) {
                const void* k = pick_kernel(s, v, sh[0], sh[1]);
                RACER_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            }
    return 0;
}

int racer_policy_f32(racer_policy* p, const float* x, int64_t x_cols, racer_acts* a, int cols, bool sample,
                     float* act, float* logp, float* value, int deterministic, int64_t step, int seed) {
    racer_device* dev = p->dev;
    const int H = p->hidden, IN = (int)p->IN;
    RACER_REQUIRE(racer_policy_f32_fits(H, IN), "the fused float32 forward does not fit this network");
    RACER_REQUIRE(racer_fused_reach(IN, x_cols, cols),
                  "float32 policy input exceeds the 32-bit offsets of the buffer loads (4 GiB)");
    if (cols <= 0) return 0;
    const bool vec = x_cols % 4 == 0 && cols % 4 == 0 && (uintptr_t)x % 16 == 0;
    const F32Layout Ly = f32_layout(H, IN);
    const dim3 grid((unsigned)((cols + FT - 1) / FT)), block(64 * ((H + 63) / 64));
    const void* kernel = pick_kernel(sample, vec, IN, H);
    const float* W2 = p->theta->ptr + p->W2;
    const float* log_std = p->theta->ptr + p->LOGSTD;
    float* h1 = a->h1->ptr;
    float* h2 = a->h2->ptr;
    float* o = a->o->ptr;
    int64_t cols64 = cols, act_cols = cols;
    const int64_t* base = dev->step_base;
    void* args[] = {(void*)&p->frag0_f32, (void*)&p->frag1_f32, (void*)&W2,     (void*)&x,          (void*)&x_cols,
                    (void*)&H,            (void*)&IN,           (void*)&cols64, (void*)&h1,         (void*)&h2,
                    (void*)&o,            (void*)&log_std,      (void*)&act,    (void*)&act_cols,   (void*)&logp,
                    (void*)&value,        (void*)&deterministic, (void*)&step,  (void*)&base,       (void*)&seed};
    racer_timed timer(dev, sample ? "policy_f32_sample" : "policy_f32");
    RACER_TRY(hipLaunchKernel(kernel, grid, block, args, Ly.total, dev->stream));
    return racer_launch_check("policy_f32");
}
