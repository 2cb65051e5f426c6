// BF16 MFMA matrix products for the policy (gfx950 v_mfma_f32_32x32x16_bf16),
// float32 accumulation.
//
// One 256-thread workgroup computes a 64 x 64 tile of C; each of its four
// wavefronts owns a 32 x 32 block and issues two 32x32x16 MFMAs per 32-wide K
// stage. Both operands are staged in LDS as [row][k] with k contiguous, so each
// lane's MFMA fragment (8 consecutive k of one row or column) is one 16-byte
// LDS read. Activations arrive as float32 and are rounded to bf16
// (round-to-nearest-even) while staging; weights are pre-rounded once per
// update into a row-padded bf16 copy (rows padded to a multiple of 8 values).
#include <type_traits>

#include "racer_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;  // LDS row stride in bf16 (80 bytes)

// Round to nearest even with gfx950's v_cvt_pk_bf16_f32 (checked against the
// integer rounding (u + 0x7fff + lsb) >> 16 on all 2^32 patterns: equal for
// every finite value; a NaN comes out quiet).
__device__ __forceinline__ uint16_t to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// B operand modes: NN reads B[K][N] (k-major rows), NT reads B[N][K].
template <bool A_BF16, bool B_NT, int EPI>
__global__ void __launch_bounds__(256) k_gemm_bf16(const void* __restrict__ Ap, int64_t lda,
                                                   const float* __restrict__ B, int64_t ldb, float* __restrict__ C,
                                                   int64_t ldc, int64_t chunk_stride, int M, int N, int K, int kchunk) {
    __shared__ __attribute__((aligned(16))) uint16_t As[TM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][LDK];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    f32x16 acc;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0f;

    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        {  // A tile: 64 rows x 32 k, 8 consecutive k per thread
            const int row = tid >> 2, kc = (tid & 3) * 8;
            const int m = m0 + row;
            uint16_t v[8];
            if constexpr (A_BF16) {
                // rows are padded to a multiple of 8 with zeros: whole 16-byte chunks are in range
                const uint16_t* a = reinterpret_cast<const uint16_t*>(Ap);
                if (m < M && k0 + kc < kend) {
                    uint4 raw = *reinterpret_cast<const uint4*>(a + (int64_t)m * lda + k0 + kc);
                    *reinterpret_cast<uint4*>(v) = raw;
                } else {
                    *reinterpret_cast<uint4*>(v) = make_uint4(0, 0, 0, 0);
                }
            } else {
                const float* a = reinterpret_cast<const float*>(Ap);
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    int k = k0 + kc + e;
                    v[e] = (m < M && k < kend) ? to_bf16(a[(int64_t)m * lda + k]) : (uint16_t)0;
                }
            }
            *reinterpret_cast<uint4*>(&As[row][kc]) = *reinterpret_cast<uint4*>(v);
        }
        if constexpr (B_NT) {  // B[N][K]: 8 consecutive k of one column per thread
            const int col = tid >> 2, kc = (tid & 3) * 8;
            const int n = n0 + col;
            uint16_t v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                int k = k0 + kc + e;
                v[e] = (n < N && k < kend) ? to_bf16(B[(int64_t)n * ldb + k]) : (uint16_t)0;
            }
            *reinterpret_cast<uint4*>(&Bs[col][kc]) = *reinterpret_cast<uint4*>(v);
        } else {  // B[K][N]: 8 consecutive columns of one k per thread, coalesced along N
            const int kk = tid >> 3, nc = (tid & 7) * 8;
            const int k = k0 + kk;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                int n = n0 + nc + e;
                Bs[nc + e][kk] = (k < kend && n < N) ? to_bf16(B[(int64_t)k * ldb + n]) : (uint16_t)0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < TK; ks += 16) {
            const int r = lane & 31, kb = (lane >> 5) * 8;
            bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + r][ks + kb]);
            bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn * 32 + r][ks + kb]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D[i][j]: j = lane % 32, i = 8 (q / 4) + 4 (lane / 32) + q % 4
    float* dst = C + (int64_t)blockIdx.z * chunk_stride;
    const int j = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = m0 + wm * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        if (i < M && j < N) {
            float v = acc[q];
            if constexpr (EPI == RACER_EPI_TANH) v = racer_tanh(v);
            dst[(int64_t)i * ldc + j] = v;
        }
    }
}

// Forward products C[M, N] = epi(A[M, K] B[K, N]) with A a row-padded bf16
// weight copy and B float32 activations (one column per sample). 64 x 64
// tiles, K in stages of 64: while the four wavefronts run a stage's eight
// MFMAs from LDS, the next stage's operands are already loading into
// registers. Each thread gathers 16 consecutive k of one column (loads that
// are contiguous across the wavefront), rounds them to bf16 and writes them to
// LDS as two 16-byte vectors, so LDS holds B as [column][k].
//
// IO16 (the PPO update's through product): B is already bf16 and C is stored
// in bf16, rounded as its consumers round it on load; the same bits on half
// the bytes. PAIRS (IO16 with even N and ldb and a 4-byte aligned B): thread t
// loads columns 2 (t % 32) and 2 (t % 32) + 1 for 8 consecutive k as 4-byte
// words and splits them into the two columns' 16-byte LDS rows, half the load
// instructions of one column's 16 k as 2-byte loads; the same LDS contents.
template <int EPI, bool IO16 = false, bool PAIRS = false>
__global__ void __launch_bounds__(256) k_gemm_bf16_nn(const uint16_t* __restrict__ A, int64_t lda,
                                                      const void* __restrict__ B, int64_t ldb, void* __restrict__ C,
                                                      int64_t ldc, int M, int N, int K, const float* __restrict__ H,
                                                      int64_t ldh) {
    constexpr int SK = 64, LD = SK + 8;
    __shared__ __attribute__((aligned(16))) uint16_t As[TM][LD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][LD];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // one-dimensional grid, tile t = column tile x row tiles + row tile: the
    // row tiles that share a column block of B are neighbours on one XCD
    const int tiles_m = (M + TM - 1) / TM, t = xcd_block();
    const int m0 = (t % tiles_m) * TM, n0 = (t / tiles_m) * TN;
    const int arow = tid >> 2, akc = (tid & 3) * 16;  // A: 16 consecutive k of one row
    const int bcol = tid & 63, bkc = (tid >> 6) * 16;  // B: 16 consecutive k of one column
    constexpr int BSZ = IO16 ? 2 : 4;  // bytes of one B element
    static_assert(!PAIRS || IO16, "column pairs are 4-byte words of bf16");
    const int pcol = (tid & 31) * 2, pkc = (tid >> 5) * 8;  // PAIRS: two columns, 8 consecutive k
    uint4 ra[2];
    using BVal = std::conditional_t<IO16, uint16_t, float>;
    BVal rb[PAIRS ? 1 : 16];
    uint32_t rp[PAIRS ? 8 : 1];
    // buffer loads: out-of-range offsets (rows past M, k past K, columns past N) read zeros
    const auto ares = racer_rsrc(A, (uint64_t)M * lda * 2);
    const auto bres = racer_rsrc(B, ((uint64_t)(K - 1) * ldb + N) * BSZ);
    auto load = [&](int k0) {
        const int m = m0 + arow;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = k0 + akc + 8 * h;  // rows are zero-padded to lda, a multiple of 8
            ra[h] = racer_ld_b128(ares, (m < M && k < K) ? (uint32_t)(((int64_t)m * lda + k) * 2) : RACER_OOB);
        }
        if constexpr (PAIRS) {
            // N even: column n + 1 exists wherever column n does
            const int n = n0 + pcol;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + pkc + e;
                const uint32_t off = (n < N && k < K) ? (uint32_t)(((int64_t)k * ldb + n) * 2) : RACER_OOB;
                rp[e] = __builtin_bit_cast(uint32_t, racer_ld_f32(bres, off));
            }
        } else {
            const int n = n0 + bcol;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int k = k0 + bkc + e;
                const uint32_t off = (n < N && k < K) ? (uint32_t)(((int64_t)k * ldb + n) * BSZ) : RACER_OOB;
                if constexpr (IO16)
                    rb[e] = racer_ld_u16(bres, off);
                else
                    rb[e] = racer_ld_f32(bres, off);
            }
        }
    };
    f32x16 acc;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
    load(0);
    // the mask's h values, loaded with the first stage (the epilogue then waits for no load)
    float hm[16];
    const int jm = n0 + wn * 32 + (lane & 31);
    if constexpr (EPI == RACER_EPI_MASK) {
        const auto hres = racer_rsrc(H, ((uint64_t)(M - 1) * ldh + N) * 4);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int i = m0 + wm * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            hm[q] = racer_ld_f32(hres, (i < M && jm < N) ? (uint32_t)(((int64_t)i * ldh + jm) * 4) : RACER_OOB);
        }
    }
    for (int k0 = 0; k0 < K; k0 += SK) {
        *reinterpret_cast<uint4*>(&As[arow][akc]) = ra[0];
        *reinterpret_cast<uint4*>(&As[arow][akc + 8]) = ra[1];
        if constexpr (PAIRS) {
            // word e holds k = pkc + e of column pcol (low half) and pcol + 1 (high half)
            uint4 c0, c1;
            c0.x = (rp[0] & 0xFFFFu) | (rp[1] << 16);
            c0.y = (rp[2] & 0xFFFFu) | (rp[3] << 16);
            c0.z = (rp[4] & 0xFFFFu) | (rp[5] << 16);
            c0.w = (rp[6] & 0xFFFFu) | (rp[7] << 16);
            c1.x = (rp[0] >> 16) | (rp[1] & 0xFFFF0000u);
            c1.y = (rp[2] >> 16) | (rp[3] & 0xFFFF0000u);
            c1.z = (rp[4] >> 16) | (rp[5] & 0xFFFF0000u);
            c1.w = (rp[6] >> 16) | (rp[7] & 0xFFFF0000u);
            *reinterpret_cast<uint4*>(&Bs[pcol][pkc]) = c0;
            *reinterpret_cast<uint4*>(&Bs[pcol + 1][pkc]) = c1;
        } else {
            uint16_t v[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                if constexpr (IO16)
                    v[e] = rb[e];
                else
                    v[e] = to_bf16(rb[e]);
            }
            *reinterpret_cast<uint4*>(&Bs[bcol][bkc]) = *reinterpret_cast<uint4*>(v);
            *reinterpret_cast<uint4*>(&Bs[bcol][bkc + 8]) = *reinterpret_cast<uint4*>(v + 8);
        }
        __syncthreads();
        if (k0 + SK < K) load(k0 + SK);
#pragma unroll
        for (int ks = 0; ks < SK; ks += 16) {
            const int r = lane & 31, kb = (lane >> 5) * 8;
            bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + r][ks + kb]);
            bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn * 32 + r][ks + kb]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int j = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = m0 + wm * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        if (i < M && j < N) {
            float v = acc[q];
            if constexpr (EPI == RACER_EPI_TANH) v = racer_tanh(v);
            if constexpr (EPI == RACER_EPI_TANH_MLP) v = racer_tanh_mlp(v);
            if constexpr (EPI == RACER_EPI_MASK) {
                const float h = hm[q];
                v = v * (1.0f - h * h);
            }
            if constexpr (IO16)
                static_cast<uint16_t*>(C)[(int64_t)i * ldc + j] = to_bf16(v);
            else
                static_cast<float*>(C)[(int64_t)i * ldc + j] = v;
        }
    }
}

// Weight gradients C[M, N] = A[M, K] B[N, K]^T (racer/ppo.mojo:121-122) with K
// the long batch dimension and both operands float32 with k contiguous, split
// over K into chunks whose partial products are summed afterwards in chunk
// order. One 256-thread workgroup owns a 128 x 128 tile of C; wavefront w
// owns the 64 x 64 block (w / 2, w % 2) as 2 x 2 v_mfma_f32_32x32x16_bf16
// accumulators. K advances in stages of 64: each thread fetches eight float4
// of A and eight of B (16 consecutive threads cover one row's 256 bytes) with
// buffer loads, rounds them to bf16 (nearest even) into LDS as [row][k], and
// the next stage's loads are issued before this stage's 16 MFMAs per
// wavefront. Rows past M or N and k past the chunk read as zeros.
//
// The grid is one-dimensional: workgroup g runs on XCD g % 8 (dispatch is
// round-robin over the XCDs), and when the chunk count is a multiple of 8
// every tile of chunk z is placed on XCD z % 8, so the tiles that share a
// chunk's A and B rows stream them through one L2 (one HBM read of each k
// stage instead of one per tile); otherwise the tiles go in plain order.
//
// AB16 / BB16: A / B are already bf16 (rounded by the same conversion), read
// as 8-byte loads of 4 k: the products are the same bits on half the bytes.
constexpr int WG_T = 128, WG_K = 64, WG_LD = WG_K + 8;

template <bool AB16, bool BB16>
__global__ void __launch_bounds__(256) k_wgrad_bf16(const void* __restrict__ Ap, int64_t lda,
                                                    const void* __restrict__ Bp, int64_t ldb, float* __restrict__ P,
                                                    int64_t ldp, int64_t chunk_stride, int M, int N, int K,
                                                    int kchunk, int tiles_x, int tiles_y, int chunks) {
    __shared__ __attribute__((aligned(16))) uint16_t As[WG_T][WG_LD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[WG_T][WG_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int tiles = tiles_x * tiles_y, g = blockIdx.x;
    int z, tile;
    if ((chunks & 7) == 0) {
        const int slot = g >> 3;
        z = (slot / tiles) * 8 + (g & 7);
        tile = slot % tiles;
    } else {
        z = g / tiles;
        tile = g % tiles;
    }
    const int m0 = (tile / tiles_x) * WG_T, n0 = (tile % tiles_x) * WG_T;
    const int kbeg = z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    // descriptors from the chunk's first k: offsets stay below 4 GB
    constexpr int ASZ = AB16 ? 2 : 4, BSZ = BB16 ? 2 : 4;  // bytes of one A / B element
    const auto ares = racer_rsrc(static_cast<const char*>(Ap) + (int64_t)kbeg * ASZ,
                                 ((uint64_t)(M - 1) * lda + (kend - kbeg)) * ASZ);
    const auto bres = racer_rsrc(static_cast<const char*>(Bp) + (int64_t)kbeg * BSZ,
                                 ((uint64_t)(N - 1) * ldb + (kend - kbeg)) * BSZ);
    using AReg = std::conditional_t<AB16, uint2, uint4>;
    using BReg = std::conditional_t<BB16, uint2, uint4>;
    AReg ra[8];
    BReg rb[8];
    auto load = [&](int k0) {  // k0 relative to kbeg
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = tid + 256 * i, r = idx >> 4, kq = (idx & 15) * 4;
            const int k = k0 + kq;
            const bool kin = kbeg + k < kend;
            const uint32_t ao = (m0 + r < M && kin) ? (uint32_t)(((int64_t)(m0 + r) * lda + k) * ASZ) : RACER_OOB;
            const uint32_t bo = (n0 + r < N && kin) ? (uint32_t)(((int64_t)(n0 + r) * ldb + k) * BSZ) : RACER_OOB;
            if constexpr (AB16)
                ra[i] = __builtin_bit_cast(uint2, racer_ld_f32x2(ares, ao));
            else
                ra[i] = racer_ld_b128(ares, ao);
            if constexpr (BB16)
                rb[i] = __builtin_bit_cast(uint2, racer_ld_f32x2(bres, bo));
            else
                rb[i] = racer_ld_b128(bres, bo);
        }
    };
    auto stage = [&](int k0) {  // registers -> LDS as bf16, zeroing k at or past the chunk's end
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = tid + 256 * i, r = idx >> 4, kq = (idx & 15) * 4;
            const int lim = kend - (kbeg + k0 + kq);  // valid elements in this float4
            const uint32_t* av = reinterpret_cast<const uint32_t*>(&ra[i]);
            const uint32_t* bv = reinterpret_cast<const uint32_t*>(&rb[i]);
            uint16_t a4[4], b4[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if constexpr (AB16)
                    a4[e] = e < lim ? (uint16_t)(av[e >> 1] >> (16 * (e & 1))) : (uint16_t)0;
                else
                    a4[e] = e < lim ? to_bf16(__builtin_bit_cast(float, av[e])) : (uint16_t)0;
                if constexpr (BB16)
                    b4[e] = e < lim ? (uint16_t)(bv[e >> 1] >> (16 * (e & 1))) : (uint16_t)0;
                else
                    b4[e] = e < lim ? to_bf16(__builtin_bit_cast(float, bv[e])) : (uint16_t)0;
            }
            *reinterpret_cast<uint2*>(&As[r][kq]) = *reinterpret_cast<uint2*>(a4);
            *reinterpret_cast<uint2*>(&Bs[r][kq]) = *reinterpret_cast<uint2*>(b4);
        }
    };
    f32x16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[a][b][q] = 0.0f;
    const int span = kend - kbeg;
    load(0);
    for (int k0 = 0; k0 < span; k0 += WG_K) {
        stage(k0);
        __syncthreads();
        load(k0 + WG_K);  // past the chunk: out-of-range offsets, zeros, never staged
        const int r = lane & 31, kb = (lane >> 5) * 8;
#pragma unroll
        for (int ks = 0; ks < WG_K; ks += 16) {
            bf16x8 a0 = *reinterpret_cast<const bf16x8*>(&As[wm * 64 + r][ks + kb]);
            bf16x8 a1 = *reinterpret_cast<const bf16x8*>(&As[wm * 64 + 32 + r][ks + kb]);
            bf16x8 b0 = *reinterpret_cast<const bf16x8*>(&Bs[wn * 64 + r][ks + kb]);
            bf16x8 b1 = *reinterpret_cast<const bf16x8*>(&Bs[wn * 64 + 32 + r][ks + kb]);
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    // D[i][j]: j = lane % 32, i = 8 (q / 4) + 4 (lane / 32) + q % 4
    float* dst = P + (int64_t)z * chunk_stride;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int j = n0 + wn * 64 + 32 * b + (lane & 31);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int i = m0 + wm * 64 + 32 * a + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                if (i < M && j < N) dst[(int64_t)i * ldp + j] = acc[a][b][q];
            }
        }
}

__global__ void k_pack_bf16(const float* src, int64_t ld_src, uint16_t* dst, int64_t ld_dst, int rows, int cols) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)rows * ld_dst) return;
    int r = (int)(idx / ld_dst), c = (int)(idx % ld_dst);
    dst[idx] = c < cols ? to_bf16(src[(int64_t)r * ld_src + c]) : (uint16_t)0;
}

__global__ void k_sum_chunks_bf16(const float* __restrict__ partial, int64_t stride, int chunks, float* __restrict__ C,
                                  int64_t ldc, int M, int N) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)M * N) return;
    int m = (int)(idx / N), n = (int)(idx % N);
    float acc = 0.0f;
    for (int c = 0; c < chunks; ++c) acc += partial[c * stride + (int64_t)m * ldc + n];
    C[(int64_t)m * ldc + n] = acc;
}

}  // namespace

int racer_pack_bf16(racer_device* dev, const float* src, int64_t ld_src, uint16_t* dst, int64_t ld_dst, int rows,
                    int cols) {
    int64_t total = (int64_t)rows * ld_dst;
    if (total <= 0) return 0;
    {
        racer_timed timer(dev, "pack_bf16");
        hipLaunchKernelGGL(k_pack_bf16, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, dev->stream, src, ld_src,
                           dst, ld_dst, rows, cols);
    }
    return racer_launch_check("pack_bf16");
}

int racer_gemm_nn_bf16(racer_device* dev, const uint16_t* A, int64_t lda, const void* B, int64_t ldb, void* C,
                       int64_t ldc, int M, int N, int K, int epi, const float* H, int64_t ldh) {
    if (M <= 0 || N <= 0) return 0;
    RACER_REQUIRE(lda % 8 == 0 && (uint64_t)M * lda * 2 < 0xFFFFFF00ull &&
                      ((uint64_t)(K - 1) * ldb + N) * 4 < 0xFFFFFF00ull,
                  "bf16 product operands exceed the 32-bit offsets of the buffer loads");
    dim3 grid((unsigned)(((N + TN - 1) / TN) * ((M + TM - 1) / TM))), block(256);
    racer_timed timer(dev, epi == RACER_EPI_TANH || epi == RACER_EPI_TANH_MLP
                               ? "gemm_bf16_tanh"
                               : (epi == RACER_EPI_MASK ? "gemm_bf16_mask" : "gemm_bf16"));
    if (epi == RACER_EPI_TANH)
        hipLaunchKernelGGL(k_gemm_bf16_nn<RACER_EPI_TANH>, grid, block, 0, dev->stream, A, lda, B, ldb, C, ldc, M, N,
                           K, H, ldh);
    else if (epi == RACER_EPI_TANH_MLP)
        hipLaunchKernelGGL(k_gemm_bf16_nn<RACER_EPI_TANH_MLP>, grid, block, 0, dev->stream, A, lda, B, ldb, C, ldc, M,
                           N, K, H, ldh);
    else if (epi == RACER_EPI_MASK)
        hipLaunchKernelGGL(k_gemm_bf16_nn<RACER_EPI_MASK>, grid, block, 0, dev->stream, A, lda, B, ldb, C, ldc, M, N,
                           K, H, ldh);
    else
        hipLaunchKernelGGL(k_gemm_bf16_nn<RACER_EPI_NONE>, grid, block, 0, dev->stream, A, lda, B, ldb, C, ldc, M, N,
                           K, H, ldh);
    return racer_launch_check("gemm_nn_bf16");
}

int racer_gemm_nn_bf16_mask16(racer_device* dev, const uint16_t* A, int64_t lda, const uint16_t* B, int64_t ldb,
                              uint16_t* C, int64_t ldc, int M, int N, int K, const float* H, int64_t ldh) {
    if (M <= 0 || N <= 0) return 0;
    RACER_REQUIRE(lda % 8 == 0 && (uint64_t)M * lda * 2 < 0xFFFFFF00ull &&
                      ((uint64_t)(K - 1) * ldb + N) * 2 < 0xFFFFFF00ull && ((uint64_t)(M - 1) * ldh + N) * 4 < 0xFFFFFF00ull,
                  "bf16 product operands exceed the 32-bit offsets of the buffer loads");
    dim3 grid((unsigned)(((N + TN - 1) / TN) * ((M + TM - 1) / TM))), block(256);
    racer_timed timer(dev, "gemm_bf16_mask");
    if (N % 2 == 0 && ldb % 2 == 0 && (uintptr_t)B % 4 == 0)
        hipLaunchKernelGGL((k_gemm_bf16_nn<RACER_EPI_MASK, true, true>), grid, block, 0, dev->stream, A, lda, B, ldb,
                           C, ldc, M, N, K, H, ldh);
    else
        hipLaunchKernelGGL((k_gemm_bf16_nn<RACER_EPI_MASK, true>), grid, block, 0, dev->stream, A, lda, B, ldb, C,
                           ldc, M, N, K, H, ldh);
    return racer_launch_check("gemm_nn_bf16");
}

// The batch splits into chunks of 4096 (both kernels); each chunk's partial
// product fills one [M, ldc] slab of the workspace. At the PPO update's
// 262144-column minibatches that is 64 chunks, 1152 workgroups for W0's
// gradient: the k loop's loads wait on HBM latency, and twice the workgroups
// of 8192-column chunks hide more of it (15.7 against 19.7 ms of weight
// gradients per update); 2048 adds more partial sums than it saves.
constexpr int NT_BF16_CHUNK = 4096;

size_t racer_gemm_nt_bf16_partials(int M, int64_t ldc, int K) {
    int chunks = K <= 0 ? 1 : (K + NT_BF16_CHUNK - 1) / NT_BF16_CHUNK;
    return (size_t)M * (size_t)ldc * (size_t)chunks;
}

// The MFMA weight-gradient kernel, then the chunks' partial products summed
// in chunk order into C.
template <bool AB16, bool BB16>
static int wgrad_launch(racer_device* dev, const void* A, int64_t lda, const void* B, int64_t ldb, float* C,
                        int64_t ldc, int M, int N, int K, float* partial) {
    const int chunk = NT_BF16_CHUNK;
    int chunks = (K + chunk - 1) / chunk;
    int64_t stride = (int64_t)M * ldc;
    const int tx = (N + WG_T - 1) / WG_T, ty = (M + WG_T - 1) / WG_T;
    dim3 grid((unsigned)(tx * ty * chunks)), block(256);
    {
        racer_timed timer(dev, "wgrad_bf16");
        hipLaunchKernelGGL((k_wgrad_bf16<AB16, BB16>), grid, block, 0, dev->stream, A, lda, B, ldb, partial, ldc,
                           stride, M, N, K, chunk, tx, ty, chunks);
    }
    if (racer_launch_check("wgrad_bf16")) return 1;
    int64_t total = (int64_t)M * N;
    {
        racer_timed timer(dev, "sum_chunks");
        hipLaunchKernelGGL(k_sum_chunks_bf16, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, dev->stream,
                           partial, stride, chunks, C, ldc, M, N);
    }
    return racer_launch_check("sum_chunks");
}

int racer_gemm_nt_bf16_io(racer_device* dev, const void* A, bool a16, int64_t lda, const void* B, bool b16,
                          int64_t ldb, float* C, int64_t ldc, int M, int N, int K, float* ws, size_t ws_len) {
    if (M <= 0 || N <= 0) return 0;
    const uintptr_t aal = a16 ? 8 : 16, bal = b16 ? 8 : 16;
    RACER_REQUIRE(K > 0 && lda % 4 == 0 && ldb % 4 == 0 && K % 4 == 0 && (uintptr_t)A % aal == 0 &&
                      (uintptr_t)B % bal == 0,
                  "bf16 weight-gradient operands must be aligned to 4 elements");
    RACER_REQUIRE(((uint64_t)(M - 1) * lda + K) * (a16 ? 2 : 4) < 0xFFFFFF00ull &&
                      ((uint64_t)(N - 1) * ldb + K) * (b16 ? 2 : 4) < 0xFFFFFF00ull,
                  "bf16 product operands exceed the 32-bit offsets of the buffer loads");
    float* partial = nullptr;
    if (racer_partials(dev, racer_gemm_nt_bf16_partials(M, ldc, K), ws, ws_len, &partial)) return 1;
    if (a16 && b16) return wgrad_launch<true, true>(dev, A, lda, B, ldb, C, ldc, M, N, K, partial);
    if (a16) return wgrad_launch<true, false>(dev, A, lda, B, ldb, C, ldc, M, N, K, partial);
    if (b16) return wgrad_launch<false, true>(dev, A, lda, B, ldb, C, ldc, M, N, K, partial);
    return wgrad_launch<false, false>(dev, A, lda, B, ldb, C, ldc, M, N, K, partial);
}

int racer_gemm_nt_bf16(racer_device* dev, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                       int64_t ldc, int M, int N, int K, float* ws, size_t ws_len) {
    if (M <= 0 || N <= 0) return 0;
    float* partial = nullptr;
    if (racer_partials(dev, racer_gemm_nt_bf16_partials(M, ldc, K), ws, ws_len, &partial)) return 1;
    // the MFMA tile kernel: 16-byte aligned rows and operands within the reach of 32-bit offsets
    const bool aligned = lda % 4 == 0 && ldb % 4 == 0 && K % 4 == 0 && (uintptr_t)A % 16 == 0 && (uintptr_t)B % 16 == 0;
    const bool reach = ((uint64_t)(M - 1) * lda + K) * 4 < 0xFFFFFF00ull && ((uint64_t)(N - 1) * ldb + K) * 4 < 0xFFFFFF00ull;
    if (aligned && reach && K > 0) return wgrad_launch<false, false>(dev, A, lda, B, ldb, C, ldc, M, N, K, partial);
    const int chunk = NT_BF16_CHUNK;
    int chunks = K <= 0 ? 1 : (K + chunk - 1) / chunk;
    int64_t stride = (int64_t)M * ldc;
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, chunks), block(256);
    {
        racer_timed timer(dev, "gemm_nt_bf16");
        hipLaunchKernelGGL((k_gemm_bf16<false, true, RACER_EPI_NONE>), grid, block, 0, dev->stream, A, lda, B, ldb,
                           partial, ldc, stride, M, N, K, chunk);
    }
    if (racer_launch_check("gemm_nt_bf16")) return 1;
    int64_t total = (int64_t)M * N;
    {
        racer_timed timer(dev, "sum_chunks");
        hipLaunchKernelGGL(k_sum_chunks_bf16, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, dev->stream,
                           partial, stride, chunks, C, ldc, M, N);
    }
    return racer_launch_check("sum_chunks");
}
