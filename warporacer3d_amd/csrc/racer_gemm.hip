// Matrix products of the policy (racer/ppo.mojo:108-138; the reference calls
// MAX matmul through racer/device.mojo:82-98).
//
// Float32 path: every output accumulates fmaf(a, b, acc) over k = 0, 1, ...
// in order, one accumulator per output, so forward results are bit-identical
// to the CPU restatement. Tiles are 64 x 64 outputs per 256-thread workgroup,
// 4 x 4 per thread, operands staged through LDS 16 k at a time.
//
// Weight gradients (K = the minibatch, up to 262144) split K into chunks of
// RACER_SPLITK columns; chunk partials are summed in chunk order.
#include "racer_internal.h"

namespace {

constexpr int BM = 64, BN = 64, BK = 16;

template <int EPI>
__device__ __forceinline__ float epilogue(float acc, const float* H, int64_t ldh, int m, int n) {
    if constexpr (EPI == RACER_EPI_TANH) return racer_tanh(acc);
    if constexpr (EPI == RACER_EPI_MASK) {
        float h = H[(int64_t)m * ldh + n];
        return acc * (1.0f - h * h);
    }
    return acc;
}

// C[M, N] = epi(A[M, K] B[K, N])
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_nn_f32(const float* __restrict__ A, int64_t lda,
                                                     const float* __restrict__ B, int64_t ldb, float* __restrict__ C,
                                                     int64_t ldc, int M, int N, int K, const float* __restrict__ H,
                                                     int64_t ldh) {
    __shared__ float As[BK][BM + 1];
    __shared__ float Bs[BK][BN];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += BK) {
        {  // A tile: 64 rows x 16 k, 16 consecutive k per row
            int kk = tid & 15, r = tid >> 4;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                int row = r + 16 * p;
                int m = m0 + row, k = k0 + kk;
                As[kk][row] = (m < M && k < K) ? A[(int64_t)m * lda + k] : 0.0f;
            }
        }
        {  // B tile: 16 k x 64 columns, coalesced along the columns
            int c = tid & 63, r = tid >> 6;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                int kk = r + 4 * p;
                int k = k0 + kk, n = n0 + c;
                Bs[kk][c] = (k < K && n < N) ? B[(int64_t)k * ldb + n] : 0.0f;
            }
        }
        __syncthreads();
        const int kv = min(BK, K - k0);
        for (int kk = 0; kk < kv; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int m = m0 + ty + 16 * i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int n = n0 + tx + 16 * j;
            if (n < N) C[(int64_t)m * ldc + n] = epilogue<EPI>(acc[i][j], H, ldh, m, n);
        }
    }
}

// C[M, N] = (A[M, 3] B[3, N]) (1 - H^2) for the output layer's K = 3 (the
// backward's d2, racer/ppo.mojo:125-130): one thread per output, row
// blockIdx.y, each the fmaf chain over k = 0, 1, 2 from 0 (the tiled kernel's
// order and start, so the same bits), streamed at HBM rate instead of through
// 64 x 64 tiles of mostly k padding. OUT16: C stored in bf16 (nearest even).
//
// VEC: 8 consecutive columns per thread with 16-byte loads and stores (every
// leading dimension and N a multiple of 8, 16-byte aligned operands), so the
// launch is an eighth of the workgroups of one thread per output.
template <bool OUT16>
__device__ __forceinline__ float mask_k3_one(float a0, float a1, float a2, float b0, float b1, float b2, float h) {
    float acc = fmaf(a0, b0, 0.0f);
    acc = fmaf(a1, b1, acc);
    acc = fmaf(a2, b2, acc);
    return acc * (1.0f - h * h);
}

template <bool OUT16, bool VEC>
__global__ void __launch_bounds__(256) k_mask_k3(const float* __restrict__ A, const float* __restrict__ B,
                                                 int64_t ldb, void* __restrict__ C, int64_t ldc, int N,
                                                 const float* __restrict__ H, int64_t ldh) {
    const int i = blockIdx.y;
    const float a0 = A[3 * i], a1 = A[3 * i + 1], a2 = A[3 * i + 2];
    if constexpr (VEC) {
        const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
        if (j >= N) return;
        float v[8];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int64_t jj = j + 4 * half;
            const float4 b0 = *reinterpret_cast<const float4*>(B + jj);
            const float4 b1 = *reinterpret_cast<const float4*>(B + ldb + jj);
            const float4 b2 = *reinterpret_cast<const float4*>(B + 2 * ldb + jj);
            const float4 h = *reinterpret_cast<const float4*>(H + (int64_t)i * ldh + jj);
            v[4 * half + 0] = mask_k3_one<OUT16>(a0, a1, a2, b0.x, b1.x, b2.x, h.x);
            v[4 * half + 1] = mask_k3_one<OUT16>(a0, a1, a2, b0.y, b1.y, b2.y, h.y);
            v[4 * half + 2] = mask_k3_one<OUT16>(a0, a1, a2, b0.z, b1.z, b2.z, h.z);
            v[4 * half + 3] = mask_k3_one<OUT16>(a0, a1, a2, b0.w, b1.w, b2.w, h.w);
        }
        if constexpr (OUT16) {
            uint16_t o[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) o[e] = __builtin_bit_cast(uint16_t, (__bf16)v[e]);
            *reinterpret_cast<uint4*>(static_cast<uint16_t*>(C) + (int64_t)i * ldc + j) =
                *reinterpret_cast<const uint4*>(o);
        } else {
            float* c = static_cast<float*>(C) + (int64_t)i * ldc + j;
            *reinterpret_cast<float4*>(c) = make_float4(v[0], v[1], v[2], v[3]);
            *reinterpret_cast<float4*>(c + 4) = make_float4(v[4], v[5], v[6], v[7]);
        }
    } else {
        const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (j >= N) return;
        const float v = mask_k3_one<OUT16>(a0, a1, a2, B[j], B[ldb + j], B[2 * ldb + j], H[(int64_t)i * ldh + j]);
        if constexpr (OUT16)
            static_cast<uint16_t*>(C)[(int64_t)i * ldc + j] = __builtin_bit_cast(uint16_t, (__bf16)v);
        else
            static_cast<float*>(C)[(int64_t)i * ldc + j] = v;
    }
}

// Partial C[chunk][M, N] = A[M, kchunk] B[N, kchunk]^T over one K chunk.
__global__ void __launch_bounds__(256) k_gemm_nt_f32(const float* __restrict__ A, int64_t lda,
                                                     const float* __restrict__ B, int64_t ldb,
                                                     float* __restrict__ out, int64_t out_stride, int64_t ldc, int M,
                                                     int N, int K, int chunk) {
    __shared__ float As[BK][BM + 1];
    __shared__ float Bs[BK][BN + 1];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kbeg = blockIdx.z * chunk;
    const int kend = min(K, kbeg + chunk);
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        int kk = tid & 15, r = tid >> 4;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            int row = r + 16 * p;
            int k = k0 + kk;
            int m = m0 + row, n = n0 + row;
            As[kk][row] = (m < M && k < kend) ? A[(int64_t)m * lda + k] : 0.0f;
            Bs[kk][row] = (n < N && k < kend) ? B[(int64_t)n * ldb + k] : 0.0f;
        }
        __syncthreads();
        const int kv = min(BK, kend - k0);
        for (int q = 0; q < kv; ++q) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[q][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[q][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    float* dst = out + (int64_t)blockIdx.z * out_stride;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int m = m0 + ty + 16 * i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int n = n0 + tx + 16 * j;
            if (n < N) dst[(int64_t)m * ldc + n] = acc[i][j];
        }
    }
}

// The same chunked partial products on gfx950's K = 1 float32 MFMA
// (v_mfma_f32_16x16x1_4b_f32 forms each output as one correctly rounded fused
// multiply-add per k, so its chain over a chunk's k in order is the fmaf chain
// of k_gemm_nt_f32, bit for bit). A workgroup owns a 64 x 64 tile of one
// chunk's partial, like k_gemm_nt_f32: stages of 32 k of both operands' 64
// rows go to LDS as [row][k] through 16-byte loads (eight lanes per row's 128
// bytes), the next stage's loads in flight under this stage's MFMAs.
// Wavefront w owns rows 16 w .. + 15 of the tile and all 64 columns as the
// instruction's four 16-column blocks: lane l reads four k of row 16 w + l %
// 16 and of column l from LDS at once. k past the chunk reads zeros: fma(0, 0,
// acc) leaves acc as it is, since a chain that starts at +0 never holds -0.
typedef float nt_f32x16 __attribute__((ext_vector_type(16)));
constexpr int NT_SK = 32, NT_LD = NT_SK + 4;
__global__ void __launch_bounds__(256) k_gemm_nt_f32_mfma(const float* __restrict__ A, int64_t lda,
                                                          const float* __restrict__ B, int64_t ldb,
                                                          float* __restrict__ out, int64_t out_stride, int64_t ldc,
                                                          int M, int N, int K, int chunk) {
    __shared__ __attribute__((aligned(16))) float As[64][NT_LD];
    __shared__ __attribute__((aligned(16))) float Bs[64][NT_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int kbeg = blockIdx.z * chunk, kend = min(K, kbeg + chunk);
    const auto ares = racer_rsrc(A, ((uint64_t)(M - 1) * lda + K) * 4);
    const auto bres = racer_rsrc(B, ((uint64_t)(N - 1) * ldb + K) * 4);
    // thread t stages rows t / 8 and t / 8 + 32, k quad t % 8, of both operands
    const int srow = tid >> 3, skq = (tid & 7) * 4;
    float4 ra[2], rb[2];
    auto load = [&](int k0) {
        const int k = k0 + skq;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int m = m0 + srow + 32 * h, n = n0 + srow + 32 * h;
            ra[h] = __builtin_bit_cast(float4, racer_ld_b128(ares, (m < M && k < kend)
                                                                       ? (uint32_t)(((int64_t)m * lda + k) * 4)
                                                                       : RACER_OOB));
            rb[h] = __builtin_bit_cast(float4, racer_ld_b128(bres, (n < N && k < kend)
                                                                       ? (uint32_t)(((int64_t)n * ldb + k) * 4)
                                                                       : RACER_OOB));
        }
    };
    nt_f32x16 acc = {};
    const int ar = 16 * wave + (lane & 15);
    load(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += NT_SK) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            *reinterpret_cast<float4*>(&As[srow + 32 * h][skq]) = ra[h];
            *reinterpret_cast<float4*>(&Bs[srow + 32 * h][skq]) = rb[h];
        }
        __syncthreads();
        if (k0 + NT_SK < kend) load(k0 + NT_SK);
#pragma unroll
        for (int q = 0; q < NT_SK / 4; ++q) {
            const float4 a = *reinterpret_cast<const float4*>(&As[ar][4 * q]);
            const float4 b = *reinterpret_cast<const float4*>(&Bs[lane][4 * q]);
            acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.x, b.x, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.y, b.y, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.z, b.z, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.w, b.w, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D of block b = r / 4 (columns 16 b .. + 15): row 4 (lane / 16) + r % 4 of
    // the wavefront's 16, column lane % 16 of the block's
    float* dst = out + (int64_t)blockIdx.z * out_stride;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = m0 + 16 * wave + 4 * (lane >> 4) + (r & 3), j = n0 + 16 * (r >> 2) + (lane & 15);
        if (i < M && j < N) dst[(int64_t)i * ldc + j] = acc[r];
    }
}

// The backward's masked through products C = (A B) (1 - H^2) on the K = 1
// float32 MFMA: every output is k_gemm_nn_f32's fmaf chain over k from 0 in
// order, then the same epilogue, bit for bit. A 64 x 64 tile per workgroup;
// stages of 32 k of A ([row][k], k contiguous) and of B ([k][column], columns
// contiguous) go to LDS through 16-byte loads, the next stage's in flight
// under this stage's MFMAs; wavefront w owns rows 16 w .. + 15 and all 64
// columns as the instruction's four column blocks, lane l reading four k of
// row 16 w + l % 16 at once and column l at each k. k past K reads zeros
// (exact, as in k_gemm_nt_f32_mfma).
__global__ void __launch_bounds__(256) k_gemm_nn_f32_mask_mfma(const float* __restrict__ A, int64_t lda,
                                                               const float* __restrict__ B, int64_t ldb,
                                                               float* __restrict__ C, int64_t ldc, int M, int N,
                                                               int K, const float* __restrict__ H, int64_t ldh) {
    __shared__ __attribute__((aligned(16))) float As[64][NT_LD];
    __shared__ __attribute__((aligned(16))) float Bs[NT_SK][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const auto ares = racer_rsrc(A, ((uint64_t)(M - 1) * lda + K) * 4);
    const auto bres = racer_rsrc(B, ((uint64_t)(K - 1) * ldb + N) * 4);
    // A: rows t / 8 and t / 8 + 32, k quad t % 8; B: k rows t / 16 and t / 16
    // + 16, column quad t % 16
    const int arow = tid >> 3, akq = (tid & 7) * 4, bk = tid >> 4, bcq = (tid & 15) * 4;
    float4 ra[2], rb[2];
    auto load = [&](int k0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int m = m0 + arow + 32 * h, k = k0 + akq;
            ra[h] = __builtin_bit_cast(float4, racer_ld_b128(ares, (m < M && k < K)
                                                                       ? (uint32_t)(((int64_t)m * lda + k) * 4)
                                                                       : RACER_OOB));
            const int kb = k0 + bk + 16 * h, n = n0 + bcq;
            rb[h] = __builtin_bit_cast(float4, racer_ld_b128(bres, (kb < K && n < N)
                                                                       ? (uint32_t)(((int64_t)kb * ldb + n) * 4)
                                                                       : RACER_OOB));
        }
    };
    nt_f32x16 acc = {};
    const int ar = 16 * wave + (lane & 15);
    load(0);
    for (int k0 = 0; k0 < K; k0 += NT_SK) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            *reinterpret_cast<float4*>(&As[arow + 32 * h][akq]) = ra[h];
            *reinterpret_cast<float4*>(&Bs[bk + 16 * h][bcq]) = rb[h];
        }
        __syncthreads();
        if (k0 + NT_SK < K) load(k0 + NT_SK);
#pragma unroll
        for (int q = 0; q < NT_SK / 4; ++q) {
            const float4 a = *reinterpret_cast<const float4*>(&As[ar][4 * q]);
            acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.x, Bs[4 * q][lane], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.y, Bs[4 * q + 1][lane], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.z, Bs[4 * q + 2][lane], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.w, Bs[4 * q + 3][lane], acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = m0 + 16 * wave + 4 * (lane >> 4) + (r & 3), j = n0 + 16 * (r >> 2) + (lane & 15);
        if (i < M && j < N) C[(int64_t)i * ldc + j] = epilogue<RACER_EPI_MASK>(acc[r], H, ldh, i, j);
    }
}

// C[m, n] = sum over chunks c in order of partial[c][m, n].
// The same chunked product for at most four rows of A (the output layer's
// gradient, M = 3): thread (j, n) of a 256-thread workgroup accumulates
// A[j, k] B[n, k] with fmaf over k in order, exactly as k_gemm_nt_f32 does,
// for 64 rows n. B arrives 64 rows x 64 k per stage through LDS with float4
// loads (16 threads per 256-byte row); the next stage is loaded while this one
// is summed.
constexpr int SK_N = 64, SK_K = 64, SK_STAGES = 4;

__global__ void __launch_bounds__(256) k_gemm_nt_f32_skinny(const float* __restrict__ A, int64_t lda,
                                                            const float* __restrict__ B, int64_t ldb,
                                                            float* __restrict__ out, int64_t out_stride, int64_t ldc,
                                                            int M, int N, int K, int chunk) {
    __shared__ float Bs[SK_K][SK_N + 1];  // [k][column]: the fmaf loop's reads are conflict-free
    __shared__ float As[4][SK_K];
    const int tid = threadIdx.x;
    const int n = tid & 63, j = tid >> 6;
    const int n0 = blockIdx.x * SK_N;
    const int kbeg = blockIdx.z * chunk;
    const int kend = min(K, kbeg + chunk);
    const auto bres = racer_rsrc(B, ((uint64_t)(N - 1) * ldb + K) * 4);
    const auto ares = racer_rsrc(A, ((uint64_t)(M - 1) * lda + K) * 4);
    // SK_STAGES stages of SK_K in flight in registers: the chunk streams once
    // from HBM, and one stage at a time leaves the wait for each stage's
    // loads exposed
    // B: 16 consecutive threads cover one column's 64 k as float4 loads (the
    // host sends only 16-byte aligned rows and chunk bounds, so a float4 is
    // wholly inside or wholly past the chunk)
    float4 rb[SK_STAGES][4];
    float ra[SK_STAGES];
    auto load = [&](int k0, int u) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int idx = tid + 256 * i, r = idx >> 4, kq = (idx & 15) * 4;
            const int k = k0 + kq;
            const uint32_t off = (n0 + r < N && k < kend) ? (uint32_t)(((int64_t)(n0 + r) * ldb + k) * 4) : RACER_OOB;
            rb[u][i] = __builtin_bit_cast(float4, racer_ld_b128(bres, off));
        }
        const int r = tid >> 6, k = k0 + (tid & 63);
        const uint32_t off = (r < M && k < kend) ? (uint32_t)(((int64_t)r * lda + k) * 4) : RACER_OOB;
        ra[u] = racer_ld_f32(ares, off);
    };
    float acc = 0.0f;
#pragma unroll
    for (int u = 0; u < SK_STAGES; ++u) load(kbeg + u * SK_K, u);
    for (int k0 = kbeg; k0 < kend; k0 += SK_STAGES * SK_K) {
#pragma unroll
        for (int u = 0; u < SK_STAGES; ++u) {
            const int ks = k0 + u * SK_K;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int idx = tid + 256 * i, r = idx >> 4, kq = (idx & 15) * 4;
                Bs[kq + 0][r] = rb[u][i].x;
                Bs[kq + 1][r] = rb[u][i].y;
                Bs[kq + 2][r] = rb[u][i].z;
                Bs[kq + 3][r] = rb[u][i].w;
            }
            As[tid >> 6][tid & 63] = ra[u];
            __syncthreads();
            // the slot is free again: this stage's successor SK_STAGES stages on
            // (past the chunk the offsets are out of range and read zeros)
            load(ks + SK_STAGES * SK_K, u);
            const int kv = min(SK_K, kend - ks);
            if (j < M) {
                if (kv == SK_K) {
                    // unrolled, so the LDS reads go out ahead of the fmaf chain
                    // (k in order, as the loop below)
#pragma unroll
                    for (int q = 0; q < SK_K; ++q) acc = fmaf(As[j][q], Bs[q][n], acc);
                } else {
                    for (int q = 0; q < kv; ++q) acc = fmaf(As[j][q], Bs[q][n], acc);
                }
            }
            __syncthreads();
        }
    }
    if (j < M && n0 + n < N) out[(int64_t)blockIdx.z * out_stride + (int64_t)j * ldc + n0 + n] = acc;
}

__global__ void k_sum_chunks(const float* __restrict__ partial, int64_t stride, int chunks, float* __restrict__ C,
                             int64_t ldc, int M, int N) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)M * N) return;
    int m = (int)(idx / N), n = (int)(idx % N);
    float acc = 0.0f;
    for (int c = 0; c < chunks; ++c) acc += partial[c * stride + (int64_t)m * ldc + n];
    C[(int64_t)m * ldc + n] = acc;
}

}  // namespace

int racer_gemm_nn_f32(racer_device* dev, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                      int64_t ldc, int M, int N, int K, int epi, const float* H, int64_t ldh) {
    if (M <= 0 || N <= 0) return 0;
    dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM), block(256);
    racer_timed timer(dev, epi == RACER_EPI_TANH ? "gemm_f32_tanh" : (epi == RACER_EPI_MASK ? "gemm_f32_mask" : "gemm_f32"));
    if (epi == RACER_EPI_TANH)
        hipLaunchKernelGGL(k_gemm_nn_f32<RACER_EPI_TANH>, grid, block, 0, dev->stream, A, lda, B, ldb, C, ldc, M, N,
                           K, H, ldh);
    else if (epi == RACER_EPI_MASK && K % 4 == 0 && N % 4 == 0 && lda % 4 == 0 && ldb % 4 == 0 &&
             (uintptr_t)A % 16 == 0 && (uintptr_t)B % 16 == 0 && ((uint64_t)(M - 1) * lda + K) * 4 < 0xFFFFFF00ull &&
             ((uint64_t)(K - 1) * ldb + N) * 4 < 0xFFFFFF00ull)
        hipLaunchKernelGGL(k_gemm_nn_f32_mask_mfma, dim3((N + 63) / 64, (M + 63) / 64), block, 0, dev->stream, A, lda,
                           B, ldb, C, ldc, M, N, K, H, ldh);
    else if (epi == RACER_EPI_MASK)
        hipLaunchKernelGGL(k_gemm_nn_f32<RACER_EPI_MASK>, grid, block, 0, dev->stream, A, lda, B, ldb, C, ldc, M, N,
                           K, H, ldh);
    else
        hipLaunchKernelGGL(k_gemm_nn_f32<RACER_EPI_NONE>, grid, block, 0, dev->stream, A, lda, B, ldb, C, ldc, M, N,
                           K, H, ldh);
    return racer_launch_check("gemm_nn_f32");
}

int racer_mask_k3(racer_device* dev, const float* A, const float* B, int64_t ldb, void* C, bool out16, int64_t ldc,
                  int M, int N, const float* H, int64_t ldh) {
    if (M <= 0 || N <= 0) return 0;
    RACER_REQUIRE(M <= 65535, "the K = 3 masked product takes at most 65535 rows");
    const bool vec = N % 8 == 0 && ldb % 8 == 0 && ldc % 8 == 0 && ldh % 8 == 0 && (uintptr_t)B % 16 == 0 &&
                     (uintptr_t)C % 16 == 0 && (uintptr_t)H % 16 == 0;
    const int per = vec ? 8 : 1;
    dim3 grid((unsigned)((N / per + 255) / 256), (unsigned)M), block(256);
    racer_timed timer(dev, "mask_k3");
    auto kernel = out16 ? (vec ? k_mask_k3<true, true> : k_mask_k3<true, false>)
                        : (vec ? k_mask_k3<false, true> : k_mask_k3<false, false>);
    hipLaunchKernelGGL(kernel, grid, block, 0, dev->stream, A, B, ldb, C, ldc, N, H, ldh);
    return racer_launch_check("mask_k3");
}

// The batch splits into chunks of 4096; each chunk's partial product fills one
// [M, ldc] slab of the workspace.
constexpr int NT_F32_CHUNK = 4096;

size_t racer_gemm_nt_f32_partials(int M, int64_t ldc, int K) {
    int chunks = K <= 0 ? 1 : (K + NT_F32_CHUNK - 1) / NT_F32_CHUNK;
    return (size_t)M * (size_t)ldc * (size_t)chunks;
}

int racer_gemm_nt_f32(racer_device* dev, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                      int64_t ldc, int M, int N, int K, float* ws, size_t ws_len) {
    if (M <= 0 || N <= 0) return 0;
    const int chunk = NT_F32_CHUNK;
    int chunks = K <= 0 ? 1 : (K + chunk - 1) / chunk;
    int64_t stride = (int64_t)M * ldc;
    float* partial = nullptr;
    if (racer_partials(dev, racer_gemm_nt_f32_partials(M, ldc, K), ws, ws_len, &partial)) return 1;
    const bool reach = ((uint64_t)(M - 1) * lda + K) * 4 < 0xFFFFFF00ull && ((uint64_t)(N - 1) * ldb + K) * 4 < 0xFFFFFF00ull;
    const bool aligned4 = ldb % 4 == 0 && K % 4 == 0 && (uintptr_t)B % 16 == 0;  // float4 loads of B
    if (M <= 4 && reach && aligned4) {
        racer_timed timer(dev, "gemm_nt_f32_skinny");
        dim3 grid((N + SK_N - 1) / SK_N, 1, chunks), block(256);
        hipLaunchKernelGGL(k_gemm_nt_f32_skinny, grid, block, 0, dev->stream, A, lda, B, ldb, partial, stride,
                           ldc, M, N, K, chunk);
    } else if (reach && lda % 4 == 0 && ldb % 4 == 0 && K % 4 == 0 && (uintptr_t)A % 16 == 0 &&
               (uintptr_t)B % 16 == 0) {
        racer_timed timer(dev, "gemm_nt_f32");
        dim3 grid((N + 63) / 64, (M + 63) / 64, chunks), block(256);
        hipLaunchKernelGGL(k_gemm_nt_f32_mfma, grid, block, 0, dev->stream, A, lda, B, ldb, partial, stride, ldc, M,
                           N, K, chunk);
    } else {
        racer_timed timer(dev, "gemm_nt_f32");
        dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, chunks), block(256);
        hipLaunchKernelGGL(k_gemm_nt_f32, grid, block, 0, dev->stream, A, lda, B, ldb, partial, stride, ldc, M,
                           N, K, chunk);
    }
    if (racer_launch_check("gemm_nt_f32")) return 1;
    int64_t total = (int64_t)M * N;
    {
        racer_timed timer(dev, "sum_chunks");
        hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, dev->stream,
                           partial, stride, chunks, C, ldc, M, N);
    }
    return racer_launch_check("sum_chunks");
}
