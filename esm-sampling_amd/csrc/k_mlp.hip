// k_mlp.hip — the learned dispatch-block proposal: a small MLP, trained on the sampler's own
// states by maximum likelihood, proposes the gas utilization of one node over a block of B
// hours; every proposal is accepted or rejected by the exact Metropolis–Hastings ratio, so the
// target distribution of the sampler is unchanged.
//
// The move (node n, block d = steps dB … dB+B−1, replica j) replaces y_n over the block by
// y' ~ q(·|c), where the context c does not depend on the block's y: the requirement the gas
// unit meets, Q_τ = r_τ + κ cl(y_τ) (the residual without the unit's output), per hour of the
// block over κ; the node's three capacity factors per hour; its six capacities relative to the
// system mean; a constant. q is a product over the block's hours of mixtures of K Gaussians
// whose weights, means (around cl(Q_τ/κ)) and log standard deviations are the MLP's output. With
// flows and every other
// unit fixed, ΔC(y') is the block's gas terms and node balances (the terms of the gas atom,
// src/local.jl:991-1020), and the move is accepted with probability
// min(1, exp(−β_v ΔC(y') + log q(y|c) − log q(y'|c))) (FP64). Instances of different (node,
// block, replica) touch disjoint values, so all of them run in one launch.
//
// Network: din = 160 inputs (6B + 7 used), three hidden layers of 256 with ReLU, dout = 3KB outputs
// rounded up to a multiple of 64, BF16 weights and activations, FP32 accumulation, FP32 master
// weights and Adam moments. The GEMMs are one LDS-tiled kernel on v_mfma_f32_32x32x16_bf16 (64 × 64 tile per
// workgroup, four waves of 32 × 32), with transposed operands for the backward pass.
#include "sweep_core.h"

namespace esm {

// ---------------------------------------------------------------------------------------------
// BF16
// ---------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint16_t f2bf(float x) {   // round to nearest even
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------------------------
// GEMM: C[i][j] = Σ_k a(i, k) b(k, j) over k < K, i < M, j < N, with
//   a(i, k) = TA ? A[k·lda + i] : A[i·lda + k],  b(k, j) = TB ? B[j·ldb + k] : B[k·ldb + j]
// epilogues: EPI_F32 store FP32 (+ bias), EPI_RELU bias + ReLU to BF16, EPI_MASK BF16 of
// C ⊙ (H > 0) (H: BF16 [M][N], the layer's activation), EPI_PART the FP32 partial sum of K-split
// blockIdx.z (K-chunk kc) stored at C[z][i][j] (ldc = N), summed over the splits in split order
// by k_sum_splits: the gradient does not depend on the order in which workgroups finish
// ---------------------------------------------------------------------------------------------
enum { EPI_F32, EPI_RELU, EPI_MASK, EPI_PART };
constexpr int GB_M = 64, GB_N = 64, GB_K = 32, G_PAD = 8;

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, int kc, const uint16_t *__restrict__ A, int lda,
                                               const uint16_t *__restrict__ B, int ldb, const float *__restrict__ bias,
                                               const uint16_t *__restrict__ H, void *__restrict__ C, int ldc) {
    __shared__ uint16_t As[GB_M][GB_K + G_PAD];   // [i][k]
    __shared__ uint16_t Bs[GB_N][GB_K + G_PAD];   // [j][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * GB_M, j0 = blockIdx.x * GB_N;
    const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
    const int wr = wave >> 1, wc = wave & 1;
    f32x16 acc = {0};
    for (int k0 = kb; k0 < ke; k0 += GB_K) {
        // stage the tiles: 64 × 32 values of each operand, 8 per thread
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int idx = tid * 8 + e;                 // 0 … 2047
            int r, c;
            if (TA) { c = idx / GB_M; r = idx % GB_M; }  // consecutive threads along i (contiguous in A)
            else { r = idx / GB_K; c = idx % GB_K; }     // along k
            const int i = i0 + r, k = k0 + c;
            uint16_t v = 0;
            if (i < M && k < ke) v = TA ? A[(size_t)k * lda + i] : A[(size_t)i * lda + k];
            As[r][c] = v;
        }
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int idx = tid * 8 + e;
            int r, c;
            if (TB) { r = idx / GB_K; c = idx % GB_K; }  // along k (contiguous in B)
            else { c = idx / GB_N; r = idx % GB_N; }     // along j
            const int j = j0 + r, k = k0 + c;
            uint16_t v = 0;
            if (j < N && k < ke) v = TB ? B[(size_t)j * ldb + k] : B[(size_t)k * ldb + j];
            Bs[r][c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 16) {
            u16x8 a, b;
            const int ra = wr * 32 + (lane & 31), rb = wc * 32 + (lane & 31), ko = kk + 8 * (lane >> 5);
#pragma unroll
            for (int v = 0; v < 8; v++) {
                a[v] = As[ra][ko + v];
                b[v] = Bs[rb][ko + v];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                          acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int i = i0 + wr * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), j = j0 + wc * 32 + (lane & 31);
        if (i >= M || j >= N) continue;
        float x = acc[v];
        if (EPI == EPI_PART) {
            ((float *)C)[((size_t)blockIdx.z * M + i) * ldc + j] = x;
            continue;
        }
        if (bias && EPI != EPI_MASK) x += bias[j];
        if (EPI == EPI_F32) ((float *)C)[(size_t)i * ldc + j] = x;
        else if (EPI == EPI_RELU) ((uint16_t *)C)[(size_t)i * ldc + j] = f2bf(fmaxf(x, 0.0f));
        else ((uint16_t *)C)[(size_t)i * ldc + j] = H[(size_t)i * ldc + j] != 0 && !(H[(size_t)i * ldc + j] & 0x8000u) ? f2bf(x) : (uint16_t)0;
    }
}

// The forward GEMM C[i][j] = Σ_k A[i][k] B[j][k] (+ bias, ReLU to BF16 or FP32) with both operands
// K-contiguous (activations [M][K], weights Wt [N][K]): 128 × 128 tile per workgroup, four waves
// of 64 × 64 (2 × 2 MFMA 32 × 32 × 16 tiles each), K in steps of 32 staged in LDS with 16-byte
// global loads and ds_read_b128 fragment reads, the next K step prefetched into registers while
// the current one is multiplied. K must be a multiple of 32 and lda, ldb multiples of 8.
constexpr int GN_M = 128, GN_N = 128, GN_K = 32, GN_LD = GN_K + 8;

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_nt(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
                                                 const uint16_t *__restrict__ B, int ldb, const float *__restrict__ bias,
                                                 void *__restrict__ C, int ldc) {
    __shared__ __attribute__((aligned(16))) uint16_t As[GN_M][GN_LD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[GN_N][GN_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * GN_M, j0 = blockIdx.x * GN_N;
    const int wr = wave >> 1, wc = wave & 1;
    // global → register staging: 2 chunks of 8 values per operand per thread (row = c / 4, k = 8 (c % 4))
    uint4 ra[2], rb[2];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int c = tid + 256 * e, r = c >> 2, kc = (c & 3) * 8;
            const int i = i0 + r, j = j0 + r;
            ra[e] = i < M ? *(const uint4 *)(A + (size_t)i * lda + k0 + kc) : make_uint4(0, 0, 0, 0);
            rb[e] = j < N ? *(const uint4 *)(B + (size_t)j * ldb + k0 + kc) : make_uint4(0, 0, 0, 0);
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int c = tid + 256 * e, r = c >> 2, kc = (c & 3) * 8;
            *(uint4 *)&As[r][kc] = ra[e];
            *(uint4 *)&Bs[r][kc] = rb[e];
        }
    };
    f32x16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x16{0};
    load(0);
    store();
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += GN_K) {
        const bool more = k0 + GN_K < K;
        if (more) load(k0 + GN_K);
#pragma unroll
        for (int kk = 0; kk < GN_K; kk += 16) {
            const int ko = kk + 8 * (lane >> 5);
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                fa[q] = *(const bf16x8 *)&As[wr * 64 + q * 32 + (lane & 31)][ko];
                fb[q] = *(const bf16x8 *)&Bs[wc * 64 + q * 32 + (lane & 31)][ko];
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            store();
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int j = j0 + wc * 64 + b * 32 + (lane & 31);
            if (j >= N) continue;
            const float bj = bias ? bias[j] : 0.0f;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int i = i0 + wr * 64 + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (i >= M) continue;
                const float x = acc[a][b][v] + bj;
                if (EPI == EPI_F32) ((float *)C)[(size_t)i * ldc + j] = x;
                else ((uint16_t *)C)[(size_t)i * ldc + j] = f2bf(fmaxf(x, 0.0f));
            }
        }
}

// The wide form for N ≥ 192: a 128 × 64·NB tile per workgroup (NB = 4: 256 columns; NB = 3: 192,
// which the 576 outputs of 8 components × 24 hours fill exactly), the activations of a row block
// read once per layer; four waves of 64 × 32·NB (2 × NB MFMA tiles: each A fragment feeds NB
// MFMAs), K in steps of 32 through a double-buffered LDS ring, one barrier per step.
constexpr int GW_N = 256;

template <int EPI, int NB>
__global__ void __launch_bounds__(256, 2) k_gemm_nt_w(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
                                                   const uint16_t *__restrict__ B, int ldb, const float *__restrict__ bias,
                                                   void *__restrict__ C, int ldc) {
    __shared__ __attribute__((aligned(16))) uint16_t As[2][GN_M][GN_LD];
    constexpr int TN = 64 * NB;
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][TN][GN_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * GN_M, j0 = blockIdx.x * TN;
    const int wr = wave >> 1, wc = wave & 1;
    // staging: A 128 rows × 4 chunks (2 per thread), B 64·NB rows × 4 chunks (NB per thread)
    uint4 ra[2], rb[NB];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int c = tid + 256 * e, r = c >> 2, kc = (c & 3) * 8, i = i0 + r;
            ra[e] = i < M ? *(const uint4 *)(A + (size_t)i * lda + k0 + kc) : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < NB; e++) {
            const int c = tid + 256 * e, r = c >> 2, kc = (c & 3) * 8, j = j0 + r;
            rb[e] = j < N ? *(const uint4 *)(B + (size_t)j * ldb + k0 + kc) : make_uint4(0, 0, 0, 0);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int c = tid + 256 * e;
            *(uint4 *)&As[buf][c >> 2][(c & 3) * 8] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < NB; e++) {
            const int c = tid + 256 * e;
            *(uint4 *)&Bs[buf][c >> 2][(c & 3) * 8] = rb[e];
        }
    };
    f32x16 acc[2][NB];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) acc[a][b] = f32x16{0};
    load(0);
    store(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < K; k0 += GN_K) {
        const bool more = k0 + GN_K < K;
        if (more) load(k0 + GN_K);
#pragma unroll
        for (int kk = 0; kk < GN_K; kk += 16) {
            const int ko = kk + 8 * (lane >> 5);
            bf16x8 fa[2], fb[NB];
#pragma unroll
            for (int q = 0; q < 2; q++) fa[q] = *(const bf16x8 *)&As[cur][wr * 64 + q * 32 + (lane & 31)][ko];
#pragma unroll
            for (int q = 0; q < NB; q++) fb[q] = *(const bf16x8 *)&Bs[cur][wc * 32 * NB + q * 32 + (lane & 31)][ko];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < NB; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
        if (more) store(cur ^ 1);   // the other buffer: its last readers passed the previous barrier
        __syncthreads();
        cur ^= 1;
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = j0 + wc * 32 * NB + b * 32 + (lane & 31);
            if (j >= N) continue;
            const float bj = bias ? bias[j] : 0.0f;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int i = i0 + wr * 64 + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (i >= M) continue;
                const float x = acc[a][b][v] + bj;
                if (EPI == EPI_F32) ((float *)C)[(size_t)i * ldc + j] = x;
                else ((uint16_t *)C)[(size_t)i * ldc + j] = f2bf(fmaxf(x, 0.0f));
            }
        }
}

template <int EPI>
static int gemm_nt(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                   const float *bias, void *C, int ldc) {
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    if (K % GN_K || lda % 8 || ldb % 8 || ((uintptr_t)A & 15) || ((uintptr_t)B & 15)) {
        set_error("gemm_nt: K must be a multiple of 32, operands 16-byte aligned");
        return ESM_EINVAL;
    }
    if (N % 192 == 0 && N % GW_N != 0) {   // 192-wide tiles fill N exactly
        const dim3 grid((unsigned)(N / 192), (unsigned)((M + GN_M - 1) / GN_M));
        hipLaunchKernelGGL((k_gemm_nt_w<EPI, 3>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, bias, C, ldc);
    } else if (N >= GW_N) {
        const dim3 grid((unsigned)((N + GW_N - 1) / GW_N), (unsigned)((M + GN_M - 1) / GN_M));
        hipLaunchKernelGGL((k_gemm_nt_w<EPI, 4>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, bias, C, ldc);
    } else {
        const dim3 grid((unsigned)((N + GN_N - 1) / GN_N), (unsigned)((M + GN_M - 1) / GN_M));
        hipLaunchKernelGGL((k_gemm_nt<EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, bias, C, ldc);
    }
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

// C[i][j] += Σ_z part[z][i][j] over the nz K-splits, in split order
__global__ void k_sum_splits(const float *__restrict__ part, int nz, int M, int N, float *__restrict__ C, int ldc) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)M * N) return;
    const int i = (int)(g / N), j = (int)(g % N);
    float acc = 0.0f;
    for (int z = 0; z < nz; z++) acc += part[((size_t)z * M + i) * N + j];
    C[(size_t)i * ldc + j] += acc;
}

constexpr int KSPLIT_MAX = 32;   // K-splits of the weight-gradient GEMMs (≥ 512 workgroups at 256 × 256)

template <bool TA, bool TB, int EPI>
static int gemm(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                const float *bias, const uint16_t *H, void *C, int ldc, int ksplit = 1, float *part = nullptr) {
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    const int kc = ((K + ksplit - 1) / ksplit + GB_K - 1) / GB_K * GB_K;
    const int nz = (K + kc - 1) / kc;
    const dim3 grid((unsigned)((N + GB_N - 1) / GB_N), (unsigned)((M + GB_M - 1) / GB_M), (unsigned)nz);
    if (EPI == EPI_PART) {
        if (!part || nz > KSPLIT_MAX) {
            set_error("gemm: split partials need a buffer of at most KSPLIT_MAX splits");
            return ESM_EINVAL;
        }
        hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, kc, A, lda, B, ldb, bias, H, part, N);
        hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)(((size_t)M * N + 255) / 256)), dim3(256), 0, s, part, nz, M, N,
                           (float *)C, ldc);
    } else {
        hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, kc, A, lda, B, ldb, bias, H, C, ldc);
    }
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------------------------
// the instances, their features and the mixture head
// ---------------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------------
// Fused forward pass: the four layers of a 128-row tile in one workgroup. The tile's activations
// stay in LDS (128 × 256 BF16, written in place by each hidden layer's epilogue after every wave
// has read them); the weights stream through a double-buffered LDS ring in K steps of 64, 64·NB
// columns per pass (NB = 4 for the hidden layers, 3 for the output layer's 192-column passes);
// eight waves of 32 rows × 32·NB columns (NB MFMA tiles), two per SIMD (141 KB of LDS per
// workgroup). Measured per 65,536-row chunk: 114 µs (four waves with 32-wide K steps: 175 µs,
// the time of the four layered launches). Each output element accumulates
// over K in the same order (steps of 32, two v_mfma_f32_32x32x16_bf16 per step) and passes the
// same bias, ReLU and BF16 rounding as in the layered kernels, so the outputs are bit for bit
// theirs; the activations never reach HBM (the layered pass writes and re-reads 3 × 128 KB per
// 256 rows and launches four kernels).
// ---------------------------------------------------------------------------------------------
constexpr int FF_M = 128, FF_LD = 256 + 8;

struct FwdArgs {
    const uint16_t *w[4];   // BF16 Wt [out][in] of each layer
    const float *b[4];      // FP32 bias [out]
    int din, dh, dout;
};

// one pass of columns j0 … j0 + 64·NB of W [nout][K] against the tile's activations [128][K]:
// FF_WAVES waves in a (FF_WAVES/2) × 2 grid, each 128/(FF_WAVES/2) rows (MT MFMA row tiles) × 32·NB
// columns
constexpr int FF_WAVES = 8, FF_MT = 128 / (FF_WAVES / 2) / 32, FF_K = 64, FF_BLD = FF_K + 8;
template <int NB>
__device__ __forceinline__ void ff_pass(uint16_t (*Act)[FF_LD], uint16_t (*Bs)[256][FF_BLD], const uint16_t *W, int K,
                                        int nout, int j0, f32x16 (&acc)[FF_MT][NB]) {
    constexpr int NT = 64 * FF_WAVES, CPR = FF_K / 8, CH = 64 * NB * CPR, E = (CH + NT - 1) / NT;   // 8-element chunks of a ring slot
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave >> 1, wc = wave & 1;
    uint4 rb[E];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int c = tid + NT * e, r = c / CPR, kc = (c % CPR) * 8, j = j0 + r;
            rb[e] = c < CH && j < nout && k0 + kc < K ? *(const uint4 *)(W + (size_t)j * K + k0 + kc)
                                                     : make_uint4(0, 0, 0, 0);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int c = tid + NT * e;
            if (c < CH) *(uint4 *)&Bs[buf][c / CPR][(c % CPR) * 8] = rb[e];
        }
    };
#pragma unroll
    for (int a = 0; a < FF_MT; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) acc[a][b] = f32x16{0};
    load(0);
    __syncthreads();   // the previous pass's readers of the ring and writers of the activations are done
    store(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < K; k0 += FF_K) {
        const bool more = k0 + FF_K < K;
        if (more) load(k0 + FF_K);
#pragma unroll
        for (int kk = 0; kk < FF_K; kk += 16) {
            if (k0 + kk >= K) break;   // K a multiple of 32: the last step may hold half
            const int ko = kk + 8 * (lane >> 5);
            bf16x8 fa[FF_MT], fb[NB];
#pragma unroll
            for (int q = 0; q < FF_MT; q++)
                fa[q] = *(const bf16x8 *)&Act[wr * 32 * FF_MT + q * 32 + (lane & 31)][k0 + ko];
#pragma unroll
            for (int q = 0; q < NB; q++) fb[q] = *(const bf16x8 *)&Bs[cur][wc * 32 * NB + q * 32 + (lane & 31)][ko];
#pragma unroll
            for (int a = 0; a < FF_MT; a++)
#pragma unroll
                for (int b = 0; b < NB; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
        if (more) store(cur ^ 1);   // the other buffer: its last readers passed the previous barrier
        __syncthreads();            // after the last step: every wave has read the activations
        cur ^= 1;
    }
}

// rows [M][din] BF16 → outputs [M][dout] FP32; din a multiple of 32 and ≤ 256, dh = 256
__global__ void __launch_bounds__(64 * FF_WAVES, 1) k_mlp_fwd(int M, FwdArgs f, const uint16_t *__restrict__ X,
                                                              float *__restrict__ OUT) {
    __shared__ __attribute__((aligned(16))) uint16_t Act[FF_M][FF_LD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][256][FF_BLD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave >> 1, wc = wave & 1;
    const int i0 = blockIdx.x * FF_M;
    const int nch = f.din / 8;   // 8-element chunks of an input row
    for (int c = tid; c < FF_M * nch; c += 64 * FF_WAVES) {
        const int r = c / nch, kc = (c % nch) * 8, i = i0 + r;
        *(uint4 *)&Act[r][kc] = i < M ? *(const uint4 *)(X + (size_t)i * f.din + kc) : make_uint4(0, 0, 0, 0);
    }
    // hidden layers: the tile's 256 outputs in one pass, ReLU'd BF16 back into the activations
    for (int l = 0; l < 3; l++) {
        f32x16 acc[FF_MT][4];
        ff_pass<4>(Act, Bs, f.w[l], l == 0 ? f.din : f.dh, f.dh, 0, acc);
#pragma unroll
        for (int a = 0; a < FF_MT; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int j = wc * 128 + b * 32 + (lane & 31);
                const float bj = f.b[l][j];
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int r = wr * 32 * FF_MT + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                    Act[r][j] = f2bf(fmaxf(acc[a][b][v] + bj, 0.0f));
                }
            }
    }
    // output layer: 192 columns per pass
    for (int j0 = 0; j0 < f.dout; j0 += 192) {
        f32x16 acc[FF_MT][3];
        ff_pass<3>(Act, Bs, f.w[3], f.dh, f.dout, j0, acc);
#pragma unroll
        for (int a = 0; a < FF_MT; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const int j = j0 + wc * 96 + b * 32 + (lane & 31);
                if (j >= f.dout) continue;
                const float bj = f.b[3][j];
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int i = i0 + wr * 32 * FF_MT + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                    if (i < M) OUT[(size_t)i * f.dout + j] = acc[a][b][v] + bj;
                }
            }
    }
}

struct MlpDims {
    int B, K, din, dh, dout;
    int nb;                  // blocks per node: T / B
    long long n_inst;        // N · nb · R
};

// instance m = (n · nb + d) · R + j
__host__ __device__ __forceinline__ void inst_of(long long m, int nb, int R, int &n, int &d, int &j) {
    j = (int)(m % R);
    const long long q = m / R;
    d = (int)(q % nb);
    n = (int)(q / nb);
}

// Instance kernels run one 32-lane group per instance (B ≤ 24 hours, so lane b < B owns hour b),
// eight groups per 256-thread block; sums over the hours are butterflies within the group.
constexpr int MLP_G = 32, MLP_GPB = 8;

__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = MLP_G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, MLP_G);
    return v;
}

// features of instances idx[t] (or m0 + t without idx), t < cnt, into X [cnt][din] (BF16), the
// block's current y into Y [cnt][B] and the centres cl(Q/κ) into MU0 [cnt][B]: features
// [Q/κ per hour (B)][capacity factors per carrier and hour (3B)][capacities over the system
// means (6)][1][the node's net line inflow over κ per hour (B)][the largest loading |F|/s of its
// lines per hour (B)][0 …]
// A block builds the features of a tile of MLP_FT consecutive instances (consecutive replicas of
// one node and block, the replica being the fastest index): thread (i, h) takes instance i of the
// tile and the hours b ≡ h (mod 256 / MLP_FT), so every state read of a warp covers consecutive
// replicas of one hour (coalesced); the tile's rows are assembled in LDS and written out as
// contiguous 16-byte stores.
constexpr int MLP_FT = 32;
constexpr int MLP_DIN = 160;

template <typename S>
__global__ void __launch_bounds__(256) k_mlp_features(DevModel<S> m, SweepState<S> st, const double *kraw, MlpDims D,
                                                      long long m0, const long long *idx, int cnt, uint16_t *X, float *Y,
                                                      float *MU0) {
    __shared__ __attribute__((aligned(16))) uint16_t xs[MLP_FT * MLP_DIN];
    __shared__ float ys[MLP_FT * 24], mus[MLP_FT * 24];
    __shared__ double kmean[ESM_NK];
    const int tid = threadIdx.x, ti = tid % MLP_FT, hp = tid / MLP_FT;
    constexpr int NHP = 256 / MLP_FT;
    const long long t0 = (long long)blockIdx.x * MLP_FT;
    const int nin = (int)min((long long)MLP_FT, (long long)cnt - t0);
    const int B = D.B, din = D.din;
    if (tid < ESM_NK) {
        double mean = 0.0;
        for (int q = 0; q < m.N; q++) mean += kraw[tid * m.N + q];
        kmean[tid] = mean / m.N;
    }
    __syncthreads();
    if (ti < nin) {
        int n, d, j;
        inst_of(idx ? idx[t0 + ti] : m0 + t0 + ti, D.nb, m.R, n, d, j);
        const size_t TR = (size_t)m.T * m.R;
        uint16_t *x = xs + ti * din;
        const S kap = m.kc[ESM_KGAS * m.N + n];
        for (int b = hp; b < B; b += NHP) {
            const size_t o = (size_t)(d * B + b) * m.R + j;
            const S y = st.v[plane(m.N, m.T, m.R) + (size_t)n * TR + o];
            const S Q = st.r[(size_t)n * TR + o] + kap * clampv(y, S(0), S(1));
            const float q = kap > S(0) ? (float)(Q / kap) : 0.0f;
            x[b] = f2bf(q);
            ys[ti * B + b] = (float)y;
            mus[ti * B + b] = fminf(fmaxf(q, 0.0f), 1.0f);
#pragma unroll
            for (int c = 0; c < 3; c++) x[B + c * B + b] = f2bf((float)st.x[((size_t)c * m.N + n) * TR + o]);
            // net line inflow / κ and the largest incident loading |F| / s
            S in = S(0), mx = S(0);
            for (int e = m.adj_ptr[n]; e < m.adj_ptr[n + 1]; e++) {
                const int l = m.adj_line[e];
                const S F = st.F[(size_t)l * TR + o];
                in += m.bus1[l] == n ? F : -F;      // F flows from bus0 to bus1
                mx = fmax(mx, fabs(F) / m.ls[l]);
            }
            x[4 * B + ESM_NK + 1 + b] = f2bf(kap > S(0) ? (float)(in / kap) : 0.0f);
            x[5 * B + ESM_NK + 1 + b] = f2bf((float)mx);
        }
        if (hp == 0) {
            for (int c = 0; c < ESM_NK; c++)
                x[4 * B + c] = f2bf(kmean[c] > 0 ? (float)(kraw[c * m.N + n] / kmean[c]) : 0.0f);
            x[4 * B + ESM_NK] = f2bf(1.0f);
        }
        for (int f = 6 * B + ESM_NK + 1 + hp; f < din; f += NHP) x[f] = f2bf(0.0f);
    }
    __syncthreads();
    // the tile's rows are contiguous in X, Y and MU0
    const uint4 *src = (const uint4 *)xs;
    uint4 *dst = (uint4 *)(X + t0 * din);
    for (int i = tid; i < nin * din / 8; i += 256) dst[i] = src[i];
    for (int i = tid; i < nin * B; i += 256) {
        Y[t0 * B + i] = ys[i];
        MU0[t0 * B + i] = mus[i];
    }
}

// the proposal density of instance row o [dout]: independent hours, each a mixture of K Gaussians
// with its own weights (log-softmax over k of the logits), means (around cl(Q/κ)) and log
// standard deviations (clamped to [−7, 2]); layout [logits K·B][means K·B][log sd K·B], index
// k·B + b. (With flows and every other unit fixed, the target of the block is a product over its
// hours, so the family is one of products too.)
// log Σ_k exp(logit_kb): the normalizer of hour b's weights
__device__ __forceinline__ float mix_lse(const float *o, int K, int B, int b) {
    float mx = -INFINITY;
    for (int q = 0; q < K; q++) mx = fmaxf(mx, o[q * B + b]);
    float s = 0.0f;
    for (int q = 0; q < K; q++) s += expf(o[q * B + b] - mx);
    return mx + logf(s);
}
__device__ __forceinline__ float mix_logw(const float *o, int B, int k, int b, float lse) { return o[k * B + b] - lse; }
__device__ __forceinline__ float mix_mu(const float *o, const float *mu0, int K, int B, int k, int b) {
    return mu0[b] + o[K * B + k * B + b];
}
__device__ __forceinline__ float mix_ls(const float *o, int K, int B, int k, int b) {
    return fminf(fmaxf(o[2 * K * B + k * B + b], -7.0f), 2.0f);
}

// log q_b(y_b) of hour b in FP64 and the responsibilities of its components
__device__ double mix_logq_hour(const float *o, const float *mu0, double y, int K, int B, int b, float lse,
                                double *resp) {
    double lk[8], mx = -INFINITY;
    for (int k = 0; k < K; k++) {
        const double ls = mix_ls(o, K, B, k, b), z = (y - (double)mix_mu(o, mu0, K, B, k, b)) * exp(-ls);
        lk[k] = (double)mix_logw(o, B, k, b, lse) - 0.5 * z * z - ls - 0.91893853320467274178;
        mx = fmax(mx, lk[k]);
    }
    double tot = 0.0;
    for (int k = 0; k < K; k++) tot += exp(lk[k] - mx);
    const double lq = mx + log(tot);
    if (resp)
        for (int k = 0; k < K; k++) resp[k] = exp(lk[k] - lq);
    return lq;
}

// loss −log q(y|c) of the rows and its gradient dL/dout (BF16 [cnt][dout], scaled by `scale`);
// lane b of a row's group owns hour b
__global__ void __launch_bounds__(256) k_mlp_head_grad(const float *OUT, const float *MU0, const float *Y, int cnt, int K,
                                                       int B, int dout, float scale, uint16_t *DOUT, double *loss) {
    const int t = blockIdx.x * MLP_GPB + (threadIdx.x >> 5), b = threadIdx.x & (MLP_G - 1);
    double l = 0.0;
    if (t < cnt) {
        const float *o = OUT + (size_t)t * dout, *mu0 = MU0 + (size_t)t * B;
        uint16_t *g = DOUT + (size_t)t * dout;
        if (b < B) {
            double resp[8];
            const double y = (double)Y[(size_t)t * B + b];
            const float lse = mix_lse(o, K, B, b);
            l = -mix_logq_hour(o, mu0, y, K, B, b, lse, resp);
            for (int k = 0; k < K; k++) {
                const int q = k * B + b;
                // d(−log q_b)/d logit = softmax − resp
                g[q] = f2bf(scale * (float)(exp((double)mix_logw(o, B, k, b, lse)) - resp[k]));
                const float ls = mix_ls(o, K, B, k, b), raw = o[2 * K * B + q];
                const double sd = exp((double)ls), z = (y - (double)mix_mu(o, mu0, K, B, k, b)) / sd;
                g[K * B + q] = f2bf(scale * (float)(-resp[k] * z / sd));                     // d/dμ
                const bool free = raw > -7.0f && raw < 2.0f;                               // inside the clamp
                g[2 * K * B + q] = f2bf(free ? scale * (float)(resp[k] * (1.0 - z * z)) : 0.0f);   // d/d log sd
            }
        }
        for (int q = 3 * K * B + b; q < dout; q += MLP_G) g[q] = 0;
    }
    // block sum of the loss, one atomic per block
    __shared__ double red[256];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[blockIdx.x] = red[0];
}

// the learned move of instances m0 … m0 + cnt − 1: draw y' from the proposal, accept with the
// exact MH ratio, commit. Lane b of the instance's group owns hour b: its draw, its terms of
// log q(y|c), log q(y'|c) and ΔC (FP64), summed over the group; lane 0 decides. Uniforms:
// Philox (replica, node·nb + block, ·, ctr) with 0x5a000000 + b/2 for the normals of hours
// b, b + 1 (Box–Muller: words 0–1 and 2–3 give the radius and the angle), 0x5e000000 + b/2 for
// their components (words 0–1 hour b, 2–3 hour b + 1), 0x5b000000 words 2–3 for the acceptance.
template <typename S>
__global__ void __launch_bounds__(256) k_mlp_propose(DevModel<S> m, SweepState<S> st, MlpDims D, long long m0, int cnt,
                                                     const float *OUT, const float *MU0, uint32_t k0, uint32_t k1,
                                                     uint32_t ctr, S beta, unsigned long long *stats, float *YP,
                                                     double *LR) {
    const int t = blockIdx.x * MLP_GPB + (threadIdx.x >> 5), b = threadIdx.x & (MLP_G - 1);
    const int B = D.B, K = D.K;
    const bool live = t < cnt, hour = live && b < B;
    int n = 0, d = 0, j = 0;
    if (live) inst_of(m0 + t, D.nb, m.R, n, d, j);
    const uint32_t site = (uint32_t)(n * D.nb + d);
    const size_t TR = (size_t)m.T * m.R, o_ = (size_t)n * TR + (size_t)(d * B + b) * m.R + j;
    double lqo = 0.0, lqn = 0.0, dC = 0.0;
    S y0 = S(0), y1 = S(0), dp = S(0);
    if (hour) {
        const float *o = OUT + (size_t)t * D.dout, *mu0 = MU0 + (size_t)t * B;
        uint32_t w[4], wc[4];
        philox4x32(k0, k1, (uint32_t)j, site, 0x5a000000u + (uint32_t)(b / 2), ctr, w);
        philox4x32(k0, k1, (uint32_t)j, site, 0x5e000000u + (uint32_t)(b / 2), ctr, wc);
        const double u1 = 1.0 - u01(w[0], w[1]), u2 = u01(w[2], w[3]);
        const double rad = sqrt(-2.0 * log(u1));
        const double z = (b & 1) ? rad * sin(6.283185307179586477 * u2) : rad * cos(6.283185307179586477 * u2);
        const double uc = (b & 1) ? u01(wc[2], wc[3]) : u01(wc[0], wc[1]);
        const float lse = mix_lse(o, K, B, b);
        int kc = K - 1;
        double acc = 0.0;
        for (int k = 0; k < K; k++) {
            acc += exp((double)mix_logw(o, B, k, b, lse));
            if (uc < acc) {
                kc = k;
                break;
            }
        }
        const float yp = mix_mu(o, mu0, K, B, kc, b) + expf(mix_ls(o, K, B, kc, b)) * (float)z;
        // the reverse density at the state's own y (not its FP32 copy)
        y0 = st.v[plane(m.N, m.T, m.R) + o_];
        y1 = (S)yp;
        lqo = mix_logq_hour(o, mu0, (double)y0, K, B, b, lse, nullptr);
        lqn = mix_logq_hour(o, mu0, (double)y1, K, B, b, lse, nullptr);
        // ΔC of the hour: the gas terms and the node's balance
        const S kap = m.kc[ESM_KGAS * m.N + n], mc = m.mc[ESM_KGAS * m.N + n], S0 = st.r[o_];
        S c0, p0, c1, p1;
        gen_cost(mc, kap, kap + m.kappa0, m.dt, m.cv, y0, c0, p0);
        gen_cost(mc, kap, kap + m.kappa0, m.dt, m.cv, y1, c1, p1);
        dp = p1 - p0;
        const S S1 = S0 - dp;
        dC = (double)(c1 - c0) + (double)m.dt * ((double)phibal(S1, m.c_shed, m.c_dump) - (double)phibal(S0, m.c_shed, m.c_dump));
        if (YP) YP[(size_t)t * B + b] = yp;
    }
    lqo = group_sum(lqo);
    lqn = group_sum(lqn);
    dC = group_sum(dC);
    int acc_ = 0;
    MoveOut out;
    if (live && b == 0) {
        uint32_t w[4];
        philox4x32(k0, k1, (uint32_t)j, site, 0x5b000000u, ctr, w);
        const double ua = u01(w[2], w[3]);
        const double lr = -(double)beta * dC + lqo - lqn;
        acc_ = isfinite(lr) && ua < exp(fmin(0.0, lr));
        if (LR) LR[t] = lr;
        out.add(0, acc_, 1);
    }
    acc_ = __shfl(acc_, 0, MLP_G);
    if (hour && acc_) {
        st.r[o_] = st.r[o_] - dp;
        st.v[plane(m.N, m.T, m.R) + o_] = y1;
    }
    add_counts(stats, out);
}

// bias gradient: db[j] += Σ_i DOUT[i][j] (BF16 [cnt][n]); one thread per column and chunk of
// COLSUM_ROWS rows writes the chunk's sum to part[chunk][j]; k_sum_splits adds the chunks in order
constexpr int COLSUM_ROWS = 64;
__global__ void k_colsum(const uint16_t *G, int cnt, int n, float *part) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int r0 = blockIdx.y * COLSUM_ROWS, r1 = min(cnt, r0 + COLSUM_ROWS);
    float s = 0.0f;
    for (int i = r0; i < r1; i++) s += bf2f(G[(size_t)i * n + j]);
    part[(size_t)blockIdx.y * n + j] = s;
}

// loss[0] += Σ_b part[b] over nb block sums, in block order
__global__ void k_sum_loss(const double *__restrict__ part, int nb, double *__restrict__ loss) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double acc = 0.0;
    for (int b = 0; b < nb; b++) acc += part[b];
    loss[0] += acc;
}

// Adam on the FP32 master parameters; the BF16 copy of every layer's Wt [out][in]
__global__ void k_adam(float *p, const float *g, float *mm, float *vv, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    mm[i] = b1 * mm[i] + (1.0f - b1) * gi;
    vv[i] = b2 * vv[i] + (1.0f - b2) * gi * gi;
    p[i] -= lr * (mm[i] / c1) / (sqrtf(vv[i] / c2) + eps);
}
__global__ void k_to_bf16(const float *wt, int n, uint16_t *wt16) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wt16[i] = f2bf(wt[i]);
}


// ---------------------------------------------------------------------------------------------
// host side: the network of a handle, its forward and backward passes over chunks of instances
// ---------------------------------------------------------------------------------------------
__global__ void k_batch_idx(uint32_t k0, uint32_t k1, uint32_t ctr, uint32_t step, long long n_inst, int cnt,
                            long long *idx) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    uint32_t w[4];
    philox4x32(k0, k1, (uint32_t)t, step, 0x5c000000u, ctr, w);
    const unsigned long long u = ((unsigned long long)w[0] << 32) | w[1];
    idx[t] = (long long)(u % (unsigned long long)n_inst);
}

constexpr int MLP_LOSS_CHUNK = 1024;
// rows per chunk of instances (and the largest explicit batch): a chunk's buffers take ≈ 0.5 GB;
// larger chunks ran the GEMMs at the same rate per row
constexpr int MLP_CAP = 65536;

struct MlpState {
    MlpDims D{};
    int lin[4], lout[4];
    size_t ow[4], ob[4], np = 0;
    float *p = nullptr, *g = nullptr, *am = nullptr, *av = nullptr;   // master params, gradient, Adam moments
    uint16_t *wt16 = nullptr;                                         // BF16 Wt of every layer (offsets ow)
    long long adam_t = 0;
    bool fused = true;   // the proposals' and explicit rows' forward in one kernel (k_mlp_fwd)
    int cap = MLP_CAP;
    uint16_t *X = nullptr, *H[3] = {nullptr, nullptr, nullptr}, *DG[4] = {nullptr, nullptr, nullptr, nullptr};
    float *OUT = nullptr, *Y = nullptr, *MU0 = nullptr, *YP = nullptr;
    double *LR = nullptr, *dloss = nullptr;   // loss sums of up to MLP_LOSS_CHUNK training steps
    double *lpart = nullptr;                  // per-block loss sums of one step
    float *part = nullptr;                    // K-split and row-chunk partial sums of the gradients
    size_t part_len = 0;
    long long *idx = nullptr;
    unsigned long long *stats = nullptr;   // [ESM_NSTAT][STAT_SLOTS]
    std::vector<void *> allocs;
};

void mlp_free(esm_handle *h) {
    if (!h->mlp) return;
    for (void *q : h->mlp->allocs) (void)hipFree(q);
    delete h->mlp;
    h->mlp = nullptr;
}

static int mlp_refresh_bf16(esm_handle *h) {
    MlpState &s = *h->mlp;
    for (int l = 0; l < 4; l++) {
        const int n = s.lin[l] * s.lout[l];
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, s.p + s.ow[l], n,
                           s.wt16 + s.ow[l]);
        ESM_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

// forward of cnt rows of s.X into s.OUT (with prof: every GEMM launch timed on its own)
// keep: the layered kernels, which leave the activations in H for the backward pass
static int mlp_forward(esm_handle *h, int cnt, bool prof = false, bool keep = false) {
    MlpState &s = *h->mlp;
    const MlpDims &D = s.D;
    int rc;
    if (!keep && s.fused && D.dh == 256 && D.din % GN_K == 0 && D.din <= 256 && cnt > 0) {
        ProfScope ps(h, prof ? ESM_KERN_MLPF : -1);
        FwdArgs f{{s.wt16 + s.ow[0], s.wt16 + s.ow[1], s.wt16 + s.ow[2], s.wt16 + s.ow[3]},
                  {s.p + s.ob[0], s.p + s.ob[1], s.p + s.ob[2], s.p + s.ob[3]}, D.din, D.dh, D.dout};
        hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)((cnt + FF_M - 1) / FF_M)), dim3(64 * FF_WAVES), 0, h->stream, cnt,
                           f, s.X, s.OUT);
        ESM_HIP_CHECK(hipGetLastError());
        return 0;
    }
    const uint16_t *in = s.X;
    for (int l = 0; l < 3; l++) {
        ProfScope ps(h, prof ? ESM_KERN_MLPF : -1);
        if ((rc = gemm_nt<EPI_RELU>(h->stream, cnt, D.dh, s.lin[l], in, s.lin[l], s.wt16 + s.ow[l], s.lin[l],
                                     s.p + s.ob[l], s.H[l], D.dh)))
            return rc;
        in = s.H[l];
    }
    ProfScope ps(h, prof ? ESM_KERN_MLPF : -1);
    return gemm_nt<EPI_F32>(h->stream, cnt, D.dout, D.dh, in, D.dh, s.wt16 + s.ow[3], D.dh, s.p + s.ob[3], s.OUT,
                            D.dout);
}

// backward of cnt rows from s.DG[3] = dL/dOUT into s.g (added)
static int mlp_backward(esm_handle *h, int cnt) {
    MlpState &s = *h->mlp;
    const MlpDims &D = s.D;
    int rc;
    // 256 rows per K-split (≥ 128 workgroups at 4096 rows), at most KSPLIT_MAX splits
    const int ksplit = std::min(KSPLIT_MAX, (cnt + 255) / 256);
    const int nchunk = (cnt + COLSUM_ROWS - 1) / COLSUM_ROWS;
    for (int l = 3; l >= 0; l--) {
        const uint16_t *act = l == 0 ? s.X : s.H[l - 1];   // the layer's input [cnt][lin]
        // dWt [lout][lin] += Σ_m DG[l][m][o] act[m][k]
        if ((rc = gemm<true, false, EPI_PART>(h->stream, s.lout[l], s.lin[l], cnt, s.DG[l], s.lout[l], act, s.lin[l],
                                               nullptr, nullptr, s.g + s.ow[l], s.lin[l], ksplit, s.part)))
            return rc;
        hipLaunchKernelGGL(k_colsum, dim3((unsigned)((s.lout[l] + 255) / 256), (unsigned)nchunk), dim3(256), 0, h->stream,
                           s.DG[l], cnt, s.lout[l], s.part);
        hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)((s.lout[l] + 255) / 256)), dim3(256), 0, h->stream, s.part, nchunk,
                           1, s.lout[l], s.g + s.ob[l], s.lout[l]);
        ESM_HIP_CHECK(hipGetLastError());
        if (l > 0)   // dact [cnt][lin] = DG[l] · Wt, masked by the ReLU of the activation
            if ((rc = gemm<false, false, EPI_MASK>(h->stream, cnt, s.lin[l], s.lout[l], s.DG[l], s.lout[l],
                                                    s.wt16 + s.ow[l], s.lin[l], nullptr, s.H[l - 1], s.DG[l - 1], s.lin[l])))
                return rc;
    }
    return 0;
}

template <typename S>
static int mlp_features(esm_handle *h, long long m0, const long long *idx, int cnt) {
    MlpState &s = *h->mlp;
    const DevModel<S> m = dev_model<S>(h);
    SweepState<S> st{(S *)h->d_v, (S *)h->d_r, (S *)h->d_F, (const S *)h->d_x};
    if (s.D.din != MLP_DIN || s.D.B > 24 || (uintptr_t)s.X % 16) {
        set_error("mlp_features: 160 inputs, at most 24 hours per block, 16-byte aligned rows");
        return ESM_EINVAL;
    }
    hipLaunchKernelGGL(k_mlp_features<S>, dim3((unsigned)((cnt + MLP_FT - 1) / MLP_FT)), dim3(256), 0, h->stream, m, st,
                       (const double *)h->d_k, s.D, m0, idx, cnt, s.X, s.Y, s.MU0);
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace esm

using namespace esm;

extern "C" {

int32_t esm_mlp_init(esm_handle *h, int32_t block_hours, int32_t ncomp, const uint32_t key[2]) {
    if (!h || !key) { set_error("esm_mlp_init: null argument"); return ESM_EINVAL; }
    if (block_hours < 1 || block_hours > 24 || block_hours > h->T || ncomp < 1 || ncomp > 8) {
        set_error("esm_mlp_init: block_hours in [1, min(24, T)], ncomp in [1, 8]");
        return ESM_EINVAL;
    }
    try {
        mlp_free(h);
        h->mlp = new MlpState();
        MlpState &s = *h->mlp;
        MlpDims &D = s.D;
        D.B = block_hours;
        D.K = ncomp;
        D.din = 160;
        D.dh = 256;
        D.dout = (3 * D.K * D.B + 63) / 64 * 64;
        D.nb = h->T / D.B;
        D.n_inst = (long long)h->N * D.nb * h->R;
        const int lin[4] = {D.din, D.dh, D.dh, D.dh}, lout[4] = {D.dh, D.dh, D.dh, D.dout};
        size_t o = 0;
        for (int l = 0; l < 4; l++) {
            s.lin[l] = lin[l];
            s.lout[l] = lout[l];
            s.ow[l] = o;
            o += (size_t)lin[l] * lout[l];
            s.ob[l] = o;
            o += lout[l];
        }
        s.np = o;
        // initial weights: uniform ±sqrt(6 / fan_in) (ReLU layers), ±0.01·sqrt(6 / fan_in) at the
        // output; biases 0 except the log standard deviations: log 0.2
        std::vector<float> p(s.np, 0.0f);
        for (int l = 0; l < 4; l++) {
            const double a = (l < 3 ? 1.0 : 0.01) * sqrt(6.0 / lin[l]);
            for (size_t i = 0; i < (size_t)lin[l] * lout[l]; i++) {
                uint32_t w[4];
                philox4x32(key[0], key[1], (uint32_t)i, (uint32_t)l, 0x5d000000u, 0u, w);
                p[s.ow[l] + i] = (float)(a * (2.0 * u01(w[0], w[1]) - 1.0));
            }
        }
        for (int q = 2 * D.K * D.B; q < 3 * D.K * D.B; q++) p[s.ob[3] + q] = (float)log(0.2);
        auto dalloc = [&](void **ptr, size_t bytes) -> int {
            ESM_HIP_CHECK(hipMalloc(ptr, bytes));
            s.allocs.push_back(*ptr);
            ESM_HIP_CHECK(hipMemsetAsync(*ptr, 0, bytes, h->stream));
            return 0;
        };
        int rc;
        const size_t c = (size_t)s.cap;
#define ALLOC(ptr, bytes) if ((rc = dalloc((void **)&(ptr), (bytes)))) return rc
        ALLOC(s.p, s.np * 4);
        ALLOC(s.g, s.np * 4);
        ALLOC(s.am, s.np * 4);
        ALLOC(s.av, s.np * 4);
        ALLOC(s.wt16, s.np * 2);
        ALLOC(s.X, c * D.din * 2);
        for (int l = 0; l < 3; l++) ALLOC(s.H[l], c * D.dh * 2);
        for (int l = 0; l < 3; l++) ALLOC(s.DG[l], c * D.dh * 2);
        ALLOC(s.DG[3], c * D.dout * 2);
        ALLOC(s.OUT, c * D.dout * 4);
        ALLOC(s.Y, c * D.B * 4);
        ALLOC(s.MU0, c * D.B * 4);
        ALLOC(s.YP, c * D.B * 4);
        ALLOC(s.LR, c * 8);
        ALLOC(s.dloss, (size_t)MLP_LOSS_CHUNK * 8);
        ALLOC(s.lpart, (c / MLP_GPB + 1) * 8);
        {
            size_t wmax = 0, omax = 0;
            for (int l = 0; l < 4; l++) {
                wmax = std::max(wmax, (size_t)lin[l] * lout[l]);
                omax = std::max(omax, (size_t)lout[l]);
            }
            s.part_len = std::max((size_t)KSPLIT_MAX * wmax, (c + COLSUM_ROWS - 1) / COLSUM_ROWS * omax);
        }
        ALLOC(s.part, s.part_len * 4);
        ALLOC(s.idx, c * 8);
        ALLOC(s.stats, (size_t)ESM_NSTAT * STAT_SLOTS * 8);
#undef ALLOC
        ESM_HIP_CHECK(hipMemcpyAsync(s.p, p.data(), s.np * 4, hipMemcpyHostToDevice, h->stream));
        if ((rc = mlp_refresh_bf16(h))) return rc;
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    } catch (const std::exception &e) {
        set_error(std::string("esm_mlp_init: ") + e.what());
        return ESM_ENOMEM;
    }
    return ESM_OK;
}

int32_t esm_mlp_dims(esm_handle *h, int32_t *dims) {
    if (!h || !dims || !h->mlp) { set_error("esm_mlp_dims: no network (esm_mlp_init)"); return ESM_EINVAL; }
    const MlpDims &D = h->mlp->D;
    dims[0] = D.B; dims[1] = D.K; dims[2] = D.din; dims[3] = D.dh; dims[4] = D.dout; dims[5] = D.nb;
    dims[6] = (int32_t)h->mlp->np;
    return ESM_OK;
}

int32_t esm_mlp_params(esm_handle *h, double *get, const double *set) {
    if (!h || !h->mlp) { set_error("esm_mlp_params: no network (esm_mlp_init)"); return ESM_EINVAL; }
    MlpState &s = *h->mlp;
    std::vector<float> p(s.np);
    if (set) {
        for (size_t i = 0; i < s.np; i++) p[i] = (float)set[i];
        ESM_HIP_CHECK(hipMemcpyAsync(s.p, p.data(), s.np * 4, hipMemcpyHostToDevice, h->stream));
        int rc;
        if ((rc = mlp_refresh_bf16(h))) return rc;
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    }
    if (get) {
        ESM_HIP_CHECK(hipMemcpyAsync(p.data(), s.p, s.np * 4, hipMemcpyDeviceToHost, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
        for (size_t i = 0; i < s.np; i++) get[i] = p[i];
    }
    return ESM_OK;
}

// the forward pass of the proposals and of esm_mlp_forward: fused (1, default) or layered (0)
int32_t esm_mlp_options(esm_handle *h, int32_t fused) {
    if (!h || !h->mlp) { set_error("esm_mlp_options: no network (esm_mlp_init)"); return ESM_EINVAL; }
    if (fused != 0 && fused != 1) { set_error("esm_mlp_options: fused is 0 or 1"); return ESM_EINVAL; }
    h->mlp->fused = fused == 1;
    return ESM_OK;
}

// Adam's state: [m (np)][v (np)][t], FP32 moments as doubles (exact both ways)
int32_t esm_mlp_optimizer(esm_handle *h, double *get, const double *set) {
    if (!h || !h->mlp) { set_error("esm_mlp_optimizer: no network (esm_mlp_init)"); return ESM_EINVAL; }
    MlpState &s = *h->mlp;
    std::vector<float> q(2 * s.np);
    if (set) {
        if (!(set[2 * s.np] >= 0) || set[2 * s.np] != nearbyint(set[2 * s.np])) {
            set_error("esm_mlp_optimizer: the step count must be a whole number >= 0");
            return ESM_EINVAL;
        }
        for (size_t i = 0; i < 2 * s.np; i++) q[i] = (float)set[i];
        ESM_HIP_CHECK(hipMemcpyAsync(s.am, q.data(), s.np * 4, hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipMemcpyAsync(s.av, q.data() + s.np, s.np * 4, hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
        s.adam_t = (long long)set[2 * s.np];
    }
    if (get) {
        ESM_HIP_CHECK(hipMemcpyAsync(q.data(), s.am, s.np * 4, hipMemcpyDeviceToHost, h->stream));
        ESM_HIP_CHECK(hipMemcpyAsync(q.data() + s.np, s.av, s.np * 4, hipMemcpyDeviceToHost, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
        for (size_t i = 0; i < 2 * s.np; i++) get[i] = q[i];
        get[2 * s.np] = (double)s.adam_t;
    }
    return ESM_OK;
}

// forward of M explicit input rows x [M][din] (rounded to BF16): out [M][dout]
int32_t esm_mlp_forward(esm_handle *h, const double *x, int64_t M, double *out) {
    if (!h || !h->mlp || !x || !out) { set_error("esm_mlp_forward: null argument or no network"); return ESM_EINVAL; }
    MlpState &s = *h->mlp;
    if (M < 1 || M > s.cap) { set_error("esm_mlp_forward: 1 <= M <= 65536"); return ESM_EINVAL; }
    const MlpDims &D = s.D;
    std::vector<uint16_t> xb((size_t)M * D.din);
    for (size_t i = 0; i < xb.size(); i++) xb[i] = f2bf((float)x[i]);
    ESM_HIP_CHECK(hipMemcpyAsync(s.X, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, h->stream));
    int rc;
    if ((rc = mlp_forward(h, (int)M))) return rc;
    std::vector<float> o((size_t)M * D.dout);
    ESM_HIP_CHECK(hipMemcpyAsync(o.data(), s.OUT, o.size() * 4, hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < o.size(); i++) out[i] = o[i];
    return ESM_OK;
}

// loss Σ −log q(y|c) of M explicit rows (x [M][din], y and mu0 [M][B]) and the gradient of its
// mean with respect to the master parameters (layout of esm_mlp_params)
int32_t esm_mlp_gradient(esm_handle *h, const double *x, const double *y, const double *mu0, int64_t M, double *grad,
                         double *loss) {
    if (!h || !h->mlp || !x || !y || !mu0) { set_error("esm_mlp_gradient: null argument or no network"); return ESM_EINVAL; }
    MlpState &s = *h->mlp;
    if (M < 1 || M > s.cap) { set_error("esm_mlp_gradient: 1 <= M <= 65536"); return ESM_EINVAL; }
    const MlpDims &D = s.D;
    std::vector<uint16_t> xb((size_t)M * D.din);
    for (size_t i = 0; i < xb.size(); i++) xb[i] = f2bf((float)x[i]);
    std::vector<float> yf((size_t)M * D.B), mf((size_t)M * D.B);
    for (size_t i = 0; i < yf.size(); i++) { yf[i] = (float)y[i]; mf[i] = (float)mu0[i]; }
    ESM_HIP_CHECK(hipMemcpyAsync(s.X, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, h->stream));
    ESM_HIP_CHECK(hipMemcpyAsync(s.Y, yf.data(), yf.size() * 4, hipMemcpyHostToDevice, h->stream));
    ESM_HIP_CHECK(hipMemcpyAsync(s.MU0, mf.data(), mf.size() * 4, hipMemcpyHostToDevice, h->stream));
    int rc;
    if ((rc = mlp_forward(h, (int)M, false, true))) return rc;
    ESM_HIP_CHECK(hipMemsetAsync(s.g, 0, s.np * 4, h->stream));
    ESM_HIP_CHECK(hipMemsetAsync(s.dloss, 0, 8, h->stream));
    const unsigned nbk = (unsigned)((M + MLP_GPB - 1) / MLP_GPB);
    hipLaunchKernelGGL(k_mlp_head_grad, dim3(nbk), dim3(256), 0, h->stream, s.OUT, s.MU0, s.Y, (int)M, D.K, D.B, D.dout,
                       1.0f / (float)M, s.DG[3], s.lpart);
    hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(64), 0, h->stream, s.lpart, (int)nbk, s.dloss);
    ESM_HIP_CHECK(hipGetLastError());
    if ((rc = mlp_backward(h, (int)M))) return rc;
    std::vector<float> g(s.np);
    double l = 0.0;
    ESM_HIP_CHECK(hipMemcpyAsync(g.data(), s.g, s.np * 4, hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipMemcpyAsync(&l, s.dloss, 8, hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (grad)
        for (size_t i = 0; i < s.np; i++) grad[i] = g[i];
    if (loss) *loss = l;
    return ESM_OK;
}

// `steps` Adam steps on minibatches of `batch` instances drawn from the current states of every
// replica (maximum likelihood of the blocks' y under q(·|c)); loss: the mean loss of each step
int32_t esm_mlp_train(esm_handle *h, int32_t steps, int32_t batch, double lr, const uint32_t key[2], uint32_t counter,
                      double *loss) {
    if (!h || !h->mlp || !key) { set_error("esm_mlp_train: null argument or no network"); return ESM_EINVAL; }
    MlpState &s = *h->mlp;
    if (steps < 0 || batch < 1 || batch > s.cap || !(lr > 0)) {
        set_error("esm_mlp_train: steps >= 0, 1 <= batch <= 65536, lr > 0");
        return ESM_EINVAL;
    }
    const MlpDims &D = s.D;
    int rc;
    for (int c0 = 0; c0 < steps; c0 += MLP_LOSS_CHUNK) {
        const int c1 = std::min(steps, c0 + MLP_LOSS_CHUNK);
        ESM_HIP_CHECK(hipMemsetAsync(s.dloss, 0, (size_t)(c1 - c0) * 8, h->stream));
        for (int it = c0; it < c1; it++) {
            ProfScope pt(h, ESM_KERN_MLPT);
            hipLaunchKernelGGL(k_batch_idx, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, h->stream, key[0], key[1],
                               counter, (uint32_t)it, D.n_inst, batch, s.idx);
            ESM_HIP_CHECK(hipGetLastError());
            rc = h->prec == ESM_FP64 ? mlp_features<double>(h, 0, s.idx, batch) : mlp_features<float>(h, 0, s.idx, batch);
            if (rc) return rc;
            if ((rc = mlp_forward(h, batch, false, true))) return rc;
            ESM_HIP_CHECK(hipMemsetAsync(s.g, 0, s.np * 4, h->stream));
            const unsigned nbk = (unsigned)((batch + MLP_GPB - 1) / MLP_GPB);
            hipLaunchKernelGGL(k_mlp_head_grad, dim3(nbk), dim3(256), 0, h->stream, s.OUT, s.MU0, s.Y, batch, D.K, D.B,
                               D.dout, 1.0f / (float)batch, s.DG[3], s.lpart);
            hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(64), 0, h->stream, s.lpart, (int)nbk, s.dloss + (it - c0));
            ESM_HIP_CHECK(hipGetLastError());
            if ((rc = mlp_backward(h, batch))) return rc;
            s.adam_t++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float k1 = 1.0f - powf(b1, (float)s.adam_t), k2 = 1.0f - powf(b2, (float)s.adam_t);
            hipLaunchKernelGGL(k_adam, dim3((unsigned)((s.np + 255) / 256)), dim3(256), 0, h->stream, s.p, s.g, s.am, s.av,
                               s.np, (float)lr, b1, b2, 1e-8f, k1, k2);
            ESM_HIP_CHECK(hipGetLastError());
            if ((rc = mlp_refresh_bf16(h))) return rc;
        }
        if (loss) {
            std::vector<double> l((size_t)(c1 - c0));
            ESM_HIP_CHECK(hipMemcpyAsync(l.data(), s.dloss, l.size() * 8, hipMemcpyDeviceToHost, h->stream));
            ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
            for (int it = c0; it < c1; it++) loss[it] = l[(size_t)(it - c0)] / batch;
        }
    }
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    return ESM_OK;
}

// one learned-proposal sweep: every (node, block, replica) once, in chunks; stats [ESM_NSTAT]
// (attempted, accepted, 0, 0, network evaluations); with out / yprop / logratio (all instances,
// [n_inst][dout], [n_inst][B], [n_inst]) the network outputs, the proposals and the MH
// log-ratios (for the checks)
int32_t esm_mlp_propose(esm_handle *h, const uint32_t key[2], uint32_t counter, int64_t *stats, double *out,
                        double *yprop, double *logratio) {
    if (!h || !h->mlp || !key) { set_error("esm_mlp_propose: null argument or no network"); return ESM_EINVAL; }
    MlpState &s = *h->mlp;
    const MlpDims &D = s.D;
    int rc;
    ESM_HIP_CHECK(hipMemsetAsync(s.stats, 0, (size_t)ESM_NSTAT * STAT_SLOTS * 8, h->stream));
    for (long long m0 = 0; m0 < D.n_inst; m0 += s.cap) {
        const int cnt = (int)std::min<long long>(s.cap, D.n_inst - m0);
        rc = h->prec == ESM_FP64 ? mlp_features<double>(h, m0, nullptr, cnt) : mlp_features<float>(h, m0, nullptr, cnt);
        if (rc) return rc;
        if ((rc = mlp_forward(h, cnt, true))) return rc;
        const dim3 grid((unsigned)((cnt + MLP_GPB - 1) / MLP_GPB)), blk(256);
        ProfScope pp(h, ESM_KERN_MLPP);
        if (h->prec == ESM_FP64) {
            const DevModel<double> m = dev_model<double>(h);
            SweepState<double> st{(double *)h->d_v, (double *)h->d_r, (double *)h->d_F, (const double *)h->d_x};
            hipLaunchKernelGGL(k_mlp_propose<double>, grid, blk, 0, h->stream, m, st, D, m0, cnt, s.OUT, s.MU0, key[0], key[1],
                               counter, h->m.beta_v, s.stats, s.YP, s.LR);
        } else {
            const DevModel<float> m = dev_model<float>(h);
            SweepState<float> st{(float *)h->d_v, (float *)h->d_r, (float *)h->d_F, (const float *)h->d_x};
            hipLaunchKernelGGL(k_mlp_propose<float>, grid, blk, 0, h->stream, m, st, D, m0, cnt, s.OUT, s.MU0, key[0], key[1],
                               counter, (float)h->m.beta_v, s.stats, s.YP, s.LR);
        }
        ESM_HIP_CHECK(hipGetLastError());
        if (out || yprop || logratio) {
            std::vector<float> o;
            if (out) {
                o.resize((size_t)cnt * D.dout);
                ESM_HIP_CHECK(hipMemcpyAsync(o.data(), s.OUT, o.size() * 4, hipMemcpyDeviceToHost, h->stream));
            }
            std::vector<float> yp;
            if (yprop) {
                yp.resize((size_t)cnt * D.B);
                ESM_HIP_CHECK(hipMemcpyAsync(yp.data(), s.YP, yp.size() * 4, hipMemcpyDeviceToHost, h->stream));
            }
            if (logratio) ESM_HIP_CHECK(hipMemcpyAsync(logratio + m0, s.LR, (size_t)cnt * 8, hipMemcpyDeviceToHost, h->stream));
            ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
            for (size_t i = 0; i < o.size(); i++) out[(size_t)m0 * D.dout + i] = o[i];
            for (size_t i = 0; i < yp.size(); i++) yprop[(size_t)m0 * D.B + i] = yp[i];
        }
    }
    std::vector<unsigned long long> st((size_t)ESM_NSTAT * STAT_SLOTS);
    ESM_HIP_CHECK(hipMemcpyAsync(st.data(), s.stats, st.size() * 8, hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (stats)
        for (int i = 0; i < ESM_NSTAT; i++) {
            unsigned long long acc = 0;
            for (int q = 0; q < STAT_SLOTS; q++) acc += st[(size_t)i * STAT_SLOTS + q];
            stats[i] = (int64_t)acc;
        }
    h->scores_valid = false;
    return ESM_OK;
}

}  // extern "C"
