// Learned slice proposal: a masked autoregressive network (MADE, Germain et
// al. 2015) over one Trotter slice, trained on slices the plain sweep emits,
// used as an independence Metropolis-Hastings proposal.
//
// Network (natural order): N -> H1 -> H2 -> N logits, ReLU hidden layers.
// Degrees: input d has degree d+1; hidden unit k has m(k) = 1 + k mod (N-1);
// masks  M1[k][d] = m1(k) >= d+1,  M2[k'][k] = m2(k') >= m1(k),
//        M3[d][k] = d+1 > m2(k)   so logit d depends on spins < d only.
// Logit l_d gives p(s_d = -1 | s_<d) = sigmoid(l_d); log q(s) = -sum_d
// [softplus(l_d) - b_d l_d] with b_d = (s_d == -1).
//
// Arithmetic: BF16 weights and activations, FP32 accumulation on MFMA
// (v_mfma_f32_16x16x32_bf16), FP32 master weights and Adam moments.  Every
// GEMM is C[M][N] = A[M][K] . B[N][K]^T with both operands K-contiguous.
//
// Exactness of the proposal density.  The sampler computes logit d with the
// same full forward pass (same kernels, same shapes, same summation order) on
// the partially sampled slice; the weights that connect logit d to spins >= d
// are exact zeros, so logit d is bit-identical to the one a forward pass over
// the finished slice produces.  log q(new) and log q(old) therefore come from
// one function q, the density the proposal was drawn from.
#include <math.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>

#include "piqmc_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ uint16_t f2bf_bits(float f)
{
    // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ------------------------------------------------------------------ GEMM
// C = A . B^T (+ bias[n]) (ReLU); A [M][K], B [N][K] bf16 (uint16 bits), row
// strides lda/ldb; outputs fp32 Cf [M][ldc] and/or bf16 Cb [M][ldc].
// Workgroup 256 threads = 4 wavefronts in 2x2; tile 64x64, K-step 32; each
// wavefront owns a 32x32 sub-tile = 2x2 MFMA 16x16x32 blocks.
constexpr int GT = 64, GK = 32, GKP = GK + 8;  // LDS row pitch (bf16) avoids bank conflicts

__global__ __launch_bounds__(256) void gemm_nt_bf16(const uint16_t *A, int lda, const uint16_t *B, int ldb,
                                                    int M, int N, int K, const float *bias, int relu,
                                                    float *Cf, uint16_t *Cb, int ldc)
{
    __shared__ __align__(16) uint16_t As[GT][GKP];
    __shared__ __align__(16) uint16_t Bs[GT][GKP];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    // each thread stages 8 consecutive bf16 of one row of each tile
    const int lr = tid >> 2, lk = (tid & 3) * 8;
    for (int k0 = 0; k0 < K; k0 += GK) {
        {
            const int gm = m0 + lr, gk = k0 + lk;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gm < M) {
                if (gk + 8 <= K && ((lda & 7) == 0)) {
                    v = *reinterpret_cast<const uint4 *>(A + (size_t)gm * lda + gk);
                } else {
                    uint16_t t[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) t[j] = (gk + j < K) ? A[(size_t)gm * lda + gk + j] : 0;
                    v = *reinterpret_cast<uint4 *>(t);
                }
            }
            *reinterpret_cast<uint4 *>(&As[lr][lk]) = v;
            const int gn = n0 + lr;
            uint4 w = make_uint4(0, 0, 0, 0);
            if (gn < N) {
                if (gk + 8 <= K && ((ldb & 7) == 0)) {
                    w = *reinterpret_cast<const uint4 *>(B + (size_t)gn * ldb + gk);
                } else {
                    uint16_t t[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) t[j] = (gk + j < K) ? B[(size_t)gn * ldb + gk + j] : 0;
                    w = *reinterpret_cast<uint4 *>(t);
                }
            }
            *reinterpret_cast<uint4 *>(&Bs[lr][lk]) = w;
        }
        __syncthreads();
        // fragments: lane holds row (lane&15), k = (lane>>4)*8 .. +7
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            uint4 va = *reinterpret_cast<const uint4 *>(&As[wm + t * 16 + (lane & 15)][(lane >> 4) * 8]);
            uint4 vb = *reinterpret_cast<const uint4 *>(&Bs[wn + t * 16 + (lane & 15)][(lane >> 4) * 8]);
            af[t] = __builtin_bit_cast(bf16x8, va);
            bfr[t] = __builtin_bit_cast(bf16x8, vb);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // accumulator layout: lane holds column (lane&15), rows (lane>>4)*4 + j
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = n0 + wn + b * 16 + (lane & 15);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = m0 + wm + a * 16 + (lane >> 4) * 4 + j;
                if (row < M && col < N) {
                    float v = acc[a][b][j];
                    if (bias) v += bias[col];
                    if (relu) v = v > 0.f ? v : 0.f;
                    if (Cf) Cf[(size_t)row * ldc + col] = v;
                    if (Cb) Cb[(size_t)row * ldc + col] = f2bf_bits(v);
                }
            }
        }
}

// Pipelined form for K, lda, ldb multiples of 64 / 8 / 8: K-stages of 64
// (two MFMA k-steps) through a double-buffered LDS tile, with the global loads
// of the next kNpf - 1 stages in flight in a register ring while the current
// stage's MFMAs run (one workgroup per 64 x 64 tile leaves a CU one wavefront
// per SIMD, so the stage loop itself has to hide the HBM latency), one barrier
// per stage.  Each thread moves 2 x 16 B of A and of B per stage; rows past
// M / N load row 0 and are zeroed when stored.
constexpr int PK = 64, PKP = PK + 8;  // stage depth; LDS row pitch (bf16)
constexpr int kNpf = 4;               // stages whose loads are issued ahead

__global__ __launch_bounds__(256) void gemm_nt_bf16_pipe(const uint16_t *A, int lda, const uint16_t *B, int ldb,
                                                         int M, int N, int K, const float *bias, int relu,
                                                         float *Cf, uint16_t *Cb, int ldc)
{
    __shared__ __align__(16) uint16_t As[2][GT][PKP];
    __shared__ __align__(16) uint16_t Bs[2][GT][PKP];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    // staging map: thread -> rows r and r+32, 8 consecutive k at column c
    const int sr = tid >> 3, sc = (tid & 7) * 8;
    const bool a0 = m0 + sr < M, a1 = m0 + sr + 32 < M, b0 = n0 + sr < N, b1 = n0 + sr + 32 < N;
    const uint16_t *pa0 = A + (size_t)(a0 ? m0 + sr : 0) * lda + sc;
    const uint16_t *pa1 = A + (size_t)(a1 ? m0 + sr + 32 : 0) * lda + sc;
    const uint16_t *pb0 = B + (size_t)(b0 ? n0 + sr : 0) * ldb + sc;
    const uint16_t *pb1 = B + (size_t)(b1 ? n0 + sr + 32 : 0) * ldb + sc;
    uint4 ra0[kNpf], ra1[kNpf], rb0[kNpf], rb1[kNpf];  // register ring: slot = stage % kNpf
    auto fetch = [&](int slot, int stage) __attribute__((always_inline)) {
        const int k0 = stage * PK;
        ra0[slot] = *reinterpret_cast<const uint4 *>(pa0 + k0);
        ra1[slot] = *reinterpret_cast<const uint4 *>(pa1 + k0);
        rb0[slot] = *reinterpret_cast<const uint4 *>(pb0 + k0);
        rb1[slot] = *reinterpret_cast<const uint4 *>(pb1 + k0);
    };
    const uint32_t ma0 = a0 ? ~0u : 0u, ma1 = a1 ? ~0u : 0u, mb0 = b0 ? ~0u : 0u, mb1 = b1 ? ~0u : 0u;
    auto masked = [](uint4 v, uint32_t m) { return make_uint4(v.x & m, v.y & m, v.z & m, v.w & m); };
    auto stash = [&](int slot, int buf) __attribute__((always_inline)) {
        *reinterpret_cast<uint4 *>(&As[buf][sr][sc]) = masked(ra0[slot], ma0);
        *reinterpret_cast<uint4 *>(&As[buf][sr + 32][sc]) = masked(ra1[slot], ma1);
        *reinterpret_cast<uint4 *>(&Bs[buf][sr][sc]) = masked(rb0[slot], mb0);
        *reinterpret_cast<uint4 *>(&Bs[buf][sr + 32][sc]) = masked(rb1[slot], mb1);
    };
    // Loads are issued unconditionally (stage indices past the end repeat the
    // last stage), so the count of loads in flight is the same at every wait.
    const int ns = K / PK;
#pragma unroll
    for (int j = 0; j < kNpf; ++j) fetch(j, min(j, ns - 1));
    stash(0, 0);
    __syncthreads();
    for (int s0 = 0; s0 < ns; s0 += kNpf) {
#pragma unroll
        for (int j = 0; j < kNpf; ++j) {  // stage s = s0 + j sits in LDS buffer s & 1
            const int s = s0 + j;
            if (s >= ns) break;
            const int buf = s & 1;
            fetch(j, min(s + kNpf, ns - 1));  // slot j is free: stage s is already in LDS
#pragma unroll
            for (int kk = 0; kk < PK; kk += 32) {
                bf16x8 af[2], bfr[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    uint4 va = *reinterpret_cast<const uint4 *>(&As[buf][wm + t * 16 + (lane & 15)][kk + (lane >> 4) * 8]);
                    uint4 vb = *reinterpret_cast<const uint4 *>(&Bs[buf][wn + t * 16 + (lane & 15)][kk + (lane >> 4) * 8]);
                    af[t] = __builtin_bit_cast(bf16x8, va);
                    bfr[t] = __builtin_bit_cast(bf16x8, vb);
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
            }
            stash((j + 1) % kNpf, buf ^ 1);  // after the last stage: unused
            __syncthreads();
        }
    }
    // accumulator layout: lane holds column (lane&15), rows (lane>>4)*4 + j
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = n0 + wn + b * 16 + (lane & 15);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = m0 + wm + a * 16 + (lane >> 4) * 4 + j;
                if (row < M && col < N) {
                    float v = acc[a][b][j];
                    if (bias) v += bias[col];
                    if (relu) v = v > 0.f ? v : 0.f;
                    if (Cf) Cf[(size_t)row * ldc + col] = v;
                    if (Cb) Cb[(size_t)row * ldc + col] = f2bf_bits(v);
                }
            }
        }
}

// Larger tiles: BM x BN per workgroup of four wavefronts (2 x 2), each
// wavefront (BM/2) x (BN/2) as (BM/32) x (BN/32) blocks of 16 x 16 MFMAs, so
// every fragment read from LDS feeds BN/32 or BM/32 MFMAs instead of two.
// K-stages of 64 through a double-buffered LDS tile; the next stage's global
// loads are in flight in registers while the current one's MFMAs run.  Rows
// past M / N read a valid row and are zeroed on the way to LDS.
template <int BM, int BN>
__global__ __launch_bounds__(256) void gemm_nt_bf16_tile(const uint16_t *A, int lda, const uint16_t *B, int ldb,
                                                         int M, int N, int K, const float *bias, int relu,
                                                         float *Cf, uint16_t *Cb, int ldc)
{
    constexpr int TM = BM / 32, TN = BN / 32;          // 16 x 16 blocks per wavefront
    constexpr int CA = BM * 8 / 256, CB = BN * 8 / 256;  // 16-byte chunks per thread per stage
    __shared__ __align__(16) uint16_t As[2][BM][PKP];
    __shared__ __align__(16) uint16_t Bs[2][BN][PKP];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wid >> 1) * (BM / 2), wn = (wid & 1) * (BN / 2);
    f32x4 acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t *pa[CA], *pb[CB];
    uint32_t ma[CA], mb[CB];
    int ra[CA], ca[CA], rb[CB], cb[CB];
#pragma unroll
    for (int u = 0; u < CA; ++u) {
        const int c = tid + 256 * u;
        ra[u] = c >> 3;
        ca[u] = (c & 7) * 8;
        const bool in = m0 + ra[u] < M;
        pa[u] = A + (size_t)(in ? m0 + ra[u] : 0) * lda + ca[u];
        ma[u] = in ? ~0u : 0u;
    }
#pragma unroll
    for (int u = 0; u < CB; ++u) {
        const int c = tid + 256 * u;
        rb[u] = c >> 3;
        cb[u] = (c & 7) * 8;
        const bool in = n0 + rb[u] < N;
        pb[u] = B + (size_t)(in ? n0 + rb[u] : 0) * ldb + cb[u];
        mb[u] = in ? ~0u : 0u;
    }
    auto masked = [](uint4 v, uint32_t m) { return make_uint4(v.x & m, v.y & m, v.z & m, v.w & m); };
    uint4 va[CA], vb[CB];
    auto fetch = [&](int stage) __attribute__((always_inline)) {
        const int k0 = stage * PK;
#pragma unroll
        for (int u = 0; u < CA; ++u) va[u] = *reinterpret_cast<const uint4 *>(pa[u] + k0);
#pragma unroll
        for (int u = 0; u < CB; ++u) vb[u] = *reinterpret_cast<const uint4 *>(pb[u] + k0);
    };
    auto stash = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < CA; ++u) *reinterpret_cast<uint4 *>(&As[buf][ra[u]][ca[u]]) = masked(va[u], ma[u]);
#pragma unroll
        for (int u = 0; u < CB; ++u) *reinterpret_cast<uint4 *>(&Bs[buf][rb[u]][cb[u]]) = masked(vb[u], mb[u]);
    };
    const int ns = K / PK;
    fetch(0);
    stash(0);
    if (ns > 1) fetch(1);
    __syncthreads();
    for (int s = 0; s < ns; ++s) {
        const int buf = s & 1;
#pragma unroll
        for (int kk = 0; kk < PK; kk += 32) {
            bf16x8 af[TM], bfr[TN];
#pragma unroll
            for (int t = 0; t < TM; ++t)
                af[t] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4 *>(&As[buf][wm + t * 16 + (lane & 15)][kk + (lane >> 4) * 8]));
#pragma unroll
            for (int t = 0; t < TN; ++t)
                bfr[t] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4 *>(&Bs[buf][wn + t * 16 + (lane & 15)][kk + (lane >> 4) * 8]));
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
        if (s + 1 < ns) {
            stash(buf ^ 1);                 // stage s + 1, loaded during stage s
            if (s + 2 < ns) fetch(s + 2);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            const int col = n0 + wn + b * 16 + (lane & 15);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = m0 + wm + a * 16 + (lane >> 4) * 4 + j;
                if (row < M && col < N) {
                    float v = acc[a][b][j];
                    if (bias) v += bias[col];
                    if (relu) v = v > 0.f ? v : 0.f;
                    if (Cf) Cf[(size_t)row * ldc + col] = v;
                    if (Cb) Cb[(size_t)row * ldc + col] = f2bf_bits(v);
                }
            }
        }
}

// K split inside the workgroup: KG groups of four wavefronts share one 64 x 64
// output tile, group g taking the K-stages s = g, g + KG, g + 2 KG, ...  Each
// group has its own double-buffered stage in LDS and a two-stage register
// ring of global loads, so a SIMD holds KG wavefronts of the same tile and
// one's LDS reads and barriers hide behind the others' MFMAs.  At the end the
// groups' partial tiles meet in LDS and are summed in group order (the result
// does not depend on timing).  Loads past the last stage repeat it, so every
// group issues the same loads; stages past the end are not computed.
template <int KG>
__global__ __launch_bounds__(256 * KG) void gemm_nt_bf16_splitk(const uint16_t *A, int lda, const uint16_t *B,
                                                              int ldb, int M, int N, int K, const float *bias,
                                                              int relu, float *Cf, uint16_t *Cb, int ldc)
{
    constexpr int kBuf = GT * PKP;                        // bf16 per stage buffer of A or of B
    constexpr int kRedPitch = GT + 4;                     // floats per row of a partial tile
    constexpr size_t kStageBytes = (size_t)KG * 4 * kBuf * 2;  // per group: A x 2 buffers, B x 2 buffers
    constexpr size_t kRedBytes = (size_t)KG * GT * kRedPitch * 4;
    __shared__ __align__(16) unsigned char smem[kStageBytes > kRedBytes ? kStageBytes : kRedBytes];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int grp = wid >> 2, gt = tid & 255, gw = wid & 3;
    uint16_t *As = reinterpret_cast<uint16_t *>(smem) + (size_t)grp * 4 * kBuf;  // [buf][row][k]
    uint16_t *Bs = As + 2 * kBuf;
    auto a_at = [&](int buf, int r, int k) -> uint16_t * { return As + buf * kBuf + r * PKP + k; };
    auto b_at = [&](int buf, int r, int k) -> uint16_t * { return Bs + buf * kBuf + r * PKP + k; };
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (gw >> 1) * 32, wn = (gw & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int sr = gt >> 3, sc = (gt & 7) * 8;
    const bool a0 = m0 + sr < M, a1 = m0 + sr + 32 < M, b0 = n0 + sr < N, b1 = n0 + sr + 32 < N;
    const uint16_t *pa0 = A + (size_t)(a0 ? m0 + sr : 0) * lda + sc;
    const uint16_t *pa1 = A + (size_t)(a1 ? m0 + sr + 32 : 0) * lda + sc;
    const uint16_t *pb0 = B + (size_t)(b0 ? n0 + sr : 0) * ldb + sc;
    const uint16_t *pb1 = B + (size_t)(b1 ? n0 + sr + 32 : 0) * ldb + sc;
    const uint32_t ma0 = a0 ? ~0u : 0u, ma1 = a1 ? ~0u : 0u, mb0 = b0 ? ~0u : 0u, mb1 = b1 ? ~0u : 0u;
    auto masked = [](uint4 v, uint32_t m) { return make_uint4(v.x & m, v.y & m, v.z & m, v.w & m); };
    const int ns = K / PK, rounds = (ns + KG - 1) / KG;
    uint4 ra0[2], ra1[2], rb0[2], rb1[2];
    auto fetch = [&](int slot, int round) __attribute__((always_inline)) {
        const int k0 = min(round * KG + grp, ns - 1) * PK;
        ra0[slot] = *reinterpret_cast<const uint4 *>(pa0 + k0);
        ra1[slot] = *reinterpret_cast<const uint4 *>(pa1 + k0);
        rb0[slot] = *reinterpret_cast<const uint4 *>(pb0 + k0);
        rb1[slot] = *reinterpret_cast<const uint4 *>(pb1 + k0);
    };
    auto stash = [&](int slot, int buf) __attribute__((always_inline)) {
        *reinterpret_cast<uint4 *>(a_at(buf, sr, sc)) = masked(ra0[slot], ma0);
        *reinterpret_cast<uint4 *>(a_at(buf, sr + 32, sc)) = masked(ra1[slot], ma1);
        *reinterpret_cast<uint4 *>(b_at(buf, sr, sc)) = masked(rb0[slot], mb0);
        *reinterpret_cast<uint4 *>(b_at(buf, sr + 32, sc)) = masked(rb1[slot], mb1);
    };
    fetch(0, 0);
    fetch(1, 1);
    stash(0, 0);
    __syncthreads();
    // round r uses register slot r & 1 and LDS buffer r & 1 (slots indexed
    // by constants, so the ring stays in registers)
    auto step = [&](int r, const int slot) __attribute__((always_inline)) {
        const int buf = slot;
        fetch(slot, r + 2);                               // slot held round r, already in LDS
        if (r * KG + grp < ns) {
#pragma unroll
            for (int kk = 0; kk < PK; kk += 32) {
                bf16x8 af[2], bfr[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    af[t] = __builtin_bit_cast(
                        bf16x8, *reinterpret_cast<const uint4 *>(a_at(buf, wm + t * 16 + (lane & 15), kk + (lane >> 4) * 8)));
                    bfr[t] = __builtin_bit_cast(
                        bf16x8, *reinterpret_cast<const uint4 *>(b_at(buf, wn + t * 16 + (lane & 15), kk + (lane >> 4) * 8)));
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
            }
        }
        stash(slot ^ 1, buf ^ 1);                         // round r + 1 (unused after the last)
        __syncthreads();
    };
    for (int r = 0; r < rounds; r += 2) {
        step(r, 0);
        if (r + 1 < rounds) step(r + 1, 1);
    }
    float *red = reinterpret_cast<float *>(smem);         // [KG][GT][kRedPitch]; stages are done
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                red[((size_t)grp * GT + wm + a * 16 + (lane >> 4) * 4 + j) * kRedPitch + wn + b * 16 + (lane & 15)] =
                    acc[a][b][j];
    __syncthreads();
    for (int idx = tid; idx < GT * GT; idx += 256 * KG) {
        const int rl = idx / GT, cl = idx % GT, row = m0 + rl, col = n0 + cl;
        if (row < M && col < N) {
            float v = 0.f;
#pragma unroll
            for (int g = 0; g < KG; ++g) v += red[((size_t)g * GT + rl) * kRedPitch + cl];
            if (bias) v += bias[col];
            if (relu) v = v > 0.f ? v : 0.f;
            if (Cf) Cf[(size_t)row * ldc + col] = v;
            if (Cb) Cb[(size_t)row * ldc + col] = f2bf_bits(v);
        }
    }
}

// Tile of the pipelined GEMM: 0 automatic, 1 64x64, 2 128x64, 3 128x128,
// 4 64x64 with K in two groups, 5 64x64 with K in four groups
// (PIQMC_GEMM_TILE forces one, for comparisons).
int gemm_tile_choice(int M, int N, int K)
{
    const char *e = std::getenv("PIQMC_GEMM_TILE");
    if (e && e[0] >= '1' && e[0] <= '5') return e[0] - '0';
    // Measured on MI355X: the K split wins while 64 x 64 tiles leave CUs
    // idle; the wide tiles win once they alone give each CU two of them.
    const long long t64 = (long long)((M + 63) / 64) * ((N + 63) / 64);
    const long long t128 = (long long)((M + 127) / 128) * ((N + 127) / 128);
    const long long t12864 = (long long)((M + 127) / 128) * ((N + 63) / 64);
    const int ns = K / PK;
    if (t64 <= 256 && ns >= 8) return 5;
    if (t64 <= 512 && ns >= 4) return 4;
    if (t128 >= 512) return 3;
    if (t12864 >= 512) return 2;
    return 1;
}

void gemm(const uint16_t *A, int lda, const uint16_t *B, int ldb, int M, int N, int K, const float *bias,
          int relu, float *Cf, uint16_t *Cb, int ldc, hipStream_t st)
{
    if (K % PK == 0 && lda % 8 == 0 && ldb % 8 == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0) {
        const int tile = gemm_tile_choice(M, N, K);
        if (tile == 4 || tile == 5) {
            dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
            if (tile == 5)
                gemm_nt_bf16_splitk<4><<<grid, 1024, 0, st>>>(A, lda, B, ldb, M, N, K, bias, relu, Cf, Cb, ldc);
            else
                gemm_nt_bf16_splitk<2><<<grid, 512, 0, st>>>(A, lda, B, ldb, M, N, K, bias, relu, Cf, Cb, ldc);
            return;
        }
        if (tile == 3) {
            dim3 grid((N + 127) / 128, (M + 127) / 128);
            gemm_nt_bf16_tile<128, 128><<<grid, 256, 0, st>>>(A, lda, B, ldb, M, N, K, bias, relu, Cf, Cb, ldc);
            return;
        }
        if (tile == 2) {
            dim3 grid((N + 63) / 64, (M + 127) / 128);
            gemm_nt_bf16_tile<128, 64><<<grid, 256, 0, st>>>(A, lda, B, ldb, M, N, K, bias, relu, Cf, Cb, ldc);
            return;
        }
        dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
        gemm_nt_bf16_pipe<<<grid, 256, 0, st>>>(A, lda, B, ldb, M, N, K, bias, relu, Cf, Cb, ldc);
        return;
    }
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
    gemm_nt_bf16<<<grid, 256, 0, st>>>(A, lda, B, ldb, M, N, K, bias, relu, Cf, Cb, ldc);
}

// ------------------------------------------------------------------ elementwise
__global__ void spins_to_bf16(const int8_t *s, size_t n, uint16_t *x)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        x[t] = s[t] < 0 ? 0xBF80u : 0x3F80u;  // -1.0 / +1.0 in bf16
}

// transpose bf16 [R][C] -> [C][R]
__global__ void transpose_bf16(const uint16_t *in, int R, int C, uint16_t *out)
{
    __shared__ uint16_t tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int y = threadIdx.y; y < 32; y += blockDim.y) {
        const int r = by + y, c = bx + threadIdx.x;
        tile[y][threadIdx.x] = (r < R && c < C) ? in[(size_t)r * C + c] : 0;
    }
    __syncthreads();
    for (int y = threadIdx.y; y < 32; y += blockDim.y) {
        const int c = bx + y, r = by + threadIdx.x;
        if (c < C && r < R) out[(size_t)c * R + r] = tile[threadIdx.x][y];
    }
}

// Output layer loss: per sample nll = sum_d softplus(l_d) - b_d l_d; grad
// dL/dl (scaled by gscale) to bf16 dlog [B][N]; per-sample nll to out.
__global__ void output_nll(const float *logits, const int8_t *s, int B, int N, float gscale, uint16_t *dlog,
                           double *nll, double sign)
{
    const int b = blockIdx.x;
    double acc = 0.0;
    for (int d = threadIdx.x; d < N; d += blockDim.x) {
        const float l = logits[(size_t)b * N + d];
        const float bit = s[(size_t)b * N + d] < 0 ? 1.f : 0.f;
        const double sp = l > 0 ? (double)l + log1p(exp(-(double)l)) : log1p(exp((double)l));
        acc += sp - (double)bit * (double)l;
        if (dlog) {
            const float sg = 1.f / (1.f + expf(-l));
            dlog[(size_t)b * N + d] = f2bf_bits((sg - bit) * gscale);
        }
    }
    __shared__ double red[16];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) t += red[w];
        nll[b] = sign * t;
    }
}

__global__ void fill_i8(int8_t *p, size_t n, int8_t v)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        p[t] = v;
}

// dH (bf16, in place) *= (H > 0): ReLU backward against the stored activation.
__global__ void relu_back(uint16_t *dh, const uint16_t *h, size_t n)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        if (!(bf2f(h[t]) > 0.f)) dh[t] = 0;
}

// column sums of bf16 [B][C] -> fp32 [C]
__global__ void col_sum_bf16(const uint16_t *g, int B, int C, float *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += bf2f(g[(size_t)b * C + c]);
    out[c] = s;
}

// Adam step on fp32 master weights (masked), refresh bf16 W [O][I] and W^T [I][O].
__global__ void adam_w(float *w, float *m, float *v, const float *g, const uint8_t *mask, int O, int I,
                       float lr, float b1, float b2, float eps, float c1, float c2, uint16_t *wb, uint16_t *wt)
{
    const size_t n = (size_t)O * I;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        float wv = 0.f;
        if (mask[t]) {
            const float gr = g[t];
            const float mv = b1 * m[t] + (1.f - b1) * gr;
            const float vv = b2 * v[t] + (1.f - b2) * gr * gr;
            m[t] = mv;
            v[t] = vv;
            wv = w[t] - lr * (mv / c1) / (sqrtf(vv / c2) + eps);
        }
        w[t] = wv;
        const uint16_t bb = f2bf_bits(wv);
        wb[t] = bb;
        const int o = (int)(t / I), i = (int)(t % I);
        wt[(size_t)i * O + o] = bb;
    }
}

__global__ void adam_b(float *b, float *m, float *v, const float *g, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gr = g[t];
    const float mv = b1 * m[t] + (1.f - b1) * gr;
    const float vv = b2 * v[t] + (1.f - b2) * gr * gr;
    m[t] = mv;
    v[t] = vv;
    b[t] -= lr * (mv / c1) / (sqrtf(vv / c2) + eps);
}

// Sampling step d: draw s_d for every row from sigmoid(logit d).  The key and
// round come from device memory (rng = {key0, key1, round}) so that the
// captured sampling graph stays valid across calls.
__global__ void sample_step(const float *logits, int B, int N, int d, const uint32_t *rng, int8_t *s, uint16_t *x)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float l = logits[(size_t)b * N + d];
    uint32_t out[4];
    philox4x32_10(rng[2], (uint32_t)b, (uint32_t)d, 0x5A4D0000u, rng[0], rng[1], out);
    const double u = ((double)out[0] + 0.5) * 2.3283064365386963e-10;
    const double p = 1.0 / (1.0 + exp(-(double)l));  // p(s_d = -1)
    const bool neg = u < p;
    s[(size_t)b * N + d] = neg ? -1 : 1;
    x[(size_t)b * N + d] = neg ? 0xBF80u : 0x3F80u;
}

}  // namespace

#include "made_state.h"

namespace {
int made_alloc_batch(piqmc_made *m, int B)
{
    if (B <= m->cap) return PIQMC_OK;
    const int N = m->N, H1 = m->H1, H2 = m->H2;
    void **bufs[] = {(void **)&m->x, (void **)&m->h1, (void **)&m->h2, (void **)&m->logits, (void **)&m->dlog,
                     (void **)&m->dh2, (void **)&m->dh1, (void **)&m->tA, (void **)&m->tB, (void **)&m->nll,
                     (void **)&m->s};
    for (auto p : bufs)
        if (*p) (void)hipFree(*p), *p = nullptr;
    if (m->gexec) (void)hipGraphExecDestroy(m->gexec), m->gexec = nullptr, m->gB = 0;  // holds old addresses
    const int Hm = std::max(std::max(H1, H2), N);
    PIQMC_HIP(hipMalloc(&m->x, 2 * (size_t)B * N));
    PIQMC_HIP(hipMalloc(&m->h1, 2 * (size_t)B * H1));
    PIQMC_HIP(hipMalloc(&m->h2, 2 * (size_t)B * H2));
    PIQMC_HIP(hipMalloc(&m->logits, 4 * (size_t)B * N));
    PIQMC_HIP(hipMalloc(&m->dlog, 2 * (size_t)B * N));
    PIQMC_HIP(hipMalloc(&m->dh2, 2 * (size_t)B * H2));
    PIQMC_HIP(hipMalloc(&m->dh1, 2 * (size_t)B * H1));
    PIQMC_HIP(hipMalloc(&m->tA, 2 * (size_t)B * Hm));
    PIQMC_HIP(hipMalloc(&m->tB, 2 * (size_t)B * Hm));
    PIQMC_HIP(hipMalloc(&m->nll, 8 * (size_t)B));
    PIQMC_HIP(hipMalloc(&m->s, (size_t)B * N));
    m->cap = B;
    return PIQMC_OK;
}

int grid1(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 65535); }

void transpose(const uint16_t *in, int R, int C, uint16_t *out, hipStream_t st)
{
    dim3 g((C + 31) / 32, (R + 31) / 32), blk(32, 8);
    transpose_bf16<<<g, blk, 0, st>>>(in, R, C, out);
}

// forward over m->x (bf16 spins) rows [0, B): h1, h2 (bf16), logits (fp32)
void forward(piqmc_made *m, int B, hipStream_t st)
{
    gemm(m->x, m->N, m->wb[0], m->N, B, m->H1, m->N, m->b[0], 1, nullptr, m->h1, m->H1, st);
    gemm(m->h1, m->H1, m->wb[1], m->H1, B, m->H2, m->H1, m->b[1], 1, nullptr, m->h2, m->H2, st);
    gemm(m->h2, m->H2, m->wb[2], m->H2, B, m->N, m->H2, m->b[2], 0, m->logits, nullptr, m->N, st);
}
}  // namespace

extern "C" int piqmc_made_create(int N, int H1, int H2, uint64_t seed, piqmc_made **out)
{
    if (!out || N < 2 || H1 < 1 || H2 < 1) return piqmc::set_error(PIQMC_EINVAL, "made_create: bad sizes");
    auto *m = new piqmc_made();
    m->N = N; m->H1 = H1; m->H2 = H2;
    const int ins[3] = {N, H1, H2}, outs[3] = {H1, H2, N};
    // degrees and masks
    std::vector<int> m1(H1), m2(H2);
    for (int k = 0; k < H1; ++k) m1[k] = 1 + k % (N - 1);
    for (int k = 0; k < H2; ++k) m2[k] = 1 + k % (N - 1);
    // deterministic init: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from a 64-bit LCG
    uint64_t rs = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        rs = rs * 6364136223846793005ull + 1442695040888963407ull;
        return ((rs >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
    };
    for (int l = 0; l < 3; ++l) {
        const int O = outs[l], I = ins[l];
        m->layers_in[l] = I; m->layers_out[l] = O;
        std::vector<uint8_t> mk((size_t)O * I);
        for (int o = 0; o < O; ++o)
            for (int i = 0; i < I; ++i) {
                bool on;
                if (l == 0) on = m1[o] >= i + 1;
                else if (l == 1) on = m2[o] >= m1[i];
                else on = o + 1 > m2[i];
                mk[(size_t)o * I + i] = on;
            }
        std::vector<float> wv((size_t)O * I);
        const double sc = 1.0 / sqrt((double)I);
        for (size_t t = 0; t < wv.size(); ++t) wv[t] = mk[t] ? (float)(uni() * sc) : 0.f;
        std::vector<float> zero_o(O, 0.f);
        size_t nw = (size_t)O * I;
        hipError_t e = hipSuccess;
        auto A = [&](void **p, size_t bytes) { if (e == hipSuccess) e = hipMalloc(p, bytes); };
        A((void **)&m->w[l], 4 * nw); A((void **)&m->mw[l], 4 * nw); A((void **)&m->vw[l], 4 * nw);
        A((void **)&m->gw[l], 4 * nw); A((void **)&m->wb[l], 2 * nw); A((void **)&m->wt[l], 2 * nw);
        A((void **)&m->mask[l], nw); A((void **)&m->b[l], 4 * O); A((void **)&m->mb[l], 4 * O);
        A((void **)&m->vb[l], 4 * O); A((void **)&m->gb[l], 4 * O);
        if (e == hipSuccess) e = hipMemcpy(m->w[l], wv.data(), 4 * nw, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(m->mask[l], mk.data(), nw, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(m->mw[l], 0, 4 * nw);
        if (e == hipSuccess) e = hipMemset(m->vw[l], 0, 4 * nw);
        if (e == hipSuccess) e = hipMemset(m->gw[l], 0, 4 * nw);
        if (e == hipSuccess) e = hipMemset(m->b[l], 0, 4 * O);
        if (e == hipSuccess) e = hipMemset(m->mb[l], 0, 4 * O);
        if (e == hipSuccess) e = hipMemset(m->vb[l], 0, 4 * O);
        if (e != hipSuccess) {
            piqmc_made_destroy(m);
            return piqmc::hip_fail(e, "made_create");
        }
        // bf16 copies: an Adam step with zero gradient and zero lr refreshes them
        adam_w<<<grid1(nw), 256>>>(m->w[l], m->mw[l], m->vw[l], m->gw[l], m->mask[l], O, I, 0.f, 0.9f, 0.999f,
                                   1e-8f, 1.f, 1.f, m->wb[l], m->wt[l]);
    }
    PIQMC_HIP(hipDeviceSynchronize());
    *out = m;
    return PIQMC_OK;
}

extern "C" int piqmc_made_destroy(piqmc_made *m)
{
    if (!m) return PIQMC_OK;
    for (int l = 0; l < 3; ++l) {
        void *ps[] = {m->w[l], m->b[l], m->mw[l], m->vw[l], m->mb[l], m->vb[l], m->wb[l], m->wt[l], m->mask[l],
                      m->gw[l], m->gb[l]};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
    void *bs[] = {m->x, m->h1, m->h2, m->logits, m->dlog, m->dh2, m->dh1, m->tA, m->tB, m->nll, m->s, m->rng,
                  m->epoch_rows, m->nll_acc};
    for (void *p : bs)
        if (p) (void)hipFree(p);
    if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
    if (m->ev_in) (void)hipEventDestroy(m->ev_in);
    if (m->ev_out) (void)hipEventDestroy(m->ev_out);
    if (m->gs) (void)hipStreamDestroy(m->gs);
    piqmc::made_ar_release(m);
    delete m;
    return PIQMC_OK;
}

// Parameters as fp32 host arrays: w0 [H1][N], b0 [H1], w1 [H2][H1], b1 [H2], w2 [N][H2], b2 [N].
extern "C" int piqmc_made_get_params(const piqmc_made *m, float *w0, float *b0, float *w1, float *b1, float *w2,
                                     float *b2)
{
    if (!m) return piqmc::set_error(PIQMC_EINVAL, "null made");
    float *ws[3] = {w0, w1, w2}, *bs[3] = {b0, b1, b2};
    for (int l = 0; l < 3; ++l) {
        const size_t nw = (size_t)m->layers_out[l] * m->layers_in[l];
        if (ws[l]) PIQMC_HIP(hipMemcpy(ws[l], m->w[l], 4 * nw, hipMemcpyDeviceToHost));
        if (bs[l]) PIQMC_HIP(hipMemcpy(bs[l], m->b[l], 4 * (size_t)m->layers_out[l], hipMemcpyDeviceToHost));
    }
    return PIQMC_OK;
}

extern "C" int piqmc_made_get_grads(const piqmc_made *m, float *w0, float *b0, float *w1, float *b1, float *w2,
                                    float *b2)
{
    if (!m) return piqmc::set_error(PIQMC_EINVAL, "null made");
    float *ws[3] = {w0, w1, w2}, *bs[3] = {b0, b1, b2};
    for (int l = 0; l < 3; ++l) {
        const size_t nw = (size_t)m->layers_out[l] * m->layers_in[l];
        if (ws[l]) PIQMC_HIP(hipMemcpy(ws[l], m->gw[l], 4 * nw, hipMemcpyDeviceToHost));
        if (bs[l]) PIQMC_HIP(hipMemcpy(bs[l], m->gb[l], 4 * (size_t)m->layers_out[l], hipMemcpyDeviceToHost));
    }
    return PIQMC_OK;
}

// log q of B slices (int8 [B][N], device) -> logq_dev [B] (float64).
extern "C" int piqmc_made_logprob(piqmc_made *m, int B, const int8_t *s_dev, double *logq_dev, void *stream)
{
    if (!m || B < 0 || (B && (!s_dev || !logq_dev))) return piqmc::set_error(PIQMC_EINVAL, "made_logprob: bad args");
    if (!B) return PIQMC_OK;
    int rc = made_alloc_batch(m, B);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    spins_to_bf16<<<grid1((size_t)B * m->N), 256, 0, st>>>(s_dev, (size_t)B * m->N, m->x);
    forward(m, B, st);
    output_nll<<<B, 256, 0, st>>>(m->logits, s_dev, B, m->N, 0.f, nullptr, m->nll, -1.0);
    PIQMC_HIP(hipMemcpyAsync(logq_dev, m->nll, 8 * (size_t)B, hipMemcpyDeviceToDevice, st));
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

// Gradients of the mean NLL over B slices into the gradient buffers; returns
// the mean NLL.  With lr > 0 also applies one Adam step.
namespace {
// One training step on the stream, no host synchronisation: the gradients
// of the mean NLL of B slices, and an Adam step when lr > 0.  Per-row NLL in
// m->nll.
int train_core(piqmc_made *m, int B, const int8_t *s_dev, float lr, hipStream_t st)
{
    const int N = m->N, H1 = m->H1, H2 = m->H2;
    spins_to_bf16<<<grid1((size_t)B * N), 256, 0, st>>>(s_dev, (size_t)B * N, m->x);
    forward(m, B, st);
    output_nll<<<B, 256, 0, st>>>(m->logits, s_dev, B, N, 1.f / (float)B, m->dlog, m->nll, 1.0);
    // layer 3: dW2 = dlog^T h2, db2 = colsum(dlog), dh2 = dlog W2
    transpose(m->dlog, B, N, m->tA, st);   // [N][B]
    transpose(m->h2, B, H2, m->tB, st);    // [H2][B]
    gemm(m->tA, B, m->tB, B, N, H2, B, nullptr, 0, m->gw[2], nullptr, H2, st);
    col_sum_bf16<<<(N + 255) / 256, 256, 0, st>>>(m->dlog, B, N, m->gb[2]);
    gemm(m->dlog, N, m->wt[2], N, B, H2, N, nullptr, 0, nullptr, m->dh2, H2, st);
    relu_back<<<grid1((size_t)B * H2), 256, 0, st>>>(m->dh2, m->h2, (size_t)B * H2);
    // layer 2
    transpose(m->dh2, B, H2, m->tA, st);   // [H2][B]
    transpose(m->h1, B, H1, m->tB, st);    // [H1][B]
    gemm(m->tA, B, m->tB, B, H2, H1, B, nullptr, 0, m->gw[1], nullptr, H1, st);
    col_sum_bf16<<<(H2 + 255) / 256, 256, 0, st>>>(m->dh2, B, H2, m->gb[1]);
    gemm(m->dh2, H2, m->wt[1], H2, B, H1, H2, nullptr, 0, nullptr, m->dh1, H1, st);
    relu_back<<<grid1((size_t)B * H1), 256, 0, st>>>(m->dh1, m->h1, (size_t)B * H1);
    // layer 1
    transpose(m->dh1, B, H1, m->tA, st);   // [H1][B]
    transpose(m->x, B, N, m->tB, st);      // [N][B]
    gemm(m->tA, B, m->tB, B, H1, N, B, nullptr, 0, m->gw[0], nullptr, N, st);
    col_sum_bf16<<<(H1 + 255) / 256, 256, 0, st>>>(m->dh1, B, H1, m->gb[0]);
    if (lr > 0.f) {
        m->step++;
        m->ar_dirty = 1;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
        for (int l = 0; l < 3; ++l) {
            const int O = m->layers_out[l], I = m->layers_in[l];
            adam_w<<<grid1((size_t)O * I), 256, 0, st>>>(m->w[l], m->mw[l], m->vw[l], m->gw[l], m->mask[l], O, I, lr,
                                                         b1, b2, eps, c1, c2, m->wb[l], m->wt[l]);
            adam_b<<<(O + 255) / 256, 256, 0, st>>>(m->b[l], m->mb[l], m->vb[l], m->gb[l], O, lr, b1, b2, eps, c1, c2);
        }
    }
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

// rows data[perm[b0 + b]] -> out [B][N]
__global__ void gather_rows(const int8_t *data, const int32_t *perm, int b0, int B, int N, int8_t *out)
{
    const size_t tot = (size_t)B * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(t / N), i = (int)(t % N);
        out[t] = data[(size_t)perm[b0 + b] * N + i];
    }
}

// acc[0] += sum of v[0..n) (one workgroup; launches on one stream run in order)
__global__ void accumulate_sum(const double *v, int n, double *acc)
{
    __shared__ double red[16];
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) t += v[i];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) s += red[w];
        acc[0] += s;
    }
}

__global__ void zero_double(double *p) { *p = 0.0; }
}  // namespace

extern "C" int piqmc_made_train_step(piqmc_made *m, int B, const int8_t *s_dev, float lr, double *mean_nll,
                                     void *stream)
{
    if (!m || B < 1 || !s_dev) return piqmc::set_error(PIQMC_EINVAL, "made_train_step: bad args");
    int rc = made_alloc_batch(m, B);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    rc = train_core(m, B, s_dev, lr, st);
    if (rc) return rc;
    if (mean_nll) {
        std::vector<double> h(B);
        PIQMC_HIP(hipMemcpyAsync(h.data(), m->nll, 8 * (size_t)B, hipMemcpyDeviceToHost, st));
        PIQMC_HIP(hipStreamSynchronize(st));
        double t = 0;
        for (double v : h) t += v;
        *mean_nll = t / B;
    }
    return PIQMC_OK;
}

// One epoch of Adam over n slices data_dev int8 [n][N] in the order perm_dev
// int32 [n]: floor(n / batch) steps of `batch` rows, launched back to back
// with no host synchronisation until the epoch's mean NLL is read at the end.
extern "C" int piqmc_made_train_epoch(piqmc_made *m, int n, const int8_t *data_dev, const int32_t *perm_dev,
                                      int batch, float lr, double *mean_nll, void *stream)
{
    if (!m || n < 1 || batch < 1 || batch > n || !data_dev || !perm_dev)
        return piqmc::set_error(PIQMC_EINVAL, "made_train_epoch: bad args");
    // the batch rows live past the sampler's rows in a buffer of its own
    int rc = made_alloc_batch(m, batch);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (!m->epoch_rows || m->epoch_cap < batch) {
        if (m->epoch_rows) PIQMC_HIP(hipFree(m->epoch_rows));
        m->epoch_rows = nullptr;
        m->epoch_cap = 0;
        PIQMC_HIP(hipMalloc(&m->epoch_rows, (size_t)batch * m->N));
        m->epoch_cap = batch;
    }
    if (!m->nll_acc) PIQMC_HIP(hipMalloc(&m->nll_acc, sizeof(double)));
    zero_double<<<1, 1, 0, st>>>(m->nll_acc);
    const int steps = n / batch;
    for (int t = 0; t < steps; ++t) {
        gather_rows<<<grid1((size_t)batch * m->N), 256, 0, st>>>(data_dev, perm_dev, t * batch, batch, m->N,
                                                                 m->epoch_rows);
        rc = train_core(m, batch, m->epoch_rows, lr, st);
        if (rc) return rc;
        accumulate_sum<<<1, 256, 0, st>>>(m->nll, batch, m->nll_acc);
    }
    PIQMC_HIP(hipGetLastError());
    if (mean_nll) {
        double h = 0.0;
        PIQMC_HIP(hipMemcpyAsync(&h, m->nll_acc, sizeof(double), hipMemcpyDeviceToHost, st));
        PIQMC_HIP(hipStreamSynchronize(st));
        *mean_nll = h / ((double)steps * batch);
    }
    return PIQMC_OK;
}

// Draw B slices from q (autoregressive, N forward passes), int8 [B][N] to
// s_out_dev, log q of each draw to logq_dev.  Draw d of row b uses
// philox(key=seed, ctr={round, b, d, 0x5A4D0000})[0].
extern "C" int piqmc_made_sample(piqmc_made *m, int B, uint64_t seed, uint32_t round, int8_t *s_out_dev,
                                 double *logq_dev, void *stream)
{
    if (!m || B < 1 || !s_out_dev) return piqmc::set_error(PIQMC_EINVAL, "made_sample: bad args");
    int rc = made_alloc_batch(m, B);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int N = m->N;
    if (!m->rng) PIQMC_HIP(hipMalloc(&m->rng, 4 * sizeof(uint32_t)));
    const uint32_t rng_h[4] = {(uint32_t)seed, (uint32_t)(seed >> 32), round, 0};
    PIQMC_HIP(hipMemcpyAsync(m->rng, rng_h, sizeof(rng_h), hipMemcpyHostToDevice, st));
    PIQMC_HIP(hipStreamSynchronize(st));  // rng_h lives on this stack frame
    // N dependent steps of (3 GEMMs + draw): the whole chain is one sequence
    // of launches that only depends on B, so it is captured once into a HIP
    // graph and replayed.
    auto chain = [&](hipStream_t s) {
        fill_i8<<<grid1((size_t)B * N), 256, 0, s>>>(m->s, (size_t)B * N, 1);  // placeholders beyond the prefix
        spins_to_bf16<<<grid1((size_t)B * N), 256, 0, s>>>(m->s, (size_t)B * N, m->x);
        for (int d = 0; d < N; ++d) {
            forward(m, B, s);
            sample_step<<<(B + 255) / 256, 256, 0, s>>>(m->logits, B, N, d, m->rng, m->s, m->x);
        }
        forward(m, B, s);
        output_nll<<<B, 256, 0, s>>>(m->logits, m->s, B, N, 0.f, nullptr, m->nll, -1.0);
    };
    if (m->use_graph) {
        if (!m->gs) {
            PIQMC_HIP(hipStreamCreateWithFlags(&m->gs, hipStreamNonBlocking));
            PIQMC_HIP(hipEventCreateWithFlags(&m->ev_in, hipEventDisableTiming));
            PIQMC_HIP(hipEventCreateWithFlags(&m->ev_out, hipEventDisableTiming));
        }
        if (!m->gexec || m->gB != B) {
            if (m->gexec) (void)hipGraphExecDestroy(m->gexec), m->gexec = nullptr;
            hipGraph_t graph = nullptr;
            PIQMC_HIP(hipStreamBeginCapture(m->gs, hipStreamCaptureModeThreadLocal));
            chain(m->gs);
            PIQMC_HIP(hipStreamEndCapture(m->gs, &graph));
            hipError_t e = hipGraphInstantiate(&m->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return piqmc::hip_fail(e, "sampling graph instantiate");
            m->gB = B;
        }
        PIQMC_HIP(hipEventRecord(m->ev_in, st));
        PIQMC_HIP(hipStreamWaitEvent(m->gs, m->ev_in, 0));
        PIQMC_HIP(hipGraphLaunch(m->gexec, m->gs));
        PIQMC_HIP(hipEventRecord(m->ev_out, m->gs));
        PIQMC_HIP(hipStreamWaitEvent(st, m->ev_out, 0));
    } else {
        chain(st);
    }
    PIQMC_HIP(hipMemcpyAsync(s_out_dev, m->s, (size_t)B * N, hipMemcpyDeviceToDevice, st));
    if (logq_dev) PIQMC_HIP(hipMemcpyAsync(logq_dev, m->nll, 8 * (size_t)B, hipMemcpyDeviceToDevice, st));
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

namespace {
// old slice of replica r -> rows [R][N]
__global__ void gather_slices(const int8_t *confs, int R, int P, int N, const int32_t *slice, int8_t *out)
{
    const size_t tot = (size_t)R * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t / N), i = (int)(t % N);
        out[t] = confs[((size_t)r * P + slice[r]) * N + i];
    }
}

// One workgroup per replica: action change of replacing slice k by the
// proposal, Metropolis-Hastings test with the proposal densities, write-back.
__global__ void mh_kernel(int8_t *confs, int R, int P, int N, int Z, const int2 *rows, const int8_t *newsl,
                          const double *lq_new, const double *lq_old, const int32_t *slice, float jperp,
                          float temp, uint32_t k0, uint32_t k1, uint32_t round, int32_t *accepted)
{
    const int r = blockIdx.x;
    const int k = slice[r];
    const int8_t *old = confs + ((size_t)r * P + k) * N;
    const int8_t *nw = newsl + (size_t)r * N;
    const int8_t *lft = confs + ((size_t)r * P + (k == 0 ? P - 1 : k - 1)) * N;
    const int8_t *rgt = confs + ((size_t)r * P + (k == P - 1 ? 0 : k + 1)) * N;
    double dE = 0.0, dT = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const double so = old[i], sn = nw[i];
        for (int e = 0; e < Z; ++e) {
            const int2 rw = rows[(size_t)i * Z + e];
            const double J = (double)__int_as_float(rw.y);
            if (rw.x == i) dE -= J * (sn - so);
            else dE -= 0.5 * J * (sn * (double)nw[rw.x] - so * (double)old[rw.x]);
        }
        dT -= (double)jperp * (sn - so) * ((double)lft[i] + (double)rgt[i]);
    }
    __shared__ double red[2][16];
    for (int o = 32; o > 0; o >>= 1) {
        dE += __shfl_xor(dE, o, 64);
        dT += __shfl_xor(dT, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = dE;
        red[1][threadIdx.x >> 6] = dT;
    }
    __syncthreads();
    __shared__ int acc_flag;
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) a += red[0][w], b += red[1][w];
        const double logA = -(a + b) / (double)temp + lq_old[r] - lq_new[r];
        uint32_t out[4];
        philox4x32_10(round, (uint32_t)r, 0u, 0x4D480000u, k0, k1, out);
        const double u = ((double)out[0] + 0.5) * 2.3283064365386963e-10;
        acc_flag = (logA >= 0.0 || log(u) < logA) ? 1 : 0;
        accepted[r] = acc_flag;
    }
    __syncthreads();
    if (acc_flag) {
        int8_t *dst = confs + ((size_t)r * P + k) * N;
        for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = nw[i];
    }
}
}  // namespace

extern "C" int piqmc_made_mh(piqmc_made *m, const piqmc_problem *pb, int R, int P, int8_t *confs,
                             const int8_t *newsl, const double *lq_new, const int32_t *slice, float jperp,
                             float temp, uint64_t seed, uint32_t round, int32_t *accepted, void *stream)
{
    if (!m || !pb || R < 1 || P < 1 || !confs || !newsl || !lq_new || !slice || !accepted)
        return piqmc::set_error(PIQMC_EINVAL, "made_mh: bad args");
    if (pb->N != m->N) return piqmc::set_error(PIQMC_EINVAL, "made_mh: network and problem sizes differ");
    hipStream_t st = (hipStream_t)stream;
    int8_t *old = nullptr;
    double *lq_old = nullptr;
    PIQMC_HIP(hipMallocAsync((void **)&old, (size_t)R * m->N, st));
    PIQMC_HIP(hipMallocAsync((void **)&lq_old, 8 * (size_t)R, st));
    gather_slices<<<grid1((size_t)R * m->N), 256, 0, st>>>(confs, R, P, m->N, slice, old);
    int rc = piqmc_made_logprob(m, R, old, lq_old, stream);
    if (rc) {
        (void)hipFreeAsync(old, st);
        (void)hipFreeAsync(lq_old, st);
        return rc;
    }
    mh_kernel<<<R, 256, 0, st>>>(confs, R, P, m->N, pb->Z, pb->d_rows, newsl, lq_new, lq_old, slice, jperp, temp,
                                 (uint32_t)seed, (uint32_t)(seed >> 32), round, accepted);
    PIQMC_HIP(hipFreeAsync(old, st));
    PIQMC_HIP(hipFreeAsync(lq_old, st));
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

// The independence slice move with the fused sampler: slice k = slice_dev[r]
// of replica r is replaced by y ~ q_beta with min(1, exp(-dS/T) q_beta(old) /
// q_beta(y)); both densities come from the fused sampler's arithmetic.
extern "C" int piqmc_made_slice_mh(piqmc_made *m, const piqmc_problem *pb, int R, int P, int8_t *confs,
                                   const int32_t *slice, float beta, float jperp, float temp, uint64_t seed,
                                   uint32_t round, int32_t *accepted, void *stream)
{
    if (!m || !pb || R < 1 || P < 1 || !confs || !slice || !accepted)
        return piqmc::set_error(PIQMC_EINVAL, "made_slice_mh: bad args");
    if (pb->N != m->N) return piqmc::set_error(PIQMC_EINVAL, "made_slice_mh: network and problem sizes differ");
    if (!(beta > 0.f) || !std::isfinite(beta)) return piqmc::set_error(PIQMC_EINVAL, "made_slice_mh: beta must be > 0");
    hipStream_t st = (hipStream_t)stream;
    int8_t *rows = nullptr;
    double *lq = nullptr;
    int rc = piqmc::made_ar_scratch(m, R, &rows, &lq);
    if (rc) return rc;
    int8_t *y = rows, *old = rows + (size_t)R * m->N;
    gather_slices<<<grid1((size_t)R * m->N), 256, 0, st>>>(confs, R, P, m->N, slice, old);
    rc = piqmc::made_ar_launch(m, R, R, old, beta, seed, round, y, lq, st);
    if (rc) return rc;
    mh_kernel<<<R, 256, 0, st>>>(confs, R, P, m->N, pb->Z, pb->d_rows, y, lq, lq + R, slice, jperp, temp,
                                 (uint32_t)seed, (uint32_t)(seed >> 32), round, accepted);
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

extern "C" int piqmc_made_set_graph(piqmc_made *m, int use_graph)
{
    if (!m) return piqmc::set_error(PIQMC_EINVAL, "null made");
    m->use_graph = use_graph ? 1 : 0;
    return PIQMC_OK;
}

// Plain GEMM entry for tests: C[M][N] = A[M][K] . B[N][K]^T (bf16 in, fp32 out).
extern "C" int piqmc_gemm_bf16_nt(const uint16_t *A, const uint16_t *B, int M, int N, int K, float *C, void *stream)
{
    if (!A || !B || !C || M < 1 || N < 1 || K < 1) return piqmc::set_error(PIQMC_EINVAL, "gemm: bad args");
    gemm(A, K, B, K, M, N, K, nullptr, 0, C, nullptr, N, (hipStream_t)stream);
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}
