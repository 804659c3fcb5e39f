// Fused autoregressive sampler of the MADE slice network, and the world-line
// move it proposes.
//
// Sampling a MADE draws spin d from p(s_d = -1 | s_<d) = sigmoid(l_d), and
// l_d only depends on the units whose inputs are all known.  With natural
// order and degrees m(k) = 1 + k mod (N-1), a first-layer unit of degree
// e + 1 is complete once s_0..s_e are known, a second-layer unit of degree
// e + 1 once every first-layer unit of degree <= e + 1 is, and logit e + 1
// once every second-layer unit of degree <= e + 1 is.  So one pass over the
// positions keeps running pre-activations and, at position e:
//   z1 += W0[:, e] s_e                           (every unit still open)
//   a_k = bf16(relu(z1_k)) for the units of degree e + 1;  z2 += W1[:, k] a_k
//   b_k = bf16(relu(z2_k)) for the units of degree e + 1;  l  += W2[:, k] b_k
//   l_{e+1} is final: draw s_{e+1} (or read it, to evaluate a given slice).
// That is the work of one forward pass instead of N of them.  Masked weights
// are exact zeros, so the column updates may run over every unit.
//
// Layout: one wavefront carries RW rows; lane ln holds, for each row, the
// pre-activations of the units of degree (q*64 + ln) + 1 (slot j*DS + q for
// the j-th unit of that degree, k = d + j (N-1)) and logit q*64 + ln + 1 in
// slot q.  The units completing at position e then sit in lane e mod 64 and a
// slot fixed for the whole run of 64 positions e / 64, so they are read with
// v_readlane from statically indexed registers.  Weight columns are stored
// per lane in that layout (bf16, 16-byte groups).
//
// The sampler defines its own density: every logit is a fixed sequence of
// fp32 operations of the spins before it, computed the same way for a drawn
// slice and for a given one, so the log q it returns with a draw equals the
// log q it evaluates for that slice, bit for bit.  q_beta(s) = prod_d
// sigmoid(+-beta l_d) is a normalized distribution for every beta > 0; beta > 1
// sharpens the proposal towards the network's modes.
#include <math.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "made_state.h"

namespace {

// rows per wavefront: 4, or 1 for long slices (register budget)
constexpr int ar_rw(int DS) { return DS >= 16 ? 1 : 4; }

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float bf16_round(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}
__device__ __forceinline__ float readlane_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

struct ArArgs {
    int N, H1, H2, B, Bs;
    float beta;
    const uint16_t *w0t;  // [N-1][64][ZSP]  input column e in the z1 layout
    const uint16_t *w1t;  // [H1][64][ZSP]   first-layer unit column in the z2 layout
    const uint16_t *w2t;  // [H2][64][DSP]   second-layer unit column in the logit layout
    const float *b0l, *b1l;  // [64][ZS]
    const float *b2l;        // [64][DS]  bias of logit q*64 + ln + 1
    const float *b2;         // b2[0]: logit 0
    const int8_t *forced;    // [B - Bs][N]
    int8_t *s_out;           // [Bs][N]
    double *logq;            // [B]
    uint32_t k0, k1, round;
};

// Unit of the j-th entry of degree d (0-based degree index d = m - 1), or -1.
__host__ __device__ inline int ar_unit(int d, int j, int N, int H)
{
    if (d > N - 2) return -1;
    const int k = d + j * (N - 1);
    return k < H ? k : -1;
}

// Tables of the fused sampler from the bf16 weights wb0 [H1][N], wb1 [H2][H1],
// wb2 [N][H2] and the fp32 biases.
__global__ void ar_tables_kernel(int N, int H1, int H2, int DS, int JM, const uint16_t *wb0, const uint16_t *wb1,
                                 const uint16_t *wb2, const float *b0, const float *b1, const float *b2,
                                 uint16_t *w0t, uint16_t *w1t, uint16_t *w2t, float *b0l, float *b1l, float *b2l)
{
    const int ZS = DS * JM, ZSP = (ZS + 7) & ~7, DSP = (DS + 7) & ~7;
    const size_t n0 = (size_t)(N - 1) * 64 * ZSP, n1 = (size_t)H1 * 64 * ZSP, n2 = (size_t)H2 * 64 * DSP;
    const size_t nb = (size_t)64 * (2 * ZS + DS);
    const size_t tot = n0 + n1 + n2 + nb;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        if (t < n0) {
            const int s = (int)(t % ZSP), ln = (int)((t / ZSP) % 64), e = (int)(t / ((size_t)64 * ZSP));
            const int k = s < ZS ? ar_unit((s % DS) * 64 + ln, s / DS, N, H1) : -1;
            w0t[t] = k >= 0 ? wb0[(size_t)k * N + e] : (uint16_t)0;
        } else if (t < n0 + n1) {
            const size_t u = t - n0;
            const int s = (int)(u % ZSP), ln = (int)((u / ZSP) % 64), k1 = (int)(u / ((size_t)64 * ZSP));
            const int k2 = s < ZS ? ar_unit((s % DS) * 64 + ln, s / DS, N, H2) : -1;
            w1t[u] = k2 >= 0 ? wb1[(size_t)k2 * H1 + k1] : (uint16_t)0;
        } else if (t < n0 + n1 + n2) {
            const size_t u = t - n0 - n1;
            const int q = (int)(u % DSP), ln = (int)((u / DSP) % 64), k2 = (int)(u / ((size_t)64 * DSP));
            const int d = q * 64 + ln;  // logit d + 1
            w2t[u] = (q < DS && d <= N - 2) ? wb2[(size_t)(d + 1) * H2 + k2] : (uint16_t)0;
        } else {
            const size_t u = t - n0 - n1 - n2;
            const int ln = (int)(u / (2 * ZS + DS)), c = (int)(u % (2 * ZS + DS));
            if (c < ZS) {
                const int k = ar_unit((c % DS) * 64 + ln, c / DS, N, H1);
                b0l[ln * ZS + c] = k >= 0 ? b0[k] : 0.f;
            } else if (c < 2 * ZS) {
                const int s = c - ZS, k = ar_unit((s % DS) * 64 + ln, s / DS, N, H2);
                b1l[ln * ZS + s] = k >= 0 ? b1[k] : 0.f;
            } else {
                const int q = c - 2 * ZS, d = q * 64 + ln;
                b2l[ln * DS + q] = d <= N - 2 ? b2[d + 1] : 0.f;
            }
        }
    }
}

// Unpack 8 bf16 of a 16-byte group into floats.
__device__ __forceinline__ void unpack8(const uint4 v, float *f)
{
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        f[2 * h] = __uint_as_float(w[h] << 16);
        f[2 * h + 1] = __uint_as_float(w[h] & 0xFFFF0000u);
    }
}

template <int DS, int JM>
struct ArLayout {
    static constexpr int ZS = DS * JM, ZSP = (ZS + 7) & ~7, DSP = (DS + 7) & ~7;
};

// Decision at position t for the wavefront's rows (lane r < RW handles row
// row0 + r): the logit of row r is lg[r] (wave-uniform).  Returns the spin
// (+1/-1, as float) of every row, wave-uniform, in xs.
template <int RW>
__device__ __forceinline__ void ar_decide(const ArArgs &a, int t, int ln, int row, bool rowok, bool sampled,
                                          int fxv, const float *lg, uint32_t *rnd, double &lq, float *xs)
{
    float my = lg[0];
#pragma unroll
    for (int r = 1; r < RW; ++r) my = ln == r ? lg[r] : my;
    const double lb = (double)my * (double)a.beta;
    const double ex = exp(-fabs(lb));  // in (0, 1]
    const double l1p = log1p(ex);
    int x = 1;
    if (sampled) {
        if ((t & 3) == 0) philox4x32_10(a.round, (uint32_t)row, (uint32_t)(t >> 2), 0x4D415200u, a.k0, a.k1, rnd);
        const uint32_t w = (t & 3) == 0 ? rnd[0] : ((t & 3) == 1 ? rnd[1] : ((t & 3) == 2 ? rnd[2] : rnd[3]));
        const double u = ((double)w + 0.5) * 2.3283064365386963e-10;
        const double p = lb >= 0.0 ? 1.0 / (1.0 + ex) : ex / (1.0 + ex);  // p(s_t = -1)
        x = u < p ? -1 : 1;
    } else {
        x = fxv < 0 ? -1 : 1;
    }
    // log sigmoid(+-lb) = -(max(-+lb, 0) + log1p(exp(-|lb|)))
    if (rowok) lq -= (x < 0 ? fmax(-lb, 0.0) : fmax(lb, 0.0)) + l1p;
    if (rowok && sampled) a.s_out[(size_t)row * a.N + t] = (int8_t)x;
    const float xf = (float)x;
#pragma unroll
    for (int r = 0; r < RW; ++r) xs[r] = readlane_f(xf, r);
}

template <int DS, int JM, int RW>
__global__ __launch_bounds__(64) void made_ar_kernel(ArArgs a)
{
    using Lay = ArLayout<DS, JM>;
    constexpr int ZS = Lay::ZS, ZSP = Lay::ZSP, DSP = Lay::DSP;
    const int ln = threadIdx.x;
    const int row0 = blockIdx.x * RW;
    const int N = a.N;
    float z1[RW][ZS], z2[RW][ZS], lp[RW][DS];
#pragma unroll
    for (int s = 0; s < ZS; ++s) {
        const float b0 = a.b0l[ln * ZS + s], b1 = a.b1l[ln * ZS + s];
#pragma unroll
        for (int r = 0; r < RW; ++r) z1[r][s] = b0, z2[r][s] = b1;
    }
#pragma unroll
    for (int q = 0; q < DS; ++q) {
        const float b2 = a.b2l[ln * DS + q];
#pragma unroll
        for (int r = 0; r < RW; ++r) lp[r][q] = b2;
    }
    // lane r < RW carries row row0 + r's draws and log q
    const int row = row0 + ln;
    const bool rowok = ln < RW && row < a.B;
    const bool sampled = rowok && row < a.Bs;
    const int8_t *frow = (rowok && !sampled) ? a.forced + (size_t)(row - a.Bs) * N : nullptr;
    uint32_t rnd[4] = {0u, 0u, 0u, 0u};
    double lq = 0.0;
    float xs[RW];
    {
        float lg[RW];
        const float l0 = a.b2[0];
#pragma unroll
        for (int r = 0; r < RW; ++r) lg[r] = l0;
        ar_decide<RW>(a, 0, ln, row, rowok, sampled, frow ? (int)frow[0] : 1, lg, rnd, lq, xs);
    }
#pragma unroll
    for (int c = 0; c < DS; ++c) {
        // given spins of this run of positions: lane ln holds position 64c + ln + 1 of each row
        int fx[RW];
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int rr = row0 + r, pos = 64 * c + ln + 1;
            fx[r] = (rr < a.B && rr >= a.Bs && pos < N) ? (int)a.forced[(size_t)(rr - a.Bs) * N + pos] : 1;
        }
        const int e_end = min(64 * c + 64, N - 1);
        for (int e = 64 * c; e < e_end; ++e) {
            const int el = e & 63;
            // z1 += W0[:, e] s_e over the units still open (degree index >= e)
            {
                const uint4 *src = reinterpret_cast<const uint4 *>(a.w0t + ((size_t)e * 64 + ln) * ZSP);
                float w[ZSP];
#pragma unroll
                for (int g = 0; g < ZSP / 8; ++g) unpack8(src[g], w + 8 * g);
#pragma unroll
                for (int s = 0; s < ZS; ++s) {
                    if (s % DS < c) continue;
#pragma unroll
                    for (int r = 0; r < RW; ++r) z1[r][s] = fmaf(w[s], xs[r], z1[r][s]);
                }
            }
            // first-layer units of degree index e complete: feed the second layer
#pragma unroll
            for (int j = 0; j < JM; ++j) {
                const int k = e + j * (N - 1);
                if (k >= a.H1) continue;
                float av[RW];
#pragma unroll
                for (int r = 0; r < RW; ++r) av[r] = bf16_round(fmaxf(readlane_f(z1[r][j * DS + c], el), 0.f));
                const uint4 *src = reinterpret_cast<const uint4 *>(a.w1t + ((size_t)k * 64 + ln) * ZSP);
                float w[ZSP];
#pragma unroll
                for (int g = 0; g < ZSP / 8; ++g) unpack8(src[g], w + 8 * g);
#pragma unroll
                for (int s = 0; s < ZS; ++s) {
                    if (s % DS < c) continue;
#pragma unroll
                    for (int r = 0; r < RW; ++r) z2[r][s] = fmaf(w[s], av[r], z2[r][s]);
                }
            }
            // second-layer units of degree index e complete: feed the logits
#pragma unroll
            for (int j = 0; j < JM; ++j) {
                const int k = e + j * (N - 1);
                if (k >= a.H2) continue;
                float bv[RW];
#pragma unroll
                for (int r = 0; r < RW; ++r) bv[r] = bf16_round(fmaxf(readlane_f(z2[r][j * DS + c], el), 0.f));
                const uint4 *src = reinterpret_cast<const uint4 *>(a.w2t + ((size_t)k * 64 + ln) * DSP);
                float w[DSP];
#pragma unroll
                for (int g = 0; g < DSP / 8; ++g) unpack8(src[g], w + 8 * g);
#pragma unroll
                for (int q = c; q < DS; ++q) {
#pragma unroll
                    for (int r = 0; r < RW; ++r) lp[r][q] = fmaf(w[q], bv[r], lp[r][q]);
                }
            }
            // logit e + 1 is final
            float lg[RW];
            int fxv = 1;
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                lg[r] = readlane_f(lp[r][c], el);
                const int v = __builtin_amdgcn_readlane(fx[r], el);
                fxv = ln == r ? v : fxv;
            }
            ar_decide<RW>(a, e + 1, ln, row, rowok, sampled, fxv, lg, rnd, lq, xs);
        }
    }
    if (rowok) a.logq[row] = lq;
}

// One workgroup per replica: the world-line move.  f = y * s^0 (y the drawn
// slice, s^0 the replica's slice 0); every slice k becomes s^k * f, which
// leaves every Trotter bond unchanged, so the action changes by
// sum_k [E(s^k f) - E(s^k)].  Accepted with min(1, exp(-dA/T) q(s^0)/q(y)).
__global__ void worldline_mh_kernel(int8_t *confs, int R, int P, int N, int Z, const int2 *rows, const int8_t *y,
                                    const double *lq, float temp, uint32_t k0, uint32_t k1, uint32_t round,
                                    int32_t *accepted)
{
    const int r = blockIdx.x;
    int8_t *base = confs + (size_t)r * P * N;
    const int8_t *yr = y + (size_t)r * N;
    double dA = 0.0;
    for (int t = threadIdx.x; t < P * N; t += blockDim.x) {
        const int k = t / N, i = t - k * N;
        const int8_t *sk = base + (size_t)k * N;
        const int fi = (int)yr[i] * (int)base[i];  // base = slice 0
        const double si = sk[i];
        for (int z = 0; z < Z; ++z) {
            const int2 rw = rows[(size_t)i * Z + z];
            const double J = (double)__int_as_float(rw.y);
            if (rw.x == i) {
                if (fi < 0) dA += 2.0 * J * si;  // -h s_i: flipping adds 2 h s_i
            } else {
                const int fj = (int)yr[rw.x] * (int)base[rw.x];
                if (fi != fj) dA += J * si * (double)sk[rw.x];  // 2 J s_i s_j, each pair seen twice
            }
        }
    }
    __shared__ double red[16];
    for (int o = 32; o > 0; o >>= 1) dA += __shfl_xor(dA, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dA;
    __syncthreads();
    __shared__ int acc_flag;
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) t += red[w];
        const double logA = -t / (double)temp + lq[R + r] - lq[r];
        uint32_t out[4];
        philox4x32_10(round, (uint32_t)r, 0u, 0x574C0000u, k0, k1, out);
        const double u = ((double)out[0] + 0.5) * 2.3283064365386963e-10;
        acc_flag = (logA >= 0.0 || log(u) < logA) ? 1 : 0;
        accepted[r] = acc_flag;
    }
    __syncthreads();
    if (acc_flag) {
        for (int t = threadIdx.x; t < P * N; t += blockDim.x) {
            const int i = t % N;
            // f_i from the untouched row y and the original slice 0: slice 0
            // itself is rewritten last, after every other slice read it
            if (t >= N) base[t] = (int8_t)(base[t] * yr[i] * base[i]);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < N; i += blockDim.x) base[i] = yr[i];
    }
}

__global__ void gather_slice0(const int8_t *confs, int R, int P, int N, int8_t *out)
{
    const size_t tot = (size_t)R * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t / N), i = (int)(t % N);
        out[t] = confs[(size_t)r * P * N + i];
    }
}

int grid1(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 65535); }

typedef void (*ArKernel)(ArArgs);

ArKernel ar_kernel_for(int DS, int JM)
{
#define AR_CASE(ds, jm) \
    if (DS == ds && JM == jm) return made_ar_kernel<ds, jm, ar_rw(ds)>;
    AR_CASE(1, 1) AR_CASE(1, 2) AR_CASE(1, 3) AR_CASE(1, 4)
    AR_CASE(2, 1) AR_CASE(2, 2)
    AR_CASE(4, 1) AR_CASE(4, 2) AR_CASE(4, 3)
    AR_CASE(27, 1)
#undef AR_CASE
    return nullptr;
}

int ar_shape(const piqmc_made *m, int *DS, int *JM)
{
    const int N = m->N;
    if (N < 2) return piqmc::set_error(PIQMC_EINVAL, "fused sampler: nspins must be >= 2");
    *DS = (N - 1 + 63) / 64;
    const int j1 = (m->H1 + N - 2) / (N - 1), j2 = (m->H2 + N - 2) / (N - 1);
    *JM = std::max(j1, j2);
    if (!ar_kernel_for(*DS, *JM))
        return piqmc::set_error(PIQMC_EINVAL, "fused sampler: no kernel for " + std::to_string(N) + " spins with " +
                                                  std::to_string(m->H1) + "/" + std::to_string(m->H2) +
                                                  " hidden units (supported: 2 <= N <= 65 with H <= 4(N-1), "
                                                  "66 <= N <= 129 with H <= 2(N-1), 194 <= N <= 257 with "
                                                  "H <= 3(N-1), 1666 <= N <= 1729 with H < N)");
    return PIQMC_OK;
}

int ensure_tables(piqmc_made *m, hipStream_t st)
{
    int DS = 0, JM = 0;
    int rc = ar_shape(m, &DS, &JM);
    if (rc) return rc;
    if (!m->ar_dirty && m->ar_w0t) return PIQMC_OK;
    const int N = m->N, ZS = DS * JM, ZSP = (ZS + 7) & ~7, DSP = (DS + 7) & ~7;
    if (!m->ar_w0t) {
        PIQMC_HIP(hipMalloc(&m->ar_w0t, 2 * (size_t)(N - 1) * 64 * ZSP));
        PIQMC_HIP(hipMalloc(&m->ar_w1t, 2 * (size_t)m->H1 * 64 * ZSP));
        PIQMC_HIP(hipMalloc(&m->ar_w2t, 2 * (size_t)m->H2 * 64 * DSP));
        PIQMC_HIP(hipMalloc(&m->ar_b0l, 4 * (size_t)64 * ZS));
        PIQMC_HIP(hipMalloc(&m->ar_b1l, 4 * (size_t)64 * ZS));
        PIQMC_HIP(hipMalloc(&m->ar_b2l, 4 * (size_t)64 * DS));
        m->ar_ds = DS;
        m->ar_jm = JM;
    }
    ar_tables_kernel<<<1024, 256, 0, st>>>(N, m->H1, m->H2, DS, JM, m->wb[0], m->wb[1], m->wb[2], m->b[0], m->b[1],
                                           m->b[2], m->ar_w0t, m->ar_w1t, m->ar_w2t, m->ar_b0l, m->ar_b1l, m->ar_b2l);
    PIQMC_HIP(hipGetLastError());
    m->ar_dirty = 0;
    return PIQMC_OK;
}

int launch_ar(piqmc_made *m, int Bs, int Bf, const int8_t *forced, float beta, uint64_t seed, uint32_t round,
              int8_t *s_out, double *logq, hipStream_t st)
{
    int rc = ensure_tables(m, st);
    if (rc) return rc;
    ArArgs a = {};
    a.N = m->N; a.H1 = m->H1; a.H2 = m->H2; a.B = Bs + Bf; a.Bs = Bs; a.beta = beta;
    a.w0t = m->ar_w0t; a.w1t = m->ar_w1t; a.w2t = m->ar_w2t;
    a.b0l = m->ar_b0l; a.b1l = m->ar_b1l; a.b2l = m->ar_b2l; a.b2 = m->b[2];
    a.forced = forced; a.s_out = s_out; a.logq = logq;
    a.k0 = (uint32_t)seed; a.k1 = (uint32_t)(seed >> 32); a.round = round;
    const int B = Bs + Bf;
    const int rw = ar_rw(m->ar_ds);
    ar_kernel_for(m->ar_ds, m->ar_jm)<<<(B + rw - 1) / rw, 64, 0, st>>>(a);
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

}  // namespace

namespace piqmc {
int made_ar_launch(piqmc_made *m, int Bs, int Bf, const int8_t *forced, float beta, uint64_t seed, uint32_t round,
                   int8_t *s_out, double *logq, hipStream_t st)
{
    return launch_ar(m, Bs, Bf, forced, beta, seed, round, s_out, logq, st);
}

int made_ar_scratch(piqmc_made *m, int R, int8_t **rows, double **lq)
{
    const int N = m->N;
    if ((size_t)R > m->ar_cap) {
        if (m->ar_rows) PIQMC_HIP(hipFree(m->ar_rows));
        if (m->ar_lq) PIQMC_HIP(hipFree(m->ar_lq));
        m->ar_rows = nullptr;
        m->ar_lq = nullptr;
        m->ar_cap = 0;
        PIQMC_HIP(hipMalloc(&m->ar_rows, 2 * (size_t)R * N));
        PIQMC_HIP(hipMalloc(&m->ar_lq, 2 * 8 * (size_t)R));
        m->ar_cap = (size_t)R;
    }
    *rows = m->ar_rows;
    *lq = m->ar_lq;
    return PIQMC_OK;
}

void made_ar_release(piqmc_made *m)
{
    void *ps[] = {m->ar_w0t, m->ar_w1t, m->ar_w2t, m->ar_b0l, m->ar_b1l, m->ar_b2l, m->ar_rows, m->ar_lq};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    m->ar_w0t = m->ar_w1t = m->ar_w2t = nullptr;
    m->ar_b0l = m->ar_b1l = m->ar_b2l = nullptr;
    m->ar_rows = nullptr;
    m->ar_lq = nullptr;
    m->ar_cap = 0;
}
}  // namespace piqmc

extern "C" int piqmc_made_ar(piqmc_made *m, int Bs, int Bf, const int8_t *forced_dev, float beta, uint64_t seed,
                             uint32_t round, int8_t *s_out_dev, double *logq_dev, void *stream)
{
    if (!m || Bs < 0 || Bf < 0 || (Bs && !s_out_dev) || (Bf && !forced_dev) || ((Bs || Bf) && !logq_dev))
        return piqmc::set_error(PIQMC_EINVAL, "made_ar: bad arguments");
    if (!(beta > 0.f) || !std::isfinite(beta)) return piqmc::set_error(PIQMC_EINVAL, "made_ar: beta must be finite and > 0");
    if (Bs + Bf == 0) return PIQMC_OK;
    return launch_ar(m, Bs, Bf, forced_dev, beta, seed, round, s_out_dev, logq_dev, (hipStream_t)stream);
}

extern "C" int piqmc_made_worldline_mh(piqmc_made *m, const piqmc_problem *pb, int R, int P, int8_t *confs,
                                       float beta, float temp, uint64_t seed, uint32_t round, int32_t *accepted,
                                       void *stream)
{
    if (!m || !pb || R < 1 || P < 1 || !confs || !accepted)
        return piqmc::set_error(PIQMC_EINVAL, "made_worldline_mh: bad arguments");
    if (pb->N != m->N) return piqmc::set_error(PIQMC_EINVAL, "made_worldline_mh: network and problem sizes differ");
    if (!(temp > 0.f) || !std::isfinite(temp)) return piqmc::set_error(PIQMC_EINVAL, "made_worldline_mh: temp must be > 0");
    if (!(beta > 0.f) || !std::isfinite(beta)) return piqmc::set_error(PIQMC_EINVAL, "made_worldline_mh: beta must be > 0");
    hipStream_t st = (hipStream_t)stream;
    const int N = m->N;
    int8_t *rows = nullptr;
    double *lq = nullptr;
    int rc = piqmc::made_ar_scratch(m, R, &rows, &lq);
    if (rc) return rc;
    int8_t *y = rows, *s0 = rows + (size_t)R * N;
    gather_slice0<<<grid1((size_t)R * N), 256, 0, st>>>(confs, R, P, N, s0);
    // rows [0, R): draws y; rows [R, 2R): the replicas' slices 0, evaluated
    rc = launch_ar(m, R, R, s0, beta, seed, round, y, lq, st);
    if (rc) return rc;
    worldline_mh_kernel<<<R, 256, 0, st>>>(confs, R, P, N, pb->Z, pb->d_rows, y, lq, temp, (uint32_t)seed,
                                          (uint32_t)(seed >> 32), round, accepted);
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}
