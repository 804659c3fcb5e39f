// Replay sweeps: qmc.QuantumAnneal (piqmc/qmc.pyx:30-136) and sa.Anneal
// (piqmc/sa.pyx:50-120) reproduced bit-for-bit on gfx950.
//
// A replica's sweep is a strictly sequential Markov chain (the reference keeps
// a running ediff across the spins of a slice, D2, and draws libc rand() only
// on the rejection branch), so one lane owns one replica and the replicas of
// a wave advance in lockstep.  A few lanes per wave own replicas (a chain's
// speed is its latency, so replicas are spread over every SIMD) and the waves
// of a workgroup share one LDS copy of the neighbour rows.  Per replica the
// kernel keeps, bit-packed in LDS with slot-interleaved words (conflict-free):
//   - the slice being swept (streamed from HBM once per slice),
//   - slices 1 and P-1, the only Trotter neighbours the reference reads (D1),
//   - the 31-word glibc TYPE_3 ring of its rand() stream.
// The permutation stream is read from HBM as [sweep][position][replica]
// uint16; each replica's permutation of a sweep is staged in LDS once and
// read by every slice of the sweep, one attempt ahead of use.
//
// Arithmetic is the reference's: ediff is float and every term is an exact
// +-2*float product, so the float add equals the reference's double add
// rounded to float; ediff/temp is IEEE float division; exp() is double; the
// comparison is against rand()/2147483647.0 in double.
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "piqmc_internal.h"

namespace {

struct ReplayArgs {
    int R, N, P, Z, Zp, S, M, Wn, sa, Lw;
    float temp;
    const float *steps;     // [S] J_perp (QA) or float temperature (SA)
    const int2 *rows;       // [N][Z] (j, J bits)
    const double *jd;       // [N][Z] double J (problems built from a dense matrix), else null
    const uint16_t *perms;  // [S*M][N][R]
    uint32_t *bits;         // [P][Wn][R]
    uint32_t *rstate;       // [R][32]
    unsigned long long *draws;
    uint32_t *ties;
};

__device__ __forceinline__ uint32_t getbit(const uint32_t *reg, int Ls, int lane, int i)
{
    return (reg[(i >> 5) * Ls + lane] >> (i & 31)) & 1u;
}

// LDS word stride of a workgroup with L replica slots: at least 8, so that a
// row of the bit-packed spins spans a multiple of 32 bytes.
__host__ __device__ inline int replay_stride(int L) { return L < 8 ? 8 : L; }

// Words of one spin region: (Wn + 1) rows of Ls words, rounded up to a power
// of two, so that regions aligned to their size hold every row offset below
// the alignment (see ORA in replay_kernel).
__host__ __device__ inline int replay_region_words(int Wn, int Ls)
{
    int rs = 1;
    while (rs < (Wn + 1) * Ls) rs <<= 1;
    return rs;
}

// LDS image of a workgroup, words slot-interleaved with stride Ls: up to one
// region of padding (the regions start aligned to their size), spin regions
// [nreg][RS] u32 of which rows [Wn + 1][Ls] are used (row Wn stays zero),
// rand() rings [31][Ls] u32, and with STAGED (8-byte aligned) the neighbour
// rows [N][Zp] int2, padded with zero fields to a multiple of 4 entries, then
// the sweep's permutations [N][Ls] u16.
__host__ __device__ inline size_t replay_words(int nreg, int Wn, int Ls)
{
    const size_t rs = (size_t)replay_region_words(Wn, Ls);
    return (rs + nreg * rs + (size_t)31 * Ls + 1) & ~(size_t)1;
}

// LDS byte address of a pointer into shared memory, and a word read at one.
typedef __attribute__((address_space(3))) const uint32_t LdsWord;
__device__ __forceinline__ uint32_t lds_addr(const void *p)
{
    return (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const char *)p;
}
__device__ __forceinline__ uint32_t lds_word(uint32_t addr)
{
    return *(LdsWord *)(uintptr_t)addr;
}

// ORA (staged float couplings, at most 8 slots per workgroup): a neighbour's
// word is at (region-aligned base + slot * 4) + (row offset & ~31), and the
// two never share a bit (the slot offset is below 32, the row offset below the
// region's alignment), so one AND-OR forms the address instead of an AND and
// an add.
template <bool F64J, bool STAGED, bool SA, bool ORA = false>
__global__ void replay_kernel(ReplayArgs a)
{
    extern __shared__ uint32_t sm[];
    // The first Lw lanes of each wave own one replica each ("slots"); a
    // replica's sweep is a dependency chain, so few lanes per wave spread the
    // replicas over every SIMD instead of stacking them in a few waves.
    const int Lw = a.Lw, L = (blockDim.x >> 6) * Lw;  // L: slots per workgroup
    const int Ls = replay_stride(L);                  // LDS word stride, a power of two
    const int lsh = 31 - __clz(Ls);
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const int lane = wv * Lw + ln;
    const int r = blockIdx.x * L + lane;
    const int N = a.N, P = a.P, Z = a.Z, Wn = a.Wn, R = a.R;
    const int nreg = SA ? 1 : (P == 2 ? 2 : 3);
    const int Zr = STAGED ? a.Zp : Z;  // row stride of the table in use
    int2 *lrows = reinterpret_cast<int2 *>(sm + replay_words(nreg, Wn, Ls));
    uint16_t *lperm = reinterpret_cast<uint16_t *>(lrows + (size_t)N * a.Zp);
    const int2 *rows = STAGED ? lrows : a.rows;
    if (STAGED) {  // every thread helps, then the only barrier
        for (int q = threadIdx.x; q < N * a.Zp; q += blockDim.x) {
            const int i = q / a.Zp, z = q - i * a.Zp;
            int2 v = z < Z ? a.rows[(size_t)i * Z + z] : make_int2(i, 0);
            if (!F64J) {
                // Float couplings are staged as -2J (exact), so the term of an
                // entry is -2J with its sign bit flipped when the two spins
                // differ, and the neighbour as the byte offset of its spin
                // word within a region (a multiple of 32: Ls >= 8) plus
                // 31 - (j mod 32), the shift that brings its bit to bit 31.
                // Self entries and padding point at the region's zero row.
                const int j = v.x;
                const uint32_t row = j == i ? (uint32_t)Wn : (uint32_t)j >> 5;
                const uint32_t sh = j == i ? 0u : 31u - ((uint32_t)j & 31u);
                v.x = (int)(((row << lsh) << 2) | sh);
                v.y = __float_as_int(-2.0f * __int_as_float(v.y));
            }
            lrows[q] = v;
        }
        __syncthreads();
    }
    if (ln >= Lw || r >= a.R) return;  // no barriers below: a slot touches only its own words
    // regions aligned to their size (RS words)
    const int RS = replay_region_words(Wn, Ls);
    const uint32_t rsb = (uint32_t)RS * 4u;
    uint32_t *base = sm + ((rsb - (lds_addr(sm) & (rsb - 1u))) & (rsb - 1u)) / 4u;
    uint32_t *cur = base, *s1, *sp, *ring;
    if (SA) {
        s1 = sp = cur;
        ring = base + RS;
    } else {
        s1 = base + RS;
        sp = (P == 2) ? s1 : base + 2 * RS;
        ring = base + nreg * RS;  // nreg spin regions, then the rand() rings
    }
    for (int g = 0; g < nreg; ++g) base[g * RS + Wn * Ls + lane] = 0u;  // zero rows
    auto load = [&](uint32_t *reg, int k) {
        for (int w = 0; w < Wn; ++w) reg[w * Ls + lane] = a.bits[((size_t)k * Wn + w) * R + r];
    };
    auto store = [&](const uint32_t *reg, int k) {
        for (int w = 0; w < Wn; ++w) a.bits[((size_t)k * Wn + w) * R + r] = reg[w * Ls + lane];
    };
    if (SA) {
        load(cur, 0);
    } else {
        load(s1, 1);
        if (P != 2) load(sp, P - 1);
    }
    for (int w = 0; w < 31; ++w) ring[w * Ls + lane] = a.rstate[(size_t)r * 32 + w];
    uint32_t f = a.rstate[(size_t)r * 32 + 31] % 31u;
    unsigned long long nd = 0;
    uint32_t nt = 0;
    const float T = a.temp;
    const int nslice = SA ? 1 : P;

    for (int s = 0; s < a.S; ++s) {
        const float sv = a.steps[s];
        const float tdiv = SA ? sv : T;
        const uint32_t svm2 = __float_as_uint(-2.0f * sv);  // exact
        for (int m = 0; m < a.M; ++m) {
            const uint16_t *pg = a.perms + (size_t)(s * a.M + m) * N * R + r;
            if (STAGED) {
                // this replica's permutation, used by every slice of the
                // sweep, into its LDS column: loads in groups of 8, clamped so
                // that none is conditional
                for (int t0 = 0; t0 < N; t0 += 8) {
                    uint16_t v[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) v[q] = pg[(size_t)min(t0 + q, N - 1) * R];
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (t0 + q < N) lperm[(size_t)(t0 + q) * Ls + lane] = v[q];
                }
            }
            for (int k = 0; k < nslice; ++k) {
                uint32_t *reg = SA ? cur : (k == 1 ? s1 : (k == P - 1 ? sp : cur));
                const bool own = !SA && reg == cur;
                if (own) load(cur, k);
                // this slot's word 0 in the slice's region
                const char *regb = reinterpret_cast<const char *>(reg + lane);
                const uint32_t rega = lds_addr(reg + lane);
                float e = 0.0f;
                int inext = STAGED ? (int)lperm[lane] : (int)pg[0];
                // the first four row entries of the next attempt's spin, read
                // one attempt ahead (the staged rows are read-only)
                int2 rwn[4];
                if constexpr (STAGED && !F64J) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) rwn[q] = rows[__umul24((uint32_t)inext, (uint32_t)Zr) + q];
                }
                for (int t = 0; t < N; ++t) {
                    const int i = inext;  // the next position's index is read one attempt ahead
                    const int tn = min(t + 1, N - 1);
                    inext = STAGED ? (int)lperm[((uint32_t)tn << lsh) + lane] : (int)pg[(size_t)tn * R];
                    // every LDS word this attempt needs that does not depend
                    // on another read is read first, together
                    const uint32_t wo = ((uint32_t)i >> 5) << lsh;
                    uint32_t *wself = reg + wo + lane;
                    const uint32_t ws = *wself;
                    const uint32_t wl = SA ? 0u : sp[wo + lane];  // slice P-1
                    const uint32_t wr = SA ? 0u : s1[wo + lane];  // slice 1
                    const uint32_t rpi = f >= 3 ? f - 3 : f + 28;
                    const uint32_t rf = ring[(f << lsh) + lane], rr = ring[(rpi << lsh) + lane];
                    const int2 *row = rows + __umul24((uint32_t)i, (uint32_t)Zr);
                    int2 rw[4];
                    if constexpr (STAGED && !F64J) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) rw[q] = rwn[q];
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    const uint32_t bsh = (uint32_t)i & 31u;
                    const uint32_t ish = ~(uint32_t)i;     // low five bits: 31 - bsh
                    const uint32_t xi = (ws >> bsh) & 1u;  // 1: s_i = -1
                    const float m2 = xi ? 2.0f : -2.0f;    // -2*s_i
                    if (SA) e = 0.0f;
                    if constexpr (F64J) {
                        // *_dense: J is read as double, so each term is
                        // added in double and the sum rounded to float
                        // (piqmc/qmc.pyx:209-218, piqmc/sa.pyx:172-179)
                        const double *jrow = a.jd + (size_t)i * Z;
                        for (int z = 0; z < Z; ++z) {
                            const int j = row[z].x;
                            const double J = jrow[z];
                            const double t2 = j == i ? (double)m2 * J
                                                     : (double)m2 * (getbit(reg, Ls, lane, j) ? -J : J);
                            e = (float)((double)e + t2);
                        }
                    } else if constexpr (STAGED) {
                        // Every term carries the factor s_i, so the sum runs
                        // on s_i * e: IEEE addition is symmetric under
                        // negation (fl(-a - b) = -fl(a + b)), so negating the
                        // running sum, adding the terms without s_i and
                        // negating back gives the reference's float sum bit
                        // for bit.  Entry (off | sh, -2J): the term is -2J,
                        // sign flipped when s_j = -1 (self entries and
                        // padding read the zero row).
                        const uint32_t xs = xi << 31;
                        e = __uint_as_float(__float_as_uint(e) ^ xs);
                        for (int z = 0; z < Zr; z += 4) {
                            if (z) {
#pragma unroll
                                for (int q = 0; q < 4; ++q) rw[q] = row[z + q];
                            }
                            uint32_t wn[4];
#pragma unroll
                            for (int q = 0; q < 4; ++q)
                                wn[q] = ORA ? lds_word(((uint32_t)rw[q].x & ~31u) | rega)
                                            : *reinterpret_cast<const uint32_t *>(regb + ((uint32_t)rw[q].x & ~31u));
                            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                const uint32_t tb = wn[q] << ((uint32_t)rw[q].x & 31u);  // s_j's bit at bit 31
                                e = e + __uint_as_float((uint32_t)rw[q].y ^ (tb & 0x80000000u));
                            }
                        }
                        if constexpr (!SA) {
                            // -2 s_i (J_perp s_l) / s_i = -2 J_perp, sign flipped when s_l = -1
                            e = e + __uint_as_float(svm2 ^ ((wl << (ish & 31u)) & 0x80000000u));
                            e = e + __uint_as_float(svm2 ^ ((wr << (ish & 31u)) & 0x80000000u));
                        }
                        e = __uint_as_float(__float_as_uint(e) ^ xs);
                    } else {
                        for (int z = 0; z < Z; ++z) {
                            const int2 r1 = row[z];
                            const float J = __int_as_float(r1.y);
                            if (r1.x == i) {
                                e = e + m2 * J;
                            } else {
                                const float sj = getbit(reg, Ls, lane, r1.x) ? -1.0f : 1.0f;
                                e = e + m2 * (J * sj);
                            }
                        }
                    }
                    bool flip;
                    if constexpr (!SA) {
                        if constexpr (F64J || !STAGED) {
                            // -2 s_i (J_perp s_l) = -2 J_perp, sign flipped when s_i != s_l
                            e = e + __uint_as_float(svm2 ^ ((xi ^ ((wl >> bsh) & 1u)) << 31));
                            e = e + __uint_as_float(svm2 ^ ((xi ^ ((wr >> bsh) & 1u)) << 31));
                        }
                        flip = e > 0.0f;
                    } else {
                        // sa.Anneal accepts ediff >= 0 (piqmc/sa.pyx:114),
                        // Anneal_dense ediff > 0 (piqmc/sa.pyx:181)
                        flip = F64J ? e > 0.0f : e >= 0.0f;
                    }
                    if (!flip) {
                        const uint32_t v = rf + rr;
                        ring[(f << lsh) + lane] = v;
                        f = f == 30 ? 0 : f + 1;
                        ++nd;
                        const uint32_t x = v >> 1;
                        // exp(e/T) <= exp(-22) < 2.8e-10 < 1/2147483647 <= u
                        // whenever e < -22.5 T (T > 0, x > 0): rejected without exp
                        const bool far = tdiv > 0.0f && tdiv < INFINITY && e < -22.5f * tdiv && x != 0u;
                        if (!far) {
                            const double u = (double)x / 2147483647.0;
                            const float q = e / tdiv;
                            const double ex = exp((double)q);
                            flip = ex > u;
                            if (ex > 0.0 && ex < INFINITY && fabs(ex - u) <= ex * 0x1p-50) ++nt;
                        }
                    }
                    if constexpr (STAGED && !F64J) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) rwn[q] = rows[__umul24((uint32_t)inext, (uint32_t)Zr) + q];
                    }
                    if (flip) *wself = ws ^ (1u << bsh);
                }
                if (own) store(cur, k);
            }
        }
    }
    if (SA) {
        store(cur, 0);
    } else {
        store(s1, 1);
        if (P != 2) store(sp, P - 1);
    }
    for (int w = 0; w < 31; ++w) a.rstate[(size_t)r * 32 + w] = ring[w * Ls + lane];
    a.rstate[(size_t)r * 32 + 31] = f;
    a.draws[r] = nd;
    a.ties[r] = nt;
}

// int8 [R][P][N] -> bits [P][Wn][R] (bit set = spin -1)
__global__ void pack_lanes_kernel(const int8_t *confs, uint32_t *bits, int R, int P, int N, int Wn)
{
    const size_t tot = (size_t)R * P * Wn;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < tot;
         x += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(x % R);
        const size_t kw = x / R;
        const int w = (int)(kw % Wn), k = (int)(kw / Wn);
        const int8_t *src = confs + ((size_t)r * P + k) * N;
        uint32_t word = 0;
        for (int b = 0; b < 32; ++b) {
            const int i = w * 32 + b;
            if (i < N && src[i] < 0) word |= 1u << b;
        }
        bits[x] = word;
    }
}

__global__ void unpack_lanes_kernel(const uint32_t *bits, int8_t *confs, int R, int P, int N, int Wn)
{
    const size_t tot = (size_t)R * P * N;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < tot;
         x += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(x % N);
        const size_t rk = x / N;
        const int k = (int)(rk % P), r = (int)(rk / P);
        const uint32_t w = bits[((size_t)k * Wn + (i >> 5)) * R + r];
        confs[x] = ((w >> (i & 31)) & 1u) ? (int8_t)-1 : (int8_t)1;
    }
}

// Flags any permutation entry outside [0, N) (the sweep indexes LDS and the
// neighbour table with it).
__global__ void check_perms_kernel(const uint16_t *perms, size_t n, int N, int *bad)
{
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n;
         x += (size_t)gridDim.x * blockDim.x)
        if ((int)perms[x] >= N) atomicOr(bad, 1);
}

int grid_for(size_t n, int block)
{
    size_t g = (n + block - 1) / block;
    return (int)(g > 65535 ? 65535 : (g < 1 ? 1 : g));
}

int run_replay(const piqmc_problem *pb, int R, int P, int S, int M, float temp, int sa,
               const float *steps_host, int8_t *confs, const uint16_t *perms, uint32_t *rand,
               uint64_t *draws, uint32_t *ties, hipStream_t st)
{
    const int N = pb->N, Wn = (N + 31) / 32;
    const int nreg = sa ? 1 : (P == 2 ? 2 : 3);
    // Launch shape: W waves per workgroup share one LDS copy of the neighbour
    // rows; Lw lanes per wave each own a replica.  Lw is the smallest power of
    // two that still gives every CU two workgroups, so replicas spread over all
    // SIMDs with two chains' latency to hide per SIMD (4096 replicas of the
    // 1024-spin lattice: 8.4e9 attempts/s at 2 lanes per wave, 5.2e9 at 4, 3.8e9
    // at 1).  The staged form (rows and the sweep's permutations in LDS) is
    // preferred; shapes shrink until it fits, else rows and permutations are
    // read from global memory.
    const int Zp = (pb->Z + 3) & ~3;
    auto lds_of = [&](int slots, bool staged) {
        const int Ls = replay_stride(slots);
        const size_t words = replay_words(nreg, Wn, Ls) * 4;
        return staged ? words + (size_t)N * Zp * sizeof(int2) + (size_t)N * Ls * 2 : words;
    };
    constexpr size_t kLds = 160 * 1024;
    constexpr int kCUs = 256;
    int W0 = 4, Lw0 = 64;
    while (Lw0 > 1 && (size_t)W0 * Lw0 * kCUs * 2 > (size_t)R) Lw0 /= 2;
    int W = W0, Lw = Lw0;
    bool staged = true;
    while (lds_of(W * Lw, true) > kLds && (W > 1 || Lw > 1)) {
        if (Lw > 1) Lw /= 2; else W /= 2;
    }
    if (lds_of(W * Lw, true) > kLds) {
        staged = false;
        W = W0;
        Lw = Lw0;
        while (lds_of(W * Lw, false) > kLds && (W > 1 || Lw > 1)) {
            if (Lw > 1) Lw /= 2; else W /= 2;
        }
    }
    const int L = W * Lw;
    if (lds_of(L, staged) > kLds)
        return piqmc::set_error(PIQMC_EINVAL, "replay: nspins too large for one workgroup's LDS");
    auto lds = [&](int l) { return lds_of(l, staged); };
    const size_t nperm = (size_t)S * M * N * R;
    if (nperm && N < 65536) {
        const int rc = piqmc::check_perms(perms, nperm, N, st);
        if (rc) return rc;
    }
    float *d_steps = nullptr;
    uint32_t *d_bits = nullptr;
    unsigned long long *d_draws = reinterpret_cast<unsigned long long *>(draws);
    PIQMC_HIP(hipMallocAsync((void **)&d_steps, sizeof(float) * (S > 0 ? S : 1), st));
    PIQMC_HIP(hipMallocAsync((void **)&d_bits, sizeof(uint32_t) * (size_t)P * Wn * R, st));
    if (S > 0) PIQMC_HIP(hipMemcpyAsync(d_steps, steps_host, sizeof(float) * S, hipMemcpyHostToDevice, st));
    pack_lanes_kernel<<<grid_for((size_t)R * P * Wn, 256), 256, 0, st>>>(confs, d_bits, R, P, N, Wn);
    ReplayArgs a;
    a.R = R; a.N = N; a.P = P; a.Z = pb->Z; a.S = S; a.M = M; a.Wn = Wn; a.sa = sa;
    a.Zp = Zp;
    a.Lw = Lw;
    a.temp = temp;
    a.steps = d_steps;
    a.rows = pb->d_rows;
    a.perms = perms;
    a.bits = d_bits;
    a.rstate = rand;
    a.draws = d_draws;
    a.ties = ties;
    a.jd = pb->dense ? pb->d_jd : nullptr;
    // the AND-OR neighbour address needs slot offsets below 32 bytes
    const bool ora = staged && !pb->dense && L <= 8;
    auto pick = [&](auto SAc) {
        constexpr bool S_ = decltype(SAc)::value;
        return pb->dense ? (staged ? replay_kernel<true, true, S_> : replay_kernel<true, false, S_>)
                         : (staged ? (ora ? replay_kernel<false, true, S_, true> : replay_kernel<false, true, S_>)
                                   : replay_kernel<false, false, S_>);
    };
    auto kern = sa ? pick(std::true_type()) : pick(std::false_type());
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds(L));
    piqmc::timing_begin(st);
    kern<<<(R + L - 1) / L, 64 * W, lds(L), st>>>(a);
    piqmc::timing_end(st);
    PIQMC_HIP(hipGetLastError());
    unpack_lanes_kernel<<<grid_for((size_t)R * P * N, 256), 256, 0, st>>>(d_bits, confs, R, P, N, Wn);
    PIQMC_HIP(hipFreeAsync(d_bits, st));
    PIQMC_HIP(hipFreeAsync(d_steps, st));
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

}  // namespace

namespace piqmc {
int check_perms(const uint16_t *perms, size_t n, int N, hipStream_t st)
{
    int *d_bad = nullptr, bad = 0;
    PIQMC_HIP(hipMallocAsync((void **)&d_bad, sizeof(int), st));
    PIQMC_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), st));
    check_perms_kernel<<<grid_for(n, 256), 256, 0, st>>>(perms, n, N, d_bad);
    PIQMC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st));
    PIQMC_HIP(hipFreeAsync(d_bad, st));
    PIQMC_HIP(hipStreamSynchronize(st));
    if (bad) return set_error(PIQMC_EINVAL, "permutation entry out of range [0, nspins)");
    return PIQMC_OK;
}
}  // namespace piqmc

extern "C" int piqmc_qa_replay(const piqmc_problem *pb, int R, int P, int S, int M, float temp,
                               const double *sched, int8_t *confs, const uint16_t *perms,
                               uint32_t *rand, uint64_t *draws, uint32_t *ties, void *stream)
{
    if (!pb || R < 1 || S < 0 || M < 0 || (S > 0 && !sched) || !confs || !rand || !draws || !ties)
        return piqmc::set_error(PIQMC_EINVAL, "qa_replay: bad arguments");
    if (P < 2) return piqmc::set_error(PIQMC_EINVAL, "QuantumAnneal needs slices >= 2 (reads slice 1)");
    if (pb->N > 65536) return piqmc::set_error(PIQMC_EINVAL, "replay supports nspins <= 65536");
    if ((long)S * M > 0 && !perms) return piqmc::set_error(PIQMC_EINVAL, "qa_replay: null perms");
    std::vector<float> jp(S > 0 ? S : 0);
    for (int s = 0; s < S; ++s) {
        if ((float)P * temp == 0.0f) return piqmc::set_error(PIQMC_EZERODIV, "float division");
        jp[s] = host_jperp(sched[s], P, temp);
    }
    if ((long)S * M > 0 && temp == 0.0f) return piqmc::set_error(PIQMC_EZERODIV, "float division");
    return run_replay(pb, R, P, S, M, temp, 0, jp.data(), confs, perms, rand, draws, ties,
                      (hipStream_t)stream);
}

extern "C" int piqmc_sa_replay(const piqmc_problem *pb, int R, int S, int M, const double *sched,
                               int8_t *svec, const uint16_t *perms, uint32_t *rand, uint64_t *draws,
                               uint32_t *ties, void *stream)
{
    if (!pb || R < 1 || S < 0 || M < 0 || (S > 0 && !sched) || !svec || !rand || !draws || !ties)
        return piqmc::set_error(PIQMC_EINVAL, "sa_replay: bad arguments");
    if (pb->N > 65536) return piqmc::set_error(PIQMC_EINVAL, "replay supports nspins <= 65536");
    if ((long)S * M > 0 && !perms) return piqmc::set_error(PIQMC_EINVAL, "sa_replay: null perms");
    std::vector<float> tv(S > 0 ? S : 0);
    for (int s = 0; s < S; ++s) tv[s] = (float)sched[s];  // piqmc/sa.pyx:96
    return run_replay(pb, R, 1, S, M, 0.0f, 1, tv.data(), svec, perms, rand, draws, ties,
                      (hipStream_t)stream);
}

// ------------------------------------------------------------------ host drop-ins
namespace {
template <typename T>
struct DevBuf {
    T *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc((void **)&p, sizeof(T) * (n ? n : 1)); }
};

// Couplings of a drop-in call: a neighbour table (QuantumAnneal, Anneal) or a
// dense matrix with element strides (QuantumAnneal_dense, Anneal_dense).
struct Couplings {
    const double *nbs = nullptr;
    int Z = 0;
    const double *J = nullptr;
    int64_t js0 = 0, js1 = 0;
};

int dropin(const double *sched, int S, int M, int P, float temp, int N, double *confs, int64_t cs0,
           int64_t cs1, const Couplings &cp, const int64_t *perms, uint32_t *rs,
           uint64_t *draws_out, uint32_t *ties_out, int sa)
{
    if (N < 1 || S < 0 || M < 0 || !confs || !rs || (cp.J ? false : (!cp.nbs || cp.Z < 1)))
        return piqmc::set_error(PIQMC_EINVAL, "bad arguments");
    if (!sa && P < 2) return piqmc::set_error(PIQMC_EINVAL, "QuantumAnneal needs slices >= 2 (reads slice 1)");
    if (rs[31] > 30) return piqmc::set_error(PIQMC_EINVAL, "bad TYPE_3 state");
    const size_t G = (size_t)S * M;
    std::vector<int8_t> c8((size_t)P * N);
    for (int k = 0; k < P; ++k)
        for (int i = 0; i < N; ++i) {
            const double v = confs[i * cs0 + k * cs1];
            if (v != 1.0 && v != -1.0)
                return piqmc::set_error(PIQMC_EINVAL, "spin configurations must hold +1/-1 values");
            c8[(size_t)k * N + i] = v < 0 ? -1 : 1;
        }
    std::vector<uint16_t> p16(G * N);
    for (size_t g = 0; g < G; ++g)
        for (int t = 0; t < N; ++t) {
            const int64_t v = perms[g * N + t];
            if (v < 0 || v >= N) return piqmc::set_error(PIQMC_EINVAL, "permutation entry out of range");
            p16[g * N + t] = (uint16_t)v;
        }
    piqmc_problem *pb = nullptr;
    int rc = cp.J ? piqmc_problem_create_dense(N, cp.J, cp.js0, cp.js1, &pb)
                  : piqmc_problem_create(N, cp.Z, cp.nbs, &pb);
    if (rc) return rc;
    struct Guard {
        piqmc_problem *p;
        ~Guard() { piqmc_problem_destroy(p); }
    } guard{pb};
    DevBuf<int8_t> dc;
    DevBuf<uint16_t> dp;
    DevBuf<uint32_t> dr, dt;
    DevBuf<uint64_t> dd;
    PIQMC_HIP(dc.alloc(c8.size()));
    PIQMC_HIP(dp.alloc(p16.size()));
    PIQMC_HIP(dr.alloc(32));
    PIQMC_HIP(dt.alloc(1));
    PIQMC_HIP(dd.alloc(1));
    PIQMC_HIP(hipMemcpy(dc.p, c8.data(), c8.size(), hipMemcpyHostToDevice));
    if (!p16.empty()) PIQMC_HIP(hipMemcpy(dp.p, p16.data(), p16.size() * 2, hipMemcpyHostToDevice));
    PIQMC_HIP(hipMemcpy(dr.p, rs, 32 * 4, hipMemcpyHostToDevice));
    rc = sa ? piqmc_sa_replay(pb, 1, S, M, sched, dc.p, dp.p, dr.p, dd.p, dt.p, nullptr)
            : piqmc_qa_replay(pb, 1, P, S, M, temp, sched, dc.p, dp.p, dr.p, dd.p, dt.p, nullptr);
    if (rc) return rc;
    PIQMC_HIP(hipDeviceSynchronize());
    uint64_t nd = 0;
    uint32_t nt = 0;
    PIQMC_HIP(hipMemcpy(c8.data(), dc.p, c8.size(), hipMemcpyDeviceToHost));
    PIQMC_HIP(hipMemcpy(rs, dr.p, 32 * 4, hipMemcpyDeviceToHost));
    PIQMC_HIP(hipMemcpy(&nd, dd.p, 8, hipMemcpyDeviceToHost));
    PIQMC_HIP(hipMemcpy(&nt, dt.p, 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < P; ++k)
        for (int i = 0; i < N; ++i) confs[i * cs0 + k * cs1] = (double)c8[(size_t)k * N + i];
    if (draws_out) *draws_out = nd;
    if (ties_out) *ties_out = nt;
    return PIQMC_OK;
}
}  // namespace

extern "C" int piqmc_quantum_anneal(const double *sched, int S, int mcsteps, int slices, float temp,
                                    int nspins, double *confs, int64_t cs0, int64_t cs1,
                                    const double *nbs, int maxnb, const int64_t *perms,
                                    uint32_t *rs, uint64_t *draws, uint32_t *ties)
{
    Couplings cp;
    cp.nbs = nbs;
    cp.Z = maxnb;
    return dropin(sched, S, mcsteps, slices, temp, nspins, confs, cs0, cs1, cp, perms, rs, draws,
                  ties, 0);
}

extern "C" int piqmc_quantum_anneal_dense(const double *sched, int S, int mcsteps, int slices,
                                          float temp, int nspins, double *confs, int64_t cs0,
                                          int64_t cs1, const double *J, int64_t js0, int64_t js1,
                                          const int64_t *perms, uint32_t *rs, uint64_t *draws,
                                          uint32_t *ties)
{
    if (!J) return piqmc::set_error(PIQMC_EINVAL, "null coupling matrix");
    Couplings cp;
    cp.J = J;
    cp.js0 = js0;
    cp.js1 = js1;
    return dropin(sched, S, mcsteps, slices, temp, nspins, confs, cs0, cs1, cp, perms, rs, draws,
                  ties, 0);
}

extern "C" int piqmc_anneal(const double *sched, int S, int mcsteps, double *svec, int64_t ss,
                            int nspins, const double *nbs, int maxnb, const int64_t *perms,
                            uint32_t *rs, uint64_t *draws, uint32_t *ties)
{
    Couplings cp;
    cp.nbs = nbs;
    cp.Z = maxnb;
    return dropin(sched, S, mcsteps, 1, 0.0f, nspins, svec, ss, 0, cp, perms, rs, draws, ties, 1);
}

extern "C" int piqmc_anneal_dense(const double *sched, int S, int mcsteps, double *svec, int64_t ss,
                                  int nspins, const double *J, int64_t js0, int64_t js1,
                                  const int64_t *perms, uint32_t *rs, uint64_t *draws, uint32_t *ties)
{
    if (!J) return piqmc::set_error(PIQMC_EINVAL, "null coupling matrix");
    Couplings cp;
    cp.J = J;
    cp.js0 = js0;
    cp.js1 = js1;
    return dropin(sched, S, mcsteps, 1, 0.0f, nspins, svec, ss, 0, cp, perms, rs, draws, ties, 1);
}
