// Parallel sweep in "fixed" semantics (periodic Trotter neighbours k-1, k+1,
// per-spin reset): graph-colored in space, parity-alternated in imaginary
// time.  It is the parallel counterpart of qmc.QuantumAnneal_parallel
// (piqmc/qmc.pyx:247-357) with the races removed: a phase (tau, c) updates
// every spin of spatial color c in every slice of parity tau, and no two such
// spins interact.
//
// Layout.  One workgroup owns one replica.  Its P slices live in LDS
// bit-packed by parity: word st[tau][i] holds spin i of slices tau, tau+2, ...
// (bit m = slice 2m+tau), so a 32-/64-bit word carries P/2 attempts that share
// their couplings.  The Trotter neighbours of those slices are the other
// parity word rotated by one bit.
//
// +-1 couplings (every nonzero coupling J=+-1 between distinct spins) run a
// bit-sliced path: unsatisfied-bond counts are added with bitwise full adders
// over the whole word, the (z, u_spatial, u_trotter) class of each bit selects
// always/never/stochastic acceptance from a per-field-step table computed on
// the host with piqmc_fexp64, and Philox4x32-10 is evaluated only for the bits
// whose class is stochastic.  Other couplings run a per-bit path with the same
// double-precision arithmetic as the CPU oracle.  Both paths take exactly the
// decisions of piqmc_fixed_class with the draw
//   philox(key=seed, ctr={sweep, spin, replica, tau<<24 | m>>2})[m&3]
// for bit m, so the result does not depend on workgroup size or scheduling.
#include <math.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "piqmc_internal.h"

namespace {

constexpr int kClassStride = 8 * 8 * 3;  // (z, u_s, u_t) thresholds per field step

// Device scratch of the allocation-free parallel sweep (piqmc_qa_colored_noalloc):
// one copy per GPU, held for the life of the process.  Calls on one GPU share
// it, so calls that overlap on one GPU race.  Every buffer carved from it is
// fully written before it is read (tables by copies, the packed state by the
// pack kernel), so nothing in it needs clearing between calls.
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kScratchMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
unsigned char *g_scratch_ptr[kScratchMaxDevices] = {};
std::mutex g_scratch_mu;
// Stochastic-word queues of the +-1 sweep, per wavefront: with several
// classes one queue of 16-byte records (index | replica << 16 | parity << 31,
// three class masks); with a single class a queue of 8-byte records (index |
// replica << 16 | parity << 31, mask) and a second one of 16-byte records for
// words still undecided after their first Philox block (the same, then
// accepted bits, undecided bits).  Fresh records are drained 64 at a time once
// 64 are queued, after every replica's append (at most 63 + 64 queued) or, in
// the fixed-plane sweep, after both replicas' (at most 63 + 2 x 64); the
// second queue holds at most 63 + 64.
constexpr int kQueueCap1 = 192, kQueueCap2 = 128;
constexpr int kQueueBytes = 8 * kQueueCap1 + 16 * kQueueCap2;  // per wavefront
// Geometry of the fixed-plane cached sweep (see colored_body's FIXED).
constexpr int kFixRep = 2, kFixPlane = 1024;

// LDS of the word-per-thread sweep: packed state [rpw][2][N] words, class
// thresholds [kClassStride], per-degree table [8], reduction scratch [16]
// doubles, then (16-byte aligned) the wavefronts' stochastic queues.
__host__ __device__ constexpr inline size_t queue_offset(int N, int rpw, int word_bytes)
{
    return ((size_t)word_bytes * 2 * (size_t)N * rpw + 4 * (kClassStride + 8) + 16 * sizeof(double) + 15) &
           ~(size_t)15;
}

template <typename WT>
struct ColArgs {
    int N, Z, P, W, S, M, ncolors, zmax, R, rpw, zuni;
    uint32_t key0, key1, replica_base, sweep_base;
    const float *jperp;        // [S]
    const float *temps;        // [S] temperature of each field step
    const int2 *rows;          // [N][Z]
    const int32_t *order;      // [N] spins by color
    const int *color_off;      // [ncolors+1]
    const int32_t *cnb;        // [zmax][N] (+-1 path)
    const uint32_t *cinfo;     // [N]
    const float *hq;           // [N] field of spin order[q] (+-1 couplings with fields)
    const double *prevj;       // [N] coupling of spin order[q] to order[q-1] (dense path)
    const uint32_t *thr;       // [S][kClassStride]
    const uint32_t *ztab;      // [S][8] per degree z: u_s thresholds (see run_colored)
    WT *packed;                // [R][2][N]
    double *etrace;            // [R][S] or null
    double *fields;            // [R][2][W][N|1] local fields (cached-field dense sweep)
};

__device__ __forceinline__ int popc(uint32_t x) { return __popc(x); }
__device__ __forceinline__ int popc(uint64_t x) { return __popcll(x); }
__device__ __forceinline__ int ctz(uint32_t x) { return __ffs(x) - 1; }
__device__ __forceinline__ int ctz(uint64_t x) { return __ffsll((unsigned long long)x) - 1; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of one value per thread; `red` holds one slot per wave.
template <typename T>
__device__ T block_sum(T v, T *red)
{
    v = wave_sum(v);
    const int wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) red[wid] = v;
    __syncthreads();
    T t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < nw; ++w) t += red[w];
    return t;
}

// Three-input XOR and majority of bit-sliced words (one gfx950 v_bitop3_b32
// each for 32-bit words: truth tables 0x96 and 0xE8, symmetric in the inputs).
template <typename WT>
__device__ __forceinline__ WT xor3(WT a, WT b, WT c)
{
    if constexpr (sizeof(WT) == 4) return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
    else return a ^ b ^ c;
}
template <typename WT>
__device__ __forceinline__ WT maj3(WT a, WT b, WT c)
{
    if constexpr (sizeof(WT) == 4) return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);
    else return (a & b) | (c & (a ^ b));
}

// t-th bit-sliced random word of a Philox4x32 output block: four 32-bit words
// for 32-bit spin words, two 64-bit words for 64-bit spin words.
template <typename WT>
__device__ __forceinline__ WT rword(const uint32_t out[4], int t)
{
    if constexpr (sizeof(WT) == 4) return out[t];
    else return (WT)out[2 * t] | ((WT)out[2 * t + 1] << 32);
}

// Smallest-count masks of a bit-sliced counter (c2 c1 c0): bit set where the
// count is >= k.  k is uniform across the wavefront on regular lattices.
template <typename WT>
__device__ __forceinline__ WT ge_pick(int k, WT c0, WT c1, WT c2)
{
    switch (k) {
    case 0: return ~WT(0);
    case 1: return c0 | c1 | c2;
    case 2: return c1 | c2;
    case 3: return c2 | (c1 & c0);
    case 4: return c2;
    case 5: return c2 & (c1 | c0);
    case 6: return c2 & c1;
    case 7: return c2 & c1 & c0;
    default: return WT(0);
    }
}

// Bit `bit` of a packed per-lane word as an all-ones or zero mask (a signed
// one-bit field extract).
template <typename WT>
__device__ __forceinline__ WT bitmask(uint32_t pk, int bit)
{
    return (WT)(int64_t)((int32_t)(pk << (31 - bit)) >> 31);  // 0 or all ones
}

// Decision classes of one spin with a field h in one field step, for +-1
// couplings: a bit's class depends only on its spin's sign sigma, its
// unsatisfied-bond count u_s (0..z) and antiparallel Trotter bonds u_t, via
//   ed = -2 * fl((sigma*h + (z - 2 u_s)) + J_perp (2 - 2 u_t)),
// which is exactly the general path's -2 s_i (h_i + sum J s_j + J_perp ts) when
// the row's partial sums are exact (fields 0 or 2^-27 <= |h| <= 2^20, the
// problem-level condition).  Classes are non-decreasing in u_s, so each
// (sigma, u_t) reduces to two thresholds; the first three stochastic
// (sigma, u_t, u_s) triples keep their draw thresholds.
struct FieldClasses {
    uint32_t ka, ks;   // nibble sigma*3 + u_t: accept from u_s >= ka, stochastic from ks
    int ns;            // stochastic triples
    uint32_t slot;     // byte t: combo (bits 0-2) and u_s (bits 4-7) of triple t < 3
    uint32_t st[3];    // their draw thresholds
};

__device__ __forceinline__ int field_class(double hs, int z, int us, double B, double T, unsigned int *thr)
{
    const double X = (hs + (double)(z - 2 * us)) + B;
    const double ed = -2.0 * X;
    if (ed > 0.0) return 1;
    if (T > 0.0 && ed < -23.5 * T) return 0;  // exp(ed/T) < e^-23 < 2^-33: never accepted
    return piqmc_fixed_class(ed, T, thr);
}

__device__ __forceinline__ void field_classes(double h, int z, float jp, double T, FieldClasses &fc)
{
    fc.ns = 0;
    fc.ka = fc.ks = fc.slot = 0;
    fc.st[0] = fc.st[1] = fc.st[2] = 0;
    for (int cb = 0; cb < 6; ++cb) {
        const int sg = cb / 3, ut = cb % 3;
        const double hs = sg ? -h : h;
        const double B = (double)jp * (double)(2 - 2 * ut);
        // Accepted counts are those with X(u_s) = (hs + (z - 2 u_s)) + B < 0, and X
        // does not increase with u_s: the boundary is estimated in closed form
        // and then checked on the exact expression (rounding moves it by one
        // at most), instead of walking down from u_s = z.
        auto Xat = [&](int u) { return (hs + (double)(z - 2 * u)) + B; };
        const double est = fmin(fmax((hs + B + (double)z) * 0.5, -1.0), (double)z + 1.0);
        int amin = min(max((int)floor(est) + 1, 0), z + 1);
        while (amin <= z && !(Xat(amin) < 0.0)) ++amin;
        while (amin > 0 && Xat(amin - 1) < 0.0) --amin;
        int smin, us = amin - 1;
        unsigned int thr = 0;
        int cls = us >= 0 ? field_class(hs, z, us, B, T, &thr) : 1;
        smin = amin;
        for (; us >= 0 && cls == 2; --us) {
            smin = us;
            if (fc.ns < 3) {
                fc.slot |= (uint32_t)(cb | (us << 4)) << (8 * fc.ns);
                if (fc.ns == 0) fc.st[0] = thr;
                else if (fc.ns == 1) fc.st[1] = thr;
                else fc.st[2] = thr;
            }
            ++fc.ns;
            if (us > 0) cls = field_class(hs, z, us - 1, B, T, &thr);
        }
        fc.ka |= (uint32_t)amin << (4 * cb);
        fc.ks |= (uint32_t)smin << (4 * cb);
    }
}

// Replicas r0 .. r0+nrep-1 of one problem, whole schedule, by one workgroup.
// UZ: every spin has the same degree a.zuni, so the class thresholds of a
// field step are wavefront-uniform and the mask selection is scalar code.
// ZU > 0: the uniform degree is known at compile time (4 for square, 6 for
// cubic lattices), so the neighbour loops have fixed trip counts.
// CACHED (with ZU > 0): two colors of at most two spins per thread each; every
// thread keeps its spins' indices and neighbour lists in registers for the
// whole schedule instead of reloading them in every phase.
// Rotation of a parity word's W slices left by tsa (right by tsb = W - tsa):
// the Trotter neighbours of the other parity.  32-bit words take one funnel
// shift: (O : O << rls) >> rsh keeps (O << tsa) | (O >> (W - tsa)) in its
// low W bits (rls = 32 - W, rsh = 32 - tsa; W = 1 with tsa = 0 is O itself,
// rls = rsh = 0).  Bits past W are masked by the caller.
template <typename WT>
__device__ inline WT rot_slices(WT O, int tsa, int tsb, int, int)
{
    return (O << tsa) | (O >> tsb);
}
template <>
__device__ inline uint32_t rot_slices<uint32_t>(uint32_t O, int, int, int rls, int rsh)
{
    return __builtin_amdgcn_alignbit(O, O << rls, (uint32_t)rsh);
}

// FIXED (cached +-1 queue sweep only): exactly kFixRep replicas per workgroup,
// each (replica, parity) plane kFixPlane words apart whatever N (N <= kFixPlane),
// and every thread holding a spin of each color in both of its slots.  All
// plane offsets are then compile-time, so every LDS access of the word body
// takes its plane as the instruction's immediate offset, and no lane is
// masked off.
template <typename WT, bool PMJ, bool UZ = false, bool FLD = false, int ZU = 0, bool CACHED = false, bool CQ = false,
          bool FIXED = false>
__device__ void colored_body(const ColArgs<WT> &a, int r0, int nrep_, unsigned char *smem)
{
    static_assert(!CQ || (UZ && PMJ && !FLD && sizeof(WT) == 4), "stochastic queue: +-1 uniform-degree 32-bit path");
    static_assert(!FIXED || (CACHED && CQ), "fixed planes: cached queue sweep");
    const int N = a.N, W = a.W, tid = threadIdx.x, bd = blockDim.x;
    const int nrep = FIXED ? kFixRep : nrep_;
    const int PW = FIXED ? kFixPlane : N;                             // words per (replica, parity) plane
    const int rpw = FIXED ? kFixRep : a.rpw;
    WT *st = reinterpret_cast<WT *>(smem);                            // [nrep][2][PW]
    uint32_t *tthr = reinterpret_cast<uint32_t *>(st + (size_t)2 * PW * rpw);  // [kClassStride]
    uint32_t *tz = tthr + kClassStride;                               // [8]
    double *red = reinterpret_cast<double *>(tz + 8);                 // [16]
    // per-wavefront queues of stochastic words (CQ), kQueueBytes each
    unsigned char *wqb = smem + queue_offset(PW, rpw, sizeof(WT)) + (size_t)(tid >> 6) * kQueueBytes;
    uint4 *wq = reinterpret_cast<uint4 *>(wqb);           // several classes
    uint2 *wq1 = reinterpret_cast<uint2 *>(wqb);          // single class: fresh words
    uint4 *wq2 = reinterpret_cast<uint4 *>(wqb + 8 * kQueueCap1);  // single class: undecided after block 0
    int qn = 0;  // queued records of this wavefront (wave-uniform)
    int q2 = 0;  // records in wq2 (wave-uniform)
    const WT valid = (W == (int)(8 * sizeof(WT))) ? ~WT(0) : ((WT(1) << W) - 1);
    WT *gst = a.packed + (size_t)r0 * 2 * N;
    if constexpr (FIXED) {
        for (int q = tid; q < 2 * N * nrep; q += bd) st[(q / N) * PW + q % N] = gst[q];
    } else {
        for (int q = tid; q < 2 * N * nrep; q += bd) st[q] = gst[q];
    }
    constexpr int ZC = ZU > 0 ? ZU : 1;
    int cq_[2][2], cI[2][2], cN[2][2][ZC];
    uint32_t cF[2][2];
    if (CACHED) {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int q = a.color_off[c] + tid + k * bd;
                const bool in = q < a.color_off[c + 1];
                const int qs = in ? q : 0;
                cq_[c][k] = in ? q : -1;
                cI[c][k] = a.order[qs];
                cF[c][k] = a.cinfo[qs];
#pragma unroll
                for (int z = 0; z < ZC; ++z) cN[c][k][z] = a.cnb[(size_t)z * N + qs] * (int)sizeof(WT);  // bytes
            }
    }

    // Queue records' first word: index | replica << 16 | parity << 31, or in the
    // fixed-plane sweep the word's byte offset in the planes, index << 2 |
    // parity << 12 | replica << 13, so that a drain needs no address arithmetic
    // to write the word back.
    static_assert(kFixPlane * 4 == (1 << 12) && kFixRep == 2, "fixed-plane record layout");
    auto rec_i = [](uint32_t x) __attribute__((always_inline)) -> int {
        if constexpr (FIXED) return (int)((x >> 2) & (uint32_t)(kFixPlane - 1));
        else return (int)(x & 0xffffu);
    };
    auto rec_rr = [](uint32_t x) __attribute__((always_inline)) -> int {
        if constexpr (FIXED) return (int)(x >> 13);
        else return (int)((x >> 16) & 0x7fffu);
    };
    auto rec_tau = [](uint32_t x) __attribute__((always_inline)) -> int {
        if constexpr (FIXED) return (int)((x >> 12) & 1u);
        else return (int)(x >> 31);
    };
    auto rec_word = [&](uint32_t x) __attribute__((always_inline)) -> WT * {
        if constexpr (FIXED) return reinterpret_cast<WT *>(reinterpret_cast<unsigned char *>(st) + x);
        else return &st[((size_t)rec_rr(x) * 2 + rec_tau(x)) * PW + rec_i(x)];
    };
    // A queue drain's accepted bits into their word, as an LDS atomic XOR with no
    // return: a drain holds each word once and no other wavefront writes it in
    // the phase, so this equals a read-modify-write without its round trip.
    auto xor_word = [](WT *w, WT v) __attribute__((always_inline)) {
        if constexpr (sizeof(WT) == 4) atomicXor(reinterpret_cast<unsigned int *>(w), (unsigned int)v);
        else *w ^= v;
    };
    // a neighbour's word: cached neighbours are byte offsets (no shift per word)
    auto nbword = [](const WT *base, int nb) __attribute__((always_inline)) -> WT {
        if constexpr (CACHED) return *reinterpret_cast<const WT *>(reinterpret_cast<const char *>(base) + nb);
        else return base[nb];
    };
    // thresholds clamp at 8 ("count >= 8" is empty).  Nine masks also keep the
    // runtime-indexed selections below as indexed register moves: with six
    // (degree 4 needs no more) the compiler expands each into a chain of
    // selects, and the 1024-spin sweep runs 9.6 -> 12.5 ms.
    constexpr int KMAX = 8;

    for (int s = 0; s < a.S; ++s) {
        __syncthreads();
        if (PMJ && !FLD) {
            for (int q = tid; q < kClassStride; q += bd) tthr[q] = a.thr[(size_t)s * kClassStride + q];
            if (tid < 8) tz[tid] = a.ztab[s * 8 + tid];
        }
        __syncthreads();
        // uniform-degree lattices: the class thresholds of this step are wave-uniform
        const uint32_t zt_u = (PMJ && UZ) ? __builtin_amdgcn_readfirstlane(a.ztab[s * 8 + a.zuni]) : 0u;
        // its six thresholds (accept / stochastic per Trotter class), clamped to the table
        const int ka0 = min((int)(zt_u & 15), KMAX), ks0 = min((int)((zt_u >> 4) & 15), KMAX);
        const int ka1 = min((int)((zt_u >> 8) & 15), KMAX), ks1 = min((int)((zt_u >> 12) & 15), KMAX);
        const int ka2 = min((int)((zt_u >> 16) & 15), KMAX), ks2 = min((int)((zt_u >> 20) & 15), KMAX);
        // and its stochastic (u_t, u_s) classes with their draw thresholds (the
        // first three; more take the table walk)
        int un = 0, ut_s[3] = {0, 0, 0}, us_s[3] = {0, 0, 0};
        uint32_t th_s[3] = {0u, 0u, 0u};
        if (PMJ && UZ && !FLD) {
#pragma unroll
            for (int ut = 0; ut < 3; ++ut) {
                const int lo = ut == 0 ? ks0 : (ut == 1 ? ks1 : ks2), hi = ut == 0 ? ka0 : (ut == 1 ? ka1 : ka2);
                for (int us = lo; us < hi; ++us) {
                    if (un < 3) {
                        const uint32_t th = a.thr[(size_t)s * kClassStride + (a.zuni * 8 + us) * 3 + ut];
                        if (un == 0) { ut_s[0] = ut; us_s[0] = us; th_s[0] = th; }
                        else if (un == 1) { ut_s[1] = ut; us_s[1] = us; th_s[1] = th; }
                        else { ut_s[2] = ut; us_s[2] = us; th_s[2] = th; }
                    }
                    ++un;
                }
            }
        }
        const float jp = a.jperp[s];
        const double Ts = (double)a.temps[s];
        // cached spins with fields: their decision classes of this field step
        FieldClasses cfc[2][2];
        if constexpr (FLD && CACHED) {
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int q = c == 0 ? (k == 0 ? cq_[0][0] : cq_[0][1]) : (k == 0 ? cq_[1][0] : cq_[1][1]);
                    field_classes(q >= 0 ? (double)a.hq[q] : 0.0, ZU, jp, Ts, cfc[c][k]);
                }
        }
        const bool cq = CQ && un <= 3;  // this step's stochastic words go through the queue
        // CQ kernels run one of three copies of the sweep, chosen per step:
        // MODE 0 without stochastic classes (every stochastic branch compiled
        // out), MODE 1 with exactly one (queue records are the stochastic
        // mask), MODE 2 otherwise.
        auto sweeps = [&](auto MODEc) __attribute__((always_inline)) {
            constexpr int MODE = decltype(MODEc)::value;
            constexpr bool DET = MODE == 0;
            // MODE 1, 3, 4: one stochastic class, in Trotter class u_t = 0, 1, 2
            constexpr bool SINGLE = MODE == 1 || MODE == 3 || MODE == 4;
            constexpr int UTC = MODE == 3 ? 1 : (MODE == 4 ? 2 : 0);
            const bool one = SINGLE || (MODE == 2 && un == 1);  // a single stochastic class
            const bool cqm = SINGLE || cq;
            for (int m = 0; m < a.M; ++m) {
                const uint32_t g = a.sweep_base + (uint32_t)(s * a.M + m);
                // The cached sweep runs all of it in one pass of this loop (three
                // phases, below); the general one a pass per Trotter parity.
                for (int tau = 0; tau < (CACHED ? 1 : 2); ++tau) {
                    // The top cnt queued words, one per lane: each word's stochastic
                    // bits meet the same bit-sliced uniforms, thresholds and early
                    // exit as in the inline path, and the accepted ones are XORed
                    // into the word (its deterministic flips are already stored).
                    // Single class, second stage: the top cnt records of wq2 continue
                    // from Philox block 1 with the usual early exit.
                    auto drain2 = [&](const int cnt) __attribute__((always_inline)) {
                        __builtin_amdgcn_wave_barrier();
                        const int lane = tid & 63;
                        if (lane < cnt) {
                            const uint4 rec = wq2[q2 - cnt + lane];
                            const int i = rec_i(rec.x), rr = rec_rr(rec.x);
                            const int tau = rec_tau(rec.x);  // the word's Trotter parity
                            const WT m0 = (WT)rec.y;
                            WT lt = (WT)rec.z, und = (WT)rec.w;
                            const uint32_t rep = a.replica_base + (uint32_t)(r0 + rr);
                            for (int qd = 1; qd < 8 && und; ++qd) {
                                uint32_t out[4];
                                philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qd, a.key0, a.key1, out);
    #pragma unroll
                                for (int t = 0; t < 4; ++t) {
                                    const WT Rb = rword<WT>(out, t);
                                    const WT tb = ((th_s[0] >> (31 - (qd * 4 + t))) & 1u) ? m0 : WT(0);
                                    lt |= und & ~Rb & tb;
                                    und &= ~(Rb ^ tb);
                                }
                            }
                            xor_word(rec_word(rec.x), lt);
                        }
                        q2 -= cnt;
                        __builtin_amdgcn_wave_barrier();
                    };
                    // Single class, first stage: the top cnt fresh records meet
                    // Philox block 0 (four bit planes); decided words are written,
                    // undecided ones move to wq2, which is resolved 64 at a time,
                    // so a word needing more blocks no longer holds up 63 others.
                    auto drain1 = [&](const int cnt) __attribute__((always_inline)) {
                        __builtin_amdgcn_wave_barrier();
                        const int lane = tid & 63;
                        bool more = false;
                        uint4 rec2 = make_uint4(0u, 0u, 0u, 0u);
                        if (lane < cnt) {
                            const uint2 r2 = wq1[qn - cnt + lane];
                            const int i = rec_i(r2.x), rr = rec_rr(r2.x);
                            const int tau = rec_tau(r2.x);
                            const WT m0 = (WT)r2.y;
                            WT lt = 0, und = m0;
                            uint32_t out[4];
                            philox4x32_10(g, (uint32_t)i, a.replica_base + (uint32_t)(r0 + rr), uint32_t(tau) << 24,
                                          a.key0, a.key1, out);
    #pragma unroll
                            for (int t = 0; t < 4; ++t) {
                                const WT Rb = rword<WT>(out, t);
                                const WT tb = ((th_s[0] >> (31 - t)) & 1u) ? m0 : WT(0);
                                lt |= und & ~Rb & tb;
                                und &= ~(Rb ^ tb);
                            }
                            if (und == 0) {
                                xor_word(rec_word(r2.x), lt);
                            } else {
                                more = true;
                                rec2 = make_uint4(r2.x, r2.y, (uint32_t)lt, (uint32_t)und);
                            }
                        }
                        qn -= cnt;
                        const unsigned long long bal = __ballot(more);
                        if (more) {
                            const int pos = q2 + (int)__builtin_amdgcn_mbcnt_hi(
                                                     (uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                            wq2[pos] = rec2;
                        }
                        q2 += (int)__popcll(bal);
                        __builtin_amdgcn_wave_barrier();
                        if (q2 >= 64) drain2(64);
                    };
                    // Single class, end of a color phase with at most 64 records
                    // left: the fresh ones (Philox block 0 onwards) and the
                    // undecided ones (block 1 onwards) share one pass, each lane
                    // starting at its own block, so the phase ends with one
                    // partial batch instead of two.
                    auto drain_last = [&]() __attribute__((always_inline)) {
                        __builtin_amdgcn_wave_barrier();
                        const int lane = tid & 63;
                        if (lane < qn + q2) {
                            int i, rr, qd0, tau;
                            uint32_t rx;
                            WT m0, lt, und;
                            if (lane < qn) {
                                const uint2 r2 = wq1[lane];
                                rx = r2.x;
                                i = rec_i(r2.x);
                                rr = rec_rr(r2.x);
                                tau = rec_tau(r2.x);
                                m0 = (WT)r2.y;
                                lt = 0;
                                und = m0;
                                qd0 = 0;
                            } else {
                                const uint4 rec = wq2[lane - qn];
                                rx = rec.x;
                                i = rec_i(rec.x);
                                rr = rec_rr(rec.x);
                                tau = rec_tau(rec.x);
                                m0 = (WT)rec.y;
                                lt = (WT)rec.z;
                                und = (WT)rec.w;
                                qd0 = 1;
                            }
                            const uint32_t rep = a.replica_base + (uint32_t)(r0 + rr);
                            for (int qd = qd0; qd < 8 && und; ++qd) {
                                uint32_t out[4];
                                philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qd, a.key0, a.key1, out);
    #pragma unroll
                                for (int t = 0; t < 4; ++t) {
                                    const WT Rb = rword<WT>(out, t);
                                    const WT tb = ((th_s[0] >> (31 - (qd * 4 + t))) & 1u) ? m0 : WT(0);
                                    lt |= und & ~Rb & tb;
                                    und &= ~(Rb ^ tb);
                                }
                            }
                            xor_word(rec_word(rx), lt);
                        }
                        qn = 0;
                        q2 = 0;
                        __builtin_amdgcn_wave_barrier();
                    };
                    // NS: classes of the step (1, or up to 3), a compile-time constant.
                    auto drain = [&](const int cnt, auto NS) __attribute__((always_inline)) {
                        __builtin_amdgcn_wave_barrier();
                        const int lane = tid & 63;
                        if (lane < cnt) {
                            uint4 rec;  // single-class records are (index | replica << 16, mask)
                            if constexpr (NS() == 1) {
                                const uint2 r2 = *reinterpret_cast<const uint2 *>(wq + (qn - cnt + lane));
                                rec = make_uint4(r2.x, r2.y, 0u, 0u);
                            } else {
                                rec = wq[qn - cnt + lane];
                            }
                            const int i = rec_i(rec.x), rr = rec_rr(rec.x);
                            const int tau = rec_tau(rec.x);
                            const WT m0 = (WT)rec.y, m1 = NS() > 1 ? (WT)rec.z : WT(0), m2 = NS() > 2 ? (WT)rec.w : WT(0);
                            const uint32_t rep = a.replica_base + (uint32_t)(r0 + rr);
                            WT lt = 0, und = m0 | m1 | m2;
                            for (int qd = 0; qd < 8 && und; ++qd) {
                                uint32_t out[4];
                                philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qd, a.key0, a.key1, out);
    #pragma unroll
                                for (int t = 0; t < 4; ++t) {
                                    const WT Rb = rword<WT>(out, t);
                                    const int sh = 31 - (qd * 4 + t);
                                    const WT tb = (((th_s[0] >> sh) & 1u) ? m0 : WT(0)) |
                                                  (((th_s[1] >> sh) & 1u) ? m1 : WT(0)) |
                                                  (((th_s[2] >> sh) & 1u) ? m2 : WT(0));
                                    lt |= und & ~Rb & tb;
                                    und &= ~(Rb ^ tb);
                                }
                            }
                            xor_word(rec_word(rec.x), lt);
                        }
                        qn -= cnt;
                        __builtin_amdgcn_wave_barrier();
                    };
                    // one spin (position q in color order, index i, neighbours nb)
                    // for every replica of the workgroup; lanes with act false take
                    // part only in the queue's wave-wide steps
                    auto spin = [&](const bool act, const int q, const int i, const uint32_t info, const int zi,
                                    const int *nb, const FieldClasses *fcp, const int tau) __attribute__((always_inline)) {
                            const int tsa = tau == 0 ? 1 : W - 1, tsb = tau == 0 ? W - 1 : 1;  // rotation by one slice
                            // the same rotation as one funnel shift of (O : O << rls) right by rsh
                            const int rls = (W == 1 && tau == 1) ? 0 : 32 - W, rsh = (32 - tsa) & 31;
                            const uint32_t zt = UZ ? zt_u : ((PMJ && !FLD) ? tz[zi] : 0u);
                            // fields: this spin's class thresholds, shared by the replicas below
                            // (cached spins: computed once per field step)
                            double hfield = 0.0;
                            FieldClasses fc;
                            if (FLD) {
                                hfield = (double)a.hq[q];
                                if (fcp) fc = *fcp;
                                else field_classes(hfield, zi, jp, Ts, fc);
                            }
                            // fixed planes: unrolled, so each replica's planes are immediates
    #pragma unroll(FIXED ? kFixRep : 1)
                            for (int rr = 0; rr < nrep; ++rr) {
                              WT qX = 0, qm0 = 0, qm1 = 0, qm2 = 0;  // stochastic word for the queue
                              if (act) {
                                WT *mine = st + ((size_t)rr * 2 + tau) * PW;
                                const WT *other = st + ((size_t)rr * 2 + 1 - tau) * PW;
                                const uint32_t rep = a.replica_base + (uint32_t)(r0 + rr);
                                const WT S_ = mine[i], O = other[i];
                                // Trotter neighbours of slice 2m+tau: bits m of O and of O
                                // rotated by one slice (left for tau 0, right for tau 1);
                                // everything below is symmetric in the two
                                const WT Lw = O, Rw = rot_slices<WT>(O, tsa, tsb, rls, rsh) & valid;
                                const WT uL = S_ ^ Lw, uR = S_ ^ Rw;  // 1 = antiparallel Trotter bond
                                WT flips = 0;
                                if (PMJ) {
                                    WT c0 = 0, c1 = 0, c2 = 0;  // bit-sliced unsatisfied-bond counter
                                    if constexpr (ZU == 4 || ZU == 6) {
                                        // carry-save adder tree: a full adder is one XOR3 and one
                                        // majority (v_bitop3 each), 6 logic ops for four bonds and
                                        // 8 for six instead of 5 per bond for a ripple counter
                                        WT x[ZU];
    #pragma unroll
                                        for (int z = 0; z < ZU; ++z)
                                            x[z] = xor3<WT>(S_, nbword(mine, nb[z]), ((info >> z) & 1u) ? ~WT(0) : WT(0));
                                        if constexpr (ZU == 4) {
                                            const WT s1 = xor3<WT>(x[0], x[1], x[2]), k1 = maj3<WT>(x[0], x[1], x[2]);
                                            c0 = s1 ^ x[3];
                                            const WT k2 = s1 & x[3];
                                            c1 = k1 ^ k2;
                                            c2 = k1 & k2;
                                        } else {
                                            const WT s1 = xor3<WT>(x[0], x[1], x[2]), k1 = maj3<WT>(x[0], x[1], x[2]);
                                            const WT s2 = xor3<WT>(x[3], x[4], x[5]), k2 = maj3<WT>(x[3], x[4], x[5]);
                                            c0 = s1 ^ s2;
                                            const WT k3 = s1 & s2;
                                            c1 = xor3<WT>(k1, k2, k3);
                                            c2 = maj3<WT>(k1, k2, k3);
                                        }
                                    } else {
    #pragma unroll
                                    for (int z = 0; z < 8; ++z) {
                                        if (ZU > 0 ? z < ZU : (z < a.zmax && z < zi)) {
                                            const WT x = S_ ^ nbword(mine, nb[z]) ^ (((info >> z) & 1u) ? ~WT(0) : WT(0));
                                            const WT k0 = c0 & x;
                                            c0 ^= x;
                                            const WT k1 = c1 & k0;
                                            c1 ^= k0;
                                            c2 ^= k1;
                                        }
                                    }
                                    }
                                    // Trotter classes; bits past W are zero in every class
                                    const WT e0 = ~(uL | uR) & valid, e1 = uL ^ uR, e2 = uL & uR;
                                    WT X;
                                    WT ge[KMAX + 1];  // "count >= k" masks (uniform-degree path)
                                    if (FLD) {
                                        // Per-spin thresholds for (sigma, u_t), nibble sigma*3 + u_t
                                        // (sigma = +1 bits are ~S_).  Each bit's threshold is built as
                                        // bit planes by selects on (S_, uL, uR) -- u_t = 1 where uL
                                        // and uR differ, else 2 or 0 by uL -- and compared with the
                                        // bond count once: the same masks as a "count >= k" per
                                        // (sigma, u_t) class, in a third of the work.
                                        const WT xr = uL ^ uR;
                                        auto kplane = [&](uint32_t pk, int j) __attribute__((always_inline)) -> WT {
                                            const WT a0 = bitmask<WT>(pk, j), a1 = bitmask<WT>(pk, 4 + j),
                                                     a2 = bitmask<WT>(pk, 8 + j);
                                            const WT b0 = bitmask<WT>(pk, 12 + j), b1 = bitmask<WT>(pk, 16 + j),
                                                     b2 = bitmask<WT>(pk, 20 + j);
                                            const WT fp = (xr & a1) | (~xr & ((uL & a2) | (~uL & a0)));
                                            const WT fn = (xr & b1) | (~xr & ((uL & b2) | (~uL & b0)));
                                            return (S_ & fn) | (~S_ & fp);
                                        };
                                        auto ge_k = [&](uint32_t pk) __attribute__((always_inline)) -> WT {
                                            const WT K0 = kplane(pk, 0), K1 = kplane(pk, 1), K2 = kplane(pk, 2);
                                            WT t = c0 | ~K0;
                                            t = (c1 & ~K1) | (~(c1 ^ K1) & t);
                                            t = (c2 & ~K2) | (~(c2 ^ K2) & t);
                                            // thresholds reach 8 ("never") only with seven couplings
                                            if constexpr (!(ZU > 0 && ZU <= 6)) t &= ~kplane(pk, 3);
                                            return t;
                                        };
                                        flips = ge_k(fc.ka) & valid;
                                        X = ge_k(fc.ks) & valid;
                                    } else if (UZ) {
                                        // thresholds are wave-uniform: every "count >= k" mask
                                        // once, then six selections by a uniform index
                                        const WT gek[9] = {~WT(0), c0 | c1 | c2, c1 | c2, c2 | (c1 & c0), c2,
                                                           c2 & (c1 | c0), c2 & c1, c2 & c1 & c0, WT(0)};
    #pragma unroll
                                        for (int kk = 0; kk < KMAX; ++kk) ge[kk] = gek[kk];
                                        ge[KMAX] = WT(0);
                                        // fixed planes (degree 4, no fields): one antiparallel
                                        // Trotter bond cancels J_perp, so that class accepts from
                                        // count 2 in every step with T > 0 (the host checks it)
                                        flips = (e0 & ge[ka0]) | (e1 & ge[FIXED ? 2 : ka1]) | (e2 & ge[ka2]);
                                        if constexpr (SINGLE) {
                                            // the step's one stochastic class (u_t, u_s): exactly the
                                            // bits of Trotter class u_t with count u_s (no accepted
                                            // bit is among them, so no mask by ~flips)
                                            const WT eu = UTC == 0 ? e0 : (UTC == 1 ? e1 : e2);
                                            X = eu & ge[us_s[0]] & ~ge[us_s[0] + 1];
                                        } else {
                                            X = DET ? WT(0) : ((e0 & ge[ks0]) | (e1 & ge[ks1]) | (e2 & ge[ks2]));
                                        }
                                    } else {
                                        flips = (e0 & ge_pick((int)(zt & 15), c0, c1, c2)) |
                                                (e1 & ge_pick((int)((zt >> 8) & 15), c0, c1, c2)) |
                                                (e2 & ge_pick((int)((zt >> 16) & 15), c0, c1, c2));
                                        X = (e0 & ge_pick((int)((zt >> 4) & 15), c0, c1, c2)) |
                                            (e1 & ge_pick((int)((zt >> 12) & 15), c0, c1, c2)) |
                                            (e2 & ge_pick((int)((zt >> 20) & 15), c0, c1, c2));
                                    }
                                    if constexpr (!SINGLE) X &= ~flips;
                                    if (CQ && cqm) {
                                        // masks of the step's uniform classes, queued with the word
                                        // (a single class holds every stochastic bit)
                                        qX = X;
                                        if (one) qm0 = X;
                                        else
    #pragma unroll
                                        for (int t = 0; t < 3; ++t) {
                                            if (t < un) {
                                                const int ut = ut_s[t], us = us_s[t];
                                                const WT eu = ut == 0 ? e0 : (ut == 1 ? e1 : e2);
                                                const WT mk = X & eu & ge[us] & ~ge[us + 1];
                                                if (t == 0) qm0 = mk;
                                                else if (t == 1) qm1 = mk;
                                                else qm2 = mk;
                                            }
                                        }
                                    } else if (X) {
                                        // Bit-sliced comparison u < thr for every stochastic bit at
                                        // once, most significant uniform bit first: random word R_b
                                        // supplies bit 31-b of all uniforms; a bit is decided at the
                                        // first b where its uniform bit differs from its threshold bit.
                                        WT lt = 0, und = X;
                                        // The stochastic classes of this word (one per field step on
                                        // a regular lattice, rarely more): masks and thresholds in
                                        // registers; more classes take the slow paths below.
                                        WT cm0 = 0, cm1 = 0, cm2 = 0;
                                        uint32_t ct0 = 0, ct1 = 0, ct2 = 0;
                                        int ns = 0;
                                        if (UZ && !FLD) {
                                            // the step's uniform classes: masks by indexed selection
                                            ns = un;
    #pragma unroll
                                            for (int t = 0; t < 3; ++t) {
                                                if (t < ns) {
                                                    const int ut = ut_s[t], us = us_s[t];
                                                    const WT eu = ut == 0 ? e0 : (ut == 1 ? e1 : e2);
                                                    const WT mk = eu & ge[us] & ~ge[us + 1];
                                                    if (t == 0) { cm0 = mk; ct0 = th_s[0]; }
                                                    else if (t == 1) { cm1 = mk; ct1 = th_s[1]; }
                                                    else { cm2 = mk; ct2 = th_s[2]; }
                                                }
                                            }
                                        } else if (FLD) {
                                            ns = fc.ns;
    #pragma unroll
                                            for (int t = 0; t < 3; ++t) {
                                                if (t < ns) {
                                                    const uint32_t sb = (fc.slot >> (8 * t)) & 255u;
                                                    const int cb = (int)(sb & 7u);
                                                    const WT eu = (cb % 3) == 0 ? e0 : ((cb % 3) == 1 ? e1 : e2);
                                                    // count == u_s (u_s <= 7: three bits at 4..6 of the slot)
                                                    const WT eq = ~((c0 ^ bitmask<WT>(sb, 4)) | (c1 ^ bitmask<WT>(sb, 5)) |
                                                                    (c2 ^ bitmask<WT>(sb, 6)));
                                                    const WT mk = (cb < 3 ? ~S_ : S_) & eu & eq;
                                                    if (t == 0) { cm0 = mk; ct0 = fc.st[0]; }
                                                    else if (t == 1) { cm1 = mk; ct1 = fc.st[1]; }
                                                    else { cm2 = mk; ct2 = fc.st[2]; }
                                                }
                                            }
                                        } else {
    #pragma unroll
                                            for (int ut = 0; ut < 3; ++ut) {
                                                const WT eu = ut == 0 ? e0 : (ut == 1 ? e1 : e2);
                                                const int amin = (int)((zt >> (8 * ut)) & 15);
                                                for (int us = (int)((zt >> (8 * ut + 4)) & 15); us < amin; ++us) {
                                                    const WT mk = eu & ge_pick(us, c0, c1, c2) & ~ge_pick(us + 1, c0, c1, c2);
                                                    const uint32_t tk = tthr[(zi * 8 + us) * 3 + ut];
                                                    if (ns == 0) { cm0 = mk; ct0 = tk; }
                                                    else if (ns == 1) { cm1 = mk; ct1 = tk; }
                                                    else if (ns == 2) { cm2 = mk; ct2 = tk; }
                                                    ++ns;
                                                }
                                            }
                                        }
                                        constexpr int kPer = sizeof(WT) == 4 ? 4 : 2;  // words per Philox call
                                        if (FLD && ns > 3) {
                                            // many stochastic triples (high temperature): each
                                            // stochastic bit assembles its own uniform and threshold
                                            // (out of line as a called function it costs more in the
                                            // call's register traffic than it saves: 297 -> 301 ms
                                            // on the 3-D L=12 field workload)
                                            for (int mm = 0; mm < W; ++mm) {
                                                if (!((X >> mm) & WT(1))) continue;
                                                const int sg = (int)((S_ >> mm) & WT(1));
                                                const int ut = (int)((uL >> mm) & WT(1)) + (int)((uR >> mm) & WT(1));
                                                const int us = (int)((c0 >> mm) & WT(1)) + 2 * (int)((c1 >> mm) & WT(1)) +
                                                               4 * (int)((c2 >> mm) & WT(1));
                                                unsigned int thr = 0;
                                                (void)field_class(sg ? -hfield : hfield, zi, us,
                                                                  (double)jp * (double)(2 - 2 * ut), Ts, &thr);
                                                uint32_t u = 0;
                                                for (int qd = 0; qd < 32 / kPer; ++qd) {
                                                    uint32_t out[4];
                                                    philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qd,
                                                                  a.key0, a.key1, out);
    #pragma unroll
                                                    for (int t = 0; t < kPer; ++t)
                                                        u |= (uint32_t)((rword<WT>(out, t) >> mm) & WT(1)) << (31 - (qd * kPer + t));
                                                }
                                                if (u < thr) lt |= WT(1) << mm;
                                            }
                                            und = 0;
                                        }
                                        for (int qd = 0; qd < 32 / kPer && und; ++qd) {
                                            uint32_t out[4];
                                            philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qd,
                                                          a.key0, a.key1, out);
    #pragma unroll
                                            for (int t = 0; t < kPer; ++t) {
                                                const WT Rb = rword<WT>(out, t);
                                                const int sh = 31 - (qd * kPer + t);
                                                WT tb;  // bits whose class threshold has bit sh set
                                                // (with fields more than three classes never get here)
                                                if (FLD || ns <= 3) {
                                                    tb = (((ct0 >> sh) & 1u) ? cm0 : WT(0)) | (((ct1 >> sh) & 1u) ? cm1 : WT(0)) |
                                                         (((ct2 >> sh) & 1u) ? cm2 : WT(0));
                                                } else {
                                                    tb = 0;
    #pragma unroll
                                                    for (int ut = 0; ut < 3; ++ut) {
                                                        const WT eu = ut == 0 ? e0 : (ut == 1 ? e1 : e2);
                                                        const int amin = (int)((zt >> (8 * ut)) & 15);
                                                        for (int us = (int)((zt >> (8 * ut + 4)) & 15); us < amin; ++us)
                                                            if ((tthr[(zi * 8 + us) * 3 + ut] >> sh) & 1u)
                                                                tb |= eu & ge_pick(us, c0, c1, c2) & ~ge_pick(us + 1, c0, c1, c2);
                                                    }
                                                }
                                                lt |= und & ~Rb & tb;
                                                und &= ~(Rb ^ tb);
                                            }
                                        }
                                        flips |= lt & X;
                                    }
                                } else {
                                    const int2 *row = a.rows + (size_t)i * a.Z;
                                    bool have_r = false;
                                    WT R[32];
                                    for (int mm = 0; mm < W; ++mm) {
                                        const double si = ((S_ >> mm) & 1) ? -1.0 : 1.0;
                                        double h = 0.0;
                                        for (int e = 0; e < a.Z; ++e) {
                                            const int2 rw = row[e];
                                            const double J = (double)__int_as_float(rw.y);
                                            if (rw.x == i) h += J;
                                            else h += J * (((mine[rw.x] >> mm) & 1) ? -1.0 : 1.0);
                                        }
                                        const double ts = (((Lw >> mm) & 1) ? -1.0 : 1.0) + (((Rw >> mm) & 1) ? -1.0 : 1.0);
                                        const double ed = -2.0 * si * (h + (double)jp * ts);
                                        unsigned int thr = 0;
                                        const int cls = piqmc_fixed_class(ed, Ts, &thr);
                                        bool fl = cls == 1;
                                        if (cls == 2) {
                                            if (!have_r) {  // the word's 32 bit-sliced random words
                                                have_r = true;
                                                constexpr int kPer = sizeof(WT) == 4 ? 4 : 2;
    #pragma unroll
                                                for (int qd = 0; qd < 32 / kPer; ++qd) {
                                                    uint32_t out[4];
                                                    philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qd,
                                                                  a.key0, a.key1, out);
    #pragma unroll
                                                    for (int t = 0; t < kPer; ++t)
                                                        R[qd * kPer + t] = rword<WT>(out, t);
                                                }
                                            }
                                            uint32_t u = 0;
    #pragma unroll
                                            for (int bb = 0; bb < 32; ++bb) u |= (uint32_t)((R[bb] >> mm) & 1) << (31 - bb);
                                            fl = u < thr;
                                        }
                                        if (fl) flips |= WT(1) << mm;
                                    }
                                }
                                mine[i] = S_ ^ flips;
                              }
                              if constexpr (CQ && !DET) {
                                  if (cqm) {
                                      // append this lane's stochastic word (if any) to the
                                      // wavefront's queue; a full batch of 64 is drained at once
                                      const unsigned long long bal = __ballot(qX != 0);
                                      if (qX != 0) {
                                          const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi(
                                                                   (uint32_t)(bal >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                                          // formed here: hoisted out of the step loop, the
                                          // records' fixed parts would hold registers
                                          uint32_t rx = FIXED ? (uint32_t)i << 2 : (uint32_t)i;
                                          asm volatile("" : "+v"(rx));
                                          if constexpr (FIXED) rx |= ((uint32_t)tau << 12) | ((uint32_t)rr << 13);
                                          else rx |= ((uint32_t)rr << 16) | ((uint32_t)tau << 31);
                                          if (SINGLE)
                                              wq1[pos] = make_uint2(rx, (uint32_t)qm0);
                                          else
                                              wq[pos] = make_uint4(rx, (uint32_t)qm0, (uint32_t)qm1, (uint32_t)qm2);
                                      }
                                      qn += (int)__popcll(bal);
                                      if (!FIXED && qn >= 64) {
                                          if constexpr (SINGLE) drain1(64);
                                          else drain(64, std::integral_constant<int, 3>());
                                      }
                                  }
                              }
                            }
                    };
                    if constexpr (CACHED) {
                        // Three phases: (parity 0, color 0); (parity 0, color 1)
                        // together with (parity 1, color 0); (parity 1, color 1).
                        // The two updates of the middle phase touch disjoint words
                        // and neither reads what the other writes (parity 0 of
                        // color 1 reads parity 0 of color 0 and parity 1 of color
                        // 1; parity 1 of color 0 reads parity 1 of color 1 and
                        // parity 0 of color 0), so running them in one phase gives
                        // the sequential order's spins, with one partial queue batch
                        // and one barrier fewer per sweep.
                        // Each slot (phase, part, k) is its own instantiation, so
                        // (parity, color, k) are compile-time whatever the size of
                        // the code and the cached registers are used directly.
                        auto slot = [&](auto PH, auto PART, auto K) __attribute__((always_inline)) {
                            constexpr int ph = decltype(PH)::value, part = decltype(PART)::value;
                            constexpr int k = decltype(K)::value;
                            constexpr int tp = ph == 0 ? 0 : (ph == 2 ? 1 : part);
                            constexpr int c = ph == 0 ? 0 : (ph == 2 ? 1 : 1 - part);
                            const int q = cq_[c][k];
                            if (!CQ && q < 0) return;
                            int nb[ZC];
    #pragma unroll
                            for (int z = 0; z < ZC; ++z) nb[z] = cN[c][k][z];
                            const FieldClasses *fcp = nullptr;
                            if constexpr (FLD) fcp = &cfc[c][k];
                            spin(FIXED || q >= 0, q, cI[c][k], cF[c][k], ZU, nb, fcp, tp);
                            if constexpr (FIXED && !DET) {
                                // one drain site for both replicas' appends
                                if (cqm) {
                                    while (qn >= 64) {
                                        if constexpr (SINGLE) drain1(64);
                                        else drain(64, std::integral_constant<int, 3>());
                                    }
                                }
                            }
                        };
                        auto phase = [&](auto PH) __attribute__((always_inline)) {
                            using I0 = std::integral_constant<int, 0>;
                            using I1 = std::integral_constant<int, 1>;
                            slot(PH, I0(), I0());
                            slot(PH, I0(), I1());
                            if constexpr (decltype(PH)::value == 1) {
                                slot(PH, I1(), I0());
                                slot(PH, I1(), I1());
                            }
                            if constexpr (CQ && SINGLE) {
                                // full batches first, then one mixed pass over the rest
                                while (qn + q2 > 64) {
                                    if (qn > 0) drain1(min(qn, 64));
                                    else drain2(64);
                                }
                                if (qn + q2 > 0) drain_last();
                            } else if (CQ && !DET && qn > 0) {
                                drain(qn, std::integral_constant<int, 3>());
                            }
                            __syncthreads();
                        };
                        phase(std::integral_constant<int, 0>());
                        phase(std::integral_constant<int, 1>());
                        phase(std::integral_constant<int, 2>());
                    } else {
                        for (int c = 0; c < a.ncolors; ++c) {
                            const int q0 = a.color_off[c], q1 = a.color_off[c + 1];
                            for (int qb = q0; qb < q1; qb += bd) {
                                // with the queue every lane runs every pass (inactive past q1)
                                const bool act = qb + tid < q1;
                                if (!CQ && !act) break;
                                const int q = act ? qb + tid : q0;
                                const int i = a.order[q];
                                uint32_t info = 0;
                                int zi = 0;
                                int nb[8];
                                if (PMJ) {
                                    info = a.cinfo[q];
                                    zi = ZU > 0 ? ZU : (UZ ? a.zuni : (int)((info >> 8) & 15));
                                    if (ZU > 0) {
    #pragma unroll
                                        for (int z = 0; z < ZU; ++z) nb[z] = a.cnb[(size_t)z * N + q];
                                    } else {
    #pragma unroll
                                        for (int z = 0; z < 8; ++z)
                                            if (z < a.zmax) nb[z] = z < zi ? a.cnb[(size_t)z * N + q] : i;
                                    }
                                }
                                spin(act, q, i, info, zi, nb, nullptr, tau);
                            }
                            if constexpr (CQ && SINGLE) {
                                // full batches first, then one mixed pass over the rest
                                while (qn + q2 > 64) {
                                    if (qn > 0) drain1(min(qn, 64));
                                    else drain2(64);
                                }
                                if (qn + q2 > 0) drain_last();
                            } else if (CQ && !DET && qn > 0) {
                                drain(qn, std::integral_constant<int, 3>());
                            }
                            __syncthreads();
                        }
                    }
                }
            }
        };
        if constexpr (CQ) {
            if (un == 0) sweeps(std::integral_constant<int, 0>());
            else if (un == 1 && ut_s[0] == 0) sweeps(std::integral_constant<int, 1>());
            else if (un == 1 && ut_s[0] == 1) sweeps(std::integral_constant<int, 3>());
            else if (un == 1) sweeps(std::integral_constant<int, 4>());
            else sweeps(std::integral_constant<int, 2>());
        } else {
            sweeps(std::integral_constant<int, 2>());
        }
        if (a.etrace) {
            // Mean over slices of the classical energy; sums of +-1 products are
            // popcounts of XORed parity words.
            for (int rr = 0; rr < nrep; ++rr) {
                const WT *s0 = st + (size_t)rr * 2 * PW, *s1 = s0 + PW;
                double acc = 0.0;
                for (int i = tid; i < N; i += bd) {
                    const WT a0 = s0[i], a1 = s1[i];
                    const int2 *row = a.rows + (size_t)i * a.Z;
                    for (int e = 0; e < a.Z; ++e) {
                        const int2 rw = row[e];
                        const double J = (double)__int_as_float(rw.y);
                        if (J == 0.0) continue;
                        if (rw.x == i) {
                            acc -= J * (double)(a.P - 2 * (popc(a0) + popc(a1)));
                        } else {
                            const int pc = popc(a0 ^ s0[rw.x]) + popc(a1 ^ s1[rw.x]);
                            acc -= 0.5 * J * (double)(a.P - 2 * pc);
                        }
                    }
                }
                const double tot = block_sum(acc, red);
                if (tid == 0) a.etrace[(size_t)(r0 + rr) * a.S + s] = tot / (double)a.P;
                __syncthreads();
            }
        }
    }
    __syncthreads();
    if constexpr (FIXED) {
        for (int q = tid; q < 2 * N * nrep; q += bd) gst[q] = st[(q / N) * PW + q % N];
    } else {
        for (int q = tid; q < 2 * N * nrep; q += bd) gst[q] = st[q];
    }
}

// flatten: the sweep body is inlined whole (its unrolled phases need the
// cached registers of the kernel's own frame)
template <typename WT, bool PMJ, bool UZ, bool FLD = false, int ZU = 0, bool CACHED = false, bool CQ = false>
__global__ void __attribute__((flatten)) colored_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int r0 = blockIdx.x * a.rpw;
    colored_body<WT, PMJ, UZ, FLD, ZU, CACHED, CQ>(a, r0, min(a.rpw, a.R - r0), smem);
}

// The field sweep with cached spins and decision classes (cubic lattices): at
// most 512 threads, so that the cached registers fit without spilling.
template <typename WT>
__global__ void __attribute__((flatten)) __launch_bounds__(512) colored_fields_cached_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int r0 = blockIdx.x * a.rpw;
    colored_body<WT, true, false, true, 6, true, false>(a, r0, min(a.rpw, a.R - r0), smem);
}

// The queue kernels (+-1 couplings, uniform degree) with a register budget of
// five wavefronts per SIMD: left to itself the compiler takes 117 registers
// and four; on the 1024-spin workload five run 15.1 -> 14.4 ms despite a
// 12-byte spill.
template <typename WT, int ZU, bool CACHED>
__global__ void __attribute__((flatten)) __attribute__((amdgpu_waves_per_eu(5))) colored_q_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int r0 = blockIdx.x * a.rpw;
    colored_body<WT, true, true, false, ZU, CACHED, true>(a, r0, min(a.rpw, a.R - r0), smem);
}

// The fixed-plane cached queue sweep (square lattices, 32-bit words): 256
// threads, two replicas, N <= kFixPlane.  Its LDS is a static array, so the
// plane offsets and the array's base are all compile-time and fold into the
// LDS instructions' immediate offsets.  Five wavefronts per SIMD (96
// registers): the word loops fit without spilling, the few spills left are in
// the per-phase queue tails (1024-spin workload: 8.91 ms, against 9.04 ms at
// four wavefronts and 106 registers).
constexpr size_t kFixBlock = 256;
constexpr size_t kFixLds = queue_offset(kFixPlane, kFixRep, 4) + kFixBlock / 64 * kQueueBytes;
__global__ void __attribute__((flatten)) __attribute__((amdgpu_waves_per_eu(5))) __launch_bounds__(kFixBlock)
colored_q_fixed_kernel(ColArgs<uint32_t> a)
{
    __shared__ __align__(16) unsigned char fsm[kFixLds];
    colored_body<uint32_t, true, true, false, 4, true, true, true>(a, blockIdx.x * kFixRep, kFixRep, fsm);
}

// Launch of the +-1 sweep for a uniform-degree lattice of compile-time degree
// ZU, with per-thread cached neighbour lists when two colors of at most two
// spins per thread allow it, and with the per-wavefront stochastic queue when
// its LDS fits (32-bit words).
template <typename WT, int ZU, bool CQ>
void launch_uniform_q(const ColArgs<WT> &a, bool cached, int grid, int block, size_t lds, hipStream_t st,
                      bool fixed = false)
{
    if constexpr (CQ) {
        if constexpr (ZU == 4 && sizeof(WT) == 4) {
            if (fixed) {
                colored_q_fixed_kernel<<<grid, (int)kFixBlock, 0, st>>>(a);
                return;
            }
        }
        if (cached) {
            (void)hipFuncSetAttribute((const void *)colored_q_kernel<WT, ZU, true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            colored_q_kernel<WT, ZU, true><<<grid, block, lds, st>>>(a);
        } else {
            (void)hipFuncSetAttribute((const void *)colored_q_kernel<WT, ZU, false>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            colored_q_kernel<WT, ZU, false><<<grid, block, lds, st>>>(a);
        }
    } else if (cached) {
        (void)hipFuncSetAttribute((const void *)colored_kernel<WT, true, true, false, ZU, true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        colored_kernel<WT, true, true, false, ZU, true><<<grid, block, lds, st>>>(a);
    } else {
        (void)hipFuncSetAttribute((const void *)colored_kernel<WT, true, true, false, ZU, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        colored_kernel<WT, true, true, false, ZU, false><<<grid, block, lds, st>>>(a);
    }
}

// The stochastic queue is on unless PIQMC_STOCHASTIC_QUEUE=0 (both forms give
// identical results; the switch exists to compare them).
inline bool queue_enabled()
{
    const char *e = std::getenv("PIQMC_STOCHASTIC_QUEUE");
    return !(e && e[0] == '0');
}

// Exactly summable dense problems sweep one replica per workgroup unless
// PIQMC_DENSE_WG=0 (identical results; the switch exists to compare).
inline bool dense_wg_enabled()
{
    const char *e = std::getenv("PIQMC_DENSE_WG");
    return !(e && e[0] == '0');
}

// The cached-register field sweep (PIQMC_CACHED_FIELDS=0 selects the general
// one, for comparisons).
inline bool cached_fields_enabled()
{
    const char *e = std::getenv("PIQMC_CACHED_FIELDS");
    return !(e && e[0] == '0');
}

// The fixed-plane cached sweep runs where it applies unless PIQMC_FIXED_PLANES=0
// (identical results; the switch exists to compare).
inline bool fixed_planes_enabled()
{
    const char *e = std::getenv("PIQMC_FIXED_PLANES");
    return !(e && e[0] == '0');
}

// The fixed-plane kernel applies to +-1 problems of uniform degree 4 with 32-bit
// words, two colors of exactly 2 x kFixBlock spins, N <= kFixPlane and an even
// replica count (at kFixRep replicas per workgroup and kFixBlock threads).
inline bool fixed_tile_applies(const piqmc_problem *pb, int R, size_t word_bytes)
{
    return fixed_planes_enabled() && queue_enabled() && word_bytes == 4 && pb->pmj && pb->zuni == 4 &&
           pb->ncolors == 2 && pb->N <= kFixPlane && R % kFixRep == 0 &&
           pb->color_off[1] == 2 * (int)kFixBlock && pb->color_off[2] - pb->color_off[1] == 2 * (int)kFixBlock;
}

// LDS of the word-per-thread sweep with the stochastic queues of block/64 wavefronts.
inline size_t queue_lds(int N, int rpw, int word_bytes, int block)
{
    return queue_offset(N, rpw, word_bytes) + (size_t)(block / 64) * kQueueBytes;
}

template <typename WT, int ZU>
void launch_uniform(const ColArgs<WT> &a, const piqmc_problem *pb, int grid, int block, size_t lds, hipStream_t st,
                    bool ka1_two = false)
{
    int biggest = 0;
    for (int c = 0; c < pb->ncolors; ++c) biggest = std::max(biggest, pb->color_off[c + 1] - pb->color_off[c]);
    const bool cached = pb->ncolors == 2 && biggest <= 2 * block;
    const size_t lq = queue_lds(a.N, a.rpw, sizeof(WT), block);
    if constexpr (sizeof(WT) == 4) {
        // fixed planes: two replicas per workgroup (R even), both colors
        // exactly two spins per thread, N within a plane
        const bool fixed = ZU == 4 && ka1_two && a.rpw == kFixRep && block == (int)kFixBlock &&
                           fixed_tile_applies(pb, a.R, sizeof(WT));
        if (fixed && queue_enabled()) {
            launch_uniform_q<WT, ZU, true>(a, cached, grid, block, 0, st, true);
            return;
        }
        if (lq <= 160 * 1024 && queue_enabled()) {
            launch_uniform_q<WT, ZU, true>(a, cached, grid, block, lq, st);
            return;
        }
    }
    launch_uniform_q<WT, ZU, false>(a, cached, grid, block, lds, st);
}

// ------------------------------------------------------------------ dense path
// Graphs whose color classes are tiny (dense couplings such as
// Sherrington-Kirkpatrick need one color per spin) leave the word-per-thread
// sweep with one busy thread per phase.  Here one wavefront owns one replica
// and its lanes own the slices of the parity word being updated: lane m
// evaluates the exact double local field of (spin i, slice 2m+tau) in row
// order -- the CPU twin's arithmetic -- and the word's flips are gathered
// with a ballot.  Stochastic lanes read their uniform from the same
// bit-sliced Philox words as the word-per-thread path.
// LDS of one dense-path wavefront: its replica's packed state [2][N], then
// (8-byte aligned) the staged row of the spin being updated -- couplings as
// double [Z] and the neighbours' packed words [Z].
__host__ __device__ inline size_t dense_wave_bytes(int N, int Z, int word_bytes)
{
    const size_t state = ((size_t)2 * N * word_bytes + 7) & ~(size_t)7;
    return state + (size_t)Z * 8 + (((size_t)Z * word_bytes + 7) & ~(size_t)7);
}

// Decision of spin i in the W slices of one parity word, lane m < W = slice
// 2m+tau, given the exact spatial field h of this lane's slice: the
// arithmetic, classes and Philox words of the lane-per-slice body.  Returns
// the flip mask (bit m = slice m flips), wave-uniform.
template <typename WT, bool PF = false>
__device__ __forceinline__ WT dense_decide(const ColArgs<WT> &a, double h, WT S_, WT O, int lane, int tau,
                                           float jp, double Ts, uint32_t g, int i, uint32_t rep)
{
    const int W = a.W;
    const WT one = WT(1);
    const bool active = lane < W;
    const int lb = active ? lane : 0;
    const int ml = tau == 0 ? (lb == 0 ? W - 1 : lb - 1) : lb;
    const int mr = tau == 0 ? lb : (lb == W - 1 ? 0 : lb + 1);
    const double si = ((S_ >> lb) & one) ? -1.0 : 1.0;
    const double ts = (((O >> ml) & one) ? -1.0 : 1.0) + (((O >> mr) & one) ? -1.0 : 1.0);
    const double ed = -2.0 * si * (h + (double)jp * ts);
    // ed < -23 T: exp(ed/T) < 1.1e-10 < 2^-33, never accepted (class 0 of
    // piqmc_fixed_class) -- decided without the division and the exponential
    const bool pos = ed > 0.0, far = Ts > 0.0 && ed < -23.0 * Ts;
    bool fl = active && pos;
    if (__ballot(active && !pos && !far)) {
        // PF: the exponential and the first two Philox blocks are independent
        // and computed together, so their latencies overlap (a win where the
        // decision is the whole chain; where other wavefronts sum the next
        // field meanwhile, the unused blocks cost more than they save)
        constexpr int kPer = sizeof(WT) == 4 ? 4 : 2;
        const uint32_t c3 = uint32_t(tau) << 24;
        uint32_t o0[4] = {0u, 0u, 0u, 0u}, o1[4] = {0u, 0u, 0u, 0u};
        if constexpr (PF) {
            philox4x32_10(g, (uint32_t)i, rep, c3, a.key0, a.key1, o0);
            philox4x32_10(g, (uint32_t)i, rep, c3 | 1u, a.key0, a.key1, o1);
        }
        unsigned int thr = 0;
        const int cls = (!active || pos || far) ? 0 : piqmc_fixed_class(ed, Ts, &thr);
        if (cls == 1) fl = true;
        if (__ballot(cls == 2)) {
            bool und = cls == 2, lt = false;
            for (int qd = 0; qd < 32 / kPer && __ballot(und); ++qd) {
                uint32_t out[4];
                if (PF && qd == 0) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) out[t] = o0[t];
                } else if (PF && qd == 1) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) out[t] = o1[t];
                } else {
                    philox4x32_10(g, (uint32_t)i, rep, c3 | (uint32_t)qd, a.key0, a.key1, out);
                }
#pragma unroll
                for (int t = 0; t < kPer; ++t) {
                    const uint32_t ub = (uint32_t)((rword<WT>(out, t) >> lb) & one);
                    const uint32_t tb = (thr >> (31 - (qd * kPer + t))) & 1u;
                    if (und && ub != tb) {
                        lt = ub < tb;
                        und = false;
                    }
                }
            }
            if (cls == 2) fl = lt;
        }
    }
    return (WT)__ballot(fl);
}

template <typename WT, bool SPLIT = false>
__device__ void dense_replica(const ColArgs<WT> &a, int r, unsigned char *base)
{
    const int N = a.N, W = a.W, Z = a.Z, lane = threadIdx.x & 63;
    WT *st = reinterpret_cast<WT *>(base);
    double *jb = reinterpret_cast<double *>(base + (((size_t)2 * N * sizeof(WT) + 7) & ~(size_t)7));
    WT *wb = reinterpret_cast<WT *>(jb + Z);
    // With exactly summable rows a slice's field is split over `parts` lanes
    // (slice lane % W, part lane / W) when W divides 16, and each lane keeps
    // four partial sums: the same double, with more lanes busy and shorter
    // dependency chains.  Lanes of part 0 decide.
    const int parts = SPLIT ? 64 / W : 1;  // SPLIT: exact rows and W dividing 16
    const int part = parts > 1 ? lane / W : 0;
    const bool active = parts > 1 ? part == 0 : lane < W;
    const WT one = WT(1);
    const uint32_t rep = a.replica_base + (uint32_t)r;
    WT *gst = a.packed + (size_t)r * 2 * N;
    for (int q = lane; q < 2 * N; q += 64) st[q] = gst[q];
    // a spin's row, one entry per lane (entries lane + 64k), read ahead of use;
    // indices are clamped so the reads are unconditional
    constexpr int kPre = 4;
    int2 pre[kPre];
    auto fetch_row = [&](int i) {
        const int2 *row = a.rows + (size_t)i * Z;
#pragma unroll
        for (int k = 0; k < kPre; ++k) pre[k] = row[min(lane + 64 * k, Z - 1)];
    };
    for (int s = 0; s < a.S; ++s) {
        const float jp = a.jperp[s];
        const double Ts = (double)a.temps[s];
        for (int m = 0; m < a.M; ++m) {
            const uint32_t g = a.sweep_base + (uint32_t)(s * a.M + m);
            for (int tau = 0; tau < 2; ++tau) {
                WT *mine = st + tau * N;
                const WT *other = st + (1 - tau) * N;
                int inext = __builtin_amdgcn_readfirstlane(a.order[0]);
                fetch_row(inext);
                for (int q = 0; q < N; ++q) {  // colors in order; spins of one color are independent
                    const int i = inext;  // wave-uniform
                    // stage the row: couplings and the current words of the neighbours
                    // (a field entry stages word 0, so its term keeps the sign of J)
#pragma unroll
                    for (int k = 0; k < kPre; ++k) {
                        const int e = lane + 64 * k;
                        if (e < Z) {
                            jb[e] = (double)__int_as_float(pre[k].y);
                            wb[e] = pre[k].x == i ? WT(0) : mine[pre[k].x];
                        }
                    }
                    for (int e = lane + 64 * kPre; e < Z; e += 64) {  // rows longer than 256
                        const int2 rw = a.rows[(size_t)i * Z + e];
                        jb[e] = (double)__int_as_float(rw.y);
                        wb[e] = rw.x == i ? WT(0) : mine[rw.x];
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    // the next index, read once for its row prefetch and carried over
                    if (q + 1 < N) {
                        inext = __builtin_amdgcn_readfirstlane(a.order[q + 1]);
                        fetch_row(inext);
                    }
                    const WT S_ = mine[i], O = other[i];
                    const int lb = parts > 1 ? lane % W : (lane < W ? lane : 0);
                    // Trotter neighbours of slice 2lb+tau: slices 2lb+tau-1 and 2lb+tau+1
                    const int ml = tau == 0 ? (lb == 0 ? W - 1 : lb - 1) : lb;
                    const int mr = tau == 0 ? lb : (lb == W - 1 ? 0 : lb + 1);
                    const double si = ((S_ >> lb) & one) ? -1.0 : 1.0;
                    // local field in row order (J*(+-1) is exact, so the select
                    // form adds the same values as h += J*s_j)
                    // J with its sign bit flipped by bit lb of w: shift that bit to
                    // bit 31 and xor it into J's high word
                    const uint32_t shl = 31u - (uint32_t)(lb & 31);
                    auto term = [&](double J, WT w) __attribute__((always_inline)) {
                        const uint32_t half = sizeof(WT) == 8 && lb >= 32 ? (uint32_t)((uint64_t)w >> 32) : (uint32_t)w;
                        return __longlong_as_double(__double_as_longlong(J) ^
                                                    ((long long)((half << shl) & 0x80000000u) << 32));
                    };
                    double h = 0.0;
                    int e = 0;
                    constexpr int G = 8;  // entries per group; the next group is read ahead
                    if (SPLIT) {
                        // any order gives the same double: this lane's entries are
                        // part, part + parts, ...; groups of G read one group
                        // ahead, four chains, then a butterfly over the parts
                        double hk[4] = {0.0, 0.0, 0.0, 0.0};
                        const int T = Z > part ? (Z - part + parts - 1) / parts : 0;
                        int t = 0;
                        if (T >= G) {
                            double jc[G];
                            WT wc[G];
#pragma unroll
                            for (int k = 0; k < G; ++k) {
                                jc[k] = jb[part + parts * k];
                                wc[k] = wb[part + parts * k];
                            }
                            for (; t + G <= T; t += G) {
                                const int tn = min(t + G, T - G);
                                double jn[G];
                                WT wn[G];
#pragma unroll
                                for (int k = 0; k < G; ++k) {
                                    jn[k] = jb[part + parts * (tn + k)];
                                    wn[k] = wb[part + parts * (tn + k)];
                                }
#pragma unroll
                                for (int k = 0; k < G; ++k) hk[k & 3] += term(jc[k], wc[k]);
#pragma unroll
                                for (int k = 0; k < G; ++k) {
                                    jc[k] = jn[k];
                                    wc[k] = wn[k];
                                }
                            }
                        }
                        for (; t < T; ++t) hk[0] += term(jb[part + parts * t], wb[part + parts * t]);
                        h = (hk[0] + hk[1]) + (hk[2] + hk[3]);
                        for (int sh = W; sh < 64 && parts > 1; sh <<= 1) h += __shfl_xor(h, sh, 64);
                        e = Z;
                    } else if (Z >= G) {
                        double jc[G];
                        WT wc[G];
#pragma unroll
                        for (int k = 0; k < G; ++k) {
                            jc[k] = jb[k];
                            wc[k] = wb[k];
                        }
                        for (; e + G <= Z; e += G) {
                            const int en = min(e + G, Z - G);
                            double jn[G];
                            WT wn[G];
#pragma unroll
                            for (int k = 0; k < G; ++k) {
                                jn[k] = jb[en + k];
                                wn[k] = wb[en + k];
                            }
#pragma unroll
                            for (int k = 0; k < G; ++k) h += term(jc[k], wc[k]);
#pragma unroll
                            for (int k = 0; k < G; ++k) {
                                jc[k] = jn[k];
                                wc[k] = wn[k];
                            }
                        }
                    }
                    for (; e < Z; ++e) h += term(jb[e], wb[e]);
                    const double ts = (((O >> ml) & one) ? -1.0 : 1.0) + (((O >> mr) & one) ? -1.0 : 1.0);
                    const double ed = -2.0 * si * (h + (double)jp * ts);
                    unsigned int thr = 0;
                    const int cls = piqmc_fixed_class(ed, Ts, &thr);
                    bool fl = active && cls == 1;
                    if (__ballot(active && cls == 2)) {
                        // u < thr, compared most significant bit first: bit 31-b of
                        // this lane's uniform is bit lb of random word b; a lane is
                        // decided at the first bit where u and thr differ (equal
                        // throughout means u == thr: rejected)
                        constexpr int kPer = sizeof(WT) == 4 ? 4 : 2;
                        bool und = active && cls == 2, lt = false;
                        for (int qd = 0; qd < 32 / kPer && __ballot(und); ++qd) {
                            uint32_t out[4];
                            philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qd, a.key0,
                                          a.key1, out);
#pragma unroll
                            for (int t = 0; t < kPer; ++t) {
                                const uint32_t ub = (uint32_t)((rword<WT>(out, t) >> lb) & one);
                                const uint32_t tb = (thr >> (31 - (qd * kPer + t))) & 1u;
                                if (und && ub != tb) {
                                    lt = ub < tb;
                                    und = false;
                                }
                            }
                        }
                        if (active && cls == 2) fl = lt;
                    }
                    const unsigned long long bal = __ballot(fl);  // part-0 lanes are slices 0..W-1
                    if (lane == 0) mine[i] = S_ ^ (WT)bal;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
        }
        if (a.etrace) {
            double acc = 0.0;
            for (int i = lane; i < N; i += 64) {
                const WT a0 = st[i], a1 = st[N + i];
                const int2 *row = a.rows + (size_t)i * a.Z;
                for (int e = 0; e < a.Z; ++e) {
                    const int2 rw = row[e];
                    const double J = (double)__int_as_float(rw.y);
                    if (J == 0.0) continue;
                    if (rw.x == i) acc -= J * (double)(a.P - 2 * (popc(a0) + popc(a1)));
                    else acc -= 0.5 * J * (double)(a.P - 2 * (popc(a0 ^ st[rw.x]) + popc(a1 ^ st[N + rw.x])));
                }
            }
            acc = wave_sum(acc);
            if (lane == 0) a.etrace[(size_t)r * a.S + s] = acc / (double)a.P;
        }
    }
    for (int q = lane; q < 2 * N; q += 64) gst[q] = st[q];
}

// LDS of the workgroup-per-replica dense sweep: state [2][N], staged row
// (couplings [Z] as double, neighbour words [Z]), partial fields [4][64].
__host__ __device__ inline size_t dense_wg_bytes(int N, int Z, int word_bytes)
{
    return dense_wave_bytes(N, Z, word_bytes) + 4 * 64 * sizeof(double);
}

// Dense couplings whose rows sum exactly in double in any order, one replica
// per workgroup of four wavefronts.  Per spin the whole workgroup stages the
// row, every wavefront sums a quarter of it for all slices (lane = slice, and
// 64/W lanes per slice when W divides 32), and wavefront 0 adds the four
// partial fields and decides exactly as the lane-per-slice body does: the
// same doubles, the same Philox words, the same flips.  Three barriers per
// spin buy four times the lanes for one replica's dependency chain.
template <typename WT>
__device__ void dense_wg_replica(const ColArgs<WT> &a, int r, unsigned char *base)
{
    const int N = a.N, W = a.W, Z = a.Z, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    WT *st = reinterpret_cast<WT *>(base);
    double *jb = reinterpret_cast<double *>(base + (((size_t)2 * N * sizeof(WT) + 7) & ~(size_t)7));
    WT *wb = reinterpret_cast<WT *>(jb + Z);
    double *ps = reinterpret_cast<double *>(base + dense_wave_bytes(N, Z, sizeof(WT)));  // [4][64]
    const int parts = (W <= 32 && 64 % W == 0) ? 64 / W : 1;
    const int lb = parts > 1 ? lane % W : (lane < W ? lane : 0);
    const int gp = wv * parts + (parts > 1 ? lane / W : 0), stride = 4 * parts;
    const bool active = lane < W;  // in wavefront 0: the lanes that decide slice lb
    const WT one = WT(1);
    const uint32_t rep = a.replica_base + (uint32_t)r;
    WT *gst = a.packed + (size_t)r * 2 * N;
    for (int q = tid; q < 2 * N; q += 256) st[q] = gst[q];
    __syncthreads();  // the staging below reads spins that other wavefronts loaded
    constexpr int kPre = 2;  // row entries tid + 256 k, read one spin ahead
    int2 pre[kPre];
    auto fetch_row = [&](int i) {
        const int2 *row = a.rows + (size_t)i * Z;
#pragma unroll
        for (int k = 0; k < kPre; ++k) pre[k] = row[min(tid + 256 * k, Z - 1)];
    };
    const uint32_t shl = 31u - (uint32_t)(lb & 31);
    auto term = [&](double J, WT w) __attribute__((always_inline)) {
        const uint32_t half = sizeof(WT) == 8 && lb >= 32 ? (uint32_t)((uint64_t)w >> 32) : (uint32_t)w;
        return __longlong_as_double(__double_as_longlong(J) ^ ((long long)((half << shl) & 0x80000000u) << 32));
    };
    for (int s = 0; s < a.S; ++s) {
        const float jp = a.jperp[s];
        const double Ts = (double)a.temps[s];
        for (int m = 0; m < a.M; ++m) {
            const uint32_t g = a.sweep_base + (uint32_t)(s * a.M + m);
            for (int tau = 0; tau < 2; ++tau) {
                WT *mine = st + tau * N;
                const WT *other = st + (1 - tau) * N;
                int inext = __builtin_amdgcn_readfirstlane(a.order[0]);
                fetch_row(inext);
                for (int q = 0; q < N; ++q) {
                    const int i = inext;
#pragma unroll
                    for (int k = 0; k < kPre; ++k) {
                        const int e = tid + 256 * k;
                        if (e < Z) {
                            jb[e] = (double)__int_as_float(pre[k].y);
                            wb[e] = pre[k].x == i ? WT(0) : mine[pre[k].x];
                        }
                    }
                    for (int e = tid + 256 * kPre; e < Z; e += 256) {
                        const int2 rw = a.rows[(size_t)i * Z + e];
                        jb[e] = (double)__int_as_float(rw.y);
                        wb[e] = rw.x == i ? WT(0) : mine[rw.x];
                    }
                    __syncthreads();
                    // the next index, read once for its row prefetch and carried over
                    if (q + 1 < N) {
                        inext = __builtin_amdgcn_readfirstlane(a.order[q + 1]);
                        fetch_row(inext);
                    }
                    // this lane's entries gp, gp + stride, ...: with one lane per
                    // slice (long runs) in groups of G read one group ahead, four
                    // chains; with several lanes per slice (runs of 8-32) two
                    // chains read in place were measured faster
                    constexpr int G = 8;
                    double hk[4] = {0.0, 0.0, 0.0, 0.0};
                    const int T = Z > gp ? (Z - gp + stride - 1) / stride : 0;
                    int t = 0;
                    if (parts > 1) {
                        int e = gp;
                        for (; e + stride < Z; e += 2 * stride) {
                            hk[0] += term(jb[e], wb[e]);
                            hk[1] += term(jb[e + stride], wb[e + stride]);
                        }
                        if (e < Z) hk[0] += term(jb[e], wb[e]);
                        t = T;
                    } else if (T >= G) {
                        double jc[G];
                        WT wc[G];
#pragma unroll
                        for (int k = 0; k < G; ++k) {
                            jc[k] = jb[gp + stride * k];
                            wc[k] = wb[gp + stride * k];
                        }
                        for (; t + G <= T; t += G) {
                            const int tn = min(t + G, T - G);
                            double jn[G];
                            WT wn[G];
#pragma unroll
                            for (int k = 0; k < G; ++k) {
                                jn[k] = jb[gp + stride * (tn + k)];
                                wn[k] = wb[gp + stride * (tn + k)];
                            }
#pragma unroll
                            for (int k = 0; k < G; ++k) hk[k & 3] += term(jc[k], wc[k]);
#pragma unroll
                            for (int k = 0; k < G; ++k) {
                                jc[k] = jn[k];
                                wc[k] = wn[k];
                            }
                        }
                    }
                    for (; t < T; ++t) hk[0] += term(jb[gp + stride * t], wb[gp + stride * t]);
                    double hp = (hk[0] + hk[1]) + (hk[2] + hk[3]);
                    for (int sh = W; sh < 64 && parts > 1; sh <<= 1) hp += __shfl_xor(hp, sh, 64);
                    if (lane < W) ps[wv * 64 + lane] = hp;
                    __syncthreads();
                    if (wv == 0) {
                        const double h = (ps[lb] + ps[64 + lb]) + (ps[128 + lb] + ps[192 + lb]);
                        const WT S_ = mine[i], O = other[i];
                        const int ml = tau == 0 ? (lb == 0 ? W - 1 : lb - 1) : lb;
                        const int mr = tau == 0 ? lb : (lb == W - 1 ? 0 : lb + 1);
                        const double si = ((S_ >> lb) & one) ? -1.0 : 1.0;
                        const double ts = (((O >> ml) & one) ? -1.0 : 1.0) + (((O >> mr) & one) ? -1.0 : 1.0);
                        const double ed = -2.0 * si * (h + (double)jp * ts);
                        unsigned int thr = 0;
                        const int cls = piqmc_fixed_class(ed, Ts, &thr);
                        bool fl = active && cls == 1;
                        if (__ballot(active && cls == 2)) {
                            constexpr int kPer = sizeof(WT) == 4 ? 4 : 2;
                            bool und = active && cls == 2, lt = false;
                            for (int qd = 0; qd < 32 / kPer && __ballot(und); ++qd) {
                                uint32_t out[4];
                                philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qd, a.key0,
                                              a.key1, out);
#pragma unroll
                                for (int t = 0; t < kPer; ++t) {
                                    const uint32_t ub = (uint32_t)((rword<WT>(out, t) >> lb) & one);
                                    const uint32_t tb = (thr >> (31 - (qd * kPer + t))) & 1u;
                                    if (und && ub != tb) {
                                        lt = ub < tb;
                                        und = false;
                                    }
                                }
                            }
                            if (active && cls == 2) fl = lt;
                        }
                        const unsigned long long bal = __ballot(fl);
                        if (lane == 0) mine[i] = S_ ^ (WT)bal;
                    }
                    __syncthreads();
                }
            }
        }
        if (a.etrace && wv == 0) {
            double acc = 0.0;
            for (int i = lane; i < N; i += 64) {
                const WT a0 = st[i], a1 = st[N + i];
                const int2 *row = a.rows + (size_t)i * a.Z;
                for (int e = 0; e < a.Z; ++e) {
                    const int2 rw = row[e];
                    const double J = (double)__int_as_float(rw.y);
                    if (J == 0.0) continue;
                    if (rw.x == i) acc -= J * (double)(a.P - 2 * (popc(a0) + popc(a1)));
                    else acc -= 0.5 * J * (double)(a.P - 2 * (popc(a0 ^ st[rw.x]) + popc(a1 ^ st[N + rw.x])));
                }
            }
            acc = wave_sum(acc);
            if (lane == 0) a.etrace[(size_t)r * a.S + s] = acc / (double)a.P;
        }
    }
    __syncthreads();
    for (int q = tid; q < 2 * N; q += 256) gst[q] = st[q];
}

template <typename WT>
__global__ void __launch_bounds__(256) dense_wg_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    dense_wg_replica<WT>(a, blockIdx.x, smem);
}

// Workgroup barrier that orders LDS accesses only: global loads issued
// before it (prefetches) stay outstanding instead of being waited for.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The same sweep pipelined over the workgroup's four wavefronts: wavefront 0
// decides spin order[q] while wavefronts 1-3 stage and sum the row of
// order[q+1], one barrier per spin.  The entries of that row for order[q]
// (whose flips are being decided) are left out of the sums, and wavefront 0
// adds them back, with order[q]'s new spins, before deciding order[q+1]
// (prevj holds their summed coupling).  Exactly summable rows make every
// partial and the completed field the exact field, so the doubles, the
// Philox words and the flips are those of the lane-per-slice body.  Each
// summing wavefront stages only its own entries (wavefront-private LDS, no
// barrier between staging and summing); partial sums are double-buffered.
// A staged entry is one 16-byte record (coupling, spin word) that a lane reads
// with one LDS instruction; with 64-bit words a second copy carries the high
// half for lanes 32-63.  Each lane group's records are contiguous.
// nsw: summing wavefronts (the workgroup's wavefronts but one)
__host__ __device__ inline int dense_pipe_slots(int Z, int nsw) { return (Z + nsw - 1) / nsw + 64; }
__host__ __device__ inline size_t dense_pipe_bytes(int N, int Z, int word_bytes, int nsw = 3)
{
    const size_t state = ((size_t)2 * N * word_bytes + 15) & ~(size_t)15;
    const size_t copies = word_bytes == 8 ? 2 : 1;
    // + the spin order and the predecessor couplings, [N] each
    return state + 2 * (size_t)nsw * 64 * sizeof(double) + (size_t)nsw * copies * dense_pipe_slots(Z, nsw) * 16 +
           (size_t)N * 12;
}

template <typename WT>
__device__ void dense_pipe_replica(const ColArgs<WT> &a, int r, unsigned char *base)
{
    const int N = a.N, W = a.W, Z = a.Z, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    WT *st = reinterpret_cast<WT *>(base);
    double *ps = reinterpret_cast<double *>(base + (((size_t)2 * N * sizeof(WT) + 15) & ~(size_t)15));  // [2][nsw][64]
    const int nsw = (int)(blockDim.x >> 6) - 1, nth = (int)blockDim.x;
    const int C = dense_pipe_slots(Z, nsw);
    constexpr int kCopies = sizeof(WT) == 8 ? 2 : 1;
    const int sw = wv - 1;  // summing wavefront 0..2 (wavefront 0 decides)
    uint4 *rec = reinterpret_cast<uint4 *>(ps + 2 * nsw * 64) + (size_t)(sw < 0 ? 0 : sw) * kCopies * C;
    double *pj = reinterpret_cast<double *>(reinterpret_cast<uint4 *>(ps + 2 * nsw * 64) + (size_t)nsw * kCopies * C);  // [N]
    int *ord = reinterpret_cast<int *>(pj + N);  // [N]
    const int parts = (W <= 32 && 64 % W == 0) ? 64 / W : 1;
    const int lb = parts > 1 ? lane % W : (lane < W ? lane : 0);
    const int grp = parts > 1 ? lane / W : 0;
    const int gp = sw * parts + grp, stride = nsw * parts;
    const int Cg = (Z + stride - 1) / stride;  // records per lane group
    const uint4 *rb = rec + (sizeof(WT) == 8 && lb >= 32 ? C : 0) + grp * Cg;
    const bool active = lane < W;  // in wavefront 0: the lanes that decide slice lb
    const WT one = WT(1);
    const uint32_t rep = a.replica_base + (uint32_t)r;
    WT *gst = a.packed + (size_t)r * 2 * N;
    for (int q = tid; q < 2 * N; q += nth) st[q] = gst[q];
    for (int q = tid; q < N; q += nth) {
        ord[q] = a.order[q];
        pj[q] = a.prevj[q];
    }
    // slot t of summing wavefront sw is row entry slot_e(t), stored as record
    // (t % parts) Cg + t / parts: lane group grp reads records grp Cg + u,
    // u = 0, 1, ... = entries gp + u stride
    auto slot_e = [&](int t) { return (t / parts) * stride + sw * parts + t % parts; };
    auto stage = [&](int t, int2 rw, int i, int ex, const WT *mine) __attribute__((always_inline)) {
        const bool x = rw.x == ex;
        const double J = x ? 0.0 : (double)__int_as_float(rw.y);
        const WT w = (x || rw.x == i) ? WT(0) : mine[rw.x];
        const int pos = (t % parts) * Cg + t / parts;
        const uint32_t jlo = (uint32_t)__double2loint(J), jhi = (uint32_t)__double2hiint(J);
        rec[pos] = make_uint4(jlo, jhi, (uint32_t)w, 0u);
        if constexpr (sizeof(WT) == 8) rec[C + pos] = make_uint4(jlo, jhi, (uint32_t)((uint64_t)w >> 32), 0u);
    };
    constexpr int kPre = 2;  // slots lane + 64 k, read one spin ahead
    int2 pre[kPre];
    auto fetch_row = [&](int i) {
        const int2 *row = a.rows + (size_t)i * Z;
#pragma unroll
        for (int k = 0; k < kPre; ++k) pre[k] = row[min(slot_e(lane + 64 * k), Z - 1)];
    };
    const uint32_t shl = 31u - (uint32_t)(lb & 31);
    auto term = [&](double J, WT w) __attribute__((always_inline)) {
        const uint32_t half = sizeof(WT) == 8 && lb >= 32 ? (uint32_t)((uint64_t)w >> 32) : (uint32_t)w;
        return __longlong_as_double(__double_as_longlong(J) ^ ((long long)((half << shl) & 0x80000000u) << 32));
    };
    // a record's coupling with the sign of this lane's spin applied
    auto rterm = [&](uint4 r) __attribute__((always_inline)) {
        return __hiloint2double((int)(r.y ^ ((r.z << shl) & 0x80000000u)), (int)r.x);
    };
    const long long K = (long long)a.S * a.M * 2 * N;
    int q = 0, tau = 0, m = 0, s = 0;   // step k
    int pq = 0, ptau = 0, pm = 0, pso = 0;  // step k - 1
    int iprev = -1;                      // spin of step k - 1 in the same half-sweep, else -1
    WT lastw = 0;                        // wavefront 0: new word of the spin it decided last
    if (sw >= 0) fetch_row(__builtin_amdgcn_readfirstlane(a.order[0]));
    lds_barrier();
    float jp = 0.0f;  // wavefront 0: J_perp and temperature of field step pso
    double Ts = 0.0;
    int js = -1;
    for (long long k = 0; k <= K; ++k) {
        if (wv == 0) {
            if (k > 0) {
                const int i = __builtin_amdgcn_readfirstlane(ord[pq]);
                WT *mine = st + ptau * N;
                const WT *other = st + (1 - ptau) * N;
                const double *pp = ps + ((k - 1) & 1) * nsw * 64;
                double h = pp[lb];
                for (int w = 1; w < nsw; ++w) h += pp[w * 64 + lb];
                if (pq > 0) h += term(pj[pq], lastw);
                if (pso != js) {
                    js = pso;
                    jp = a.jperp[pso];
                    Ts = (double)a.temps[pso];
                }
                const uint32_t g = a.sweep_base + (uint32_t)(pso * a.M + pm);
                const WT S_ = mine[i], O = other[i];
                lastw = S_ ^ dense_decide<WT>(a, h, S_, O, lane, ptau, jp, Ts, g, (int)i, rep);
                if (lane == 0) mine[i] = lastw;
                if (a.etrace && pq == N - 1 && ptau == 1 && pm == a.M - 1) {
                    double acc = 0.0;
                    for (int i2 = lane; i2 < N; i2 += 64) {
                        const WT a0 = st[i2], a1 = st[N + i2];
                        const int2 *row = a.rows + (size_t)i2 * a.Z;
                        for (int e = 0; e < a.Z; ++e) {
                            const int2 rw = row[e];
                            const double J = (double)__int_as_float(rw.y);
                            if (J == 0.0) continue;
                            if (rw.x == i2) acc -= J * (double)(a.P - 2 * (popc(a0) + popc(a1)));
                            else acc -= 0.5 * J * (double)(a.P - 2 * (popc(a0 ^ st[rw.x]) + popc(a1 ^ st[N + rw.x])));
                        }
                    }
                    acc = wave_sum(acc);
                    if (lane == 0) a.etrace[(size_t)r * a.S + pso] = acc / (double)a.P;
                }
            }
        } else if (k < K) {
            const int i = __builtin_amdgcn_readfirstlane(ord[q]);
            const WT *mine = st + tau * N;
            // stage this wavefront's entries of row i; the entries for iprev
            // (decided concurrently by wavefront 0) count zero here
#pragma unroll
            for (int kk = 0; kk < kPre; ++kk) {
                const int t = lane + 64 * kk;
                if (slot_e(t) < Z) stage(t, pre[kk], i, iprev, mine);
            }
            for (int t = lane + 64 * kPre; slot_e(t) < Z; t += 64)
                stage(t, a.rows[(size_t)i * Z + slot_e(t)], i, iprev, mine);
            if (k + 1 < K) fetch_row(__builtin_amdgcn_readfirstlane(ord[q + 1 < N ? q + 1 : 0]));
            // the slots were written by this wavefront: LDS keeps its order
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            constexpr int G = 8;
            double hk[4] = {0.0, 0.0, 0.0, 0.0};
            const int T = Z > gp ? (Z - gp + stride - 1) / stride : 0;
            int u = 0;
            if (T >= G) {
                // groups of G records, the next group read while this one is summed
                uint4 rr[G];
#pragma unroll
                for (int kk = 0; kk < G; ++kk) rr[kk] = rb[kk];
                for (; u + G <= T; u += G) {
                    const int un = min(u + G, T - G);
                    uint4 rn[G];
#pragma unroll
                    for (int kk = 0; kk < G; ++kk) rn[kk] = rb[un + kk];
#pragma unroll
                    for (int kk = 0; kk < G; ++kk) hk[kk & 3] += rterm(rr[kk]);
#pragma unroll
                    for (int kk = 0; kk < G; ++kk) rr[kk] = rn[kk];
                }
            }
            for (; u < T; ++u) hk[0] += rterm(rb[u]);
            double hp = (hk[0] + hk[1]) + (hk[2] + hk[3]);
            for (int sh = W; sh < 64 && parts > 1; sh <<= 1) hp += __shfl_xor(hp, sh, 64);
            if (lane < W) ps[(k & 1) * nsw * 64 + sw * 64 + lane] = hp;
        }
        // the wavefronts exchange data through LDS only, so the row prefetch
        // stays in flight across the barrier
        lds_barrier();
        pq = q; ptau = tau; pm = m; pso = s;
        iprev = __builtin_amdgcn_readfirstlane(ord[q]);
        if (++q == N) {
            q = 0;
            iprev = -1;
            if (++tau == 2) {
                tau = 0;
                if (++m == a.M) {
                    m = 0;
                    ++s;
                }
            }
        }
    }
    for (int q2 = tid; q2 < 2 * N; q2 += nth) gst[q2] = st[q2];
}

template <typename WT>
__global__ void __launch_bounds__(512) dense_pipe_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    dense_pipe_replica<WT>(a, blockIdx.x, smem);
}

// Wavefronts per pipelined workgroup: eight (seven summing) when a lane's
// share of the row with three summing wavefronts is 64 terms or more and the
// replicas are few enough for one workgroup per CU (256-spin SK at P = 128,
// 4-256 replicas: 58 -> 46 ms); four otherwise (eight were 5-70 % slower at
// P <= 64 or 1024 replicas).  PIQMC_DENSE_PIPE_WAVES = 4 or 8 forces a count.
inline int dense_pipe_waves(int Z, int W, int R)
{
    const char *e = std::getenv("PIQMC_DENSE_PIPE_WAVES");
    if (e && (e[0] == '4' || e[0] == '8')) return e[0] - '0';
    const int parts = (W <= 32 && 64 % W == 0) ? 64 / W : 1;
    return ((Z + 3 * parts - 1) / (3 * parts) >= 64 && R <= 256) ? 8 : 4;
}

inline bool dense_pipe_enabled()
{
    const char *e = std::getenv("PIQMC_DENSE_PIPE");
    return !(e && e[0] == '0');
}

// One replica per wavefront, four wavefronts per workgroup, no barriers.
template <typename WT>
__global__ void dense_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = threadIdx.x >> 6, r = blockIdx.x * (blockDim.x >> 6) + w;
    if (r >= a.R) return;
    dense_replica<WT>(a, r, smem + (size_t)w * dense_wave_bytes(a.N, a.Z, sizeof(WT)));
}

// The same with each slice's field split over 64 / W lanes (exactly summable
// rows, W dividing 32); at most four wavefronts per workgroup, which leaves
// the split sums their registers.
constexpr int kDenseMaxBlock = 256;
template <typename WT>
__global__ void __launch_bounds__(kDenseMaxBlock) dense_split_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = threadIdx.x >> 6, r = blockIdx.x * (blockDim.x >> 6) + w;
    if (r >= a.R) return;
    dense_replica<WT, true>(a, r, smem + (size_t)w * dense_wave_bytes(a.N, a.Z, sizeof(WT)));
}

// ------------------------------------------------------------------ dense path, cached fields
// For rows that sum exactly in any order and list each neighbour once, with
// the same coupling in both directions (piqmc_problem's fields_ok), the local
// field of every spin in every slice of the parity being swept stays in LDS:
// a decision reads one field per slice instead of summing a row, and a flip
// of spin i in slice m adds 2 J_ij s_i^new to the field of each neighbour j
// in slice m.  Every field is a signed sum of its row's couplings, so with
// exact rows each update lands on exactly the double a fresh row sum gives:
// the decisions, Philox words and flips are those of the other dense bodies.
// A parity's fields change only while that parity is swept, so the two
// tables live in global memory and the active one is staged per half sweep.
//
// LDS of one replica: state [2][N] words, then (8-byte aligned) fields [W][NP]
// doubles with NP = N rounded up to odd: a decision reads one field per slice
// at a stride of NP doubles, an update writes consecutive spins of one slice,
// and neither conflicts on the banks.
__host__ __device__ inline int fields_pitch(int N) { return N | 1; }
// Rows the cached-field update covers: four entries per lane of a wavefront.
constexpr int kFieldsMaxRow = 256;
__host__ __device__ inline size_t dense_fields_bytes(int N, int W, int word_bytes)
{
    const size_t state = ((size_t)2 * N * word_bytes + 7) & ~(size_t)7;
    return state + (size_t)W * fields_pitch(N) * sizeof(double) + (((size_t)N * sizeof(int) + 7) & ~(size_t)7);
}

// Wavefronts that share one replica of the cached-field sweep.  Standalone
// (waves_per_wg 0): eight, four or two when only one replica's tables fit a
// workgroup's LDS.  In a workgroup of waves_per_wg wavefronts holding
// replicas_per_wg replicas: waves_per_wg / replicas_per_wg when that divides
// evenly and splits the slices evenly.  1: one wavefront per replica
// (PIQMC_FIELDS_SPLIT=0 forces it).
inline int fields_split_waves(int W, int replicas_per_wg, int waves_per_wg = 0)
{
    const char *e = std::getenv("PIQMC_FIELDS_SPLIT");
    if (e && e[0] == '0') return 1;
    if (waves_per_wg > 0) {
        if (replicas_per_wg < 1 || waves_per_wg % replicas_per_wg) return 1;
        const int nw = waves_per_wg / replicas_per_wg;
        return nw >= 2 && W % nw == 0 && W / nw >= 2 ? nw : 1;
    }
    if (replicas_per_wg > 1) return 1;
    // eight wavefronts (a 512-thread workgroup) where the slices allow: each
    // updates fewer flipped slices per spin on the shared chain (256-spin SK,
    // 4 replicas: P=16 57.2 -> 51.5 ms, P=128 16.8 -> 14.1 ms against four)
    if (W % 8 == 0 && W >= 8) return 8;
    if (W % 4 == 0 && W >= 8) return 4;
    if (W % 2 == 0 && W >= 4) return 2;
    return 1;
}

// Initial local fields of the cached-field sweep, one thread per (replica,
// parity, slice, spin): the row's sum with a field entry counted as +J, in
// row order (exact rows give the same double in any order).
template <typename WT>
__global__ void fields_init_kernel(ColArgs<WT> a)
{
    const int N = a.N, W = a.W, Z = a.Z, NP = fields_pitch(N);
    const size_t tot = (size_t)a.R * 2 * W * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(t % N);
        const size_t rest = t / N;
        const int m = (int)(rest % W), tau = (int)((rest / W) % 2), r = (int)(rest / (2 * (size_t)W));
        const WT *w = a.packed + ((size_t)r * 2 + tau) * N;
        double h = 0.0;
        for (int e = 0; e < Z; ++e) {
            const int2 rw = a.rows[(size_t)j * Z + e];
            const double J = (double)__int_as_float(rw.y);
            const WT wj = rw.x == j ? WT(0) : w[rw.x];
            h += ((wj >> m) & WT(1)) ? -J : J;
        }
        a.fields[(((size_t)r * 2 + tau) * W + m) * NP + j] = h;
    }
}

template <typename WT>
__device__ void dense_fields_replica(const ColArgs<WT> &a, int r, unsigned char *base)
{
    const int N = a.N, W = a.W, Z = a.Z, lane = threadIdx.x & 63, NP = fields_pitch(N);
    WT *st = reinterpret_cast<WT *>(base);
    double *tab = reinterpret_cast<double *>(base + (((size_t)2 * N * sizeof(WT) + 7) & ~(size_t)7));
    const WT one = WT(1);
    const int lb = lane < W ? lane : 0;
    const uint32_t rep = a.replica_base + (uint32_t)r;
    WT *gst = a.packed + (size_t)r * 2 * N;
    double *gt = a.fields + (size_t)r * 2 * W * NP;  // [2][W][NP]
    // the spin order, then the state, into LDS (the fields of both parities
    // were computed by fields_init_kernel)
    int *ord = reinterpret_cast<int *>(tab + (size_t)W * NP);
    for (int q = lane; q < N; q += 64) ord[q] = a.order[q];
    for (int q = lane; q < 2 * N; q += 64) st[q] = gst[q];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr int kPre = 4;  // row entries per lane held in registers (rows up to 256)
    int2 pre[kPre], cur[kPre];
    auto fetch_row = [&](int i) {
        const int2 *row = a.rows + (size_t)i * Z;
#pragma unroll
        for (int k = 0; k < kPre; ++k) pre[k] = row[min(lane + 64 * k, Z - 1)];
    };
    for (int s = 0; s < a.S; ++s) {
        const float jp = a.jperp[s];
        const double Ts = (double)a.temps[s];
        for (int m = 0; m < a.M; ++m) {
            const uint32_t g = a.sweep_base + (uint32_t)(s * a.M + m);
            for (int tau = 0; tau < 2; ++tau) {
                WT *mine = st + tau * N;
                const WT *other = st + (1 - tau) * N;
                double *gtt = gt + (size_t)tau * W * NP;
                for (int t = lane; t < W * NP; t += 64) tab[t] = gtt[t];
                int inext = __builtin_amdgcn_readfirstlane(ord[0]);
                fetch_row(inext);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (int q = 0; q < N; ++q) {
                    const int i = inext;
#pragma unroll
                    for (int k = 0; k < kPre; ++k) cur[k] = pre[k];
                    // the next index, read once for its row prefetch and carried over
                    // (re-reading it put one more LDS round trip on the decision chain)
                    if (q + 1 < N) { inext = __builtin_amdgcn_readfirstlane(ord[q + 1]); fetch_row(inext); }
                    const double h = tab[(size_t)lb * NP + i];
                    const WT S_ = mine[i], O = other[i];
                    const WT bal = dense_decide<WT, true>(a, h, S_, O, lane, tau, jp, Ts, g, i, rep);
                    if (lane == 0) mine[i] = S_ ^ bal;
                    // neighbours' fields in the slices where spin i flipped:
                    // + 2 J s_i^new (old bit set = -1, so the new spin is +1).
                    // Up to kFl flipped slices per pass: all their reads issue
                    // before any add, so a pass costs one LDS round trip.
                    constexpr int kFl = 4;
                    int jj[kPre];
                    double jv[kPre];
#pragma unroll
                    for (int k = 0; k < kPre; ++k) {
                        const float J = __int_as_float(cur[k].y);
                        const bool use = lane + 64 * k < Z && cur[k].x != i && J != 0.0f;
                        jj[k] = use ? cur[k].x : -1;
                        jv[k] = 2.0 * (double)J;
                    }
                    WT rem = bal;
                    while (rem) {
                        int mm[kFl];
                        double sg[kFl];
#pragma unroll
                        for (int f = 0; f < kFl; ++f) {
                            mm[f] = rem ? ctz(rem) : -1;
                            sg[f] = mm[f] >= 0 && ((S_ >> mm[f]) & one) ? 1.0 : -1.0;
                            rem &= rem - (rem ? WT(1) : WT(0));
                        }
                        double v[kFl][kPre];
#pragma unroll
                        for (int f = 0; f < kFl; ++f)
#pragma unroll
                            for (int k = 0; k < kPre; ++k)
                                v[f][k] = (mm[f] >= 0 && jj[k] >= 0) ? tab[(size_t)mm[f] * NP + jj[k]] : 0.0;
#pragma unroll
                        for (int f = 0; f < kFl; ++f)
#pragma unroll
                            for (int k = 0; k < kPre; ++k)
                                if (mm[f] >= 0 && jj[k] >= 0) tab[(size_t)mm[f] * NP + jj[k]] = v[f][k] + sg[f] * jv[k];
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
                for (int t = lane; t < W * NP; t += 64) gtt[t] = tab[t];
            }
        }
        if (a.etrace) {
            double acc = 0.0;
            for (int i = lane; i < N; i += 64) {
                const WT a0 = st[i], a1 = st[N + i];
                const int2 *row = a.rows + (size_t)i * a.Z;
                for (int e = 0; e < a.Z; ++e) {
                    const int2 rw = row[e];
                    const double J = (double)__int_as_float(rw.y);
                    if (J == 0.0) continue;
                    if (rw.x == i) acc -= J * (double)(a.P - 2 * (popc(a0) + popc(a1)));
                    else acc -= 0.5 * J * (double)(a.P - 2 * (popc(a0 ^ st[rw.x]) + popc(a1 ^ st[N + rw.x])));
                }
            }
            acc = wave_sum(acc);
            if (lane == 0) a.etrace[(size_t)r * a.S + s] = acc / (double)a.P;
        }
    }
    for (int q = lane; q < 2 * N; q += 64) gst[q] = st[q];
}

// One replica over the nw wavefronts of a workgroup: wavefront wv owns slices
// [wv*Ws, (wv+1)*Ws), Ws = W/nw, and their rows of the field table (the
// layout of dense_fields_replica: the table's W rows split among the
// wavefronts).  Slices of one parity do not interact within a half sweep, so
// the wavefronts decide their slices of every spin independently: each reads
// the shared state word (only its own bits matter to it), XORs its flips in
// atomically, and updates its own rows.  They meet at a workgroup barrier
// between half sweeps, where the other parity becomes the Trotter input.
// Every wavefront draws the word's Philox blocks itself (the counter depends on
// the word, not the slice), so the decisions equal the one-wavefront sweep's.
// Used where one replica's tables fill the LDS and the other wavefronts of the
// workgroup would idle: the chain of N decisions per half sweep, which sets a
// lone replica's time, is then shared out.  FL: flipped slices per update pass,
// at most the wavefront's slices (W / nw), so a pass issues no reads for slots
// that can never hold a flip.
template <typename WT, int FL = 4>
__device__ void dense_fields_split(const ColArgs<WT> &a, int r, unsigned char *base, int wv, int nw)
{
    const int N = a.N, W = a.W, Z = a.Z, lane = threadIdx.x & 63, NP = fields_pitch(N);
    const int Ws = W / nw, s0 = wv * Ws;
    WT *st = reinterpret_cast<WT *>(base);
    double *tab0 = reinterpret_cast<double *>(base + (((size_t)2 * N * sizeof(WT) + 7) & ~(size_t)7));
    double *tab = tab0 + (size_t)s0 * NP;  // this wavefront's rows
    const WT one = WT(1);
    const int ll = lane < Ws ? lane : 0;  // local slice
    const int sl = lane < Ws ? s0 + lane : W;  // global slice (W: inactive lane)
    const uint32_t rep = a.replica_base + (uint32_t)r;
    WT *gst = a.packed + (size_t)r * 2 * N;
    double *gt = a.fields + (size_t)r * 2 * W * NP;  // [2][W][NP]
    int *ord = reinterpret_cast<int *>(tab0 + (size_t)W * NP);
    const int tid = wv * 64 + lane, bd = nw * 64;  // this replica's wavefronts (a workgroup may hold several)
    for (int q = tid; q < N; q += bd) ord[q] = a.order[q];
    for (int q = tid; q < 2 * N; q += bd) st[q] = gst[q];
    __syncthreads();
    constexpr int kPre = 4;  // row entries per lane held in registers (rows up to 256)
    int2 pre[kPre], cur[kPre];
    auto fetch_row = [&](int i) {
        const int2 *row = a.rows + (size_t)i * Z;
#pragma unroll
        for (int k = 0; k < kPre; ++k) pre[k] = row[min(lane + 64 * k, Z - 1)];
    };
    for (int s = 0; s < a.S; ++s) {
        const float jp = a.jperp[s];
        const double Ts = (double)a.temps[s];
        for (int m = 0; m < a.M; ++m) {
            const uint32_t g = a.sweep_base + (uint32_t)(s * a.M + m);
            for (int tau = 0; tau < 2; ++tau) {
                WT *mine = st + tau * N;
                const WT *other = st + (1 - tau) * N;
                double *gtt = gt + (size_t)tau * W * NP + (size_t)s0 * NP;
                for (int t = lane; t < Ws * NP; t += 64) tab[t] = gtt[t];
                int inext = __builtin_amdgcn_readfirstlane(ord[0]);
                fetch_row(inext);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (int q = 0; q < N; ++q) {
                    const int i = inext;
#pragma unroll
                    for (int k = 0; k < kPre; ++k) cur[k] = pre[k];
                    // the next index, read once for its row prefetch and carried over
                    // (re-reading it put one more LDS round trip on the decision chain)
                    if (q + 1 < N) { inext = __builtin_amdgcn_readfirstlane(ord[q + 1]); fetch_row(inext); }
                    const double h = tab[(size_t)ll * NP + i];
                    const WT S_ = mine[i], O = other[i];
                    // the ballot is over lanes: lane l decided slice s0 + l
                    const WT bal = dense_decide<WT, true>(a, h, S_, O, sl, tau, jp, Ts, g, i, rep) << s0;
                    if (lane == 0 && bal) atomicXor(&mine[i], bal);
                    constexpr int kFl = FL;
                    int jj[kPre];
                    double jv[kPre];
#pragma unroll
                    for (int k = 0; k < kPre; ++k) {
                        const float J = __int_as_float(cur[k].y);
                        const bool use = lane + 64 * k < Z && cur[k].x != i && J != 0.0f;
                        jj[k] = use ? cur[k].x : -1;
                        jv[k] = 2.0 * (double)J;
                    }
                    WT rem = bal;
                    while (rem) {
                        int mm[kFl];
                        double sg[kFl];
#pragma unroll
                        for (int f = 0; f < kFl; ++f) {
                            mm[f] = rem ? ctz(rem) : -1;
                            sg[f] = mm[f] >= 0 && ((S_ >> mm[f]) & one) ? 1.0 : -1.0;
                            rem &= rem - (rem ? WT(1) : WT(0));
                        }
                        double v[kFl][kPre];
#pragma unroll
                        for (int f = 0; f < kFl; ++f)
#pragma unroll
                            for (int k = 0; k < kPre; ++k)
                                v[f][k] = (mm[f] >= 0 && jj[k] >= 0) ? tab[(size_t)(mm[f] - s0) * NP + jj[k]] : 0.0;
#pragma unroll
                        for (int f = 0; f < kFl; ++f)
#pragma unroll
                            for (int k = 0; k < kPre; ++k)
                                if (mm[f] >= 0 && jj[k] >= 0)
                                    tab[(size_t)(mm[f] - s0) * NP + jj[k]] = v[f][k] + sg[f] * jv[k];
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
                for (int t = lane; t < Ws * NP; t += 64) gtt[t] = tab[t];
                __syncthreads();  // this parity complete: the Trotter input of the next half sweep
            }
        }
        if (a.etrace) {
            // wavefront 0 sums the energy; the others wait before changing the state
            if (wv == 0) {
                double acc = 0.0;
                for (int i = lane; i < N; i += 64) {
                    const WT a0 = st[i], a1 = st[N + i];
                    const int2 *row = a.rows + (size_t)i * a.Z;
                    for (int e = 0; e < a.Z; ++e) {
                        const int2 rw = row[e];
                        const double J = (double)__int_as_float(rw.y);
                        if (J == 0.0) continue;
                        if (rw.x == i) acc -= J * (double)(a.P - 2 * (popc(a0) + popc(a1)));
                        else acc -= 0.5 * J * (double)(a.P - 2 * (popc(a0 ^ st[rw.x]) + popc(a1 ^ st[N + rw.x])));
                    }
                }
                acc = wave_sum(acc);
                if (lane == 0) a.etrace[(size_t)r * a.S + s] = acc / (double)a.P;
            }
            __syncthreads();
        }
    }
    for (int q = tid; q < 2 * N; q += bd) gst[q] = st[q];
}

// One replica per wavefront; as many wavefronts per workgroup as the launch
// gives (the host keeps the LDS of all of them within 160 KiB).
template <typename WT>
__global__ void __launch_bounds__(256) dense_fields_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = threadIdx.x >> 6, r = blockIdx.x * (blockDim.x >> 6) + w;
    if (r >= a.R) return;
    dense_fields_replica<WT>(a, r, smem + (size_t)w * dense_fields_bytes(a.N, a.W, sizeof(WT)));
}

// One replica per workgroup, its slices split over the workgroup's wavefronts.
template <typename WT, int FL>
__global__ void __launch_bounds__(512) dense_fields_split_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    dense_fields_split<WT, FL>(a, blockIdx.x, smem, threadIdx.x >> 6, blockDim.x >> 6);
}

// Flipped slices per update pass of the split sweep for a wavefront's slices.
template <typename WT>
const void *fields_split_kernel_for(int slices)
{
    if constexpr (sizeof(WT) == 4) {
        if (slices <= 1) return (const void *)dense_fields_split_kernel<WT, 1>;
        if (slices <= 2) return (const void *)dense_fields_split_kernel<WT, 2>;
    }
    return (const void *)dense_fields_split_kernel<WT, 4>;
}

// int8 [R][P][N] -> packed [R][2][N] (bit m of parity tau = slice 2m+tau, set = -1)
template <typename WT>
__global__ void pack_parity_kernel(const int8_t *confs, WT *packed, int R, int P, int N)
{
    const size_t tot = (size_t)R * 2 * N;
    const int W = P / 2;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < tot;
         x += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(x % N);
        const size_t rt = x / N;
        const int tau = (int)(rt % 2), r = (int)(rt / 2);
        WT w = 0;
        for (int m = 0; m < W; ++m)
            if (confs[((size_t)r * P + 2 * m + tau) * N + i] < 0) w |= WT(1) << m;
        packed[x] = w;
    }
}

template <typename WT>
__global__ void unpack_parity_kernel(const WT *packed, int8_t *confs, int R, int P, int N)
{
    const size_t tot = (size_t)R * P * N;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < tot;
         x += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(x % N);
        const size_t rk = x / N;
        const int k = (int)(rk % P), r = (int)(rk / P);
        const WT w = packed[((size_t)r * 2 + (k & 1)) * N + i];
        confs[x] = ((w >> (k >> 1)) & WT(1)) ? (int8_t)-1 : (int8_t)1;
    }
}

// The same two layouts for 16-spin groups: one thread owns spins 16q..16q+15
// of one replica for all P slices, so every slice row moves as one 16-byte
// access per lane (a wavefront covers 1 KiB contiguous) and the packed words
// are read or written once.  Needs N % 16 == 0 and 16-byte aligned spins.
template <typename WT>
__global__ void pack_parity16_kernel(const int8_t *confs, WT *packed, int R, int P, int N)
{
    const int G = N / 16;
    const size_t tot = (size_t)R * G;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < tot;
         x += (size_t)gridDim.x * blockDim.x) {
        const int q = (int)(x % G), r = (int)(x / G);
        const int8_t *src = confs + (size_t)r * P * N + 16 * (size_t)q;
        WT w0[16], w1[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) w0[j] = w1[j] = 0;
        // four slice pairs per trip: eight 16-byte loads in flight per lane
        // (two left the pack at 3.5 TB/s on 4096 x 64 x 1024 spins)
#pragma unroll 4
        for (int k = 0; k < P; k += 2) {
            const uint4 a = *reinterpret_cast<const uint4 *>(src + (size_t)k * N);
            const uint4 b = *reinterpret_cast<const uint4 *>(src + (size_t)(k + 1) * N);
            const uint32_t av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
            const int m = k >> 1;
#pragma unroll
            for (int j = 0; j < 16; ++j) {  // a spin is -1 iff its byte's sign bit is set
                w0[j] |= (WT)((av[j >> 2] >> (8 * (j & 3) + 7)) & 1u) << m;
                w1[j] |= (WT)((bv[j >> 2] >> (8 * (j & 3) + 7)) & 1u) << m;
            }
        }
        WT *d0 = packed + (size_t)r * 2 * N + 16 * (size_t)q, *d1 = d0 + N;
        // 16-byte stores: the words of 16 spins are 64 or 128 aligned bytes
        constexpr int kV = 16 / (int)sizeof(WT);
#pragma unroll
        for (int j = 0; j < 16; j += kV) {
            uint4 v0, v1;
            __builtin_memcpy(&v0, &w0[j], 16);
            __builtin_memcpy(&v1, &w1[j], 16);
            *reinterpret_cast<uint4 *>(d0 + j) = v0;
            *reinterpret_cast<uint4 *>(d1 + j) = v1;
        }
    }
}

template <typename WT>
__global__ void unpack_parity16_kernel(const WT *packed, int8_t *confs, int R, int P, int N)
{
    const int G = N / 16;
    const size_t tot = (size_t)R * G;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < tot;
         x += (size_t)gridDim.x * blockDim.x) {
        const int q = (int)(x % G), r = (int)(x / G);
        const WT *s0 = packed + (size_t)r * 2 * N + 16 * (size_t)q, *s1 = s0 + N;
        WT w0[16], w1[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            w0[j] = s0[j];
            w1[j] = s1[j];
        }
        int8_t *dst = confs + (size_t)r * P * N + 16 * (size_t)q;
        for (int k = 0; k < P; k += 2) {
            const int m = k >> 1;
            uint32_t av[4] = {0u, 0u, 0u, 0u}, bv[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int j = 0; j < 16; ++j) {  // bit set: -1 (0xFF), clear: +1 (0x01)
                av[j >> 2] |= (1u | (uint32_t)((w0[j] >> m) & WT(1)) * 0xFEu) << (8 * (j & 3));
                bv[j >> 2] |= (1u | (uint32_t)((w1[j] >> m) & WT(1)) * 0xFEu) << (8 * (j & 3));
            }
            *reinterpret_cast<uint4 *>(dst + (size_t)k * N) = make_uint4(av[0], av[1], av[2], av[3]);
            *reinterpret_cast<uint4 *>(dst + (size_t)(k + 1) * N) = make_uint4(bv[0], bv[1], bv[2], bv[3]);
        }
    }
}

int grid_for(size_t n, int block)
{
    size_t g = (n + block - 1) / block;
    return (int)(g > 65535 ? 65535 : (g < 1 ? 1 : g));
}

// int8 spins [R][P][N] <-> parity-packed words [R][2][N] on stream st.
template <typename WT>
void launch_pack(const int8_t *confs, WT *packed, int R, int P, int N, hipStream_t st)
{
    if (N % 16 == 0 && ((uintptr_t)confs & 15) == 0 && ((uintptr_t)packed & 15) == 0)
        pack_parity16_kernel<WT><<<grid_for((size_t)R * (N / 16), 256), 256, 0, st>>>(confs, packed, R, P, N);
    else
        pack_parity_kernel<WT><<<grid_for((size_t)R * 2 * N, 256), 256, 0, st>>>(confs, packed, R, P, N);
}

template <typename WT>
void launch_unpack(const WT *packed, int8_t *confs, int R, int P, int N, hipStream_t st)
{
    if (N % 16 == 0 && ((uintptr_t)confs & 15) == 0)
        unpack_parity16_kernel<WT><<<grid_for((size_t)R * (N / 16), 256), 256, 0, st>>>(packed, confs, R, P, N);
    else
        unpack_parity_kernel<WT><<<grid_for((size_t)R * P * N, 256), 256, 0, st>>>(packed, confs, R, P, N);
}

// Per-field-step tables for the +-1 path.  The decision class of a bit is a
// function of (z, u_s, u_t) and, at fixed (z, u_t), monotone in the
// unsatisfied-bond count u_s (ediff grows by 4 per unit), so each (z, u_t) is
// summarised by two thresholds: accept for u_s >= amin, draw for
// smin <= u_s < amin, reject below.  ztab packs amin/smin of u_t = 0,1,2 in
// nibbles (amin, smin) at bits 0,4 / 8,12 / 16,20; thr holds the draw
// thresholds.  Returns false if some class is not monotone (then the per-bit
// path is used).
bool pmj_tables(const piqmc_problem *pb, int S, const std::vector<float> &jp,
                const std::vector<float> &temps, std::vector<uint32_t> &thr, std::vector<uint32_t> &ztab)
{
    thr.assign((size_t)S * kClassStride, 0);
    ztab.assign((size_t)S * 8, 0);
    // a uniform-degree lattice reads only its own degree's entries
    const int zlo = pb->zuni >= 0 ? pb->zuni : 0, zhi = pb->zuni >= 0 ? pb->zuni : pb->zmax;
    for (int s = 0; s < S; ++s)
        for (int z = zlo; z <= zhi; ++z)
            for (int ut = 0; ut < 3; ++ut) {
                int cls[9];
                for (int us = 0; us <= z; ++us) {
                    // -2 s_i (h + Jp*ts) with s_i*h = z - 2u_s, s_i*ts = 2 - 2u_t (exact)
                    const double ed = -2.0 * ((double)(z - 2 * us) + (double)jp[s] * (double)(2 - 2 * ut));
                    unsigned int t = 0;
                    cls[us] = piqmc_fixed_class(ed, (double)temps[s], &t);
                    if (cls[us] == 2) thr[(size_t)s * kClassStride + (z * 8 + us) * 3 + ut] = t;
                }
                int amin = z + 1, smin = z + 1;
                for (int us = z; us >= 0 && cls[us] == 1; --us) amin = us;
                for (int us = amin - 1; us >= 0 && cls[us] == 2; --us) smin = us;
                if (smin > amin) smin = amin;
                for (int us = 0; us < smin; ++us)
                    if (cls[us] != 0) return false;
                ztab[(size_t)s * 8 + z] |= ((uint32_t)amin | ((uint32_t)smin << 4)) << (8 * ut);
            }
    return true;
}

bool use_dense(const piqmc_problem *pb)
{
    return pb->sweep_path == 2 || (pb->sweep_path == 0 && pb->N < 8 * pb->ncolors);
}

// The device buffers of one parallel-sweep call.  Either they are carved,
// 256-byte aligned, out of a region the caller owns (base != null: nothing is
// allocated), or each is allocated on the call's stream and freed there when
// the call returns, on its error paths included.
struct CallBuffers {
    unsigned char *base = nullptr;
    size_t cap = 0, used = 0;
    hipStream_t st = nullptr;
    void *owned[8] = {};
    int nowned = 0;

    static size_t round_up(size_t b) { return (b + 255) & ~(size_t)255; }
    int take(void **p, size_t bytes)
    {
        *p = nullptr;
        if (base) {
            if (used + round_up(bytes) > cap) return piqmc::set_error(PIQMC_EINVAL, "parallel sweep: scratch region too small");
            *p = base + used;
            used += round_up(bytes);
            return PIQMC_OK;
        }
        if (nowned == 8) return piqmc::set_error(PIQMC_EINVAL, "parallel sweep: too many buffers");
        PIQMC_HIP(hipMallocAsync(p, bytes, st));
        owned[nowned++] = *p;
        return PIQMC_OK;
    }
    ~CallBuffers()
    {
        for (int k = nowned - 1; k >= 0; --k) (void)hipFreeAsync(owned[k], st);
    }
};

// Device bytes a parallel-sweep call takes from a CallBuffers region (each
// buffer rounded up to 256 bytes): J_perp and temperatures [2S] floats, the
// +-1 path's per-step tables, the color offsets and the packed state
// [R][2][N] words.
size_t colored_scratch_bytes(const piqmc_problem *pb, int R, int P, int S)
{
    const size_t wb = P / 2 <= 32 ? 4 : 8;
    size_t b = CallBuffers::round_up(sizeof(float) * (S ? 2 * S : 1));
    if (use_dense(pb) && pb->Z <= kFieldsMaxRow &&
        (pb->dense_variant == 6 || (pb->dense_variant == 0 && pb->exact_rows && pb->fields_ok)))
        b += CallBuffers::round_up(sizeof(double) * 2 * (size_t)(P / 2) * (pb->N | 1) * R);
    if (!use_dense(pb)) {
        if (pb->pmj && S)
            b += CallBuffers::round_up(sizeof(uint32_t) * (size_t)S * kClassStride) +
                 CallBuffers::round_up(sizeof(uint32_t) * (size_t)S * 8);
        b += CallBuffers::round_up(sizeof(int) * (pb->ncolors + 1));
    }
    return b + CallBuffers::round_up(wb * 2 * (size_t)pb->N * R);
}

template <typename WT>
int run_colored(const piqmc_problem *pb, int R, int P, int S, int M,
                const std::vector<float> &jp, const std::vector<float> &temps, int8_t *confs, uint64_t seed, uint32_t rbase,
                uint32_t sbase, double *etrace, int block, int rpw, hipStream_t st,
                unsigned char *scratch = nullptr, size_t scratch_cap = 0)
{
    const int N = pb->N;
    CallBuffers buf;
    buf.base = scratch;
    buf.cap = scratch_cap;
    buf.st = st;
    int rc;
    // dense path when color classes average fewer than 8 spins (or forced)
    if (use_dense(pb)) {
        // one replica per wavefront; small replica counts get one wavefront per
        // workgroup so that they spread over the CUs
        if (block <= 0) block = R >= 4 * 256 ? 256 : (R >= 2 * 256 ? 128 : 64);
        if (block > 1024 || (block & 63)) return piqmc::set_error(PIQMC_EINVAL, "block must be a multiple of 64, <= 1024");
        const int wpb = block / 64;
        const size_t lds = dense_wave_bytes(N, pb->Z, sizeof(WT)) * wpb;
        const int W = P / 2;
        const bool split_ok = pb->exact_rows && W <= 16 && 16 % W == 0 && block <= kDenseMaxBlock;
        bool wg = false, pipe = false, split = false;
        int pw = 4;
        size_t lwg = 0;
        bool fields_auto = false;
        if (pb->dense_variant == 0) {
            // cached fields while replicas are few (256-spin SK, P=128: 33 ms
            // against 46 ms pipelined at 4 replicas, 20 against 18 at 64)
            fields_auto = pb->exact_rows && pb->fields_ok && pb->Z <= kFieldsMaxRow && R <= 32 &&
                          dense_wg_enabled() && dense_fields_bytes(N, W, sizeof(WT)) <= 160 * 1024;
            // a workgroup per replica while replicas are few (1.2-1.9x up to 1024
            // 256-spin replicas); from 4096 on, a wavefront each is 1.3x faster
            const bool wg0 = !fields_auto && pb->exact_rows && dense_wg_enabled() && R < 2048;
            pw = dense_pipe_waves(pb->Z, W, R);
            pipe = wg0 && dense_pipe_enabled() && dense_pipe_bytes(N, pb->Z, sizeof(WT), pw - 1) <= 160 * 1024;
            lwg = pipe ? dense_pipe_bytes(N, pb->Z, sizeof(WT), pw - 1) : dense_wg_bytes(N, pb->Z, sizeof(WT));
            wg = wg0 && lwg <= 160 * 1024;
            split = !wg && split_ok;
        } else if (pb->dense_variant == 2) {
            if (!split_ok)
                return piqmc::set_error(PIQMC_EINVAL, "dense variant 2 (slice split over lanes) needs exactly summable "
                                                      "rows, P/2 dividing 16 and block <= 256");
            split = true;
        } else if (pb->dense_variant == 6) {
            if (!pb->exact_rows || !pb->fields_ok || pb->Z > kFieldsMaxRow ||
                dense_fields_bytes(N, W, sizeof(WT)) > 160 * 1024)
                return piqmc::set_error(PIQMC_EINVAL, "dense variant 6 (cached fields) needs exactly summable rows "
                                                      "of at most 256 entries that list each neighbour once with the "
                                                      "same coupling both ways, and W x N fields that fit the LDS");
        } else if (pb->dense_variant >= 3) {
            pipe = pb->dense_variant >= 4;
            pw = pb->dense_variant == 5 ? 8 : 4;
            lwg = pipe ? dense_pipe_bytes(N, pb->Z, sizeof(WT), pw - 1) : dense_wg_bytes(N, pb->Z, sizeof(WT));
            if (!pb->exact_rows || lwg > 160 * 1024)
                return piqmc::set_error(PIQMC_EINVAL, "dense variants 3-5 (a workgroup per replica) need exactly "
                                                      "summable rows and a row that fits the LDS");
            wg = true;
        }
        const bool fields = pb->dense_variant == 6 || fields_auto;
        if (!wg && !fields && lds > 160 * 1024) return piqmc::set_error(PIQMC_EINVAL, "dense sweep: state exceeds LDS");
        float *d_jp = nullptr;  // [S] J_perp, then [S] temperatures
        WT *d_packed = nullptr;
        double *d_fields = nullptr;
        if ((rc = buf.take((void **)&d_jp, sizeof(float) * (S ? 2 * S : 1)))) return rc;
        if (S) PIQMC_HIP(hipMemcpyAsync(d_jp, jp.data(), sizeof(float) * S, hipMemcpyHostToDevice, st));
        if (S) PIQMC_HIP(hipMemcpyAsync(d_jp + S, temps.data(), sizeof(float) * S, hipMemcpyHostToDevice, st));
        if ((rc = buf.take((void **)&d_packed, sizeof(WT) * 2 * (size_t)N * R))) return rc;
        if (fields && (rc = buf.take((void **)&d_fields, sizeof(double) * 2 * (size_t)W * fields_pitch(N) * R)))
            return rc;
        launch_pack<WT>(confs, d_packed, R, P, N, st);
        ColArgs<WT> a = {};
        a.N = N; a.Z = pb->Z; a.P = P; a.W = P / 2; a.S = S; a.M = M; a.R = R;
        a.ncolors = pb->ncolors; a.temps = d_jp + S;
        a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
        a.replica_base = rbase; a.sweep_base = sbase;
        a.jperp = d_jp; a.rows = pb->d_rows; a.order = pb->d_order; a.prevj = pb->d_prevj;
        a.packed = d_packed; a.etrace = etrace; a.fields = d_fields;
        piqmc::timing_begin(st);
        if (fields) {
            // one replica per wavefront, as many per workgroup as block and LDS allow
            const size_t per = dense_fields_bytes(N, W, sizeof(WT));
            const int wf = std::max(1, std::min(std::min(wpb, 4), (int)((160 * 1024) / per)));
            fields_init_kernel<WT><<<grid_for((size_t)R * 2 * W * N, 256), 256, 0, st>>>(a);
            const int ns = fields_split_waves(W, wf);
            if (ns > 1) {
                // one replica's tables fill the LDS: its slices over ns wavefronts
                const void *kf = fields_split_kernel_for<WT>(W / ns);
                (void)hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)per);
                void *kargs[] = {(void *)&a};
                PIQMC_HIP(hipLaunchKernel(kf, dim3(R), dim3(64 * ns), kargs, per, st));
            } else {
                (void)hipFuncSetAttribute((const void *)dense_fields_kernel<WT>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)(per * wf));
                dense_fields_kernel<WT><<<(R + wf - 1) / wf, 64 * wf, per * wf, st>>>(a);
            }
        } else if (wg && pipe) {
            // one replica per workgroup, one spin decided while the next is summed
            (void)hipFuncSetAttribute((const void *)dense_pipe_kernel<WT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lwg);
            dense_pipe_kernel<WT><<<R, 64 * pw, lwg, st>>>(a);
        } else if (wg) {
            // one replica per workgroup (the field split over four wavefronts)
            (void)hipFuncSetAttribute((const void *)dense_wg_kernel<WT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lwg);
            dense_wg_kernel<WT><<<R, 256, lwg, st>>>(a);
        } else if (split) {
            // split when it gives at least four lanes per slice (P = 16: 2.2x,
            // P = 32: 1.4x on 256-spin SK; two lanes per slice measured 4 % slower)
            (void)hipFuncSetAttribute((const void *)dense_split_kernel<WT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            dense_split_kernel<WT><<<(R + wpb - 1) / wpb, block, lds, st>>>(a);
        } else {
            (void)hipFuncSetAttribute((const void *)dense_kernel<WT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            dense_kernel<WT><<<(R + wpb - 1) / wpb, block, lds, st>>>(a);
        }
        piqmc::timing_end(st);
        PIQMC_HIP(hipGetLastError());
        launch_unpack<WT>(d_packed, confs, R, P, N, st);
        PIQMC_HIP(hipGetLastError());
        return PIQMC_OK;
    }
    std::vector<uint32_t> thr, ztab;
    const bool pmj = pb->pmj && pmj_tables(pb, S, jp, temps, thr, ztab);
    if (rpw <= 0 && pb->pmjf && !pb->pmj && pb->ncolors == 2 && pb->zuni_nf == 6 && cached_fields_enabled()) {
        // the cached field sweep evaluates each step's decision classes once per
        // workgroup for all of its replicas: as many replicas per workgroup as
        // put every workgroup on the device at once (one 512-thread workgroup
        // per CU), within the LDS (3-D L=12 field workload, 1024 replicas: 4
        // per workgroup 251 ms, 2 298 ms, 3 400 ms, 1 392 ms)
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 256;
        const size_t per = sizeof(WT) * 2 * (size_t)N, fixed_part = 4 * (kClassStride + 8) + 16 * sizeof(double);
        const int lds_max = (int)std::max<size_t>(1, (160 * 1024 - fixed_part) / per);
        rpw = std::max(1, std::min(lds_max, (R + ncu - 1) / ncu));
    }
    if (rpw <= 0 && block <= 0 && fixed_tile_applies(pb, R, sizeof(WT))) {
        // the fixed-plane kernel's tile (1024-spin workload: 9.25 ms against
        // 9.6 ms for the best general tile)
        rpw = kFixRep;
        block = (int)kFixBlock;
    }
    if (rpw <= 0) rpw = R >= 2048 ? 4 : (R >= 1024 ? 2 : 1);
    if (rpw > 64) return piqmc::set_error(PIQMC_EINVAL, "replicas per workgroup must be <= 64");
    const size_t lds = sizeof(WT) * 2 * (size_t)N * rpw + 4 * (kClassStride + 8) + 16 * sizeof(double);
    if (lds > 160 * 1024)
        return piqmc::set_error(PIQMC_EINVAL, "colored sweep: replicas per workgroup x 2 x nspins x P/2 bits exceed one workgroup's LDS");
    if (block <= 0) {
        // 256 threads; 512 when a color class would take more than two passes
        // of 256 (measured on the 1728-spin 3-D lattice: 1.08e12 vs 0.84e12
        // attempts/s)
        int biggest = 0;
        for (int c = 0; c < pb->ncolors; ++c) biggest = std::max(biggest, pb->color_off[c + 1] - pb->color_off[c]);
        block = biggest > 512 ? 512 : 256;
    }
    if (block > 1024 || (block & 63)) return piqmc::set_error(PIQMC_EINVAL, "block must be a multiple of 64, <= 1024");
    float *d_jp = nullptr;  // [S] J_perp, then [S] temperatures
    uint32_t *d_thr = nullptr, *d_zt = nullptr;
    int *d_coff = nullptr;
    WT *d_packed = nullptr;
    if ((rc = buf.take((void **)&d_jp, sizeof(float) * (S ? 2 * S : 1)))) return rc;
    if (S) PIQMC_HIP(hipMemcpyAsync(d_jp, jp.data(), sizeof(float) * S, hipMemcpyHostToDevice, st));
    if (S) PIQMC_HIP(hipMemcpyAsync(d_jp + S, temps.data(), sizeof(float) * S, hipMemcpyHostToDevice, st));
    if (pb->pmj && S) {
        // taken whenever the problem is +-1 (as colored_scratch_bytes counts),
        // filled only when the tables apply
        if ((rc = buf.take((void **)&d_thr, sizeof(uint32_t) * (size_t)S * kClassStride))) return rc;
        if ((rc = buf.take((void **)&d_zt, sizeof(uint32_t) * (size_t)S * 8))) return rc;
        if (pmj) {
            PIQMC_HIP(hipMemcpyAsync(d_thr, thr.data(), sizeof(uint32_t) * thr.size(), hipMemcpyHostToDevice, st));
            PIQMC_HIP(hipMemcpyAsync(d_zt, ztab.data(), sizeof(uint32_t) * ztab.size(), hipMemcpyHostToDevice, st));
        }
    }
    if ((rc = buf.take((void **)&d_coff, sizeof(int) * (pb->ncolors + 1)))) return rc;
    PIQMC_HIP(hipMemcpyAsync(d_coff, pb->color_off.data(), sizeof(int) * (pb->ncolors + 1), hipMemcpyHostToDevice, st));
    if ((rc = buf.take((void **)&d_packed, sizeof(WT) * 2 * (size_t)N * R))) return rc;
    launch_pack<WT>(confs, d_packed, R, P, N, st);

    ColArgs<WT> a;
    a.N = N; a.Z = pb->Z; a.P = P; a.W = P / 2; a.S = S; a.M = M;
    a.ncolors = pb->ncolors; a.zmax = pb->zmax; a.temps = d_jp + S;
    a.R = R; a.rpw = rpw; a.zuni = pmj ? pb->zuni : -1;
    a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
    a.replica_base = rbase; a.sweep_base = sbase;
    a.jperp = d_jp; a.rows = pb->d_rows; a.order = pb->d_order; a.color_off = d_coff;
    a.cnb = pb->d_cnb; a.cinfo = pb->d_cinfo; a.thr = pmj ? d_thr : nullptr; a.ztab = pmj ? d_zt : nullptr; a.hq = pb->d_hq; a.prevj = pb->d_prevj;
    a.packed = d_packed; a.etrace = etrace;
    const int grid = (R + rpw - 1) / rpw;
    piqmc::timing_begin(st);
    if (!pmj && pb->pmjf) {
        a.zuni = -1;
        int biggest = 0;
        for (int c = 0; c < pb->ncolors; ++c) biggest = std::max(biggest, pb->color_off[c + 1] - pb->color_off[c]);
        // two colors of at most two spins per thread and six couplings per
        // spin besides the field (cubic lattices): spins, neighbours and
        // decision classes cached in registers
        if (pb->ncolors == 2 && biggest <= 2 * block && block <= 512 && pb->zuni_nf == 6 && cached_fields_enabled()) {
            (void)hipFuncSetAttribute((const void *)colored_fields_cached_kernel<WT>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            colored_fields_cached_kernel<WT><<<grid, block, lds, st>>>(a);
        } else {
            (void)hipFuncSetAttribute((const void *)colored_kernel<WT, true, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            colored_kernel<WT, true, false, true><<<grid, block, lds, st>>>(a);
        }
    } else if (pmj) {
        if (a.zuni == 4) {
            // every step's accept threshold for one antiparallel Trotter bond
            // is count >= 2 (the fixed-plane kernel takes it as a constant)
            bool ka1_two = true;
            for (int s = 0; s < S; ++s) ka1_two = ka1_two && ((ztab[(size_t)s * 8 + 4] >> 8) & 15u) == 2u;
            launch_uniform<WT, 4>(a, pb, grid, block, lds, st, ka1_two);
        } else if (a.zuni == 6) {
            launch_uniform<WT, 6>(a, pb, grid, block, lds, st);
        } else if (a.zuni >= 0) {
            const size_t lq = queue_lds(N, rpw, sizeof(WT), block);
            bool queued = false;
            if constexpr (sizeof(WT) == 4) {
                if (lq <= 160 * 1024 && queue_enabled()) {
                    (void)hipFuncSetAttribute((const void *)colored_q_kernel<WT, 0, false>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lq);
                    colored_q_kernel<WT, 0, false><<<grid, block, lq, st>>>(a);
                    queued = true;
                }
            }
            if (!queued) {
                (void)hipFuncSetAttribute((const void *)colored_kernel<WT, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                colored_kernel<WT, true, true><<<grid, block, lds, st>>>(a);
            }
        } else {
            (void)hipFuncSetAttribute((const void *)colored_kernel<WT, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            colored_kernel<WT, true, false><<<grid, block, lds, st>>>(a);
        }
    } else {
        (void)hipFuncSetAttribute((const void *)colored_kernel<WT, false, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        colored_kernel<WT, false, false><<<grid, block, lds, st>>>(a);
    }
    piqmc::timing_end(st);
    PIQMC_HIP(hipGetLastError());
    launch_unpack<WT>(d_packed, confs, R, P, N, st);
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

}  // namespace

namespace {
int qa_colored_impl(const piqmc_problem *pb, int R, int P, int S, int M, float temp, const double *sched,
                    int8_t *confs, uint64_t seed, uint32_t rbase, uint32_t sbase, double *etrace, int block,
                    int rpw, hipStream_t st, bool use_scratch);
}

extern "C" int piqmc_qa_colored(const piqmc_problem *pb, int R, int P, int S, int M, float temp,
                                const double *sched, int8_t *confs, uint64_t seed, uint32_t rbase,
                                uint32_t sbase, double *etrace, int block, int rpw, void *stream)
{
    return qa_colored_impl(pb, R, P, S, M, temp, sched, confs, seed, rbase, sbase, etrace, block, rpw,
                           (hipStream_t)stream, false);
}

extern "C" int piqmc_qa_colored_noalloc(const piqmc_problem *pb, int R, int P, int S, int M, float temp,
                                        const double *sched, int8_t *confs, uint64_t seed, uint32_t rbase,
                                        uint32_t sbase, double *etrace, int block, int rpw, void *stream)
{
    return qa_colored_impl(pb, R, P, S, M, temp, sched, confs, seed, rbase, sbase, etrace, block, rpw,
                           (hipStream_t)stream, true);
}

namespace {

// Address of this device's copy of `scratch`, looked up once per device with
// that device current.  The host-side name of the array is not a device
// address; only these pointers are passed to kernels and copies.
int scratch_region(unsigned char **out)
{
    int dev = -1;
    PIQMC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchMaxDevices)
        return piqmc::set_error(PIQMC_EINVAL, "qa_colored_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return piqmc::hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return piqmc::set_error(PIQMC_EHIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return PIQMC_OK;
}

int qa_colored_impl(const piqmc_problem *pb, int R, int P, int S, int M, float temp, const double *sched,
                    int8_t *confs, uint64_t seed, uint32_t rbase, uint32_t sbase, double *etrace, int block,
                    int rpw, hipStream_t st, bool use_scratch)
{
    if (!pb || R < 1 || S < 0 || M < 0 || (S > 0 && !sched) || !confs)
        return piqmc::set_error(PIQMC_EINVAL, "qa_colored: bad arguments");
    if (P < 2 || (P & 1) || P > 128)
        return piqmc::set_error(PIQMC_EINVAL, "parallel sweep needs an even number of slices, 2 <= P <= 128");
    std::vector<float> jp(S > 0 ? S : 0);
    for (int s = 0; s < S; ++s) {
        if ((float)P * temp == 0.0f) return piqmc::set_error(PIQMC_EZERODIV, "float division");
        jp[s] = host_jperp(sched[s], P, temp);
    }
    if ((long)S * M > 0 && temp == 0.0f) return piqmc::set_error(PIQMC_EZERODIV, "float division");
    unsigned char *region = nullptr;
    size_t cap = 0;
    if (use_scratch) {
        const size_t need = colored_scratch_bytes(pb, R, P, S);
        if (need > kScratchBytes)
            return piqmc::set_error(PIQMC_EINVAL, "qa_colored_noalloc: this call needs " + std::to_string(need) +
                                                      " bytes of device scratch; the maximum is " +
                                                      std::to_string(kScratchBytes) +
                                                      " (use piqmc_qa_colored, or fewer replicas per call)");
        int rc = scratch_region(&region);
        if (rc) return rc;
        cap = kScratchBytes;
    }
    const std::vector<float> temps(jp.size(), temp);
    if (P / 2 <= 32)
        return run_colored<uint32_t>(pb, R, P, S, M, jp, temps, confs, seed, rbase, sbase, etrace, block, rpw, st,
                                     region, cap);
    return run_colored<uint64_t>(pb, R, P, S, M, jp, temps, confs, seed, rbase, sbase, etrace, block, rpw, st,
                                 region, cap);
}

}  // namespace

// Classical annealing of C independent chains (C even) with the parallel
// sweep: chains are laid out as the slices of C/P' replicas (P' the largest
// even divisor of C up to 64) whose Trotter coupling is exactly zero, so each
// chain sees only its spatial couplings, at temperature sched[s] in step s.
extern "C" int piqmc_sa_colored(const piqmc_problem *pb, int C, int S, int M, const double *sched,
                                int8_t *spins, uint64_t seed, uint32_t rbase, uint32_t sbase,
                                int block, int rpw, void *stream)
{
    if (!pb || C < 2 || (C & 1) || S < 0 || M < 0 || (S > 0 && !sched) || !spins)
        return piqmc::set_error(PIQMC_EINVAL, "sa_colored: bad arguments (the chain count must be even)");
    int P = 64;
    while (C % P) P -= 2;
    std::vector<float> jp(S > 0 ? S : 0, 0.0f), temps(S > 0 ? S : 0);
    for (int s = 0; s < S; ++s) {
        temps[s] = (float)sched[s];
        if (!(temps[s] > 0.0f) || !std::isfinite(temps[s]))
            return piqmc::set_error(PIQMC_EINVAL, "sa_colored: temperatures must be finite and > 0");
    }
    hipStream_t st = (hipStream_t)stream;
    return run_colored<uint32_t>(pb, C / P, P, S, M, jp, temps, spins, seed, rbase, sbase, nullptr, block,
                                 rpw, st);
}

#include "batch_part.inc"
