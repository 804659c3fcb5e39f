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

#include "piqmc_internal.h"

namespace {

constexpr int kClassStride = 8 * 8 * 3;  // (z, u_s, u_t) thresholds per field step

template <typename WT>
struct ColArgs {
    int N, Z, P, W, S, M, ncolors, zmax;
    float temp;
    uint32_t key0, key1, replica_base, sweep_base;
    const float *jperp;        // [S]
    const int2 *rows;          // [N][Z]
    const int32_t *order;      // [N] spins by color
    const int *color_off;      // [ncolors+1]
    const int32_t *cnb;        // [zmax][N] (+-1 path)
    const uint32_t *cinfo;     // [N]
    const uint32_t *thr;       // [S][kClassStride]
    const uint32_t *accm;      // [S][8]
    const uint32_t *stom;      // [S][8]
    WT *packed;                // [R][2][N]
    double *etrace;            // [R][S] or null
};

template <typename WT>
__device__ __forceinline__ WT rotl1(WT x, int W, WT valid)
{
    return ((x << 1) | (x >> (W - 1))) & valid;
}
template <typename WT>
__device__ __forceinline__ WT rotr1(WT x, int W)
{
    return (x >> 1) | ((x & WT(1)) << (W - 1));
}
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

template <typename WT, bool PMJ>
__global__ void colored_kernel(ColArgs<WT> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, W = a.W, tid = threadIdx.x, bd = blockDim.x;
    WT *st = reinterpret_cast<WT *>(smem);                       // [2][N]
    uint32_t *tthr = reinterpret_cast<uint32_t *>(st + 2 * N);    // [kClassStride]
    uint32_t *tacc = tthr + kClassStride;                         // [8]
    uint32_t *tsto = tacc + 8;                                    // [8]
    double *red = reinterpret_cast<double *>(tsto + 8);           // [16]
    const WT valid = (W == (int)(8 * sizeof(WT))) ? ~WT(0) : ((WT(1) << W) - 1);
    const int b = blockIdx.x;
    const uint32_t rep = a.replica_base + (uint32_t)b;
    WT *gst = a.packed + (size_t)b * 2 * N;
    for (int q = tid; q < 2 * N; q += bd) st[q] = gst[q];

    for (int s = 0; s < a.S; ++s) {
        __syncthreads();
        if (PMJ) {
            for (int q = tid; q < kClassStride; q += bd) tthr[q] = a.thr[(size_t)s * kClassStride + q];
            if (tid < 8) {
                tacc[tid] = a.accm[s * 8 + tid];
                tsto[tid] = a.stom[s * 8 + tid];
            }
        }
        __syncthreads();
        const float jp = a.jperp[s];
        for (int m = 0; m < a.M; ++m) {
            const uint32_t g = a.sweep_base + (uint32_t)(s * a.M + m);
            for (int tau = 0; tau < 2; ++tau) {
                WT *mine = st + tau * N;
                const WT *other = st + (1 - tau) * N;
                for (int c = 0; c < a.ncolors; ++c) {
                    const int q0 = a.color_off[c], q1 = a.color_off[c + 1];
                    for (int q = q0 + tid; q < q1; q += bd) {
                        const int i = a.order[q];
                        const WT S_ = mine[i], O = other[i];
                        const WT Lw = tau == 0 ? rotl1(O, W, valid) : O;
                        const WT Rw = tau == 0 ? O : rotr1(O, W);
                        const WT uL = S_ ^ Lw, uR = S_ ^ Rw;  // 1 = antiparallel Trotter bond
                        WT flips = 0;
                        if (PMJ) {
                            const uint32_t info = a.cinfo[q];
                            const int zi = (info >> 8) & 15;
                            WT c0 = 0, c1 = 0, c2 = 0;  // bit-sliced unsatisfied-bond counter
                            for (int z = 0; z < a.zmax; ++z) {
                                if (z < zi) {
                                    const int j = a.cnb[(size_t)z * N + q];
                                    const WT x = S_ ^ mine[j] ^ (((info >> z) & 1u) ? ~WT(0) : WT(0));
                                    const WT k0 = c0 & x;
                                    c0 ^= x;
                                    const WT k1 = c1 & k0;
                                    c1 ^= k0;
                                    c2 ^= k1;
                                }
                            }
                            const uint32_t am = tacc[zi], sm = tsto[zi];
                            const WT et[3] = {~(uL | uR), uL ^ uR, uL & uR};
                            WT X = 0;
                            for (int us = 0; us <= a.zmax; ++us) {
                                const WT es = ((us & 4) ? c2 : ~c2) & ((us & 2) ? c1 : ~c1) & ((us & 1) ? c0 : ~c0);
#pragma unroll
                                for (int ut = 0; ut < 3; ++ut) {
                                    const int bp = us * 3 + ut;
                                    if ((am >> bp) & 1u) flips |= es & et[ut];
                                    else if ((sm >> bp) & 1u) X |= es & et[ut];
                                }
                            }
                            flips &= valid;
                            X &= valid;
                            while (X) {
                                const int qd = ctz(X) >> 2;
                                uint32_t out[4];
                                philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qd, a.key0,
                                              a.key1, out);
                                const WT qm = (X >> (4 * qd)) & WT(15);
#pragma unroll
                                for (int bb = 0; bb < 4; ++bb) {
                                    if ((qm >> bb) & WT(1)) {
                                        const int mm = 4 * qd + bb;
                                        const int us = (int)(((c2 >> mm) & 1) << 2 | ((c1 >> mm) & 1) << 1 | ((c0 >> mm) & 1));
                                        const int ut = (int)((uL >> mm) & 1) + (int)((uR >> mm) & 1);
                                        if (out[bb] < tthr[(zi * 8 + us) * 3 + ut]) flips |= WT(1) << mm;
                                    }
                                }
                                X &= ~(WT(15) << (4 * qd));
                            }
                        } else {
                            const int2 *row = a.rows + (size_t)i * a.Z;
                            int qcur = -1;
                            uint32_t out[4] = {0, 0, 0, 0};
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
                                const int cls = piqmc_fixed_class(ed, (double)a.temp, &thr);
                                bool fl = cls == 1;
                                if (cls == 2) {
                                    if ((mm >> 2) != qcur) {
                                        qcur = mm >> 2;
                                        philox4x32_10(g, (uint32_t)i, rep, (uint32_t(tau) << 24) | (uint32_t)qcur,
                                                      a.key0, a.key1, out);
                                    }
                                    fl = out[mm & 3] < thr;
                                }
                                if (fl) flips |= WT(1) << mm;
                            }
                        }
                        mine[i] = S_ ^ flips;
                    }
                    __syncthreads();
                }
            }
        }
        if (a.etrace) {
            // Mean over slices of the classical energy; sums of +-1 products are
            // popcounts of XORed parity words.
            double acc = 0.0;
            for (int i = tid; i < N; i += bd) {
                const WT a0 = st[i], a1 = st[N + i];
                const int2 *row = a.rows + (size_t)i * a.Z;
                for (int e = 0; e < a.Z; ++e) {
                    const int2 rw = row[e];
                    const double J = (double)__int_as_float(rw.y);
                    if (J == 0.0) continue;
                    if (rw.x == i) {
                        acc -= J * (double)(a.P - 2 * (popc(a0) + popc(a1)));
                    } else {
                        const int pc = popc(a0 ^ st[rw.x]) + popc(a1 ^ st[N + rw.x]);
                        acc -= 0.5 * J * (double)(a.P - 2 * pc);
                    }
                }
            }
            const double tot = block_sum(acc, red);
            if (tid == 0) a.etrace[(size_t)b * a.S + s] = tot / (double)a.P;
        }
    }
    __syncthreads();
    for (int q = tid; q < 2 * N; q += bd) gst[q] = st[q];
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

int grid_for(size_t n, int block)
{
    size_t g = (n + block - 1) / block;
    return (int)(g > 65535 ? 65535 : (g < 1 ? 1 : g));
}

template <typename WT>
int run_colored(const piqmc_problem *pb, int R, int P, int S, int M, float temp,
                const std::vector<float> &jp, int8_t *confs, uint64_t seed, uint32_t rbase,
                uint32_t sbase, double *etrace, int block, hipStream_t st)
{
    const int N = pb->N;
    const bool pmj = pb->pmj != 0;
    const size_t lds = sizeof(WT) * 2 * (size_t)N + 4 * (kClassStride + 16) + 16 * sizeof(double);
    if (lds > 160 * 1024)
        return piqmc::set_error(PIQMC_EINVAL, "colored sweep: 2*nspins*P/2 bits exceed one workgroup's LDS");
    // Per-field-step tables for the +-1 path.
    std::vector<uint32_t> thr, accm, stom;
    if (pmj) {
        thr.assign((size_t)S * kClassStride, 0);
        accm.assign((size_t)S * 8, 0);
        stom.assign((size_t)S * 8, 0);
        for (int s = 0; s < S; ++s)
            for (int z = 0; z <= pb->zmax; ++z)
                for (int us = 0; us <= z; ++us)
                    for (int ut = 0; ut < 3; ++ut) {
                        // -2 s_i (h + Jp*ts) with s_i*h = z - 2u_s, s_i*ts = 2 - 2u_t (exact)
                        const double ed = -2.0 * ((double)(z - 2 * us) + (double)jp[s] * (double)(2 - 2 * ut));
                        unsigned int t = 0;
                        const int cls = piqmc_fixed_class(ed, (double)temp, &t);
                        const int bp = us * 3 + ut;
                        if (cls == 1) accm[(size_t)s * 8 + z] |= 1u << bp;
                        if (cls == 2) {
                            stom[(size_t)s * 8 + z] |= 1u << bp;
                            thr[(size_t)s * kClassStride + (z * 8 + us) * 3 + ut] = t;
                        }
                    }
    }
    float *d_jp = nullptr;
    uint32_t *d_thr = nullptr, *d_acc = nullptr, *d_sto = nullptr;
    int *d_coff = nullptr;
    WT *d_packed = nullptr;
    PIQMC_HIP(hipMallocAsync((void **)&d_jp, sizeof(float) * (S ? S : 1), st));
    if (S) PIQMC_HIP(hipMemcpyAsync(d_jp, jp.data(), sizeof(float) * S, hipMemcpyHostToDevice, st));
    if (pmj && S) {
        PIQMC_HIP(hipMallocAsync((void **)&d_thr, sizeof(uint32_t) * thr.size(), st));
        PIQMC_HIP(hipMallocAsync((void **)&d_acc, sizeof(uint32_t) * accm.size(), st));
        PIQMC_HIP(hipMallocAsync((void **)&d_sto, sizeof(uint32_t) * stom.size(), st));
        PIQMC_HIP(hipMemcpyAsync(d_thr, thr.data(), sizeof(uint32_t) * thr.size(), hipMemcpyHostToDevice, st));
        PIQMC_HIP(hipMemcpyAsync(d_acc, accm.data(), sizeof(uint32_t) * accm.size(), hipMemcpyHostToDevice, st));
        PIQMC_HIP(hipMemcpyAsync(d_sto, stom.data(), sizeof(uint32_t) * stom.size(), hipMemcpyHostToDevice, st));
    }
    PIQMC_HIP(hipMallocAsync((void **)&d_coff, sizeof(int) * (pb->ncolors + 1), st));
    PIQMC_HIP(hipMemcpyAsync(d_coff, pb->color_off.data(), sizeof(int) * (pb->ncolors + 1), hipMemcpyHostToDevice, st));
    PIQMC_HIP(hipMallocAsync((void **)&d_packed, sizeof(WT) * 2 * (size_t)N * R, st));
    pack_parity_kernel<WT><<<grid_for((size_t)R * 2 * N, 256), 256, 0, st>>>(confs, d_packed, R, P, N);

    ColArgs<WT> a;
    a.N = N; a.Z = pb->Z; a.P = P; a.W = P / 2; a.S = S; a.M = M;
    a.ncolors = pb->ncolors; a.zmax = pb->zmax; a.temp = temp;
    a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
    a.replica_base = rbase; a.sweep_base = sbase;
    a.jperp = d_jp; a.rows = pb->d_rows; a.order = pb->d_order; a.color_off = d_coff;
    a.cnb = pb->d_cnb; a.cinfo = pb->d_cinfo; a.thr = d_thr; a.accm = d_acc; a.stom = d_sto;
    a.packed = d_packed; a.etrace = etrace;
    if (block <= 0) block = 256;
    if (block > 1024 || (block & 63)) return piqmc::set_error(PIQMC_EINVAL, "block must be a multiple of 64, <= 1024");
    piqmc::timing_begin(st);
    if (pmj) {
        (void)hipFuncSetAttribute((const void *)colored_kernel<WT, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        colored_kernel<WT, true><<<R, block, lds, st>>>(a);
    } else {
        (void)hipFuncSetAttribute((const void *)colored_kernel<WT, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        colored_kernel<WT, false><<<R, block, lds, st>>>(a);
    }
    piqmc::timing_end(st);
    PIQMC_HIP(hipGetLastError());
    unpack_parity_kernel<WT><<<grid_for((size_t)R * P * N, 256), 256, 0, st>>>(d_packed, confs, R, P, N);
    PIQMC_HIP(hipFreeAsync(d_packed, st));
    PIQMC_HIP(hipFreeAsync(d_coff, st));
    if (d_thr) PIQMC_HIP(hipFreeAsync(d_thr, st));
    if (d_acc) PIQMC_HIP(hipFreeAsync(d_acc, st));
    if (d_sto) PIQMC_HIP(hipFreeAsync(d_sto, st));
    PIQMC_HIP(hipFreeAsync(d_jp, st));
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

}  // namespace

extern "C" int piqmc_qa_colored(const piqmc_problem *pb, int R, int P, int S, int M, float temp,
                                const double *sched, int8_t *confs, uint64_t seed, uint32_t rbase,
                                uint32_t sbase, double *etrace, int block, void *stream)
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
    hipStream_t st = (hipStream_t)stream;
    if (P / 2 <= 32)
        return run_colored<uint32_t>(pb, R, P, S, M, temp, jp, confs, seed, rbase, sbase, etrace, block, st);
    return run_colored<uint64_t>(pb, R, P, S, M, temp, jp, confs, seed, rbase, sbase, etrace, block, st);
}
