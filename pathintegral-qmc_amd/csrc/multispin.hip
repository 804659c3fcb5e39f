// sa.Anneal_multispin (piqmc/sa.pyx:282-405) on gfx950: 64 classical
// configurations packed one bit each into a uint64 per spin (configuration k
// at bit 63-k, bit 1 = spin -1, piqmc/sa.pyx:339-345).
//
// One wavefront owns one packed set and lane k owns configuration k, so the
// reference's 64-iteration bit loops become one lane-parallel step: the
// spin index, the neighbour row and the packed words are wave-uniform (LDS
// broadcast reads), each lane accumulates its own double ediff, compares
// exp(ediff/temp) with its own uniform, and a ballot, bit-reversed to the
// 63-k order, is the flip mask written back to the spin's word.
//
// Arithmetic follows piqmc/sa.pyx:355-393: ediff is double, reset per
// attempt, and moves by 2*(double)(float)J per neighbour entry in row order;
// temp is the float temperature of the step; exp() is double.  The uniforms
// are supplied by the caller (the reference draws them from numpy's
// RandomState, rng.rand(64) per attempt, piqmc/sa.pyx:331,398).
#include <math.h>

#include "piqmc_internal.h"

namespace {

struct MsArgs {
    int R, N, Z, G;
    const float *temps;     // [G] temperature of each sweep
    const int2 *rows;       // [N][Z] (j, float bits of J)
    const uint16_t *perms;  // [G][R][N]
    const double *U;        // [G][N][R][64]
    unsigned long long *bits;  // [R][N] packed configurations, in/out
    uint32_t *ties;            // [R] accumulated near ties
};

constexpr int kMaxSetsPerBlock = 4;

__global__ void __launch_bounds__(64 * kMaxSetsPerBlock) multispin_kernel(MsArgs a)
{
    extern __shared__ unsigned long long sv[];  // [sets per block][N]
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = blockIdx.x * (int)(blockDim.x >> 6) + w;
    const int N = a.N, Z = a.Z, R = a.R;
    unsigned long long *s = sv + (size_t)w * N;
    if (r < R)
        for (int q = lane; q < N; q += 64) s[q] = a.bits[(size_t)r * N + q];
    __syncthreads();
    if (r >= R) return;  // whole waves leave; no barrier follows
    uint32_t nt = 0;
    for (int g = 0; g < a.G; ++g) {
        const double T = (double)a.temps[g];
        const uint16_t *perm = a.perms + ((size_t)g * R + r) * N;
        for (int t = 0; t < N; ++t) {
            const int i = __builtin_amdgcn_readfirstlane((int)perm[t]);
            const unsigned long long wi = s[i];
            const int2 *row = a.rows + (size_t)i * Z;
            double e = 0.0;
            for (int z = 0; z < Z; ++z) {
                const int j = __builtin_amdgcn_readfirstlane(row[z].x);
                const float J = __int_as_float(__builtin_amdgcn_readfirstlane(row[z].y));
                const unsigned long long x = j == i ? wi : (wi ^ s[j]);
                const double J2 = 2.0 * (double)J;
                e = ((x >> (63 - lane)) & 1ull) ? e + J2 : e - J2;
            }
            const double u = a.U[(((size_t)g * N + t) * R + r) * 64 + lane];
            const double ex = exp(e / T);
            const bool acc = ex > u;
            if (ex > 0.0 && ex < INFINITY && fabs(ex - u) <= ex * 0x1p-50) ++nt;
            const unsigned long long fm = __builtin_bitreverse64(__ballot(acc));
            s[i] = wi ^ fm;  // every lane stores the same word
        }
    }
    for (int q = lane; q < N; q += 64) a.bits[(size_t)r * N + q] = s[q];
    for (int off = 32; off > 0; off >>= 1) nt += __shfl_xor(nt, off, 64);
    if (lane == 0) a.ties[r] += nt;
}

}  // namespace

extern "C" int piqmc_sa_multispin(const piqmc_problem *pb, int R, int G, const float *temps_host,
                                  uint64_t *bits_dev, const uint16_t *perms_dev,
                                  const double *U_dev, uint32_t *ties_dev, void *stream)
{
    if (!pb || R < 1 || G < 0 || (G > 0 && (!temps_host || !perms_dev || !U_dev)) || !bits_dev || !ties_dev)
        return piqmc::set_error(PIQMC_EINVAL, "sa_multispin: bad arguments");
    const int N = pb->N;
    if (N > 65536) return piqmc::set_error(PIQMC_EINVAL, "sa_multispin supports nspins <= 65536");
    int spb = kMaxSetsPerBlock;  // sets (wavefronts) per workgroup, as many as the LDS holds
    while (spb > 1 && sizeof(unsigned long long) * (size_t)N * spb > 160 * 1024) --spb;
    const size_t lds = sizeof(unsigned long long) * (size_t)N * spb;
    if (lds > 160 * 1024)
        return piqmc::set_error(PIQMC_EINVAL, "sa_multispin: nspins too large for one workgroup's LDS");
    if (G == 0) return PIQMC_OK;
    hipStream_t st = (hipStream_t)stream;
    const size_t nperm = (size_t)G * N * R;
    if (N < 65536) {  // permutation entries index LDS: reject out-of-range ones first
        const int rc = piqmc::check_perms(perms_dev, nperm, N, st);
        if (rc) return rc;
    }
    float *d_t = nullptr;
    PIQMC_HIP(hipMallocAsync((void **)&d_t, sizeof(float) * G, st));
    PIQMC_HIP(hipMemcpyAsync(d_t, temps_host, sizeof(float) * G, hipMemcpyHostToDevice, st));
    MsArgs a;
    a.R = R; a.N = N; a.Z = pb->Z; a.G = G;
    a.temps = d_t;
    a.rows = pb->d_rows;
    a.perms = perms_dev;
    a.U = U_dev;
    a.bits = reinterpret_cast<unsigned long long *>(bits_dev);
    a.ties = ties_dev;
    (void)hipFuncSetAttribute((const void *)multispin_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    piqmc::timing_begin(st);
    multispin_kernel<<<(R + spb - 1) / spb, 64 * spb, lds, st>>>(a);
    piqmc::timing_end(st);
    PIQMC_HIP(hipGetLastError());
    PIQMC_HIP(hipFreeAsync(d_t, st));
    return PIQMC_OK;
}
