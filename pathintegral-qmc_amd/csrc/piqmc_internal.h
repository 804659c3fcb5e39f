// Internal declarations shared by the HIP translation units of libpiqmc_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/piqmc_amd.h"
#include "piqmc_fexp64.h"

// ------------------------------------------------------------------ errors
namespace piqmc {
int set_error(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
// PIQMC_EINVAL (with a message) if any of the n uint16 permutation entries on
// the device is >= N; synchronizes the stream.
int check_perms(const uint16_t *perms, size_t n, int N, hipStream_t st);
}  // namespace piqmc

#define PIQMC_HIP(call)                                                   \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return piqmc::hip_fail(e_, #call);          \
    } while (0)

// ------------------------------------------------------------------ problem
// A neighbour table (tools.GenerateNeighbors format) resident in HBM, with the
// analysis the colored sweep needs.
struct piqmc_problem {
    int N = 0, Z = 0;
    std::vector<int32_t> idx;  // [N][Z]  (int)nbs[i][z][0]
    std::vector<float> J;      // [N][Z]  (float)nbs[i][z][1]
    int2 *d_rows = nullptr;    // [N][Z]  (j, float bits of J), replay + generic path
    // Problems built from a dense coupling matrix replay the *_dense functions,
    // which read J in double precision: d_jd holds those values, [N][Z].
    int dense = 0;
    double *d_jd = nullptr;

    // Graph coloring of the nonzero off-diagonal couplings.
    int ncolors = 0;
    std::vector<int> color;       // [N]
    std::vector<int> color_off;   // [ncolors+1] offsets into order
    int32_t *d_order = nullptr;   // [N] spins sorted by color

    // +-1 coupling path: every nonzero entry is J=+-1 to another spin.
    int pmj = 0;
    int zmax = 0;                 // max nonzero off-diagonal entries per row
    int zuni = -1;                // common degree when every row has zmax entries, else -1
    int zuni_nf = -1;             // common number of non-field couplings per row, else -1
    int sweep_path = 0;           // parallel sweep: 0 auto, 1 word per thread, 2 dense (lane per slice)
    int dense_variant = 0;        // dense sweep layout: 0 auto, 1..5 (piqmc_problem_set_dense_variant)
    int32_t *d_cnb = nullptr;     // [zmax][N] neighbour of position q (color order)
    uint32_t *d_cinfo = nullptr;  // [N] bits 0..7 negative-coupling mask, 8..11 z
    // +-1 couplings with fields: as above, plus at most one field per spin with
    // 2^-27 <= |h| <= 2^20 (so the local field's partial sums are exact)
    int pmjf = 0;
    float *d_hq = nullptr;        // [N] field of spin order[q]
    // Every row's local field sums exactly in double, in any order (all of
    // its couplings are multiples of 2^L and all partial sums stay below 2^E
    // with E - L <= 52): the lane-per-slice sweep may then split and
    // reassociate the sum without changing a bit.
    int exact_rows = 0;
    // Every row lists each neighbour at most once with a nonzero coupling, and
    // that coupling appears, equal, in the neighbour's row: a flip of spin i
    // then changes exactly the fields its own row names (cached-field sweep).
    int fields_ok = 0;
    // [N] sum of the couplings of spin order[q] to spin order[q-1] (0 at q = 0):
    // the pipelined dense sweep adds that neighbour's term after its decision
    double *d_prevj = nullptr;
};

// ------------------------------------------------------------------ device helpers
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32x32->64 product each (v_mad_u64_u32) rather than separate
        // low and high multiplies: measured 5 % faster on the +-1 sweep
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        // three-input XORs as one gfx950 v_bitop3_b32 each (truth table 0x96)
        uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
        uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Host J_perp, piqmc/qmc.pyx:95 (qmc.c:2055-2075).
float host_jperp(double gamma, int P, float temp);

// HIP events bracketing the main sweep kernel of the last call on this host
// thread (read back by piqmc_last_kernel_ms).
namespace piqmc {
void timing_begin(hipStream_t st);
void timing_end(hipStream_t st);
}  // namespace piqmc
