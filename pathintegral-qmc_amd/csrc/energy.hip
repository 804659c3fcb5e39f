// Classical Ising energies on the device.
//  - piqmc_slice_energies: every Trotter slice / replica configuration from the
//    neighbour table (the min-over-slices and residual-energy metric of
//    examples/spinglass32.py:130-137).
//  - piqmc_classical_energy: sa.ClassicalIsingEnergy (piqmc/sa.pyx:25-44),
//    E = -s.(J_off s) - sum(diag(J) s), over a COO coupling list.
// One workgroup per configuration; per-thread partial sums are reduced with
// wavefront shuffles and one LDS slot per wave.
#include <mutex>

#include "piqmc_internal.h"

namespace {

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ double block_sum_d(double v)
{
    __shared__ double red[16];
    v = wave_sum_d(v);
    const int wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < nw; ++w) t += red[w];
    return t;
}

__global__ void slice_energy_kernel(const int2 *rows, int N, int Z, const int8_t *spins, double *E)
{
    const int8_t *s = spins + (size_t)blockIdx.x * N;
    double acc = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const double si = (double)s[i];
        for (int e = 0; e < Z; ++e) {
            const int2 rw = rows[(size_t)i * Z + e];
            const double J = (double)__int_as_float(rw.y);
            if (rw.x == i) acc -= J * si;
            else acc -= 0.5 * J * si * (double)s[rw.x];
        }
    }
    const double t = block_sum_d(acc);
    if (threadIdx.x == 0) E[blockIdx.x] = t;
}

__global__ void coo_energy_kernel(const int64_t *ri, const int64_t *ci, const double *v, int64_t nnz,
                                  int N, const double *spins, double *E)
{
    const double *s = spins + (size_t)blockIdx.x * N;
    double off = 0.0, diag = 0.0;
    for (int64_t t = threadIdx.x; t < nnz; t += blockDim.x) {
        const int64_t i = ri[t], j = ci[t];
        if (i == j) diag += v[t] * s[i];
        else off += s[i] * v[t] * s[j];
    }
    const double o = block_sum_d(off);
    __syncthreads();
    const double d = block_sum_d(diag);
    if (threadIdx.x == 0) E[blockIdx.x] = -o - d;
}

}  // namespace

extern "C" int piqmc_slice_energies(const piqmc_problem *pb, int B, const int8_t *spins, double *E,
                                    void *stream)
{
    if (!pb || B < 0 || (B && (!spins || !E))) return piqmc::set_error(PIQMC_EINVAL, "slice_energies: bad arguments");
    if (B == 0) return PIQMC_OK;
    slice_energy_kernel<<<B, 256, 0, (hipStream_t)stream>>>(pb->d_rows, pb->N, pb->Z, spins, E);
    PIQMC_HIP(hipGetLastError());
    return PIQMC_OK;
}

namespace {
// Device workspace of piqmc_classical_energy, kept between calls (one per
// device, grown when a call needs more) so that repeated energy evaluations
// do not allocate; calls are serialised by its mutex.
struct EnergyWorkspace {
    std::mutex mu;
    void *buf = nullptr;
    size_t cap = 0;
    hipStream_t st = nullptr;
};
constexpr int kEnergyMaxDevices = 64;
EnergyWorkspace g_energy_ws[kEnergyMaxDevices];
}  // namespace

extern "C" int piqmc_classical_energy(int N, int64_t nnz, const int64_t *rows, const int64_t *cols,
                                      const double *vals, int B, const double *spins, double *E)
{
    if (N < 0 || nnz < 0 || B < 0 || (nnz && (!rows || !cols || !vals)) || (B && (!spins || !E)))
        return piqmc::set_error(PIQMC_EINVAL, "classical_energy: bad arguments");
    for (int64_t t = 0; t < nnz; ++t)
        if (rows[t] < 0 || rows[t] >= N || cols[t] < 0 || cols[t] >= N)
            return piqmc::set_error(PIQMC_EINVAL, "coupling index out of range");
    if (B == 0) return PIQMC_OK;
    int dev = -1;
    PIQMC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kEnergyMaxDevices) return piqmc::set_error(PIQMC_EINVAL, "classical_energy: device ordinal");
    EnergyWorkspace &ws = g_energy_ws[dev];
    std::lock_guard<std::mutex> lock(ws.mu);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t nz = nnz ? (size_t)nnz : 1;
    const size_t o_c = up(8 * nz), o_v = o_c + up(8 * nz), o_s = o_v + up(8 * nz);
    const size_t o_e = o_s + up(8 * (size_t)B * (N ? N : 1)), need = o_e + up(8 * (size_t)B);
    if (!ws.st) PIQMC_HIP(hipStreamCreateWithFlags(&ws.st, hipStreamNonBlocking));
    if (need > ws.cap) {
        if (ws.buf) PIQMC_HIP(hipFree(ws.buf));
        ws.buf = nullptr;
        ws.cap = 0;
        PIQMC_HIP(hipMalloc(&ws.buf, need));
        ws.cap = need;
    }
    char *base = static_cast<char *>(ws.buf);
    int64_t *d_r = reinterpret_cast<int64_t *>(base), *d_c = reinterpret_cast<int64_t *>(base + o_c);
    double *d_v = reinterpret_cast<double *>(base + o_v), *d_s = reinterpret_cast<double *>(base + o_s);
    double *d_e = reinterpret_cast<double *>(base + o_e);
    if (nnz) {
        PIQMC_HIP(hipMemcpyAsync(d_r, rows, 8 * nz, hipMemcpyHostToDevice, ws.st));
        PIQMC_HIP(hipMemcpyAsync(d_c, cols, 8 * nz, hipMemcpyHostToDevice, ws.st));
        PIQMC_HIP(hipMemcpyAsync(d_v, vals, 8 * nz, hipMemcpyHostToDevice, ws.st));
    }
    if (N) PIQMC_HIP(hipMemcpyAsync(d_s, spins, 8 * (size_t)B * N, hipMemcpyHostToDevice, ws.st));
    coo_energy_kernel<<<B, 256, 0, ws.st>>>(d_r, d_c, d_v, nnz, N, d_s, d_e);
    PIQMC_HIP(hipGetLastError());
    PIQMC_HIP(hipMemcpyAsync(E, d_e, 8 * (size_t)B, hipMemcpyDeviceToHost, ws.st));
    PIQMC_HIP(hipStreamSynchronize(ws.st));
    return PIQMC_OK;
}
