// Hand-written part of the vlxmi C ABI: version, Boys tables, status strings,
// and a Boys-function probe kernel used by the parity tests.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

#include "vlxmi_device.hpp"

extern "C" int vlxmi_version(void) { return 1; }

extern "C" int vlxmi_boys_table_size(uint32_t* nrow, uint32_t* ncol, uint32_t* nft)
{
    if (!nrow || !ncol || !nft) return VLXMI_ERR_ARG;
    *nrow = VLX_BOYS_NROW;
    *ncol = VLX_BOYS_NCOL;
    *nft = 32;
    return VLXMI_OK;
}

// F_M(T) by its convergent series exp(-T) sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1)),
// then F_m for m < M by downward recursion; long double throughout.
static void boys_series(int M, long double T, long double* F)
{
    long double term = 1.0L / (2.0L * M + 1.0L), sum = term;
    for (int k = 1; k < 4000; ++k) {
        term *= 2.0L * T / (2.0L * M + 2.0L * k + 1.0L);
        sum += term;
        if (term < sum * 1e-22L) break;
    }
    const long double et = expl(-T);
    F[M] = et * sum;
    for (int m = M - 1; m >= 0; --m) F[m] = (2.0L * T * F[m + 1] + et) / (2.0L * m + 1.0L);
}

// Fills the host buffers of the table layout documented in vlxmi_device.hpp.
extern "C" int vlxmi_boys_table_fill(double* table, double* ft)
{
    if (!table || !ft) return VLXMI_ERR_ARG;
    long double F[VLX_BOYS_NCOL];
    for (int k = 0; k < VLX_BOYS_NROW; ++k) {
        const long double T = (long double)k * (long double)VLX_BOYS_DT;
        boys_series(VLX_BOYS_NCOL - 2, T, F);
        for (int m = 0; m < VLX_BOYS_NCOL - 1; ++m) table[k * VLX_BOYS_NCOL + m] = (double)F[m];
        table[k * VLX_BOYS_NCOL + VLX_BOYS_NCOL - 1] = (double)expl(-T);
    }
    for (int m = 0; m < 32; ++m) ft[m] = 1.0 / (2.0 * m + 1.0);
    return VLXMI_OK;
}

extern "C" const char* vlxmi_status_string(int code)
{
    switch (code) {
        case VLXMI_OK: return "ok";
        case VLXMI_ERR_ARG: return "invalid argument (null pointer or bad count)";
        case VLXMI_ERR_LAYOUT: return "list length is not a whole number of primitive-pair blocks";
        case VLXMI_ERR_SIZE: return "size exceeds the entry point's limit";
        default: return hipGetErrorString((hipError_t)code);
    }
}

template <int L>
__global__ void boys_probe_kernel(const double* __restrict__ T, uint32_t n, const double* __restrict__ table,
                                  double* __restrict__ out)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double F[L + 1];
    vlxmi::boys<L>(T[t], table, F);
#pragma unroll
    for (int m = 0; m <= L; ++m) out[(size_t)t * (L + 1) + m] = F[m];
}

// out[t*(L+1)+m] = F_m(T[t]) for the device Boys function of order L (0..8).
extern "C" int vlxmi_boys_eval(const double* T, uint32_t n, uint32_t L, const double* table, double* out,
                               void* stream)
{
    if (n == 0) return VLXMI_OK;
    if (!T || !table || !out || L > 8) return VLXMI_ERR_ARG;
    dim3 g((n + 63) / 64), b(64);
    hipStream_t s = (hipStream_t)stream;
    switch (L) {
        case 0: hipLaunchKernelGGL(boys_probe_kernel<0>, g, b, 0, s, T, n, table, out); break;
        case 1: hipLaunchKernelGGL(boys_probe_kernel<1>, g, b, 0, s, T, n, table, out); break;
        case 2: hipLaunchKernelGGL(boys_probe_kernel<2>, g, b, 0, s, T, n, table, out); break;
        case 3: hipLaunchKernelGGL(boys_probe_kernel<3>, g, b, 0, s, T, n, table, out); break;
        case 4: hipLaunchKernelGGL(boys_probe_kernel<4>, g, b, 0, s, T, n, table, out); break;
        case 5: hipLaunchKernelGGL(boys_probe_kernel<5>, g, b, 0, s, T, n, table, out); break;
        case 6: hipLaunchKernelGGL(boys_probe_kernel<6>, g, b, 0, s, T, n, table, out); break;
        case 7: hipLaunchKernelGGL(boys_probe_kernel<7>, g, b, 0, s, T, n, table, out); break;
        default: hipLaunchKernelGGL(boys_probe_kernel<8>, g, b, 0, s, T, n, table, out); break;
    }
    return (int)hipGetLastError();
}
