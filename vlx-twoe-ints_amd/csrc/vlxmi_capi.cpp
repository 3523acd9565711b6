// Hand-written part of the vlxmi C ABI: version, Boys tables, status strings,
// and a Boys-function probe kernel used by the parity tests.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include "vlxmi_device.hpp"

extern "C" int vlxmi_version(void) { return 1; }

extern "C" int vlxmi_boys_table_size(uint32_t* nrow, uint32_t* ncol, uint32_t* nft)
{
    if (!nrow || !ncol || !nft) return VLXMI_ERR_ARG;
    *nrow = VLX_BOYS_NROW + VLX_BOYS_CROWS;
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
    for (int k = 0; k < VLX_BOYS_CROWS; ++k) {
        const long double T = (long double)k * (long double)VLX_BOYS_CDT;
        boys_series(VLX_BOYS_NCOL - 1, T, F);
        for (int m = 0; m < VLX_BOYS_NCOL; ++m) table[VLX_BOYS_COARSE_OFF + k * VLX_BOYS_NCOL + m] = (double)F[m];
    }
    for (int m = 0; m < 32; ++m) ft[m] = 1.0 / (2.0 * m + 1.0);
    return VLXMI_OK;
}

extern "C" const char* vlxmi_status_string(int code)
{
    switch (code) {
        case VLXMI_OK: return "ok";
        case VLXMI_ERR_ARG: return "invalid argument (null pointer or bad count)";
        case VLXMI_ERR_LAYOUT: return "list is not in the primitive-pair block order the kernels need";
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

template <int L>
__global__ void boys_coarse_probe_kernel(const double* __restrict__ T, uint32_t n,
                                         const double* __restrict__ table, double* __restrict__ out)
{
    __shared__ double lt[VLX_BOYS_CROWS * VLX_BOYS_CTERMS];
    vlxmi::boys_lds_fill<L>(table, lt);
    __syncthreads();
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double F[L + 1];
    vlxmi::boys_lds<L>(T[t], lt, F);
#pragma unroll
    for (int m = 0; m <= L; ++m) out[(size_t)t * (L + 1) + m] = F[m];
}

// As vlxmi_boys_eval, through the coarse rows copied to LDS (the evaluation
// some light-class kernels use); L = 0..4.
extern "C" int vlxmi_boys_eval_coarse(const double* T, uint32_t n, uint32_t L, const double* table, double* out,
                                      void* stream)
{
    if (n == 0) return VLXMI_OK;
    if (!T || !table || !out || L > 4) return VLXMI_ERR_ARG;
    dim3 g((n + 63) / 64), b(64);
    hipStream_t s = (hipStream_t)stream;
    switch (L) {
        case 0: hipLaunchKernelGGL(boys_coarse_probe_kernel<0>, g, b, 0, s, T, n, table, out); break;
        case 1: hipLaunchKernelGGL(boys_coarse_probe_kernel<1>, g, b, 0, s, T, n, table, out); break;
        case 2: hipLaunchKernelGGL(boys_coarse_probe_kernel<2>, g, b, 0, s, T, n, table, out); break;
        case 3: hipLaunchKernelGGL(boys_coarse_probe_kernel<3>, g, b, 0, s, T, n, table, out); break;
        default: hipLaunchKernelGGL(boys_coarse_probe_kernel<4>, g, b, 0, s, T, n, table, out); break;
    }
    return (int)hipGetLastError();
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

// Number of primitive shell quartets of one K class that pass the block screen
// QKmax_IJ * QKmax_KL * maxD_bd > thr (the N_q of the algorithmic FLOP model).
// qa/qc: partner-block maxima in CSR order (non-increasing within a primitive),
// *_displ/*_cnt: first block and block count per primitive.  out is added to.
__global__ void count_k_kernel(const uint32_t* __restrict__ ikI, const uint32_t* __restrict__ ikK, uint32_t nblk,
                               const double* __restrict__ qa, const uint32_t* __restrict__ qa_displ,
                               const uint32_t* __restrict__ qa_cnt, const double* __restrict__ qc,
                               const uint32_t* __restrict__ qc_displ, const uint32_t* __restrict__ qc_cnt,
                               double lim, unsigned long long* __restrict__ out)
{
    const uint32_t b = blockIdx.x;
    if (b >= nblk) return;
    const uint32_t I = ikI[b], K = ikK[b];
    const double* qaI = qa + qa_displ[I];
    const uint32_t nJ = qa_cnt[I];
    const double* qcK = qc + qc_displ[K];
    const uint32_t nL = qc_cnt[K];
    unsigned long long c = 0;
    for (uint32_t j = threadIdx.x; j < nJ; j += blockDim.x) {
        const double q = qaI[j];
        uint32_t lo = 0, hi = nL;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (q * qcK[mid] > lim) lo = mid + 1; else hi = mid;
        }
        c += lo;
    }
    if (c) atomicAdd(out, c);
}

extern "C" int vlxmi_count_k_quartets(const uint32_t* ikI, const uint32_t* ikK, uint32_t nblk,
                                      const double* qa, const uint32_t* qa_displ, const uint32_t* qa_cnt,
                                      const double* qc, const uint32_t* qc_displ, const uint32_t* qc_cnt,
                                      double thr_over_maxD, unsigned long long* out, void* stream)
{
    if (nblk == 0) return VLXMI_OK;
    if (!ikI || !ikK || !qa || !qa_displ || !qa_cnt || !qc || !qc_displ || !qc_cnt || !out) return VLXMI_ERR_ARG;
    hipLaunchKernelGGL(count_k_kernel, dim3(nblk), dim3(64), 0, (hipStream_t)stream, ikI, ikK, nblk, qa, qa_displ,
                       qa_cnt, qc, qc_displ, qc_cnt, thr_over_maxD, out);
    return (int)hipGetLastError();
}

// Deterministic AO assembly: out[seg_out[s]] = sum over t in [seg_off[s], seg_off[s+1])
// of values[perm[t]], summed in list order (no atomics, one writer per output).
__global__ void segment_sum_kernel(const double* __restrict__ values, const uint32_t* __restrict__ perm,
                                   const uint32_t* __restrict__ seg_off, const uint32_t* __restrict__ seg_out,
                                   uint32_t nseg, double* __restrict__ out)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    double acc = 0.0;
    const uint32_t e = seg_off[s + 1];
    for (uint32_t t = seg_off[s]; t < e; ++t) acc += values[perm[t]];
    out[seg_out[s]] = acc;
}

extern "C" int vlxmi_segment_sum(const double* values, const uint32_t* perm, const uint32_t* seg_off,
                                 const uint32_t* seg_out, uint32_t nseg, double* out, void* stream)
{
    if (nseg == 0) return VLXMI_OK;
    if (!values || !perm || !seg_off || !seg_out || !out) return VLXMI_ERR_ARG;
    hipLaunchKernelGGL(segment_sum_kernel, dim3((nseg + 255) / 256), dim3(256), 0, (hipStream_t)stream, values,
                       perm, seg_off, seg_out, nseg, out);
    return (int)hipGetLastError();
}

// Fock-matrix assembly: out[seg_out[s]] = sum over t in [seg_off[s], seg_off[s+1])
// of coef[src[t]] * values[perm[t]], in list order (one writer per output, no
// atomics).  With the Coulomb and exchange per-pair outputs concatenated in
// ``values`` and src = 0 / 1 selecting c_J / c_K, one pass gives the
// two-electron part c_J J + c_K K of the Fock matrix in the Cartesian AO basis.
__global__ void fock_assemble_kernel(const double* __restrict__ values, const uint32_t* __restrict__ perm,
                                     const uint8_t* __restrict__ src, const uint32_t* __restrict__ seg_off,
                                     const uint32_t* __restrict__ seg_out, uint32_t nseg, double cJ, double cK,
                                     double* __restrict__ out)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    double acc = 0.0;
    const uint32_t e = seg_off[s + 1];
    for (uint32_t t = seg_off[s]; t < e; ++t) acc += (src[t] ? cK : cJ) * values[perm[t]];
    out[seg_out[s]] = acc;
}

extern "C" int vlxmi_fock_assemble(const double* values, const uint32_t* perm, const uint8_t* src,
                                   const uint32_t* seg_off, const uint32_t* seg_out, uint32_t nseg, double cJ,
                                   double cK, double* out, void* stream)
{
    if (nseg == 0) return VLXMI_OK;
    if (!values || !perm || !src || !seg_off || !seg_out || !out) return VLXMI_ERR_ARG;
    hipLaunchKernelGGL(fock_assemble_kernel, dim3((nseg + 255) / 256), dim3(256), 0, (hipStream_t)stream, values,
                       perm, src, seg_off, seg_out, nseg, cJ, cK, out);
    return (int)hipGetLastError();
}

// p[0..n) = 0 with a kernel (graph-capture safe: records no memset node).
__global__ void fill_zero_kernel(double* __restrict__ p, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = i; k < n; k += stride) p[k] = 0.0;
}

extern "C" int vlxmi_fill_zero(double* p, uint64_t n, void* stream)
{
    if (n == 0) return VLXMI_OK;
    if (!p) return VLXMI_ERR_ARG;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, p, n);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ket-split partial sums of the Coulomb launchers.
//
// A Coulomb launch with ket_split = n > 1 runs n chunks of the ket list in
// parallel (grid.y).  Chunk y writes its per-entry partial sums to
// part[y * count + e]; ket_split_reduce_kernel then adds, for every entry e,
// the n partials in chunk order with Neumaier's compensated summation, so the
// result does not depend on the order the chunks ran in (no atomics: the build
// is bit-reproducible) and carries about one rounding of the total instead of
// one per chunk.
// ---------------------------------------------------------------------------

namespace {
struct Workspace {
    double* p = nullptr;
    size_t n = 0;
};
std::mutex g_ws_mutex;
std::map<std::pair<int, void*>, Workspace> g_ws;   // (device, stream) -> buffer
}  // namespace

// A device buffer of at least ndoubles doubles reserved for launches on
// (current device, stream): launches on one stream run in order, so they can
// share it; another stream gets its own.  Grown on demand (after synchronizing
// the stream, whose earlier launches may still read the old buffer) and kept
// until vlxmi_release_workspaces().
extern "C" int vlxmi_ket_workspace(size_t ndoubles, void* stream, double** out)
{
    if (!out) return VLXMI_ERR_ARG;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    Workspace& w = g_ws[std::make_pair(dev, stream)];
    if (w.n < ndoubles) {
        if (w.p) {
            e = hipStreamSynchronize((hipStream_t)stream);
            if (e != hipSuccess) return (int)e;
            (void)hipFree(w.p);
            w.p = nullptr;
            w.n = 0;
        }
        const size_t n = ndoubles + ndoubles / 4;
        e = hipMalloc((void**)&w.p, n * sizeof(double));
        if (e != hipSuccess) {
            w.p = nullptr;
            return (int)e;
        }
        w.n = n;
    }
    *out = w.p;
    return VLXMI_OK;
}

// Frees every ket-split workspace (after a device synchronization).
extern "C" int vlxmi_release_workspaces(void)
{
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if (!g_ws.empty()) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) return (int)e;
    }
    for (auto& kv : g_ws)
        if (kv.second.p) (void)hipFree(kv.second.p);
    g_ws.clear();
    return VLXMI_OK;
}

__device__ __forceinline__ void neumaier_add(double& s, double& c, double x)
{
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
    s = t;
}

// 256 threads = 64 entries x 4 interleaved chunk subsets; the four partial
// (sum, compensation) pairs of an entry are combined in a fixed order.
__global__ void __launch_bounds__(256)
ket_split_reduce_kernel(const double* __restrict__ part, uint32_t nsplit, uint32_t count,
                        double* __restrict__ out)
{
    __shared__ double ls[4][64], lc[4][64];
    const uint32_t el = threadIdx.x & 63, q = threadIdx.x >> 6;
    const uint32_t e = blockIdx.x * 64 + el;
    double s = 0.0, c = 0.0;
    if (e < count)
        for (uint32_t y = q; y < nsplit; y += 4) neumaier_add(s, c, part[(size_t)y * count + e]);
    ls[q][el] = s;
    lc[q][el] = c;
    __syncthreads();
    if (q == 0 && e < count) {
        double S = ls[0][el], Cc = lc[0][el];
        for (int k = 1; k < 4; ++k) {
            neumaier_add(S, Cc, ls[k][el]);
            Cc += lc[k][el];
        }
        out[e] += S + Cc;
    }
}

// out[e] += sum_{y < nsplit} part[y * count + e], e < count (see above).
extern "C" int vlxmi_ket_split_reduce(const double* part, uint32_t nsplit, uint32_t count, double* out, void* stream)
{
    if (count == 0 || nsplit == 0) return VLXMI_OK;
    if (!part || !out) return VLXMI_ERR_ARG;
    hipLaunchKernelGGL(ket_split_reduce_kernel, dim3((count + 63) / 64), dim3(256), 0, (hipStream_t)stream, part,
                       nsplit, count, out);
    return (int)hipGetLastError();
}
