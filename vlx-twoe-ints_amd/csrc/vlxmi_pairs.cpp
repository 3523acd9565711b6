// Device side of the pair-list builder (the step before the J/K kernels:
// VeloxChem's host builds these lists for the reference kernels, whose
// arguments they are, gen_coulomb.py:74-93, gen_exchange.py:75-115,161-168).
//
// vlxmi_pair_candidates: every candidate primitive-pair block (I, J) of a pair
//   class with its overlap factors S_ij = f_i f_j (pi/zeta)^(3/2)
//   exp(-a_i a_j |AB|^2 / zeta) per component pair (a-major);
// vlxmi_block_absmax: max |x| over each block of nab values (Schwarz bound of
//   a block);
// vlxmi_kcsr_fill: the entries of an exchange CSR from sorted ordered blocks.
// Selection and sorting run as device sorts driven by the host module
// (vlxmi/devpairs.py); the Schwarz factors come from vlxmi_schwarz_{ab}.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "vlxmi_device.hpp"

namespace {
inline uint32_t ncart(uint32_t l) { return (l + 1) * (l + 2) / 2; }
}

// One workgroup per primitive I of type a; threads over the partners J (J >= I
// for one type).  Block index: I*count_b + J, or for one type the row-major
// upper triangle I*count - I*(I-1)/2 + (J - I).
__global__ void __launch_bounds__(256)
pair_candidates_kernel(uint32_t na, uint32_t nb, int same, const double* __restrict__ info_a, uint32_t count_a,
                       const double* __restrict__ info_b, uint32_t count_b, const double* __restrict__ fac_a,
                       const double* __restrict__ fac_b, uint32_t* __restrict__ Iout, uint32_t* __restrict__ Jout,
                       double* __restrict__ S)
{
    // the host builder's operation order, unfused (vlxmi/pairs.py)
#pragma clang fp contract(off)
    const uint32_t I = blockIdx.x;
    if (I >= count_a) return;
    const double ai = info_a[I];
    const double Ax = info_a[I + 2 * count_a], Ay = info_a[I + 3 * count_a], Az = info_a[I + 4 * count_a];
    const uint32_t j0 = same ? I : 0;
    const uint64_t row = same ? (uint64_t)I * count_b - (uint64_t)I * (I - 1) / 2 : (uint64_t)I * count_b;
    for (uint32_t J = j0 + threadIdx.x; J < count_b; J += blockDim.x) {
        const double aj = info_b[J];
        const double dx = Ax - info_b[J + 2 * count_b], dy = Ay - info_b[J + 3 * count_b],
                     dz = Az - info_b[J + 4 * count_b];
        const double z = ai + aj;
        const double ab2 = dx * dx + dy * dy + dz * dz;
        const double base = pow(M_PI / z, 1.5) * exp(-ai * aj / z * ab2);
        const uint64_t b = row + (J - j0);
        Iout[b] = I;
        Jout[b] = J;
        for (uint32_t ca = 0; ca < na; ++ca)
            for (uint32_t cb = 0; cb < nb; ++cb)
                S[(b * na + ca) * nb + cb] = base * fac_a[(size_t)I * na + ca] * fac_b[(size_t)J * nb + cb];
    }
}

extern "C" int vlxmi_pair_candidates(uint32_t la, uint32_t lb, const double* info_a, uint32_t count_a,
                                     const double* info_b, uint32_t count_b, const double* fac_a, const double* fac_b,
                                     uint32_t* I, uint32_t* J, double* S, void* stream)
{
    if (la > 2 || lb > 2 || la > lb) return VLXMI_ERR_ARG;
    if (count_a == 0 || count_b == 0) return VLXMI_OK;
    if (!info_a || !info_b || !fac_a || !fac_b || !I || !J || !S) return VLXMI_ERR_ARG;
    if (la == lb && count_a != count_b) return VLXMI_ERR_ARG;
    hipLaunchKernelGGL(pair_candidates_kernel, dim3(count_a), dim3(256), 0, (hipStream_t)stream, ncart(la),
                       ncart(lb), (int)(la == lb), info_a, count_a, info_b, count_b, fac_a, fac_b, I, J, S);
    return (int)hipGetLastError();
}

__global__ void block_absmax_kernel(const double* __restrict__ x, uint64_t nblk, uint32_t nab,
                                    double* __restrict__ out)
{
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    double m = 0.0;
    for (uint32_t t = 0; t < nab; ++t) m = fmax(m, fabs(x[b * nab + t]));
    out[b] = m;
}

extern "C" int vlxmi_block_absmax(const double* x, uint64_t nblk, uint32_t nab, double* out, void* stream)
{
    if (nblk == 0) return VLXMI_OK;
    if (!x || !out || nab == 0) return VLXMI_ERR_ARG;
    hipLaunchKernelGGL(block_absmax_kernel, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       x, nblk, nab, out);
    return (int)hipGetLastError();
}

// Exchange CSR entries.  The ordered blocks (I of type a, J of type b) are
// given by src (index into a pair set's blocks) and trans (the pair set holds
// them as (J, I), so S and Q are read transposed), sorted by I, then by
// non-increasing block max, then J; rank[p] is the block's position among I's
// partners.  Entry of (I, ca, rank, cb): base[I] + ca * seg[I] + rank * nb + cb,
// seg[I] = nb * (number of partners of I).
__global__ void kcsr_fill_kernel(uint32_t na, uint32_t nb, const int64_t* __restrict__ src,
                                 const uint8_t* __restrict__ trans, const int64_t* __restrict__ Iord,
                                 const int64_t* __restrict__ Jord, const int64_t* __restrict__ rank,
                                 const int64_t* __restrict__ base, const int64_t* __restrict__ seg, uint64_t nord,
                                 const double* __restrict__ S, const double* __restrict__ Q,
                                 double* __restrict__ Qe, double* __restrict__ Se, uint32_t* __restrict__ De)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nord) return;
    const int64_t b = src[p];
    const int tr = trans[p];
    const int64_t I = Iord[p], J = Jord[p];
    for (uint32_t ca = 0; ca < na; ++ca)
        for (uint32_t cb = 0; cb < nb; ++cb) {
            const size_t s = tr ? ((size_t)b * nb + cb) * na + ca : ((size_t)b * na + ca) * nb + cb;
            const size_t e = (size_t)(base[I] + ca * seg[I] + rank[p] * nb + cb);
            Qe[e] = Q[s];
            Se[e] = S[s];
            De[e] = (uint32_t)(J * nb + cb);
        }
}

extern "C" int vlxmi_kcsr_fill(uint32_t la, uint32_t lb, const int64_t* src, const uint8_t* trans,
                               const int64_t* Iord, const int64_t* Jord, const int64_t* rank, const int64_t* base,
                               const int64_t* seg, uint64_t nord, const double* S, const double* Q, double* Qe,
                               double* Se, uint32_t* De, void* stream)
{
    if (la > 2 || lb > 2) return VLXMI_ERR_ARG;
    if (nord == 0) return VLXMI_OK;
    if (!src || !trans || !Iord || !Jord || !rank || !base || !seg || !S || !Q || !Qe || !Se || !De)
        return VLXMI_ERR_ARG;
    hipLaunchKernelGGL(kcsr_fill_kernel, dim3((unsigned)((nord + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       ncart(la), ncart(lb), src, trans, Iord, Jord, rank, base, seg, nord, S, Q, Qe, Se, De);
    return (int)hipGetLastError();
}
