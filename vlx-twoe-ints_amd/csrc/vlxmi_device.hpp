// Device-side constants and the Boys function for the J/K kernels (gfx950).
//
// Boys function F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt, m = 0..L (L <= 8).
// The reference calls VeloxChem's gpu::computeBoysFunction(F, T, L,
// boys_func_table, boys_func_ft) (gen_coulomb.py:257-259); that function is not
// in the reference tree, so the table format below is this project's own:
//
//   boys_func_table: VLX_BOYS_NROW fine rows x VLX_BOYS_NCOL doubles, row k
//     holds F_0..F_14 at T_k = k * VLX_BOYS_DT (columns 0..14) and exp(-T_k)
//     (column 15); one 128-byte row per grid point; then VLX_BOYS_CROWS coarse
//     rows (below).
//   boys_func_ft: 1/(2m+1), m = 0..31 (the generated kernels fold these
//     constants at compile time; the array is kept for the signature).
//
// For T < VLX_BOYS_TMAX: a 6-term Taylor expansion of F_L about the nearest
// grid point, |T - T_k| <= DT/2 = 0.01, so the remainder is below
// F_L * 0.01^6 / 6! = 1.4e-15 F_L (F_{m+1} <= F_m); the six coefficients
// F_L..F_{L+5}(T_k) are contiguous, three 16-byte loads per lane -- the table
// gather is what limits the light classes, so the row is kept that short.
// exp(-T): for L <= 4 from the device exp (those classes are bound by the
// table gathers, not by instructions); for L >= 5, whose kernels re-evaluate
// the prologue once per component pass and are not gather-bound, as
// exp(-T_k) (column 15) times a 7-term Taylor of exp(T_k - T) (remainder
// < 0.01^7/7! = 2e-18).  Then downward recursion
// F_{m} = (2T F_{m+1} + exp(-T)) / (2m+1).
// For T >= VLX_BOYS_TMAX: F_0 = sqrt(pi/T)/2 (Newton-refined 1/sqrt) and upward recursion
// F_{m+1} = ((2m+1) F_m - exp(-T)) / (2T); there exp(-T) < 5e-18 and the
// single-precision hardware exponential is accurate enough for it.
//
// Coarse rows (after the fine table, VLX_BOYS_CROWS rows of VLX_BOYS_NCOL
// doubles): F_0..F_15 at T_k = k * VLX_BOYS_CDT.  Kernels of a few light
// classes copy the 12 columns F_L..F_{L+11} of these rows, scaled by 1/j!, into
// LDS once per workgroup (9.7 KB) and evaluate F_L by a 12-term Taylor expansion
// about the nearest coarse point, |T - T_k| <= 0.2, remainder below
// 0.2^12 / 12! = 8.6e-18 (relative to F_L): no global gather per quartet.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vlxmi_status.h"

// List-order validation the class launchers run when it is switched on
// (vlxmi_lists.cpp).
extern "C" int vlxmi_list_validation(void);
extern "C" int vlxmi_validate_j_list(uint32_t la, uint32_t lb, const uint32_t* first, const uint32_t* second,
                                     const double* mat_Q, const double* mat_D, uint32_t count, void* stream);
extern "C" int vlxmi_validate_ik_list(uint32_t la, uint32_t lc, const uint32_t* ik_i, const uint32_t* ik_k,
                                      uint32_t count, void* stream);
extern "C" int vlxmi_validate_k_csr(uint32_t la, uint32_t lb, uint32_t count_a, const uint32_t* displs,
                                    const uint32_t* counts, const double* Q, const uint32_t* D_inds, void* stream);

// Ket-split partial sums of the Coulomb launchers (vlxmi_capi.cpp).
extern "C" int vlxmi_ket_workspace(size_t ndoubles, void* stream, double** out);
extern "C" int vlxmi_ket_split_reduce(const double* part, uint32_t nsplit, uint32_t count, double* out,
                                      void* stream);

#define VLX_BOYS_NCOL 16
#define VLX_BOYS_NROW 2001
#define VLX_BOYS_DT 0.02
#define VLX_BOYS_INV_DT 50.0
#define VLX_BOYS_TMAX 40.0
#define VLX_BOYS_CROWS 101
#define VLX_BOYS_CDT 0.4
#define VLX_BOYS_INV_CDT 2.5
#define VLX_BOYS_CTERMS 12
#define VLX_BOYS_COARSE_OFF (VLX_BOYS_NROW * VLX_BOYS_NCOL)

#define MATH_CONST_PI 3.14159265358979323846
#define MATH_CONST_INV_PI 0.31830988618379067154

// one wavefront per workgroup for both kernel families
#define TILE_SIZE_J 64
#define TILE_SIZE_K 64

namespace vlxmi {

// Orders one wavefront's LDS accesses around this point: the ket tile a
// wavefront stages is read only by its own lanes, so kernels that run several
// independent wavefronts per workgroup synchronise each wavefront on its own.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 1/x and 1/sqrt(x) for positive, normal x well inside the double range (the
// Gaussian exponent sums and Boys arguments these are used on): the hardware
// estimate refined by two Newton steps, each of which squares the relative
// error, so the result is within a few ulp; no denormal/overflow scaling.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double rsq_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double t = fma(-h * y, y, 0.5);
    y = fma(y, t, y);
    t = fma(-h * y, y, 0.5);
    return fma(y, t, y);
}

template <int L>
__device__ __forceinline__ void boys(double T, const double* __restrict__ table, double* F)
{
    if (T < VLX_BOYS_TMAX) {
        const int k = (int)(T * VLX_BOYS_INV_DT + 0.5);
        const double nd = (double)k * VLX_BOYS_DT - T;   // -(T - T_k)
        const double* row = table + k * VLX_BOYS_NCOL;
        // F_L(T) = sum_j F_{L+j}(T_k) (-d)^j / j!, j = 0..5 (Horner)
        double s = row[L + 5];
        s = fma(s, nd * (1.0 / 5.0), row[L + 4]);
        s = fma(s, nd * (1.0 / 4.0), row[L + 3]);
        s = fma(s, nd * (1.0 / 3.0), row[L + 2]);
        s = fma(s, nd * (1.0 / 2.0), row[L + 1]);
        s = fma(s, nd, row[L]);
        F[L] = s;
        if (L > 0) {
            double et;
            if constexpr (L >= 5) {
                double e = 1.0 / 720.0;
                e = fma(e, nd, 1.0 / 120.0);
                e = fma(e, nd, 1.0 / 24.0);
                e = fma(e, nd, 1.0 / 6.0);
                e = fma(e, nd, 0.5);
                e = fma(e, nd, 1.0);
                e = fma(e, nd, 1.0);
                et = row[15] * e;
            } else {
                et = exp(-T);
            }
            const double T2 = 2.0 * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = fma(T2, F[m + 1], et) * (1.0 / (2.0 * m + 1.0));
        }
    } else {
        const double inv_T = rcp_nr(T);
        F[0] = (0.5 * 1.77245385090551602730) * rsq_nr(T);   // sqrt(pi/T) / 2
        if (L > 0) {
            const double et = (double)__expf(-(float)T);
            const double h = 0.5 * inv_T;
#pragma unroll
            for (int m = 0; m < L; ++m) F[m + 1] = ((2.0 * m + 1.0) * F[m] - et) * h;
        }
    }
}

// The LDS copy of the coarse rows for order L: lt[k * 12 + j] = F_{L+j}(T_k) / j!.
// Every thread of the workgroup calls it; the caller synchronises afterwards.
template <int L>
__device__ __forceinline__ void boys_lds_fill(const double* __restrict__ table, double* lt)
{
    static_assert(L + VLX_BOYS_CTERMS - 1 < VLX_BOYS_NCOL, "coarse rows hold F_0..F_15");
    constexpr double inv_fact[VLX_BOYS_CTERMS] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720,
                                                  1.0 / 5040, 1.0 / 40320, 1.0 / 362880, 1.0 / 3628800,
                                                  1.0 / 39916800};
    for (int t = threadIdx.x; t < VLX_BOYS_CROWS * VLX_BOYS_CTERMS; t += blockDim.x) {
        const int r = t / VLX_BOYS_CTERMS, j = t - r * VLX_BOYS_CTERMS;
        lt[t] = table[VLX_BOYS_COARSE_OFF + r * VLX_BOYS_NCOL + L + j] * inv_fact[j];
    }
}

// F_0..F_L from the LDS copy of the coarse rows (T < VLX_BOYS_TMAX), exp(-T)
// and downward recursion; the T >= VLX_BOYS_TMAX branch is that of boys<L>.
template <int L>
__device__ __forceinline__ void boys_lds(double T, const double* lt, double* F)
{
    if (T < VLX_BOYS_TMAX) {
        const int k = (int)(T * VLX_BOYS_INV_CDT + 0.5);
        const double nd = (double)k * VLX_BOYS_CDT - T;   // -(T - T_k)
        const double* row = lt + k * VLX_BOYS_CTERMS;
        double s = row[VLX_BOYS_CTERMS - 1];
#pragma unroll
        for (int j = VLX_BOYS_CTERMS - 2; j >= 0; --j) s = fma(s, nd, row[j]);
        F[L] = s;
        if (L > 0) {
            const double et = exp(-T);
            const double T2 = 2.0 * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = fma(T2, F[m + 1], et) * (1.0 / (2.0 * m + 1.0));
        }
    } else {
        const double inv_T = rcp_nr(T);
        F[0] = (0.5 * 1.77245385090551602730) * rsq_nr(T);
        if (L > 0) {
            const double et = (double)__expf(-(float)T);
            const double h = 0.5 * inv_T;
#pragma unroll
            for (int m = 0; m < L; ++m) F[m + 1] = ((2.0 * m + 1.0) * F[m] - et) * h;
        }
    }
}

}  // namespace vlxmi
