// Device-side constants and the Boys function for the J/K kernels (gfx950).
//
// Boys function F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt, m = 0..L (L <= 8).
// The reference calls VeloxChem's gpu::computeBoysFunction(F, T, L,
// boys_func_table, boys_func_ft) (gen_coulomb.py:257-259); that function is not
// in the reference tree, so the table format below is this project's own:
//
//   boys_func_table: VLX_BOYS_NROW rows x VLX_BOYS_NCOL doubles, row k holds
//     F_0..F_14 at T_k = k * VLX_BOYS_DT (columns 0..14) and exp(-T_k)
//     (column 15); one 128-byte row per grid point.
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
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vlxmi_status.h"

#define VLX_BOYS_NCOL 16
#define VLX_BOYS_NROW 2001
#define VLX_BOYS_DT 0.02
#define VLX_BOYS_INV_DT 50.0
#define VLX_BOYS_TMAX 40.0

#define MATH_CONST_PI 3.14159265358979323846
#define MATH_CONST_INV_PI 0.31830988618379067154

// one wavefront per workgroup for both kernel families
#define TILE_SIZE_J 64
#define TILE_SIZE_K 64

namespace vlxmi {

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

}  // namespace vlxmi
