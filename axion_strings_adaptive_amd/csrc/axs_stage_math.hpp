// axs_stage_math.hpp -- per-cell arithmetic of the fused RHS + RK4 stage,
// shared by the uniform-level kernel (axs_rk.hip) and the box work-list
// kernel (axs_boxes.hip).
//
// FP64 follows the CPU oracle's operation order with contraction off, which
// makes FP64 results bit-identical to it: the published 4th-order weights
// (1/12, 4/3, 5/2) of GRTeclyn's FourthOrderDerivatives::diff2 summed x, y,
// z, then AxionStringsRHS.impl.hpp:105-117 term by term.  FP32 writes the
// Laplacian on differences from the centre value (exact by Sterbenz' lemma
// for a smooth field), removing the cancellation of the -30 f0 term when
// |psi| ~ R >> 1.  Every FP32 expression is also written out with explicit
// fmaf and contraction off, so FP32 results do not depend on how a kernel's
// shape (tile, cells per lane) lets the compiler fuse operations: every tile
// variant and the box kernel produce identical bits.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/axs.h"

namespace axs
{

enum
{
    MODE_RHS = 0, // write k = f(y) (AxionStringsRHS over the level)
    MODE_S1 = 1,  // acc = k1, y1 = y + (dt/2) k1
    MODE_S23 = 2, // acc += 2 k, y_{s} = y + a k
    MODE_S4 = 3   // y_{n+1} = y + (dt/6)(acc + k4)  [+ Kahan for FP32C]
};

__device__ __forceinline__ double diff2_ref(double fm2, double fm1, double f0,
                                            double fp1, double fp2, double i2)
{
#pragma clang fp contract(off)
    const double w_far = 8.33333333333333333333e-2;
    const double w_near = 1.33333333333333333333e+0;
    const double w_local = 2.50000000000000000000e+0;
    return (-w_far * fm2 + w_near * fm1 - w_local * f0 + w_near * fp1 -
            w_far * fp2) *
           i2;
}

// The 13-point Laplacian of one component: x, y, z neighbours at -2, -1,
// +1, +2 and the centre.
struct Star
{
    double xm2, xm1, xp1, xp2, ym2, ym1, yp1, yp2, zm2, zm1, zp1, zp2, c;
};
struct StarF
{
    float xm2, xm1, xp1, xp2, ym2, ym1, yp1, yp2, zm2, zm1, zp1, zp2, c;
};

__device__ __forceinline__ double star_lap(const Star &s, double i2)
{
    double l = 0.0;
    l += diff2_ref(s.xm2, s.xm1, s.c, s.xp1, s.xp2, i2);
    l += diff2_ref(s.ym2, s.ym1, s.c, s.yp1, s.yp2, i2);
    l += diff2_ref(s.zm2, s.zm1, s.c, s.zp1, s.zp2, i2);
    return l;
}

// The weights carry 1/dx^2: (4/3) i2 and (1/12) i2 are formed once per
// kernel (i2 is invariant), so a Laplacian costs one multiply and one fma
// after the sums.  For dx a power of two (every level here: dx = 0.5 and
// its refinements) the products are exact and the result is bitwise the
// unscaled form times i2.
__device__ __forceinline__ float star_lap(const StarF &s, float i2)
{
#pragma clang fp contract(off)
    const float w_near = (4.0f / 3.0f) * i2, w_far = (1.0f / 12.0f) * i2;
    const float c = s.c;
    const float nsum = ((s.xm1 - c) + (s.xp1 - c)) + ((s.ym1 - c) + (s.yp1 - c)) +
                       ((s.zm1 - c) + (s.zp1 - c));
    const float fsum = ((s.xm2 - c) + (s.xp2 - c)) + ((s.ym2 - c) + (s.yp2 - c)) +
                       ((s.zm2 - c) + (s.zp2 - c));
    return fmaf(w_near, nsum, -(w_far * fsum));
}

// Pi' = lap psi + c psi - (lambda/2)(|psi|^2 - R^2) psi
__device__ __forceinline__ void force(double lap1, double lap2, double psi1,
                                      double psi2, double cc, double lam,
                                      double R2, double &F1, double &F2)
{
#pragma clang fp contract(off)
    const double d = psi1 * psi1 + psi2 * psi2 - R2;
    const double pc = 0.5 * lam * d;
    F1 = lap1 + cc * psi1 - pc * psi1;
    F2 = lap2 + cc * psi2 - pc * psi2;
}

// FP32 force with the stage scalars already rounded: ccf = (float)c,
// hlam = 0.5f * (float)lambda, r2f = (float)R^2.  The two linear terms
// share one coefficient, c - (lambda/2)(|psi|^2 - R^2), formed by one fma
// (five operations for both components); where c = 0 (radiation era) it is
// exactly -(lambda/2)(|psi|^2 - R^2) and the bits are those of the
// term-by-term form.
__device__ __forceinline__ void force_f(float lap1, float lap2, float psi1,
                                        float psi2, float ccf, float hlam,
                                        float r2f, float &F1, float &F2)
{
#pragma clang fp contract(off)
    const float d = fmaf(psi1, psi1, fmaf(psi2, psi2, -r2f));
    const float cm = fmaf(-hlam, d, ccf);
    F1 = fmaf(cm, psi1, lap1);
    F2 = fmaf(cm, psi2, lap2);
}

__device__ __forceinline__ void force(float lap1, float lap2, float psi1,
                                      float psi2, double cc, double lam,
                                      double R2, float &F1, float &F2)
{
    force_f(lap1, lap2, psi1, psi2, (float)cc, 0.5f * (float)lam, (float)R2,
            F1, F2);
}

__device__ __forceinline__ double axpy_f64(double y, double a, double k)
{
#pragma clang fp contract(off)
    return y + a * k;
}
__device__ __forceinline__ double acc2_f64(double acc, double k)
{
#pragma clang fp contract(off)
    return acc + 2.0 * k;
}
__device__ __forceinline__ double final_f64(double y, double h6, double acc,
                                            double k)
{
#pragma clang fp contract(off)
    return y + h6 * (acc + k);
}

// y + a k, as stages 1-3 form their outputs (the fused stage pairs also
// form the next stage's psi at stencil points with it, so both agree to the
// bit).
template <typename T, int PREC>
__device__ __forceinline__ T stage_axpy(T y, double a, T k)
{
#pragma clang fp contract(off)
    if constexpr (PREC == AXS_FP64)
        return axpy_f64(y, a, k);
    else
        return fmaf((float)a, k, y);
}

// FP32 / FP32C stage update with the scalars already rounded to float:
// a_f = (float)a, h6_f = (float)(dt/6), h6_lo the low part of dt/6.
template <int PREC, int MODE>
__device__ __forceinline__ void stage_update_f(const float k[4],
                                               const float yn[4],
                                               const float ac[4], float lo[4],
                                               float a_f, float h6_f,
                                               float h6_lo, float acc_out[4],
                                               float y_out[4])
{
#pragma clang fp contract(off)
    if constexpr (MODE == MODE_RHS)
    {
        for (int c = 0; c < 4; ++c)
            y_out[c] = k[c];
    }
    else if constexpr (MODE == MODE_S1)
    {
        for (int c = 0; c < 4; ++c)
        {
            acc_out[c] = k[c];
            y_out[c] = fmaf(a_f, k[c], yn[c]);
        }
    }
    else if constexpr (MODE == MODE_S23)
    {
        for (int c = 0; c < 4; ++c)
        {
            acc_out[c] = fmaf(2.0f, k[c], ac[c]);
            y_out[c] = fmaf(a_f, k[c], yn[c]);
        }
    }
    else
    {
        for (int c = 0; c < 4; ++c)
        {
            if constexpr (PREC == AXS_FP32)
                y_out[c] = fmaf(h6_f, ac[c] + k[c], yn[c]);
            else
            {
                // Kahan: true y_n = hi + lo; add delta = (dt/6)(acc + k)
                // with dt/6 carried as hi + lo floats.
                const float sv = ac[c] + k[c];
                const float delta = fmaf(h6_f, sv, h6_lo * sv);
                const float yk = delta + lo[c];
                const float t = yn[c] + yk;
                lo[c] = yk - (t - yn[c]);
                y_out[c] = t;
            }
        }
    }
}

// Stage update of one cell.  k[4] = (Pi1, Pi2, F1, F2) of the stage input;
// yn[4] = y_n (stage 1: the stage input itself), ac[4] the accumulator,
// lo[4] the FP32C low part of y_n.  Outputs: acc_out (stages 1-3),
// y_out (the next stage input, or y_{n+1} at stage 4), lo (stage 4 FP32C).
template <typename T, int PREC, int MODE>
__device__ __forceinline__ void stage_update(const T k[4], const T yn[4],
                                             const T ac[4], float lo[4],
                                             double a, double h6, float h6_lo,
                                             T acc_out[4], T y_out[4])
{
#pragma clang fp contract(off)
    if constexpr (PREC != AXS_FP64)
    {
        stage_update_f<PREC, MODE>(k, yn, ac, lo, (float)a, (float)h6, h6_lo,
                                   acc_out, y_out);
    }
    else if constexpr (MODE == MODE_RHS)
    {
        for (int c = 0; c < 4; ++c)
            y_out[c] = k[c];
    }
    else if constexpr (MODE == MODE_S1)
    {
        for (int c = 0; c < 4; ++c)
        {
            acc_out[c] = k[c];
            y_out[c] = axpy_f64(yn[c], a, k[c]);
        }
    }
    else if constexpr (MODE == MODE_S23)
    {
        for (int c = 0; c < 4; ++c)
        {
            acc_out[c] = acc2_f64(ac[c], k[c]);
            y_out[c] = axpy_f64(yn[c], a, k[c]);
        }
    }
    else
    {
        for (int c = 0; c < 4; ++c)
            y_out[c] = final_f64(yn[c], h6, ac[c], k[c]);
    }
}

} // namespace axs
