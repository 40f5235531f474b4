// axs_rk.hip -- fused RHS + RK4-stage kernels for a uniform periodic level.
//
// One launch per RK stage.  Each workgroup owns a 64 x TY column tile of the
// x-y plane and streams through a chunk of z planes:
//   * psi1/psi2 of the stage input are read once per cell per stage: the z
//     neighbours (radius 2) live in a 5-deep register queue per thread, the
//     x/y neighbours come from a double-buffered LDS plane with a 2-cell halo
//     (one barrier per plane);
//   * the pointwise RHS (AxionStringsRHS.impl.hpp:83-118) and the RK4 stage
//     axpy are applied in registers and the stage outputs written once.
// Byte traffic per cell and stage (4 components of S bytes each):
//   stage 1: read y_n, write acc, write y_1                     = 3 S
//   stage 2,3: read y_s, y_n, acc, write acc, y_{s+1}             = 5 S
//   stage 4: read y_3, y_n, acc, write y_{n+1} (+ read/write lo)  = 4 S (6 S)
// = 17 S per RK4 step in FP64 (544 B/cell), 19 S in FP32C (304 B/cell).
//
// FP64 arithmetic follows the CPU oracle's operation order with contraction
// off, so FP64 results are bit-identical to it.  FP32 arithmetic writes the
// 4th-order Laplacian on differences from the centre value (exact by
// Sterbenz' lemma for a smooth field), which removes the cancellation of the
// -30 f0 term when |psi| ~ R >> 1.
#include "axs_internal.hpp"

namespace axs
{

enum
{
    MODE_RHS = 0,
    MODE_S1 = 1,
    MODE_S23 = 2,
    MODE_S4 = 3
};

template <typename T> struct StageParams
{
    const T *yin;  // stage input (4 comps)
    const T *yn;   // y_n (4 comps), stages 2..4
    T *acc;        // accumulator (4 comps)
    T *yout;       // output (4 comps); stage 4: the y_n buffer itself
    float *comp;   // FP32C low part of y_n (stage 4)
    const double *coeff; // stage-scalar table [step][4][3]
    const int *step;     // device step index or nullptr
    int coeff_stage;
    double a;      // stage axpy factor: dt/2, dt/2, dt
    double h6;     // dt/6
    float h6_lo;   // dt/6 - (float)(dt/6)
    int nx, ny, nz;
    uint32_t n;
    double inv_dx2;
    int z_chunk;
    int tiles_x, tiles_y, tiles_z;
    int xcd_remap;
};

__device__ __forceinline__ int wrap_lo(int i, int n) { return i < 0 ? i + n : i; }
__device__ __forceinline__ int wrap_hi(int i, int n) { return i >= n ? i - n : i; }

// ---------------------------------------------------------------- FP64 --
// Exact restatement of the oracle's arithmetic (contraction off).
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

__device__ __forceinline__ void rhs_pi_f64(double lap1, double lap2,
                                           double psi1, double psi2,
                                           double cc, double lam, double R2,
                                           double &F1, double &F2)
{
#pragma clang fp contract(off)
    const double d = psi1 * psi1 + psi2 * psi2 - R2;
    const double pc = 0.5 * lam * d;
    F1 = lap1 + cc * psi1 - pc * psi1;
    F2 = lap2 + cc * psi2 - pc * psi2;
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

template <typename T, int PREC, int MODE, int TY>
__global__ void __launch_bounds__(64 * TY)
    k_stage(StageParams<T> p)
{
    constexpr int TX = 64;
    constexpr int HX = TX + 4;
    constexpr int HY = TY + 4;
    __shared__ T sm[2][2][HY][HX];

    // ---- tile (XCD-aware: consecutive tiles land on one XCD's L2) ----
    int bid = blockIdx.x;
    if (p.xcd_remap)
    {
        const int ntiles = p.tiles_x * p.tiles_y * p.tiles_z;
        bid = (bid & 7) * (ntiles >> 3) + (bid >> 3);
    }
    const int bx = bid % p.tiles_x;
    const int rem = bid / p.tiles_x;
    const int by = rem % p.tiles_y;
    const int bz = rem / p.tiles_y;

    const int tx = threadIdx.x;
    const int ty = threadIdx.y;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t n = p.n;
    const uint32_t plane = (uint32_t)nx * (uint32_t)ny;
    const int gx = bx * TX + tx;
    const int gy = by * TY + ty;
    const bool active = gx < nx && gy < ny;
    const int xw = gx % nx;
    const int yw = gy % ny;
    const uint32_t rowc = (uint32_t)yw * nx + xw;

    // halo roles
    const bool hxl = tx < 2, hxr = tx >= TX - 2;
    const bool hyl = ty < 2, hyr = ty >= TY - 2;
    uint32_t rowhx = 0, rowhy = 0;
    if (hxl || hxr)
    {
        int hx = hxl ? (bx * TX - 2 + tx) : (bx * TX + TX + (tx - (TX - 2)));
        hx = ((hx % nx) + nx) % nx;
        rowhx = (uint32_t)yw * nx + hx;
    }
    if (hyl || hyr)
    {
        int hy = hyl ? (by * TY - 2 + ty) : (by * TY + TY + (ty - (TY - 2)));
        hy = ((hy % ny) + ny) % ny;
        rowhy = (uint32_t)hy * nx + xw;
    }

    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);

    const T *psi1 = p.yin;
    const T *psi2 = p.yin + n;

    auto zoff = [&](int z) -> uint32_t {
        z = wrap_lo(wrap_hi(z, nz), nz);
        return (uint32_t)z * plane;
    };

    // register queues: psi at z-2 .. z+2
    T q1[5], q2[5];
#pragma unroll
    for (int i = 0; i < 5; ++i)
    {
        const uint32_t o = zoff(z_begin - 2 + i) + rowc;
        q1[i] = psi1[o];
        q2[i] = psi2[o];
    }
    T hx1 = 0, hx2 = 0, hy1 = 0, hy2 = 0;
    {
        const uint32_t o = zoff(z_begin);
        if (hxl || hxr)
        {
            hx1 = psi1[o + rowhx];
            hx2 = psi2[o + rowhx];
        }
        if (hyl || hyr)
        {
            hy1 = psi1[o + rowhy];
            hy2 = psi2[o + rowhy];
        }
    }

    // stage scalars
    const double *cf =
        p.coeff + 12 * (p.step ? *p.step : 0) + 3 * p.coeff_stage;
    const double cc_d = cf[0], lam_d = cf[1], R2_d = cf[2];

    for (int z = z_begin; z < z_end; ++z)
    {
        const int b = (z - z_begin) & 1;
        // ---- stage the plane into LDS ----
        sm[b][0][ty + 2][tx + 2] = q1[2];
        sm[b][1][ty + 2][tx + 2] = q2[2];
        if (hxl)
        {
            sm[b][0][ty + 2][tx] = hx1;
            sm[b][1][ty + 2][tx] = hx2;
        }
        else if (hxr)
        {
            sm[b][0][ty + 2][tx + 4] = hx1;
            sm[b][1][ty + 2][tx + 4] = hx2;
        }
        if (hyl)
        {
            sm[b][0][ty][tx + 2] = hy1;
            sm[b][1][ty][tx + 2] = hy2;
        }
        else if (hyr)
        {
            sm[b][0][ty + 4][tx + 2] = hy1;
            sm[b][1][ty + 4][tx + 2] = hy2;
        }

        // ---- prefetch next plane (queue head + halos) ----
        T nq1 = q1[4], nq2 = q2[4];
        const bool more = z + 1 < z_end;
        if (more)
        {
            const uint32_t o3 = zoff(z + 3) + rowc;
            nq1 = psi1[o3];
            nq2 = psi2[o3];
            const uint32_t o1 = zoff(z + 1);
            if (hxl || hxr)
            {
                hx1 = psi1[o1 + rowhx];
                hx2 = psi2[o1 + rowhx];
            }
            if (hyl || hyr)
            {
                hy1 = psi1[o1 + rowhy];
                hy2 = psi2[o1 + rowhy];
            }
        }

        // ---- centre loads of this plane ----
        const uint32_t oc = (uint32_t)z * plane + rowc;
        T pi1 = 0, pi2 = 0;
        T yn0 = 0, yn1 = 0, yn2 = 0, yn3 = 0;
        T ac0 = 0, ac1 = 0, ac2 = 0, ac3 = 0;
        float lo0 = 0, lo1 = 0, lo2 = 0, lo3 = 0;
        if (active)
        {
            pi1 = p.yin[2 * n + oc];
            pi2 = p.yin[3 * n + oc];
            if (MODE == MODE_S23 || MODE == MODE_S4)
            {
                yn0 = p.yn[oc];
                yn1 = p.yn[n + oc];
                yn2 = p.yn[2 * n + oc];
                yn3 = p.yn[3 * n + oc];
                ac0 = p.acc[oc];
                ac1 = p.acc[n + oc];
                ac2 = p.acc[2 * n + oc];
                ac3 = p.acc[3 * n + oc];
            }
            if (MODE == MODE_S4 && PREC == AXS_FP32C)
            {
                lo0 = p.comp[oc];
                lo1 = p.comp[n + oc];
                lo2 = p.comp[2 * n + oc];
                lo3 = p.comp[3 * n + oc];
            }
        }

        __syncthreads();

        if (active)
        {
            const int cy = ty + 2, cx = tx + 2;
            T F1, F2;
            if constexpr (PREC == AXS_FP64)
            {
                const double i2 = p.inv_dx2;
                double l1 = 0.0, l2 = 0.0;
                l1 += diff2_ref(sm[b][0][cy][cx - 2], sm[b][0][cy][cx - 1],
                                q1[2], sm[b][0][cy][cx + 1],
                                sm[b][0][cy][cx + 2], i2);
                l1 += diff2_ref(sm[b][0][cy - 2][cx], sm[b][0][cy - 1][cx],
                                q1[2], sm[b][0][cy + 1][cx],
                                sm[b][0][cy + 2][cx], i2);
                l1 += diff2_ref(q1[0], q1[1], q1[2], q1[3], q1[4], i2);
                l2 += diff2_ref(sm[b][1][cy][cx - 2], sm[b][1][cy][cx - 1],
                                q2[2], sm[b][1][cy][cx + 1],
                                sm[b][1][cy][cx + 2], i2);
                l2 += diff2_ref(sm[b][1][cy - 2][cx], sm[b][1][cy - 1][cx],
                                q2[2], sm[b][1][cy + 1][cx],
                                sm[b][1][cy + 2][cx], i2);
                l2 += diff2_ref(q2[0], q2[1], q2[2], q2[3], q2[4], i2);
                double f1, f2;
                rhs_pi_f64(l1, l2, q1[2], q2[2], cc_d, lam_d, R2_d, f1, f2);
                F1 = f1;
                F2 = f2;
            }
            else
            {
                const float w_near = 4.0f / 3.0f, w_far = 1.0f / 12.0f;
                const float i2 = (float)p.inv_dx2;
                const float c1 = q1[2], c2 = q2[2];
                const float n1 = ((sm[b][0][cy][cx - 1] - c1) +
                                  (sm[b][0][cy][cx + 1] - c1)) +
                                 ((sm[b][0][cy - 1][cx] - c1) +
                                  (sm[b][0][cy + 1][cx] - c1)) +
                                 ((q1[1] - c1) + (q1[3] - c1));
                const float f1 = ((sm[b][0][cy][cx - 2] - c1) +
                                  (sm[b][0][cy][cx + 2] - c1)) +
                                 ((sm[b][0][cy - 2][cx] - c1) +
                                  (sm[b][0][cy + 2][cx] - c1)) +
                                 ((q1[0] - c1) + (q1[4] - c1));
                const float n2 = ((sm[b][1][cy][cx - 1] - c2) +
                                  (sm[b][1][cy][cx + 1] - c2)) +
                                 ((sm[b][1][cy - 1][cx] - c2) +
                                  (sm[b][1][cy + 1][cx] - c2)) +
                                 ((q2[1] - c2) + (q2[3] - c2));
                const float f2 = ((sm[b][1][cy][cx - 2] - c2) +
                                  (sm[b][1][cy][cx + 2] - c2)) +
                                 ((sm[b][1][cy - 2][cx] - c2) +
                                  (sm[b][1][cy + 2][cx] - c2)) +
                                 ((q2[0] - c2) + (q2[4] - c2));
                const float lap1 = (w_near * n1 - w_far * f1) * i2;
                const float lap2 = (w_near * n2 - w_far * f2) * i2;
                const float d = c1 * c1 + c2 * c2 - (float)R2_d;
                const float pc = 0.5f * (float)lam_d * d;
                const float ccf = (float)cc_d;
                F1 = lap1 + ccf * c1 - pc * c1;
                F2 = lap2 + ccf * c2 - pc * c2;
            }

            // k = (Pi1, Pi2, F1, F2)
            if constexpr (MODE == MODE_RHS)
            {
                p.yout[oc] = pi1;
                p.yout[n + oc] = pi2;
                p.yout[2 * n + oc] = F1;
                p.yout[3 * n + oc] = F2;
            }
            else if constexpr (MODE == MODE_S1)
            {
                // y_n == y_in at stage 1
                p.acc[oc] = pi1;
                p.acc[n + oc] = pi2;
                p.acc[2 * n + oc] = F1;
                p.acc[3 * n + oc] = F2;
                if constexpr (PREC == AXS_FP64)
                {
                    const double a = p.a;
                    p.yout[oc] = axpy_f64(q1[2], a, pi1);
                    p.yout[n + oc] = axpy_f64(q2[2], a, pi2);
                    p.yout[2 * n + oc] = axpy_f64(pi1, a, F1);
                    p.yout[3 * n + oc] = axpy_f64(pi2, a, F2);
                }
                else
                {
                    const float a = (float)p.a;
                    p.yout[oc] = q1[2] + a * pi1;
                    p.yout[n + oc] = q2[2] + a * pi2;
                    p.yout[2 * n + oc] = pi1 + a * F1;
                    p.yout[3 * n + oc] = pi2 + a * F2;
                }
            }
            else if constexpr (MODE == MODE_S23)
            {
                if constexpr (PREC == AXS_FP64)
                {
                    const double a = p.a;
                    p.acc[oc] = acc2_f64(ac0, pi1);
                    p.acc[n + oc] = acc2_f64(ac1, pi2);
                    p.acc[2 * n + oc] = acc2_f64(ac2, F1);
                    p.acc[3 * n + oc] = acc2_f64(ac3, F2);
                    p.yout[oc] = axpy_f64(yn0, a, pi1);
                    p.yout[n + oc] = axpy_f64(yn1, a, pi2);
                    p.yout[2 * n + oc] = axpy_f64(yn2, a, F1);
                    p.yout[3 * n + oc] = axpy_f64(yn3, a, F2);
                }
                else
                {
                    const float a = (float)p.a;
                    p.acc[oc] = ac0 + 2.0f * pi1;
                    p.acc[n + oc] = ac1 + 2.0f * pi2;
                    p.acc[2 * n + oc] = ac2 + 2.0f * F1;
                    p.acc[3 * n + oc] = ac3 + 2.0f * F2;
                    p.yout[oc] = yn0 + a * pi1;
                    p.yout[n + oc] = yn1 + a * pi2;
                    p.yout[2 * n + oc] = yn2 + a * F1;
                    p.yout[3 * n + oc] = yn3 + a * F2;
                }
            }
            else // MODE_S4
            {
                if constexpr (PREC == AXS_FP64)
                {
                    const double h6 = p.h6;
                    p.yout[oc] = final_f64(yn0, h6, ac0, pi1);
                    p.yout[n + oc] = final_f64(yn1, h6, ac1, pi2);
                    p.yout[2 * n + oc] = final_f64(yn2, h6, ac2, F1);
                    p.yout[3 * n + oc] = final_f64(yn3, h6, ac3, F2);
                }
                else if constexpr (PREC == AXS_FP32)
                {
                    const float h6 = (float)p.h6;
                    p.yout[oc] = yn0 + h6 * (ac0 + pi1);
                    p.yout[n + oc] = yn1 + h6 * (ac1 + pi2);
                    p.yout[2 * n + oc] = yn2 + h6 * (ac2 + F1);
                    p.yout[3 * n + oc] = yn3 + h6 * (ac3 + F2);
                }
                else
                {
                    // Kahan: true y_n = hi + lo; add delta = (dt/6)(acc + k)
                    // with dt/6 carried as hi + lo floats.
                    const float h6 = (float)p.h6, h6l = p.h6_lo;
                    float hv[4] = {yn0, yn1, yn2, yn3};
                    float lv[4] = {lo0, lo1, lo2, lo3};
                    const float sv[4] = {ac0 + pi1, ac1 + pi2, ac2 + F1,
                                         ac3 + F2};
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                    {
                        const float delta = h6 * sv[c] + h6l * sv[c];
                        const float yk = delta + lv[c];
                        const float t = hv[c] + yk;
                        lv[c] = yk - (t - hv[c]);
                        hv[c] = t;
                    }
                    p.yout[oc] = hv[0];
                    p.yout[n + oc] = hv[1];
                    p.yout[2 * n + oc] = hv[2];
                    p.yout[3 * n + oc] = hv[3];
                    p.comp[oc] = lv[0];
                    p.comp[n + oc] = lv[1];
                    p.comp[2 * n + oc] = lv[2];
                    p.comp[3 * n + oc] = lv[3];
                }
            }
        }

        // ---- advance the queue ----
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            q1[i] = q1[i + 1];
            q2[i] = q2[i + 1];
        }
        q1[4] = nq1;
        q2[4] = nq2;
    }
}

__global__ void k_step_advance(int *step, int reset)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *step = reset ? 0 : *step + 1;
}

// ------------------------------------------------------------ launchers --
static const int kTileTY[3] = {4, 8, 16};

int num_tile_variants() { return 3; }

template <typename T, int PREC, int MODE, int TY>
static int launch_t(const StageParams<T> &p, int ntiles, hipStream_t s)
{
    hipLaunchKernelGGL((k_stage<T, PREC, MODE, TY>), dim3(ntiles),
                       dim3(64, TY), 0, s, p);
    return AXS_OK;
}

template <typename T, int PREC, int MODE>
static int launch_ty(int variant, const StageParams<T> &p, int ntiles,
                     hipStream_t s)
{
    switch (variant)
    {
    case 0:
        return launch_t<T, PREC, MODE, 4>(p, ntiles, s);
    case 1:
        return launch_t<T, PREC, MODE, 8>(p, ntiles, s);
    default:
        return launch_t<T, PREC, MODE, 16>(p, ntiles, s);
    }
}

template <typename T, int PREC>
static int launch_mode(int mode, int variant, const StageParams<T> &p,
                       int ntiles, hipStream_t s)
{
    switch (mode)
    {
    case MODE_RHS:
        return launch_ty<T, PREC, MODE_RHS>(variant, p, ntiles, s);
    case MODE_S1:
        return launch_ty<T, PREC, MODE_S1>(variant, p, ntiles, s);
    case MODE_S23:
        return launch_ty<T, PREC, MODE_S23>(variant, p, ntiles, s);
    default:
        return launch_ty<T, PREC, MODE_S4>(variant, p, ntiles, s);
    }
}

template <typename T>
static int fill_and_launch(axs_level *L, int mode, const void *yin, void *yout,
                           const double *coeff, const int *step,
                           int coeff_stage, double a, double dt)
{
    StageParams<T> p{};
    p.yin = static_cast<const T *>(yin);
    p.yn = static_cast<const T *>(L->y);
    p.acc = static_cast<T *>(L->acc);
    p.yout = static_cast<T *>(yout);
    p.comp = L->comp;
    p.coeff = coeff;
    p.step = step;
    p.coeff_stage = coeff_stage;
    p.a = a;
    p.h6 = dt / 6.0;
    p.h6_lo = (float)(p.h6 - (double)(float)p.h6);
    p.nx = L->g.nx;
    p.ny = L->g.ny;
    p.nz = L->g.nz;
    p.n = L->g.ncell;
    p.inv_dx2 = 1.0 / (L->dx * L->dx);
    const int variant =
        (L->tile_variant >= 0 && L->tile_variant < 3) ? L->tile_variant : 1;
    const int TY = kTileTY[variant];
    p.tiles_x = (p.nx + 63) / 64;
    p.tiles_y = (p.ny + TY - 1) / TY;
    int zc = L->z_chunk;
    if (zc <= 0)
    {
        zc = 64;
        while (zc > 4 && (long)p.tiles_x * p.tiles_y * ((p.nz + zc - 1) / zc) <
                             4096)
            zc >>= 1;
    }
    if (zc > p.nz)
        zc = p.nz;
    p.z_chunk = zc;
    p.tiles_z = (p.nz + zc - 1) / zc;
    const long ntiles = (long)p.tiles_x * p.tiles_y * p.tiles_z;
    p.xcd_remap = (ntiles % 8 == 0) ? 1 : 0;
    int rc;
    if constexpr (sizeof(T) == 8)
        rc = launch_mode<T, AXS_FP64>(mode, variant, p, (int)ntiles,
                                      L->ctx->stream);
    else if (L->precision == AXS_FP32C)
        rc = launch_mode<T, AXS_FP32C>(mode, variant, p, (int)ntiles,
                                       L->ctx->stream);
    else
        rc = launch_mode<T, AXS_FP32>(mode, variant, p, (int)ntiles,
                                      L->ctx->stream);
    if (rc != AXS_OK)
        return rc;
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

int launch_rk_stage(axs_level *L, int stage, const void *yin, void *yout,
                    const double *coeff_table, const int *step_ptr,
                    int coeff_stage, double dt)
{
    const int mode = stage == 1 ? MODE_S1 : (stage == 4 ? MODE_S4 : MODE_S23);
    const double a = stage == 3 ? dt : 0.5 * dt;
    if (L->precision == AXS_FP64)
        return fill_and_launch<double>(L, mode, yin, yout, coeff_table,
                                       step_ptr, coeff_stage, a, dt);
    return fill_and_launch<float>(L, mode, yin, yout, coeff_table, step_ptr,
                                  coeff_stage, a, dt);
}

int launch_rhs_eval(axs_level *L, const double *coeff3_dev, void *rhs)
{
    if (L->precision == AXS_FP64)
        return fill_and_launch<double>(L, MODE_RHS, L->y, rhs, coeff3_dev,
                                       nullptr, 0, 0.0, 0.0);
    return fill_and_launch<float>(L, MODE_RHS, L->y, rhs, coeff3_dev, nullptr,
                                  0, 0.0, 0.0);
}

int launch_step_advance(axs_level *L, int reset)
{
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, L->ctx->stream,
                       L->d_step, reset);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

} // namespace axs
