// axs_rk.hip -- fused RHS + RK4-stage kernels for a uniform periodic level.
//
// One launch per RK stage.  Each workgroup owns a 64 x TY column tile of the
// x-y plane and streams through a chunk of z planes:
//   * psi1/psi2 of the stage input are read once per cell per stage: the z
//     neighbours (radius 2) live in a 5-deep register queue per thread, the
//     x/y neighbours come from a double-buffered LDS plane with a 2-cell halo
//     (one barrier per plane); periodic ghosts are index wrapping;
//   * the pointwise RHS (AxionStringsRHS.impl.hpp:83-118) and the RK4 stage
//     axpy are applied in registers (axs_stage_math.hpp) and the stage
//     outputs written once.
// Byte traffic per cell and stage (4 components of S bytes each):
//   stage 1: read y_n, write acc, write y_1                     = 3 S
//   stage 2,3: read y_s, y_n, acc, write acc, y_{s+1}             = 5 S
//   stage 4: read y_3, y_n, acc, write y_{n+1} (+ read/write lo)  = 4 S (6 S)
// = 17 S per RK4 step in FP64 (544 B/cell), 19 S in FP32C (304 B/cell).
// Tiles are mapped XCD-aware: consecutive tiles (sharing halo rows) are
// dispatched to the same XCD and its L2.
#include "axs_internal.hpp"
#include "axs_stage_math.hpp"

namespace axs
{

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

template <typename T> struct StarOf;
template <> struct StarOf<double>
{
    using type = Star;
};
template <> struct StarOf<float>
{
    using type = StarF;
};

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
    const double cc = cf[0], lam = cf[1], R2 = cf[2];
    const T i2 = (T)p.inv_dx2;

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
        if (z + 1 < z_end)
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
        T pi[2] = {0, 0};
        T yn[4] = {0, 0, 0, 0}, ac[4] = {0, 0, 0, 0};
        float lo[4] = {0, 0, 0, 0};
        if (active)
        {
            pi[0] = p.yin[2 * n + oc];
            pi[1] = p.yin[3 * n + oc];
            if (MODE == MODE_S23 || MODE == MODE_S4)
            {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                {
                    yn[c] = p.yn[c * n + oc];
                    ac[c] = p.acc[c * n + oc];
                }
            }
            if (MODE == MODE_S4 && PREC == AXS_FP32C)
            {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    lo[c] = p.comp[c * n + oc];
            }
        }

        __syncthreads();

        if (active)
        {
            const int cy = ty + 2, cx = tx + 2;
            using S = typename StarOf<T>::type;
            const S s1{sm[b][0][cy][cx - 2], sm[b][0][cy][cx - 1],
                       sm[b][0][cy][cx + 1], sm[b][0][cy][cx + 2],
                       sm[b][0][cy - 2][cx], sm[b][0][cy - 1][cx],
                       sm[b][0][cy + 1][cx], sm[b][0][cy + 2][cx],
                       q1[0], q1[1], q1[3], q1[4], q1[2]};
            const S s2{sm[b][1][cy][cx - 2], sm[b][1][cy][cx - 1],
                       sm[b][1][cy][cx + 1], sm[b][1][cy][cx + 2],
                       sm[b][1][cy - 2][cx], sm[b][1][cy - 1][cx],
                       sm[b][1][cy + 1][cx], sm[b][1][cy + 2][cx],
                       q2[0], q2[1], q2[3], q2[4], q2[2]};
            T F1, F2;
            force(star_lap(s1, i2), star_lap(s2, i2), q1[2], q2[2], cc, lam,
                  R2, F1, F2);
            const T k[4] = {pi[0], pi[1], F1, F2};
            if (MODE == MODE_S1)
            {
                // y_n is the stage input itself
                yn[0] = q1[2];
                yn[1] = q2[2];
                yn[2] = pi[0];
                yn[3] = pi[1];
            }
            T acc_out[4], y_out[4];
            stage_update<T, PREC, MODE>(k, yn, ac, lo, p.a, p.h6, p.h6_lo,
                                        acc_out, y_out);
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                p.yout[c * n + oc] = y_out[c];
                if (MODE == MODE_S1 || MODE == MODE_S23)
                    p.acc[c * n + oc] = acc_out[c];
                if (MODE == MODE_S4 && PREC == AXS_FP32C)
                    p.comp[c * n + oc] = lo[c];
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
