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
#include "axs_vec.hpp"

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
    int nontemporal; // stream the stage outputs past the caches
};

// Each workgroup: a (64 VX) x TY column tile, streaming a chunk of z.
// PF: the centre loads (Pi, y_n, acc, lo) of plane z+1 are issued before
// plane z is computed, so they are in flight for a whole plane.
template <typename T, int PREC, int MODE, int TY, int VX, bool PF>
__global__ void __launch_bounds__(64 * TY)
    k_stage(StageParams<T> p)
{
    constexpr int TX = 64 * VX; // tile width in cells
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
    const int gx = bx * TX + tx * VX; // first of this lane's VX cells
    const int gy = by * TY + ty;
    const bool active = gx < nx && gy < ny; // nx % VX == 0
    const int xw = gx % nx;
    const int yw = gy % ny;
    const uint32_t rowc = (uint32_t)yw * nx + xw;
    const bool nt = p.nontemporal != 0;

    // halo roles: VX == 1 -> lanes 0,1 / 62,63 one cell each;
    // VX >= 2 -> lane 0 / lane 63 one aligned pair each
    constexpr int HL = VX == 1 ? 2 : 1;
    const bool hxl = tx < HL, hxr = tx >= 64 - HL;
    const bool hyl = ty < 2, hyr = ty >= TY - 2;
    uint32_t rowhx = 0, rowhy = 0;
    if (hxl || hxr)
    {
        int hx;
        if constexpr (VX == 1)
            hx = hxl ? (bx * TX - 2 + tx) : (bx * TX + TX + (tx - 62));
        else
            hx = hxl ? bx * TX - 2 : bx * TX + TX;
        hx = ((hx % nx) + nx) % nx;
        rowhx = (uint32_t)yw * nx + hx;
    }
    if (hyl || hyr)
    {
        int hy = hyl ? (by * TY - 2 + ty) : (by * TY + TY + (ty - (TY - 2)));
        hy = ((hy % ny) + ny) % ny;
        rowhy = (uint32_t)hy * nx + xw;
    }
    constexpr int HW = VX == 1 ? 1 : 2; // halo words per x-halo lane

    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);

    const T *psi1 = p.yin;
    const T *psi2 = p.yin + n;

    auto zoff = [&](int z) -> uint32_t {
        z = wrap_lo(wrap_hi(z, nz), nz);
        return (uint32_t)z * plane;
    };

    // register queues: psi at z-2 .. z+2
    T q1[5][VX], q2[5][VX];
#pragma unroll
    for (int i = 0; i < 5; ++i)
    {
        const uint32_t o = zoff(z_begin - 2 + i) + rowc;
        vload<T, VX>(psi1 + o, q1[i]);
        vload<T, VX>(psi2 + o, q2[i]);
    }
    T hx1[HW], hx2[HW], hy1[VX], hy2[VX];
    {
        const uint32_t o = zoff(z_begin);
        if (hxl || hxr)
        {
            vload<T, HW>(psi1 + o + rowhx, hx1);
            vload<T, HW>(psi2 + o + rowhx, hx2);
        }
        if (hyl || hyr)
        {
            vload<T, VX>(psi1 + o + rowhy, hy1);
            vload<T, VX>(psi2 + o + rowhy, hy2);
        }
    }

    // stage scalars
    const double *cf =
        p.coeff + 12 * (p.step ? *p.step : 0) + 3 * p.coeff_stage;
    const double cc = cf[0], lam = cf[1], R2 = cf[2];
    const T i2 = (T)p.inv_dx2;

    // centre values of the current plane (and, with PF, of the next one)
    T pi[2][VX], yn[4][VX], ac[4][VX];
    float lo[4][VX];
    T pin[2][VX], ynn[4][VX], acn[4][VX];
    float lon[4][VX];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < VX; ++v)
        {
            yn[c][v] = ynn[c][v] = 0;
            ac[c][v] = acn[c][v] = 0;
            lo[c][v] = lon[c][v] = 0.0f;
            if (c < 2)
                pi[c][v] = pin[c][v] = 0;
        }
    auto load_centre = [&](int zz, T(&pi_)[2][VX], T(&yn_)[4][VX],
                           T(&ac_)[4][VX], float(&lo_)[4][VX]) {
        if (!active)
            return;
        const uint32_t o = (uint32_t)zz * plane + rowc;
        vload<T, VX>(p.yin + 2 * n + o, pi_[0]);
        vload<T, VX>(p.yin + 3 * n + o, pi_[1]);
        if constexpr (MODE == MODE_S23 || MODE == MODE_S4)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                vload<T, VX>(p.yn + c * n + o, yn_[c]);
                vload<T, VX>(p.acc + c * n + o, ac_[c]);
            }
        }
        if constexpr (MODE == MODE_S4 && PREC == AXS_FP32C)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                vload<float, VX>(p.comp + c * n + o, lo_[c]);
        }
    };
    if constexpr (PF)
        load_centre(z_begin, pi, yn, ac, lo);

    for (int z = z_begin; z < z_end; ++z)
    {
        const int b = (z - z_begin) & 1;
        const int cx0 = tx * VX + 2;
        // ---- stage the plane into LDS ----
#pragma unroll
        for (int v = 0; v < VX; ++v)
        {
            sm[b][0][ty + 2][cx0 + v] = q1[2][v];
            sm[b][1][ty + 2][cx0 + v] = q2[2][v];
        }
        if constexpr (VX == 1)
        {
            if (hxl)
            {
                sm[b][0][ty + 2][tx] = hx1[0];
                sm[b][1][ty + 2][tx] = hx2[0];
            }
            else if (hxr)
            {
                sm[b][0][ty + 2][tx + 4] = hx1[0];
                sm[b][1][ty + 2][tx + 4] = hx2[0];
            }
        }
        else
        {
            if (hxl)
            {
                sm[b][0][ty + 2][0] = hx1[0];
                sm[b][0][ty + 2][1] = hx1[1];
                sm[b][1][ty + 2][0] = hx2[0];
                sm[b][1][ty + 2][1] = hx2[1];
            }
            else if (hxr)
            {
                sm[b][0][ty + 2][TX + 2] = hx1[0];
                sm[b][0][ty + 2][TX + 3] = hx1[1];
                sm[b][1][ty + 2][TX + 2] = hx2[0];
                sm[b][1][ty + 2][TX + 3] = hx2[1];
            }
        }
        if (hyl)
        {
#pragma unroll
            for (int v = 0; v < VX; ++v)
            {
                sm[b][0][ty][cx0 + v] = hy1[v];
                sm[b][1][ty][cx0 + v] = hy2[v];
            }
        }
        else if (hyr)
        {
#pragma unroll
            for (int v = 0; v < VX; ++v)
            {
                sm[b][0][ty + 4][cx0 + v] = hy1[v];
                sm[b][1][ty + 4][cx0 + v] = hy2[v];
            }
        }

        // ---- prefetch next plane (queue head + halos) ----
        T nq1[VX], nq2[VX];
#pragma unroll
        for (int v = 0; v < VX; ++v)
        {
            nq1[v] = q1[4][v];
            nq2[v] = q2[4][v];
        }
        if (z + 1 < z_end)
        {
            const uint32_t o3 = zoff(z + 3) + rowc;
            vload<T, VX>(psi1 + o3, nq1);
            vload<T, VX>(psi2 + o3, nq2);
            const uint32_t o1 = zoff(z + 1);
            if (hxl || hxr)
            {
                vload<T, HW>(psi1 + o1 + rowhx, hx1);
                vload<T, HW>(psi2 + o1 + rowhx, hx2);
            }
            if (hyl || hyr)
            {
                vload<T, VX>(psi1 + o1 + rowhy, hy1);
                vload<T, VX>(psi2 + o1 + rowhy, hy2);
            }
        }

        // ---- centre loads: this plane, or (PF) the next one ----
        const uint32_t oc = (uint32_t)z * plane + rowc;
        if constexpr (PF)
        {
            if (z + 1 < z_end)
                load_centre(z + 1, pin, ynn, acn, lon);
        }
        else
            load_centre(z, pi, yn, ac, lo);

        __syncthreads();

        if (active)
        {
            const int cy = ty + 2;
            T out_y[4][VX], out_a[4][VX];
#pragma unroll
            for (int v = 0; v < VX; ++v)
            {
                const int cx = cx0 + v;
                using S = typename StarOf<T>::type;
                const S s1{sm[b][0][cy][cx - 2], sm[b][0][cy][cx - 1],
                           sm[b][0][cy][cx + 1], sm[b][0][cy][cx + 2],
                           sm[b][0][cy - 2][cx], sm[b][0][cy - 1][cx],
                           sm[b][0][cy + 1][cx], sm[b][0][cy + 2][cx],
                           q1[0][v], q1[1][v], q1[3][v], q1[4][v], q1[2][v]};
                const S s2{sm[b][1][cy][cx - 2], sm[b][1][cy][cx - 1],
                           sm[b][1][cy][cx + 1], sm[b][1][cy][cx + 2],
                           sm[b][1][cy - 2][cx], sm[b][1][cy - 1][cx],
                           sm[b][1][cy + 1][cx], sm[b][1][cy + 2][cx],
                           q2[0][v], q2[1][v], q2[3][v], q2[4][v], q2[2][v]};
                T F1, F2;
                force(star_lap(s1, i2), star_lap(s2, i2), q1[2][v], q2[2][v],
                      cc, lam, R2, F1, F2);
                const T k[4] = {pi[0][v], pi[1][v], F1, F2};
                T ynv[4], acv[4];
                float lov[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                {
                    ynv[c] = yn[c][v];
                    acv[c] = ac[c][v];
                    lov[c] = lo[c][v];
                }
                if (MODE == MODE_S1)
                {
                    // y_n is the stage input itself
                    ynv[0] = q1[2][v];
                    ynv[1] = q2[2][v];
                    ynv[2] = pi[0][v];
                    ynv[3] = pi[1][v];
                }
                T acc_out[4], y_out[4];
                stage_update<T, PREC, MODE>(k, ynv, acv, lov, p.a, p.h6,
                                            p.h6_lo, acc_out, y_out);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                {
                    out_y[c][v] = y_out[c];
                    out_a[c][v] = acc_out[c];
                    lo[c][v] = lov[c];
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                vstore<T, VX>(p.yout + c * n + oc, out_y[c], nt);
                if (MODE == MODE_S1 || MODE == MODE_S23)
                    vstore<T, VX>(p.acc + c * n + oc, out_a[c], nt);
                if (MODE == MODE_S4 && PREC == AXS_FP32C)
                    vstore<float, VX>(p.comp + c * n + oc, lo[c], nt);
            }
        }

        if constexpr (PF)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int v = 0; v < VX; ++v)
                {
                    yn[c][v] = ynn[c][v];
                    ac[c][v] = acn[c][v];
                    lo[c][v] = lon[c][v];
                    if (c < 2)
                        pi[c][v] = pin[c][v];
                }
        }

        // ---- advance the queue ----
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < VX; ++v)
            {
                q1[i][v] = q1[i + 1][v];
                q2[i][v] = q2[i + 1][v];
            }
#pragma unroll
        for (int v = 0; v < VX; ++v)
        {
            q1[4][v] = nq1[v];
            q2[4][v] = nq2[v];
        }
    }
}

__global__ void k_step_advance(int *step, int reset)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *step = reset ? 0 : *step + 1;
}

// ------------------------------------------------------------ launchers --
// Tile variants: (TY, VX) pairs.  variant = base + kBaseVariants * g with
// g = 0 plain, 1 non-temporal stores of the stage outputs, 2 centre-load
// prefetch (PF), 3 both.
struct TileShape
{
    int ty, vx;
};
static const TileShape kTiles[] = {{4, 1}, {8, 1}, {16, 1}, {4, 2},
                                   {8, 2}, {4, 4}, {8, 4}};
constexpr int kBaseVariants = 7;

int num_tile_variants() { return 4 * kBaseVariants; }

template <typename T, int PREC, int MODE, int TY, int VX>
static int launch_t(const StageParams<T> &p, int ntiles, hipStream_t s,
                    bool pf)
{
    if (pf)
        hipLaunchKernelGGL((k_stage<T, PREC, MODE, TY, VX, true>),
                           dim3(ntiles), dim3(64, TY), 0, s, p);
    else
        hipLaunchKernelGGL((k_stage<T, PREC, MODE, TY, VX, false>),
                           dim3(ntiles), dim3(64, TY), 0, s, p);
    return AXS_OK;
}

template <typename T, int PREC, int MODE>
static int launch_ty(int base, const StageParams<T> &p, int ntiles,
                     hipStream_t s, bool pf)
{
    switch (base)
    {
    case 0:
        return launch_t<T, PREC, MODE, 4, 1>(p, ntiles, s, pf);
    case 1:
        return launch_t<T, PREC, MODE, 8, 1>(p, ntiles, s, pf);
    case 2:
        return launch_t<T, PREC, MODE, 16, 1>(p, ntiles, s, pf);
    case 3:
        return launch_t<T, PREC, MODE, 4, 2>(p, ntiles, s, pf);
    case 4:
        return launch_t<T, PREC, MODE, 8, 2>(p, ntiles, s, pf);
    default:
        // FP64 x4 tiles need more than 256 VGPRs (spills): run x2 instead
        if constexpr (sizeof(T) == 8)
            return launch_t<T, PREC, MODE, 8, 2>(p, ntiles, s, pf);
        else if (base == 5)
            return launch_t<T, PREC, MODE, 4, 4>(p, ntiles, s, pf);
        else
            return launch_t<T, PREC, MODE, 8, 4>(p, ntiles, s, pf);
    }
}

template <typename T, int PREC>
static int launch_mode(int mode, int base, const StageParams<T> &p, int ntiles,
                       hipStream_t s, bool pf)
{
    switch (mode)
    {
    case MODE_RHS:
        return launch_ty<T, PREC, MODE_RHS>(base, p, ntiles, s, pf);
    case MODE_S1:
        return launch_ty<T, PREC, MODE_S1>(base, p, ntiles, s, pf);
    case MODE_S23:
        return launch_ty<T, PREC, MODE_S23>(base, p, ntiles, s, pf);
    default:
        return launch_ty<T, PREC, MODE_S4>(base, p, ntiles, s, pf);
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
    // default: the 1024^3 autotuner choices (8-row tiles with centre-load
    // prefetch; 4 cells per lane in FP32, 2 in FP64)
    const int tuned = sizeof(T) == 8 ? 2 * kBaseVariants + 4
                                     : 2 * kBaseVariants + 6;
    int variant = (L->tile_variant >= 0 && L->tile_variant < num_tile_variants())
                      ? L->tile_variant
                      : tuned;
    const int group = variant / kBaseVariants;
    p.nontemporal = (group & 1) ? 1 : 0;
    const bool pf = (group & 2) != 0;
    int base = variant % kBaseVariants;
    if (sizeof(T) == 8 && kTiles[base].vx == 4)
        base = 4; // FP64 x4 tiles need more than 256 VGPRs (spills)
    // vector variants need whole vectors per row
    if (p.nx % kTiles[base].vx != 0)
        base = 1;
    const int TY = kTiles[base].ty, VX = kTiles[base].vx;
    p.tiles_x = (p.nx + 64 * VX - 1) / (64 * VX);
    p.tiles_y = (p.ny + TY - 1) / TY;
    int zc = L->z_chunk;
    if (zc <= 0)
    {
        zc = 32;
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
        rc = launch_mode<T, AXS_FP64>(mode, base, p, (int)ntiles,
                                      L->ctx->stream, pf);
    else if (L->precision == AXS_FP32C)
        rc = launch_mode<T, AXS_FP32C>(mode, base, p, (int)ntiles,
                                       L->ctx->stream, pf);
    else
        rc = launch_mode<T, AXS_FP32>(mode, base, p, (int)ntiles,
                                      L->ctx->stream, pf);
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
