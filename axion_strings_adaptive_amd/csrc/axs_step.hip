// axs_step.hip -- one launch per RK4 step on a uniform periodic level (FP32
// and FP32C storage).
//
// Every stage input the four stages need is formed from y_n on the fly:
//   psi_1 = psi_n + (dt/2) Pi_n                   (pointwise)
//   Pi_1  = Pi_n + (dt/2) F(psi_n)                (stencil of psi_n)
//   Pi_2  = Pi_n + (dt/2) F(psi_1)                (stencil of psi_1)
//   psi_2 = psi_n + (dt/2) Pi_1,  psi_3 = psi_n + dt Pi_2
//   k3 = (Pi_2, F(psi_2)), k4 = (Pi_n + dt F(psi_2), F(psi_3)) at the cell.
// F(psi_2) and F(psi_3) need psi_2, psi_3 on the 13-point star, hence
// F(psi_n) and F(psi_1) on the star as well: each workgroup evaluates those
// on its tile grown by 2 cells in x and y (one thread per point of the grown
// tile) from an LDS plane of psi_n, psi_1 grown by 4, streams z with
// register queues, and stores only y_{n+1}: y_n read and y_{n+1} written,
// 2 S, plus 2 S for the FP32C low words = 64 B per cell and step (112 B for
// the fused pairs without an accumulator, 304 B for one launch per stage;
// S = 4 words).  Every value is formed by the expressions of the
// per-stage kernels (stage_update, stage_axpy, force, star_lap), so the bits
// are the same.  The output must not be y_n's buffer (callers alternate two
// state buffers).
#include "axs_internal.hpp"
#include "axs_stage_math.hpp"
#include "axs_vec.hpp"

namespace axs
{

struct StepParams
{
    const float *yn; // y_n (4 comps)
    float *yout;     // y_{n+1} (4 comps; not y_n's buffer)
    float *comp;     // FP32C low part of y_n (centre only, in place)
    const double *coeff;
    const int *step;
    double h;     // dt
    double h6;    // dt/6
    float h6_lo;  // dt/6 - (float)(dt/6)
    int nx, ny, nz;
    uint32_t n;
    double inv_dx2;
    int z_chunk;
    int tiles_x, tiles_y, tiles_z;
    int xcd_remap;
};

constexpr int kStepTX = 60; // interior columns per tile (64 lanes - 2 x 2)

// Block of 64 (TY + 4) threads (AL: 68 (TY + 4)): each owns one point of
// the tile grown by 2.  Loop variable q: the plane where F(psi_n), F(psi_1)
// and psi_2, psi_3 are formed; the final stage runs at plane q - 2.
template <int PREC, int TY, bool AL>
__global__ void __launch_bounds__((AL ? 68 : 64) * (TY + 4)) k_step(StepParams p)
{
    // AL = false: 64 x (TY + 4) threads, interior 60 x TY (rows of the tile
    // grown by 2 are the 64 lanes of a wave); AL = true: interior 64 x TY
    // (rows start on cache-line boundaries), the 4 x-halo columns of the
    // grown tile taken by 4 (TY + 4) more threads
    constexpr int TXI = AL ? 64 : kStepTX; // interior columns
    constexpr int EXB = TXI + 4;           // grown by 2
    constexpr int EXA = TXI + 8;           // grown by 4
    constexpr int HYA = TY + 8, HYB = TY + 4;
    __shared__ float sa[2][4][HYA][EXA]; // [buffer][psn1 psn2 ps11 ps12]
    __shared__ float sb[3][4][HYB][EXB]; // [q % 3][ps21 ps22 ps31 ps32]

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

    // the own point (ec, er) of the tile grown by 2: x = X0 - 2 + ec,
    // y = Y0 - 2 + er
    const int t = threadIdx.x;
    int ec, er;
    if (!AL || t < 64 * (TY + 4))
    {
        ec = t % 64 + (AL ? 2 : 0);
        er = t / 64;
    }
    else
    {
        const int u = t - 64 * (TY + 4);
        ec = u % 4 < 2 ? u % 4 : u % 4 + 64;
        er = u / 4;
    }
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t n = p.n;
    const uint32_t plane = (uint32_t)nx * (uint32_t)ny;
    const int X0 = bx * TXI, Y0 = by * TY;
    auto wrapx = [&](int x) { return ((x % nx) + nx) % nx; };
    auto wrapy = [&](int y) { return ((y % ny) + ny) % ny; };
    const int gx = X0 - 2 + ec, gy = Y0 - 2 + er;
    const uint32_t row = (uint32_t)wrapy(gy) * nx + wrapx(gx);
    const bool interior = ec >= 2 && ec < TXI + 2 && er >= 2 && er < TY + 2;
    const bool active = interior && gx < nx && gy < ny;
    // halo points of the plane grown by 4 (a plus shape: no corners)
    const bool hyl = er < 2, hyr = er >= TY + 2;
    const bool hxl = ec < 2, hxr = ec >= TXI + 2;
    uint32_t rowhy = 0, rowhx = 0;
    int yah = 0, xah = 0;
    if (hyl || hyr)
    {
        yah = hyl ? er : er + 4;
        rowhy = (uint32_t)wrapy(Y0 - 4 + yah) * nx + wrapx(gx);
    }
    if (hxl || hxr)
    {
        xah = hxl ? ec : ec + 4;
        rowhx = (uint32_t)wrapy(gy) * nx + wrapx(X0 - 4 + xah);
    }

    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);
    auto zoff = [&](int z) -> uint32_t {
        z = ((z % nz) + nz) % nz;
        return (uint32_t)z * plane;
    };
    const float *src[4] = {p.yn, p.yn + n, p.yn + 2 * n, p.yn + 3 * n};
    const double ah = 0.5 * p.h;
    auto psi1_of = [&](float ps, float pi) {
        return stage_axpy<float, PREC>(ps, ah, pi);
    };

    // stage scalars (c, lambda, R^2) of the four stages, read once
    const double *cfp = p.coeff + 12 * (p.step ? *p.step : 0);
    double cf[12];
#pragma unroll
    for (int i = 0; i < 12; ++i)
        cf[i] = cfp[i];
    const float i2 = (float)p.inv_dx2;

    // register queues of the own point: psi_n, psi_1, Pi_n at q-2..q+2;
    // psi_2, psi_3 at q-4..q; F(psi_n), F(psi_1) at q-2..q
    float qn[5][2], q1[5][2], qpn[5][2], q2[5][2], q3[5][2];
    float qfn[3][2], qf1[3][2];
#pragma unroll
    for (int i = 0; i < 5; ++i)
    {
        const uint32_t o = zoff(z_begin - 4 + i) + row;
#pragma unroll
        for (int c = 0; c < 2; ++c)
        {
            qn[i][c] = src[c][o];
            qpn[i][c] = src[2 + c][o];
            q1[i][c] = psi1_of(qn[i][c], qpn[i][c]);
            q2[i][c] = q3[i][c] = 0.0f;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 2; ++c)
            qfn[i][c] = qf1[i][c] = 0.0f;
    // halo words (psi_n 1, 2, Pi_n 1, 2) of plane q
    float hy[4] = {0, 0, 0, 0}, hx[4] = {0, 0, 0, 0};
    auto load_halo = [&](int z) {
        const uint32_t o = zoff(z);
        if (hyl || hyr)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                hy[c] = src[c][o + rowhy];
        }
        if (hxl || hxr)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                hx[c] = src[c][o + rowhx];
        }
    };
    load_halo(z_begin - 2);
    float lo[4] = {0, 0, 0, 0};

    for (int q = z_begin - 2; q < z_end + 2; ++q)
    {
        const int ba = (q - z_begin + 2) & 1;
        const int sbq = (q - z_begin + 2) % 3;  // slot of plane q
        const int sbz = (q - z_begin + 3) % 3;  // slot of plane q - 2
        // ---- psi_n, psi_1 of plane q into LDS (grown by 4) ----
        {
            const int ya = er + 2, xa = ec + 2;
#pragma unroll
            for (int c = 0; c < 2; ++c)
            {
                sa[ba][c][ya][xa] = qn[2][c];
                sa[ba][2 + c][ya][xa] = q1[2][c];
            }
            if (hyl || hyr)
            {
#pragma unroll
                for (int c = 0; c < 2; ++c)
                {
                    sa[ba][c][yah][xa] = hy[c];
                    sa[ba][2 + c][yah][xa] = psi1_of(hy[c], hy[2 + c]);
                }
            }
            if (hxl || hxr)
            {
#pragma unroll
                for (int c = 0; c < 2; ++c)
                {
                    sa[ba][c][ya][xah] = hx[c];
                    sa[ba][2 + c][ya][xah] = psi1_of(hx[c], hx[2 + c]);
                }
            }
        }
        // ---- prefetch: own point of plane q+3, halos of plane q+1 ----
        float nw[4];
        {
            const uint32_t o3 = zoff(q + 3) + row;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                nw[c] = src[c][o3];
        }
        load_halo(q + 1);
        const int z = q - 2; // plane of the final stage
        if (PREC == AXS_FP32C && active && z >= z_begin)
        {
            const uint32_t oz = (uint32_t)z * plane + row;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                lo[c] = p.comp[c * n + oz];
        }

        __syncthreads();

        // ---- F(psi_n), F(psi_1) at the own point of plane q; Pi_1, Pi_2,
        //      psi_2, psi_3 there ----
        {
            const int ya = er + 2, xa = ec + 2;
            // the Laplacian of field k of the LDS plane, z neighbours zq
            auto lap_a = [&](int k, const float(&zq)[5][2], int c) {
                const StarF s{sa[ba][k][ya][xa - 2], sa[ba][k][ya][xa - 1],
                              sa[ba][k][ya][xa + 1], sa[ba][k][ya][xa + 2],
                              sa[ba][k][ya - 2][xa], sa[ba][k][ya - 1][xa],
                              sa[ba][k][ya + 1][xa], sa[ba][k][ya + 2][xa],
                              zq[0][c], zq[1][c], zq[3][c], zq[4][c], zq[2][c]};
                return star_lap(s, i2);
            };
            float F[2][2];
            force(lap_a(0, qn, 0), lap_a(1, qn, 1), qn[2][0], qn[2][1], cf[0],
                  cf[1], cf[2], F[0][0], F[0][1]);
            force(lap_a(2, q1, 0), lap_a(3, q1, 1), q1[2][0], q1[2][1], cf[3],
                  cf[4], cf[5], F[1][0], F[1][1]);
            float p1[2], p2[2], s2[2], s3[2];
#pragma unroll
            for (int c = 0; c < 2; ++c)
            {
                p1[c] = stage_axpy<float, PREC>(qpn[2][c], ah, F[0][c]);
                p2[c] = stage_axpy<float, PREC>(qpn[2][c], ah, F[1][c]);
                s2[c] = stage_axpy<float, PREC>(qn[2][c], ah, p1[c]);
                s3[c] = stage_axpy<float, PREC>(qn[2][c], p.h, p2[c]);
                sb[sbq][c][er][ec] = s2[c];
                sb[sbq][2 + c][er][ec] = s3[c];
            }
            // queues of plane-q values (index 2 = plane q)
#pragma unroll
            for (int c = 0; c < 2; ++c)
            {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                {
                    q2[i][c] = q2[i + 1][c];
                    q3[i][c] = q3[i + 1][c];
                }
                q2[4][c] = s2[c];
                q3[4][c] = s3[c];
#pragma unroll
                for (int i = 0; i < 2; ++i)
                {
                    qfn[i][c] = qfn[i + 1][c];
                    qf1[i][c] = qf1[i + 1][c];
                }
                qfn[2][c] = F[0][c];
                qf1[2][c] = F[1][c];
            }
        }

        // ---- stages 3, 4 and the update at plane z = q - 2 ----
        // (psi_2, psi_3 of plane z sit in slot sbz, written two planes
        // ago; the z neighbours in q2 / q3 at planes z-2..z+2)
        if (z >= z_begin && active)
        {
            auto lap_b = [&](int k, const float(&zq)[5][2], int c) {
                const StarF s{sb[sbz][k][er][ec - 2], sb[sbz][k][er][ec - 1],
                              sb[sbz][k][er][ec + 1], sb[sbz][k][er][ec + 2],
                              sb[sbz][k][er - 2][ec], sb[sbz][k][er - 1][ec],
                              sb[sbz][k][er + 1][ec], sb[sbz][k][er + 2][ec],
                              zq[0][c], zq[1][c], zq[3][c], zq[4][c], zq[2][c]};
                return star_lap(s, i2);
            };
            float F2[2], F3[2];
            force(lap_b(0, q2, 0), lap_b(1, q2, 1), q2[2][0], q2[2][1], cf[6],
                  cf[7], cf[8], F2[0], F2[1]);
            force(lap_b(2, q3, 0), lap_b(3, q3, 1), q3[2][0], q3[2][1], cf[9],
                  cf[10], cf[11], F3[0], F3[1]);
            // y_n at plane z: psi_n = qn[0], Pi_n = qpn[0]
            const float ynv[4] = {qn[0][0], qn[0][1], qpn[0][0], qpn[0][1]};
            float none[4] = {0, 0, 0, 0};
            float lov[4] = {lo[0], lo[1], lo[2], lo[3]};
            const float k1[4] = {qpn[0][0], qpn[0][1], qfn[0][0], qfn[0][1]};
            float acc1[4], y1[4];
            stage_update<float, PREC, MODE_S1>(k1, ynv, none, lov, ah, p.h6,
                                               p.h6_lo, acc1, y1);
            const float k2[4] = {y1[2], y1[3], qf1[0][0], qf1[0][1]};
            float acc2[4], y2[4];
            stage_update<float, PREC, MODE_S23>(k2, ynv, acc1, lov, ah, p.h6,
                                                p.h6_lo, acc2, y2);
            const float k3[4] = {y2[2], y2[3], F2[0], F2[1]};
            float acc3[4], y3[4];
            stage_update<float, PREC, MODE_S23>(k3, ynv, acc2, lov, p.h, p.h6,
                                                p.h6_lo, acc3, y3);
            const float k4[4] = {y3[2], y3[3], F3[0], F3[1]};
            float unused[4], y4[4];
            stage_update<float, PREC, MODE_S4>(k4, ynv, acc3, lov, p.h, p.h6,
                                               p.h6_lo, unused, y4);
            const uint32_t oz = (uint32_t)z * plane + row;
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                p.yout[c * n + oz] = y4[c];
                if (PREC == AXS_FP32C)
                    p.comp[c * n + oz] = lov[c];
            }
        }

        // ---- advance the psi_n / psi_1 / Pi_n queues to plane q+1 ----
#pragma unroll
        for (int c = 0; c < 2; ++c)
        {
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
                qn[i][c] = qn[i + 1][c];
                q1[i][c] = q1[i + 1][c];
                qpn[i][c] = qpn[i + 1][c];
            }
            qn[4][c] = nw[c];
            qpn[4][c] = nw[2 + c];
            q1[4][c] = psi1_of(nw[c], nw[2 + c]);
        }
    }
}

template <int TY, bool AL>
static void launch_step_ty(const StepParams &p, int prec, long ntiles,
                           hipStream_t s)
{
    const dim3 block((AL ? 68 : 64) * (TY + 4));
    if (prec == AXS_FP32C)
        hipLaunchKernelGGL((k_step<AXS_FP32C, TY, AL>), dim3(ntiles), block, 0,
                           s, p);
    else
        hipLaunchKernelGGL((k_step<AXS_FP32, TY, AL>), dim3(ntiles), block, 0,
                           s, p);
}

// One RK4 step in one launch: y_n = yin -> y_{n+1} = yout (yout != yin),
// FP32C lo in place.  Interior tile 60 x TY: the level's pair variant
// selects the tile: -1 (default) and 3: 64 x 11 with rows aligned to cache
// lines (1020 threads, 4 waves per SIMD; z chunk 128), the 1024^3 sweep's
// fastest; 2: 60 x 12 (1024 threads); 0: 60 x 8; 1: 60 x 4.
int launch_rk_step(axs_level *L, const void *yin, void *yout,
                   const double *coeff_table, const int *step_ptr, double dt)
{
    if (L->precision == AXS_FP64)
        return set_error(AXS_ERR_UNSUPPORTED,
                         "one launch per RK4 step is FP32 / FP32C only");
    // default (-1) and variant 3: 64-column interior (rows aligned to cache
    // lines), 11 rows, 1020 threads -- the 1024^3 sweep's fastest
    const bool al = L->pair_variant == 3 || L->pair_variant < 0;
    const int TY =
        al ? 11 : (L->pair_variant == 1 ? 4 : (L->pair_variant == 0 ? 8 : 12));
    const int txi = al ? 64 : kStepTX;
    StepParams p{};
    p.yn = static_cast<const float *>(yin);
    p.yout = static_cast<float *>(yout);
    p.comp = L->comp;
    p.coeff = coeff_table;
    p.step = step_ptr;
    p.h = dt;
    p.h6 = dt / 6.0;
    p.h6_lo = (float)(p.h6 - (double)(float)p.h6);
    p.nx = L->g.nx;
    p.ny = L->g.ny;
    p.nz = L->g.nz;
    p.n = L->g.ncell;
    p.inv_dx2 = 1.0 / (L->dx * L->dx);
    p.tiles_x = (p.nx + txi - 1) / txi;
    p.tiles_y = (p.ny + TY - 1) / TY;
    int zc = L->pair_z_chunk > 0 ? L->pair_z_chunk : (al ? 128 : 64);
    while (zc > 4 && (long)p.tiles_x * p.tiles_y * ((p.nz + zc - 1) / zc) < 4096)
        zc >>= 1;
    if (zc > p.nz)
        zc = p.nz;
    p.z_chunk = zc;
    p.tiles_z = (p.nz + zc - 1) / zc;
    const long ntiles = (long)p.tiles_x * p.tiles_y * p.tiles_z;
    p.xcd_remap = (ntiles % 8 == 0) ? 1 : 0;
    if (al)
        launch_step_ty<11, true>(p, L->precision, ntiles, L->ctx->stream);
    else if (TY == 4)
        launch_step_ty<4, false>(p, L->precision, ntiles, L->ctx->stream);
    else if (TY == 12)
        launch_step_ty<12, false>(p, L->precision, ntiles, L->ctx->stream);
    else
        launch_step_ty<8, false>(p, L->precision, ntiles, L->ctx->stream);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

} // namespace axs
