// axs_pair.hip -- RK4 stages fused in pairs on a uniform periodic level:
// stages 1 + 2 in one launch, stages 3 + 4 in another.
//
// The second stage of a pair reads its input psi on the 13-point star, and
// that psi is pointwise in what the first stage reads:
//   stage-2 input psi = psi_n + (dt/2) Pi_n     (the psi part of y_1)
//   stage-4 input psi = psi_n + dt Pi_2         (the psi part of y_3)
// while the second stage needs its input Pi and the accumulator at the
// centre only.  So one z-streaming pass with the same radius-2 halo computes
// both stages: it stages two star fields per plane (A, the first stage's
// psi; B, the second stage's psi formed at every stencil point) and never
// stores y_1 or y_3.  Bytes per cell and RK4 step (S = 4 components of one
// word):
//   pair 1: read y_n, write acc, write y_2                        = 3 S
//   pair 2: read y_2, y_n, acc, write y_{n+1} (+ lo read/write)    = 4 S (6 S)
// = 7 S in FP64 (224 B/cell) and 9 S in FP32C (144 B/cell), against 17 S and
// 19 S with one launch per stage.  Every value is formed by the expressions
// of the per-stage kernels (stage_update, stage_axpy, force, star_lap), so
// the two paths give the same bits.
//
// Pair 2 reads psi_n on the star while other workgroups store y_{n+1}, so its
// output must not be y_n's buffer: callers alternate two state buffers.
//
// Modes 3 and 4 (accumulator-free pairs): mode 3 is pair 1 without the
// accumulator store; mode 4 is pair 2 reading Pi_n on the star as well and
// recomputing acc = k1 + 2 k2 at the centre (k1 from the star of psi_n,
// k2 from that of psi_1 = psi_n + (dt/2) Pi_n, formed at every stencil point)
// instead of loading it: 2 S + 3 S per step (FP64, FP32; FP32C + 2 S), with
// the same expressions and so the same bits.
#include "axs_internal.hpp"
#include "axs_stage_math.hpp"
#include "axs_vec.hpp"

#include <type_traits>

namespace axs
{

template <typename T> struct PairParams
{
    const T *yin;   // first stage's input: y_n (pair 1) or y_2 (pair 2)
    const T *yn;    // y_n (pair 2)
    T *acc;         // accumulator: written by pair 1, read by pair 2
    T *yout;        // y_2 (pair 1) or y_{n+1} (pair 2; not y_n's buffer)
    float *comp;    // FP32C low part of y_n (pair 2; centre only, in place)
    const double *coeff; // stage-scalar table [step][4][3]
    const int *step;     // device step index or nullptr
    int coeff_stage0;    // the first stage's row: 0 (pairs 1, 3) or 2 (2, 4)
    double a0;           // first stage's axpy factor: dt/2 (pairs 1, 3), dt (2, 4)
    double a1;           // dt/2 (stages 1 and 2)
    double h6;           // dt/6
    float h6_lo;         // dt/6 - (float)(dt/6)
    int nx, ny, nz;
    uint32_t n;
    double inv_dx2;
    int z_chunk;
    int tiles_x, tiles_y, tiles_z;
    int xcd_remap;
    int nontemporal;
    int y_major; // tile ids run along y first (tiles sharing y-halo rows
                 // run together on one XCD)
};

// Each workgroup: a (64 VX) x TY column tile, streaming a chunk of z planes.
// Register queues per lane: A and B at z-2..z+2, the first stage's input Pi
// (and psi_n for pair 2) at z..z+2.  PF: pair 2's centre loads (Pi_n, acc,
// lo) of plane z+1 are issued before plane z is computed.
template <typename T, int PREC, int PAIR, int TY, int VX, bool PF>
__global__ void __launch_bounds__(64 * TY) k_pair(PairParams<T> p)
{
    constexpr int TX = 64 * VX;
    constexpr int HX = TX + 4;
    constexpr int HY = TY + 4;
    // words loaded per stencil point and star fields per plane
    constexpr int NB = PAIR == 2 ? 6 : (PAIR == 4 ? 8 : 4);
    constexpr int NF = PAIR == 4 ? 8 : 4;
    constexpr bool FIRST = PAIR == 1 || PAIR == 3; // stages 1 + 2
    __shared__ T sm[2][NF][HY][HX]; // [buffer][A1 A2 B1 B2 (C1 C2 D1 D2)][y][x]

    int bid = blockIdx.x;
    if (p.xcd_remap)
    {
        const int ntiles = p.tiles_x * p.tiles_y * p.tiles_z;
        bid = (bid & 7) * (ntiles >> 3) + (bid >> 3);
    }
    int bx, by, bz;
    if (p.y_major)
    {
        by = bid % p.tiles_y;
        const int rem = bid / p.tiles_y;
        bx = rem % p.tiles_x;
        bz = rem / p.tiles_x;
    }
    else
    {
        bx = bid % p.tiles_x;
        const int rem = bid / p.tiles_x;
        by = rem % p.tiles_y;
        bz = rem / p.tiles_y;
    }

    const int tx = threadIdx.x;
    const int ty = threadIdx.y;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t n = p.n;
    const uint32_t plane = (uint32_t)nx * (uint32_t)ny;
    const int gx = bx * TX + tx * VX;
    const int gy = by * TY + ty;
    const bool active = gx < nx && gy < ny; // nx % VX == 0
    const int xw = gx % nx;
    const int yw = gy % ny;
    const uint32_t rowc = (uint32_t)yw * nx + xw;
    const bool nt = p.nontemporal != 0;

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
    constexpr int HW = VX == 1 ? 1 : 2;

    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);

    // the words of one stencil point: psi1, psi2, Pi1, Pi2 of the first
    // stage's input, then (pairs 2, 4) psi_n 1, 2, then (pair 4) Pi_n 1, 2
    const T *src[8] = {p.yin,         p.yin + n,     p.yin + 2 * n,
                       p.yin + 3 * n, p.yn,          p.yn + n,
                       p.yn + 2 * n,  p.yn + 3 * n};
    auto zoff = [&](int z) -> uint32_t {
        z = wrap_lo(wrap_hi(z, nz), nz);
        return (uint32_t)z * plane;
    };
    // B (the second stage's psi) at a point from its loaded words
    auto form_b = [&](const T *w, int c) -> T {
        const T base = FIRST ? w[c] : w[4 + c];
        return stage_axpy<T, PREC>(base, p.a0, w[2 + c]);
    };
    // D (pair 4: stage 2's psi = psi_n + (dt/2) Pi_n) at a point
    auto form_d = [&](const T *w, int c) -> T {
        return stage_axpy<T, PREC>(w[4 + c], p.a1, w[6 + c]);
    };

    T qa1[5][VX], qa2[5][VX], qb1[5][VX], qb2[5][VX];
    T qp1[3][VX], qp2[3][VX], qn1[3][VX], qn2[3][VX];
    // pair 4: C = psi_n and D = psi_1 at z-2..z+2, Pi_n at z..z+2
    T qc1[5][VX], qc2[5][VX], qd1[5][VX], qd2[5][VX];
    T qm1[3][VX], qm2[3][VX];
#pragma unroll
    for (int i = 0; i < 5; ++i)
    {
        const uint32_t o = zoff(z_begin - 2 + i) + rowc;
        T w[8][VX];
#pragma unroll
        for (int c = 0; c < NB; ++c)
            vload<T, VX>(src[c] + o, w[c]);
#pragma unroll
        for (int v = 0; v < VX; ++v)
        {
            T pt[8];
#pragma unroll
            for (int c = 0; c < NB; ++c)
                pt[c] = w[c][v];
            qa1[i][v] = pt[0];
            qa2[i][v] = pt[1];
            qb1[i][v] = form_b(pt, 0);
            qb2[i][v] = form_b(pt, 1);
            if constexpr (PAIR == 4)
            {
                qc1[i][v] = pt[4];
                qc2[i][v] = pt[5];
                qd1[i][v] = form_d(pt, 0);
                qd2[i][v] = form_d(pt, 1);
            }
            else
            {
                qc1[i][v] = qc2[i][v] = qd1[i][v] = qd2[i][v] = 0;
            }
            if (i >= 2)
            {
                qp1[i - 2][v] = pt[2];
                qp2[i - 2][v] = pt[3];
                if constexpr (PAIR == 4)
                {
                    qm1[i - 2][v] = pt[6];
                    qm2[i - 2][v] = pt[7];
                }
                else
                {
                    qm1[i - 2][v] = qm2[i - 2][v] = 0;
                }
                if constexpr (PAIR == 2)
                {
                    qn1[i - 2][v] = pt[4];
                    qn2[i - 2][v] = pt[5];
                }
                else
                {
                    qn1[i - 2][v] = 0;
                    qn2[i - 2][v] = 0;
                }
            }
        }
    }
    // halo words of the current plane
    T hx[8][HW], hy[8][VX];
    auto load_halo = [&](int z) {
        const uint32_t o = zoff(z);
        if (hxl || hxr)
        {
#pragma unroll
            for (int c = 0; c < NB; ++c)
                vload<T, HW>(src[c] + o + rowhx, hx[c]);
        }
        if (hyl || hyr)
        {
#pragma unroll
            for (int c = 0; c < NB; ++c)
                vload<T, VX>(src[c] + o + rowhy, hy[c]);
        }
    };
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
#pragma unroll
        for (int v = 0; v < HW; ++v)
            hx[c][v] = 0;
#pragma unroll
        for (int v = 0; v < VX; ++v)
            hy[c][v] = 0;
    }
    load_halo(z_begin);

    const double *cfs = p.coeff + 12 * (p.step ? *p.step : 0);
    const double *cf0 = cfs + 3 * p.coeff_stage0;
    const double cc0 = cf0[0], lam0 = cf0[1], R20 = cf0[2];
    const double cc1 = cf0[3], lam1 = cf0[4], R21 = cf0[5];
    // pair 4: stages 1 and 2 as well
    const double ccs1 = cfs[0], lams1 = cfs[1], R2s1 = cfs[2];
    const double ccs2 = cfs[3], lams2 = cfs[4], R2s2 = cfs[5];
    const T i2 = (T)p.inv_dx2;

    // pair 2 centre values: Pi_n, acc, lo (and with PF those of plane z+1)
    T pn[2][VX], ac[4][VX], pnn[2][VX], acn[4][VX];
    float lo[4][VX], lon[4][VX];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < VX; ++v)
        {
            ac[c][v] = acn[c][v] = 0;
            lo[c][v] = lon[c][v] = 0.0f;
            if (c < 2)
                pn[c][v] = pnn[c][v] = 0;
        }
    auto load_centre = [&](int zz, T(&pn_)[2][VX], T(&ac_)[4][VX],
                           float(&lo_)[4][VX]) {
        if constexpr (PAIR == 4 && PREC == AXS_FP32C)
        {
            if (!active)
                return;
            const uint32_t o = (uint32_t)zz * plane + rowc;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                vload_nt<float, VX>(p.comp + c * n + o, lo_[c]);
        }
        if constexpr (PAIR == 2)
        {
            if (!active)
                return;
            const uint32_t o = (uint32_t)zz * plane + rowc;
            vload_nt<T, VX>(p.yn + 2 * n + o, pn_[0]);
            vload_nt<T, VX>(p.yn + 3 * n + o, pn_[1]);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                vload_nt<T, VX>(p.acc + c * n + o, ac_[c]);
            if constexpr (PREC == AXS_FP32C)
            {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    vload_nt<float, VX>(p.comp + c * n + o, lo_[c]);
            }
        }
    };
    if constexpr (PF)
        load_centre(z_begin, pn, ac, lo);

    for (int z = z_begin; z < z_end; ++z)
    {
        const int b = (z - z_begin) & 1;
        const int cx0 = tx * VX + 2;
        // ---- stage A and B of plane z into LDS ----
#pragma unroll
        for (int v = 0; v < VX; ++v)
        {
            sm[b][0][ty + 2][cx0 + v] = qa1[2][v];
            sm[b][1][ty + 2][cx0 + v] = qa2[2][v];
            sm[b][2][ty + 2][cx0 + v] = qb1[2][v];
            sm[b][3][ty + 2][cx0 + v] = qb2[2][v];
            if constexpr (PAIR == 4)
            {
                sm[b][4 % NF][ty + 2][cx0 + v] = qc1[2][v];
                sm[b][5 % NF][ty + 2][cx0 + v] = qc2[2][v];
                sm[b][6 % NF][ty + 2][cx0 + v] = qd1[2][v];
                sm[b][7 % NF][ty + 2][cx0 + v] = qd2[2][v];
            }
        }
        if (hxl || hxr)
        {
            // VX == 1: lanes 0, 1 -> columns 0, 1; lanes 62, 63 -> TX+2, TX+3
            // VX >= 2: lane 0 -> columns 0, 1; lane 63 -> TX+2, TX+3
            int col0;
            if constexpr (VX == 1)
                col0 = hxl ? tx : tx + 4;
            else
                col0 = hxl ? 0 : TX + 2;
#pragma unroll
            for (int v = 0; v < HW; ++v)
            {
                T pt[8];
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    pt[c] = hx[c][v];
                sm[b][0][ty + 2][col0 + v] = pt[0];
                sm[b][1][ty + 2][col0 + v] = pt[1];
                sm[b][2][ty + 2][col0 + v] = form_b(pt, 0);
                sm[b][3][ty + 2][col0 + v] = form_b(pt, 1);
                if constexpr (PAIR == 4)
                {
                    sm[b][4 % NF][ty + 2][col0 + v] = pt[4];
                    sm[b][5 % NF][ty + 2][col0 + v] = pt[5];
                    sm[b][6 % NF][ty + 2][col0 + v] = form_d(pt, 0);
                    sm[b][7 % NF][ty + 2][col0 + v] = form_d(pt, 1);
                }
            }
        }
        if (hyl || hyr)
        {
            const int row = hyl ? ty : ty + 4;
#pragma unroll
            for (int v = 0; v < VX; ++v)
            {
                T pt[8];
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    pt[c] = hy[c][v];
                sm[b][0][row][cx0 + v] = pt[0];
                sm[b][1][row][cx0 + v] = pt[1];
                sm[b][2][row][cx0 + v] = form_b(pt, 0);
                sm[b][3][row][cx0 + v] = form_b(pt, 1);
                if constexpr (PAIR == 4)
                {
                    sm[b][4 % NF][row][cx0 + v] = pt[4];
                    sm[b][5 % NF][row][cx0 + v] = pt[5];
                    sm[b][6 % NF][row][cx0 + v] = form_d(pt, 0);
                    sm[b][7 % NF][row][cx0 + v] = form_d(pt, 1);
                }
            }
        }

        // ---- prefetch: queue heads at z+3, halos of plane z+1 ----
        T nw[8][VX];
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int v = 0; v < VX; ++v)
                nw[c][v] = 0;
        if (z + 1 < z_end)
        {
            const uint32_t o3 = zoff(z + 3) + rowc;
#pragma unroll
            for (int c = 0; c < NB; ++c)
                vload<T, VX>(src[c] + o3, nw[c]);
            load_halo(z + 1);
        }

        const uint32_t oc = (uint32_t)z * plane + rowc;
        if constexpr (PF)
        {
            if (z + 1 < z_end)
                load_centre(z + 1, pnn, acn, lon);
        }
        else
            load_centre(z, pn, ac, lo);

        __syncthreads();

        if (active)
        {
            const int cy = ty + 2;
            T out_y[4][VX], out_a[4][VX];
            float out_lo[4][VX];
#pragma unroll
            for (int v = 0; v < VX; ++v)
            {
                const int cx = cx0 + v;
                using S = typename StarOf<T>::type;
                const S sA1{sm[b][0][cy][cx - 2], sm[b][0][cy][cx - 1],
                            sm[b][0][cy][cx + 1], sm[b][0][cy][cx + 2],
                            sm[b][0][cy - 2][cx], sm[b][0][cy - 1][cx],
                            sm[b][0][cy + 1][cx], sm[b][0][cy + 2][cx],
                            qa1[0][v], qa1[1][v], qa1[3][v], qa1[4][v], qa1[2][v]};
                const S sA2{sm[b][1][cy][cx - 2], sm[b][1][cy][cx - 1],
                            sm[b][1][cy][cx + 1], sm[b][1][cy][cx + 2],
                            sm[b][1][cy - 2][cx], sm[b][1][cy - 1][cx],
                            sm[b][1][cy + 1][cx], sm[b][1][cy + 2][cx],
                            qa2[0][v], qa2[1][v], qa2[3][v], qa2[4][v], qa2[2][v]};
                const S sB1{sm[b][2][cy][cx - 2], sm[b][2][cy][cx - 1],
                            sm[b][2][cy][cx + 1], sm[b][2][cy][cx + 2],
                            sm[b][2][cy - 2][cx], sm[b][2][cy - 1][cx],
                            sm[b][2][cy + 1][cx], sm[b][2][cy + 2][cx],
                            qb1[0][v], qb1[1][v], qb1[3][v], qb1[4][v], qb1[2][v]};
                const S sB2{sm[b][3][cy][cx - 2], sm[b][3][cy][cx - 1],
                            sm[b][3][cy][cx + 1], sm[b][3][cy][cx + 2],
                            sm[b][3][cy - 2][cx], sm[b][3][cy - 1][cx],
                            sm[b][3][cy + 1][cx], sm[b][3][cy + 2][cx],
                            qb2[0][v], qb2[1][v], qb2[3][v], qb2[4][v], qb2[2][v]};
                T FA1, FA2, FB1, FB2;
                force(star_lap(sA1, i2), star_lap(sA2, i2), qa1[2][v],
                      qa2[2][v], cc0, lam0, R20, FA1, FA2);
                force(star_lap(sB1, i2), star_lap(sB2, i2), qb1[2][v],
                      qb2[2][v], cc1, lam1, R21, FB1, FB2);
                float lov[4];
                T ynv[4], acv[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                {
                    lov[c] = lo[c][v];
                    acv[c] = ac[c][v];
                }
                if constexpr (FIRST)
                {
                    // stage 1 on y_n, stage 2 on y_1 (y_n is the input)
                    ynv[0] = qa1[2][v];
                    ynv[1] = qa2[2][v];
                    ynv[2] = qp1[0][v];
                    ynv[3] = qp2[0][v];
                    const T k1[4] = {qp1[0][v], qp2[0][v], FA1, FA2};
                    T acc1[4], y1[4];
                    stage_update<T, PREC, MODE_S1>(k1, ynv, acv, lov, p.a0,
                                                   p.h6, p.h6_lo, acc1, y1);
                    const T k2[4] = {y1[2], y1[3], FB1, FB2};
                    T acc2[4], y2[4];
                    stage_update<T, PREC, MODE_S23>(k2, ynv, acc1, lov, p.a1,
                                                    p.h6, p.h6_lo, acc2, y2);
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                    {
                        out_y[c][v] = y2[c];
                        out_a[c][v] = acc2[c];
                        out_lo[c][v] = 0.0f;
                    }
                }
                else
                {
                    if constexpr (PAIR == 4)
                    {
                        // acc = k1 + 2 k2 recomputed: stage 1 on y_n (star
                        // C), stage 2 on y_1 (star D), as pair 1 forms it
                        ynv[0] = qc1[2][v];
                        ynv[1] = qc2[2][v];
                        ynv[2] = qm1[0][v];
                        ynv[3] = qm2[0][v];
                        const S sC1{sm[b][4 % NF][cy][cx - 2], sm[b][4 % NF][cy][cx - 1],
                                    sm[b][4 % NF][cy][cx + 1], sm[b][4 % NF][cy][cx + 2],
                                    sm[b][4 % NF][cy - 2][cx], sm[b][4 % NF][cy - 1][cx],
                                    sm[b][4 % NF][cy + 1][cx], sm[b][4 % NF][cy + 2][cx],
                                    qc1[0][v], qc1[1][v], qc1[3][v], qc1[4][v], qc1[2][v]};
                        const S sC2{sm[b][5 % NF][cy][cx - 2], sm[b][5 % NF][cy][cx - 1],
                                    sm[b][5 % NF][cy][cx + 1], sm[b][5 % NF][cy][cx + 2],
                                    sm[b][5 % NF][cy - 2][cx], sm[b][5 % NF][cy - 1][cx],
                                    sm[b][5 % NF][cy + 1][cx], sm[b][5 % NF][cy + 2][cx],
                                    qc2[0][v], qc2[1][v], qc2[3][v], qc2[4][v], qc2[2][v]};
                        const S sD1{sm[b][6 % NF][cy][cx - 2], sm[b][6 % NF][cy][cx - 1],
                                    sm[b][6 % NF][cy][cx + 1], sm[b][6 % NF][cy][cx + 2],
                                    sm[b][6 % NF][cy - 2][cx], sm[b][6 % NF][cy - 1][cx],
                                    sm[b][6 % NF][cy + 1][cx], sm[b][6 % NF][cy + 2][cx],
                                    qd1[0][v], qd1[1][v], qd1[3][v], qd1[4][v], qd1[2][v]};
                        const S sD2{sm[b][7 % NF][cy][cx - 2], sm[b][7 % NF][cy][cx - 1],
                                    sm[b][7 % NF][cy][cx + 1], sm[b][7 % NF][cy][cx + 2],
                                    sm[b][7 % NF][cy - 2][cx], sm[b][7 % NF][cy - 1][cx],
                                    sm[b][7 % NF][cy + 1][cx], sm[b][7 % NF][cy + 2][cx],
                                    qd2[0][v], qd2[1][v], qd2[3][v], qd2[4][v], qd2[2][v]};
                        T FC1, FC2, FD1, FD2;
                        force(star_lap(sC1, i2), star_lap(sC2, i2), qc1[2][v],
                              qc2[2][v], ccs1, lams1, R2s1, FC1, FC2);
                        force(star_lap(sD1, i2), star_lap(sD2, i2), qd1[2][v],
                              qd2[2][v], ccs2, lams2, R2s2, FD1, FD2);
                        const T k1[4] = {qm1[0][v], qm2[0][v], FC1, FC2};
                        T acc1[4], y1[4];
                        stage_update<T, PREC, MODE_S1>(k1, ynv, acv, lov, p.a1,
                                                       p.h6, p.h6_lo, acc1, y1);
                        const T k2[4] = {y1[2], y1[3], FD1, FD2};
                        T y2u[4];
                        stage_update<T, PREC, MODE_S23>(k2, ynv, acc1, lov, p.a1,
                                                        p.h6, p.h6_lo, acv, y2u);
                    }
                    else
                    {
                        ynv[0] = qn1[0][v];
                        ynv[1] = qn2[0][v];
                        ynv[2] = pn[0][v];
                        ynv[3] = pn[1][v];
                    }
                    // stage 3 on y_2, stage 4 on y_3
                    const T k3[4] = {qp1[0][v], qp2[0][v], FA1, FA2};
                    T acc3[4], y3[4];
                    stage_update<T, PREC, MODE_S23>(k3, ynv, acv, lov, p.a0,
                                                    p.h6, p.h6_lo, acc3, y3);
                    const T k4[4] = {y3[2], y3[3], FB1, FB2};
                    T unused[4], y4[4];
                    stage_update<T, PREC, MODE_S4>(k4, ynv, acc3, lov, p.a1,
                                                   p.h6, p.h6_lo, unused, y4);
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                    {
                        out_y[c][v] = y4[c];
                        out_a[c][v] = 0;
                        out_lo[c][v] = lov[c];
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                vstore<T, VX>(p.yout + c * n + oc, out_y[c], nt);
                if constexpr (PAIR == 1)
                    vstore<T, VX>(p.acc + c * n + oc, out_a[c], nt);
                if constexpr (!FIRST && PREC == AXS_FP32C)
                    vstore<float, VX>(p.comp + c * n + oc, out_lo[c], nt);
            }
        }

        if constexpr (PF)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int v = 0; v < VX; ++v)
                {
                    ac[c][v] = acn[c][v];
                    lo[c][v] = lon[c][v];
                    if (c < 2)
                        pn[c][v] = pnn[c][v];
                }
        }

        // ---- advance the queues ----
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < VX; ++v)
            {
                qa1[i][v] = qa1[i + 1][v];
                qa2[i][v] = qa2[i + 1][v];
                qb1[i][v] = qb1[i + 1][v];
                qb2[i][v] = qb2[i + 1][v];
                qc1[i][v] = qc1[i + 1][v];
                qc2[i][v] = qc2[i + 1][v];
                qd1[i][v] = qd1[i + 1][v];
                qd2[i][v] = qd2[i + 1][v];
            }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int v = 0; v < VX; ++v)
            {
                qp1[i][v] = qp1[i + 1][v];
                qp2[i][v] = qp2[i + 1][v];
                qn1[i][v] = qn1[i + 1][v];
                qn2[i][v] = qn2[i + 1][v];
                qm1[i][v] = qm1[i + 1][v];
                qm2[i][v] = qm2[i + 1][v];
            }
#pragma unroll
        for (int v = 0; v < VX; ++v)
        {
            T pt[8];
#pragma unroll
            for (int c = 0; c < 8; ++c)
                pt[c] = nw[c][v];
            qa1[4][v] = pt[0];
            qa2[4][v] = pt[1];
            qb1[4][v] = form_b(pt, 0);
            qb2[4][v] = form_b(pt, 1);
            qp1[2][v] = pt[2];
            qp2[2][v] = pt[3];
            qn1[2][v] = pt[4];
            qn2[2][v] = pt[5];
            if constexpr (PAIR == 4)
            {
                qc1[4][v] = pt[4];
                qc2[4][v] = pt[5];
                qd1[4][v] = form_d(pt, 0);
                qd2[4][v] = form_d(pt, 1);
                qm1[2][v] = pt[6];
                qm2[2][v] = pt[7];
            }
        }
    }
}

// ------------------------------------------------------------ launchers --
// Pair tile variants: (TY, VX).  variant = base + kPairBase * g, g bit 0:
// centre-load prefetch (pair 2), bit 1: non-temporal stores of the outputs,
// bit 2: y-major tile order.
struct PairShape
{
    int ty, vx;
};
static const PairShape kPairTiles[] = {{4, 1}, {8, 1}, {4, 2},
                                       {8, 2}, {4, 4}, {16, 1}};
constexpr int kPairBase = 6;

int num_pair_variants() { return 8 * kPairBase; }

template <typename T, int PREC, int PAIR, int TY, int VX>
static void launch_pt(const PairParams<T> &p, int ntiles, hipStream_t s,
                      bool pf)
{
    if (pf)
        hipLaunchKernelGGL((k_pair<T, PREC, PAIR, TY, VX, true>),
                           dim3(ntiles), dim3(64, TY), 0, s, p);
    else
        hipLaunchKernelGGL((k_pair<T, PREC, PAIR, TY, VX, false>),
                           dim3(ntiles), dim3(64, TY), 0, s, p);
}

template <typename T, int PREC, int PAIR>
static void launch_pshape(int base, const PairParams<T> &p, int ntiles,
                          hipStream_t s, bool pf)
{
    if constexpr (PAIR >= 3)
    {
        // the tiles launch_rk_pair maps these modes to (no spilling shapes)
        if constexpr (PAIR == 3 && sizeof(T) == 4)
        {
            if (base == 3)
            {
                launch_pt<T, PREC, PAIR, 8, 2>(p, ntiles, s, pf);
                return;
            }
        }
        if (base == 2)
            launch_pt<T, PREC, PAIR, 4, 2>(p, ntiles, s, pf);
        else if (base == 1)
            launch_pt<T, PREC, PAIR, 8, 1>(p, ntiles, s, pf);
        else
            launch_pt<T, PREC, PAIR, 4, 1>(p, ntiles, s, pf);
    }
    else
        switch (base)
        {
    case 0:
        launch_pt<T, PREC, PAIR, 4, 1>(p, ntiles, s, pf);
        break;
    case 1:
        launch_pt<T, PREC, PAIR, 8, 1>(p, ntiles, s, pf);
        break;
    case 2:
        launch_pt<T, PREC, PAIR, 4, 2>(p, ntiles, s, pf);
        break;
    case 3:
        // FP64 8 x 2 tiles need more than 256 VGPRs (spills): run 4 x 2
        if constexpr (sizeof(T) == 8)
            launch_pt<T, PREC, PAIR, 4, 2>(p, ntiles, s, pf);
        else
            launch_pt<T, PREC, PAIR, 8, 2>(p, ntiles, s, pf);
        break;
    case 4:
        if constexpr (sizeof(T) == 8)
            launch_pt<T, PREC, PAIR, 4, 2>(p, ntiles, s, pf);
        else
            launch_pt<T, PREC, PAIR, 4, 4>(p, ntiles, s, pf);
        break;
    default:
        // 16-row tiles: a quarter of the rows are y-halo reads, not half
        // (FP64 would spill at 1024 threads: 8 rows)
        if constexpr (sizeof(T) == 8)
            launch_pt<T, PREC, PAIR, 8, 1>(p, ntiles, s, pf);
        else
            launch_pt<T, PREC, PAIR, 16, 1>(p, ntiles, s, pf);
        break;
        }
}

template <typename T, int PREC>
static void launch_pmode(int pair, int base, const PairParams<T> &p,
                         int ntiles, hipStream_t s, bool pf)
{
    switch (pair)
    {
    case 1:
        launch_pshape<T, PREC, 1>(base, p, ntiles, s, pf);
        break;
    case 2:
        launch_pshape<T, PREC, 2>(base, p, ntiles, s, pf);
        break;
    case 3:
        launch_pshape<T, PREC, 3>(base, p, ntiles, s, pf);
        break;
    default:
        launch_pshape<T, PREC, 4>(base, p, ntiles, s, pf);
        break;
    }
}

// Pair 1 (stages 1, 2): y_n = yin -> acc, y_2 = yout.  Pair 2 (stages 3, 4):
// y_2 = yin, y_n = yn, acc, lo -> y_{n+1} = yout (yout != yn).  Pairs 3
// and 4: the same without the accumulator (pair 4 recomputes it; acc is
// not used).
int launch_rk_pair(axs_level *L, int pair, const void *yin, const void *yn,
                   void *yout, void *acc, const double *coeff_table,
                   const int *step_ptr, double dt)
{
    const bool f64 = L->precision == AXS_FP64;
    const int nx = L->g.nx, ny = L->g.ny, nz = L->g.nz;
    // defaults: the 1024^3 sweep's fastest (FP32: 8 x 2 tiles, z chunk 32;
    // FP64: 4 x 2 tiles, z chunk 64)
    const bool dflt =
        !(L->pair_variant >= 0 && L->pair_variant < num_pair_variants());
    // (accumulator-free pairs: FP32 8 x 1 for both launches)
    const int variant =
        dflt ? ((pair >= 3 && !f64) ? 1 : (f64 ? 2 : 3)) : L->pair_variant;
    const int group = variant / kPairBase;
    const bool pf = (group & 1) != 0;
    int base = variant % kPairBase;
    // FP64 tiles that would spill run smaller: 8 x 2 and 4 x 4 as 4 x 2,
    // 16 x 1 as 8 x 1 (the same mapping as launch_pshape, so that the grid
    // below is sized for the kernel that runs)
    if (f64 && (base == 3 || base == 4))
        base = 2;
    if (f64 && base == 5)
        base = 1;
    // accumulator-free pairs in FP32 / FP32C 4 x 4 tiles are not launched:
    // one GPU test of them (with the centre-load prefetch) gave wrong
    // results and the cause has not been shown
    if (pair >= 3 && base == 4 && !f64)
        return set_error(AXS_ERR_UNSUPPORTED,
                         "fused pairs without an accumulator (stage fusion 2) "
                         "do not run in 4 x 4 tiles (pair variants 4 + 6 g)");
    // the stages 3+4 kernel of those pairs holds four star fields: its 8 x 2
    // and 16 x 1 tiles spill, so they run as 4 x 2 / 8 x 1, and the stages
    // 1+2 kernel's 16 x 1 too (FP64 4 x 4 runs as 4 x 2, above)
    if (pair == 4 && base == 3)
        base = 2;
    if (pair >= 3 && base == 5)
        base = 1;
    if (nx % kPairTiles[base].vx != 0)
        base = 1;
    const int TY = kPairTiles[base].ty, VX = kPairTiles[base].vx;
    const int tiles_x = (nx + 64 * VX - 1) / (64 * VX);
    const int tiles_y = (ny + TY - 1) / TY;
    int zc = L->pair_z_chunk;
    if (zc <= 0)
    {
        zc = (f64 && dflt) ? 64 : 32;
        while (zc > 4 &&
               (long)tiles_x * tiles_y * ((nz + zc - 1) / zc) < 4096)
            zc >>= 1;
    }
    if (zc > nz)
        zc = nz;
    const int tiles_z = (nz + zc - 1) / zc;
    const long ntiles = (long)tiles_x * tiles_y * tiles_z;
    hipStream_t s = L->ctx->stream;

    auto fill = [&](auto &p) {
        using TT = typename std::remove_cv<
            typename std::remove_pointer<decltype(p.yn)>::type>::type;
        p.yin = static_cast<const TT *>(yin);
        p.yn = static_cast<const TT *>(yn);
        p.acc = static_cast<TT *>(acc);
        p.yout = static_cast<TT *>(yout);
        p.comp = L->comp;
        p.coeff = coeff_table;
        p.step = step_ptr;
        const bool first = pair == 1 || pair == 3;
        p.coeff_stage0 = first ? 0 : 2;
        p.a0 = first ? 0.5 * dt : dt;
        p.a1 = 0.5 * dt;
        p.h6 = dt / 6.0;
        p.h6_lo = (float)(p.h6 - (double)(float)p.h6);
        p.nx = nx;
        p.ny = ny;
        p.nz = nz;
        p.n = L->g.ncell;
        p.inv_dx2 = 1.0 / (L->dx * L->dx);
        p.z_chunk = zc;
        p.tiles_x = tiles_x;
        p.tiles_y = tiles_y;
        p.tiles_z = tiles_z;
        p.xcd_remap = (ntiles % 8 == 0) ? 1 : 0;
        p.nontemporal = (group & 2) ? 1 : 0;
        p.y_major = (group & 4) ? 1 : 0;
    };
    if (f64)
    {
        PairParams<double> p{};
        fill(p);
        launch_pmode<double, AXS_FP64>(pair, base, p, (int)ntiles, s, pf);
    }
    else
    {
        PairParams<float> p{};
        fill(p);
        if (L->precision == AXS_FP32C)
            launch_pmode<float, AXS_FP32C>(pair, base, p, (int)ntiles, s, pf);
        else
            launch_pmode<float, AXS_FP32>(pair, base, p, (int)ntiles, s, pf);
    }
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

} // namespace axs
