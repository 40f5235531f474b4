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
    int xcd_remap; // 0: none; 1: XCD k takes the k-th eighth of the tiles;
                   // G >= 2: bands of G consecutive tiles per XCD, the 8
                   // XCDs' bands adjacent
    int patch; // tile order: 0 x-major; 1 patches of 4 x 8 tiles (x, y)
};

// Tile (bx, by, bz) of position s in the launch order.  x-major: x fastest,
// then y, then z.  Patches: the tile rows are cut into bands of 8 (the last
// band may be shorter), and a band is walked in groups of 4 columns, each
// group row by row -- so 32 consecutive positions (one XCD's workgroups at
// a time, with the XCD remap) cover a 4 x 8 patch of tiles whose halo
// rows and columns are each other's interior.  Needs tiles_x % 4 == 0.
__device__ __forceinline__ void step_tile(const StepParams &p, int s, int &bx,
                                          int &by, int &bz)
{
    const int per_z = p.tiles_x * p.tiles_y;
    bz = s / per_z;
    int r = s - bz * per_z;
    if (!p.patch)
    {
        bx = r % p.tiles_x;
        by = r / p.tiles_x;
        return;
    }
    const int band = r / (p.tiles_x * 8);
    const int h = min(8, p.tiles_y - band * 8);
    r -= band * p.tiles_x * 8;
    const int g = r / (4 * h), w = r % (4 * h);
    by = band * 8 + w / 4;
    bx = g * 4 + (w & 3);
}

// Launch position b -> tile position.  Workgroups are dealt to the 8 XCDs
// round robin (b mod 8).  Mode 1 gives XCD k the k-th eighth of the tile
// list, so the 32 workgroups one XCD runs at a time are consecutive tiles
// (two tile rows of a 1024-wide level), whose x and inner y halos are each
// other's interior and meet in that XCD's L2.  Mode G >= 2 deals the list in
// super-bands of 8 G tiles: XCD k takes the k-th band of G consecutive tiles
// of each, so the 8 XCDs work on 8 adjacent bands of one z chunk at the
// same time and the rows between bands are read by two XCDs at about the
// same time (the last partial super-band keeps launch order).  A bijection
// on [0, ntiles) either way.
__device__ __forceinline__ int xcd_tile(const StepParams &p, int b)
{
    const int ntiles = p.tiles_x * p.tiles_y * p.tiles_z;
    if (p.xcd_remap == 1)
        return (b & 7) * (ntiles >> 3) + (b >> 3);
    if (p.xcd_remap >= 2)
    {
        const int G = p.xcd_remap;
        if (b < ntiles / (8 * G) * (8 * G))
        {
            const int j = b >> 3;
            return (j / G) * (8 * G) + (b & 7) * G + j % G;
        }
    }
    return b;
}

constexpr int kStepTX = 60; // interior columns per tile (64 lanes - 2 x 2)
constexpr int kStepBand = 32; // tiles per XCD band (variants 7, 8)

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

    const int bid = xcd_tile(p, blockIdx.x);
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

// ---------------------------------------------------------------------
// k_step_r: the same step, the same tile (64 x TY cells, rows aligned to
// cache lines, one thread per point of the tile grown by 2, the 4 x-halo
// columns on an extra wave) and the same expressions, with 23% fewer vector
// instructions than k_step (6.65e9 against 8.62e9 per 1024^3 launch):
//  * every per-thread queue along z is a ring whose length divides 6 and
//    the z loop is unrolled by 6, so that a ring index is a compile-time
//    constant in every plane and no queue moves data;
//  * psi_2, psi_3 of six planes stay in an LDS ring (their z neighbours are
//    the thread's own earlier writes, so they need no barrier), psi_1 at
//    the z neighbours is formed again from psi_n and Pi_n, and the grown
//    tile's corner points load no y-halo (nothing formed there is read):
//    117 registers, no spills;
//  * loads and stores are buffer instructions on per-plane descriptors
//    built in scalar registers: a thread's address is its 32-bit row
//    offset, fixed for the whole launch, so no vector address arithmetic;
//  * psi pairs live in LDS as float2 (one 64-bit access per star point),
//    and the stage scalars are rounded to float once, in scalar registers.
// Measured at 1024^3 on MI355X it runs at the speed of k_step (19 ms per
// step, within the placement spread), with more HBM traffic (62-70 against
// 42 B per cell read): its time per plane is not set by the vector
// instructions it saved.
// ---------------------------------------------------------------------
// A uniform address kept in scalar registers (readfirstlane stops the
// compiler from building it, or anything derived from it, in vector
// registers).
__device__ __forceinline__ uint64_t uniform_addr(const void *p)
{
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return ((uint64_t)hi << 32) | lo;
}
// Raw buffer over one plane at byte address base + 4 off (base, off
// uniform): the range only has to exceed the largest row offset (checked by
// the launcher).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(uint64_t base,
                                                             uint32_t off)
{
    void *u = reinterpret_cast<void *>(base + 4 * (uint64_t)off);
    return __builtin_amdgcn_make_buffer_rsrc(u, (short)0, (int)0xffffffff,
                                             0x00020000);
}
// AUX is the instruction's cache-policy field: 0 for data that other
// workgroups read again (y_n: halos), 2 (non-temporal) for data touched
// once per step (the FP32C low words, read and written at the own point,
// and y_{n+1}, written once).  Non-temporal traffic leaves the caches to
// the halo rows: measured at 1024^3 FP32C, 4-5 % off the step.
template <int AUX = 0>
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX));
}
template <int AUX = 0>
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                       float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r,
                                          off, 0, AUX);
}
constexpr int kNonTemporal = 2;
__device__ __forceinline__ float uniform_f(float v)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// D: psi_n, Pi_n of the own point and the halo words are loaded D planes
// before their first use (D = 1 or 2); the FP32C low words one plane ahead.
template <int PREC, int TY, int D>
__global__ void __launch_bounds__(68 * (TY + 4)) k_step_r(StepParams p)
{
    constexpr int EXB = 68, EXA = 72;
    constexpr int HYA = TY + 8, HYB = TY + 4;
    constexpr int NW = 64 * (TY + 4); // threads of the full-row waves
    // LDS planes, flattened: sa[buffer][psi_n, psi_1][HYA][EXA],
    // sb[plane % 6][psi_2, psi_3][HYB][EXB]
    __shared__ float2 sa[2 * 2 * HYA * EXA];
    __shared__ float2 sb[6 * 2 * HYB * EXB];

    const int bid = xcd_tile(p, blockIdx.x);
    int bx, by, bz;
    step_tile(p, bid, bx, by, bz);

    // own point (ec, er) of the tile grown by 2: x = X0 - 2 + ec,
    // y = Y0 - 2 + er
    const int t = threadIdx.x;
    int ec, er;
    if (t < NW)
    {
        ec = t % 64 + 2;
        er = t / 64;
    }
    else
    {
        const int u = t - NW;
        ec = u % 4 < 2 ? u % 4 : u % 4 + 64;
        er = u / 4;
    }
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t n = p.n;
    const uint32_t plane = (uint32_t)nx * (uint32_t)ny;
    const int X0 = bx * 64, Y0 = by * TY;
    auto wrapx = [&](int x) { return ((x % nx) + nx) % nx; };
    auto wrapy = [&](int y) { return ((y % ny) + ny) % ny; };
    const int gx = X0 - 2 + ec, gy = Y0 - 2 + er;
    const uint32_t rowb = 4u * ((uint32_t)wrapy(gy) * nx + wrapx(gx));
    const bool interior = ec >= 2 && ec < 66 && er >= 2 && er < TY + 2;
    const bool active = interior && gx < nx && gy < ny;
    // halo points of the plane grown by 4 (a plus shape): the y-halo rows
    // above and below the grown tile's columns 2..65, the x-halo columns
    // beside its rows.  The grown tile's corner points (x-halo wave, top
    // and bottom two rows) need no y-halo: what is formed there is never
    // read, since every later star is a plus.  So each thread has at most
    // one halo point (none for interior threads) at LDS (yh, xh), global
    // byte offset rowhb.
    const bool hxl = ec < 2, hxr = ec >= 66;
    const bool hx_on = hxl || hxr;
    const bool hyl = er < 2, hyr = er >= TY + 2;
    const bool hy_on = (hyl || hyr) && !hx_on;
    const bool h_on = hx_on || hy_on;
    uint32_t rowhb = 0;
    int yh = er + 2, xh = ec + 2;
    if (hy_on)
    {
        yh = hyl ? er : er + 4;
        rowhb = 4u * ((uint32_t)wrapy(Y0 - 4 + yh) * nx + wrapx(gx));
    }
    if (hx_on)
    {
        xh = hxl ? ec : ec + 4;
        rowhb = 4u * ((uint32_t)wrapy(gy) * nx + wrapx(X0 - 4 + xh));
    }
    const int ya = er + 2, xa = ec + 2;

    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);
    // planes z in [-4, nz + 5] are read: one add or subtract wraps them
    // when nz >= 8 (a scalar branch; the modulo otherwise)
    auto zoff = [&](int z) -> uint32_t {
        if (nz >= 8)
            z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
        else
            z = ((z % nz) + nz) % nz;
        return (uint32_t)z * plane;
    };
    uint64_t bn[4], bo[4], bl[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
    {
        bn[c] = uniform_addr(p.yn + (size_t)c * n);
        bo[c] = uniform_addr(p.yout + (size_t)c * n);
        bl[c] = PREC == AXS_FP32C ? uniform_addr(p.comp + (size_t)c * n) : 0;
    }
    // loads / stores of component c at plane offset o (cells), byte offset
    // off within the plane
    auto ld_n = [&](int c, uint32_t o, uint32_t off) {
        return bload(plane_rsrc(bn[c], o), off);
    };
    auto ld_l = [&](int c, uint32_t o, uint32_t off) {
        return bload<kNonTemporal>(plane_rsrc(bl[c], o), off);
    };
    auto st_o = [&](int c, uint32_t o, uint32_t off, float v) {
        bstore<kNonTemporal>(plane_rsrc(bo[c], o), off, v);
    };
    auto st_l = [&](int c, uint32_t o, uint32_t off, float v) {
        bstore<kNonTemporal>(plane_rsrc(bl[c], o), off, v);
    };

    // stage scalars of the four stages, rounded as force() rounds them
    const double *cfp = p.coeff + 12 * (p.step ? *p.step : 0);
    float ccf[4], hlf[4], r2f[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
    {
        ccf[s] = uniform_f((float)cfp[3 * s]);
        hlf[s] = uniform_f(0.5f * (float)cfp[3 * s + 1]);
        r2f[s] = uniform_f((float)cfp[3 * s + 2]);
    }
    const float i2 = uniform_f((float)p.inv_dx2);
    const float ahf = uniform_f((float)(0.5 * p.h)); // (float)(dt/2)
    const float hf = uniform_f((float)p.h);          // (float)dt
    const float h6f = uniform_f((float)p.h6);        // (float)(dt/6)
    const float h6lo = p.h6_lo;
    // psi_1 = psi_n + (dt/2) Pi_n, as stage_axpy forms it
    auto psi1_of = [&](float ps, float pi) { return fmaf(ahf, pi, ps); };

    // Rings: plane q in slot (q - q0) % 6 (q0 = z_begin - 2) of N, P
    // (psi_n, Pi_n at q-2..q+2) and of the LDS ring sb; in slot
    // (q - q0) % 3 of FN, F1 (F(psi_n), F(psi_1) at q-2..q).  In flight:
    // psi_n, Pi_n of the next D planes in staging slot T[(plane - q0) % D],
    // halo words of the next D planes in HW[(plane - q0) % D], the
    // FP32C low words of plane z in LO[(z - z_begin) % 2]
    static_assert(D == 1 || D == 2, "prefetch distance 1 or 2");
    float N[6][2], P[6][2], FN[3][2], F1[3][2];
    float T[D][4], HW[D][4], LO[2][4];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            T[i][c] = HW[i][c] = 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            LO[i][c] = 0.0f;
    const int q0 = z_begin - 2;
    auto load_own = [&](int z, float (&w)[4]) {
        const uint32_t o = zoff(z);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            w[c] = ld_n(c, o, rowb);
    };
    auto load_halo = [&](int z, float (&w)[4]) {
        const uint32_t o = zoff(z);
        if (h_on)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                w[c] = ld_n(c, o, rowhb);
        }
    };
    // prologue: planes q0-2..q0+1 into N, P; q0+2..q0+1+D in flight
    // (staging slot (plane - q0) % D); halos of q0..q0+D-1.  The low words
    // of plane z_begin are loaded in the loop, one plane ahead.
#pragma unroll
    for (int k = -2; k <= 1; ++k)
    {
        float w[4];
        load_own(q0 + k, w);
        const int s = (k + 6) % 6;
        N[s][0] = w[0];
        N[s][1] = w[1];
        P[s][0] = w[2];
        P[s][1] = w[3];
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
    {
        load_own(q0 + 2 + i, T[i]);
        load_halo(q0 + i, HW[i]);
    }

    // one plane q, ring phase PH = (q - q0) % 6
    auto plane_step = [&]<int PH>(int q) {
        constexpr int s0 = PH;            // plane q in the 6-rings
        constexpr int sm2 = (PH + 4) % 6; // q - 2
        constexpr int sm1 = (PH + 5) % 6; // q - 1
        constexpr int sp1 = (PH + 1) % 6; // q + 1
        constexpr int sp2 = (PH + 2) % 6; // q + 2
        constexpr int sm4 = (PH + 2) % 6; // q - 4 (sb only)
        constexpr int sm3 = (PH + 3) % 6; // q - 3 (sb only)
        constexpr int t0 = PH % 3;        // plane q in FN, F1
        constexpr int tm2 = (PH + 1) % 3; // q - 2
        constexpr int ba = PH & 1;        // LDS buffer of plane q
        constexpr int tq = PH % D;        // T slot of planes q + 2, q + 2 + D
        constexpr int hq = PH % D;        // HW slot of planes q, q + D
        constexpr int lz = PH & 1;        // LO slot of plane q - 2
        constexpr int lz1 = (PH + 1) & 1; // LO slot of plane q - 1
        // the own point's LDS indices, taken through an empty register
        // constraint once per plane: the compiler forms the addresses of
        // every ring slot from them in this plane instead of keeping each
        // slot's address live across the loop
        int ia = ya * EXA + xa, ib = er * EXB + ec;
        asm volatile("" : "+v"(ia), "+v"(ib));
        auto SA = [&](int buf, int f, int dy, int dx) -> float2 & {
            return sa[(buf * 2 + f) * HYA * EXA + ia + dy * EXA + dx];
        };
        auto SB = [&](int slot, int f, int dy, int dx) -> float2 & {
            return sb[(slot * 2 + f) * HYB * EXB + ib + dy * EXB + dx];
        };

        __builtin_amdgcn_sched_barrier(0);
        // plane q + 2 from its staging slot (loaded two planes ago)
        N[sp2][0] = T[tq][0];
        N[sp2][1] = T[tq][1];
        P[sp2][0] = T[tq][2];
        P[sp2][1] = T[tq][3];
        // psi_n, psi_1 of plane q into LDS (grown by 4)
        SA(ba, 0, 0, 0) = make_float2(N[s0][0], N[s0][1]);
        SA(ba, 1, 0, 0) = make_float2(psi1_of(N[s0][0], P[s0][0]),
                                      psi1_of(N[s0][1], P[s0][1]));
        if (h_on)
        {
            const float *h = HW[hq];
            sa[((ba * 2 + 0) * HYA + yh) * EXA + xh] = make_float2(h[0], h[1]);
            sa[((ba * 2 + 1) * HYA + yh) * EXA + xh] =
                make_float2(psi1_of(h[0], h[2]), psi1_of(h[1], h[3]));
        }
        // loads: own point of plane q + 2 + D into the slot just emptied (D
        // = 1; with D = 2 at the end of the plane, see there), lo of plane
        // q - 1
        if constexpr (D == 1)
            load_own(q + 2 + D, T[tq]);
        const int z = q - 2; // plane of the final stage
        if constexpr (PREC == AXS_FP32C)
        {
            const int zl = z + 1;
            if (active && zl >= z_begin && zl < z_end)
            {
                const uint32_t ol = (uint32_t)zl * plane;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    LO[lz1][c] = ld_l(c, ol, rowb);
            }
        }

        __syncthreads();

        // F(psi_n), F(psi_1) at the own point of plane q (one field at a
        // time); Pi_1, Pi_2, psi_2, psi_3 there
        {
            const float2 xn2 = SA(ba, 0, 0, -2), xn1 = SA(ba, 0, 0, -1);
            const float2 xp1 = SA(ba, 0, 0, 1), xp2 = SA(ba, 0, 0, 2);
            const float2 yn2 = SA(ba, 0, -2, 0), yn1 = SA(ba, 0, -1, 0);
            const float2 yp1 = SA(ba, 0, 1, 0), yp2 = SA(ba, 0, 2, 0);
            const StarF a1{xn2.x, xn1.x, xp1.x, xp2.x, yn2.x, yn1.x, yp1.x, yp2.x,
                           N[sm2][0], N[sm1][0], N[sp1][0], N[sp2][0], N[s0][0]};
            const StarF a2{xn2.y, xn1.y, xp1.y, xp2.y, yn2.y, yn1.y, yp1.y, yp2.y,
                           N[sm2][1], N[sm1][1], N[sp1][1], N[sp2][1], N[s0][1]};
            force_f(star_lap(a1, i2), star_lap(a2, i2), N[s0][0], N[s0][1],
                    ccf[0], hlf[0], r2f[0], FN[t0][0], FN[t0][1]);
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            const float2 xn2 = SA(ba, 1, 0, -2), xn1 = SA(ba, 1, 0, -1);
            const float2 xp1 = SA(ba, 1, 0, 1), xp2 = SA(ba, 1, 0, 2);
            const float2 yn2 = SA(ba, 1, -2, 0), yn1 = SA(ba, 1, -1, 0);
            const float2 yp1 = SA(ba, 1, 1, 0), yp2 = SA(ba, 1, 2, 0);
            const float2 c0 = SA(ba, 1, 0, 0);
            float om2[2], om1[2], op1[2], op2[2];
#pragma unroll
            for (int c = 0; c < 2; ++c)
            {
                om2[c] = psi1_of(N[sm2][c], P[sm2][c]);
                om1[c] = psi1_of(N[sm1][c], P[sm1][c]);
                op1[c] = psi1_of(N[sp1][c], P[sp1][c]);
                op2[c] = psi1_of(N[sp2][c], P[sp2][c]);
            }
            const StarF b1{xn2.x, xn1.x, xp1.x, xp2.x, yn2.x, yn1.x, yp1.x, yp2.x,
                           om2[0], om1[0], op1[0], op2[0], c0.x};
            const StarF b2{xn2.y, xn1.y, xp1.y, xp2.y, yn2.y, yn1.y, yp1.y, yp2.y,
                           om2[1], om1[1], op1[1], op2[1], c0.y};
            force_f(star_lap(b1, i2), star_lap(b2, i2), c0.x, c0.y, ccf[1],
                    hlf[1], r2f[1], F1[t0][0], F1[t0][1]);
        }
        {
            float s2[2], s3[2];
#pragma unroll
            for (int c = 0; c < 2; ++c)
            {
                // stage_axpy(y, a, k) = fmaf((float)a, k, y)
                const float p1 = fmaf(ahf, FN[t0][c], P[s0][c]);
                const float p2 = fmaf(ahf, F1[t0][c], P[s0][c]);
                s2[c] = fmaf(ahf, p1, N[s0][c]);
                s3[c] = fmaf(hf, p2, N[s0][c]);
            }
            SB(s0, 0, 0, 0) = make_float2(s2[0], s2[1]);
            SB(s0, 1, 0, 0) = make_float2(s3[0], s3[1]);
        }

        __builtin_amdgcn_sched_barrier(0);
        // stages 3, 4 and the update at plane z = q - 2: psi_2, psi_3 of
        // plane z sit in LDS slot sm2 (written two planes ago), their z
        // neighbours at the own point in slots sm4, sm3, sm1, s0 (planes
        // z-2..z+2 = q-4..q, written by this thread)
        if (z >= z_begin && active)
        {
            const uint32_t oz = (uint32_t)z * plane;
            float lov[4] = {LO[lz][0], LO[lz][1], LO[lz][2], LO[lz][3]};
            if constexpr (D == 1)
            {
                // halo words are never used by an interior thread: mark
                // them dead on this path so that their registers serve the
                // stage (with D = 2 they stay put: registers that may still
                // be the targets of loads in flight would need a drain)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    HW[hq][c] = __builtin_nondeterministic_value(HW[hq][c]);
            }
            float Fs2[2], Fs3[2];
            {
                const float2 xn2 = SB(sm2, 0, 0, -2), xn1 = SB(sm2, 0, 0, -1);
                const float2 xp1 = SB(sm2, 0, 0, 1), xp2 = SB(sm2, 0, 0, 2);
                const float2 yn2 = SB(sm2, 0, -2, 0), yn1 = SB(sm2, 0, -1, 0);
                const float2 yp1 = SB(sm2, 0, 1, 0), yp2 = SB(sm2, 0, 2, 0);
                const float2 c0 = SB(sm2, 0, 0, 0);
                const float2 zn2 = SB(sm4, 0, 0, 0), zn1 = SB(sm3, 0, 0, 0);
                const float2 zp1 = SB(sm1, 0, 0, 0), zp2 = SB(s0, 0, 0, 0);
                const StarF a1{xn2.x, xn1.x, xp1.x, xp2.x, yn2.x, yn1.x, yp1.x,
                               yp2.x, zn2.x, zn1.x, zp1.x, zp2.x, c0.x};
                const StarF a2{xn2.y, xn1.y, xp1.y, xp2.y, yn2.y, yn1.y, yp1.y,
                               yp2.y, zn2.y, zn1.y, zp1.y, zp2.y, c0.y};
                force_f(star_lap(a1, i2), star_lap(a2, i2), c0.x, c0.y, ccf[2],
                        hlf[2], r2f[2], Fs2[0], Fs2[1]);
            }
            __builtin_amdgcn_sched_barrier(0);
            {
                const float2 xn2 = SB(sm2, 1, 0, -2), xn1 = SB(sm2, 1, 0, -1);
                const float2 xp1 = SB(sm2, 1, 0, 1), xp2 = SB(sm2, 1, 0, 2);
                const float2 yn2 = SB(sm2, 1, -2, 0), yn1 = SB(sm2, 1, -1, 0);
                const float2 yp1 = SB(sm2, 1, 1, 0), yp2 = SB(sm2, 1, 2, 0);
                const float2 c0 = SB(sm2, 1, 0, 0);
                const float2 zn2 = SB(sm4, 1, 0, 0), zn1 = SB(sm3, 1, 0, 0);
                const float2 zp1 = SB(sm1, 1, 0, 0), zp2 = SB(s0, 1, 0, 0);
                const StarF a1{xn2.x, xn1.x, xp1.x, xp2.x, yn2.x, yn1.x, yp1.x,
                               yp2.x, zn2.x, zn1.x, zp1.x, zp2.x, c0.x};
                const StarF a2{xn2.y, xn1.y, xp1.y, xp2.y, yn2.y, yn1.y, yp1.y,
                               yp2.y, zn2.y, zn1.y, zp1.y, zp2.y, c0.y};
                force_f(star_lap(a1, i2), star_lap(a2, i2), c0.x, c0.y, ccf[3],
                        hlf[3], r2f[3], Fs3[0], Fs3[1]);
            }
            // y_n at plane z: psi_n, Pi_n in slot sm2 of N, P
            const float ynv[4] = {N[sm2][0], N[sm2][1], P[sm2][0], P[sm2][1]};
            float none[4] = {0, 0, 0, 0};
            const float k1[4] = {P[sm2][0], P[sm2][1], FN[tm2][0], FN[tm2][1]};
            float acc1[4], y1[4];
            stage_update_f<PREC, MODE_S1>(k1, ynv, none, lov, ahf, h6f, h6lo,
                                          acc1, y1);
            const float k2[4] = {y1[2], y1[3], F1[tm2][0], F1[tm2][1]};
            float acc2[4], y2[4];
            stage_update_f<PREC, MODE_S23>(k2, ynv, acc1, lov, ahf, h6f, h6lo,
                                           acc2, y2);
            const float k3[4] = {y2[2], y2[3], Fs2[0], Fs2[1]};
            float acc3[4], y3[4];
            stage_update_f<PREC, MODE_S23>(k3, ynv, acc2, lov, hf, h6f, h6lo,
                                           acc3, y3);
            const float k4[4] = {y3[2], y3[3], Fs3[0], Fs3[1]};
            float unused[4], y4[4];
            stage_update_f<PREC, MODE_S4>(k4, ynv, acc3, lov, hf, h6f, h6lo,
                                          unused, y4);
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                st_o(c, oz, rowb, y4[c]);
                if constexpr (PREC == AXS_FP32C)
                    st_l(c, oz, rowb, lov[c]);
            }
        }
        else
        {
            // halos of plane q + D into the slot plane q used
            load_halo(q + D, HW[hq]);
        }
        // D = 2: the own point's loads are the last memory instructions of
        // every plane on every path.  The wait before a loaded value is used
        // counts the memory instructions issued after it, the least over
        // all paths (stores, the low words and the halo words are issued on
        // some paths only); with these loads last, two planes ahead, the
        // waits at the top of a plane leave the newest loads in flight
        // instead of draining every store and load
        if constexpr (D == 2)
            load_own(q + 2 + D, T[tq]);
    };

    const int q_end = z_end + 2;
    for (int q = q0;;)
    {
        plane_step.template operator()<0>(q);
        if (++q >= q_end)
            break;
        plane_step.template operator()<1>(q);
        if (++q >= q_end)
            break;
        plane_step.template operator()<2>(q);
        if (++q >= q_end)
            break;
        plane_step.template operator()<3>(q);
        if (++q >= q_end)
            break;
        plane_step.template operator()<4>(q);
        if (++q >= q_end)
            break;
        plane_step.template operator()<5>(q);
        if (++q >= q_end)
            break;
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
// FP32C lo in place.  The level's pair variant selects the kernel and tile:
// -1 (default): k_step_r over 64 x 11 tiles with rows aligned to cache
// lines (1020 threads, z chunk 128), loading one plane ahead; 4: two planes
// ahead; 6 / 5: -1 / 4 with the tiles in 4 x 8 patches; 8 / 7: -1 / 4 with
// the XCDs on adjacent bands of 32 tiles; 3: k_step over the same tiles; 2:
// k_step over 60 x 12 tiles (1024 threads); 0: 60 x 8; 1: 60 x 4.  All give
// the same bits.
int launch_rk_step(axs_level *L, const void *yin, void *yout,
                   const double *coeff_table, const int *step_ptr, double dt)
{
    if (L->precision == AXS_FP64)
        return set_error(AXS_ERR_UNSUPPORTED,
                         "one launch per RK4 step is FP32 / FP32C only");
    // ring kernel: -1 (loads one plane ahead), 4 (two), 5 / 6 (4 / -1 with
    // the tiles in 4 x 8 patches)
    const int pv = L->pair_variant;
    const bool ring = pv < 0 || (pv >= 4 && pv <= 8);
    const bool al = ring || L->pair_variant == 3;
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
    // the ring kernel addresses a plane by 32-bit byte offsets
    if (ring && (uint64_t)p.nx * (uint64_t)p.ny * 4u >= (uint64_t(1) << 31))
        return set_error(AXS_ERR_UNSUPPORTED,
                         "one launch per RK4 step: x-y plane above 2^29 cells");
    p.tiles_x = (p.nx + txi - 1) / txi;
    p.tiles_y = (p.ny + TY - 1) / TY;
    int zc = L->pair_z_chunk;
    if (zc <= 0)
    {
        // automatic: 128 planes (64 for the 60-column tiles), halved on
        // small grids until there are 4096 tiles
        zc = al ? 128 : 64;
        while (zc > 4 &&
               (long)p.tiles_x * p.tiles_y * ((p.nz + zc - 1) / zc) < 4096)
            zc >>= 1;
    }
    if (zc > p.nz)
        zc = p.nz;
    p.z_chunk = zc;
    p.tiles_z = (p.nz + zc - 1) / zc;
    const long ntiles = (long)p.tiles_x * p.tiles_y * p.tiles_z;
    p.xcd_remap = (ntiles % 8 == 0) ? 1 : 0;
    if (pv == 7 || pv == 8)
        p.xcd_remap = kStepBand;
    p.patch = (ring && (pv == 5 || pv == 6) && p.tiles_x % 4 == 0) ? 1 : 0;
    hipStream_t s = L->ctx->stream;
    if (ring)
    {
        const dim3 block(68 * (11 + 4));
        // variants 4, 5, 7: loads two planes ahead; else one
        const bool far = pv == 4 || pv == 5 || pv == 7;
        if (L->precision == AXS_FP32C && far)
            hipLaunchKernelGGL((k_step_r<AXS_FP32C, 11, 2>), dim3(ntiles),
                               block, 0, s, p);
        else if (L->precision == AXS_FP32C)
            hipLaunchKernelGGL((k_step_r<AXS_FP32C, 11, 1>), dim3(ntiles),
                               block, 0, s, p);
        else if (far)
            hipLaunchKernelGGL((k_step_r<AXS_FP32, 11, 2>), dim3(ntiles),
                               block, 0, s, p);
        else
            hipLaunchKernelGGL((k_step_r<AXS_FP32, 11, 1>), dim3(ntiles),
                               block, 0, s, p);
    }
    else if (al)
        launch_step_ty<11, true>(p, L->precision, ntiles, s);
    else if (TY == 4)
        launch_step_ty<4, false>(p, L->precision, ntiles, s);
    else if (TY == 12)
        launch_step_ty<12, false>(p, L->precision, ntiles, s);
    else
        launch_step_ty<8, false>(p, L->precision, ntiles, s);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

} // namespace axs
