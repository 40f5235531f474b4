// axs_diag.hip -- string finding, tagging and energy reductions on a uniform
// periodic level.
//
// Windings (StringFinder.hpp:45-122): every vertex phase atan2(psi2, psi1)
// is evaluated once, in FP64 on the FP64-promoted state exactly like the
// reference (StringFinder.hpp:49-53), into a 3-deep ring of LDS planes;
// each cell then reads the 7 corner phases of its 3 low-corner plaquettes
// from LDS.  Pierced-plaquette counts are formed per wavefront with
// __ballot / __popcll and one 64-bit integer atomic per wavefront, so the
// counts are exact and order independent.
//
// Energies (EnergyKernel.hpp:106-233): FP64 per-cell arithmetic, per-thread
// accumulation, then a fixed-order block tree and a fixed-order second pass
// over the block partials: the 14 sums are bitwise reproducible run to run.
#include "axs_internal.hpp"

#include <cfloat>
#include <cmath>
#include <vector>

namespace axs
{

template <typename T>
__device__ __forceinline__ double ld(const T *p, const float *lo, uint32_t i)
{
    return lo ? (double)p[i] + (double)lo[i] : (double)p[i];
}

__device__ __forceinline__ int wrapc(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// ---- PlaquetteWinding.hpp:24-51, on the device -------------------------
__device__ __forceinline__ double wrap_angle_diff_d(double delta)
{
    const double pi = 3.14159265358979323846;
    const double two_pi = 2.0 * pi;
    // fmod(delta, 2 pi) returns delta itself when |delta| < 2 pi, always the
    // case for differences of two atan2 results except the +-2 pi corner:
    // only then is the (slow, iterative) fmod evaluated.
    if (!(fabs(delta) < two_pi))
        delta = fmod(delta, two_pi);
    if (delta <= -pi)
        delta += two_pi;
    else if (delta > pi)
        delta -= two_pi;
    return delta;
}

__device__ __forceinline__ int plaquette_winding_d(double t1, double t2,
                                                   double t3, double t4)
{
#pragma clang fp contract(off)
    const double two_pi = 2.0 * 3.14159265358979323846;
    const double total = wrap_angle_diff_d(t2 - t1) + wrap_angle_diff_d(t3 - t2) +
                         wrap_angle_diff_d(t4 - t3) + wrap_angle_diff_d(t1 - t4);
    return (int)lround(total / two_pi);
}

// FourthOrderDerivatives diff1 / diff2 (published 4th-order forms)
__device__ __forceinline__ double d1_ref(double fm2, double fm1, double fp1,
                                         double fp2, double i1)
{
#pragma clang fp contract(off)
    const double w_far = 8.33333333333333333333e-2;
    const double w_near = 6.66666666666666666667e-1;
    return (w_far * fm2 - w_near * fm1 + w_near * fp1 - w_far * fp2) * i1;
}
__device__ __forceinline__ double d2_ref(double fm2, double fm1, double f0,
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

struct StarLoader
{
    int nx, ny, nz;
    uint32_t plane;
    template <typename T>
    __device__ __forceinline__ double at(const T *c, const float *lo, int x,
                                         int y, int z) const
    {
        const uint32_t i = (uint32_t)wrapc(z, nz) * plane +
                           (uint32_t)wrapc(y, ny) * nx + wrapc(x, nx);
        return ld(c, lo, i);
    }
};

// --------------------------------------------------- windings / tagging --
struct WindParams
{
    int nx, ny, nz;
    uint32_t n;
    int z_chunk, tiles_x, tiles_y, tiles_z;
    const uint8_t *mask;    // optional: skip cells with mask == 0
    int8_t *wfield;         // optional: per-cell windings
    uint8_t *tags;          // optional: tagger output (only sets)
    unsigned long long *counts; // [0] plain, [1] weighted
    // tagger extras
    double dx, grad_thr, rgrad_thr, amp_thr, R;
    int use_lap, use_rgrad;
    // string velocities at pierced corners (VelocityKernel.hpp:37-141)
    int do_vel;
    double vel_R, vel_RpR, vel_mr;
    double *vel_partials; // [gridDim.x]
};

// gamma_sq_v_sq_from_phi, Velocity.hpp:62-78
__device__ __forceinline__ double gamma_sq_v_sq_d(double phi1, double phi2,
                                                  double pd1, double pd2,
                                                  double m_r)
{
    const double c1 = 0.41222;
    const double e1 = -0.025763;
    const double phi_sq = phi1 * phi1 + phi2 * phi2;
    const double phi_dot_sq = pd1 * pd1 + pd2 * pd2;
    const double re = phi1 * pd1 + phi2 * pd2;
    const double m_r_sq = m_r * m_r;
    const double c1_sq = c1 * c1;
    const double term1 = (phi_dot_sq / (2.0 * m_r_sq * c1_sq)) *
                         (1.0 - e1 * phi_sq / (c1 * c1_sq));
    const double term2 = (e1 / (4.0 * m_r_sq * c1_sq * c1_sq * c1)) * re * re;
    return term1 - term2;
}

constexpr int WTY = 8;
// what a windings pass produces besides the counts (compile-time, so the
// plain count pass carries none of the other paths' registers)
enum
{
    WM_VEL = 1,   // string velocities at pierced corners
    WM_TAG = 2,   // tagger output
    WM_TAGD = 4,  // tagger derivative criteria (Laplacian, radial gradient)
    WM_FIELD = 8, // per-cell windings
    WM_ALL = 15
};

// Each workgroup owns a 64 x WTY column tile and streams a chunk of z: the
// phases atan2(psi2, psi1) of a plane (tile + one halo column / row) go
// into a 3-deep LDS ring, each computed once per vertex; the raw psi of the
// plane after next is loaded into registers one plane ahead, so its memory
// latency overlaps the windings of the current plane.
template <typename T, int WM>
__global__ void __launch_bounds__(64 * WTY)
    k_windings(const T *y, const float *lo, WindParams p)
{
    constexpr int TX = 64;
    __shared__ double th[3][WTY + 1][TX + 1];

    const int bid = blockIdx.x;
    const int bx = bid % p.tiles_x;
    const int rem = bid / p.tiles_x;
    const int by = rem % p.tiles_y;
    const int bz = rem / p.tiles_y;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t n = p.n;
    const uint32_t plane = (uint32_t)nx * ny;
    const int gx = bx * TX + tx, gy = by * WTY + ty;
    const bool active = gx < nx && gy < ny;
    const int xw = gx % nx, yw = gy % ny;
    const int xh = (bx * TX + TX) % nx;         // x halo column
    const int yh = (by * WTY + WTY) % ny;       // y halo row
    const T *p1 = y;
    const T *p2 = y + n;
    const float *l1 = lo, *l2 = lo ? lo + n : nullptr;
    const bool hx = tx == 0, hy = ty == 0;

    // raw psi of the points this thread stages: own cell, x halo, y halo,
    // corner
    double r[4][4]; // psi1, psi2 (+ Pi1, Pi2 for the velocities)
    // gamma^2 v^2 of every vertex of the ring planes (velocity passes)
    __shared__ double gv[(WM & WM_VEL) ? 3 : 1][WTY + 1][TX + 1];
    const T *q1 = y + 2 * (size_t)n, *q2 = y + 3 * (size_t)n;
    const float *m1 = lo ? lo + 2 * (size_t)n : nullptr;
    const float *m2 = lo ? lo + 3 * (size_t)n : nullptr;
    auto raw = [&](int xx, int yy, int zz, double(&v)[4]) {
        const uint32_t i = (uint32_t)zz * plane + (uint32_t)yy * nx + xx;
        v[0] = ld(p1, l1, i);
        v[1] = ld(p2, l2, i);
        if constexpr ((WM & WM_VEL) != 0)
        {
            v[2] = ld(q1, m1, i);
            v[3] = ld(q2, m2, i);
        }
    };
    // gamma^2 v^2 at one plaquette corner, VelocityKernel.hpp:82-96
    auto corner_of = [&](const double(&v)[4]) -> double {
        const double R = p.vel_R, RpR = p.vel_RpR;
        return gamma_sq_v_sq_d(v[0] / R, v[1] / R, (v[2] - RpR * v[0]) / (R * R),
                               (v[3] - RpR * v[1]) / (R * R), p.vel_mr);
    };
    auto load_plane = [&](int zz) {
        raw(xw, yw, zz, r[0]);
        if (hx)
            raw(xh, yw, zz, r[1]);
        if (hy)
        {
            raw(xw, yh, zz, r[2]);
            if (hx)
                raw(xh, yh, zz, r[3]);
        }
    };
    auto store_plane = [&](int s) {
        th[s][ty][tx] = atan2(r[0][1], r[0][0]);
        if (hx)
            th[s][ty][TX] = atan2(r[1][1], r[1][0]);
        if (hy)
        {
            th[s][WTY][tx] = atan2(r[2][1], r[2][0]);
            if (hx)
                th[s][WTY][TX] = atan2(r[3][1], r[3][0]);
        }
        if constexpr ((WM & WM_VEL) != 0)
        {
            if (p.do_vel)
            {
                gv[s][ty][tx] = corner_of(r[0]);
                if (hx)
                    gv[s][ty][TX] = corner_of(r[1]);
                if (hy)
                {
                    gv[s][WTY][tx] = corner_of(r[2]);
                    if (hx)
                        gv[s][WTY][TX] = corner_of(r[3]);
                }
            }
        }
    };

    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);
    load_plane(z_begin);
    store_plane(z_begin % 3);
    load_plane(wrapc(z_begin + 1, nz));

    unsigned long long plain_w = 0; // per wavefront (uniform)
    unsigned long long weighted = 0; // per lane
    unsigned long long vcount = 0;   // per lane
    double vsum = 0.0;               // per lane
    const StarLoader S{nx, ny, nz, plane};
    const double i1 = 1.0 / p.dx, i2 = 1.0 / (p.dx * p.dx);

    for (int z = z_begin; z < z_end; ++z)
    {
        const int s0 = z % 3, s1 = (z + 1) % 3;
        store_plane(s1);
        if (z + 1 < z_end)
            load_plane(wrapc(z + 2, nz));
        __syncthreads();
        int wxy = 0, wyz = 0, wzx = 0;
        bool in = active;
        const uint32_t id = (uint32_t)z * plane + (uint32_t)gy * nx + gx;
        if (in && p.mask && p.mask[id] == 0)
            in = false;
        if (in)
        {
            const double t000 = th[s0][ty][tx];
            const double t100 = th[s0][ty][tx + 1];
            const double t110 = th[s0][ty + 1][tx + 1];
            const double t010 = th[s0][ty + 1][tx];
            const double t001 = th[s1][ty][tx];
            const double t101 = th[s1][ty][tx + 1];
            const double t011 = th[s1][ty + 1][tx];
            wxy = plaquette_winding_d(t000, t100, t110, t010);
            wyz = plaquette_winding_d(t000, t010, t011, t001);
            wzx = plaquette_winding_d(t000, t001, t101, t100);
            weighted += (unsigned long long)(abs(wxy) + abs(wyz) + abs(wzx));
            if constexpr ((WM & WM_VEL) != 0)
            {
                if (p.do_vel && (wxy != 0 || wyz != 0 || wzx != 0))
                {
                    // corners from the ring: s0 = plane z, s1 = plane z + 1
                    const double c000 = gv[s0][ty][tx];
                    if (wxy != 0)
                    {
                        vsum += c000 + gv[s0][ty][tx + 1] +
                                gv[s0][ty + 1][tx + 1] + gv[s0][ty + 1][tx];
                        vcount += 4;
                    }
                    if (wyz != 0)
                    {
                        vsum += c000 + gv[s0][ty + 1][tx] +
                                gv[s1][ty + 1][tx] + gv[s1][ty][tx];
                        vcount += 4;
                    }
                    if (wzx != 0)
                    {
                        vsum += c000 + gv[s1][ty][tx] + gv[s1][ty][tx + 1] +
                                gv[s0][ty][tx + 1];
                        vcount += 4;
                    }
                }
            }
            if constexpr ((WM & WM_FIELD) != 0)
            {
                if (p.wfield)
                {
                p.wfield[3 * (size_t)id] = (int8_t)wxy;
                p.wfield[3 * (size_t)id + 1] = (int8_t)wyz;
                p.wfield[3 * (size_t)id + 2] = (int8_t)wzx;
                }
            }
            if constexpr ((WM & WM_TAG) != 0)
            {
                if (p.tags)
                {
                bool tag = (wxy != 0 || wyz != 0 || wzx != 0);
                const double psi1 = ld(p1, l1, id), psi2 = ld(p2, l2, id);
                if (!tag && p.amp_thr > 0.0 &&
                    sqrt(psi1 * psi1 + psi2 * psi2) / p.R < p.amp_thr)
                    tag = true;
                if constexpr ((WM & WM_TAGD) != 0)
                {
                    if (!tag && p.use_lap)
                    {
                        const double dx2 = p.dx * p.dx;
                        double lap[2];
                        for (int c = 0; c < 2; ++c)
                        {
                            const T *q = c ? p2 : p1;
                            const float *ql = c ? l2 : l1;
                            const double f0 = ld(q, ql, id);
                            double rr = 0.0;
                            rr += d2_ref(S.at(q, ql, gx - 2, gy, z),
                                         S.at(q, ql, gx - 1, gy, z), f0,
                                         S.at(q, ql, gx + 1, gy, z),
                                         S.at(q, ql, gx + 2, gy, z), i2);
                            rr += d2_ref(S.at(q, ql, gx, gy - 2, z),
                                         S.at(q, ql, gx, gy - 1, z), f0,
                                         S.at(q, ql, gx, gy + 1, z),
                                         S.at(q, ql, gx, gy + 2, z), i2);
                            rr += d2_ref(S.at(q, ql, gx, gy, z - 2),
                                         S.at(q, ql, gx, gy, z - 1), f0,
                                         S.at(q, ql, gx, gy, z + 1),
                                         S.at(q, ql, gx, gy, z + 2), i2);
                            lap[c] = rr;
                        }
                        if (dx2 * fabs(lap[0]) > p.grad_thr ||
                            dx2 * fabs(lap[1]) > p.grad_thr)
                            tag = true;
                    }
                    if (!tag && p.use_rgrad)
                    {
                        const double psq = psi1 * psi1 + psi2 * psi2;
                        double g1sq = 0.0, g2sq = 0.0, gtsq = 0.0;
                        for (int d = 0; d < 3; ++d)
                        {
                            const int ex = d == 0, ey = d == 1, ez = d == 2;
                            const double a = d1_ref(
                                S.at(p1, l1, gx - 2 * ex, gy - 2 * ey, z - 2 * ez),
                                S.at(p1, l1, gx - ex, gy - ey, z - ez),
                                S.at(p1, l1, gx + ex, gy + ey, z + ez),
                                S.at(p1, l1, gx + 2 * ex, gy + 2 * ey, z + 2 * ez),
                                i1);
                            const double b = d1_ref(
                                S.at(p2, l2, gx - 2 * ex, gy - 2 * ey, z - 2 * ez),
                                S.at(p2, l2, gx - ex, gy - ey, z - ez),
                                S.at(p2, l2, gx + ex, gy + ey, z + ez),
                                S.at(p2, l2, gx + 2 * ex, gy + 2 * ey, z + 2 * ez),
                                i1);
                            g1sq += a * a;
                            g2sq += b * b;
                            const double dth =
                                psq > 0.0 ? (psi1 * b - psi2 * a) / psq : 0.0;
                            gtsq += dth * dth;
                        }
                        if (p.dx * sqrt(g1sq + g2sq - psq * gtsq) > p.rgrad_thr)
                            tag = true;
                    }
                }
                if (tag)
                    p.tags[id] = 1;
                }
            }
        }
        // pierced plaquettes of this wavefront, by ballot / popcount
        plain_w += __popcll(__ballot(wxy != 0)) + __popcll(__ballot(wyz != 0)) +
                   __popcll(__ballot(wzx != 0));
    }
    // wavefront reduction of the per-lane counts
    for (int o = 32; o > 0; o >>= 1)
    {
        weighted += __shfl_xor(weighted, o, 64);
        if constexpr ((WM & WM_VEL) != 0)
            vcount += __shfl_xor(vcount, o, 64);
    }
    if ((threadIdx.x & 63) == 0)
    {
        atomicAdd(&p.counts[0], plain_w);
        atomicAdd(&p.counts[1], weighted);
        if constexpr ((WM & WM_VEL) != 0)
            atomicAdd(&p.counts[2], vcount);
    }
    if constexpr ((WM & WM_VEL) != 0)
    {
        // fixed-order block reduction of the velocity sum -> one partial
        __shared__ double red[64 * WTY];
        const int t = threadIdx.y * 64 + threadIdx.x;
        red[t] = vsum;
        __syncthreads();
        for (int o = 32 * WTY; o > 0; o >>= 1)
        {
            if (t < o)
                red[t] += red[t + o];
            __syncthreads();
        }
        if (t == 0)
            p.vel_partials[blockIdx.x] = red[0];
    }
}

template <typename T, int WM>
static void launch_windings_t(int ntiles, hipStream_t s, const T *y,
                              const float *lo, const WindParams &p)
{
    hipLaunchKernelGGL((k_windings<T, WM>), dim3(ntiles), dim3(64, WTY), 0, s,
                       y, lo, p);
}

template <typename T>
static void launch_windings(int wm, int ntiles, hipStream_t s, const T *y,
                            const float *lo, const WindParams &p)
{
    switch (wm)
    {
    case 0:
        return launch_windings_t<T, 0>(ntiles, s, y, lo, p);
    case WM_VEL:
        return launch_windings_t<T, WM_VEL>(ntiles, s, y, lo, p);
    case WM_FIELD:
        return launch_windings_t<T, WM_FIELD>(ntiles, s, y, lo, p);
    case WM_TAG:
        return launch_windings_t<T, WM_TAG>(ntiles, s, y, lo, p);
    case WM_TAG | WM_TAGD:
        return launch_windings_t<T, WM_TAG | WM_TAGD>(ntiles, s, y, lo, p);
    default:
        return launch_windings_t<T, WM_ALL>(ntiles, s, y, lo, p);
    }
}

__global__ void k_zero_counts(unsigned long long *c)
{
    if (threadIdx.x < 8)
        c[threadIdx.x] = 0ull;
}

// ------------------------------------------------------------- energies --
struct EnergyParams
{
    int nx, ny, nz;
    uint32_t n;
    double dx, R, lambda, RpR;
    int scheme;
    double thr;
    const uint8_t *mask;
    double *partials; // [gridDim.x][14]
    int z_chunk, tiles_x, tiles_y, tiles_z;
};

__device__ __forceinline__ double masking_weight_d(int scheme, double thr,
                                                   double psi1, double psi2,
                                                   double R)
{
    if (scheme == AXS_MASK_NONE)
        return 1.0;
    const double mod = sqrt(psi1 * psi1 + psi2 * psi2) / R;
    if (scheme == AXS_MASK_B)
        return (mod >= thr) ? 1.0 : 0.0;
    const double f_a = 1.4142135623730951;
    const double r = mod - 1.0;
    const double f = 1.0 + r / f_a;
    return f * f;
}

// masked_a_dot, Masking.hpp:96-118 (contraction off: the numerator
// psi1 Pi2 - Pi1 psi2 cancels, so it is rounded exactly as written)
__device__ __forceinline__ double masked_a_dot_d(int scheme, double thr,
                                                 double psi1, double psi2,
                                                 double Pi1, double Pi2,
                                                 double R)
{
#pragma clang fp contract(off)
    const double raw = psi1 * Pi2 - Pi1 * psi2;
    if (scheme == AXS_MASK_A)
        return raw;
    if (scheme == AXS_MASK_B &&
        masking_weight_d(scheme, thr, psi1, psi2, R) == 0.0)
        return 0.0;
    return raw / (psi1 * psi1 + psi2 * psi2);
}

constexpr int EB = 256;
constexpr int ETY = 8; // energy tile: 64 x ETY cells, one wavefront per row

// compute_total_energy_sums (EnergyKernel.hpp:42-339) per cell, Energy.hpp
// pointwise forms.  Each workgroup owns a 64 x ETY column tile and streams
// a chunk of z planes: psi1/psi2 (FP64, FP32C as hi + lo) of the current
// plane sit in LDS with a 2-cell x/y halo, the z neighbours in a 5-deep
// register queue, so each psi value is read from HBM once per plane
// instead of once per stencil use.  Per cell the 14 terms are formed
// exactly as the direct form (same d1 operand order); only the order of the
// (reproducible, fixed-order) sums differs.
template <typename T>
__global__ void __launch_bounds__(64 * ETY)
    k_energy(const T *y, const float *lo, EnergyParams p)
{
    __shared__ double sp[2][ETY + 4][64 + 4];
    __shared__ double red[64 * ETY];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int bid = blockIdx.x;
    const int bx = bid % p.tiles_x;
    const int rem = bid / p.tiles_x;
    const int by = rem % p.tiles_y;
    const int bz = rem / p.tiles_y;
    const uint32_t n = p.n;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t plane = (uint32_t)nx * ny;
    const int gx = bx * 64 + tx, gy = by * ETY + ty;
    const bool active = gx < nx && gy < ny;
    const int xw = gx % nx, yw = gy % ny;
    const uint32_t rowc = (uint32_t)yw * nx + xw;
    const T *p1 = y, *p2 = y + n, *q1 = y + 2 * (size_t)n, *q2 = y + 3 * (size_t)n;
    const float *l1 = lo, *l2 = lo ? lo + n : nullptr;
    const float *m1 = lo ? lo + 2 * (size_t)n : nullptr;
    const float *m2 = lo ? lo + 3 * (size_t)n : nullptr;
    // halo roles: lanes 0,1 / 62,63 load the x halo, rows 0,1 / ETY-2,ETY-1
    // the y halo (ETY >= 4, so no row holds both y-halo roles)
    const bool hxl = tx < 2, hxr = tx >= 62;
    const bool hyl = ty < 2, hyr = ty >= ETY - 2;
    uint32_t rowhx = 0, rowhy = 0;
    int sx = 0, sy = 0;
    if (hxl || hxr)
    {
        const int hx = hxl ? bx * 64 - 2 + tx : bx * 64 + 64 + (tx - 62);
        rowhx = (uint32_t)yw * nx + (uint32_t)(((hx % nx) + nx) % nx);
        sx = hxl ? tx : 64 + 2 + (tx - 62);
    }
    if (hyl || hyr)
    {
        const int hy = hyl ? by * ETY - 2 + ty : by * ETY + ETY + (ty - (ETY - 2));
        rowhy = (uint32_t)(((hy % ny) + ny) % ny) * nx + xw;
        sy = hyl ? ty : ETY + 2 + (ty - (ETY - 2));
    }
    const double i1 = 1.0 / p.dx;
    const double R = p.R, R2 = R * R, R4 = R2 * R2;
    const double f_a = 1.4142135623730951;
    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);
    auto zoff = [&](int z) -> uint32_t {
        return (uint32_t)wrapc(wrapc(z, nz), nz) * plane;
    };
    double a1[5], a2[5];
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        const uint32_t o = zoff(z_begin - 2 + i) + rowc;
        a1[i + 1] = ld(p1, l1, o);
        a2[i + 1] = ld(p2, l2, o);
    }
    double acc[14];
#pragma unroll
    for (int q = 0; q < 14; ++q)
        acc[q] = 0.0;
    // halo psi and centre Pi of the next plane, loaded one plane ahead
    double hxv[2] = {0.0, 0.0}, hyv[2] = {0.0, 0.0}, piv[2] = {0.0, 0.0};
    auto load_next = [&](int z) {
        const uint32_t oz = zoff(z);
        if (hxl || hxr)
        {
            hxv[0] = ld(p1, l1, oz + rowhx);
            hxv[1] = ld(p2, l2, oz + rowhx);
        }
        if (hyl || hyr)
        {
            hyv[0] = ld(p1, l1, oz + rowhy);
            hyv[1] = ld(p2, l2, oz + rowhy);
        }
        if (active)
        {
            piv[0] = ld(q1, m1, oz + rowc);
            piv[1] = ld(q2, m2, oz + rowc);
        }
    };
    load_next(z_begin);
    for (int z = z_begin; z < z_end; ++z)
    {
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            a1[i] = a1[i + 1];
            a2[i] = a2[i + 1];
        }
        const uint32_t oz = zoff(z);
        {
            const uint32_t o = zoff(z + 2) + rowc;
            a1[4] = ld(p1, l1, o);
            a2[4] = ld(p2, l2, o);
        }
        sp[0][ty + 2][tx + 2] = a1[2];
        sp[1][ty + 2][tx + 2] = a2[2];
        if (hxl || hxr)
        {
            sp[0][ty + 2][sx] = hxv[0];
            sp[1][ty + 2][sx] = hxv[1];
        }
        if (hyl || hyr)
        {
            sp[0][sy][tx + 2] = hyv[0];
            sp[1][sy][tx + 2] = hyv[1];
        }
        const double Pi1 = piv[0], Pi2 = piv[1];
        if (z + 1 < z_end)
            load_next(z + 1);
        __syncthreads();
        const uint32_t id = oz + rowc;
        if (active && !(p.mask && p.mask[id] == 0))
        {
            const double psi1 = a1[2], psi2 = a2[2];
            const double psi_sq = psi1 * psi1 + psi2 * psi2;
            double g1sq = 0.0, g2sq = 0.0, gtsq = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                double a, b;
                if (d == 0)
                {
                    const double *r1 = &sp[0][ty + 2][tx + 2];
                    const double *r2 = &sp[1][ty + 2][tx + 2];
                    a = d1_ref(r1[-2], r1[-1], r1[1], r1[2], i1);
                    b = d1_ref(r2[-2], r2[-1], r2[1], r2[2], i1);
                }
                else if (d == 1)
                {
                    a = d1_ref(sp[0][ty][tx + 2], sp[0][ty + 1][tx + 2],
                               sp[0][ty + 3][tx + 2], sp[0][ty + 4][tx + 2], i1);
                    b = d1_ref(sp[1][ty][tx + 2], sp[1][ty + 1][tx + 2],
                               sp[1][ty + 3][tx + 2], sp[1][ty + 4][tx + 2], i1);
                }
                else
                {
                    a = d1_ref(a1[0], a1[1], a1[3], a1[4], i1);
                    b = d1_ref(a2[0], a2[1], a2[3], a2[4], i1);
                }
                g1sq += a * a;
                g2sq += b * b;
                const double dth =
                    (psi_sq > 0.0) ? (psi1 * b - psi2 * a) / psi_sq : 0.0;
                gtsq += dth * dth;
            }
            // Energy.hpp:200-325
            const double c1 = Pi1 - p.RpR * psi1, c2 = Pi2 - p.RpR * psi2;
            const double dpot = psi1 * psi1 + psi2 * psi2 - R * R;
            const double pot = 0.25 * p.lambda * dpot * dpot;
            const double rho = (c1 * c1 + c2 * c2 + (g1sq + g2sq) + pot) / R4;
            const double tp =
                (psi_sq > 0.0) ? (psi1 * Pi2 - Pi1 * psi2) / psi_sq : 0.0;
            const double a_dot = f_a * tp / R;
            const double rak = 0.5 * a_dot * a_dot;
            const double rag = 0.5 * f_a * f_a * gtsq / R2;
            double rrk = 0.0;
            if (psi_sq > 0.0)
            {
                const double rd = (psi1 * c1 + psi2 * c2) / sqrt(psi_sq);
                rrk = (rd * rd) / R4;
            }
            const double rrg = (g1sq + g2sq - psi_sq * gtsq) / R4;
            const double rrm = pot / R4;
            const double w = masking_weight_d(p.scheme, p.thr, psi1, psi2, R);
            acc[0] += rho;
            acc[1] += rho * w;
            acc[2] += rak;
            acc[3] += rak * w;
            acc[4] += rag;
            acc[5] += rag * w;
            acc[6] += rrk;
            acc[7] += rrk * w;
            acc[8] += rrg;
            acc[9] += rrg * w;
            acc[10] += rrm;
            acc[11] += rrm * w;
            acc[12] += w;
            acc[13] += 1.0;
        }
        __syncthreads();
    }
    // fixed-order block reduction of each sum -> one row of partials
    const int t = ty * 64 + tx;
    for (int q = 0; q < 14; ++q)
    {
        red[t] = acc[q];
        __syncthreads();
        for (int o = 32 * ETY; o > 0; o >>= 1)
        {
            if (t < o)
                red[t] += red[t + o];
            __syncthreads();
        }
        if (t == 0)
            p.partials[(size_t)blockIdx.x * 14 + q] = red[0];
        __syncthreads();
    }
}

// Sums nrow rows of `width` partials column-wise, in fixed order.
__global__ void k_sum_partials(const double *part, int nrow, int width,
                               double *out)
{
    __shared__ double red[256];
    const int col = blockIdx.x;
    double s = 0.0;
    for (int r = threadIdx.x; r < nrow; r += 256)
        s += part[(size_t)r * width + col];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int)threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[col] = red[0];
}

// ------------------------------------------------------------ field fill --
template <typename T, typename B>
__global__ void k_fill(const T *y, const float *lo, uint32_t n, int kind,
                       int scheme, double thr, double R, B *buf)
{
    for (uint32_t id = blockIdx.x * 256 + threadIdx.x; id < n;
         id += gridDim.x * 256)
    {
        double v;
        if (kind == AXS_FIELD_MASKED_A_DOT)
        {
            // masked_a_dot, Masking.hpp:96-118
            const double psi1 = ld(y, lo, id);
            const double psi2 = ld(y + n, lo ? lo + n : nullptr, id);
            const double Pi1 = ld(y + 2 * n, lo ? lo + 2 * n : nullptr, id);
            const double Pi2 = ld(y + 3 * n, lo ? lo + 3 * n : nullptr, id);
            v = masked_a_dot_d(scheme, thr, psi1, psi2, Pi1, Pi2, R);
        }
        else
        {
            const uint32_t off = (uint32_t)(kind + 1) * n;
            v = ld(y + off, lo ? lo + off : nullptr, id);
        }
        buf[id] = (B)v;
    }
}

static int nblocks(size_t n, int bs, int cap)
{
    size_t b = (n + bs - 1) / bs;
    if (b > (size_t)cap)
        b = cap;
    return b < 1 ? 1 : (int)b;
}

// shared launcher for windings / tags
struct VelArgs
{
    double R, RpR, m_r;
    double *sum;
    int64_t *count;
};

static int run_windings(axs_level *L, const uint8_t *mask, int8_t *wfield,
                        uint8_t *tags, const axs_tagger_params *tp,
                        int64_t out[2], const VelArgs *vel = nullptr)
{
    hipStream_t s = L->ctx->stream;
    WindParams p{};
    p.nx = L->g.nx;
    p.ny = L->g.ny;
    p.nz = L->g.nz;
    p.n = L->g.ncell;
    p.mask = mask;
    p.wfield = wfield;
    p.tags = tags;
    p.counts = L->ctx->d_counts;
    p.dx = L->dx;
    p.grad_thr = tp ? tp->gradient_threshold : DBL_MAX;
    p.rgrad_thr = tp ? tp->radial_gradient_threshold : DBL_MAX;
    p.amp_thr = tp ? tp->amplitude_threshold : 0.0;
    p.R = tp ? tp->R : 1.0;
    p.use_lap = (tp && tp->gradient_threshold < DBL_MAX) ? 1 : 0;
    p.use_rgrad = (tp && tp->radial_gradient_threshold < DBL_MAX) ? 1 : 0;
    p.tiles_x = (p.nx + 63) / 64;
    p.tiles_y = (p.ny + WTY - 1) / WTY;
    int zc = 32;
    while (zc > 2 && (long)p.tiles_x * p.tiles_y * ((p.nz + zc - 1) / zc) < 4096)
        zc >>= 1;
    if (zc > p.nz)
        zc = p.nz;
    p.z_chunk = zc;
    p.tiles_z = (p.nz + zc - 1) / zc;
    const int ntiles = p.tiles_x * p.tiles_y * p.tiles_z;
    if (vel)
    {
        AXS_CHECK_ARG(ntiles <= kMaxPartials * 16,
                      "axs_velocity_at_pierced_corners: grid too large");
        p.do_vel = 1;
        p.vel_R = vel->R;
        p.vel_RpR = vel->RpR;
        p.vel_mr = vel->m_r;
        p.vel_partials = L->ctx->d_partials;
    }
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, s, p.counts);
    const int wm = (p.do_vel ? WM_VEL : 0) | (tags ? WM_TAG : 0) |
                   ((tags && (p.use_lap || p.use_rgrad)) ? WM_TAGD : 0) |
                   (wfield ? WM_FIELD : 0);
    if (L->precision == AXS_FP64)
        launch_windings<double>(wm, ntiles, s, (const double *)L->y,
                                (const float *)nullptr, p);
    else
        launch_windings<float>(wm, ntiles, s, (const float *)L->y,
                               L->precision == AXS_FP32C
                                   ? (const float *)L->comp
                                   : (const float *)nullptr,
                               p);
    AXS_HIP_TRY(hipGetLastError());
    if (vel)
    {
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s,
                           (const double *)p.vel_partials, ntiles, 1,
                           L->ctx->d_result);
        AXS_HIP_TRY(hipGetLastError());
        AXS_HIP_TRY(hipMemcpyAsync(L->ctx->h_result, L->ctx->d_result,
                                   sizeof(double), hipMemcpyDeviceToHost, s));
    }
    if (out || vel)
    {
        AXS_HIP_TRY(hipMemcpyAsync(L->ctx->h_counts, p.counts,
                                   3 * sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, s));
        AXS_HIP_TRY(hipStreamSynchronize(s));
        if (out)
        {
            out[0] = (int64_t)L->ctx->h_counts[0];
            out[1] = (int64_t)L->ctx->h_counts[1];
        }
        if (vel)
        {
            *vel->sum = L->ctx->h_result[0];
            *vel->count = (int64_t)L->ctx->h_counts[2];
        }
    }
    return AXS_OK;
}

} // namespace axs

using namespace axs;

extern "C" {

int axs_count_plaquettes(axs_level *L, const uint8_t *mask, int64_t out[2])
{
    AXS_CHECK_ARG(L && out, "axs_count_plaquettes: NULL argument");
    return run_windings(L, mask, nullptr, nullptr, nullptr, out);
}

int axs_velocity_at_pierced_corners(axs_level *L, double R, double RpR,
                                    double m_r, const uint8_t *mask,
                                    double *sum_gamma_sq_v_sq, int64_t *count,
                                    int64_t counts[2])
{
    AXS_CHECK_ARG(L && sum_gamma_sq_v_sq && count,
                  "axs_velocity_at_pierced_corners: NULL argument");
    AXS_CHECK_ARG(R > 0.0 && m_r > 0.0,
                  "axs_velocity_at_pierced_corners: R and m_r must be > 0");
    VelArgs v{R, RpR, m_r, sum_gamma_sq_v_sq, count};
    int64_t c[2];
    return run_windings(L, mask, nullptr, nullptr, nullptr, counts ? counts : c,
                        &v);
}

int axs_windings_field(axs_level *L, int8_t *w)
{
    AXS_CHECK_ARG(L && w, "axs_windings_field: NULL argument");
    int64_t c[2];
    return run_windings(L, nullptr, w, nullptr, nullptr, c);
}

int axs_tag_cells(axs_level *L, const axs_tagger_params *tp, uint8_t *tags,
                  int64_t counts[2])
{
    AXS_CHECK_ARG(L && tp && tags, "axs_tag_cells: NULL argument");
    AXS_CHECK_ARG(tp->gradient_threshold > 0.0 &&
                      tp->radial_gradient_threshold > 0.0,
                  "axs_tag_cells: thresholds must be > 0");
    AXS_CHECK_ARG(tp->amplitude_threshold <= 0.0 || tp->R > 0.0,
                  "axs_tag_cells: amplitude criterion needs R > 0");
    int64_t c[2];
    int rc = run_windings(L, nullptr, nullptr, tags, tp, counts ? counts : c);
    return rc;
}

int axs_energy_sums(axs_level *L, double R, double lambda, double RpR,
                    const axs_masking *m, const uint8_t *mask, double out[14])
{
    AXS_CHECK_ARG(L && m && out, "axs_energy_sums: NULL argument");
    AXS_CHECK_ARG(R > 0.0, "axs_energy_sums: R must be > 0");
    AXS_CHECK_ARG(m->scheme >= 0 && m->scheme <= 2,
                  "axs_energy_sums: unknown masking scheme");
    hipStream_t s = L->ctx->stream;
    EnergyParams p{};
    p.nx = L->g.nx;
    p.ny = L->g.ny;
    p.nz = L->g.nz;
    p.n = L->g.ncell;
    p.dx = L->dx;
    p.R = R;
    p.lambda = lambda;
    p.RpR = RpR;
    p.scheme = m->scheme;
    p.thr = m->threshold;
    p.mask = mask;
    p.partials = L->ctx->d_partials;
    p.tiles_x = (p.nx + 63) / 64;
    p.tiles_y = (p.ny + ETY - 1) / ETY;
    // z chunk: enough tiles to fill the device, at most kMaxPartials rows
    int zc = 8;
    while ((long)p.tiles_x * p.tiles_y * ((p.nz + zc - 1) / zc) > kMaxPartials)
        zc *= 2;
    while (zc < 64 &&
           (long)p.tiles_x * p.tiles_y * ((p.nz + 2 * zc - 1) / (2 * zc)) >= 8192)
        zc *= 2;
    if (zc > p.nz)
        zc = p.nz;
    p.z_chunk = zc;
    p.tiles_z = (p.nz + zc - 1) / zc;
    const long nbl = (long)p.tiles_x * p.tiles_y * p.tiles_z;
    AXS_CHECK_ARG(nbl <= kMaxPartials, "axs_energy_sums: grid too large");
    const int nb = (int)nbl;
    if (L->precision == AXS_FP64)
        hipLaunchKernelGGL(k_energy<double>, dim3(nb), dim3(64, ETY), 0, s,
                           (const double *)L->y, (const float *)nullptr, p);
    else
        hipLaunchKernelGGL(k_energy<float>, dim3(nb), dim3(64, ETY), 0, s,
                           (const float *)L->y,
                           L->precision == AXS_FP32C ? (const float *)L->comp
                                                     : (const float *)nullptr,
                           p);
    AXS_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3(14), dim3(256), 0, s, p.partials,
                       nb, 14, L->ctx->d_result);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipMemcpyAsync(L->ctx->h_result, L->ctx->d_result,
                               14 * sizeof(double), hipMemcpyDeviceToHost, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    for (int q = 0; q < 14; ++q)
        out[q] = L->ctx->h_result[q];
    return AXS_OK;
}

int axs_fill_field(axs_level *L, int kind, const axs_masking *m, double R,
                   void *buf)
{
    AXS_CHECK_ARG(L && m && buf, "axs_fill_field: NULL argument");
    AXS_CHECK_ARG(kind >= 0 && kind <= 2, "axs_fill_field: unknown kind");
    AXS_CHECK_ARG(m->scheme >= 0 && m->scheme <= 2,
                  "axs_fill_field: unknown masking scheme");
    hipStream_t s = L->ctx->stream;
    const uint32_t n = L->g.ncell;
    const int nb = nblocks(n, 256, 8192);
    if (L->precision == AXS_FP64)
        hipLaunchKernelGGL((k_fill<double, double>), dim3(nb), dim3(256), 0, s,
                           (const double *)L->y, (const float *)nullptr, n, kind,
                           m->scheme, m->threshold, R, (double *)buf);
    else
        hipLaunchKernelGGL((k_fill<float, float>), dim3(nb), dim3(256), 0, s,
                           (const float *)L->y, (const float *)L->comp, n, kind,
                           m->scheme, m->threshold, R, (float *)buf);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

} // extern "C"
