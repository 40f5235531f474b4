// axs_diag.hip -- string finding, tagging and energy reductions on a uniform
// periodic level.
//
// Windings (StringFinder.hpp:45-122): psi (FP64, FP32C promoted as hi +
// lo) of a z plane sits in a 3-deep ring of LDS planes; each cell reads the
// 7 corners of its 3 low-corner plaquettes from it and forms the windings
// from sign tests that provably agree with the reference's FP64 atan2 phase
// arithmetic (StringFinder.hpp:49-53), falling back to that arithmetic for
// the near-antipodal edges where the two could differ.  Pierced-plaquette counts are formed per wavefront with
// __ballot / __popcll and one 64-bit integer atomic per wavefront, so the
// counts are exact and order independent.
//
// Energies (EnergyKernel.hpp:106-233): FP64 per-cell arithmetic, per-thread
// accumulation, then a fixed-order block tree and a fixed-order second pass
// over the block partials: the 14 sums are bitwise reproducible run to run.
#include "axs_internal.hpp"

#include <cfloat>
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>
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
// Tile of a workgroup: with 8 | ntiles, the consecutive tiles k (n/8) ..
// (k+1)(n/8) - 1 go to XCD k (workgroups are dealt to the 8 XCDs round
// robin), so tiles that share halo rows and columns run on one XCD at about
// the same time and meet in its L2; else the identity.
__device__ __forceinline__ int xcd_tile(int ntiles)
{
    const int b = (int)blockIdx.x;
    if (ntiles & 7)
        return b;
    return (b & 7) * (ntiles >> 3) + (b >> 3);
}

struct WindParams
{
    int nx, ny, nz;
    uint32_t n;
    int z_chunk, tiles_x, tiles_y, tiles_z;
    const uint8_t *mask;    // optional: skip cells with mask == 0
    int8_t *wfield;         // optional: per-cell windings
    uint8_t *tags;          // optional: tagger output (only sets)
    unsigned long long *counts; // [0] plain, [1] weighted
    // cadence passes (axs_rk4_advance_windings): counts go to
    // counts + 3 * (*seg), and the amplitude criterion takes R_tab[*seg]
    const int *seg;
    const double *R_tab;
    // tagger extras
    double dx, grad_thr, rgrad_thr, amp_thr, R;
    int use_lap, use_rgrad;
    // string velocities at pierced corners (VelocityKernel.hpp:37-141)
    int do_vel;
    double vel_RpR;
    // host-side reciprocals for gamma_sq_v_sq_k: 1/R, 1/R^2,
    // 1/(2 m_r^2 c1^2), e1/(c1 c1^2), e1/(4 m_r^2 c1^4 c1)
    double vel_iR, vel_iR2, vel_k1, vel_kc, vel_k3;
    double *vel_partials; // [gridDim.x]
};

// gamma_sq_v_sq_from_phi (Velocity.hpp:62-78) with its divisions by
// uniform constants taken as products with host-computed reciprocals (one
// rounding apart per term; the sums over corners move by ~1e-16 relative)
__device__ __forceinline__ double gamma_sq_v_sq_k(double phi1, double phi2,
                                                  double pd1, double pd2,
                                                  double k1, double kc,
                                                  double k3)
{
    const double phi_sq = phi1 * phi1 + phi2 * phi2;
    const double phi_dot_sq = pd1 * pd1 + pd2 * pd2;
    const double re = phi1 * pd1 + phi2 * pd2;
    return (phi_dot_sq * k1) * (1.0 - kc * phi_sq) - k3 * re * re;
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

// Winding of an edge a -> b, without the phases.  With theta = atan2(y, x)
// in [-pi, pi], the wrapped difference of StringFinder.hpp's
// plaquette_winding is (theta_b - theta_a) - 2 pi k, k = +1 when
// theta_b - theta_a > pi, -1 when it is <= -pi, else 0; round the sum over a
// closed plaquette the raw differences cancel, so the winding is -sum(k).
// k != 0 only when a and b lie on opposite sides of the x axis and the
// short arc between them crosses the negative x axis: k = +1 for a below, b
// above and a clockwise turn (cross(a, b) < 0), k = -1 for the mirror case.
// These sign tests agree with the FP64 phase arithmetic unless the edge is
// within ~1e-9 rad of antipodal: elsewhere the computed cross product has
// the sign of the exact one (its rounding error is ~1e-16 |a||b|) and the
// phase difference is >~1e-9 away from +-pi, far beyond the ~1e-16 rad
// error of atan2.  Near-antipodal edges (dot < 0 and |cross| <= 1e-9 |dot|,
// where |dot| ~ |a||b|), vertices on the x axis (y == +-0, where the sign of
// zero picks the phase +-pi) and non-finite or extreme values are flagged,
// and the caller then evaluates the plaquettes with atan2 exactly as the
// reference does.
struct Vtx
{
    double x, y;
};

__device__ __forceinline__ bool vtx_ok(const Vtx &v)
{
    const double m = fabs(v.x) + fabs(v.y);
    return v.y != 0.0 && m >= 1e-140 && m <= 1e140;
}

__device__ __forceinline__ int edge_k(const Vtx &a, const Vtx &b, bool &amb)
{
#pragma clang fp contract(off)
    // branch free: every lane evaluates the same instructions
    const bool sa = a.y < 0.0, sb = b.y < 0.0;
    const double d = a.x * b.x + a.y * b.y;
    const double c = a.x * b.y - a.y * b.x;
    // d < 0 and |c| <= 1e-9 |d|, i.e. |c| + 1e-9 d <= 0
    amb = amb || fma(1e-9, d, fabs(c)) <= 0.0;
    const int up = (sa && !sb && c < 0.0) ? 1 : 0;
    const int down = (!sa && sb && c > 0.0) ? 1 : 0;
    return up - down;
}

// The reference's phase arithmetic for the 3 plaquettes of a cell (flagged
// cells only).  Inlined: as a call, the values live across it would be kept
// in scratch memory on the common path.
__device__ __forceinline__ void windings_by_phase(const Vtx *v, int *w)
{
    // v: 000, 100, 110, 010, 001, 101, 011
    double t[7];
    for (int i = 0; i < 7; ++i)
        t[i] = atan2(v[i].y, v[i].x);
    w[0] = plaquette_winding_d(t[0], t[1], t[2], t[3]);
    w[1] = plaquette_winding_d(t[0], t[3], t[6], t[4]);
    w[2] = plaquette_winding_d(t[0], t[4], t[5], t[1]);
}

// Each workgroup owns a 64 x WTY column tile and streams a chunk of z: psi
// (and Pi for the velocity pass; FP64, FP32C as hi + lo) of a plane -- the
// tile plus one halo column and row -- goes into a 3-deep LDS ring; each
// cell reads the 7 corners of its 3 low-corner plaquettes from it and forms
// the windings from the edge sign tests above (atan2 only for flagged
// cells).  The stored words of the next WD planes are in flight in
// registers while a plane is processed (each thread loads its own cell and
// at most one halo point per plane), so enough bytes are outstanding to
// stream at HBM rate; velocities read the Pi of pierced corners from the
// ring.
template <typename T, int WM, bool HASLO>
__global__ void __launch_bounds__(64 * WTY)
    __attribute__((amdgpu_waves_per_eu((WM & ~WM_FIELD) == 0 ? 4 : 1)))
    k_windings(const T *y, const float *lo, WindParams p)
{
    constexpr int TX = 64;
    constexpr bool VEL = (WM & WM_VEL) != 0;
    constexpr int NC = VEL ? 4 : 2; // components staged per point
    constexpr int WD = VEL ? 2 : 3; // planes of loads in flight
    __shared__ double ring[NC][3][WTY + 1][TX + 1];
    // per vertex: k + 1 of its +x edge (bits 0-1) and +y edge (bits 2-3) in
    // the plane, flags of those edges (bits 4, 5) and of the vertex (bit 6);
    // planes z and z + 1.  Per vertex of plane z: k + 1 of its +z edge
    // (bits 0-1) and that edge's flag (bit 2).
    __shared__ uint32_t pw[2][WTY + 1][TX + 1];
    __shared__ uint32_t zw[WTY + 1][TX + 1];

    const int bid = xcd_tile(p.tiles_x * p.tiles_y * p.tiles_z);
    const int bx = bid % p.tiles_x;
    const int rem = bid / p.tiles_x;
    const int by = rem % p.tiles_y;
    const int bz = rem / p.tiles_y;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int t = ty * TX + tx;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t n = p.n;
    const uint32_t plane = (uint32_t)nx * ny;
    const int gx = bx * TX + tx, gy = by * WTY + ty;
    const bool active = gx < nx && gy < ny;
    const int xh = (bx * TX + TX) % nx;   // x halo column
    const int yh = (by * WTY + WTY) % ny; // y halo row
    const uint32_t own = (uint32_t)(gy % ny) * nx + (uint32_t)(gx % nx);
    // the one halo point this thread stages: the y halo row, then the x
    // halo column, then the corner
    const bool has_halo = t < TX + WTY + 1;
    int hr = WTY, hc = TX;
    uint32_t hoff = 0;
    if (t < TX)
    {
        hc = t;
        hoff = (uint32_t)yh * nx + (uint32_t)((bx * TX + t) % nx);
    }
    else if (t < TX + WTY)
    {
        hr = t - TX;
        hoff = (uint32_t)((by * WTY + hr) % ny) * nx + (uint32_t)xh;
    }
    else if (has_halo)
        hoff = (uint32_t)yh * nx + (uint32_t)xh;
    else
        hoff = own; // loaded but not stored: keeps the load unconditional
    const T *p1 = y;
    const T *p2 = y + n;
    const float *l1 = lo, *l2 = lo ? lo + n : nullptr;

    T oh[WD][NC], hh[WD][NC];
    float ol[WD][NC], hl[WD][NC];
    auto issue = [&](int q, int zz) {
        const uint32_t oz = (uint32_t)zz * plane;
#pragma unroll
        for (int c = 0; c < NC; ++c)
        {
            // unconditional loads: a conditional one would send the
            // register queue to scratch memory
            oh[q][c] = y[(size_t)c * n + oz + own];
            hh[q][c] = y[(size_t)c * n + oz + hoff];
            if constexpr (HASLO)
            {
                ol[q][c] = lo[(size_t)c * n + oz + own];
                hl[q][c] = lo[(size_t)c * n + oz + hoff];
            }
        }
    };
    auto store = [&](int q, int s) {
#pragma unroll
        for (int c = 0; c < NC; ++c)
        {
            double v = (double)oh[q][c], w = (double)hh[q][c];
            if constexpr (HASLO)
            {
                v += (double)ol[q][c];
                w += (double)hl[q][c];
            }
            ring[c][s][ty][tx] = v;
            if (has_halo)
                ring[c][s][hr][hc] = w;
        }
    };
    // gamma^2 v^2 at one plaquette corner, VelocityKernel.hpp:82-96
    auto corner = [&](int s, int oy, int ox) -> double {
        const double v0 = ring[0][s][ty + oy][tx + ox];
        const double v1 = ring[1][s][ty + oy][tx + ox];
        const double v2 = ring[NC - 2][s][ty + oy][tx + ox];
        const double v3 = ring[NC - 1][s][ty + oy][tx + ox];
        const double RpR = p.vel_RpR, iR = p.vel_iR, iR2 = p.vel_iR2;
        return gamma_sq_v_sq_k(v0 * iR, v1 * iR, (v2 - RpR * v0) * iR2,
                               (v3 - RpR * v1) * iR2, p.vel_k1, p.vel_kc,
                               p.vel_k3);
    };

    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);
    const int len = z_end - z_begin;
    unsigned long long plain_w = 0; // per wavefront (uniform)
    unsigned long long weighted = 0; // per lane
    unsigned long long vcount = 0;   // per lane
    double vsum = 0.0;               // per lane
    const StarLoader S{nx, ny, nz, plane};
    const double i1 = 1.0 / p.dx, i2 = 1.0 / (p.dx * p.dx);

    auto vtx = [&](int s, int r, int c) -> Vtx {
        return Vtx{ring[0][s][r][c], ring[1][s][r][c]};
    };
    // Edge words of plane zz (ring slot s1; its in-plane +x / +y edges and
    // vertex flags) and, with s0 >= 0, the +z edges from plane slot s0 to
    // s1.  Each edge is evaluated once, by the thread of its low vertex;
    // the first WTY + TX + 1 threads also cover the halo column / row.
    auto edges = [&](int zz, int s0, int s1) {
        const int pq = zz & 1;
        auto put = [&](int r, int c, bool wx, bool wy) {
            const Vtx a = vtx(s1, r, c);
            bool fx = false, fy = false, fz = false;
            const int kx = wx ? edge_k(a, vtx(s1, r, c + 1), fx) : 0;
            const int ky = wy ? edge_k(a, vtx(s1, r + 1, c), fy) : 0;
            pw[pq][r][c] = (uint32_t)(kx + 1) | (uint32_t)(ky + 1) << 2 |
                           (uint32_t)fx << 4 | (uint32_t)fy << 5 |
                           (uint32_t)!vtx_ok(a) << 6;
            if (s0 >= 0)
            {
                const int kz = edge_k(vtx(s0, r, c), a, fz);
                zw[r][c] = (uint32_t)(kz + 1) | (uint32_t)fz << 2;
            }
        };
        put(ty, tx, true, true);
        if (t < TX)
            put(WTY, t, true, false);
        else if (t < TX + WTY)
            put(t - TX, TX, false, true);
        else if (t == TX + WTY)
            put(WTY, TX, false, false);
    };
    auto kx_of = [](uint32_t w) { return (int)(w & 3u) - 1; };
    auto ky_of = [](uint32_t w) { return (int)((w >> 2) & 3u) - 1; };

    auto process = [&](int z) {
        const int s0 = z % 3, s1 = (z + 1) % 3;
        int wxy = 0, wyz = 0, wzx = 0;
        bool in = active;
        const uint32_t id = (uint32_t)z * plane + (uint32_t)gy * nx + gx;
        if (in && p.mask && p.mask[id] == 0)
            in = false;
        if (in)
        {
            const int c0 = z & 1, c1 = c0 ^ 1;
            const uint32_t P00 = pw[c0][ty][tx], P10 = pw[c0][ty][tx + 1];
            const uint32_t P01 = pw[c0][ty + 1][tx], P11 = pw[c0][ty + 1][tx + 1];
            const uint32_t Q00 = pw[c1][ty][tx], Q10 = pw[c1][ty][tx + 1];
            const uint32_t Q01 = pw[c1][ty + 1][tx];
            const uint32_t Z00 = zw[ty][tx], Z10 = zw[ty][tx + 1];
            const uint32_t Z01 = zw[ty + 1][tx];
            // the 9 edges and 7 vertices of the cell's 3 plaquettes
            const uint32_t flags = (P00 & 0x70u) | (P10 & 0x60u) |
                                   (P01 & 0x50u) | (P11 & 0x40u) |
                                   (Q00 & 0x70u) | (Q10 & 0x40u) |
                                   (Q01 & 0x40u) | ((Z00 | Z10 | Z01) & 4u);
            if (flags == 0)
            {
                // plaquette_winding(t000, t100, t110, t010) etc. as -sum(k)
                const int kz000 = kx_of(Z00), kz100 = kx_of(Z10);
                const int kz010 = kx_of(Z01);
                wxy = -(kx_of(P00) + ky_of(P10) - kx_of(P01) - ky_of(P00));
                wyz = -(ky_of(P00) + kz010 - ky_of(Q00) - kz000);
                wzx = -(kz000 + kx_of(Q00) - kz100 - kx_of(P00));
            }
            else
            {
                const Vtx v000 = vtx(s0, ty, tx), v100 = vtx(s0, ty, tx + 1);
                const Vtx v110 = vtx(s0, ty + 1, tx + 1);
                const Vtx v010 = vtx(s0, ty + 1, tx);
                const Vtx v001 = vtx(s1, ty, tx), v101 = vtx(s1, ty, tx + 1);
                const Vtx v011 = vtx(s1, ty + 1, tx);
                const Vtx v[7] = {v000, v100, v110, v010, v001, v101, v011};
                int w[3];
                windings_by_phase(v, w);
                wxy = w[0];
                wyz = w[1];
                wzx = w[2];
            }
            weighted += (unsigned long long)(abs(wxy) + abs(wyz) + abs(wzx));
            if constexpr (VEL)
            {
                if (p.do_vel && (wxy != 0 || wyz != 0 || wzx != 0))
                {
                    // corners: s0 = plane z, s1 = plane z + 1
                    const double c000 = corner(s0, 0, 0);
                    const double c100 = corner(s0, 0, 1);
                    const double c010 = corner(s0, 1, 0);
                    const double c001 = corner(s1, 0, 0);
                    if (wxy != 0)
                    {
                        vsum += c000 + c100 + corner(s0, 1, 1) + c010;
                        vcount += 4;
                    }
                    if (wyz != 0)
                    {
                        vsum += c000 + c010 + corner(s1, 1, 0) + c001;
                        vcount += 4;
                    }
                    if (wzx != 0)
                    {
                        vsum += c000 + c001 + corner(s1, 0, 1) + c100;
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
                const double Ramp = p.R_tab ? p.R_tab[*p.seg] : p.R;
                if (!tag && p.amp_thr > 0.0 &&
                    sqrt(psi1 * psi1 + psi2 * psi2) / Ramp < p.amp_thr)
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
    };

    // plane rel (0 .. len, the last one the upper neighbour of the chunk)
    // sits in register slot rel % WD until it is stored to the ring
    issue(0, z_begin);
    store(0, z_begin % 3);
#pragma unroll
    for (int j = 1; j <= WD; ++j)
        if (j <= len)
            issue(j % WD, wrapc(z_begin + j, nz));
    __syncthreads();
    edges(z_begin, -1, z_begin % 3);
    for (int k = 0; k < len; k += WD)
    {
#pragma unroll
        for (int u = 0; u < WD; ++u)
        {
            const int rel = k + u;
            if (rel < len)
            {
                const int z = z_begin + rel;
                store((u + 1) % WD, (z + 1) % 3);
                if (rel + 1 + WD <= len)
                    issue((u + 1) % WD, wrapc(z + 1 + WD, nz));
                __syncthreads();
                edges(z + 1, z % 3, (z + 1) % 3);
                __syncthreads();
                process(z);
            }
        }
    }
    // wavefront reduction of the per-lane counts
    for (int o = 32; o > 0; o >>= 1)
    {
        weighted += __shfl_xor(weighted, o, 64);
        if constexpr ((WM & WM_VEL) != 0)
            vcount += __shfl_xor(vcount, o, 64);
    }
    // then across the block's wavefronts: one atomic per block and count
    // (all blocks adding into one address serialise in L2)
    __shared__ unsigned long long wcnt[WTY][3];
    if (tx == 0)
    {
        wcnt[ty][0] = plain_w;
        wcnt[ty][1] = weighted;
        wcnt[ty][2] = vcount;
    }
    __syncthreads();
    if (t < 3)
    {
        unsigned long long v = 0;
        for (int w = 0; w < WTY; ++w)
            v += wcnt[w][t];
        unsigned long long *cnt = p.seg ? p.counts + 3 * *p.seg : p.counts;
        if (t < 2 || (WM & WM_VEL) != 0)
            atomicAdd(&cnt[t], v);
    }
    if constexpr ((WM & WM_VEL) != 0)
    {
        // fixed-order block reduction of the velocity sum -> one partial
        __shared__ double red[64 * WTY];
        red[t] = vsum;
        __syncthreads();
        for (int o = 32 * WTY; o > 0; o >>= 1)
        {
            if (t < o)
                red[t] += red[t + o];
            __syncthreads();
        }
        if (t == 0)
            p.vel_partials[bid] = red[0];
    }
}

template <typename T, int WM>
static void launch_windings_t(int ntiles, hipStream_t s, const T *y,
                              const float *lo, const WindParams &p)
{
    if constexpr (std::is_same<T, float>::value)
    {
        if (lo)
        {
            hipLaunchKernelGGL((k_windings<T, WM, true>), dim3(ntiles),
                               dim3(64, WTY), 0, s, y, lo, p);
            return;
        }
    }
    hipLaunchKernelGGL((k_windings<T, WM, false>), dim3(ntiles), dim3(64, WTY),
                       0, s, y, lo, p);
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

// ---------------------------------------- register-streaming windings pass --
// Counts, velocities and the windings field without LDS or barriers.  A
// wavefront owns 63 consecutive x cells of one y row (lane 63 is the x halo
// of lane 62) and streams a chunk of z; each lane holds psi of its column
// at rows y and y+1 for planes z and z+1 in registers (row y+1 is the next
// wavefront's row y, read again through the cache) and takes the x+1
// column from the next lane by a shuffle.  Per plane each lane evaluates
// the 5 edges it owns -- +x at rows y and y+1, +y, and +z at rows y and y+1
// -- with the sign tests above and packs their k values and flags into two
// words, which the lane to its left reads by a shuffle; a cell's windings
// are then sums of packed k values, with the atan2 path for flagged cells.
// The stored words of the next RD planes are in flight while one is
// processed.
constexpr int RWY = 4;  // wavefronts (rows) per workgroup
constexpr int RTX = 63; // cells per wavefront row

// edge word of a plane, per lane: +x edge at row y (bits 0-1, k + 1),
// +x at row y+1 (2-3), +y at row y (4-5); flags of those edges (6, 7, 8);
// vertex flags of rows y and y+1 (9, 10).  z word: +z edges at rows y
// (bits 0-1) and y+1 (2-3), their flags (4, 5).
__device__ __forceinline__ int kfield(uint32_t w, int sh)
{
    return (int)((w >> sh) & 3u) - 1;
}

template <typename T, int WM, bool HASLO>
__global__ void __launch_bounds__(64 * RWY)
    __attribute__((amdgpu_waves_per_eu((WM & WM_VEL) != 0 ? 3 : 4)))
    k_windings_r(const T *y, const float *lo, WindParams p)
{
    constexpr bool VEL = (WM & WM_VEL) != 0;
    constexpr int NC = VEL ? 4 : 2; // components per point
    constexpr int RD = 2;           // planes of loads in flight
    const int bid = xcd_tile(p.tiles_x * p.tiles_y * p.tiles_z);
    const int bx = bid % p.tiles_x;
    const int rem = bid / p.tiles_x;
    const int by = rem % p.tiles_y;
    const int bz = rem / p.tiles_y;
    const int lane = threadIdx.x, ty = threadIdx.y;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t n = p.n;
    const uint32_t plane = (uint32_t)nx * ny;
    const int gx = bx * RTX + lane; // lane 63: the halo column
    const int gy = by * RWY + ty;
    const bool active = lane < RTX && gx < nx && gy < ny;
    const int xw = gx % nx;
    const int yw = gy % ny, yw1 = (gy + 1) % ny;
    const uint32_t ia = (uint32_t)yw * nx + xw, ib = (uint32_t)yw1 * nx + xw;

    // stored words in flight: [slot][point a / b][component]
    T qw[RD][2][NC];
    float ql[RD][2][NC];
    auto issue = [&](int q, int zz) {
        const uint32_t oz = (uint32_t)zz * plane;
#pragma unroll
        for (int c = 0; c < NC; ++c)
        {
            qw[q][0][c] = y[(size_t)c * n + oz + ia];
            qw[q][1][c] = y[(size_t)c * n + oz + ib];
            if constexpr (HASLO)
            {
                ql[q][0][c] = lo[(size_t)c * n + oz + ia];
                ql[q][1][c] = lo[(size_t)c * n + oz + ib];
            }
        }
    };
    auto val = [&](int q, int pt, int c) -> double {
        if constexpr (HASLO)
            return (double)qw[q][pt][c] + (double)ql[q][pt][c];
        else
            return (double)qw[q][pt][c];
    };
    // gamma^2 v^2 at a vertex, VelocityKernel.hpp:82-96 (velocity passes
    // evaluate it once per vertex; a pierced plaquette sums its corners)
    auto gamma_of = [&](const double *v) -> double {
        const double RpR = p.vel_RpR, iR = p.vel_iR, iR2 = p.vel_iR2;
        return gamma_sq_v_sq_k(v[0] * iR, v[1] * iR,
                               (v[NC - 2] - RpR * v[0]) * iR2,
                               (v[NC - 1] - RpR * v[1]) * iR2, p.vel_k1,
                               p.vel_kc, p.vel_k3);
    };
    // plane values: own column (a: row y, b: row y+1) and psi of the x+1
    // column; gamma^2 v^2 of the same four vertices
    double A[NC], B[NC], A1[2], B1[2];     // plane z + 1 (newest)
    double Ac[NC], Bc[NC], A1c[2], B1c[2]; // plane z
    double gA = 0, gB = 0, gA1 = 0, gB1 = 0, gAc = 0, gBc = 0, gA1c = 0,
           gB1c = 0;
    uint32_t W = 0, W1 = 0, Wc = 0, W1c = 0; // edge words, planes z+1, z
    auto land = [&](int q) {
#pragma unroll
        for (int c = 0; c < NC; ++c)
        {
            A[c] = val(q, 0, c);
            B[c] = val(q, 1, c);
        }
#pragma unroll
        for (int c = 0; c < 2; ++c)
        {
            A1[c] = __shfl_down(A[c], 1, 64);
            B1[c] = __shfl_down(B[c], 1, 64);
        }
        if constexpr (VEL)
        {
            gA = gamma_of(A);
            gB = gamma_of(B);
            gA1 = __shfl_down(gA, 1, 64);
            gB1 = __shfl_down(gB, 1, 64);
        }
        const Vtx a{A[0], A[1]}, b{B[0], B[1]};
        const Vtx a1{A1[0], A1[1]}, b1{B1[0], B1[1]};
        bool fx0 = false, fx1 = false, fy0 = false;
        const int kx0 = edge_k(a, a1, fx0);
        const int kx1 = edge_k(b, b1, fx1);
        const int ky0 = edge_k(a, b, fy0);
        W = (uint32_t)(kx0 + 1) | (uint32_t)(kx1 + 1) << 2 |
            (uint32_t)(ky0 + 1) << 4 | (uint32_t)fx0 << 6 |
            (uint32_t)fx1 << 7 | (uint32_t)fy0 << 8 |
            (uint32_t)!vtx_ok(a) << 9 | (uint32_t)!vtx_ok(b) << 10;
        W1 = __shfl_down(W, 1, 64);
    };

    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);
    const int len = z_end - z_begin;
    uint32_t plain = 0, weighted = 0, vcount = 0; // per lane
    double vsum = 0.0;

    auto process = [&](int z) {
        // +z edges from plane z to z + 1 at rows y, y+1, and their x+1
        bool fz0 = false, fz1 = false;
        const int kz0 = edge_k(Vtx{Ac[0], Ac[1]}, Vtx{A[0], A[1]}, fz0);
        const int kz1 = edge_k(Vtx{Bc[0], Bc[1]}, Vtx{B[0], B[1]}, fz1);
        const uint32_t Z = (uint32_t)(kz0 + 1) | (uint32_t)(kz1 + 1) << 2 |
                           (uint32_t)fz0 << 4 | (uint32_t)fz1 << 5;
        const uint32_t Z1 = __shfl_down(Z, 1, 64);
        if (!active)
            return;
        const uint32_t id = (uint32_t)z * plane + ia;
        if (p.mask && p.mask[id] == 0)
            return;
        // flags of the 9 edges and 7 vertices of the cell's plaquettes
        // Wc: +x (y, y+1), +y edges, vertices 000, 010; W1c: +y edge at
        // x+1, vertices 100, 110; W: +x, +y edges, vertices 001, 011;
        // W1: vertex 101; Z: +z edges at 000, 010; Z1: +z edge at 100
        const uint32_t flags = (Wc & 0x7c0u) | (W1c & 0x700u) |
                               (W & 0x740u) | (W1 & 0x200u) |
                               (Z & 0x30u) | (Z1 & 0x10u);
        int wxy, wyz, wzx;
        if (flags == 0)
        {
            const int kx000 = kfield(Wc, 0), kx010 = kfield(Wc, 2);
            const int ky000 = kfield(Wc, 4), ky100 = kfield(W1c, 4);
            const int kx001 = kfield(W, 0), ky001 = kfield(W, 4);
            const int kz000 = kz0, kz010 = kz1, kz100 = kfield(Z1, 0);
            wxy = -(kx000 + ky100 - kx010 - ky000);
            wyz = -(ky000 + kz010 - ky001 - kz000);
            wzx = -(kz000 + kx001 - kz100 - kx000);
        }
        else
        {
            const Vtx v[7] = {Vtx{Ac[0], Ac[1]},   Vtx{A1c[0], A1c[1]},
                              Vtx{B1c[0], B1c[1]}, Vtx{Bc[0], Bc[1]},
                              Vtx{A[0], A[1]},     Vtx{A1[0], A1[1]},
                              Vtx{B[0], B[1]}};
            int w[3];
            windings_by_phase(v, w);
            wxy = w[0];
            wyz = w[1];
            wzx = w[2];
        }
        plain += (wxy != 0) + (wyz != 0) + (wzx != 0);
        weighted += abs(wxy) + abs(wyz) + abs(wzx);
        if constexpr (VEL)
        {
            if (p.do_vel)
            {
                // corners 000, 100, 110, 010 (plane z), 001, 101, 011
                if (wxy != 0)
                {
                    vsum += gAc + gA1c + gB1c + gBc;
                    vcount += 4;
                }
                if (wyz != 0)
                {
                    vsum += gAc + gBc + gB + gA;
                    vcount += 4;
                }
                if (wzx != 0)
                {
                    vsum += gAc + gA + gA1 + gA1c;
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
    };

    // plane rel (0 .. len) sits in slot rel % RD until it lands
    issue(0, z_begin);
#pragma unroll
    for (int j = 1; j < RD; ++j)
        issue(j, wrapc(z_begin + j, nz)); // j <= 2 < len + 1 or wraps
    land(0);
    issue(0, wrapc(z_begin + RD, nz));
    for (int k = 0; k < len; k += RD)
    {
#pragma unroll
        for (int u = 0; u < RD; ++u)
        {
            const int rel = k + u;
            if (rel < len)
            {
                const int z = z_begin + rel;
#pragma unroll
                for (int c = 0; c < NC; ++c)
                {
                    Ac[c] = A[c];
                    Bc[c] = B[c];
                }
#pragma unroll
                for (int c = 0; c < 2; ++c)
                {
                    A1c[c] = A1[c];
                    B1c[c] = B1[c];
                }
                gAc = gA;
                gBc = gB;
                gA1c = gA1;
                gB1c = gB1;
                Wc = W;
                W1c = W1;
                land((u + 1) % RD); // plane z + 1
                if (rel + 1 + RD <= len)
                    issue((u + 1) % RD, wrapc(z + 1 + RD, nz));
                process(z);
            }
        }
    }
    // wavefront, then block reduction of the counts: one atomic per block
    unsigned long long cp = plain, cw = weighted, cv = vcount;
    for (int o = 32; o > 0; o >>= 1)
    {
        cp += __shfl_xor(cp, o, 64);
        cw += __shfl_xor(cw, o, 64);
        cv += __shfl_xor(cv, o, 64);
    }
    __shared__ unsigned long long wcnt[RWY][3];
    if (lane == 0)
    {
        wcnt[ty][0] = cp;
        wcnt[ty][1] = cw;
        wcnt[ty][2] = cv;
    }
    __shared__ double red[64 * RWY];
    const int t = ty * 64 + lane;
    red[t] = vsum;
    __syncthreads();
    if (t < 3)
    {
        unsigned long long v = 0;
        for (int w = 0; w < RWY; ++w)
            v += wcnt[w][t];
        unsigned long long *cnt = p.seg ? p.counts + 3 * *p.seg : p.counts;
        if (t < 2 || VEL)
            atomicAdd(&cnt[t], v);
    }
    if constexpr (VEL)
    {
        // fixed-order block reduction of the velocity sum -> one partial
        for (int o = 32 * RWY; o > 0; o >>= 1)
        {
            if (t < o)
                red[t] += red[t + o];
            __syncthreads();
        }
        if (t == 0)
            p.vel_partials[bid] = red[0];
    }
}

template <typename T, int WM>
static void launch_windings_rt(int ntiles, hipStream_t s, const T *y,
                               const float *lo, const WindParams &p)
{
    if constexpr (std::is_same<T, float>::value)
    {
        if (lo)
        {
            hipLaunchKernelGGL((k_windings_r<T, WM, true>), dim3(ntiles),
                               dim3(64, RWY), 0, s, y, lo, p);
            return;
        }
    }
    hipLaunchKernelGGL((k_windings_r<T, WM, false>), dim3(ntiles),
                       dim3(64, RWY), 0, s, y, lo, p);
}

template <typename T>
static void launch_windings_r(int wm, int ntiles, hipStream_t s, const T *y,
                              const float *lo, const WindParams &p)
{
    if (wm == WM_VEL)
        launch_windings_rt<T, WM_VEL>(ntiles, s, y, lo, p);
    else if (wm == WM_FIELD)
        launch_windings_rt<T, WM_FIELD>(ntiles, s, y, lo, p);
    else
        launch_windings_rt<T, 0>(ntiles, s, y, lo, p);
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
    double iR, iR2, iR4; // 1 / R, 1 / R^2, 1 / R^4 (kernel arguments, so
                         // they sit in scalar registers)
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
template <typename T, bool HASLO>
__global__ void __launch_bounds__(64 * ETY)
    __attribute__((amdgpu_waves_per_eu(4)))
    k_energy(const T *y, const float *lo, EnergyParams p)
{
    constexpr int ED = 2; // planes of loads in flight
    __shared__ double sp[2][ETY + 4][64 + 4];
    __shared__ double red[64 * ETY];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int t = ty * 64 + tx;
    const int bid = xcd_tile(p.tiles_x * p.tiles_y * p.tiles_z);
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
    // the one halo point this thread stages (2-cell halo, no corners):
    // threads 0..255 the y halo rows, 256..287 the x halo columns; the rest
    // load their own cell again (unconditional loads keep the register
    // queue out of scratch memory) and store nothing
    const bool has_halo = t < 4 * 64 + 4 * ETY;
    int hr = 0, hc = 0;
    uint32_t hrow = rowc;
    if (t < 4 * 64)
    {
        const int k = t / 64, c = t % 64;
        hr = k < 2 ? k : ETY + k;            // LDS rows 0, 1, ETY+2, ETY+3
        hc = c + 2;
        const int hy = k < 2 ? by * ETY - 2 + k : by * ETY + ETY + (k - 2);
        hrow = (uint32_t)(((hy % ny) + ny) % ny) * nx +
               (uint32_t)((bx * 64 + c) % nx);
    }
    else if (has_halo)
    {
        const int u = t - 4 * 64, r = u / 4, k = u % 4;
        hr = r + 2;
        hc = k < 2 ? k : 64 + k;             // LDS columns 0, 1, 66, 67
        const int hx = k < 2 ? bx * 64 - 2 + k : bx * 64 + 64 + (k - 2);
        hrow = (uint32_t)((by * ETY + r) % ny) * nx +
               (uint32_t)(((hx % nx) + nx) % nx);
    }
    const double i1 = 1.0 / p.dx;
    const double R = p.R;
    const double iR = p.iR, iR2 = p.iR2, iR4 = p.iR4;
    const double f_a = 1.4142135623730951;
    const int z_begin = bz * p.z_chunk;
    const int z_end = min(z_begin + p.z_chunk, nz);
    const int len = z_end - z_begin;
    auto zoff = [&](int z) -> uint32_t {
        return (uint32_t)wrapc(wrapc(z, nz), nz) * plane;
    };
    // stored words in flight: psi of the own column WD planes ahead of
    // its use as the z + 2 neighbour; halo psi and centre Pi of a plane
    // ED planes ahead of the plane's turn
    T cw[ED][2], hw[ED][2], pw[ED][2];
    float cl[ED][2], hl[ED][2], pl[ED][2];
    auto issue_c = [&](int q, int z) { // own psi of plane z
        const uint32_t o = zoff(z) + rowc;
#pragma unroll
        for (int c = 0; c < 2; ++c)
        {
            cw[q][c] = y[(size_t)c * n + o];
            if constexpr (HASLO)
                cl[q][c] = lo[(size_t)c * n + o];
        }
    };
    auto issue_h = [&](int q, int z) { // halo psi and own Pi of plane z
        const uint32_t oz = zoff(z);
#pragma unroll
        for (int c = 0; c < 2; ++c)
        {
            hw[q][c] = y[(size_t)c * n + oz + hrow];
            pw[q][c] = y[(size_t)(c + 2) * n + oz + rowc];
            if constexpr (HASLO)
            {
                hl[q][c] = lo[(size_t)c * n + oz + hrow];
                pl[q][c] = lo[(size_t)(c + 2) * n + oz + rowc];
            }
        }
    };
    auto val = [&](T w, float l) -> double {
        if constexpr (HASLO)
            return (double)w + (double)l;
        else
            return (double)w;
    };
    double a1[5], a2[5];
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        const uint32_t o = zoff(z_begin - 2 + i) + rowc;
        a1[i + 1] = val(y[o], HASLO ? lo[o] : 0.0f);
        a2[i + 1] = val(y[(size_t)n + o], HASLO ? lo[(size_t)n + o] : 0.0f);
    }
    double acc[14];
#pragma unroll
    for (int q = 0; q < 14; ++q)
        acc[q] = 0.0;
    auto cellwork = [&](int z, double Pi1, double Pi2) {
        const uint32_t id = zoff(z) + rowc;
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
            // the divisions by powers of R as products with their
            // reciprocals (one rounding apart; positive terms, so the sums
            // move by ~1e-16 relative)
            const double rho = (c1 * c1 + c2 * c2 + (g1sq + g2sq) + pot) * iR4;
            const double tp =
                (psi_sq > 0.0) ? (psi1 * Pi2 - Pi1 * psi2) / psi_sq : 0.0;
            const double a_dot = f_a * tp * iR;
            const double rak = 0.5 * a_dot * a_dot;
            const double rag = 0.5 * f_a * f_a * gtsq * iR2;
            double rrk = 0.0;
            if (psi_sq > 0.0)
            {
                const double rd = (psi1 * c1 + psi2 * c2) / sqrt(psi_sq);
                rrk = (rd * rd) * iR4;
            }
            const double rrg = (g1sq + g2sq - psi_sq * gtsq) * iR4;
            const double rrm = pot * iR4;
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
    };

    // own psi of plane z + 2 + q sits in slot q % ED until it enters the
    // z queue; halo / Pi of plane z in slot (z - z_begin) % ED
#pragma unroll
    for (int q = 0; q < ED; ++q)
    {
        issue_c(q, z_begin + 2 + q);
        if (q < len)
            issue_h(q, z_begin + q);
    }
    for (int k = 0; k < len; k += ED)
    {
#pragma unroll
        for (int u = 0; u < ED; ++u)
        {
            const int rel = k + u;
            if (rel < len)
            {
                const int z = z_begin + rel;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                {
                    a1[i] = a1[i + 1];
                    a2[i] = a2[i + 1];
                }
                a1[4] = val(cw[u][0], cl[u][0]);
                a2[4] = val(cw[u][1], cl[u][1]);
                if (rel + ED < len)
                    issue_c(u, z + 2 + ED);
                sp[0][ty + 2][tx + 2] = a1[2];
                sp[1][ty + 2][tx + 2] = a2[2];
                if (has_halo)
                {
                    sp[0][hr][hc] = val(hw[u][0], hl[u][0]);
                    sp[1][hr][hc] = val(hw[u][1], hl[u][1]);
                }
                const double Pi1 = val(pw[u][0], pl[u][0]);
                const double Pi2 = val(pw[u][1], pl[u][1]);
                if (rel + ED < len)
                    issue_h(u, z + ED);
                __syncthreads();
                cellwork(z, Pi1, Pi2);
                __syncthreads();
            }
        }
    }
    // fixed-order block reduction of each sum -> one row of partials
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
            p.partials[(size_t)bid * 14 + q] = red[0];
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

// The same fill four cells per lane for FP32 / FP32C levels (n a multiple of
// 4, 16-byte aligned arrays): one 16-byte load per component (and low word)
// and one 16-byte store per lane, the per-cell arithmetic of k_fill.
template <bool HASLO>
__global__ void k_fill_f4(const float *y, const float *lo, uint32_t n, int kind,
                          int scheme, double thr, double R, float *buf)
{
    const uint32_t n4 = n / 4;
    for (uint32_t q = blockIdx.x * 256 + threadIdx.x; q < n4;
         q += gridDim.x * 256)
    {
        const uint32_t id = 4 * q;
        auto get = [&](uint32_t off, double (&v)[4]) {
            const float4 h = *reinterpret_cast<const float4 *>(y + off + id);
            v[0] = h.x;
            v[1] = h.y;
            v[2] = h.z;
            v[3] = h.w;
            if constexpr (HASLO)
            {
                const float4 l = *reinterpret_cast<const float4 *>(lo + off + id);
                v[0] = (double)h.x + (double)l.x;
                v[1] = (double)h.y + (double)l.y;
                v[2] = (double)h.z + (double)l.z;
                v[3] = (double)h.w + (double)l.w;
            }
        };
        float out[4];
        if (kind == AXS_FIELD_MASKED_A_DOT)
        {
            double psi1[4], psi2[4], Pi1[4], Pi2[4];
            get(0, psi1);
            get(n, psi2);
            get(2 * n, Pi1);
            get(3 * n, Pi2);
            for (int e = 0; e < 4; ++e)
                out[e] = (float)masked_a_dot_d(scheme, thr, psi1[e], psi2[e],
                                               Pi1[e], Pi2[e], R);
        }
        else
        {
            double v[4];
            get((uint32_t)(kind + 1) * n, v);
            for (int e = 0; e < 4; ++e)
                out[e] = (float)v[e];
        }
        *reinterpret_cast<float4 *>(buf + id) =
            make_float4(out[0], out[1], out[2], out[3]);
    }
}

static bool aligned16(const void *p)
{
    return ((uintptr_t)p & 15) == 0;
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

// enq: a cadence pass -- enqueue only, counts into enq->counts + 3 * seg
struct CadenceArgs
{
    unsigned long long *counts;
    const int *seg;
    const double *R_tab;
};

static int run_windings(axs_level *L, const uint8_t *mask, int8_t *wfield,
                        uint8_t *tags, const axs_tagger_params *tp,
                        int64_t out[2], const VelArgs *vel = nullptr,
                        const CadenceArgs *enq = nullptr)
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
    const int wm = (vel ? WM_VEL : 0) | (tags ? WM_TAG : 0) |
                   ((tags && (p.use_lap || p.use_rgrad)) ? WM_TAGD : 0) |
                   (wfield ? WM_FIELD : 0);
    // the register pass for counts / velocities / windings field, the LDS
    // pass for the tagger outputs
    const bool reg = (wm & (WM_TAG | WM_TAGD)) == 0 &&
                     (wm & (WM_VEL | WM_FIELD)) != (WM_VEL | WM_FIELD);
    p.tiles_x = reg ? (p.nx + RTX - 1) / RTX : (p.nx + 63) / 64;
    p.tiles_y = reg ? (p.ny + RWY - 1) / RWY : (p.ny + WTY - 1) / WTY;
    int zc = reg ? 64 : 32;
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
        p.vel_RpR = vel->RpR;
        // the constants of gamma_sq_v_sq_from_phi (Velocity.hpp:62-78)
        const double c1 = 0.41222, e1 = -0.025763;
        const double m_r_sq = vel->m_r * vel->m_r, c1_sq = c1 * c1;
        p.vel_iR = 1.0 / vel->R;
        p.vel_iR2 = 1.0 / (vel->R * vel->R);
        p.vel_k1 = 1.0 / (2.0 * m_r_sq * c1_sq);
        p.vel_kc = e1 / (c1 * c1_sq);
        p.vel_k3 = e1 / (4.0 * m_r_sq * c1_sq * c1_sq * c1);
        p.vel_partials = L->ctx->d_partials;
    }
    if (enq)
    {
        p.counts = enq->counts;
        p.seg = enq->seg;
        p.R_tab = enq->R_tab;
    }
    else
        hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, s, p.counts);
    const float *lo = L->precision == AXS_FP32C ? (const float *)L->comp
                                                : (const float *)nullptr;
    if (L->precision == AXS_FP64)
    {
        if (reg)
            launch_windings_r<double>(wm, ntiles, s, (const double *)L->y,
                                      nullptr, p);
        else
            launch_windings<double>(wm, ntiles, s, (const double *)L->y,
                                    nullptr, p);
    }
    else if (reg)
        launch_windings_r<float>(wm, ntiles, s, (const float *)L->y, lo, p);
    else
        launch_windings<float>(wm, ntiles, s, (const float *)L->y, lo, p);
    AXS_HIP_TRY(hipGetLastError());
    if (enq)
        return AXS_OK;
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

} // extern "C"

namespace axs
{

__global__ void k_fill_u8(uint8_t *p, size_t n, uint8_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_seg_advance(int *seg, int reset)
{
    if (threadIdx.x == 0)
        *seg = reset ? 0 : *seg + 1;
}

// One cadence pass on the level's current state: clear the tags (a fill
// kernel, so that the pass can be captured into a graph), the windings /
// tagger pass adding its counts at counts + 3 * (*seg), then *seg += 1.
int launch_windings_pass(axs_level *L, const axs_tagger_params *tp,
                         uint8_t *tags, unsigned long long *counts,
                         int *seg, const double *R_tab)
{
    hipStream_t s = L->ctx->stream;
    if (tags)
    {
        const size_t n = L->g.ncell;
        const unsigned blocks =
            (unsigned)std::min<size_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_fill_u8, dim3(blocks), dim3(256), 0, s, tags, n,
                           (uint8_t)0);
        AXS_HIP_TRY(hipGetLastError());
    }
    const CadenceArgs enq{counts, seg, R_tab};
    int rc = run_windings(L, nullptr, nullptr, tags, tags ? tp : nullptr,
                          nullptr, nullptr, &enq);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_seg_advance, dim3(1), dim3(64), 0, s, seg, 0);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

int launch_seg_reset(axs_level *L, int *seg)
{
    hipLaunchKernelGGL(k_seg_advance, dim3(1), dim3(64), 0, L->ctx->stream,
                       seg, 1);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

} // namespace axs


// The energy pass's launch parameters for an nx x ny x nz periodic grid;
// *nblocks receives the grid size (one partial row of 14 per block).
static EnergyParams energy_params(int nx, int ny, int nz, double dx, double R,
                                  double lambda, double RpR,
                                  const axs_masking *m, const uint8_t *mask,
                                  long *nblocks)
{
    EnergyParams p{};
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.n = (uint32_t)((long long)nx * ny * nz);
    p.dx = dx;
    p.R = R;
    p.lambda = lambda;
    p.RpR = RpR;
    p.iR = 1.0 / R;
    p.iR2 = 1.0 / (R * R);
    p.iR4 = 1.0 / ((R * R) * (R * R));
    p.scheme = m->scheme;
    p.thr = m->threshold;
    p.mask = mask;
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
    *nblocks = (long)p.tiles_x * p.tiles_y * p.tiles_z;
    return p;
}

// The energy pass and the fixed-order reduction of its partials into
// result[14] (device), on stream s.
static int launch_energy(EnergyParams p, int precision, const void *y,
                         const float *comp, double *partials, int nb,
                         double *result, hipStream_t s)
{
    p.partials = partials;
    if (precision == AXS_FP64)
        hipLaunchKernelGGL((k_energy<double, false>), dim3(nb), dim3(64, ETY), 0,
                           s, (const double *)y, (const float *)nullptr, p);
    else if (precision == AXS_FP32C)
        hipLaunchKernelGGL((k_energy<float, true>), dim3(nb), dim3(64, ETY), 0,
                           s, (const float *)y, comp, p);
    else
        hipLaunchKernelGGL((k_energy<float, false>), dim3(nb), dim3(64, ETY), 0,
                           s, (const float *)y, (const float *)nullptr, p);
    AXS_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3(14), dim3(256), 0, s, partials, nb,
                       14, result);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

// Reduction partials of axs_energy_sums_noalloc: one row of 14 per block of
// the energy pass, at most kNoAllocEnergyBlocks blocks (32768: a 1024^3
// grid).  Every block writes its whole row before the reduction reads it,
// so the array needs no clearing between calls.  Host code never uses this
// name as an address: it takes the device address from
// hipGetSymbolAddress, once per device (scratch_on_device).
constexpr long kNoAllocEnergyBlocks = 32768;
static __device__ __align__(256) double scratch[kNoAllocEnergyBlocks * 14];

static int scratch_on_device(double **out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double *ptr[kMaxDevices] = {};
    int dev = -1;
    AXS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(AXS_ERR_UNSUPPORTED,
                         "axs_energy_sums_noalloc: device index above 63");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev])
    {
        void *p = nullptr;
        // with device dev current: its own copy of the array
        AXS_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return set_error(AXS_ERR_HIP,
                             "axs_energy_sums_noalloc: no address for the "
                             "scratch array");
        ptr[dev] = static_cast<double *>(p);
    }
    *out = ptr[dev];
    return AXS_OK;
}

extern "C" {

int axs_energy_sums(axs_level *L, double R, double lambda, double RpR,
                    const axs_masking *m, const uint8_t *mask, double out[14])
{
    AXS_CHECK_ARG(L && m && out, "axs_energy_sums: NULL argument");
    AXS_CHECK_ARG(R > 0.0, "axs_energy_sums: R must be > 0");
    AXS_CHECK_ARG(m->scheme >= 0 && m->scheme <= 2,
                  "axs_energy_sums: unknown masking scheme");
    hipStream_t s = L->ctx->stream;
    long nbl = 0;
    const EnergyParams p = energy_params(L->g.nx, L->g.ny, L->g.nz, L->dx, R,
                                         lambda, RpR, m, mask, &nbl);
    AXS_CHECK_ARG(nbl <= kMaxPartials, "axs_energy_sums: grid too large");
    int rc = launch_energy(p, L->precision, L->y, L->comp, L->ctx->d_partials,
                           (int)nbl, L->ctx->d_result, s);
    if (rc)
        return rc;
    AXS_HIP_TRY(hipMemcpyAsync(L->ctx->h_result, L->ctx->d_result,
                               14 * sizeof(double), hipMemcpyDeviceToHost, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    for (int q = 0; q < 14; ++q)
        out[q] = L->ctx->h_result[q];
    return AXS_OK;
}

int axs_energy_sums_noalloc(const void *state, const float *comp, int nx,
                            int ny, int nz, int precision, double dx, double R,
                            double lambda, double R_prime_over_R,
                            const axs_masking *m, const uint8_t *mask,
                            double *out_dev, void *stream)
{
    // every argument and the size limit are checked before any HIP call
    AXS_CHECK_ARG(m, "axs_energy_sums_noalloc: NULL masking");
    AXS_CHECK_ARG(precision == AXS_FP64 || precision == AXS_FP32C ||
                      precision == AXS_FP32,
                  "axs_energy_sums_noalloc: unknown precision");
    AXS_CHECK_ARG(nx >= 4 && ny >= 4 && nz >= 4,
                  "axs_energy_sums_noalloc: every dimension must be >= 4");
    AXS_CHECK_ARG((long long)nx * ny * nz * 4 <= (1LL << 32),
                  "axs_energy_sums_noalloc: 4*nx*ny*nz must be <= 2^32");
    AXS_CHECK_ARG(dx > 0.0 && R > 0.0,
                  "axs_energy_sums_noalloc: dx and R must be > 0");
    AXS_CHECK_ARG(m->scheme >= 0 && m->scheme <= 2,
                  "axs_energy_sums_noalloc: unknown masking scheme");
    long nbl = 0;
    const EnergyParams p = energy_params(nx, ny, nz, dx, R, lambda,
                                         R_prime_over_R, m, mask, &nbl);
    if (nbl > kNoAllocEnergyBlocks)
        return set_error(AXS_ERR_ARG,
                         "axs_energy_sums_noalloc: the grid needs " +
                             std::to_string(nbl) +
                             " energy blocks, above this entry point's fixed "
                             "scratch of 32768 (a 1024^3 grid); use "
                             "axs_energy_sums");
    AXS_CHECK_ARG(state && out_dev, "axs_energy_sums_noalloc: NULL pointer");
    AXS_CHECK_ARG((precision == AXS_FP32C) == (comp != nullptr),
                  "axs_energy_sums_noalloc: comp is required for FP32C and "
                  "only for it");
    double *part = nullptr;
    int rc = scratch_on_device(&part);
    if (rc)
        return rc;
    return launch_energy(p, precision, state, comp, part, (int)nbl, out_dev,
                         static_cast<hipStream_t>(stream));
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
    else if (n % 4 == 0 && aligned16(L->y) && aligned16(L->comp) &&
             aligned16(buf))
    {
        const int nb4 = nblocks(n / 4, 256, 8192);
        if (L->comp)
            hipLaunchKernelGGL(k_fill_f4<true>, dim3(nb4), dim3(256), 0, s,
                               (const float *)L->y, (const float *)L->comp, n,
                               kind, m->scheme, m->threshold, R, (float *)buf);
        else
            hipLaunchKernelGGL(k_fill_f4<false>, dim3(nb4), dim3(256), 0, s,
                               (const float *)L->y, (const float *)nullptr, n,
                               kind, m->scheme, m->threshold, R, (float *)buf);
    }
    else
        hipLaunchKernelGGL((k_fill<float, float>), dim3(nb), dim3(256), 0, s,
                           (const float *)L->y, (const float *)L->comp, n, kind,
                           m->scheme, m->threshold, R, (float *)buf);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

} // extern "C"
