// axs_boxes.hip -- AMR box levels: one refinement level stored as a list
// of boxes with 3 ghost layers each, advanced by persistent work-list
// launches (the MultiFab-of-boxes path of the reference's AMR hierarchy,
// AxionStringsLevel.cpp and GRTeclyn's GRAmrLevel RK4 with FillPatch).
//
// Layout: component-major over the whole level; box b's grown region
// (n + 2 g cells per edge) starts at off[b] of every component array.
// A block-owner map (8^3-cell blocks, the blocking factor) gives, for any
// cell of the periodic domain, the box that owns it, so ghost fills and
// coarse lookups need no host-built index lists.
//
// Work-list: every box is cut into items of H rows x full box width W
// (W * H <= 512 lanes), each item streams through the box's z extent with
// the same register-queue + LDS-plane scheme as the uniform kernel, and a
// persistent grid strides over the items.
#include "axs_internal.hpp"
#include "axs_stage_math.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace axs
{

constexpr int kBlock = 8;     // blocking factor of box corners
constexpr int kBoxLanes = 512; // threads of a work-list block: one cell of an
                               // item's W x H plane each
constexpr int kBoxLds = 816;   // (W + 4) (H + 4) <= 816 for 8 <= W <= 64,
                               // 4 <= H <= kBoxLanes / W

struct BoxD
{
    int lo[3];
    int n[3];
    uint32_t off; // grown-region start in every component array
    int erow;     // the box's first row (y, z) in the x-edge arrays
};

struct WorkItem
{
    int box, y0, h, w;
};

} // namespace axs

struct axs_boxlevel
{
    axs_ctx *ctx;
    int nb;
    int dom[3];
    int ng;
    double dx;
    int precision;
    size_t word;
    std::vector<axs::BoxD> hbox;
    axs::BoxD *dbox;
    int *downer;  // [dom/8]^3 block owners (-1 = not covered)
    int bdim[3];  // dom / 8
    uint32_t total; // grown cells per component
    int64_t nvalid;
    void *y, *acc, *ya, *yb;
    float *comp;
    void *alloc[4]; // allocations behind y, acc, ya|yb, comp (from the
                    // context's pool)
    size_t alloc_bytes[4];
    axs::WorkItem *ditems;
    int nitems;
    int item_rows;   // rows per work item asked for (clamped per box)
    int grid_blocks; // cap on the persistent grid of the stage launches
    double *dcoeff;
    double *h_coeff; // pinned staging of the stage scalars (12)
    void *snap;      // coarse snapshot (y [+ comp]) for time interpolation
    float *snap_lo;
    size_t snap_bytes; // bytes in use
    size_t snap_cap;   // bytes of the pooled block
    unsigned long long snap_uid; // uid of the level the snapshot is of
    unsigned long long *derr;
    // identity of this level for the graph caches of finer levels: unique
    // over the process, unlike an address that a later level may reuse
    unsigned long long uid;
    // RK4 stage launches: 0 one per stage, 1 fused pairs (stages 1+2, 3+4;
    // the state then alternates between the arrays behind y and acc)
    int fusion;
    // x-edge arrays of the fused pairs: columns 0, 1, W-2, W-1 of every
    // box row (y, z), 4 components, packed (4 values per row, rows of a box
    // consecutive, z major), so the ghost fills read the x-face sources of
    // their neighbours contiguously instead of 2 cells per 128-byte line.
    // edge_n: the state's (valid when edge_gen == ygen and edge_state ==
    // y; the second pair writes it for y_{n+1}); edge_2: y_2's (the first
    // pair writes it).  ygen counts writes of the state's valid cells.
    void *edge_n, *edge_2;
    void *edge_alloc;
    size_t edge_alloc_bytes;
    size_t edge_rows;
    unsigned long long ygen, edge_gen;
    const void *edge_state;
    // captured step graphs, keyed by (dt, coarse level uid, t0_frac,
    // dt_frac) and the state arrays they read (this level's and the coarse
    // level's, which the fused pairs alternate); uid 0 = no coarse level
    struct GraphEntry
    {
        double dt, t0, df;
        unsigned long long coarse_uid;
        const void *state, *coarse_state;
        int fusion;
        hipGraph_t g;
        hipGraphExec_t ge;
    };
    std::vector<GraphEntry> graphs;
};

namespace axs
{

// ------------------------------------------------------------- views ----
struct LevelView
{
    const BoxD *box;
    const int *owner;
    int dom[3], bdim[3];
    int ng;
    uint32_t total;
};

static LevelView view_of(const axs_boxlevel *b)
{
    LevelView v{};
    v.box = b->dbox;
    v.owner = b->downer;
    for (int d = 0; d < 3; ++d)
    {
        v.dom[d] = b->dom[d];
        v.bdim[d] = b->bdim[d];
    }
    v.ng = b->ng;
    v.total = b->total;
    return v;
}

__device__ __forceinline__ int pmod(int a, int n)
{
    a %= n;
    return a < 0 ? a + n : a;
}

// storage index (component 0) of global cell (x, y, z) inside box `bi`'s
// grown region, the cell given in that box's index frame
__device__ __forceinline__ uint32_t sidx(const BoxD &B, int ng, int lx, int ly,
                                         int lz)
{
    const uint32_t gx = B.n[0] + 2 * ng, gy = B.n[1] + 2 * ng;
    return B.off + ((uint32_t)(lz + ng) * gy + (uint32_t)(ly + ng)) * gx +
           (uint32_t)(lx + ng);
}

// index (component 0) in the x-edge arrays of column col (0, 1, 2, 3 for
// x = 0, 1, W-2, W-1) of row (y, z) of box B
__device__ __forceinline__ size_t edge_idx(const BoxD &B, int y, int z, int col)
{
    return ((size_t)B.erow + (size_t)z * B.n[1] + y) * 4 + col;
}

// owner box of a periodic global cell and its storage index, or -1
__device__ __forceinline__ int locate(const LevelView &v, int x, int y, int z,
                                      uint32_t &idx)
{
    x = pmod(x, v.dom[0]);
    y = pmod(y, v.dom[1]);
    z = pmod(z, v.dom[2]);
    const int o = v.owner[((z / kBlock) * v.bdim[1] + y / kBlock) * v.bdim[0] +
                          x / kBlock];
    if (o < 0)
        return -1;
    const BoxD B = v.box[o];
    idx = sidx(B, v.ng, x - B.lo[0], y - B.lo[1], z - B.lo[2]);
    return o;
}

// 5-point Lagrange weights L_m(t) = prod_{k != m} (t - k)/(m - k), nodes
// m = -2..2, at offset t = -1/4 (even fine index) and +1/4 (odd) from the
// nearest coarse cell (exact binary fractions).
__constant__ double kQuartW[2][5] = {
    {-0.02197265625, 0.205078125, 0.9228515625, -0.123046875, 0.01708984375},
    {0.01708984375, -0.123046875, 0.9228515625, 0.205078125, -0.02197265625}};

template <typename T> struct CoarseSrc
{
    LevelView v;
    const T *cur;
    const float *cur_lo;
    const T *snap;
    const float *snap_lo;
    double t_frac;
};

template <typename T>
__device__ double coarse_value(const CoarseSrc<T> &c, int comp, int cx, int cy,
                               int cz, unsigned long long *err)
{
    uint32_t i;
    if (locate(c.v, cx, cy, cz, i) < 0)
    {
        atomicAdd(err, 1ull);
        return 0.0;
    }
    const size_t o = (size_t)comp * c.v.total + i;
    double v = (double)c.cur[o] + (c.cur_lo ? (double)c.cur_lo[o] : 0.0);
    if (c.snap && c.t_frac < 1.0)
    {
        const double s =
            (double)c.snap[o] + (c.snap_lo ? (double)c.snap_lo[o] : 0.0);
        v = (1.0 - c.t_frac) * s + c.t_frac * v;
    }
    return v;
}

// Interior prolongation.  A block takes a tile of 8 x 8 x 4 coarse cells
// under a fine box's valid region (box bounds are even): it stages the
// tile's coarse neighbourhood (the tile grown by 2, clipped to what the
// tile's cells use), each cell located once and its 4 components (time
// blended, FP64) put in LDS, then one thread per coarse cell computes its 8
// fine children from there.  Per child the tensor-product sums run in the
// same order as interp_quartic's (x innermost, then y, then z), with the
// same coarse values (0 where a coarse cell has no owner, which is
// counted), so the values are the same.
constexpr int kProTX = 8, kProTY = 8, kProTZ = 4;
constexpr int kProStage = (kProTX + 4) * (kProTY + 4) * (kProTZ + 4);

template <typename T>
__global__ void __launch_bounds__(256)
    k_prolong(LevelView f, T *buf, float *lo, CoarseSrc<T> c,
              unsigned long long *err)
{
    __shared__ double st[4][kProStage];
    const BoxD B = f.box[blockIdx.y];
    const int cn0 = B.n[0] / 2, cn1 = B.n[1] / 2, cn2 = B.n[2] / 2;
    const int ntx = (cn0 + kProTX - 1) / kProTX, nty = (cn1 + kProTY - 1) / kProTY,
              ntz = (cn2 + kProTZ - 1) / kProTZ;
    const int tile = (int)blockIdx.x;
    if (tile >= ntx * nty * ntz) // the whole block: before any barrier
        return;
    const int tx0 = (tile % ntx) * kProTX, ty0 = ((tile / ntx) % nty) * kProTY,
              tz0 = (tile / (ntx * nty)) * kProTZ;
    const int ex = min(kProTX, cn0 - tx0), ey = min(kProTY, cn1 - ty0),
              ez = min(kProTZ, cn2 - tz0);
    const int sx = ex + 4, sy = ey + 4, sz = ez + 4;
    // coarse global coordinates of the staged region's corner
    const int gx0 = (B.lo[0] >> 1) + tx0 - 2, gy0 = (B.lo[1] >> 1) + ty0 - 2,
              gz0 = (B.lo[2] >> 1) + tz0 - 2;
    for (int s = threadIdx.x; s < sx * sy * sz; s += blockDim.x)
    {
        const int lx = s % sx, ly = (s / sx) % sy, lz = s / (sx * sy);
        uint32_t ci;
        const bool ok = locate(c.v, gx0 + lx, gy0 + ly, gz0 + lz, ci) >= 0;
        if (!ok)
            atomicAdd(err, 1ull);
        for (int q = 0; q < 4; ++q)
        {
            double v = 0.0;
            if (ok)
            {
                const size_t o = (size_t)q * c.v.total + ci;
                v = (double)c.cur[o] + (c.cur_lo ? (double)c.cur_lo[o] : 0.0);
                if (c.snap && c.t_frac < 1.0)
                {
                    const double sv =
                        (double)c.snap[o] + (c.snap_lo ? (double)c.snap_lo[o] : 0.0);
                    v = (1.0 - c.t_frac) * sv + c.t_frac * v;
                }
            }
            st[q][s] = v;
        }
    }
    __syncthreads();
    const int ix = threadIdx.x % kProTX, iy = (threadIdx.x / kProTX) % kProTY,
              iz = threadIdx.x / (kProTX * kProTY);
    if (ix >= ex || iy >= ey || iz >= ez)
        return;
    double acc[4][2][2][2]; // comp, pz, py, px
    for (int q = 0; q < 4; ++q)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                acc[q][a][b][0] = acc[q][a][b][1] = 0.0;
    for (int k = 0; k < 5; ++k)
    {
        double ay[4][2][2]; // comp, py, px
        for (int q = 0; q < 4; ++q)
            ay[q][0][0] = ay[q][0][1] = ay[q][1][0] = ay[q][1][1] = 0.0;
        for (int j = 0; j < 5; ++j)
        {
            double ax[4][2]; // comp, px
            for (int q = 0; q < 4; ++q)
                ax[q][0] = ax[q][1] = 0.0;
            const int row = ((iz + k) * sy + (iy + j)) * sx + ix;
            for (int i = 0; i < 5; ++i)
                for (int q = 0; q < 4; ++q)
                {
                    const double v = st[q][row + i];
                    ax[q][0] += kQuartW[0][i] * v;
                    ax[q][1] += kQuartW[1][i] * v;
                }
            for (int q = 0; q < 4; ++q)
                for (int py = 0; py < 2; ++py)
                {
                    ay[q][py][0] += kQuartW[py][j] * ax[q][0];
                    ay[q][py][1] += kQuartW[py][j] * ax[q][1];
                }
        }
        for (int q = 0; q < 4; ++q)
            for (int pz = 0; pz < 2; ++pz)
                for (int py = 0; py < 2; ++py)
                {
                    acc[q][pz][py][0] += kQuartW[pz][k] * ay[q][py][0];
                    acc[q][pz][py][1] += kQuartW[pz][k] * ay[q][py][1];
                }
    }
    const int bx = tx0 + ix, by = ty0 + iy, bz = tz0 + iz; // coarse local
    for (int pz = 0; pz < 2; ++pz)
        for (int py = 0; py < 2; ++py)
            for (int px = 0; px < 2; ++px)
            {
                const uint32_t dst =
                    sidx(B, f.ng, 2 * bx + px, 2 * by + py, 2 * bz + pz);
                for (int q = 0; q < 4; ++q)
                {
                    const double v = acc[q][pz][py][px];
                    const T h = (T)v;
                    buf[(size_t)q * f.total + dst] = h;
                    if (lo)
                        lo[(size_t)q * f.total + dst] = (float)(v - (double)h);
                }
            }
}

// Ghost index t of a grown box -> local cell (lx, ly, lz): the ghost shell
// enumerated as the two z slabs (full grown planes), then the two y slabs of
// each valid plane (full grown rows), then the two x runs of each valid
// row, so consecutive threads walk along x wherever the shell allows.
__device__ __forceinline__ void ghost_cell(uint32_t t, int nx, int ny, int nz,
                                           int ng, int &lx, int &ly, int &lz)
{
    const uint32_t gx = nx + 2 * ng, gy = ny + 2 * ng;
    const uint32_t A = (uint32_t)ng * gy * gx; // one z slab
    if (t < 2 * A)
    {
        const uint32_t r = t % A;
        lz = (t < A ? -ng : nz) + (int)(r / (gy * gx));
        ly = (int)((r % (gy * gx)) / gx) - ng;
        lx = (int)(r % gx) - ng;
        return;
    }
    t -= 2 * A;
    const uint32_t Bp = 2u * ng * gx; // y slabs of one valid plane
    if (t < Bp * (uint32_t)nz)
    {
        lz = (int)(t / Bp);
        const uint32_t r = t % Bp, half = (uint32_t)ng * gx;
        ly = (r < half ? -ng : ny) + (int)((r % half) / gx);
        lx = (int)(r % gx) - ng;
        return;
    }
    t -= Bp * (uint32_t)nz;
    const uint32_t row = t / (2u * ng), k = t % (2u * ng);
    lz = (int)(row / (uint32_t)ny);
    ly = (int)(row % (uint32_t)ny);
    lx = k < (uint32_t)ng ? (int)k - ng : nx + (int)k - ng;
}

// Tensor-product quartic interpolation of the 4 coarse components at the
// fine cell (fx, fy, fz) (ratio 2, cell centred): each coarse neighbour is
// located once; per component the sums run x innermost, then y, then z.
template <typename T, int NC>
__device__ void interp_quartic(const CoarseSrc<T> &c, int fx, int fy, int fz,
                               double (&out)[NC], unsigned long long *err)
{
    const int cx = fx >> 1, cy = fy >> 1, cz = fz >> 1;
    const double *wx = kQuartW[fx & 1], *wy = kQuartW[fy & 1],
                 *wz = kQuartW[fz & 1];
    double acc[NC] = {};
    for (int k = 0; k < 5; ++k)
    {
        double ay[NC] = {};
        for (int j = 0; j < 5; ++j)
        {
            double ax[NC] = {};
            for (int i = 0; i < 5; ++i)
            {
                uint32_t ci;
                const bool ok =
                    locate(c.v, cx + i - 2, cy + j - 2, cz + k - 2, ci) >= 0;
                if (!ok)
                    atomicAdd(err, 1ull);
                for (int q = 0; q < NC; ++q)
                {
                    double v = 0.0;
                    if (ok)
                    {
                        const size_t o = (size_t)q * c.v.total + ci;
                        v = (double)c.cur[o] +
                            (c.cur_lo ? (double)c.cur_lo[o] : 0.0);
                        if (c.snap && c.t_frac < 1.0)
                        {
                            const double sv =
                                (double)c.snap[o] +
                                (c.snap_lo ? (double)c.snap_lo[o] : 0.0);
                            v = (1.0 - c.t_frac) * sv + c.t_frac * v;
                        }
                    }
                    ax[q] += wx[i] * v;
                }
            }
            for (int q = 0; q < NC; ++q)
                ay[q] += wy[j] * ax[q];
        }
        for (int q = 0; q < NC; ++q)
            acc[q] += wz[k] * ay[q];
    }
    for (int q = 0; q < NC; ++q)
        out[q] = acc[q];
}

// ghost fill of one buffer: each ghost cell of the shell of width w (<= the
// stored ghost width) from the box that owns it on this level, else
// (coarse-fine boundary) by quartic interpolation from the coarse level;
// components 0 .. NC-1.  The RK stages fill only what their stencil reads,
// psi (NC = 2) in a shell of 2; the full fill is NC = 4, w = ng.
template <typename T, int NC>
__global__ void k_fill(LevelView f, T *buf, float *lo, int have_coarse,
                       CoarseSrc<T> c, int w, unsigned long long *err)
{
    const BoxD B = f.box[blockIdx.y];
    const int ng = f.ng;
    const uint32_t gx = B.n[0] + 2 * w, gy = B.n[1] + 2 * w,
                   gz = B.n[2] + 2 * w;
    const uint32_t cells =
        gx * gy * gz - (uint32_t)B.n[0] * B.n[1] * B.n[2];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += gridDim.x * blockDim.x)
    {
        int lx, ly, lz;
        ghost_cell(t, B.n[0], B.n[1], B.n[2], w, lx, ly, lz);
        const uint32_t dst = sidx(B, ng, lx, ly, lz);
        const int X = B.lo[0] + lx, Y = B.lo[1] + ly, Z = B.lo[2] + lz;
        uint32_t src;
        if (locate(f, X, Y, Z, src) >= 0)
        {
            for (int comp = 0; comp < NC; ++comp)
            {
                buf[(size_t)comp * f.total + dst] = buf[(size_t)comp * f.total + src];
                if (lo)
                    lo[(size_t)comp * f.total + dst] = lo[(size_t)comp * f.total + src];
            }
            continue;
        }
        if (!have_coarse)
        {
            atomicAdd(err, 1ull);
            continue;
        }
        const int px = pmod(X, f.dom[0]), py = pmod(Y, f.dom[1]),
                  pz = pmod(Z, f.dom[2]);
        double v4[NC];
        interp_quartic<T, NC>(c, px, py, pz, v4, err);
        for (int comp = 0; comp < NC; ++comp)
        {
            const double v = v4[comp];
            const T h = (T)v;
            buf[(size_t)comp * f.total + dst] = h;
            if (lo)
                lo[(size_t)comp * f.total + dst] = (float)(v - (double)h);
        }
    }
}

// --------------------------------------------------- work-list stage ----
// Items of box b: rows y0, y0 + H, ... of H = clamp(rows, 4, 512 / W)
// rows each (W = the box's x edge, 8..64); a lane row of the item loads the
// y halo below (its first two rows) or above (its last two), so H >= 4.
static std::vector<WorkItem> make_items(const std::vector<BoxD> &boxes,
                                        int rows)
{
    std::vector<WorkItem> items;
    for (size_t i = 0; i < boxes.size(); ++i)
    {
        const int W = boxes[i].n[0];
        const int H = std::max(4, std::min(rows, kBoxLanes / W));
        for (int y0 = 0; y0 < boxes[i].n[1]; y0 += H)
            items.push_back(WorkItem{(int)i, y0, H, W});
    }
    return items;
}

template <typename T> struct BoxStageParams
{
    const BoxD *box;
    const WorkItem *items;
    int nitems;
    int ng;
    uint32_t total;
    const T *yin;
    const T *yn;
    T *acc;
    T *yout;
    float *comp;
    const double *coeff; // 12 doubles: stage s at [3s .. 3s+2]
    int coeff_stage;
    double a, h6;
    float h6_lo;
    double inv_dx2;
};

template <typename T> struct StarOfB;
template <> struct StarOfB<double>
{
    using type = Star;
};
template <> struct StarOfB<float>
{
    using type = StarF;
};

template <typename T, int PREC, int MODE>
__global__ void __launch_bounds__(kBoxLanes) k_box_stage(BoxStageParams<T> p)
{
    __shared__ T sm[2][2][kBoxLds];
    const double *cf = p.coeff + 3 * p.coeff_stage;
    const double cc = cf[0], lam = cf[1], R2 = cf[2];
    const T i2 = (T)p.inv_dx2;
    const size_t tot = p.total;
    const int ng = p.ng;
    for (int it = blockIdx.x; it < p.nitems; it += gridDim.x)
    {
        const WorkItem w = p.items[it];
        const BoxD B = p.box[w.box];
        const int W = w.w, H = w.h, nyb = B.n[1], nzb = B.n[2];
        const int SW = W + 4;
        const int t = threadIdx.x;
        const int lx = t % W, ly = t / W;
        const bool lane = ly < H; // has a tile slot
        const int yy = w.y0 + ly;
        const bool active = lane && yy < nyb;
        const int yl = min(yy, nyb + 1); // clamped load row (in the ghosts)
        const uint32_t gxs = W + 2 * ng, gys = nyb + 2 * ng;
        const uint32_t plane = gxs * gys;
        auto at = [&](int x, int y, int z) -> uint32_t {
            return B.off + ((uint32_t)(z + ng) * gys + (uint32_t)(y + ng)) * gxs +
                   (uint32_t)(x + ng);
        };
        const bool hxl = lane && lx < 2, hxr = lane && lx >= W - 2;
        const bool hyl = lane && ly < 2, hyr = lane && ly >= H - 2;
        const int hxx = hxl ? lx - 2 : W + (lx - (W - 2));
        const int hyy = hyl ? w.y0 + ly - 2
                            : min(w.y0 + H + (ly - (H - 2)), nyb + 1);
        const T *p1 = p.yin, *p2 = p.yin + tot;
        T q1[5] = {0, 0, 0, 0, 0}, q2[5] = {0, 0, 0, 0, 0};
        if (lane)
            for (int i = 0; i < 5; ++i)
            {
                const uint32_t o = at(lx, yl, i - 2);
                q1[i] = p1[o];
                q2[i] = p2[o];
            }
        T hx1 = 0, hx2 = 0, hy1 = 0, hy2 = 0;
        if (hxl || hxr)
        {
            hx1 = p1[at(hxx, yl, 0)];
            hx2 = p2[at(hxx, yl, 0)];
        }
        if (hyl || hyr)
        {
            hy1 = p1[at(lx, hyy, 0)];
            hy2 = p2[at(lx, hyy, 0)];
        }
        for (int z = 0; z < nzb; ++z)
        {
            const int b = z & 1;
            T *s0 = sm[b][0], *s1 = sm[b][1];
            if (lane)
            {
                s0[(ly + 2) * SW + lx + 2] = q1[2];
                s1[(ly + 2) * SW + lx + 2] = q2[2];
                if (hxl)
                {
                    s0[(ly + 2) * SW + lx] = hx1;
                    s1[(ly + 2) * SW + lx] = hx2;
                }
                else if (hxr)
                {
                    s0[(ly + 2) * SW + lx + 4] = hx1;
                    s1[(ly + 2) * SW + lx + 4] = hx2;
                }
                if (hyl)
                {
                    s0[ly * SW + lx + 2] = hy1;
                    s1[ly * SW + lx + 2] = hy2;
                }
                else if (hyr)
                {
                    s0[(ly + 4) * SW + lx + 2] = hy1;
                    s1[(ly + 4) * SW + lx + 2] = hy2;
                }
            }
            T nq1 = q1[4], nq2 = q2[4];
            if (lane && z + 1 < nzb)
            {
                const uint32_t o3 = at(lx, yl, z + 3);
                nq1 = p1[o3];
                nq2 = p2[o3];
                if (hxl || hxr)
                {
                    hx1 = p1[at(hxx, yl, z + 1)];
                    hx2 = p2[at(hxx, yl, z + 1)];
                }
                if (hyl || hyr)
                {
                    hy1 = p1[at(lx, hyy, z + 1)];
                    hy2 = p2[at(lx, hyy, z + 1)];
                }
            }
            const uint32_t oc = active ? at(lx, yy, z) : 0;
            T pi[2] = {0, 0}, yn[4] = {0, 0, 0, 0}, ac[4] = {0, 0, 0, 0};
            float lo[4] = {0, 0, 0, 0};
            if (active)
            {
                pi[0] = p.yin[2 * tot + oc];
                pi[1] = p.yin[3 * tot + oc];
                if (MODE == MODE_S23 || MODE == MODE_S4)
                    for (int c = 0; c < 4; ++c)
                    {
                        yn[c] = p.yn[c * tot + oc];
                        ac[c] = p.acc[c * tot + oc];
                    }
                if (MODE == MODE_S4 && PREC == AXS_FP32C)
                    for (int c = 0; c < 4; ++c)
                        lo[c] = p.comp[c * tot + oc];
            }
            __syncthreads();
            if (active)
            {
                const int r = (ly + 2) * SW + lx + 2;
                using S = typename StarOfB<T>::type;
                const S a1{s0[r - 2], s0[r - 1], s0[r + 1], s0[r + 2],
                           s0[r - 2 * SW], s0[r - SW], s0[r + SW], s0[r + 2 * SW],
                           q1[0], q1[1], q1[3], q1[4], q1[2]};
                const S a2{s1[r - 2], s1[r - 1], s1[r + 1], s1[r + 2],
                           s1[r - 2 * SW], s1[r - SW], s1[r + SW], s1[r + 2 * SW],
                           q2[0], q2[1], q2[3], q2[4], q2[2]};
                T F1, F2;
                force(star_lap(a1, i2), star_lap(a2, i2), q1[2], q2[2], cc, lam,
                      R2, F1, F2);
                const T k[4] = {pi[0], pi[1], F1, F2};
                if (MODE == MODE_S1)
                {
                    yn[0] = q1[2];
                    yn[1] = q2[2];
                    yn[2] = pi[0];
                    yn[3] = pi[1];
                }
                T acc_out[4], y_out[4];
                stage_update<T, PREC, MODE>(k, yn, ac, lo, p.a, p.h6, p.h6_lo,
                                            acc_out, y_out);
                for (int c = 0; c < 4; ++c)
                {
                    p.yout[c * tot + oc] = y_out[c];
                    if (MODE == MODE_S1 || MODE == MODE_S23)
                        p.acc[c * tot + oc] = acc_out[c];
                    if (MODE == MODE_S4 && PREC == AXS_FP32C)
                        p.comp[c * tot + oc] = lo[c];
                }
            }
            for (int i = 0; i < 4; ++i)
            {
                q1[i] = q1[i + 1];
                q2[i] = q2[i + 1];
            }
            q1[4] = nq1;
            q2[4] = nq2;
            (void)plane;
        }
        __syncthreads(); // LDS reuse by the next item
    }
}

// ------------------------------------------------- fused stage pairs --
// Two launches per RK4 step (stages 1+2, stages 3+4) instead of four, with
// the bits of the per-stage launches.  The second stage of each pair needs
// its input psi on the 13-point star.  At a cell of the box (and at the
// star's points inside the box) that psi is pointwise in what the pair
// reads -- psi_1 = psi_n + (dt/2) Pi_n, psi_3 = psi_n + dt Pi_2, formed by
// the expression the first stage stores -- and at the star's ghost points
// it comes from the ghost shell of a scratch array, filled before the pair
// with what the per-stage fill of that stage's input writes there: a ghost
// cell owned by a box of this level gets the value that box's first stage
// forms there, stage_axpy(psi, a, Pi) of the owner's psi and Pi; a
// coarse-fine ghost gets the quartic interpolation at the stage time.  The
// first stage of a pair reads its input as the per-stage kernel does (psi_n,
// or y_2, whose ghosts the same fill writes as k_fill<T, 2> would).
//
// Both ghost fills of a fused pair in one pass over the psi shells (width
// 2): the first stage's psi into dstA's shell (what k_fill<T, 2> of srcA
// writes at tA) and the second stage's psi into dstB's shell
// (stage_axpy(psi of ps, a, Pi of pp) of the owner cell, or the coarse
// interpolation at tB).  Each ghost cell's owner is looked up once; at a
// coarse-fine cell the coarse stencil is read once and blended at both
// times, each sum in the order of interp_quartic, so both outputs have the
// bits of the per-stage fills of those stage inputs.
//
// With x-edge arrays (eA, eps, epp: the x-edge columns of srcA, ps and pp;
// null without), a ghost cell of an x face (y and z inside the box) owned
// by a box of this level reads its source there: the same values, packed.
template <typename T, int PREC>
__global__ void k_fill_pair(LevelView f, T *dstA, const T *srcA, T *dstB,
                            const T *ps, const T *pp, double a,
                            int have_coarse, CoarseSrc<T> c, double tA,
                            double tB, const T *eA, const T *eps,
                            const T *epp, size_t es, unsigned long long *err)
{
    const BoxD B = f.box[blockIdx.y];
    const int ng = f.ng, w = 2;
    const uint32_t gx = B.n[0] + 2 * w, gy = B.n[1] + 2 * w,
                   gz = B.n[2] + 2 * w;
    const uint32_t cells =
        gx * gy * gz - (uint32_t)B.n[0] * B.n[1] * B.n[2];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += gridDim.x * blockDim.x)
    {
        int lx, ly, lz;
        ghost_cell(t, B.n[0], B.n[1], B.n[2], w, lx, ly, lz);
        const uint32_t d = sidx(B, ng, lx, ly, lz);
        const int X = B.lo[0] + lx, Y = B.lo[1] + ly, Z = B.lo[2] + lz;
        const int xw = pmod(X, f.dom[0]), yw = pmod(Y, f.dom[1]),
                  zw = pmod(Z, f.dom[2]);
        const int own = f.owner[((zw / kBlock) * f.bdim[1] + yw / kBlock) *
                                    f.bdim[0] +
                                xw / kBlock];
        if (own >= 0)
        {
            const BoxD S = f.box[own];
            const int sx = xw - S.lo[0], sy = yw - S.lo[1], sz = zw - S.lo[2];
            const bool xface = ly >= 0 && ly < B.n[1] && lz >= 0 && lz < B.n[2];
            if (eA && xface && (sx < 2 || sx >= S.n[0] - 2))
            {
                const size_t e = edge_idx(S, sy, sz, sx < 2 ? sx : sx - S.n[0] + 4);
                for (int comp = 0; comp < 2; ++comp)
                {
                    const size_t o = (size_t)comp * f.total;
                    dstA[o + d] = eA[(size_t)comp * es + e];
                    dstB[o + d] = stage_axpy<T, PREC>(
                        eps[(size_t)comp * es + e], a,
                        epp[(size_t)(2 + comp) * es + e]);
                }
                continue;
            }
            const uint32_t src = sidx(S, ng, sx, sy, sz);
            for (int comp = 0; comp < 2; ++comp)
            {
                const size_t o = (size_t)comp * f.total;
                dstA[o + d] = srcA[o + src];
                dstB[o + d] = stage_axpy<T, PREC>(
                    ps[o + src], a, pp[(size_t)(2 + comp) * f.total + src]);
            }
            continue;
        }
        if (!have_coarse)
        {
            atomicAdd(err, 2ull);
            continue;
        }
        const int cx = pmod(X, f.dom[0]) >> 1, cy = pmod(Y, f.dom[1]) >> 1,
                  cz = pmod(Z, f.dom[2]) >> 1;
        const double *wx = kQuartW[pmod(X, f.dom[0]) & 1],
                     *wy = kQuartW[pmod(Y, f.dom[1]) & 1],
                     *wz = kQuartW[pmod(Z, f.dom[2]) & 1];
        double accA[2] = {}, accB[2] = {};
        for (int k = 0; k < 5; ++k)
        {
            double ayA[2] = {}, ayB[2] = {};
            for (int j = 0; j < 5; ++j)
            {
                double axA[2] = {}, axB[2] = {};
                for (int i = 0; i < 5; ++i)
                {
                    uint32_t ci;
                    const bool ok = locate(c.v, cx + i - 2, cy + j - 2,
                                           cz + k - 2, ci) >= 0;
                    if (!ok)
                        atomicAdd(err, 2ull);
                    for (int q = 0; q < 2; ++q)
                    {
                        double vA = 0.0, vB = 0.0;
                        if (ok)
                        {
                            const size_t o = (size_t)q * c.v.total + ci;
                            const double cur = (double)c.cur[o] +
                                               (c.cur_lo ? (double)c.cur_lo[o] : 0.0);
                            vA = vB = cur;
                            if (c.snap)
                            {
                                const double sv =
                                    (double)c.snap[o] +
                                    (c.snap_lo ? (double)c.snap_lo[o] : 0.0);
                                if (tA < 1.0)
                                    vA = (1.0 - tA) * sv + tA * cur;
                                if (tB < 1.0)
                                    vB = (1.0 - tB) * sv + tB * cur;
                            }
                        }
                        axA[q] += wx[i] * vA;
                        axB[q] += wx[i] * vB;
                    }
                }
                for (int q = 0; q < 2; ++q)
                {
                    ayA[q] += wy[j] * axA[q];
                    ayB[q] += wy[j] * axB[q];
                }
            }
            for (int q = 0; q < 2; ++q)
            {
                accA[q] += wz[k] * ayA[q];
                accB[q] += wz[k] * ayB[q];
            }
        }
        for (int comp = 0; comp < 2; ++comp)
        {
            dstA[(size_t)comp * f.total + d] = (T)accA[comp];
            dstB[(size_t)comp * f.total + d] = (T)accB[comp];
        }
    }
}

// The x-edge columns (0, 1, W-2, W-1) of every row of every box of y, 4
// components, into the packed x-edge arrays (grid: x over a box's rows x 4,
// y over the boxes).
template <typename T>
__global__ void k_edges(const BoxD *box, int ng, uint32_t total, const T *y,
                        T *edge, size_t es)
{
    const BoxD B = box[blockIdx.y];
    const uint32_t n = 4u * (uint32_t)B.n[1] * (uint32_t)B.n[2];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += gridDim.x * blockDim.x)
    {
        const int col = (int)(t & 3);
        const uint32_t row = t >> 2;
        const int sy = (int)(row % (uint32_t)B.n[1]),
                  sz = (int)(row / (uint32_t)B.n[1]);
        const int sx = col < 2 ? col : B.n[0] - 4 + col;
        const uint32_t src = sidx(B, ng, sx, sy, sz);
        const size_t e = edge_idx(B, sy, sz, col);
        for (int c = 0; c < 4; ++c)
            edge[c * es + e] = y[(size_t)c * total + src];
    }
}

template <typename T> struct BoxPairParams
{
    const BoxD *box;
    const WorkItem *items;
    int nitems;
    int ng;
    uint32_t total;
    const T *yn;  // y_n
    const T *y2;  // pair 2: y_2, the stage-3 input (psi ghosts filled)
    T *y2out;     // pair 1: y_2 written
    const T *gb;  // psi ghost shell of the second stage's input
    T *acc;       // pair 1: k1 + 2 k2 written; pair 2: read, then y_{n+1}
                  // written over it (own cell only)
    float *comp;  // FP32C low words of y_n (pair 2, own cell, in place)
    const double *coeff; // 12 doubles: stage s at [3s .. 3s+2]
    double dt, h6;
    float h6_lo;
    double inv_dx2;
    T *edge;   // x-edge arrays of the pair's output (pair 1: y_2; pair 2:
               // y_{n+1}), or null
    size_t es; // their component stride
};

// PAIR 1: stages 1+2 (star fields psi_n and psi_1); PAIR 2: stages 3+4
// (psi_2 and psi_3).  Work items, LDS planes and register queues as in
// k_box_stage, with two star fields.
template <typename T, int PREC, int PAIR>
__global__ void __launch_bounds__(kBoxLanes) k_box_pair(BoxPairParams<T> p)
{
    __shared__ T sm[2][4][kBoxLds]; // [buffer][A1 A2 B1 B2]
    const int sa_ = PAIR == 1 ? 0 : 2, sb_ = sa_ + 1; // coefficient stages
    const double *cfa = p.coeff + 3 * sa_, *cfb = p.coeff + 3 * sb_;
    const T i2 = (T)p.inv_dx2;
    const size_t tot = p.total;
    const int ng = p.ng;
    // first stage's step (a of its RK update) and the second stage's psi
    // formation: pair 1 a = dt/2 for both; pair 2 a = dt
    const double a = PAIR == 1 ? 0.5 * p.dt : p.dt;
    const T *pA = PAIR == 1 ? p.yn : p.y2; // first stage's input
    for (int it = blockIdx.x; it < p.nitems; it += gridDim.x)
    {
        const WorkItem w = p.items[it];
        const BoxD B = p.box[w.box];
        const int W = w.w, H = w.h, nyb = B.n[1], nzb = B.n[2];
        const int SW = W + 4;
        const int t = threadIdx.x;
        const int lx = t % W, ly = t / W;
        const bool lane = ly < H;
        const int yy = w.y0 + ly;
        const bool active = lane && yy < nyb;
        const int yl = min(yy, nyb + 1);
        const uint32_t gxs = W + 2 * ng, gys = nyb + 2 * ng;
        auto at = [&](int x, int y, int z) -> uint32_t {
            return B.off + ((uint32_t)(z + ng) * gys + (uint32_t)(y + ng)) * gxs +
                   (uint32_t)(x + ng);
        };
        // field A (the first stage's psi) and B (the second stage's psi)
        // of the point (x, y, z) of the box's grown region
        auto load = [&](int x, int y, int z, T (&va)[2], T (&vb)[2]) {
            const uint32_t o = at(x, y, z);
            const bool ghost =
                x < 0 || x >= W || y < 0 || y >= nyb || z < 0 || z >= nzb;
            for (int c = 0; c < 2; ++c)
            {
                va[c] = pA[c * tot + o];
                if (ghost)
                    vb[c] = p.gb[c * tot + o];
                else if (PAIR == 1)
                    vb[c] = stage_axpy<T, PREC>(va[c], a, p.yn[(2 + c) * tot + o]);
                else
                    vb[c] = stage_axpy<T, PREC>(p.yn[c * tot + o], a,
                                                p.y2[(2 + c) * tot + o]);
            }
        };
        // the same for the lane's own column, also keeping the values read
        // to form field B (pair 1: Pi_n; pair 2: psi_n, Pi_2), which the
        // stage update of that cell uses again when it reaches the centre
        // of the queue (so they are not read twice).  FP64 pair 2 reads them
        // again instead: its queue would take it past 128 VGPRs.
        constexpr int NR = PAIR == 1 ? 2 : 4;
        constexpr bool keep = PAIR == 1 || sizeof(T) == 4;
        auto load_own = [&](int x, int y, int z, T (&va)[2], T (&vb)[2],
                            T (&vr)[NR]) {
            const uint32_t o = at(x, y, z);
            const bool ghost =
                x < 0 || x >= W || y < 0 || y >= nyb || z < 0 || z >= nzb;
            for (int c = 0; c < 2; ++c)
            {
                va[c] = pA[c * tot + o];
                if (ghost)
                    vb[c] = p.gb[c * tot + o];
                else if constexpr (PAIR == 1)
                {
                    const T pin = p.yn[(2 + c) * tot + o];
                    vr[c] = pin;
                    vb[c] = stage_axpy<T, PREC>(va[c], a, pin);
                }
                else
                {
                    const T psn = p.yn[c * tot + o], pi2v = p.y2[(2 + c) * tot + o];
                    if constexpr (keep)
                    {
                        vr[c] = psn;
                        vr[2 + c] = pi2v;
                    }
                    vb[c] = stage_axpy<T, PREC>(psn, a, pi2v);
                }
            }
        };
        const bool hxl = lane && lx < 2, hxr = lane && lx >= W - 2;
        const bool hyl = lane && ly < 2, hyr = lane && ly >= H - 2;
        const int hxx = hxl ? lx - 2 : W + (lx - (W - 2));
        const int hyy = hyl ? w.y0 + ly - 2
                            : min(w.y0 + H + (ly - (H - 2)), nyb + 1);
        T qa[5][2] = {}, qb[5][2] = {}, qr[5][NR] = {};
        if (lane)
            for (int i = 0; i < 5; ++i)
                load_own(lx, yl, i - 2, qa[i], qb[i], qr[i]);
        T hxa[2] = {0, 0}, hxb[2] = {0, 0}, hya[2] = {0, 0}, hyb[2] = {0, 0};
        if (hxl || hxr)
            load(hxx, yl, 0, hxa, hxb);
        if (hyl || hyr)
            load(lx, hyy, 0, hya, hyb);
        for (int z = 0; z < nzb; ++z)
        {
            const int bf = z & 1;
            T *s[4] = {sm[bf][0], sm[bf][1], sm[bf][2], sm[bf][3]};
            if (lane)
            {
                const int r = (ly + 2) * SW + lx + 2;
                for (int c = 0; c < 2; ++c)
                {
                    s[c][r] = qa[2][c];
                    s[2 + c][r] = qb[2][c];
                }
                if (hxl || hxr)
                {
                    const int rx = (ly + 2) * SW + (hxl ? lx : lx + 4);
                    for (int c = 0; c < 2; ++c)
                    {
                        s[c][rx] = hxa[c];
                        s[2 + c][rx] = hxb[c];
                    }
                }
                if (hyl || hyr)
                {
                    const int ry = (hyl ? ly : ly + 4) * SW + lx + 2;
                    for (int c = 0; c < 2; ++c)
                    {
                        s[c][ry] = hya[c];
                        s[2 + c][ry] = hyb[c];
                    }
                }
            }
            T nqa[2] = {qa[4][0], qa[4][1]}, nqb[2] = {qb[4][0], qb[4][1]};
            T nqr[NR];
            for (int c = 0; c < NR; ++c)
                nqr[c] = qr[4][c];
            if (lane && z + 1 < nzb)
            {
                load_own(lx, yl, z + 3, nqa, nqb, nqr);
                if (hxl || hxr)
                    load(hxx, yl, z + 1, hxa, hxb);
                if (hyl || hyr)
                    load(lx, hyy, z + 1, hya, hyb);
            }
            const uint32_t oc = active ? at(lx, yy, z) : 0;
            T yn[4] = {0, 0, 0, 0}, ac[4] = {0, 0, 0, 0}, pi2[2] = {0, 0};
            float lo[4] = {0, 0, 0, 0};
            if (active)
            {
                // the cell's own y_n (and Pi_2): from the queues where the
                // loads of field B put them, Pi_n of pair 2 from memory
                if constexpr (PAIR == 1)
                {
                    yn[0] = qa[2][0];
                    yn[1] = qa[2][1];
                    yn[2] = qr[2][0];
                    yn[3] = qr[2][1];
                }
                else if constexpr (keep)
                {
                    yn[0] = qr[2][0];
                    yn[1] = qr[2][1];
                    yn[2] = p.yn[2 * tot + oc];
                    yn[3] = p.yn[3 * tot + oc];
                }
                else
                    for (int c = 0; c < 4; ++c)
                        yn[c] = p.yn[c * tot + oc];
                if (PAIR == 2)
                {
                    if constexpr (keep)
                    {
                        pi2[0] = qr[2][2];
                        pi2[1] = qr[2][3];
                    }
                    else
                    {
                        pi2[0] = p.y2[2 * tot + oc];
                        pi2[1] = p.y2[3 * tot + oc];
                    }
                    for (int c = 0; c < 4; ++c)
                        ac[c] = p.acc[c * tot + oc];
                    if (PREC == AXS_FP32C)
                        for (int c = 0; c < 4; ++c)
                            lo[c] = p.comp[c * tot + oc];
                }
            }
            __syncthreads();
            if (active)
            {
                const int r = (ly + 2) * SW + lx + 2;
                using S = typename StarOfB<T>::type;
                auto star = [&](const T *f, const T (&q)[5][2], int c) {
                    return S{f[r - 2], f[r - 1], f[r + 1], f[r + 2],
                             f[r - 2 * SW], f[r - SW], f[r + SW], f[r + 2 * SW],
                             q[0][c], q[1][c], q[3][c], q[4][c], q[2][c]};
                };
                T FA1, FA2, FB1, FB2;
                force(star_lap(star(s[0], qa, 0), i2),
                      star_lap(star(s[1], qa, 1), i2), qa[2][0], qa[2][1],
                      cfa[0], cfa[1], cfa[2], FA1, FA2);
                force(star_lap(star(s[2], qb, 0), i2),
                      star_lap(star(s[3], qb, 1), i2), qb[2][0], qb[2][1],
                      cfb[0], cfb[1], cfb[2], FB1, FB2);
                if constexpr (PAIR == 1)
                {
                    // stage 1 (input y_n), stage 2 (input y_1)
                    const T k1[4] = {yn[2], yn[3], FA1, FA2};
                    T acc1[4], y1[4];
                    stage_update<T, PREC, MODE_S1>(k1, yn, ac, lo, 0.5 * p.dt,
                                                   p.h6, p.h6_lo, acc1, y1);
                    const T k2[4] = {y1[2], y1[3], FB1, FB2};
                    T acc2[4], y2[4];
                    stage_update<T, PREC, MODE_S23>(k2, yn, acc1, lo,
                                                    0.5 * p.dt, p.h6, p.h6_lo,
                                                    acc2, y2);
                    for (int c = 0; c < 4; ++c)
                    {
                        p.y2out[c * tot + oc] = y2[c];
                        p.acc[c * tot + oc] = acc2[c];
                    }
                    if (p.edge && (lx < 2 || lx >= W - 2))
                    {
                        const size_t e =
                            edge_idx(B, yy, z, lx < 2 ? lx : lx - W + 4);
                        for (int c = 0; c < 4; ++c)
                            p.edge[c * p.es + e] = y2[c];
                    }
                }
                else
                {
                    // stage 3 (input y_2), stage 4 (input y_3)
                    const T k3[4] = {pi2[0], pi2[1], FA1, FA2};
                    T acc3[4], y3[4];
                    stage_update<T, PREC, MODE_S23>(k3, yn, ac, lo, p.dt, p.h6,
                                                    p.h6_lo, acc3, y3);
                    const T k4[4] = {y3[2], y3[3], FB1, FB2};
                    T unused[4], y4[4];
                    stage_update<T, PREC, MODE_S4>(k4, yn, acc3, lo, p.dt,
                                                   p.h6, p.h6_lo, unused, y4);
                    for (int c = 0; c < 4; ++c)
                    {
                        p.acc[c * tot + oc] = y4[c];
                        if (PREC == AXS_FP32C)
                            p.comp[c * tot + oc] = lo[c];
                    }
                    if (p.edge && (lx < 2 || lx >= W - 2))
                    {
                        const size_t e =
                            edge_idx(B, yy, z, lx < 2 ? lx : lx - W + 4);
                        for (int c = 0; c < 4; ++c)
                            p.edge[c * p.es + e] = y4[c];
                    }
                }
            }
            for (int i = 0; i < 4; ++i)
            {
                for (int c = 0; c < 2; ++c)
                {
                    qa[i][c] = qa[i + 1][c];
                    qb[i][c] = qb[i + 1][c];
                }
                for (int c = 0; c < NR; ++c)
                    qr[i][c] = qr[i + 1][c];
            }
            for (int c = 0; c < 2; ++c)
            {
                qa[4][c] = nqa[c];
                qb[4][c] = nqb[c];
            }
            for (int c = 0; c < NR; ++c)
                qr[4][c] = nqr[c];
        }
        __syncthreads(); // LDS reuse by the next item
    }
}

// -------------------------------------------------- windings / tagging --
__device__ __forceinline__ double wrap_diff(double delta)
{
    const double pi = 3.14159265358979323846;
    const double two_pi = 2.0 * pi;
    delta = fmod(delta, two_pi);
    if (delta <= -pi)
        delta += two_pi;
    else if (delta > pi)
        delta -= two_pi;
    return delta;
}
__device__ __forceinline__ int pwind(double t1, double t2, double t3, double t4)
{
#pragma clang fp contract(off)
    const double two_pi = 2.0 * 3.14159265358979323846;
    const double total = wrap_diff(t2 - t1) + wrap_diff(t3 - t2) +
                         wrap_diff(t4 - t3) + wrap_diff(t1 - t4);
    return (int)lround(total / two_pi);
}

// FourthOrderDerivatives diff1 / diff2 (published 4th-order forms), the
// same arithmetic as the uniform-level tagger
__device__ __forceinline__ double tag_d1(double fm2, double fm1, double fp1,
                                         double fp2, double i1)
{
#pragma clang fp contract(off)
    const double w_far = 8.33333333333333333333e-2;
    const double w_near = 6.66666666666666666667e-1;
    return (w_far * fm2 - w_near * fm1 + w_near * fp1 - w_far * fp2) * i1;
}
__device__ __forceinline__ double tag_d2(double fm2, double fm1, double f0,
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

struct BoxTagParams
{
    double amp_thr, R;
    // StringTagger.hpp:83-218 derivative criteria (DBL_MAX = off)
    double grad_thr, rgrad_thr, dx;
};

// The windings of the three plaquettes at a cell's low corner
// (StringFinder.hpp:76-122) and, when `criteria`, the tagger's criteria at
// the cell (StringTagger.hpp:83-274, plus the amplitude extension).
// at(c, a, b, d): component c at offset (a, b, d) from the cell, as hi + lo
// in double.
template <typename At>
__device__ __forceinline__ bool tag_cell(const At &at, const BoxTagParams &tp,
                                         bool criteria, int &wxy, int &wyz,
                                         int &wzx)
{
    auto ph = [&](int a, int b, int c) -> double {
        return atan2(at(1, a, b, c), at(0, a, b, c));
    };
    const double t000 = ph(0, 0, 0), t100 = ph(1, 0, 0), t110 = ph(1, 1, 0),
                 t010 = ph(0, 1, 0), t001 = ph(0, 0, 1), t101 = ph(1, 0, 1),
                 t011 = ph(0, 1, 1);
    wxy = pwind(t000, t100, t110, t010);
    wyz = pwind(t000, t010, t011, t001);
    wzx = pwind(t000, t001, t101, t100);
    if (!criteria)
        return false;
    bool tag = wxy != 0 || wyz != 0 || wzx != 0;
    const double psi1 = at(0, 0, 0, 0), psi2 = at(1, 0, 0, 0);
    if (!tag && tp.amp_thr > 0.0)
        tag = sqrt(psi1 * psi1 + psi2 * psi2) / tp.R < tp.amp_thr;
    const double i1 = 1.0 / tp.dx, i2 = 1.0 / (tp.dx * tp.dx);
    if (!tag && tp.grad_thr < DBL_MAX)
    {
        // dx^2 |laplacian psi_c| > threshold
        const double dx2 = tp.dx * tp.dx;
        double lap[2];
        for (int c = 0; c < 2; ++c)
        {
            const double f0 = at(c, 0, 0, 0);
            double rr = 0.0;
            rr += tag_d2(at(c, -2, 0, 0), at(c, -1, 0, 0), f0, at(c, 1, 0, 0),
                         at(c, 2, 0, 0), i2);
            rr += tag_d2(at(c, 0, -2, 0), at(c, 0, -1, 0), f0, at(c, 0, 1, 0),
                         at(c, 0, 2, 0), i2);
            rr += tag_d2(at(c, 0, 0, -2), at(c, 0, 0, -1), f0, at(c, 0, 0, 1),
                         at(c, 0, 0, 2), i2);
            lap[c] = rr;
        }
        if (dx2 * fabs(lap[0]) > tp.grad_thr || dx2 * fabs(lap[1]) > tp.grad_thr)
            tag = true;
    }
    if (!tag && tp.rgrad_thr < DBL_MAX)
    {
        // dx sqrt(|grad psi|^2 - |psi|^2 |grad theta|^2)
        const double psq = psi1 * psi1 + psi2 * psi2;
        double g1sq = 0.0, g2sq = 0.0, gtsq = 0.0;
        for (int d = 0; d < 3; ++d)
        {
            const int ex = d == 0, ey = d == 1, ez = d == 2;
            const double a =
                tag_d1(at(0, -2 * ex, -2 * ey, -2 * ez), at(0, -ex, -ey, -ez),
                       at(0, ex, ey, ez), at(0, 2 * ex, 2 * ey, 2 * ez), i1);
            const double b =
                tag_d1(at(1, -2 * ex, -2 * ey, -2 * ez), at(1, -ex, -ey, -ez),
                       at(1, ex, ey, ez), at(1, 2 * ex, 2 * ey, 2 * ez), i1);
            g1sq += a * a;
            g2sq += b * b;
            const double dth = psq > 0.0 ? (psi1 * b - psi2 * a) / psq : 0.0;
            gtsq += dth * dth;
        }
        if (tp.dx * sqrt(g1sq + g2sq - psq * gtsq) > tp.rgrad_thr)
            tag = true;
    }
    return tag;
}

// per-wave plaquette counts -> one atomic pair per wave
__device__ __forceinline__ void add_counts(unsigned long long plain,
                                           unsigned long long weighted,
                                           unsigned long long *counts)
{
    for (int o = 32; o > 0; o >>= 1)
        weighted += __shfl_xor(weighted, o, 64);
    if ((threadIdx.x & 63) == 0)
    {
        atomicAdd(&counts[0], plain);
        atomicAdd(&counts[1], weighted);
    }
}

template <typename T>
__global__ void k_box_tag(LevelView v, const T *y, const float *lo,
                          uint8_t *tags, const uint8_t *cover, BoxTagParams tp,
                          unsigned long long *counts)
{
    const BoxD B = v.box[blockIdx.y];
    const uint32_t cells = (uint32_t)B.n[0] * B.n[1] * B.n[2];
    const size_t tot = v.total;
    unsigned long long plain = 0, weighted = 0;
    const int iters = (cells + gridDim.x * blockDim.x - 1) / (gridDim.x * blockDim.x);
    for (int r = 0; r < iters; ++r)
    {
        const uint32_t t = (r * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
        int wxy = 0, wyz = 0, wzx = 0;
        if (t < cells)
        {
            const int lx = t % B.n[0], ly = (t / B.n[0]) % B.n[1],
                      lz = t / ((uint32_t)B.n[0] * B.n[1]);
            const int X = B.lo[0] + lx, Y = B.lo[1] + ly, Z = B.lo[2] + lz;
            const uint32_t dense =
                ((uint32_t)Z * v.dom[1] + Y) * v.dom[0] + X;
            if (!cover || cover[dense])
            {
                // component c at offset (a, b, d) from the cell
                auto at = [&](int c, int a, int b, int d) -> double {
                    const uint32_t i = sidx(B, v.ng, lx + a, ly + b, lz + d);
                    const size_t o = (size_t)c * tot + i;
                    return (double)y[o] + (lo ? (double)lo[o] : 0.0);
                };
                const bool tag = tag_cell(at, tp, tags != nullptr, wxy, wyz, wzx);
                weighted += abs(wxy) + abs(wyz) + abs(wzx);
                if (tag)
                    tags[dense] = 1;
            }
        }
        plain += __popcll(__ballot(wxy != 0)) + __popcll(__ballot(wyz != 0)) +
                 __popcll(__ballot(wzx != 0));
    }
    add_counts(plain, weighted, counts);
}

// ------------------------------------------------ borrowed box views ----
// An axs_array4 as the kernels see it: element (i, j, k, n) of a view at
// p[(i - b0) + (j - b1) js + (k - b2) ks + n ns].
template <typename T> struct View4
{
    T *p;
    int64_t js, ks, ns;
    int b0, b1, b2;
    __device__ __forceinline__ T &operator()(int i, int j, int k, int n) const
    {
        return p[(int64_t)(i - b0) + (int64_t)(j - b1) * js +
                 (int64_t)(k - b2) * ks + (int64_t)n * ns];
    }
};

template <typename T> static View4<T> view4(const axs_array4 &a)
{
    return View4<T>{(T *)a.p, a.jstride, a.kstride, a.nstride,
                    a.begin[0],  a.begin[1], a.begin[2]};
}

// AxionStringsRHS over the cells of one box: one thread per (i, j) column
// of a 64 x 4 tile, streaming z with psi's z neighbours in a register queue
// and the x / y neighbours read from the view (the L1 / L2 serve them).
// The per-cell arithmetic is that of the level kernels, so the bits are.
template <typename T>
__global__ void __launch_bounds__(256)
    k_rhs_array4(View4<const T> in, View4<T> out, int lo0, int lo1, int lo2,
                 int nx, int ny, int nz, int zchunk, double cc, double lam,
                 double R2, double inv_dx2)
{
    const int x = blockIdx.x * 64 + (int)threadIdx.x;
    const int y = blockIdx.y * 4 + (int)threadIdx.y;
    if (x >= nx || y >= ny)
        return;
    const int i = lo0 + x, j = lo1 + y;
    const int k0 = lo2 + (int)blockIdx.z * zchunk;
    const int k1 = min(k0 + zchunk, lo2 + nz);
    const T i2 = (T)inv_dx2;
    using S = typename std::conditional<std::is_same<T, double>::value, Star,
                                        StarF>::type;
    T q1[5], q2[5];
    for (int d = 0; d < 4; ++d)
    {
        q1[d] = in(i, j, k0 - 2 + d, 0);
        q2[d] = in(i, j, k0 - 2 + d, 1);
    }
    for (int k = k0; k < k1; ++k)
    {
        q1[4] = in(i, j, k + 2, 0);
        q2[4] = in(i, j, k + 2, 1);
        const S a1{in(i - 2, j, k, 0), in(i - 1, j, k, 0), in(i + 1, j, k, 0),
                   in(i + 2, j, k, 0), in(i, j - 2, k, 0), in(i, j - 1, k, 0),
                   in(i, j + 1, k, 0), in(i, j + 2, k, 0), q1[0],
                   q1[1],              q1[3],              q1[4],
                   q1[2]};
        const S a2{in(i - 2, j, k, 1), in(i - 1, j, k, 1), in(i + 1, j, k, 1),
                   in(i + 2, j, k, 1), in(i, j - 2, k, 1), in(i, j - 1, k, 1),
                   in(i, j + 1, k, 1), in(i, j + 2, k, 1), q2[0],
                   q2[1],              q2[3],              q2[4],
                   q2[2]};
        T F1, F2;
        force(star_lap(a1, i2), star_lap(a2, i2), q1[2], q2[2], cc, lam, R2, F1,
              F2);
        out(i, j, k, 0) = in(i, j, k, 2);
        out(i, j, k, 1) = in(i, j, k, 3);
        out(i, j, k, 2) = F1;
        out(i, j, k, 3) = F2;
        for (int d = 0; d < 4; ++d)
        {
            q1[d] = q1[d + 1];
            q2[d] = q2[d + 1];
        }
    }
}

// The tagger / windings pass over the cells of one box of a borrowed view
template <typename T>
__global__ void __launch_bounds__(256)
    k_tag_array4(View4<const T> in, View4<const float> lo, int has_lo,
                 View4<uint8_t> tags, int has_tags, int lo0, int lo1, int lo2,
                 int nx, int ny, int nz, BoxTagParams tp,
                 unsigned long long *counts)
{
    const uint32_t cells = (uint32_t)nx * ny * nz;
    unsigned long long plain = 0, weighted = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    const int iters = (int)((cells + stride - 1) / stride);
    for (int r = 0; r < iters; ++r)
    {
        const uint32_t t = (uint32_t)r * stride + blockIdx.x * blockDim.x +
                           threadIdx.x;
        int wxy = 0, wyz = 0, wzx = 0;
        if (t < cells)
        {
            const int i = lo0 + (int)(t % nx), j = lo1 + (int)((t / nx) % ny),
                      k = lo2 + (int)(t / ((uint32_t)nx * ny));
            auto at = [&](int c, int a, int b, int d) -> double {
                return (double)in(i + a, j + b, k + d, c) +
                       (has_lo ? (double)lo(i + a, j + b, k + d, c) : 0.0);
            };
            const bool tag = tag_cell(at, tp, has_tags != 0, wxy, wyz, wzx);
            weighted += abs(wxy) + abs(wyz) + abs(wzx);
            if (tag)
                tags(i, j, k, 0) = 1;
        }
        plain += __popcll(__ballot(wxy != 0)) + __popcll(__ballot(wyz != 0)) +
                 __popcll(__ballot(wzx != 0));
    }
    add_counts(plain, weighted, counts);
}

// ------------------------------------------------------------- energy ----
__device__ __forceinline__ double d1r(double fm2, double fm1, double fp1,
                                      double fp2, double i1)
{
#pragma clang fp contract(off)
    const double w_far = 8.33333333333333333333e-2;
    const double w_near = 6.66666666666666666667e-1;
    return (w_far * fm2 - w_near * fm1 + w_near * fp1 - w_far * fp2) * i1;
}

// compute_total_energy_sums over a box level's valid cells
// (EnergyKernel.hpp:106-233, Energy.hpp pointwise forms): a block streams an item's
// W x H rows through the box's planes with psi (FP64, FP32C as hi + lo) of
// the current plane in LDS with a 2-cell x / y halo (from the filled
// ghosts) and the z neighbours in a 5-deep register queue, so each psi value
// is read about once per plane instead of once per stencil use.  Per cell
// the 14 terms are the reference's, cells the coverage mask excludes
// skipped; the sums are fixed-order per block, then over blocks.
template <typename T>
__global__ void __launch_bounds__(kBoxLanes)
    k_box_energy_items(const BoxD *box, const WorkItem *items, int nitems,
                       int ng, uint32_t total, int domx, int domy,
                       const T *y, const float *lo, const uint8_t *cover,
                       double dx, double R, double lambda, double RpR,
                       int scheme, double thr, double *partials)
{
    __shared__ double sp[2][kBoxLds];
    __shared__ double red[kBoxLanes];
    const size_t tot = total;
    const double i1 = 1.0 / dx, R2 = R * R, R4 = R2 * R2, f_a = 1.4142135623730951;
    double acc[14];
    for (int q = 0; q < 14; ++q)
        acc[q] = 0.0;
    auto val = [&](int comp, uint32_t i) -> double {
        const size_t o = (size_t)comp * tot + i;
        return (double)y[o] + (lo ? (double)lo[o] : 0.0);
    };
    for (int it = blockIdx.x; it < nitems; it += gridDim.x)
    {
        const WorkItem w = items[it];
        const BoxD B = box[w.box];
        const int W = w.w, H = w.h, nyb = B.n[1], nzb = B.n[2];
        const int SW = W + 4;
        const int t = threadIdx.x;
        const int lx = t % W, ly = t / W;
        const bool lane = ly < H;
        const int yy = w.y0 + ly;
        const bool active = lane && yy < nyb;
        const int yl = min(yy, nyb + 1);
        const bool hxl = lane && lx < 2, hxr = lane && lx >= W - 2;
        const bool hyl = lane && ly < 2, hyr = lane && ly >= H - 2;
        const int hxx = hxl ? lx - 2 : W + (lx - (W - 2));
        const int hyy = hyl ? w.y0 + ly - 2
                            : min(w.y0 + H + (ly - (H - 2)), nyb + 1);
        double q1[5] = {}, q2[5] = {}; // own psi at planes z-2 .. z+2
        if (lane)
            for (int i = 0; i < 5; ++i)
            {
                const uint32_t o = sidx(B, ng, lx, yl, i - 2);
                q1[i] = val(0, o);
                q2[i] = val(1, o);
            }
        for (int z = 0; z < nzb; ++z)
        {
            if (lane)
            {
                const int r = (ly + 2) * SW + lx + 2;
                sp[0][r] = q1[2];
                sp[1][r] = q2[2];
                if (hxl || hxr)
                {
                    const uint32_t o = sidx(B, ng, hxx, yl, z);
                    const int rx = (ly + 2) * SW + (hxl ? lx : lx + 4);
                    sp[0][rx] = val(0, o);
                    sp[1][rx] = val(1, o);
                }
                if (hyl || hyr)
                {
                    const uint32_t o = sidx(B, ng, lx, hyy, z);
                    const int ry = (hyl ? ly : ly + 4) * SW + lx + 2;
                    sp[0][ry] = val(0, o);
                    sp[1][ry] = val(1, o);
                }
            }
            double n1 = 0.0, n2 = 0.0; // own psi of plane z + 3
            if (lane && z + 1 < nzb)
            {
                const uint32_t o = sidx(B, ng, lx, yl, z + 3);
                n1 = val(0, o);
                n2 = val(1, o);
            }
            __syncthreads();
            const uint32_t dense =
                active ? ((uint32_t)(B.lo[2] + z) * domy + (B.lo[1] + yy)) * domx +
                             (B.lo[0] + lx)
                       : 0;
            if (active && !(cover && !cover[dense]))
            {
                const uint32_t c = sidx(B, ng, lx, yy, z);
                const double psi1 = q1[2], psi2 = q2[2], Pi1 = val(2, c),
                             Pi2 = val(3, c);
                const double psi_sq = psi1 * psi1 + psi2 * psi2;
                const int r = (ly + 2) * SW + lx + 2;
                double g1sq = 0.0, g2sq = 0.0, gtsq = 0.0;
                for (int d = 0; d < 3; ++d)
                {
                    double a, b;
                    if (d == 0)
                    {
                        a = d1r(sp[0][r - 2], sp[0][r - 1], sp[0][r + 1],
                                sp[0][r + 2], i1);
                        b = d1r(sp[1][r - 2], sp[1][r - 1], sp[1][r + 1],
                                sp[1][r + 2], i1);
                    }
                    else if (d == 1)
                    {
                        a = d1r(sp[0][r - 2 * SW], sp[0][r - SW], sp[0][r + SW],
                                sp[0][r + 2 * SW], i1);
                        b = d1r(sp[1][r - 2 * SW], sp[1][r - SW], sp[1][r + SW],
                                sp[1][r + 2 * SW], i1);
                    }
                    else
                    {
                        a = d1r(q1[0], q1[1], q1[3], q1[4], i1);
                        b = d1r(q2[0], q2[1], q2[3], q2[4], i1);
                    }
                    g1sq += a * a;
                    g2sq += b * b;
                    const double dth =
                        psi_sq > 0.0 ? (psi1 * b - psi2 * a) / psi_sq : 0.0;
                    gtsq += dth * dth;
                }
                const double c1 = Pi1 - RpR * psi1, c2 = Pi2 - RpR * psi2;
                const double dp = psi_sq - R2;
                const double pot = 0.25 * lambda * dp * dp;
                const double rho = (c1 * c1 + c2 * c2 + (g1sq + g2sq) + pot) / R4;
                const double tp =
                    psi_sq > 0.0 ? (psi1 * Pi2 - Pi1 * psi2) / psi_sq : 0.0;
                const double adot = f_a * tp / R;
                const double rak = 0.5 * adot * adot;
                const double rag = 0.5 * f_a * f_a * gtsq / R2;
                double rrk = 0.0;
                if (psi_sq > 0.0)
                {
                    const double rd = (psi1 * c1 + psi2 * c2) / sqrt(psi_sq);
                    rrk = rd * rd / R4;
                }
                const double rrg = (g1sq + g2sq - psi_sq * gtsq) / R4;
                const double rrm = pot / R4;
                double wgt = 1.0;
                if (scheme != AXS_MASK_NONE)
                {
                    const double mod = sqrt(psi_sq) / R;
                    if (scheme == AXS_MASK_B)
                        wgt = mod >= thr ? 1.0 : 0.0;
                    else
                    {
                        const double f = 1.0 + (mod - 1.0) / f_a;
                        wgt = f * f;
                    }
                }
                const double vals[14] = {rho, rho * wgt, rak, rak * wgt, rag,
                                         rag * wgt, rrk, rrk * wgt, rrg,
                                         rrg * wgt, rrm, rrm * wgt, wgt, 1.0};
                for (int q = 0; q < 14; ++q)
                    acc[q] += vals[q];
            }
            __syncthreads(); // the plane's LDS is rewritten next
            for (int i = 0; i < 4; ++i)
            {
                q1[i] = q1[i + 1];
                q2[i] = q2[i + 1];
            }
            q1[4] = n1;
            q2[4] = n2;
        }
    }
    for (int q = 0; q < 14; ++q)
    {
        red[threadIdx.x] = acc[q];
        __syncthreads();
        for (int o = kBoxLanes / 2; o > 0; o >>= 1)
        {
            if ((int)threadIdx.x < o)
                red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            partials[(size_t)blockIdx.x * 14 + q] = red[0];
        __syncthreads();
    }
}

__global__ void k_box_sum_cols(const double *part, int nrow, int width,
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

// ------------------------------------------- masks, average down, copies --
__global__ void k_set_u8(uint8_t *m, size_t n, uint8_t v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        m[i] = v;
}

__global__ void k_coverage(LevelView f, int cdx, int cdy, uint8_t *mask)
{
    const BoxD B = f.box[blockIdx.y];
    const int cx = B.n[0] / 2, cy = B.n[1] / 2, cz = B.n[2] / 2;
    const uint32_t cells = (uint32_t)cx * cy * cz;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += gridDim.x * blockDim.x)
    {
        const int x = B.lo[0] / 2 + t % cx, y = B.lo[1] / 2 + (t / cx) % cy,
                  z = B.lo[2] / 2 + t / ((uint32_t)cx * cy);
        mask[((size_t)z * cdy + y) * cdx + x] = 0;
    }
}

template <typename T>
__global__ void k_average_down(LevelView f, const T *fy, const float *flo,
                               LevelView c, T *cy_, float *clo,
                               unsigned long long *err)
{
    const BoxD B = f.box[blockIdx.y];
    const int nx = B.n[0] / 2, ny = B.n[1] / 2, nz = B.n[2] / 2;
    const uint32_t cells = (uint32_t)nx * ny * nz;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += gridDim.x * blockDim.x)
    {
        const int ix = t % nx, iy = (t / nx) % ny, iz = t / ((uint32_t)nx * ny);
        uint32_t ci;
        if (locate(c, B.lo[0] / 2 + ix, B.lo[1] / 2 + iy, B.lo[2] / 2 + iz, ci) <
            0)
        {
            atomicAdd(err, 1ull);
            continue;
        }
        for (int comp = 0; comp < 4; ++comp)
        {
            double s = 0.0;
            for (int k = 0; k < 2; ++k)
                for (int j = 0; j < 2; ++j)
                    for (int i = 0; i < 2; ++i)
                    {
                        const size_t o = (size_t)comp * f.total +
                                           sidx(B, f.ng, 2 * ix + i, 2 * iy + j,
                                                2 * iz + k);
                        s += (double)fy[o] + (flo ? (double)flo[o] : 0.0);
                    }
            const double v = 0.125 * s;
            const T h = (T)v;
            cy_[(size_t)comp * c.total + ci] = h;
            if (clo)
                clo[(size_t)comp * c.total + ci] = (float)(v - (double)h);
        }
    }
}

template <typename T>
__global__ void k_copy_level(LevelView v, T *by, float *blo, T *uy, float *ulo,
                             uint32_t un, int to_level)
{
    const BoxD B = v.box[blockIdx.y];
    const uint32_t cells = (uint32_t)B.n[0] * B.n[1] * B.n[2];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += gridDim.x * blockDim.x)
    {
        const int lx = t % B.n[0], ly = (t / B.n[0]) % B.n[1],
                  lz = t / ((uint32_t)B.n[0] * B.n[1]);
        const uint32_t u = ((uint32_t)(B.lo[2] + lz) * v.dom[1] +
                            (B.lo[1] + ly)) *
                               v.dom[0] +
                           (B.lo[0] + lx);
        const uint32_t b = sidx(B, v.ng, lx, ly, lz);
        for (int comp = 0; comp < 4; ++comp)
        {
            if (to_level)
            {
                uy[(size_t)comp * un + u] = by[(size_t)comp * v.total + b];
                if (ulo && blo)
                    ulo[(size_t)comp * un + u] = blo[(size_t)comp * v.total + b];
            }
            else
            {
                by[(size_t)comp * v.total + b] = uy[(size_t)comp * un + u];
                if (ulo && blo)
                    blo[(size_t)comp * v.total + b] = ulo[(size_t)comp * un + u];
            }
        }
    }
}

// Valid cells of every destination box that some source box also owns
// (regridding: the new level keeps the old level's data where they
// overlap).
template <typename T>
__global__ void k_copy_boxes(LevelView dv, T *dy, float *dlo, LevelView sv,
                             const T *sy, const float *slo,
                             unsigned long long *copied)
{
    const BoxD B = dv.box[blockIdx.y];
    const uint32_t cells = (uint32_t)B.n[0] * B.n[1] * B.n[2];
    unsigned long long cnt = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += gridDim.x * blockDim.x)
    {
        const int lx = t % B.n[0], ly = (t / B.n[0]) % B.n[1],
                  lz = t / ((uint32_t)B.n[0] * B.n[1]);
        uint32_t si;
        if (locate(sv, B.lo[0] + lx, B.lo[1] + ly, B.lo[2] + lz, si) < 0)
            continue;
        const uint32_t di = sidx(B, dv.ng, lx, ly, lz);
        for (int comp = 0; comp < 4; ++comp)
        {
            dy[(size_t)comp * dv.total + di] = sy[(size_t)comp * sv.total + si];
            if (dlo && slo)
                dlo[(size_t)comp * dv.total + di] =
                    slo[(size_t)comp * sv.total + si];
        }
        ++cnt;
    }
    if (cnt)
        atomicAdd(copied, cnt);
}

__global__ void k_split(const double *src, float *hi, float *lo, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
    {
        const double v = src[i];
        const float h = (float)v;
        hi[i] = h;
        if (lo)
            lo[i] = (float)(v - (double)h);
    }
}

__global__ void k_join(const float *hi, const float *lo, double *dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (double)hi[i] + (lo ? (double)lo[i] : 0.0);
}

__global__ void k_zero_err(unsigned long long *e)
{
    if (threadIdx.x < 4)
        e[threadIdx.x] = 0ull;
}

static int gblk(size_t n, int cap = 4096)
{
    size_t b = (n + 255) / 256;
    return (int)std::max<size_t>(1, std::min<size_t>(b, cap));
}

// per-box 2D grid: x = chunks of the box's cells, y = boxes
static dim3 box_grid(const axs_boxlevel *b, size_t max_cells)
{
    const size_t chunks = std::min<size_t>((max_cells + 255) / 256, 64);
    return dim3((unsigned)std::max<size_t>(1, chunks), (unsigned)b->nb);
}

// ghost cells of the widest shell of width w over the boxes
static size_t max_shell(const axs_boxlevel *b, int w)
{
    size_t m = 0;
    for (const BoxD &B : b->hbox)
        m = std::max(m, (size_t)(B.n[0] + 2 * w) * (B.n[1] + 2 * w) *
                                (B.n[2] + 2 * w) -
                            (size_t)B.n[0] * B.n[1] * B.n[2]);
    return m;
}

static size_t max_valid(const axs_boxlevel *b)
{
    size_t m = 0;
    for (const BoxD &B : b->hbox)
        m = std::max(m, (size_t)B.n[0] * B.n[1] * B.n[2]);
    return m;
}

static int check_err(axs_boxlevel *bl, const char *what)
{
    unsigned long long e = 0;
    AXS_HIP_TRY(hipMemcpyAsync(&e, bl->derr, sizeof(e), hipMemcpyDeviceToHost,
                               bl->ctx->stream));
    AXS_HIP_TRY(hipStreamSynchronize(bl->ctx->stream));
    if (e)
        return set_error(AXS_ERR_ARG,
                         std::string(what) + ": " + std::to_string(e) +
                             " cells had no owning box (proper nesting / "
                             "coarse coverage violated)");
    return AXS_OK;
}

// psi_only: the stage fills (components 0, 1 in a shell of 2, no lo)
template <typename T>
static int fill_t(axs_boxlevel *bl, void *buf, float *lo,
                  const axs_boxlevel *coarse, double t_frac, int interior,
                  bool psi_only)
{
    CoarseSrc<T> c{};
    if (coarse)
    {
        c.v = view_of(coarse);
        c.cur = (const T *)coarse->y;
        c.cur_lo = coarse->comp;
        if (bl->snap && bl->snap_uid == coarse->uid)
        {
            c.snap = (const T *)bl->snap;
            c.snap_lo = bl->snap_lo;
        }
        c.t_frac = t_frac;
    }
    if (interior)
    {
        // valid cells from the coarse level, then the ghosts as usual
        int tiles = 1;
        for (const BoxD &B : bl->hbox)
            tiles = std::max(tiles, ((B.n[0] / 2 + kProTX - 1) / kProTX) *
                                        ((B.n[1] / 2 + kProTY - 1) / kProTY) *
                                        ((B.n[2] / 2 + kProTZ - 1) / kProTZ));
        hipLaunchKernelGGL(k_prolong<T>, dim3((unsigned)tiles, (unsigned)bl->nb),
                           dim3(256), 0, bl->ctx->stream, view_of(bl),
                           (T *)buf, lo, c, bl->derr);
        AXS_HIP_TRY(hipGetLastError());
    }
    if (psi_only)
        hipLaunchKernelGGL((k_fill<T, 2>), box_grid(bl, max_shell(bl, 2)),
                           dim3(256), 0, bl->ctx->stream, view_of(bl),
                           (T *)buf, (float *)nullptr, coarse ? 1 : 0, c, 2,
                           bl->derr);
    else
        hipLaunchKernelGGL((k_fill<T, 4>), box_grid(bl, max_shell(bl, bl->ng)),
                           dim3(256), 0, bl->ctx->stream, view_of(bl),
                           (T *)buf, lo, coarse ? 1 : 0, c, bl->ng, bl->derr);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

static int fill(axs_boxlevel *bl, void *buf, float *lo,
                const axs_boxlevel *coarse, double t_frac, int interior,
                bool psi_only = false)
{
    if (bl->precision == AXS_FP64)
        return fill_t<double>(bl, buf, lo, coarse, t_frac, interior, psi_only);
    return fill_t<float>(bl, buf, lo, coarse, t_frac, interior, psi_only);
}

template <typename T, int PREC>
static int stage_t(axs_boxlevel *bl, int stage, const void *yin, void *yout,
                   double dt)
{
    BoxStageParams<T> p{};
    p.box = bl->dbox;
    p.items = bl->ditems;
    p.nitems = bl->nitems;
    p.ng = bl->ng;
    p.total = bl->total;
    p.yin = (const T *)yin;
    p.yn = (const T *)bl->y;
    p.acc = (T *)bl->acc;
    p.yout = (T *)yout;
    p.comp = bl->comp;
    p.coeff = bl->dcoeff;
    p.coeff_stage = stage - 1;
    p.a = stage == 3 ? dt : 0.5 * dt;
    p.h6 = dt / 6.0;
    p.h6_lo = (float)(p.h6 - (double)(float)p.h6);
    p.inv_dx2 = 1.0 / (bl->dx * bl->dx);
    const int grid = std::min(bl->nitems, bl->grid_blocks);
    hipStream_t s = bl->ctx->stream;
    switch (stage)
    {
    case 1:
        hipLaunchKernelGGL((k_box_stage<T, PREC, MODE_S1>), dim3(grid),
                           dim3(kBoxLanes), 0, s, p);
        break;
    case 4:
        hipLaunchKernelGGL((k_box_stage<T, PREC, MODE_S4>), dim3(grid),
                           dim3(kBoxLanes), 0, s, p);
        break;
    default:
        hipLaunchKernelGGL((k_box_stage<T, PREC, MODE_S23>), dim3(grid),
                           dim3(kBoxLanes), 0, s, p);
    }
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

static int stage(axs_boxlevel *bl, int st, const void *yin, void *yout,
                 double dt)
{
    if (bl->precision == AXS_FP64)
        return stage_t<double, AXS_FP64>(bl, st, yin, yout, dt);
    if (bl->precision == AXS_FP32C)
        return stage_t<float, AXS_FP32C>(bl, st, yin, yout, dt);
    return stage_t<float, AXS_FP32>(bl, st, yin, yout, dt);
}

// both fills of a pair in one launch: srcA's psi into dstA's shell at tA,
// stage_axpy(psi of ps, a, Pi of pp) into dstB's shell at tB
// (eA, eps, epp: the x-edge arrays of srcA, ps and pp)
template <typename T, int PREC>
static int fill_pair_t(axs_boxlevel *bl, void *dstA, const void *srcA,
                       void *dstB, const void *ps, const void *pp, double a,
                       const axs_boxlevel *coarse, double tA, double tB,
                       const void *eA, const void *eps, const void *epp)
{
    CoarseSrc<T> c{};
    if (coarse)
    {
        c.v = view_of(coarse);
        c.cur = (const T *)coarse->y;
        c.cur_lo = coarse->comp;
        if (bl->snap && bl->snap_uid == coarse->uid)
        {
            c.snap = (const T *)bl->snap;
            c.snap_lo = bl->snap_lo;
        }
    }
    hipLaunchKernelGGL((k_fill_pair<T, PREC>), box_grid(bl, max_shell(bl, 2)),
                       dim3(256), 0, bl->ctx->stream, view_of(bl), (T *)dstA,
                       (const T *)srcA, (T *)dstB, (const T *)ps,
                       (const T *)pp, a, coarse ? 1 : 0, c, tA, tB,
                       (const T *)eA, (const T *)eps, (const T *)epp,
                       4 * bl->edge_rows, bl->derr);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

static int fill_pair(axs_boxlevel *bl, void *dstA, const void *srcA,
                     void *dstB, const void *ps, const void *pp, double a,
                     const axs_boxlevel *coarse, double tA, double tB,
                     const void *eA, const void *eps, const void *epp)
{
    if (bl->precision == AXS_FP64)
        return fill_pair_t<double, AXS_FP64>(bl, dstA, srcA, dstB, ps, pp, a,
                                             coarse, tA, tB, eA, eps, epp);
    if (bl->precision == AXS_FP32C)
        return fill_pair_t<float, AXS_FP32C>(bl, dstA, srcA, dstB, ps, pp, a,
                                             coarse, tA, tB, eA, eps, epp);
    return fill_pair_t<float, AXS_FP32>(bl, dstA, srcA, dstB, ps, pp, a,
                                        coarse, tA, tB, eA, eps, epp);
}

// The x-edge arrays of the state, from y, unless they already hold the
// state's (the last fused-pair step wrote them and nothing has written the
// state since).  Run outside any capture, before a fused-pair step.
template <typename T>
static int ensure_edges_t(axs_boxlevel *bl)
{
    if (bl->edge_gen == bl->ygen && bl->edge_state == bl->y)
        return AXS_OK;
    size_t rows = 0;
    for (const BoxD &B : bl->hbox)
        rows = std::max(rows, (size_t)B.n[1] * B.n[2]);
    hipLaunchKernelGGL(k_edges<T>, box_grid(bl, 4 * rows), dim3(256), 0,
                       bl->ctx->stream, bl->dbox, bl->ng, bl->total,
                       (const T *)bl->y, (T *)bl->edge_n, 4 * bl->edge_rows);
    AXS_HIP_TRY(hipGetLastError());
    bl->edge_gen = bl->ygen;
    bl->edge_state = bl->y;
    return AXS_OK;
}

static int ensure_edges(axs_boxlevel *bl)
{
    if (bl->fusion != 1)
        return AXS_OK;
    return bl->precision == AXS_FP64 ? ensure_edges_t<double>(bl)
                                     : ensure_edges_t<float>(bl);
}

template <typename T, int PREC>
static int pair_t(axs_boxlevel *bl, int pair, double dt)
{
    BoxPairParams<T> p{};
    p.box = bl->dbox;
    p.items = bl->ditems;
    p.nitems = bl->nitems;
    p.ng = bl->ng;
    p.total = bl->total;
    p.yn = (const T *)bl->y;
    p.y2 = (const T *)bl->ya;
    p.y2out = (T *)bl->ya;
    p.gb = (const T *)bl->yb;
    p.acc = (T *)bl->acc;
    p.comp = bl->comp;
    p.coeff = bl->dcoeff;
    p.dt = dt;
    p.h6 = dt / 6.0;
    p.h6_lo = (float)(p.h6 - (double)(float)p.h6);
    p.inv_dx2 = 1.0 / (bl->dx * bl->dx);
    p.edge = (T *)(pair == 1 ? bl->edge_2 : bl->edge_n);
    p.es = 4 * bl->edge_rows;
    const int grid = std::min(bl->nitems, bl->grid_blocks);
    hipStream_t s = bl->ctx->stream;
    if (pair == 1)
        hipLaunchKernelGGL((k_box_pair<T, PREC, 1>), dim3(grid), dim3(kBoxLanes),
                           0, s, p);
    else
        hipLaunchKernelGGL((k_box_pair<T, PREC, 2>), dim3(grid), dim3(kBoxLanes),
                           0, s, p);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

static int pair(axs_boxlevel *bl, int pr, double dt)
{
    if (bl->precision == AXS_FP64)
        return pair_t<double, AXS_FP64>(bl, pr, dt);
    if (bl->precision == AXS_FP32C)
        return pair_t<float, AXS_FP32C>(bl, pr, dt);
    return pair_t<float, AXS_FP32>(bl, pr, dt);
}

// The launches of one RK4 step on the level's stream (direct or under
// capture).  Stage times of the coarse interval: t0, t0 + df/2, t0 + df/2,
// t0 + df.  One launch per stage: the ghost fill of each stage input, then
// the stage (inputs y, ya, yb, ya).  Fused pairs: one fill of psi_n ghosts
// of y and psi_1 ghosts into yb's shell, stages 1+2 (acc, y_2 into ya); one
// fill of psi_2 ghosts of ya and psi_3 ghosts into yb's shell, stages 3+4
// (y_{n+1} into acc, which the caller then makes the state).  The fills'
// fine-fine x-face sources come from the x-edge arrays: the state's
// (edge_n, made current by ensure_edges before the step) and y_2's
// (edge_2, written by the first pair); the second pair writes y_{n+1}'s
// into edge_n.
static int step_launches(axs_boxlevel *b, const axs_boxlevel *coarse,
                         double dt, double t0, double df)
{
    const double tf[4] = {t0, t0 + 0.5 * df, t0 + 0.5 * df, t0 + df};
    int rc = AXS_OK;
    if (b->fusion == 1)
    {
        rc = fill_pair(b, b->y, b->y, b->yb, b->y, b->y, 0.5 * dt, coarse,
                       tf[0], tf[1], b->edge_n, b->edge_n, b->edge_n);
        if (!rc)
            rc = pair(b, 1, dt);
        if (!rc)
            rc = fill_pair(b, b->ya, b->ya, b->yb, b->y, b->ya, dt, coarse,
                           tf[2], tf[3], b->edge_2, b->edge_n, b->edge_2);
        if (!rc)
            rc = pair(b, 2, dt);
        return rc;
    }
    void *in[4] = {b->y, b->ya, b->yb, b->ya};
    void *out[4] = {b->ya, b->yb, b->ya, b->y};
    for (int st = 0; st < 4 && rc == AXS_OK; ++st)
    {
        rc = fill(b, in[st], nullptr, coarse, tf[st], 0, true);
        if (rc == AXS_OK)
            rc = stage(b, st + 1, in[st], out[st], dt);
    }
    return rc;
}

// after a fused-pair step the state is in the array behind acc, and
// edge_n holds its x-edge columns
static void step_done(axs_boxlevel *b)
{
    ++b->ygen;
    if (b->fusion == 1)
    {
        std::swap(b->y, b->acc);
        b->edge_gen = b->ygen;
        b->edge_state = b->y;
    }
}

} // namespace axs

using namespace axs;

extern "C" {

int axs_boxlevel_create(axs_ctx *ctx, int nboxes, const int *lo, const int *hi,
                        const int domain[3], int nghost, double dx,
                        int precision, axs_boxlevel **out)
{
    AXS_CHECK_ARG(ctx && lo && hi && domain && out,
                  "axs_boxlevel_create: NULL argument");
    *out = nullptr;
    AXS_CHECK_ARG(nboxes >= 1 && nboxes <= 65535,
                  "axs_boxlevel_create: 1 <= nboxes <= 65535");
    AXS_CHECK_ARG(nghost == 3, "axs_boxlevel_create: nghost must be 3");
    AXS_CHECK_ARG(dx > 0.0, "axs_boxlevel_create: dx must be > 0");
    AXS_CHECK_ARG(precision == AXS_FP64 || precision == AXS_FP32C ||
                      precision == AXS_FP32,
                  "axs_boxlevel_create: unknown precision");
    for (int d = 0; d < 3; ++d)
        AXS_CHECK_ARG(domain[d] >= kBlock && domain[d] % kBlock == 0,
                      "axs_boxlevel_create: domain must be a multiple of 8");
    static std::atomic<unsigned long long> next_uid{1};
    axs_boxlevel *b = new axs_boxlevel{};
    b->uid = next_uid.fetch_add(1);
    b->ctx = ctx;
    b->nb = nboxes;
    b->ng = nghost;
    b->dx = dx;
    b->precision = precision;
    b->word = precision == AXS_FP64 ? 8 : 4;
    for (int d = 0; d < 3; ++d)
    {
        b->dom[d] = domain[d];
        b->bdim[d] = domain[d] / kBlock;
    }
    std::vector<int> owner((size_t)b->bdim[0] * b->bdim[1] * b->bdim[2], -1);
    uint64_t off = 0, erows = 0;
    int64_t nvalid = 0;
    for (int i = 0; i < nboxes; ++i)
    {
        BoxD B{};
        for (int d = 0; d < 3; ++d)
        {
            B.lo[d] = lo[3 * i + d];
            B.n[d] = hi[3 * i + d] - lo[3 * i + d] + 1;
            if (B.lo[d] < 0 || B.lo[d] % kBlock || B.n[d] % kBlock ||
                B.n[d] < kBlock || B.n[d] > 64 || B.lo[d] + B.n[d] > domain[d])
            {
                delete b;
                return set_error(AXS_ERR_ARG,
                                 "axs_boxlevel_create: box " + std::to_string(i) +
                                     " must lie in the domain with corners and "
                                     "edges on multiples of 8, edges 8..64");
            }
        }
        B.off = (uint32_t)off;
        off += (uint64_t)(B.n[0] + 2 * nghost) * (B.n[1] + 2 * nghost) *
               (B.n[2] + 2 * nghost);
        B.erow = (int)erows;
        erows += (uint64_t)B.n[1] * B.n[2];
        nvalid += (int64_t)B.n[0] * B.n[1] * B.n[2];
        for (int z = B.lo[2] / kBlock; z < (B.lo[2] + B.n[2]) / kBlock; ++z)
            for (int y = B.lo[1] / kBlock; y < (B.lo[1] + B.n[1]) / kBlock; ++y)
                for (int x = B.lo[0] / kBlock; x < (B.lo[0] + B.n[0]) / kBlock;
                     ++x)
                {
                    int &o = owner[((size_t)z * b->bdim[1] + y) * b->bdim[0] + x];
                    if (o >= 0)
                    {
                        delete b;
                        return set_error(AXS_ERR_ARG,
                                         "axs_boxlevel_create: boxes overlap");
                    }
                    o = i;
                }
        b->hbox.push_back(B);
    }
    // offsets inside one component array are 32-bit; component strides
    // are applied in 64-bit
    if (off >= (1ull << 32))
    {
        delete b;
        return set_error(AXS_ERR_ARG, "axs_boxlevel_create: level too large");
    }
    b->total = (uint32_t)off;
    b->nvalid = nvalid;
    b->edge_rows = (size_t)erows; // < 2^32 / 64 since off < 2^32
    // work shape: whole-width items of up to 512 lanes and a grid of up to
    // 2048 blocks unless the tuning table has one for this precision
    b->item_rows = kBoxLanes;
    b->grid_blocks = 2048;
    b->fusion = 1; // fused stage pairs
    (void)tuned_box(precision, &b->item_rows, &b->grid_blocks);
    const std::vector<WorkItem> items = make_items(b->hbox, b->item_rows);
    b->nitems = (int)items.size();
    const size_t bytes4 = (size_t)4 * off * b->word;
    // staggered array bases, as for uniform levels (axs_level_create)
    const size_t S = kArrayStagger, pad = 8 * S;
    b->alloc[0] = pool_alloc(ctx, bytes4 + pad, &b->alloc_bytes[0]);
    b->alloc[1] = pool_alloc(ctx, bytes4 + pad, &b->alloc_bytes[1]);
    b->alloc[2] = pool_alloc(ctx, 2 * (bytes4 + pad), &b->alloc_bytes[2]);
    bool ok = b->alloc[0] && b->alloc[1] && b->alloc[2] &&
              hipMalloc(&b->dbox, nboxes * sizeof(BoxD)) == hipSuccess &&
              hipMalloc(&b->downer, owner.size() * sizeof(int)) == hipSuccess &&
              hipMalloc(&b->ditems, items.size() * sizeof(WorkItem)) ==
                  hipSuccess &&
              hipMalloc(&b->dcoeff, 12 * sizeof(double)) == hipSuccess &&
              hipHostMalloc(&b->h_coeff, 12 * sizeof(double)) == hipSuccess &&
              hipMalloc(&b->derr, 4 * sizeof(unsigned long long)) == hipSuccess;
    if (ok && precision == AXS_FP32C)
    {
        b->alloc[3] = pool_alloc(ctx, bytes4 + pad, &b->alloc_bytes[3]);
        ok = b->alloc[3] != nullptr;
    }
    const size_t ebytes = (size_t)16 * b->edge_rows * b->word; // one edge array
    if (ok)
    {
        b->edge_alloc = pool_alloc(ctx, 2 * ebytes, &b->edge_alloc_bytes);
        ok = b->edge_alloc != nullptr;
    }
    if (!ok)
    {
        for (int i = 0; i < 4; ++i)
            pool_release(ctx, b->alloc[i], b->alloc_bytes[i]);
        pool_release(ctx, b->edge_alloc, b->edge_alloc_bytes);
        (void)hipFree(b->dbox);
        (void)hipFree(b->downer);
        (void)hipFree(b->ditems);
        (void)hipFree(b->dcoeff);
        (void)hipHostFree(b->h_coeff);
        (void)hipFree(b->derr);
        delete b;
        return set_error(AXS_ERR_NOMEM, "axs_boxlevel_create: out of memory");
    }
    b->y = b->alloc[0];
    b->acc = (char *)b->alloc[1] + S;
    b->ya = (char *)b->alloc[2] + 2 * S;
    b->yb = (char *)b->alloc[2] + bytes4 + 11 * S;
    if (b->alloc[3])
        b->comp = (float *)((char *)b->alloc[3] + 4 * S);
    b->edge_n = b->edge_alloc;
    b->edge_2 = (char *)b->edge_alloc + ebytes;
    b->ygen = 1; // edge_gen 0: the x-edge arrays hold nothing yet
    b->edge_gen = 0;
    b->edge_state = nullptr;
    hipStream_t s = ctx->stream;
    AXS_HIP_TRY(hipMemcpyAsync(b->dbox, b->hbox.data(), nboxes * sizeof(BoxD),
                               hipMemcpyHostToDevice, s));
    AXS_HIP_TRY(hipMemcpyAsync(b->downer, owner.data(),
                               owner.size() * sizeof(int), hipMemcpyHostToDevice,
                               s));
    AXS_HIP_TRY(hipMemcpyAsync(b->ditems, items.data(),
                               items.size() * sizeof(WorkItem),
                               hipMemcpyHostToDevice, s));
    AXS_HIP_TRY(hipMemsetAsync(b->y, 0, bytes4, s));
    if (b->comp)
        AXS_HIP_TRY(hipMemsetAsync(b->comp, 0, bytes4, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    *out = b;
    return AXS_OK;
}

// does view a hold cells lo - glo .. hi + ghi with at least nc components?
static bool view_holds(const axs_array4 *a, const int lo[3], const int hi[3],
                       int glo, int ghi, int nc)
{
    if (!a || !a->p || a->ncomp < nc)
        return false;
    for (int d = 0; d < 3; ++d)
        if (lo[d] > hi[d] || lo[d] - glo < a->begin[d] ||
            hi[d] + ghi >= a->end[d])
            return false;
    return true;
}

int axs_rhs_array4(const axs_array4 *state, axs_array4 *rhs, const int lo[3],
                   const int hi[3], const axs_rhs_coeffs *c, double dx,
                   int precision, void *stream)
{
    AXS_CHECK_ARG(state && rhs && lo && hi && c,
                  "axs_rhs_array4: NULL argument");
    AXS_CHECK_ARG(precision == AXS_FP64 || precision == AXS_FP32,
                  "axs_rhs_array4: precision must be AXS_FP64 (double views) "
                  "or AXS_FP32 (float views)");
    AXS_CHECK_ARG(dx > 0.0, "axs_rhs_array4: dx must be > 0");
    AXS_CHECK_ARG(view_holds(state, lo, hi, 2, 2, 4),
                  "axs_rhs_array4: the state view must hold lo-2 .. hi+2 with "
                  "4 components");
    AXS_CHECK_ARG(view_holds(rhs, lo, hi, 0, 0, 4),
                  "axs_rhs_array4: the rhs view must hold lo .. hi with 4 "
                  "components");
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1,
              nz = hi[2] - lo[2] + 1;
    const long cols = (long)((nx + 63) / 64) * ((ny + 3) / 4);
    int zc = 64;
    while (zc > 4 && cols * ((nz + zc - 1) / zc) < 4096)
        zc >>= 1;
    const dim3 grid((nx + 63) / 64, (ny + 3) / 4, (nz + zc - 1) / zc);
    hipStream_t s = (hipStream_t)stream;
    const double i2 = 1.0 / (dx * dx);
    if (precision == AXS_FP64)
    {
        const View4<double> in = view4<double>(*state);
        hipLaunchKernelGGL(k_rhs_array4<double>, grid, dim3(64, 4), 0, s,
                           View4<const double>{in.p, in.js, in.ks, in.ns, in.b0,
                                               in.b1, in.b2},
                           view4<double>(*rhs), lo[0], lo[1], lo[2], nx, ny, nz,
                           zc, c->curvature_coeff, c->lambda, c->R_squared, i2);
    }
    else
    {
        const View4<float> in = view4<float>(*state);
        hipLaunchKernelGGL(k_rhs_array4<float>, grid, dim3(64, 4), 0, s,
                           View4<const float>{in.p, in.js, in.ks, in.ns, in.b0,
                                              in.b1, in.b2},
                           view4<float>(*rhs), lo[0], lo[1], lo[2], nx, ny, nz,
                           zc, c->curvature_coeff, c->lambda, c->R_squared, i2);
    }
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

int axs_tag_array4(const axs_array4 *state, const axs_array4 *state_lo,
                   const int lo[3], const int hi[3],
                   const axs_tagger_params *p, double dx, int precision,
                   axs_array4 *tags, unsigned long long *counts_dev,
                   void *stream)
{
    AXS_CHECK_ARG(state && lo && hi && counts_dev,
                  "axs_tag_array4: NULL argument");
    AXS_CHECK_ARG(precision == AXS_FP64 || precision == AXS_FP32,
                  "axs_tag_array4: precision must be AXS_FP64 (double views) "
                  "or AXS_FP32 (float views, FP32C with state_lo)");
    AXS_CHECK_ARG(dx > 0.0, "axs_tag_array4: dx must be > 0");
    AXS_CHECK_ARG(!state_lo || precision == AXS_FP32,
                  "axs_tag_array4: state_lo goes with float views");
    const bool deriv = tags && p &&
                       (p->gradient_threshold < DBL_MAX ||
                        p->radial_gradient_threshold < DBL_MAX);
    // windings read the cell and its +1 neighbours; the derivative criteria
    // a radius-2 star
    AXS_CHECK_ARG(view_holds(state, lo, hi, deriv ? 2 : 0, deriv ? 2 : 1, 2),
                  "axs_tag_array4: the state view must hold lo .. hi+1 (lo-2 "
                  ".. hi+2 with the derivative criteria) with 2 components");
    AXS_CHECK_ARG(!state_lo || (state_lo->p && state_lo->jstride == state->jstride &&
                                state_lo->kstride == state->kstride &&
                                state_lo->nstride == state->nstride &&
                                state_lo->begin[0] == state->begin[0] &&
                                state_lo->begin[1] == state->begin[1] &&
                                state_lo->begin[2] == state->begin[2]),
                  "axs_tag_array4: state_lo must have the state view's shape");
    AXS_CHECK_ARG(!tags || view_holds(tags, lo, hi, 0, 0, 1),
                  "axs_tag_array4: the tags view must hold lo .. hi");
    if (p && tags)
    {
        AXS_CHECK_ARG(p->gradient_threshold > 0.0 &&
                          p->radial_gradient_threshold > 0.0,
                      "axs_tag_array4: thresholds must be > 0");
        AXS_CHECK_ARG(p->amplitude_threshold <= 0.0 || p->R > 0.0,
                      "axs_tag_array4: amplitude criterion needs R > 0");
    }
    BoxTagParams tp{};
    tp.amp_thr = p ? p->amplitude_threshold : 0.0;
    tp.R = p ? p->R : 1.0;
    tp.grad_thr = p ? p->gradient_threshold : DBL_MAX;
    tp.rgrad_thr = p ? p->radial_gradient_threshold : DBL_MAX;
    tp.dx = dx;
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1,
              nz = hi[2] - lo[2] + 1;
    const size_t cells = (size_t)nx * ny * nz;
    AXS_CHECK_ARG(cells < (size_t(1) << 31), "axs_tag_array4: box too large");
    const unsigned blocks = (unsigned)std::min<size_t>((cells + 255) / 256, 4096);
    hipStream_t s = (hipStream_t)stream;
    const View4<uint8_t> tv = tags ? view4<uint8_t>(*tags) : View4<uint8_t>{};
    const View4<float> lv = state_lo ? view4<float>(*state_lo) : View4<float>{};
    const View4<const float> lc{lv.p, lv.js, lv.ks, lv.ns, lv.b0, lv.b1, lv.b2};
    if (precision == AXS_FP64)
    {
        const View4<double> in = view4<double>(*state);
        hipLaunchKernelGGL(k_tag_array4<double>, dim3(blocks), dim3(256), 0, s,
                           View4<const double>{in.p, in.js, in.ks, in.ns, in.b0,
                                               in.b1, in.b2},
                           lc, 0, tv, tags ? 1 : 0, lo[0], lo[1], lo[2], nx, ny,
                           nz, tp, counts_dev);
    }
    else
    {
        const View4<float> in = view4<float>(*state);
        hipLaunchKernelGGL(k_tag_array4<float>, dim3(blocks), dim3(256), 0, s,
                           View4<const float>{in.p, in.js, in.ks, in.ns, in.b0,
                                              in.b1, in.b2},
                           lc, state_lo ? 1 : 0, tv, tags ? 1 : 0, lo[0], lo[1],
                           lo[2], nx, ny, nz, tp, counts_dev);
    }
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

int axs_boxlevel_destroy(axs_boxlevel *b)
{
    if (!b)
        return AXS_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    for (auto &e : b->graphs)
    {
        (void)hipGraphExecDestroy(e.ge);
        (void)hipGraphDestroy(e.g);
    }
    for (int i = 0; i < 4; ++i)
        pool_release(b->ctx, b->alloc[i], b->alloc_bytes[i]);
    pool_release(b->ctx, b->edge_alloc, b->edge_alloc_bytes);
    (void)hipFree(b->dbox);
    (void)hipFree(b->downer);
    (void)hipFree(b->ditems);
    (void)hipFree(b->dcoeff);
    (void)hipHostFree(b->h_coeff);
    (void)hipFree(b->derr);
    pool_release(b->ctx, b->snap, b->snap_cap);
    delete b;
    return AXS_OK;
}

int64_t axs_boxlevel_num_cells(axs_boxlevel *b) { return b ? b->nvalid : 0; }

int axs_boxlevel_set_work_shape(axs_boxlevel *b, int item_rows, int grid_blocks)
{
    AXS_CHECK_ARG(b, "axs_boxlevel_set_work_shape: NULL level");
    AXS_CHECK_ARG(item_rows >= 1 && item_rows <= kBoxLanes,
                  "axs_boxlevel_set_work_shape: 1 <= item_rows <= 512");
    AXS_CHECK_ARG(grid_blocks >= 1 && grid_blocks <= 65536,
                  "axs_boxlevel_set_work_shape: 1 <= grid_blocks <= 65536");
    if (item_rows == b->item_rows && grid_blocks == b->grid_blocks)
        return AXS_OK;
    hipStream_t s = b->ctx->stream;
    AXS_HIP_TRY(hipStreamSynchronize(s));
    // captured step graphs hold the old item list and grid
    for (auto &e : b->graphs)
    {
        (void)hipGraphExecDestroy(e.ge);
        (void)hipGraphDestroy(e.g);
    }
    b->graphs.clear();
    const std::vector<WorkItem> items = make_items(b->hbox, item_rows);
    if ((int)items.size() != b->nitems)
    {
        WorkItem *d = nullptr;
        if (hipMalloc(&d, items.size() * sizeof(WorkItem)) != hipSuccess)
            return set_error(AXS_ERR_NOMEM,
                             "axs_boxlevel_set_work_shape: out of memory");
        (void)hipFree(b->ditems);
        b->ditems = d;
    }
    AXS_HIP_TRY(hipMemcpyAsync(b->ditems, items.data(),
                               items.size() * sizeof(WorkItem),
                               hipMemcpyHostToDevice, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    b->nitems = (int)items.size();
    b->item_rows = item_rows;
    b->grid_blocks = grid_blocks;
    return AXS_OK;
}

int axs_boxlevel_set_stage_fusion(axs_boxlevel *b, int mode)
{
    AXS_CHECK_ARG(b, "axs_boxlevel_set_stage_fusion: NULL level");
    AXS_CHECK_ARG(mode == 0 || mode == 1,
                  "axs_boxlevel_set_stage_fusion: mode 0 (one launch per "
                  "stage) or 1 (fused stage pairs)");
    b->fusion = mode;
    return AXS_OK;
}

int axs_boxlevel_get_stage_fusion(const axs_boxlevel *b, int *mode)
{
    AXS_CHECK_ARG(b && mode, "axs_boxlevel_get_stage_fusion: NULL argument");
    *mode = b->fusion;
    return AXS_OK;
}

int axs_boxlevel_get_work_shape(const axs_boxlevel *b, int *item_rows,
                                int *grid_blocks, int *nitems)
{
    AXS_CHECK_ARG(b, "axs_boxlevel_get_work_shape: NULL level");
    if (item_rows)
        *item_rows = b->item_rows;
    if (grid_blocks)
        *grid_blocks = b->grid_blocks;
    if (nitems)
        *nitems = b->nitems;
    return AXS_OK;
}

int axs_boxlevel_upload_dense(axs_boxlevel *b, const double *dense)
{
    AXS_CHECK_ARG(b && dense, "axs_boxlevel_upload_dense: NULL argument");
    ++b->ygen; // writes the state
    const size_t n4 = (size_t)4 * b->total;
    const size_t nd = (size_t)b->dom[0] * b->dom[1] * b->dom[2];
    std::vector<double> st(n4, 0.0);
    for (const BoxD &B : b->hbox)
    {
        const int gx = B.n[0] + 2 * b->ng, gy = B.n[1] + 2 * b->ng;
        for (int c = 0; c < 4; ++c)
            for (int z = 0; z < B.n[2]; ++z)
                for (int y = 0; y < B.n[1]; ++y)
                    for (int x = 0; x < B.n[0]; ++x)
                        st[(size_t)c * b->total + B.off +
                           ((size_t)(z + b->ng) * gy + (y + b->ng)) * gx +
                           (x + b->ng)] =
                            dense[c * nd +
                                  ((size_t)(B.lo[2] + z) * b->dom[1] +
                                   (B.lo[1] + y)) *
                                      b->dom[0] +
                                  (B.lo[0] + x)];
    }
    hipStream_t s = b->ctx->stream;
    if (b->precision == AXS_FP64)
        AXS_HIP_TRY(hipMemcpyAsync(b->y, st.data(), n4 * 8,
                                   hipMemcpyHostToDevice, s));
    else
    {
        AXS_HIP_TRY(hipMemcpyAsync(b->ya, st.data(), n4 * 8,
                                   hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_split, dim3(gblk(n4)), dim3(256), 0, s,
                           (const double *)b->ya, (float *)b->y, b->comp, n4);
        AXS_HIP_TRY(hipGetLastError());
    }
    AXS_HIP_TRY(hipStreamSynchronize(s));
    return AXS_OK;
}

int axs_boxlevel_download_dense(axs_boxlevel *b, double *dense)
{
    AXS_CHECK_ARG(b && dense, "axs_boxlevel_download_dense: NULL argument");
    const size_t n4 = (size_t)4 * b->total;
    const size_t nd = (size_t)b->dom[0] * b->dom[1] * b->dom[2];
    std::vector<double> st(n4);
    hipStream_t s = b->ctx->stream;
    if (b->precision == AXS_FP64)
        AXS_HIP_TRY(hipMemcpyAsync(st.data(), b->y, n4 * 8,
                                   hipMemcpyDeviceToHost, s));
    else
    {
        hipLaunchKernelGGL(k_join, dim3(gblk(n4)), dim3(256), 0, s,
                           (const float *)b->y, (const float *)b->comp,
                           (double *)b->ya, n4);
        AXS_HIP_TRY(hipGetLastError());
        AXS_HIP_TRY(hipMemcpyAsync(st.data(), b->ya, n4 * 8,
                                   hipMemcpyDeviceToHost, s));
    }
    AXS_HIP_TRY(hipStreamSynchronize(s));
    for (const BoxD &B : b->hbox)
    {
        const int gx = B.n[0] + 2 * b->ng, gy = B.n[1] + 2 * b->ng;
        for (int c = 0; c < 4; ++c)
            for (int z = 0; z < B.n[2]; ++z)
                for (int y = 0; y < B.n[1]; ++y)
                    for (int x = 0; x < B.n[0]; ++x)
                        dense[c * nd +
                              ((size_t)(B.lo[2] + z) * b->dom[1] + (B.lo[1] + y)) *
                                  b->dom[0] +
                              (B.lo[0] + x)] =
                            st[(size_t)c * b->total + B.off +
                               ((size_t)(z + b->ng) * gy + (y + b->ng)) * gx +
                               (x + b->ng)];
    }
    return AXS_OK;
}

static int copy_level(const axs_boxlevel *b, axs_level *L, int to_level)
{
    AXS_CHECK_ARG(b && L, "axs_boxlevel_copy: NULL argument");
    AXS_CHECK_ARG(L->g.nx == b->dom[0] && L->g.ny == b->dom[1] &&
                      L->g.nz == b->dom[2],
                  "axs_boxlevel_copy: domains differ");
    AXS_CHECK_ARG(L->precision == b->precision,
                  "axs_boxlevel_copy: precisions differ");
    AXS_CHECK_ARG(L->ctx == b->ctx, "axs_boxlevel_copy: contexts differ");
    hipStream_t s = b->ctx->stream;
    const dim3 g = box_grid(b, max_valid(b));
    if (b->precision == AXS_FP64)
        hipLaunchKernelGGL(k_copy_level<double>, g, dim3(256), 0, s, view_of(b),
                           (double *)b->y, (float *)nullptr, (double *)L->y,
                           (float *)nullptr, L->g.ncell, to_level);
    else
        hipLaunchKernelGGL(k_copy_level<float>, g, dim3(256), 0, s, view_of(b),
                           (float *)b->y, b->comp, (float *)L->y, L->comp,
                           L->g.ncell, to_level);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

int axs_boxlevel_copy_from_level(axs_boxlevel *b, const axs_level *src)
{
    if (b)
        ++b->ygen; // writes the state
    return copy_level(b, const_cast<axs_level *>(src), 0);
}

int axs_boxlevel_copy_to_level(const axs_boxlevel *b, axs_level *dst)
{
    return copy_level(b, dst, 1);
}

int axs_boxlevel_copy_valid(axs_boxlevel *dst, const axs_boxlevel *src,
                            int64_t *copied)
{
    AXS_CHECK_ARG(dst && src, "axs_boxlevel_copy_valid: NULL argument");
    AXS_CHECK_ARG(dst != src, "axs_boxlevel_copy_valid: same level");
    ++dst->ygen; // writes the state
    for (int d = 0; d < 3; ++d)
        AXS_CHECK_ARG(dst->dom[d] == src->dom[d],
                      "axs_boxlevel_copy_valid: domains differ");
    AXS_CHECK_ARG(dst->precision == src->precision,
                  "axs_boxlevel_copy_valid: precisions differ");
    AXS_CHECK_ARG(dst->ctx == src->ctx,
                  "axs_boxlevel_copy_valid: contexts differ");
    hipStream_t s = dst->ctx->stream;
    unsigned long long *cnt = dst->ctx->d_counts;
    hipLaunchKernelGGL(k_zero_err, dim3(1), dim3(64), 0, s, cnt);
    const dim3 g = box_grid(dst, max_valid(dst));
    if (dst->precision == AXS_FP64)
        hipLaunchKernelGGL(k_copy_boxes<double>, g, dim3(256), 0, s,
                           view_of(dst), (double *)dst->y, (float *)nullptr,
                           view_of(src), (const double *)src->y,
                           (const float *)nullptr, cnt);
    else
        hipLaunchKernelGGL(k_copy_boxes<float>, g, dim3(256), 0, s,
                           view_of(dst), (float *)dst->y, dst->comp,
                           view_of(src), (const float *)src->y,
                           (const float *)src->comp, cnt);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipMemcpyAsync(dst->ctx->h_counts, cnt,
                               sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    if (copied)
        *copied = (int64_t)dst->ctx->h_counts[0];
    return AXS_OK;
}

int axs_boxlevel_snapshot_coarse(axs_boxlevel *b, const axs_boxlevel *c)
{
    AXS_CHECK_ARG(b && c, "axs_boxlevel_snapshot_coarse: NULL argument");
    AXS_CHECK_ARG(c->precision == b->precision,
                  "axs_boxlevel_snapshot_coarse: precisions differ");
    const size_t sb = (size_t)4 * c->total * c->word;
    const size_t lb = c->comp ? (size_t)4 * c->total * 4 : 0;
    if (b->snap_bytes != sb + lb)
    {
        // captured graphs hold the old snapshot address
        AXS_HIP_TRY(hipStreamSynchronize(b->ctx->stream));
        for (auto &e : b->graphs)
        {
            (void)hipGraphExecDestroy(e.ge);
            (void)hipGraphDestroy(e.g);
        }
        b->graphs.clear();
        pool_release(b->ctx, b->snap, b->snap_cap);
        b->snap = pool_alloc(b->ctx, sb + lb, &b->snap_cap);
        if (!b->snap)
        {
            b->snap_bytes = b->snap_cap = 0;
            return set_error(AXS_ERR_NOMEM,
                             "axs_boxlevel_snapshot_coarse: out of memory");
        }
        b->snap_bytes = sb + lb;
    }
    b->snap_lo = lb ? (float *)((char *)b->snap + sb) : nullptr;
    b->snap_uid = c->uid;
    AXS_HIP_TRY(hipMemcpyAsync(b->snap, c->y, sb, hipMemcpyDeviceToDevice,
                               b->ctx->stream));
    if (lb)
        AXS_HIP_TRY(hipMemcpyAsync(b->snap_lo, c->comp, lb,
                                   hipMemcpyDeviceToDevice, b->ctx->stream));
    return AXS_OK;
}

int axs_boxlevel_fill_ghosts(axs_boxlevel *b, const axs_boxlevel *coarse,
                             double t_frac, int interior)
{
    AXS_CHECK_ARG(b, "axs_boxlevel_fill_ghosts: NULL level");
    if (interior)
        ++b->ygen; // writes the state's valid cells
    AXS_CHECK_ARG(!interior || coarse,
                  "axs_boxlevel_fill_ghosts: interior fill needs a coarse level");
    if (coarse)
    {
        AXS_CHECK_ARG(coarse->precision == b->precision,
                      "axs_boxlevel_fill_ghosts: precisions differ");
        for (int d = 0; d < 3; ++d)
            AXS_CHECK_ARG(coarse->dom[d] * 2 == b->dom[d],
                          "axs_boxlevel_fill_ghosts: coarse domain must be half");
    }
    hipLaunchKernelGGL(k_zero_err, dim3(1), dim3(64), 0, b->ctx->stream, b->derr);
    int rc = fill(b, b->y, b->comp, coarse, t_frac, interior);
    if (rc)
        return rc;
    return check_err(b, "axs_boxlevel_fill_ghosts");
}

} // extern "C"

// Stage scalars into the level's device table through its pinned staging
// buffer: an asynchronous copy, ordered before the stage launches on the
// stream.  Every step call ends with a stream synchronisation (the error
// count read-back), so the staging buffer is free again by the next call.
static int upload_stage_scalars(axs_boxlevel *b, const axs_rhs_coeffs coeffs[4])
{
    for (int s = 0; s < 4; ++s)
    {
        b->h_coeff[3 * s] = coeffs[s].curvature_coeff;
        b->h_coeff[3 * s + 1] = coeffs[s].lambda;
        b->h_coeff[3 * s + 2] = coeffs[s].R_squared;
    }
    AXS_HIP_TRY(hipMemcpyAsync(b->dcoeff, b->h_coeff, 12 * sizeof(double),
                               hipMemcpyHostToDevice, b->ctx->stream));
    return AXS_OK;
}

extern "C" {

int axs_boxlevel_rk4_step(axs_boxlevel *b, const axs_rhs_coeffs coeffs[4],
                          double dt, const axs_boxlevel *coarse, double t0,
                          double df)
{
    AXS_CHECK_ARG(b && coeffs, "axs_boxlevel_rk4_step: NULL argument");
    if (coarse)
        for (int d = 0; d < 3; ++d)
            AXS_CHECK_ARG(coarse->dom[d] * 2 == b->dom[d] &&
                              coarse->precision == b->precision,
                          "axs_boxlevel_rk4_step: coarse level mismatch");
    hipStream_t s = b->ctx->stream;
    int rc0 = upload_stage_scalars(b, coeffs);
    if (rc0)
        return rc0;
    if ((rc0 = ensure_edges(b)))
        return rc0;
    hipLaunchKernelGGL(k_zero_err, dim3(1), dim3(64), 0, s, b->derr);
    const int rc = step_launches(b, coarse, dt, t0, df);
    if (rc)
    {
        ++b->ygen; // launches may have run part of the step
        return rc;
    }
    step_done(b);
    return check_err(b, "axs_boxlevel_rk4_step");
}

int axs_boxlevel_rk4_step_graph(axs_boxlevel *b, const axs_rhs_coeffs coeffs[4],
                                double dt, const axs_boxlevel *coarse,
                                double t0, double df)
{
    AXS_CHECK_ARG(b && coeffs, "axs_boxlevel_rk4_step_graph: NULL argument");
    if (coarse)
        for (int d = 0; d < 3; ++d)
            AXS_CHECK_ARG(coarse->dom[d] * 2 == b->dom[d] &&
                              coarse->precision == b->precision,
                          "axs_boxlevel_rk4_step_graph: coarse level mismatch");
    hipStream_t s = b->ctx->stream;
    int rc0 = upload_stage_scalars(b, coeffs);
    if (rc0)
        return rc0;
    if ((rc0 = ensure_edges(b)))
        return rc0;
    axs_boxlevel::GraphEntry *hit = nullptr;
    const unsigned long long cuid = coarse ? coarse->uid : 0ull;
    const void *cstate = coarse ? coarse->y : nullptr;
    for (auto &e : b->graphs)
        if (e.dt == dt && e.t0 == t0 && e.df == df && e.coarse_uid == cuid &&
            e.state == b->y && e.coarse_state == cstate &&
            e.fusion == b->fusion)
            hit = &e;
    if (!hit)
    {
        AXS_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = step_launches(b, coarse, dt, t0, df);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != AXS_OK || e != hipSuccess)
        {
            if (g)
                (void)hipGraphDestroy(g);
            return rc != AXS_OK
                       ? rc
                       : set_error(AXS_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                                    hipGetErrorString(e));
        }
        hipGraphExec_t ge = nullptr;
        if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess)
        {
            (void)hipGraphDestroy(g);
            return set_error(AXS_ERR_HIP, "hipGraphInstantiate failed");
        }
        b->graphs.push_back({dt, t0, df, cuid, b->y, cstate, b->fusion, g, ge});
        hit = &b->graphs.back();
    }
    // the ghost fills of the replay count cells without an owning box in
    // derr, cleared by a kernel before the replay (not a memset node) and
    // read back after it, as the direct path does
    hipLaunchKernelGGL(k_zero_err, dim3(1), dim3(64), 0, s, b->derr);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipGraphLaunch(hit->ge, s));
    step_done(b);
    return check_err(b, "axs_boxlevel_rk4_step_graph");
}

int axs_boxlevel_tag(axs_boxlevel *b, const axs_tagger_params *p,
                     uint8_t *dense_tags, const uint8_t *cover,
                     int64_t counts[2])
{
    AXS_CHECK_ARG(b && counts, "axs_boxlevel_tag: NULL argument");
    AXS_CHECK_ARG(b->ng >= 2, "axs_boxlevel_tag: needs >= 2 ghost cells");
    hipStream_t s = b->ctx->stream;
    hipLaunchKernelGGL(k_zero_err, dim3(1), dim3(64), 0, s, b->ctx->d_counts);
    const dim3 g = box_grid(b, max_valid(b));
    BoxTagParams tp{};
    tp.amp_thr = p ? p->amplitude_threshold : 0.0;
    tp.R = p ? p->R : 1.0;
    tp.grad_thr = p ? p->gradient_threshold : DBL_MAX;
    tp.rgrad_thr = p ? p->radial_gradient_threshold : DBL_MAX;
    tp.dx = b->dx;
    if (b->precision == AXS_FP64)
        hipLaunchKernelGGL(k_box_tag<double>, g, dim3(256), 0, s, view_of(b),
                           (const double *)b->y, (const float *)nullptr,
                           dense_tags, cover, tp, b->ctx->d_counts);
    else
        hipLaunchKernelGGL(k_box_tag<float>, g, dim3(256), 0, s, view_of(b),
                           (const float *)b->y, (const float *)b->comp,
                           dense_tags, cover, tp, b->ctx->d_counts);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipMemcpyAsync(b->ctx->h_counts, b->ctx->d_counts,
                               2 * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    counts[0] = (int64_t)b->ctx->h_counts[0];
    counts[1] = (int64_t)b->ctx->h_counts[1];
    return AXS_OK;
}

int axs_boxlevel_energy_sums(axs_boxlevel *b, double R, double lambda,
                             double RpR, const axs_masking *m,
                             const uint8_t *cover, double out[14])
{
    AXS_CHECK_ARG(b && m && out, "axs_boxlevel_energy_sums: NULL argument");
    AXS_CHECK_ARG(R > 0.0, "axs_boxlevel_energy_sums: R must be > 0");
    hipStream_t s = b->ctx->stream;
    // one row of 14 partials per block of the work-item pass
    const size_t cap_rows = ((size_t)kMaxPartials * 16 - 14) / 14;
    double *part = b->ctx->d_partials;
    const size_t rows = (size_t)std::max(
        1, std::min<int>(b->nitems, (int)std::min<size_t>(cap_rows, 4096)));
    if (b->precision == AXS_FP64)
        hipLaunchKernelGGL(k_box_energy_items<double>, dim3((unsigned)rows),
                           dim3(kBoxLanes), 0, s, b->dbox, b->ditems, b->nitems,
                           b->ng, b->total, b->dom[0], b->dom[1],
                           (const double *)b->y, (const float *)nullptr, cover,
                           b->dx, R, lambda, RpR, m->scheme, m->threshold, part);
    else
        hipLaunchKernelGGL(k_box_energy_items<float>, dim3((unsigned)rows),
                           dim3(kBoxLanes), 0, s, b->dbox, b->ditems, b->nitems,
                           b->ng, b->total, b->dom[0], b->dom[1],
                           (const float *)b->y, (const float *)b->comp, cover,
                           b->dx, R, lambda, RpR, m->scheme, m->threshold, part);
    hipLaunchKernelGGL(k_box_sum_cols, dim3(14), dim3(256), 0, s,
                       (const double *)part, (int)rows, 14, b->ctx->d_result);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipMemcpyAsync(b->ctx->h_result, b->ctx->d_result,
                               14 * sizeof(double), hipMemcpyDeviceToHost, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    for (int q = 0; q < 14; ++q)
        out[q] = b->ctx->h_result[q];
    return AXS_OK;
}

int axs_boxlevel_coverage_mask(const axs_boxlevel *f, const int cd[3],
                               uint8_t *mask)
{
    AXS_CHECK_ARG(f && cd && mask, "axs_boxlevel_coverage_mask: NULL argument");
    for (int d = 0; d < 3; ++d)
        AXS_CHECK_ARG(cd[d] * 2 == f->dom[d],
                      "axs_boxlevel_coverage_mask: coarse domain must be half");
    hipStream_t s = f->ctx->stream;
    const size_t n = (size_t)cd[0] * cd[1] * cd[2];
    hipLaunchKernelGGL(k_set_u8, dim3(gblk(n)), dim3(256), 0, s, mask, n,
                       (uint8_t)1);
    hipLaunchKernelGGL(k_coverage, box_grid(f, max_valid(f) / 8), dim3(256), 0, s,
                       view_of(f), cd[0], cd[1], mask);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

int axs_boxlevel_average_down(const axs_boxlevel *f, axs_boxlevel *c)
{
    AXS_CHECK_ARG(f && c, "axs_boxlevel_average_down: NULL argument");
    ++c->ygen; // writes the state
    for (int d = 0; d < 3; ++d)
        AXS_CHECK_ARG(c->dom[d] * 2 == f->dom[d],
                      "axs_boxlevel_average_down: coarse domain must be half");
    AXS_CHECK_ARG(c->precision == f->precision,
                  "axs_boxlevel_average_down: precisions differ");
    hipStream_t s = f->ctx->stream;
    hipLaunchKernelGGL(k_zero_err, dim3(1), dim3(64), 0, s, c->derr);
    const dim3 g = box_grid(f, max_valid(f) / 8);
    if (f->precision == AXS_FP64)
        hipLaunchKernelGGL(k_average_down<double>, g, dim3(256), 0, s, view_of(f),
                           (const double *)f->y, (const float *)nullptr,
                           view_of(c), (double *)c->y, (float *)nullptr, c->derr);
    else
        hipLaunchKernelGGL(k_average_down<float>, g, dim3(256), 0, s, view_of(f),
                           (const float *)f->y, (const float *)f->comp,
                           view_of(c), (float *)c->y, c->comp, c->derr);
    AXS_HIP_TRY(hipGetLastError());
    return check_err(c, "axs_boxlevel_average_down");
}

} // extern "C"
