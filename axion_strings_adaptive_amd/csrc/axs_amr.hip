// axs_amr.hip -- the AMR hierarchy's host logic over box levels: tagging
// into blocks and boxes, subcycled stepping, regridding and the composite
// diagnostics (the role of AMReX's Amr driver around AxionStringsLevel in
// the reference: tag_cells AxionStringsLevel.cpp:1916-2005, the subcycled
// RK4 of GRTeclyn's GRAmrLevel, collapse_to_level0 :110-134, the composite
// diagnostics :1256-1361).  Host code only: every device step is one of the
// axs_boxlevel_* entry points.
#include "axs_internal.hpp"

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <vector>

namespace
{

struct BoxSpec
{
    int lo[3], hi[3];
};

using Clock = std::chrono::steady_clock;

// ------------------------------------------------ blocks and boxes ----
// tags [z][y][x] of size dom -> blocks of `block` cells that hold a tag
// within r cells (periodic); dilation by a cube and "any over a block" both
// act axis by axis, so each axis is reduced on its own: block b of an axis
// takes cells [b block - r, (b + 1) block + r).
std::vector<uint8_t> tagged_blocks(const uint8_t *tags, const int dom[3],
                                   int block, int r)
{
    const int nx = dom[0], ny = dom[1], nz = dom[2];
    const int bx = nx / block, by = ny / block, bz = nz / block;
    auto any_window = [&](auto get, int n, int b) {
        const int k0 = b * block - r, k1 = (b + 1) * block + r;
        if (k0 >= 0 && k1 <= n)
        {
            for (int i = k0; i < k1; ++i)
                if (get(i))
                    return true;
            return false;
        }
        for (int k = k0; k < k1; ++k)
            if (get(((k % n) + n) % n))
                return true;
        return false;
    };
    // x: [z][y][bx]
    std::vector<uint8_t> ax((size_t)nz * ny * bx);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
        {
            const uint8_t *row = tags + ((size_t)z * ny + y) * nx;
            for (int b = 0; b < bx; ++b)
                ax[((size_t)z * ny + y) * bx + b] =
                    any_window([&](int i) { return row[i] != 0; }, nx, b);
        }
    // y: [z][by][bx]
    std::vector<uint8_t> ay((size_t)nz * by * bx);
    for (int z = 0; z < nz; ++z)
        for (int b = 0; b < by; ++b)
            for (int x = 0; x < bx; ++x)
                ay[((size_t)z * by + b) * bx + x] = any_window(
                    [&](int i) { return ax[((size_t)z * ny + i) * bx + x] != 0; },
                    ny, b);
    // z: [bz][by][bx]
    std::vector<uint8_t> out((size_t)bz * by * bx);
    for (int b = 0; b < bz; ++b)
        for (int y = 0; y < by; ++y)
            for (int x = 0; x < bx; ++x)
                out[((size_t)b * by + y) * bx + x] = any_window(
                    [&](int i) { return ay[((size_t)i * by + y) * bx + x] != 0; },
                    nz, b);
    return out;
}

// tagged_blocks on the device, from device tags: block (bx, by, bz) is set
// where a tag lies in [b block - r, (b + 1) block + r) along every axis
// (periodic), the set the axis-by-axis reduction above forms.  One
// wavefront per block, its lanes striding the (block + 2r)^3 window; only
// the block mask (dom / block cells) then goes to the host.
__global__ void k_tagged_blocks(const uint8_t *tags, int nx, int ny, int nz,
                                int block, int r, uint8_t *out)
{
    const int bx = nx / block, by = ny / block, bz = nz / block;
    const size_t nb = (size_t)bx * by * bz;
    const int lane = threadIdx.x & 63;
    const size_t nw = ((size_t)gridDim.x * blockDim.x) >> 6;
    const int w = block + 2 * r, w3 = w * w * w;
    for (size_t b = (((size_t)blockIdx.x * blockDim.x) + threadIdx.x) >> 6;
         b < nb; b += nw)
    {
        const int ix = (int)(b % bx), iy = (int)((b / bx) % by),
                  iz = (int)(b / ((size_t)bx * by));
        bool any = false;
        for (int s = lane; s < w3 && !any; s += 64)
        {
            const int dx = s % w, dy = (s / w) % w, dz = s / (w * w);
            int x = ix * block - r + dx, y = iy * block - r + dy,
                z = iz * block - r + dz;
            x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
            y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
            z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
            any = tags[((size_t)z * ny + y) * nx + x] != 0;
        }
        const bool set = __ballot(any) != 0ull;
        if (lane == 0)
            out[b] = set ? 1 : 0;
    }
}

// a block mask kept where its whole periodic (2r+1)^3 neighbourhood is set
std::vector<uint8_t> erode(const uint8_t *m, const int bd[3], int r)
{
    const int bx = bd[0], by = bd[1], bz = bd[2];
    std::vector<uint8_t> out((size_t)bx * by * bz, 0);
    for (int z = 0; z < bz; ++z)
        for (int y = 0; y < by; ++y)
            for (int x = 0; x < bx; ++x)
            {
                bool all = true;
                for (int dz = -r; dz <= r && all; ++dz)
                    for (int dy = -r; dy <= r && all; ++dy)
                        for (int dx = -r; dx <= r && all; ++dx)
                        {
                            const int X = ((x + dx) % bx + bx) % bx;
                            const int Y = ((y + dy) % by + by) % by;
                            const int Z = ((z + dz) % bz + bz) % bz;
                            all = m[((size_t)Z * by + Y) * bx + X] != 0;
                        }
                out[((size_t)z * by + y) * bx + x] = all;
            }
    return out;
}

// Fine-level boxes over the set blocks: runs of up to max_blocks blocks
// along x; runs with the same x extent in consecutive rows merged along y,
// then slabs with the same x, y extent in consecutive planes merged along
// z, each merge again at most max_blocks long.  Sorted by the low corner's
// z, y, x.
std::vector<BoxSpec> boxes_from_blocks(const uint8_t *blk, const int bd[3],
                                       int fine_block, int max_blocks)
{
    const int bx = bd[0], by = bd[1], bz = bd[2];
    using Run = std::pair<int, int>;                   // i0, i1 (exclusive)
    using Slab = std::tuple<int, int, int, int>;       // i0, i1, j0, j1
    std::vector<std::vector<std::set<Run>>> runs(
        bz, std::vector<std::set<Run>>(by));
    for (int k = 0; k < bz; ++k)
        for (int j = 0; j < by; ++j)
        {
            const uint8_t *row = blk + ((size_t)k * by + j) * bx;
            int i = 0;
            while (i < bx)
            {
                if (!row[i])
                {
                    ++i;
                    continue;
                }
                const int i0 = i;
                while (i < bx && row[i] && i - i0 < max_blocks)
                    ++i;
                runs[k][j].insert({i0, i});
            }
        }
    std::vector<std::set<Slab>> slabs(bz);
    for (int k = 0; k < bz; ++k)
    {
        std::map<Run, int> open; // run -> first row
        for (int j = 0; j <= by; ++j)
        {
            static const std::set<Run> none;
            const std::set<Run> &cur = j < by ? runs[k][j] : none;
            for (auto it = open.begin(); it != open.end();)
            {
                if (!cur.count(it->first) || j - it->second == max_blocks)
                {
                    slabs[k].insert({it->first.first, it->first.second,
                                     it->second, j});
                    it = open.erase(it);
                }
                else
                    ++it;
            }
            for (const Run &r : cur)
                if (!open.count(r))
                    open[r] = j;
        }
    }
    std::vector<BoxSpec> boxes;
    std::map<Slab, int> open; // slab -> first plane
    for (int k = 0; k <= bz; ++k)
    {
        static const std::set<Slab> none;
        const std::set<Slab> &cur = k < bz ? slabs[k] : none;
        for (auto it = open.begin(); it != open.end();)
        {
            if (!cur.count(it->first) || k - it->second == max_blocks)
            {
                const auto [i0, i1, j0, j1] = it->first;
                const int k0 = it->second;
                boxes.push_back(BoxSpec{
                    {i0 * fine_block, j0 * fine_block, k0 * fine_block},
                    {i1 * fine_block - 1, j1 * fine_block - 1,
                     k * fine_block - 1}});
                it = open.erase(it);
            }
            else
                ++it;
        }
        for (const Slab &s : cur)
            if (!open.count(s))
                open[s] = k;
    }
    std::sort(boxes.begin(), boxes.end(), [](const BoxSpec &a, const BoxSpec &b) {
        return std::make_tuple(a.lo[2], a.lo[1], a.lo[0]) <
               std::make_tuple(b.lo[2], b.lo[1], b.lo[0]);
    });
    return boxes;
}

bool same_boxes(std::vector<BoxSpec> a, std::vector<BoxSpec> b)
{
    if (a.size() != b.size())
        return false;
    auto key = [](const BoxSpec &x) {
        return std::make_tuple(x.lo[0], x.lo[1], x.lo[2], x.hi[0], x.hi[1],
                               x.hi[2]);
    };
    auto less = [&](const BoxSpec &x, const BoxSpec &y) { return key(x) < key(y); };
    std::sort(a.begin(), a.end(), less);
    std::sort(b.begin(), b.end(), less);
    for (size_t i = 0; i < a.size(); ++i)
        if (key(a[i]) != key(b[i]))
            return false;
    return true;
}

#define AXS_TRY(expr)                                                         \
    do                                                                        \
    {                                                                         \
        const int rc_ = (expr);                                               \
        if (rc_ != AXS_OK)                                                    \
            return rc_;                                                       \
    } while (0)

} // namespace

struct axs_hierarchy
{
    axs_ctx *ctx;
    axs_amr_params p;
    axs_background bg;
    axs_coeff_blend blend;
    int precision;
    double tau, dt0;
    std::vector<axs_boxlevel *> levels;
    std::vector<std::array<int, 3>> domains;
    std::vector<double> dx;
    std::vector<std::vector<BoxSpec>> boxes;
    int64_t nsteps, nregrids;
    double regrid_s[6];
    void *dtags;  // dense uint8 tags of the largest tagged domain
    size_t dtags_bytes;
    std::vector<void *> dcover; // coverage masks of levels 0..n-2
    std::vector<size_t> dcover_bytes;
};

namespace
{

size_t cells_of(const std::array<int, 3> &d)
{
    return (size_t)d[0] * d[1] * d[2];
}

int device_buffer(axs_hierarchy *h, void **buf, size_t *have, size_t need)
{
    if (*have >= need)
        return AXS_OK;
    AXS_TRY(axs_device_free(h->ctx, *buf));
    *buf = nullptr;
    *have = 0;
    AXS_TRY(axs_device_alloc(h->ctx, need, buf));
    *have = need;
    return AXS_OK;
}

// stage scalars at tau: the background's, blended after a handoff
int coeffs_at(const axs_hierarchy *h, double tau, axs_rhs_coeffs *c)
{
    AXS_TRY(axs_background_rhs_coeffs(&h->bg, tau, c));
    const axs_coeff_blend &b = h->blend;
    if (b.dtau > 0.0)
    {
        const double t = (tau - b.tau_start) / b.dtau;
        if (t < 1.0)
        {
            const double tc = std::clamp(t, 0.0, 1.0);
            const double s = tc * tc * (3.0 - 2.0 * tc);
            c->lambda = (1.0 - s) * b.lambda_start + s * c->lambda;
            c->curvature_coeff =
                (1.0 - s) * b.curvature_start + s * c->curvature_coeff;
        }
    }
    return AXS_OK;
}

// tag_cells' schedule gate (AxionStringsLevel.cpp:1957-1983): level l may
// exist once log(m_r/H) has reached entry l-1 of the schedule
bool level_allowed(const axs_hierarchy *h, int l)
{
    if (h->p.force_full_refinement || h->p.n_schedule <= 0)
        return true;
    if (l > h->p.n_schedule)
        return false;
    double hm = 0.0;
    if (axs_background_eval(&h->bg, 5, h->tau, &hm) != AXS_OK)
        return false;
    return -std::log(hm) >= h->p.level_add_log_mr_over_h[l - 1];
}

// Boxes of level l+1 from the tags of level l (its ghosts filled): tag,
// grow by n_error_buf, cut into blocks of blocking_factor / 2 cells of
// level l, keep (l >= 1) those one block inside level l's coverage, merge.
int fine_boxes(axs_hierarchy *h, int l, const std::vector<uint8_t> &covered,
               std::vector<BoxSpec> *boxes, std::vector<uint8_t> *blk)
{
    boxes->clear();
    blk->clear();
    if (!level_allowed(h, l + 1))
        return AXS_OK;
    const std::array<int, 3> &cd = h->domains[l];
    const size_t n = cells_of(cd);
    const int cblock = h->p.blocking_factor / 2;
    const int bd[3] = {cd[0] / cblock, cd[1] / cblock, cd[2] / cblock};
    const size_t nb = (size_t)bd[0] * bd[1] * bd[2];
    blk->assign(nb, 1); // every block tagged under force_full_refinement
    if (!h->p.force_full_refinement)
    {
        double R = 1.0;
        AXS_TRY(axs_background_eval(&h->bg, 0, h->tau, &R));
        // dense tags, then the block mask behind them
        AXS_TRY(device_buffer(h, &h->dtags, &h->dtags_bytes, n + nb));
        AXS_TRY(axs_device_memset(h->ctx, h->dtags, 0, n));
        const axs_tagger_params tp{h->p.gradient_threshold,
                                   h->p.radial_gradient_threshold,
                                   h->p.amplitude_threshold, R};
        int64_t counts[2];
        AXS_TRY(axs_boxlevel_tag(h->levels[l], &tp, (uint8_t *)h->dtags,
                                 nullptr, counts));
        uint8_t *dmask = (uint8_t *)h->dtags + n;
        hipLaunchKernelGGL(k_tagged_blocks,
                           dim3((unsigned)std::min<size_t>(4096, (nb + 3) / 4)),
                           dim3(256), 0, h->ctx->stream,
                           (const uint8_t *)h->dtags, cd[0], cd[1], cd[2],
                           cblock, h->p.n_error_buf, dmask);
        AXS_HIP_TRY(hipGetLastError());
        AXS_TRY(axs_device_download(h->ctx, blk->data(), dmask, nb));
    }
    if (l >= 1)
    {
        const std::vector<uint8_t> inner = erode(covered.data(), bd, 1);
        for (size_t i = 0; i < blk->size(); ++i)
            (*blk)[i] = (*blk)[i] && inner[i];
    }
    *boxes = boxes_from_blocks(blk->data(), bd, h->p.blocking_factor,
                               h->p.max_grid_size / h->p.blocking_factor);
    return AXS_OK;
}

// the block mask of the next finer level: each block split in 2 per axis
std::vector<uint8_t> refine_blocks(const std::vector<uint8_t> &blk,
                                   const int bd[3])
{
    std::vector<uint8_t> out((size_t)8 * blk.size());
    const int fx = 2 * bd[0], fy = 2 * bd[1];
    for (int z = 0; z < 2 * bd[2]; ++z)
        for (int y = 0; y < fy; ++y)
            for (int x = 0; x < fx; ++x)
                out[((size_t)z * fy + y) * fx + x] =
                    blk[((size_t)(z / 2) * bd[1] + y / 2) * bd[0] + x / 2];
    return out;
}

int create_level(axs_hierarchy *h, const std::vector<BoxSpec> &boxes,
                 const std::array<int, 3> &dom, double dx, axs_boxlevel **out)
{
    std::vector<int> lo(3 * boxes.size()), hi(3 * boxes.size());
    for (size_t b = 0; b < boxes.size(); ++b)
        for (int d = 0; d < 3; ++d)
        {
            lo[3 * b + d] = boxes[b].lo[d];
            hi[3 * b + d] = boxes[b].hi[d];
        }
    const int rc = axs_boxlevel_create(h->ctx, (int)boxes.size(), lo.data(),
                                       hi.data(), dom.data(), 3, dx,
                                       h->precision, out);
    if (rc == AXS_OK && h->p.per_stage_launches)
        return axs_boxlevel_set_stage_fusion(*out, 0);
    return rc;
}

std::array<int, 3> block_dims(const axs_hierarchy *h, int l)
{
    const int cblock = h->p.blocking_factor / 2;
    return {h->domains[l][0] / cblock, h->domains[l][1] / cblock,
            h->domains[l][2] / cblock};
}

int advance(axs_hierarchy *h, int l, double tau, double dt, double t0,
            double df)
{
    axs_boxlevel *L = h->levels[l];
    axs_boxlevel *coarse = l > 0 ? h->levels[l - 1] : nullptr;
    const bool finer = l + 1 < (int)h->levels.size();
    if (finer)
        AXS_TRY(axs_boxlevel_snapshot_coarse(h->levels[l + 1], L));
    axs_rhs_coeffs c[4];
    const double ts[4] = {tau, tau + 0.5 * dt, tau + 0.5 * dt, tau + dt};
    for (int s = 0; s < 4; ++s)
        AXS_TRY(coeffs_at(h, ts[s], &c[s]));
    AXS_TRY(h->p.use_graph
                ? axs_boxlevel_rk4_step_graph(L, c, dt, coarse, t0, df)
                : axs_boxlevel_rk4_step(L, c, dt, coarse, t0, df));
    if (finer)
    {
        AXS_TRY(advance(h, l + 1, tau, 0.5 * dt, 0.0, 0.5));
        AXS_TRY(advance(h, l + 1, tau + 0.5 * dt, 0.5 * dt, 0.5, 0.5));
        AXS_TRY(axs_boxlevel_average_down(h->levels[l + 1], L));
    }
    return AXS_OK;
}

// coverage masks of every level but the finest (0 = a finer level covers
// the cell); the masks live until the hierarchy changes
int coverage(axs_hierarchy *h)
{
    const size_t nl = h->levels.size();
    if (h->dcover.size() < nl)
    {
        h->dcover.resize(nl, nullptr);
        h->dcover_bytes.resize(nl, 0);
    }
    for (size_t l = 0; l + 1 < nl; ++l)
    {
        const size_t n = cells_of(h->domains[l]);
        AXS_TRY(device_buffer(h, &h->dcover[l], &h->dcover_bytes[l], n));
        AXS_TRY(axs_boxlevel_coverage_mask(h->levels[l + 1],
                                           h->domains[l].data(),
                                           (uint8_t *)h->dcover[l]));
    }
    return AXS_OK;
}

int fill_all_ghosts(axs_hierarchy *h)
{
    for (size_t l = 0; l < h->levels.size(); ++l)
        AXS_TRY(axs_boxlevel_fill_ghosts(h->levels[l],
                                         l ? h->levels[l - 1] : nullptr, 1.0,
                                         0));
    return AXS_OK;
}

} // namespace

using namespace axs;

extern "C" {

int axs_hierarchy_create(axs_ctx *ctx, const axs_level *base,
                         const axs_amr_params *p, const axs_background *bg,
                         const axs_coeff_blend *blend, double tau, double dt0,
                         axs_hierarchy **out)
{
    AXS_CHECK_ARG(ctx && base && p && bg && out,
                  "axs_hierarchy_create: NULL argument");
    *out = nullptr;
    AXS_CHECK_ARG(p->max_level >= 1 && p->max_level <= AXS_MAX_LEVELS,
                  "axs_hierarchy_create: max_level must be 1..8");
    const int bf = p->blocking_factor;
    AXS_CHECK_ARG(bf == 8 || bf == 16 || bf == 32 || bf == 64,
                  "axs_hierarchy_create: blocking_factor must be 8, 16, 32 "
                  "or 64");
    AXS_CHECK_ARG(p->max_grid_size >= bf && p->max_grid_size <= 64 &&
                      p->max_grid_size % bf == 0,
                  "axs_hierarchy_create: max_grid_size must be a multiple of "
                  "blocking_factor and at most 64 (box edges of 8..64 cells)");
    AXS_CHECK_ARG(p->n_error_buf >= 0 && p->n_error_buf <= bf / 2,
                  "axs_hierarchy_create: 0 <= n_error_buf <= "
                  "blocking_factor / 2");
    AXS_CHECK_ARG(p->n_schedule >= 0 && p->n_schedule <= AXS_MAX_LEVELS,
                  "axs_hierarchy_create: n_schedule must be 0..8");
    AXS_CHECK_ARG(dt0 > 0.0, "axs_hierarchy_create: dt0 must be > 0");
    const std::array<int, 3> d0 = {base->g.nx, base->g.ny, base->g.nz};
    const int e0 = std::min(p->max_grid_size, std::min({d0[0], d0[1], d0[2]}));
    for (int d = 0; d < 3; ++d)
        AXS_CHECK_ARG(d0[d] % bf == 0 && d0[d] % 8 == 0,
                      "axs_hierarchy_create: the base level's edges must be "
                      "multiples of blocking_factor");
    axs_hierarchy *h = new axs_hierarchy{};
    h->ctx = ctx;
    h->p = *p;
    h->bg = *bg;
    if (blend)
        h->blend = *blend;
    h->precision = base->precision;
    h->tau = tau;
    h->dt0 = dt0;
    auto fail = [&](int rc) {
        axs_hierarchy_destroy(h);
        return rc;
    };
    // level 0: the base grid in boxes of e0 cells per edge, z, y, x order
    std::vector<BoxSpec> b0;
    for (int z0 = 0; z0 < d0[2]; z0 += e0)
        for (int y0 = 0; y0 < d0[1]; y0 += e0)
            for (int x0 = 0; x0 < d0[0]; x0 += e0)
                b0.push_back(BoxSpec{{x0, y0, z0},
                                     {std::min(x0 + e0, d0[0]) - 1,
                                      std::min(y0 + e0, d0[1]) - 1,
                                      std::min(z0 + e0, d0[2]) - 1}});
    axs_boxlevel *L0 = nullptr;
    int rc = create_level(h, b0, d0, base->dx, &L0);
    if (rc)
        return fail(rc);
    h->levels.push_back(L0);
    h->domains.push_back(d0);
    h->dx.push_back(base->dx);
    h->boxes.push_back(b0);
    if ((rc = axs_boxlevel_copy_from_level(L0, base)) ||
        (rc = axs_boxlevel_fill_ghosts(L0, nullptr, 1.0, 0)))
        return fail(rc);
    std::vector<uint8_t> covered;
    for (int l = 1; l <= p->max_level; ++l)
    {
        std::vector<BoxSpec> boxes;
        std::vector<uint8_t> blk;
        if ((rc = fine_boxes(h, l - 1, covered, &boxes, &blk)))
            return fail(rc);
        if (boxes.empty())
            break;
        const std::array<int, 3> fd = {2 * h->domains.back()[0],
                                       2 * h->domains.back()[1],
                                       2 * h->domains.back()[2]};
        const double fdx = h->dx.back() / 2;
        axs_boxlevel *Lf = nullptr;
        if ((rc = create_level(h, boxes, fd, fdx, &Lf)))
            return fail(rc);
        h->levels.push_back(Lf);
        h->domains.push_back(fd);
        h->dx.push_back(fdx);
        h->boxes.push_back(boxes);
        if ((rc = axs_boxlevel_fill_ghosts(Lf, h->levels[l - 1], 1.0, 1)))
            return fail(rc);
        const std::array<int, 3> bd = block_dims(h, l - 1);
        covered = refine_blocks(blk, bd.data());
    }
    *out = h;
    return AXS_OK;
}

int axs_hierarchy_destroy(axs_hierarchy *h)
{
    if (!h)
        return AXS_OK;
    for (axs_boxlevel *L : h->levels)
        axs_boxlevel_destroy(L);
    axs_device_free(h->ctx, h->dtags);
    for (void *c : h->dcover)
        axs_device_free(h->ctx, c);
    delete h;
    return AXS_OK;
}

int axs_hierarchy_regrid(axs_hierarchy *h)
{
    AXS_CHECK_ARG(h, "axs_hierarchy_regrid: NULL hierarchy");
    auto lap = [&](int phase, Clock::time_point &t) -> int {
        AXS_TRY(axs_ctx_synchronize(h->ctx));
        const Clock::time_point now = Clock::now();
        h->regrid_s[phase] += std::chrono::duration<double>(now - t).count();
        t = now;
        return AXS_OK;
    };
    AXS_TRY(axs_ctx_synchronize(h->ctx));
    const Clock::time_point start = Clock::now();
    Clock::time_point t = start;
    std::vector<uint8_t> covered;
    for (int l = 1; l <= h->p.max_level; ++l)
    {
        axs_boxlevel *coarse = h->levels[l - 1];
        AXS_TRY(axs_boxlevel_fill_ghosts(coarse, l >= 2 ? h->levels[l - 2]
                                                        : nullptr,
                                         1.0, 0));
        std::vector<BoxSpec> boxes;
        std::vector<uint8_t> blk;
        AXS_TRY(fine_boxes(h, l - 1, covered, &boxes, &blk));
        AXS_TRY(lap(1, t));
        const bool have_old = l < (int)h->levels.size();
        if (boxes.empty())
        {
            while ((int)h->levels.size() > l)
            {
                axs_boxlevel_destroy(h->levels.back());
                h->levels.pop_back();
                h->domains.pop_back();
                h->dx.pop_back();
                h->boxes.pop_back();
            }
            break;
        }
        const std::array<int, 3> bd = block_dims(h, l - 1);
        std::vector<uint8_t> covered_next = refine_blocks(blk, bd.data());
        if (have_old && same_boxes(h->boxes[l], boxes))
        {
            covered.swap(covered_next);
            continue;
        }
        const std::array<int, 3> fd = {2 * h->domains[l - 1][0],
                                       2 * h->domains[l - 1][1],
                                       2 * h->domains[l - 1][2]};
        const double fdx = h->dx[l - 1] / 2;
        axs_boxlevel *Lf = nullptr;
        AXS_TRY(create_level(h, boxes, fd, fdx, &Lf));
        AXS_TRY(lap(2, t));
        int rc = axs_boxlevel_fill_ghosts(Lf, coarse, 1.0, 1);
        if (rc)
        {
            axs_boxlevel_destroy(Lf);
            return rc;
        }
        AXS_TRY(lap(3, t));
        if (have_old)
        {
            axs_boxlevel *old = h->levels[l];
            if ((rc = axs_boxlevel_copy_valid(Lf, old, nullptr)))
            {
                axs_boxlevel_destroy(Lf);
                return rc;
            }
            AXS_TRY(lap(4, t));
            axs_boxlevel_destroy(old);
            AXS_TRY(lap(5, t));
            h->levels[l] = Lf;
            h->boxes[l] = boxes;
        }
        else
        {
            h->levels.push_back(Lf);
            h->domains.push_back(fd);
            h->dx.push_back(fdx);
            h->boxes.push_back(boxes);
        }
        AXS_TRY(axs_boxlevel_fill_ghosts(Lf, coarse, 1.0, 0));
        AXS_TRY(lap(3, t));
        covered.swap(covered_next);
    }
    AXS_TRY(axs_ctx_synchronize(h->ctx));
    h->regrid_s[0] +=
        std::chrono::duration<double>(Clock::now() - start).count();
    h->nregrids += 1;
    return AXS_OK;
}

int axs_hierarchy_step(axs_hierarchy *h)
{
    AXS_CHECK_ARG(h, "axs_hierarchy_step: NULL hierarchy");
    AXS_TRY(advance(h, 0, h->tau, h->dt0, 0.0, 1.0));
    h->tau += h->dt0;
    h->nsteps += 1;
    if (h->p.regrid_int > 0 && h->nsteps % h->p.regrid_int == 0)
        AXS_TRY(axs_hierarchy_regrid(h));
    return AXS_OK;
}

int axs_hierarchy_info(const axs_hierarchy *h, int *nlevels, double *tau,
                       int64_t *nsteps, int64_t *nregrids,
                       double regrid_seconds[6])
{
    AXS_CHECK_ARG(h, "axs_hierarchy_info: NULL hierarchy");
    if (nlevels)
        *nlevels = (int)h->levels.size();
    if (tau)
        *tau = h->tau;
    if (nsteps)
        *nsteps = h->nsteps;
    if (nregrids)
        *nregrids = h->nregrids;
    if (regrid_seconds)
        std::memcpy(regrid_seconds, h->regrid_s, sizeof(h->regrid_s));
    return AXS_OK;
}

int axs_hierarchy_level(const axs_hierarchy *h, int l, axs_boxlevel **lvl,
                        int domain[3], double *dx, int *nboxes)
{
    AXS_CHECK_ARG(h && lvl, "axs_hierarchy_level: NULL argument");
    AXS_CHECK_ARG(l >= 0 && l < (int)h->levels.size(),
                  "axs_hierarchy_level: no such level");
    *lvl = h->levels[l];
    if (domain)
        for (int d = 0; d < 3; ++d)
            domain[d] = h->domains[l][d];
    if (dx)
        *dx = h->dx[l];
    if (nboxes)
        *nboxes = (int)h->boxes[l].size();
    return AXS_OK;
}

int axs_hierarchy_boxes(const axs_hierarchy *h, int l, int *lo, int *hi)
{
    AXS_CHECK_ARG(h && lo && hi, "axs_hierarchy_boxes: NULL argument");
    AXS_CHECK_ARG(l >= 0 && l < (int)h->levels.size(),
                  "axs_hierarchy_boxes: no such level");
    const std::vector<BoxSpec> &b = h->boxes[l];
    for (size_t i = 0; i < b.size(); ++i)
        for (int d = 0; d < 3; ++d)
        {
            lo[3 * i + d] = b[i].lo[d];
            hi[3 * i + d] = b[i].hi[d];
        }
    return AXS_OK;
}

int axs_hierarchy_energy_sums(axs_hierarchy *h, double R, double lambda,
                              double R_prime_over_R, const axs_masking *m,
                              double out[14])
{
    AXS_CHECK_ARG(h && m && out, "axs_hierarchy_energy_sums: NULL argument");
    AXS_TRY(fill_all_ghosts(h));
    AXS_TRY(coverage(h));
    double acc[14] = {0.0};
    for (size_t l = 0; l < h->levels.size(); ++l)
    {
        double e[14];
        const uint8_t *cover = l + 1 < h->levels.size()
                                   ? (const uint8_t *)h->dcover[l]
                                   : nullptr;
        AXS_TRY(axs_boxlevel_energy_sums(h->levels[l], R, lambda,
                                         R_prime_over_R, m, cover, e));
        const double vol = std::pow(h->dx[l], 3.0);
        for (int i = 0; i < 14; ++i)
            acc[i] += e[i] * vol;
    }
    std::memcpy(out, acc, sizeof(acc));
    return AXS_OK;
}

int axs_hierarchy_plaquette_counts(axs_hierarchy *h, int64_t *counts)
{
    AXS_CHECK_ARG(h && counts, "axs_hierarchy_plaquette_counts: NULL argument");
    AXS_TRY(fill_all_ghosts(h));
    AXS_TRY(coverage(h));
    for (size_t l = 0; l < h->levels.size(); ++l)
    {
        const uint8_t *cover = l + 1 < h->levels.size()
                                   ? (const uint8_t *)h->dcover[l]
                                   : nullptr;
        const axs_tagger_params none{DBL_MAX, DBL_MAX, 0.0, 1.0};
        AXS_TRY(axs_boxlevel_tag(h->levels[l], &none, nullptr, cover,
                                 counts + 2 * l));
    }
    return AXS_OK;
}

int axs_hierarchy_collapse(axs_hierarchy *h, axs_level *dst)
{
    AXS_CHECK_ARG(h && dst, "axs_hierarchy_collapse: NULL argument");
    AXS_CHECK_ARG(dst->g.nx == h->domains[0][0] && dst->g.ny == h->domains[0][1] &&
                      dst->g.nz == h->domains[0][2] &&
                      dst->precision == h->precision,
                  "axs_hierarchy_collapse: dst must have level 0's domain "
                  "and precision");
    // copies of levels n-2 .. 0, each taking the average of the (collapsed)
    // level above it where that level covers it
    const axs_boxlevel *src = h->levels.back();
    std::vector<axs_boxlevel *> tmp;
    int rc = AXS_OK;
    for (int l = (int)h->levels.size() - 2; l >= 0 && rc == AXS_OK; --l)
    {
        axs_boxlevel *c = nullptr;
        rc = create_level(h, h->boxes[l], h->domains[l], h->dx[l], &c);
        if (rc)
            break;
        tmp.push_back(c);
        if ((rc = axs_boxlevel_copy_valid(c, h->levels[l], nullptr)) ||
            (rc = axs_boxlevel_average_down(src, c)))
            break;
        src = c;
    }
    if (rc == AXS_OK)
        rc = axs_boxlevel_copy_to_level(src, dst);
    if (rc == AXS_OK)
        rc = axs_ctx_synchronize(h->ctx);
    for (axs_boxlevel *c : tmp)
        axs_boxlevel_destroy(c);
    return rc;
}

int axs_amr_tagged_blocks(const uint8_t *tags, const int dom[3], int block,
                          int r, uint8_t *out)
{
    AXS_CHECK_ARG(tags && dom && out, "axs_amr_tagged_blocks: NULL argument");
    AXS_CHECK_ARG(block >= 1 && r >= 0 && r <= block,
                  "axs_amr_tagged_blocks: need block >= 1, 0 <= r <= block");
    for (int d = 0; d < 3; ++d)
        AXS_CHECK_ARG(dom[d] > 0 && dom[d] % block == 0,
                      "axs_amr_tagged_blocks: dom must be multiples of block");
    const std::vector<uint8_t> b = tagged_blocks(tags, dom, block, r);
    std::memcpy(out, b.data(), b.size());
    return AXS_OK;
}

int axs_amr_tagged_blocks_device(axs_ctx *ctx, const uint8_t *tags_dev,
                                 const int dom[3], int block, int r,
                                 uint8_t *out_dev)
{
    AXS_CHECK_ARG(ctx && tags_dev && dom && out_dev,
                  "axs_amr_tagged_blocks_device: NULL argument");
    AXS_CHECK_ARG(block >= 1 && r >= 0 && r <= block,
                  "axs_amr_tagged_blocks_device: need block >= 1, 0 <= r <= block");
    for (int d = 0; d < 3; ++d)
        AXS_CHECK_ARG(dom[d] > 0 && dom[d] % block == 0,
                      "axs_amr_tagged_blocks_device: dom must be multiples of block");
    const size_t nb = (size_t)(dom[0] / block) * (dom[1] / block) * (dom[2] / block);
    hipLaunchKernelGGL(k_tagged_blocks,
                       dim3((unsigned)std::min<size_t>(4096, (nb + 3) / 4)),
                       dim3(256), 0, ctx->stream, tags_dev, dom[0], dom[1],
                       dom[2], block, r, out_dev);
    AXS_HIP_TRY(hipGetLastError());
    return axs_ctx_synchronize(ctx);
}

int axs_amr_erode(const uint8_t *mask, const int bdim[3], int r, uint8_t *out)
{
    AXS_CHECK_ARG(mask && bdim && out, "axs_amr_erode: NULL argument");
    AXS_CHECK_ARG(r >= 0 && bdim[0] > 0 && bdim[1] > 0 && bdim[2] > 0,
                  "axs_amr_erode: bad size");
    const std::vector<uint8_t> e = erode(mask, bdim, r);
    std::memcpy(out, e.data(), e.size());
    return AXS_OK;
}

int axs_amr_boxes_from_blocks(const uint8_t *blk, const int bdim[3],
                              int fine_block, int max_blocks, int max_boxes,
                              int *nboxes, int *lo, int *hi)
{
    AXS_CHECK_ARG(blk && bdim && nboxes, "axs_amr_boxes_from_blocks: NULL");
    AXS_CHECK_ARG(fine_block >= 1 && max_blocks >= 1 && bdim[0] > 0 &&
                      bdim[1] > 0 && bdim[2] > 0,
                  "axs_amr_boxes_from_blocks: bad size");
    const std::vector<BoxSpec> b =
        boxes_from_blocks(blk, bdim, fine_block, max_blocks);
    *nboxes = (int)b.size();
    if ((int)b.size() > max_boxes || !lo || !hi)
        return (int)b.size() > max_boxes
                   ? set_error(AXS_ERR_ARG,
                               "axs_amr_boxes_from_blocks: more than "
                               "max_boxes boxes")
                   : AXS_OK;
    for (size_t i = 0; i < b.size(); ++i)
        for (int d = 0; d < 3; ++d)
        {
            lo[3 * i + d] = b[i].lo[d];
            hi[3 * i + d] = b[i].hi[d];
        }
    return AXS_OK;
}

} // extern "C"
