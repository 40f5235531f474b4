// axs_internal.hpp -- internal types shared by the HIP translation units of
// libaxs_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/axs.h"

namespace axs
{

// Records the message returned by axs_last_error() and yields the code.
int set_error(int code, const std::string &msg);

#define AXS_HIP_TRY(expr)                                                     \
    do                                                                        \
    {                                                                         \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess)                                                 \
            return ::axs::set_error(AXS_ERR_HIP,                              \
                                    std::string(#expr) + ": " +               \
                                        hipGetErrorString(e_));               \
    } while (0)

#define AXS_CHECK_ARG(cond, msg)                                              \
    do                                                                        \
    {                                                                         \
        if (!(cond))                                                          \
            return ::axs::set_error(AXS_ERR_ARG, (msg));                      \
    } while (0)

// Maximum RK4 steps whose stage scalars one table upload holds.
constexpr int kCoeffTableSteps = 4096;
// Maximum partial-sum slots (blocks) a reduction pass may produce.
constexpr int kMaxPartials = 65536;
// Offset step between the state arrays of a level (see axs_level_create).
constexpr size_t kArrayStagger = (size_t(1) << 20) + (size_t(36) << 10) + 768;

// Device allocation for box-level arrays: the smallest pooled block of at
// least `bytes` (and at most twice that), else hipMalloc of `bytes` plus an
// eighth (room for the next level to grow into), releasing the pool and
// retrying once if that fails.  *got receives the block's size.
void *pool_alloc(axs_ctx *c, size_t bytes, size_t *got);
// Keeps a block of >= 64 MiB for reuse (at most 12 blocks), else frees it.
void pool_release(axs_ctx *c, void *p, size_t bytes);
// Frees every pooled block.
void pool_trim(axs_ctx *c);

// Launch shapes from the tuning table (axs_tuning_load).  tuned_level
// fills the four values for a level of this shape, precision and stage
// fusion mode and returns true when the table has them; tuned_box likewise
// for box levels of a precision.
bool tuned_level(int nx, int ny, int nz, int precision, int fused,
                 int *tile_variant, int *tile_z_chunk, int *pair_variant,
                 int *pair_z_chunk);
bool tuned_box(int precision, int *item_rows, int *grid_blocks);

struct Grid
{
    int nx, ny, nz;
    uint32_t ncell; // nx*ny*nz, < 2^31 (checked at creation)
};

} // namespace axs

struct axs_ctx
{
    int device;
    hipStream_t stream;
    // small device scratch: counters and reduction partials
    unsigned long long *d_counts; // [8]
    double *d_partials;           // [kMaxPartials * 16]
    double *d_result;             // [64]
    double *h_result;             // pinned [64]
    unsigned long long *h_counts; // pinned [8]
    // large device arrays of destroyed box levels, kept for the next box
    // level (regrids create a level while the one it replaces still lives)
    std::vector<std::pair<void *, size_t>> pool;
};

struct axs_level
{
    axs_ctx *ctx;
    axs::Grid g;
    double dx;
    int precision;
    size_t word;        // bytes per stored word (8 or 4)
    void *y;            // y_n: 4 comps (the current state)
    float *comp;        // FP32C: low part of y_n (4 comps), else nullptr
    void *acc;          // RK accumulator of the per-stage launches (4
                        // comps); the fused steps write y_{n+1} here and
                        // then swap it with y
    void *ya;           // stage state A (4 comps)
    void *yb;           // stage state B (4 comps)
    void *alloc[4];     // allocations behind y, acc, ya|yb, comp
    double *d_coeff;    // [kCoeffTableSteps][4][3]
    int *d_step;        // device step counter (graph replays)
    int tile_variant;   // stage-kernel tile variant
    int z_chunk;        // z planes per block
    int fused;          // RK4 stages: 0 per stage, 1 fused pairs, 2 pairs
                        // without a stored accumulator, 3 one launch
                        // per step (FP32 / FP32C)
    int pair_variant;   // fused-pair kernel tile variant (-1 = default)
    int pair_z_chunk;   // z planes per block of the pair kernels (0 = auto)
    // captured step graphs, one per state buffer the replay starts from
    // (fused steps alternate y between two arrays)
    hipGraph_t graph[2];
    hipGraphExec_t graph_exec[2];
    int graph_valid[2];
    double graph_key[2][2]; // dt and steps per replay at capture
    const void *graph_y[2]; // the state buffer at capture
    int graph_next;         // slot the next capture replaces
    // windings / tagger cadence of axs_rk4_advance_windings: the device
    // segment counter, counts (3 per segment) and R per segment, with room
    // for diag_cap segments; its graph of one or two segments
    int *d_seg;
    unsigned long long *d_dcounts;
    double *d_dR;
    int diag_cap;
    hipGraph_t dgraph;
    hipGraphExec_t dgraph_exec;
    int dgraph_valid;
    double dgraph_key[6]; // dt, every, segments per replay, the tagger's
                          // three thresholds
    const void *dgraph_y;    // the state buffer at capture
    const void *dgraph_tags; // the tags buffer at capture
    // FFT plans (lazily created)
    void *fft_r2c;  // hipfftHandle stored as void*
    void *fft_c2r;
    size_t fft_work_bytes;
    // stage timing
    int timing;
    hipEvent_t ev[5];
    double stage_ms[4];
    long long timed_steps;
};

namespace axs
{
// Launchers implemented in the .hip files.
int launch_rk_stage(axs_level *L, int stage, const void *yin, void *yout,
                    const double *coeff_table, const int *step_ptr,
                    int coeff_stage, double dt);
int launch_rk_pair(axs_level *L, int pair, const void *yin, const void *yn,
                   void *yout, void *acc, const double *coeff_table,
                   const int *step_ptr, double dt);
int num_pair_variants();
int launch_rk_step(axs_level *L, const void *yin, void *yout,
                   const double *coeff_table, const int *step_ptr, double dt);
int launch_rhs_eval(axs_level *L, const double *coeff3_dev, void *rhs);
int launch_step_advance(axs_level *L, int reset);
int num_tile_variants();
// cadence passes (axs_diag.hip): tags cleared, windings / tagger counts
// added at counts + 3 * (*seg) with R_tab[*seg] for the amplitude
// criterion, then *seg += 1; launch_seg_reset sets *seg = 0
int launch_windings_pass(axs_level *L, const axs_tagger_params *tp,
                         uint8_t *tags, unsigned long long *counts, int *seg,
                         const double *R_tab);
int launch_seg_reset(axs_level *L, int *seg);
} // namespace axs
