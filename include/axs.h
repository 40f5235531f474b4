/*
 * axs.h -- C ABI of libaxs_hip.so, the MI355X (gfx950) implementation of the
 * axion-string simulator's per-step hot path.
 *
 * Every entry point returns an int status (AXS_OK = 0) and never throws.
 * The text of the last error is available from axs_last_error().  Device
 * memory is allocated by the *_create calls, by axs_level_spectrum and
 * axs_level_fourier_ic on their first use (the FFT plans), by
 * axs_level_place (trial placements of the level's arrays), by
 * axs_adjoint_grad_energy and axs_forward_grad_energy (their checkpoint and
 * adjoint arrays, per call), and by box levels (their arrays, from a
 * context pool); axs_energy_sums_noalloc allocates nothing.
 *
 * Data model.  A level (axs_level) is one uniform, periodic grid of
 * nx*ny*nz cells with spacing dx that owns its state in HBM, structure of
 * arrays, component-major, x fastest:
 *     state[c][z][y][x],  c in {psi1, psi2, Pi1, Pi2}
 * (the reference's StateVariables.hpp:9-17 component order).  Periodic
 * ghost cells are realised by index wrapping inside the kernels, which is
 * what the reference's FillBoundary(periodicity) provides
 * (AxionStringsLevel.cpp:820,1255).  Precision modes:
 *   AXS_FP64  : double storage and arithmetic;
 *   AXS_FP32C : float storage, float stage arithmetic, and a float
 *               compensation array carrying the low part of y_n, updated
 *               by Kahan summation at each RK4 step (state = hi + lo);
 *   AXS_FP32  : plain float (the reference's PRECISION=FLOAT build).
 * A box level (axs_boxlevel) is a list of boxes of one AMR level, each with
 * its own 3 ghost layers, advanced by one persistent launch per RK stage.
 *
 * All calls on one context are ordered on that context's HIP stream.
 * Calls that return host values (counts, sums, spectra) synchronise that
 * stream.  A context is not thread-safe.
 */
#ifndef AXS_H_
#define AXS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------ status ------------------------------ */
#define AXS_OK 0
#define AXS_ERR_ARG 1         /* invalid argument / shape / precision */
#define AXS_ERR_HIP 2         /* HIP runtime or hipFFT failure */
#define AXS_ERR_NOMEM 3       /* device allocation failed */
#define AXS_ERR_UNSUPPORTED 4 /* valid request this build does not serve */

#define AXS_FP64 0
#define AXS_FP32C 1
#define AXS_FP32 2

/* Masking schemes, in the reference's enum order (Masking.hpp:39-44). */
#define AXS_MASK_NONE 0
#define AXS_MASK_A 1
#define AXS_MASK_B 2

/* Field kinds written into the spectrum buffer. */
#define AXS_FIELD_MASKED_A_DOT 0 /* fill_masked_a_dot_buffer (reference) */
#define AXS_FIELD_PI1 1          /* Pi1, for the |Pi|^2 spectrum */
#define AXS_FIELD_PI2 2          /* Pi2 */

typedef struct axs_ctx axs_ctx;
typedef struct axs_level axs_level;
typedef struct axs_boxlevel axs_boxlevel;

/* Background of Background.hpp:21-58 (CTauSchedule + a_inv), plus the
 * flat-space background of FlatBackground.hpp:26-44. */
typedef struct
{
    double a_inv;
    double c0;
    double c1;
    double tau_switch;
    int has_switch;
    int flat_space;  /* 1: R = 1, curvature 0, lambda = m_r^2 */
    double flat_m_r; /* FlatBackground m_r (flat_space only) */
} axs_background;

/* The three scalars of AxionStringsRHS's constructor
 * (AxionStringsRHS.hpp:32-37): curvature_coeff, lambda, R_squared. */
typedef struct
{
    double curvature_coeff;
    double lambda;
    double R_squared;
} axs_rhs_coeffs;

/* StringTaggerParams (StringTagger.hpp:83-111) plus one extension: an
 * amplitude criterion |psi|/R < amplitude_threshold (0 = off, the default,
 * which reproduces the reference tagger exactly). */
typedef struct
{
    double gradient_threshold;        /* DBL_MAX = off (reference default) */
    double radial_gradient_threshold; /* DBL_MAX = off (reference default) */
    double amplitude_threshold;       /* extension; 0 = off */
    double R;                         /* R(tau), used by the extension */
} axs_tagger_params;

/* MaskingParams (Masking.hpp:46-54). */
typedef struct
{
    int scheme;       /* AXS_MASK_* */
    double threshold; /* scheme B top-hat on |psi|/R */
} axs_masking;

/* ------------------------- context / errors ------------------------- */
/* Creates a context on HIP device `device` with its own stream. */
int axs_ctx_create(int device, axs_ctx **out);
int axs_ctx_destroy(axs_ctx *ctx);
int axs_ctx_synchronize(axs_ctx *ctx);
/* Frees the device arrays the context keeps from destroyed box levels for
 * reuse by the next box level (a regrid creates the new level while the one
 * it replaces still exists).  No counterpart in the reference (AMReX's arena
 * plays this role there); destroying the context frees them too. */
int axs_ctx_trim(axs_ctx *ctx);
/* The HIP stream the context launches on (a hipStream_t). */
void *axs_ctx_stream(axs_ctx *ctx);
/* Device memory for callers built without HIP (the C++ driver's tag and
 * coverage buffers, FFI bindings); no counterpart in the reference (AMReX's
 * arenas play this role there).  alloc: uninitialised bytes on the
 * context's device; free waits for the context's stream first; memset is
 * ordered on the stream; download copies after the stream's earlier work
 * and returns when the bytes are on the host. */
int axs_device_alloc(axs_ctx *ctx, size_t bytes, void **out);
int axs_device_free(axs_ctx *ctx, void *ptr);
int axs_device_memset(axs_ctx *ctx, void *ptr, int value, size_t bytes);
int axs_device_download(axs_ctx *ctx, void *host, const void *dev,
                        size_t bytes);
const char *axs_last_error(void);
/* Number of compiled stage-kernel tile variants (for the autotuner). */
int axs_num_tile_variants(void);
/* Loads a tuning table (the autotuner's output, a text file) that replaces
 * the one loaded before.  No counterpart in the reference (which picks
 * launch shapes inside AMReX's ParallelFor).  Lines, '#' to end of line a
 * comment:
 *   level <nx> <ny> <nz> <precision> <fusion> <tile_variant> <tile_z_chunk>
 *         <pair_variant> <pair_z_chunk>
 *   box <precision> <item_rows> <grid_blocks>
 * A level created afterwards, and a level switched to another stage fusion
 * mode, takes the tile, pair variant and z chunks of the `level` line of
 * its shape, precision and fusion mode; a box level created afterwards
 * takes the work shape of the `box` line of its precision.  Levels without
 * a line keep the built-in defaults.  A malformed line fails the call with
 * AXS_ERR_ARG naming it and leaves the table as it was. */
int axs_tuning_load(const char *path);

/* ------------------------- host background -------------------------- */
/* Background.hpp:53-133.  which: 0 R, 1 R', 2 curvature_term_coeff,
 * 3 lambda, 4 H, 5 H_over_mr_direct, 6 H_over_mr_closed_form,
 * 7 tau_from_log_mr_over_h (x is log(m_r/H)), 8 R'/R. */
int axs_background_eval(const axs_background *bg, int which, double x,
                        double *out);
/* The AxionStringsRHS scalars at conformal time tau, as
 * AxionStringsLevel::specific_eval_rhs computes them
 * (AxionStringsLevel.cpp:859-885, evolving phase or flat space). */
int axs_background_rhs_coeffs(const axs_background *bg, double tau,
                              axs_rhs_coeffs *out);

/* ------------------------------ levels ------------------------------ */
int axs_level_create(axs_ctx *ctx, int nx, int ny, int nz, double dx,
                     int precision, axs_level **out);
int axs_level_destroy(axs_level *lvl);
/* Copies a host double SoA state (4*nx*ny*nz) in / out.  FP32C stores the
 * value as hi + lo (lo in the compensation array); download returns
 * hi + lo. */
int axs_level_upload(axs_level *lvl, const double *host_state);
int axs_level_download(axs_level *lvl, double *host_state);
/* Download of the raw stored words: FP64 -> 4n doubles; FP32/FP32C -> 4n
 * floats of hi (and, for FP32C, 4n floats of lo into `lo` if non-NULL). */
int axs_level_download_raw(axs_level *lvl, void *hi, void *lo);
/* The inverse of axs_level_download_raw (exact restore of a checkpoint,
 * the role of AmrLevel::restart for the state data): FP32C needs `lo`. */
int axs_level_upload_raw(axs_level *lvl, const void *hi, const void *lo);
/* Device pointers of the state (y_n) and compensation arrays.  With stage
 * fusion on, every RK4 step writes y_{n+1} into a second array of the level
 * and the level's state then names that array: the state pointer holds
 * until the next step call. */
int axs_level_device_ptrs(axs_level *lvl, void **state, void **comp);
/* Homogeneous exact solution psi = (R, 0), Pi = (R', 0)
 * (AxionStringsLevel.cpp:824-847). */
int axs_level_set_homogeneous(axs_level *lvl, double R, double R_prime);
/* The pre-evolution -> main handoff rescale of the state in place
 * (AxionStringsLevel::apply_pre_evolution_to_main_rescale,
 * AxionStringsLevel.cpp:2027-2125): psi <- kappa psi and
 * Pi <- kappa^2 Pi + (psi / R_pre) common_term, with kappa = R_main / R_pre
 * and common_term = R_main' - kappa^2 R_pre' computed by the caller. */
int axs_level_handoff_rescale(axs_level *lvl, double kappa,
                              double common_term, double R_pre);
/* Selects the stage-kernel tile variant (autotuner output); -1 = default
 * (the 1024^3 autotuner choice: 8-row tiles, 4 cells per lane in FP32 and
 * 2 in FP64, centre loads one plane ahead; z-chunk 32, halved on small
 * grids until there are 4096 tiles); z_chunk 0 = default. */
int axs_level_set_tile(axs_level *lvl, int variant, int z_chunk);
int axs_level_get_tile(axs_level *lvl, int *variant, int *z_chunk);
/* Stage fusion; no counterpart in the reference (how the RK4 step of
 * AxionStringsLevel.cpp:850-921 is launched).  on = 1 (FP64's default):
 * each RK4 step is two launches, stages 1+2 and stages 3+4, which form the
 * second stage's psi at every stencil point from the first stage's inputs
 * (psi_n + (dt/2) Pi_n, psi_n + dt Pi_2) and so never store y_1 or y_3:
 * 9 S instead of 19 S (FP32C) and 7 S instead of 17 S (FP64, FP32) of HBM
 * traffic per cell, S = 4 words.  on = 2: the pairs without a stored
 * accumulator -- the stages 3+4 launch reads Pi_n on the stencil too and
 * recomputes acc = k1 + 2 k2 at each cell from the stars of psi_n and
 * psi_1 = psi_n + (dt/2) Pi_n: 7 S (FP32C), 5 S (FP64, FP32) per cell, for
 * twice the RHS work in that launch.  on = 3 (FP32 / FP32C only, their
 * default): the whole step in one launch, which forms every stage input
 * from y_n on the fly -- the stage-1 and stage-2 RHS on the tile grown by
 * 2 cells from an LDS plane of psi_n grown by 4 -- and stores only
 * y_{n+1}: 4 S (FP32C), 2 S (FP32) per cell.  on = 0: one launch per
 * stage.  Every mode gives the same bits.  The fused modes write y_{n+1}
 * into a second array of the level, which then becomes the state (see
 * axs_level_device_ptrs); every step, also a single one, runs fused. */
int axs_level_set_stage_fusion(axs_level *lvl, int on);
/* Tile variant of the fused-pair kernels (-1 = default, the 1024^3 sweep's
 * fastest: 8-row tiles of 2 cells per lane with z chunk 32 in FP32 (of 1
 * cell per lane for the pairs without an accumulator), 4-row tiles of 2
 * cells per lane with z chunk 64 in FP64; for the one launch per step the
 * tile: -1 and 3 = 64 x 11 cells with rows aligned to cache lines, z chunk
 * 128; 2 = 60 x 12; 0 = 60 x 8; 1 = 60 x 4) and their z chunk
 * (0 = default; on small grids halved until there are 4096 tiles);
 * axs_num_pair_variants() variants. */
int axs_num_pair_variants(void);
int axs_level_set_pair_tile(axs_level *lvl, int variant, int z_chunk);
/* The level's stage fusion mode, pair variant and pair z chunk (any
 * pointer may be NULL). */
int axs_level_get_launch(axs_level *lvl, int *fusion, int *pair_variant,
                         int *pair_z_chunk);

/* ------------------- RHS and RK4 (the hot path) --------------------- */
/* AxionStringsRHS<>::operator() over every cell of a periodic level
 * (AxionStringsRHS.impl.hpp:25-60 launched as in
 * AxionStringsLevel.cpp:909-918): rhs = (Pi1, Pi2, lap psi1 + c psi1 -
 * (lambda/2)(|psi|^2 - R^2) psi1, ...).  `rhs` is a device array of
 * 4*nx*ny*nz words of the level's precision (float for FP32C/FP32). */
int axs_rhs_eval(axs_level *lvl, const axs_rhs_coeffs *c, void *rhs);
/* One classic RK4 step with explicit stage scalars (stage s at tau,
 * tau+dt/2, tau+dt/2, tau+dt), launched in the level's stage-fusion mode
 * (FP32 / FP32C: one launch; FP64: two).  This replaces GRTeclyn's RK4
 * stage driver around specific_eval_rhs (AxionStringsLevel.cpp:850-921). */
int axs_rk4_step(axs_level *lvl, const axs_rhs_coeffs coeffs[4], double dt);
/* nsteps RK4 steps from conformal time tau with the background's stage
 * scalars (host FP64), uploaded once as a table.  use_graph = 1 replays a
 * captured hipGraph of two steps (fused modes; one graph per state array a
 * replay starts from) or one step (per-stage launches), each followed by
 * the step-counter advance; an odd last step runs as a direct launch. */
int axs_rk4_advance(axs_level *lvl, const axs_background *bg, double tau,
                    double dt, int nsteps, int use_graph);

/* nsteps RK4 steps with caller-supplied stage scalars: coeffs[4*i + s] are
 * the scalars of stage s of step i.  Covers the stage-scalar selections of
 * specific_eval_rhs (AxionStringsLevel.cpp:850-911) that are not a
 * Background: the pre-evolution phase (PreEvolutionBackground.hpp, :869-876)
 * and the post-handoff smoothstep blend of lambda and the curvature
 * coefficient (:889-907).  Same launches as axs_rk4_advance. */
int axs_rk4_advance_coeffs(axs_level *lvl, const axs_rhs_coeffs *coeffs,
                           double dt, int nsteps, int use_graph);
/* nsteps RK4 steps as axs_rk4_advance (nsteps a multiple of `every`, at
 * most 4096), each run of `every` steps followed by the plaquette windings
 * pass of axs_count_plaquettes -- and, when tags != NULL, by the
 * StringTagger pass of axs_tag_cells into tags (cleared first), with R(tau)
 * of that point for the amplitude criterion (tp->R is not used; tp NULL =
 * the reference tagger) -- the post-step sequence of the reference
 * (AxionStringsLevel.cpp:1316 counts, :1999-2003 tags) at a step cadence.
 * counts[2 j], counts[2 j + 1]: plain and weighted pierced plaquettes after
 * segment j (nsteps / every segments); tags hold the last segment's tags.
 * With use_graph, one graph replays a segment -- its steps, the tag clear
 * and the windings / tagger pass (two segments when `every` is odd and the
 * steps are fused, so that a replay ends in the state buffer it starts
 * from) -- captured on first use; the counts are read back once, at the
 * end.  The same bits as the direct launches. */
int axs_rk4_advance_windings(axs_level *lvl, const axs_background *bg,
                             double tau, double dt, int nsteps, int every,
                             const axs_tagger_params *tp, uint8_t *tags,
                             int64_t *counts, int use_graph);

/* Placement trials; no counterpart in the reference (an MI355X-specific
 * set-up step).  The physical pages a level's arrays land on decide between
 * two HBM bandwidth regimes of the stage kernels (1024^3 FP32C: about 57 or
 * 65 ms per RK4 step with per-stage launches, 30.1 to 33.8 ms with fused
 * pairs, for the same virtual layout).  Tries up to `trials`
 * placements (1..16) of the level's arrays, timing two RK4 steps on each.
 * While free memory holds a second copy of the arrays the fastest placement
 * seen is kept; otherwise each placement is freed before the next is drawn
 * and, from the third draw on, the first within 2% of the fastest seen is
 * kept (else the last).  Either way the trials stop as soon as the kept
 * placement streams the RK4 step's algorithmic bytes at the fast regime's
 * rate: per-stage launches 304 B per cell (FP32C), 544 (FP64), 272 (FP32)
 * at >= 69% of the 8 TB/s HBM peak; fused pairs 144, 224, 112 B at >= 63%
 * (FP64: 54%); pairs without a stored accumulator 112, 160, 80 B at >= 53%;
 * one launch per step 64 (FP32C), 32 (FP32) B at >= 42%.  The state is reset to zero: call it before
 * uploading data.  ms_per_step (optional, `trials` entries) receives each
 * trial's time and *used (optional) the number of placements tried. */
int axs_level_place(axs_level *lvl, int trials, double *ms_per_step, int *used);

/* Per-stage kernel timing with HIP events recorded on the context stream
 * around each stage launch of axs_rk4_advance (forces direct launches while
 * on).  ms[s] accumulates the milliseconds of stage s+1; with fused pairs
 * (stage fusion 1, 2), ms[0] and ms[1] accumulate those of the stages 1+2
 * and 3+4 launches and ms[2], ms[3] stay 0; with one launch per step (3),
 * ms[0] accumulates the step launch.  *nsteps counts the steps timed; axs_level_stage_times resets the
 * accumulators. */
int axs_level_enable_timing(axs_level *lvl, int on);
int axs_level_stage_times(axs_level *lvl, double ms[4], int64_t *nsteps);

/* ------------------------ strings / tagging ------------------------- */
/* count_plaquettes (StringFinder.hpp:76-122): out[0] = n_p_plain,
 * out[1] = n_p_weighted.  mask: optional device uint8 per cell (0 skips). */
int axs_count_plaquettes(axs_level *lvl, const uint8_t *mask, int64_t out[2]);
/* compute_velocity_at_pierced_corners (VelocityKernel.hpp:37-141):
 * gamma^2 v^2 (Velocity.hpp:62-78) summed over the 4 corners of every
 * pierced plaquette (*sum_gamma_sq_v_sq) and the number of corner
 * evaluations (*count), in the same pass as the plaquette counts
 * (counts, optional).  mask: optional device uint8 per cell. */
int axs_velocity_at_pierced_corners(axs_level *lvl, double R,
                                    double R_prime_over_R, double m_r,
                                    const uint8_t *mask,
                                    double *sum_gamma_sq_v_sq, int64_t *count,
                                    int64_t counts[2]);
/* plaquette_windings_at for every cell (StringFinder.hpp:45-68), into a
 * device int8 array w[cell*3 + {xy, yz, zx}]. */
int axs_windings_field(axs_level *lvl, int8_t *w);
/* StringTagger::operator() for every cell (StringTagger.hpp:240-274):
 * sets tags[cell] = 1 where a criterion holds, never clears.  tags is a
 * device uint8 array.  counts (optional, host) receives the plaquette
 * counts of the same pass. */
int axs_tag_cells(axs_level *lvl, const axs_tagger_params *p, uint8_t *tags,
                  int64_t counts[2]);

/* ----------------------------- energies ----------------------------- */
/* compute_total_energy_sums (EnergyKernel.hpp:106-233): the 14 raw sums in
 * TotalEnergySums order. mask: optional device uint8 per cell. */
int axs_energy_sums(axs_level *lvl, double R, double lambda,
                    double R_prime_over_R, const axs_masking *m,
                    const uint8_t *mask, double out[14]);

/* compute_total_energy_sums over caller-owned device state, allocating
 * nothing.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid it).  Same kernels, sums and order as
 * axs_energy_sums, hence the same bits: state is 4*nx*ny*nz words of the
 * precision's type in the level layout, comp the FP32C low words (NULL
 * otherwise), mask an optional device uint8 per cell.  The 14 sums are
 * written to out_dev (device, 14 doubles) on `stream` (a hipStream_t; NULL
 * = the default stream); the call returns without synchronising.  The
 * reduction partials live in a fixed scratch array of the library (3.7 MB
 * per GPU, held until the process ends): grids needing more than 32768
 * energy blocks (beyond 1024^3) are refused with AXS_ERR_ARG before any GPU
 * work.  Calls that overlap on one GPU share that array and race: order
 * them on one stream. */
int axs_energy_sums_noalloc(const void *state, const float *comp, int nx,
                            int ny, int nz, int precision, double dx, double R,
                            double lambda, double R_prime_over_R,
                            const axs_masking *m, const uint8_t *mask,
                            double *out_dev, void *stream);

/* ------------------------- spectrum / masking ----------------------- */
/* fill_masked_a_dot_buffer (MaskedFieldBuffer.hpp:15-35) and the Pi1/Pi2
 * fill variants, into a device array of nx*ny*nz words of the level's
 * arithmetic type (double for FP64, float otherwise). */
int axs_fill_field(axs_level *lvl, int kind, const axs_masking *m, double R,
                   void *buf);
/* compute_spectrum (SpectrumKernel.hpp:52-153) of the level's field `kind`
 * (filled internally into the level's stage scratch, R2C on the device,
 * then shell binning on the device).  Accumulates (+=) into
 * shell_sum[n_shells], shell_count[n_shells], tot[2] = {full_cube,
 * inscribed_sphere}; n_shells = max(nx,ny,nz)/2 + 1.  Also returns the
 * real-space sum of squares of the filled buffer in *sum_sq if non-NULL. */
int axs_level_spectrum(axs_level *lvl, int kind, const axs_masking *m,
                       double R, double *shell_sum, int64_t *shell_count,
                       double tot[2], double *sum_sq);

/* ---------------------------- initial data -------------------------- */
/* FourierIC::generate (FourierIC.hpp:237-285) into component `comp`
 * (0 psi1 / 1 psi2) of the level, on the device: hash-seeded spectral
 * fill, Hermitian projection of the self-conjugate planes, C2R, real-space
 * rescale to target_mean_square.  Uses the stage scratch as FFT memory. */
int axs_level_fourier_ic(axs_level *lvl, int comp, double k_max_cells,
                         double target_mean_square, uint64_t seed,
                         int comp_seed_offset);
/* Sets components c0..c0+ncomp-1 to zero (the reference's
 * state_new.setVal(0.0, c_Pi1, 2, 0), AxionStringsLevel.cpp:818). */
int axs_level_zero_components(axs_level *lvl, int c0, int ncomp);

/* --------------------- comparisons (FP32C vs FP64) ------------------ */
/* max over cells of |psi_a - psi_b| (complex modulus of the difference)
 * and the max of |psi_b|, for two levels of equal shape. */
int axs_level_compare(axs_level *a, axs_level *b, double *max_dpsi,
                      double *max_psi_b);

/* -------------------------- AMR box levels -------------------------- */
/* A refinement level made of nboxes boxes, box b covering cells
 * lo[3b..3b+2] to hi[3b..3b+2] (inclusive) of a periodic index space of
 * size domain[0..2] (the level's own index space; ratio 2 to the next
 * coarser level), each box stored with nghost (3) ghost layers.  Box
 * corners and the domain must be multiples of 8 cells (the blocking
 * factor), boxes must not overlap, and box edges must be 8..64 cells
 * (the reference's blocking_factor / max_grid_size range).  This is the
 * MultiFab of one AMR level of the reference (AxionStringsLevel's state,
 * boxes from amr.blocking_factor / amr.max_grid_size). */
int axs_boxlevel_create(axs_ctx *ctx, int nboxes, const int *lo,
                        const int *hi, const int domain[3], int nghost,
                        double dx, int precision, axs_boxlevel **out);
int axs_boxlevel_destroy(axs_boxlevel *bl);
/* Number of valid cells (without ghosts) over all boxes. */
int64_t axs_boxlevel_num_cells(axs_boxlevel *bl);
/* Work shape of the box level's RK4 stage launches; no counterpart in the
 * reference (AMReX's ParallelFor picks it there).  Every box is cut into
 * items of whole box rows, H = clamp(item_rows, 4, 512 / box x edge) rows
 * each (512-thread blocks), and a grid of min(items, grid_blocks) blocks
 * strides over them.
 * Defaults: item_rows 512 (the widest items), grid_blocks 2048, or the
 * tuning table's `box` line.  Every shape gives the same bits.  Setting a
 * new shape waits for the level's stream, drops its captured graphs and
 * may reallocate the item list. */
int axs_boxlevel_set_work_shape(axs_boxlevel *bl, int item_rows,
                                int grid_blocks);
/* Current item_rows, grid_blocks and number of items (NULL skips). */
int axs_boxlevel_get_work_shape(const axs_boxlevel *bl, int *item_rows,
                                int *grid_blocks, int *nitems);
/* Launches of the box level's RK4 step; no counterpart in the reference
 * (GRTeclyn runs one RK stage per pass over the boxes).  mode 0: one
 * persistent launch per stage, each after the ghost fill of its input;
 * mode 1: two launches, stages 1+2 and 3+4, whose second stage forms its
 * psi (psi_n + (dt/2) Pi_n, psi_n + dt Pi_2) on the stencil inside the
 * boxes and reads it at the ghost cells from a fill that writes what the
 * per-stage fill of that stage's input writes (the owner's value, or the
 * coarse interpolation at the stage time) -- the same bits, 144 instead of
 * 304 bytes per FP32C cell and step.  After a mode-1 step the state lives
 * in the other of two arrays; every entry point follows it.  Default 1. */
int axs_boxlevel_set_stage_fusion(axs_boxlevel *bl, int mode);
int axs_boxlevel_get_stage_fusion(const axs_boxlevel *bl, int *mode);
/* Copies values from / to a dense periodic host array of the whole domain
 * (4 * domain cells, doubles, x fastest); cells outside every box are not
 * read / not written. */
int axs_boxlevel_upload_dense(axs_boxlevel *bl, const double *dense);
int axs_boxlevel_download_dense(axs_boxlevel *bl, double *dense);
/* Copies every box's valid cells from a uniform level of the same domain
 * and precision (level-0 initial data). */
int axs_boxlevel_copy_from_level(axs_boxlevel *bl, const axs_level *src);
/* Snapshot of the coarse level's current state, the t = 0 end of the time
 * interpolation used by fill_ghosts / rk4_step (take it before the coarse
 * level advances). */
int axs_boxlevel_snapshot_coarse(axs_boxlevel *bl, const axs_boxlevel *coarse);
/* FillPatch of the state: each box's ghost cells from the valid cells of
 * the level's boxes (periodic) where they exist, else from `coarse` (ratio
 * 2) by tensor-product quartic (5-point Lagrange) interpolation of the
 * coarse state interpolated linearly in time between the snapshot
 * (t_frac = 0) and the coarse level's current state (t_frac = 1).
 * coarse may be NULL when the boxes tile the whole periodic domain.
 * interior = 1 also interpolates the valid cells from the coarse level
 * (initial data of a new level). */
int axs_boxlevel_fill_ghosts(axs_boxlevel *bl, const axs_boxlevel *coarse,
                             double t_frac, int interior);
/* One RK4 step of every box, each stage one persistent launch over the
 * box work-list, preceded by the ghost fill of its input at the stage time
 * (t0_frac, t0_frac + dt_frac/2, t0_frac + dt_frac/2, t0_frac + dt_frac of
 * the coarse interval; a ratio-2 subcycled fine step has dt_frac = 0.5).
 * These fills write what the stage stencil reads, psi in the 2 ghost layers
 * next to each box; the state's other ghost values are left as they were
 * (axs_boxlevel_fill_ghosts fills all of them, as tagging and the energy
 * sums need). */
int axs_boxlevel_rk4_step(axs_boxlevel *bl, const axs_rhs_coeffs coeffs[4],
                          double dt, const axs_boxlevel *coarse,
                          double t0_frac, double dt_frac);
/* The same step replayed from a hipGraph captured on first use for this
 * (dt, coarse, t0_frac, dt_frac): the 4 ghost fills and 4 work-list stage
 * launches are one graph launch; the stage scalars are copied to the device
 * before it, and cells without an owning box are reported after it, as by
 * axs_boxlevel_rk4_step (AXS_ERR_ARG). */
int axs_boxlevel_rk4_step_graph(axs_boxlevel *bl,
                                const axs_rhs_coeffs coeffs[4], double dt,
                                const axs_boxlevel *coarse, double t0_frac,
                                double dt_frac);
/* StringTagger over every valid box cell into a dense device uint8 array
 * of the level's domain (sets only), all criteria of axs_tag_cells
 * (windings, Laplacian, radial gradient, amplitude); plaquette counts of
 * the same pass.  Ghosts must be filled.  cover: optional dense device
 * uint8 coverage mask (0 = covered by a finer level, excluded from the
 * counts and the tags). */
/* dense_tags may be NULL: the plaquette counts only. */
int axs_boxlevel_tag(axs_boxlevel *bl, const axs_tagger_params *p,
                     uint8_t *dense_tags, const uint8_t *cover,
                     int64_t counts[2]);
/* compute_total_energy_sums over the level's valid cells, excluding cells
 * where the dense coverage mask is 0 (EnergyKernel.hpp:106-233). */
int axs_boxlevel_energy_sums(axs_boxlevel *bl, double R, double lambda,
                             double R_prime_over_R, const axs_masking *m,
                             const uint8_t *cover, double out[14]);
/* Dense uint8 coverage mask of the coarse domain (ratio 2): 0 where a
 * coarse cell is covered by a box of `fine`, 1 elsewhere
 * (amrex::makeFineMask, AxionStringsLevel.cpp:33-85). */
int axs_boxlevel_coverage_mask(const axs_boxlevel *fine,
                               const int coarse_domain[3], uint8_t *mask);
/* average_down of the fine level's valid cells onto the covered valid
 * cells of the coarse box level (mean of the 8 children;
 * collapse_to_level0, AxionStringsLevel.cpp:110-134). */
int axs_boxlevel_average_down(const axs_boxlevel *fine, axs_boxlevel *coarse);
/* Copies every box's valid cells into a uniform level of the same domain
 * and precision (cells outside the boxes are left as they are); with the
 * spectrum entry points this gives the level-0 FFT of the reference. */
int axs_boxlevel_copy_to_level(const axs_boxlevel *bl, axs_level *dst);
/* Regridding: copies the valid cells of `dst` that a valid box of `src`
 * (same domain, precision and context) also covers, leaving the others
 * as they are (AMReX's FillPatch from the old level inside AmrLevel::init
 * after a regrid, AxionStringsLevel.cpp:1916-2005 tags the new grids).
 * *copied (optional) receives the number of cells copied. */
int axs_boxlevel_copy_valid(axs_boxlevel *dst, const axs_boxlevel *src,
                            int64_t *copied);

/* ---------------------- borrowed box data (Array4) ------------------- */
/* A view of one box's data that the caller owns in device memory, as
 * AMReX's Array4 (AMReX_Array4.H): element (i, j, k, n) at
 * p[(i - begin[0]) + (j - begin[1]) jstride + (k - begin[2]) kstride +
 *   n nstride] for begin <= (i, j, k) < end, 0 <= n < ncomp; strides in
 * elements.  With these entry points the kernels run inside the reference's
 * per-box loops on its own MultiFab data, with no copy in or out. */
typedef struct
{
    void *p;
    int64_t jstride, kstride, nstride;
    int begin[3];
    int end[3];
    int ncomp;
} axs_array4;
/* AxionStringsRHS::operator() over the cells lo..hi (inclusive) of one box,
 * the ParallelFor of AxionStringsLevel.cpp:913-918
 * (AxionStringsRHS.impl.hpp:25-60): rhs(.., 0..3) = (Pi1, Pi2,
 * lap psi1 + c psi1 - (lambda/2)(|psi|^2 - R^2) psi1, ...).  The state view
 * holds lo-2 .. hi+2 (ghosts filled by the caller, as FillPatch does) with 4
 * components; the rhs view holds lo .. hi.  precision AXS_FP64 (double
 * views) or AXS_FP32 (float views): the arithmetic and bits of
 * axs_rhs_eval.  Asynchronous on `stream` (a hipStream_t, NULL = the null
 * stream); allocates nothing. */
int axs_rhs_array4(const axs_array4 *state, axs_array4 *rhs, const int lo[3],
                   const int hi[3], const axs_rhs_coeffs *c, double dx,
                   int precision, void *stream);
/* StringTagger / count_plaquettes over the cells lo..hi of one box
 * (StringTagger.hpp:83-274, StringFinder.hpp:76-122; the ParallelFor of
 * AxionStringsLevel.cpp:1999-2003): the windings of the plaquettes at each
 * cell's low corner, added to counts_dev[0] (plain) and counts_dev[1]
 * (weighted), unsigned 64-bit device counters the caller clears; with a
 * tags view (uint8, 1 component) tags are set to 1 where a criterion of p
 * holds (never cleared; p NULL = the reference tagger).  The state view
 * holds lo .. hi+1 (lo-2 .. hi+2 with the derivative criteria); FP32C
 * levels pass their compensation words as state_lo (same shape, float),
 * else NULL.  The bits of axs_tag_cells / axs_boxlevel_tag.  Asynchronous
 * on `stream`; allocates nothing. */
int axs_tag_array4(const axs_array4 *state, const axs_array4 *state_lo,
                   const int lo[3], const int hi[3],
                   const axs_tagger_params *p, double dx, int precision,
                   axs_array4 *tags, unsigned long long *counts_dev,
                   void *stream);

/* -------------------------- AMR hierarchy --------------------------- */
/* The reference's AMR driver around the hot path (AMReX's Amr with
 * AxionStringsLevel as its AmrLevel): a base level and up to max_level
 * ratio-2 levels of boxes, built by tagging (tag_cells,
 * AxionStringsLevel.cpp:1916-2005), buffering by amr.n_error_buf, cutting
 * into amr.blocking_factor blocks and merging them into boxes of at most
 * amr.max_grid_size cells; advanced by subcycling (level l+1 takes two
 * half steps per step of level l, its coarse-fine ghosts interpolated in
 * time, then averaged down onto level l); rebuilt from fresh tags every
 * amr.regrid_int coarse steps.  Box generation is block merging (runs of
 * blocks along x, merged along y, then z), not AMReX's Berger-Rigoutsos
 * clustering: the boxes differ from the reference's, the nesting,
 * buffering and coverage rules are the same. */
#define AXS_MAX_LEVELS 8
typedef struct axs_hierarchy axs_hierarchy;
typedef struct axs_amr_params
{
    int max_level;       /* amr.max_level, 1..AXS_MAX_LEVELS */
    int n_error_buf;     /* amr.n_error_buf: tags grown by this many cells */
    int blocking_factor; /* amr.blocking_factor: 8, 16, 32 or 64 fine cells */
    int max_grid_size;   /* amr.max_grid_size: box edges <= this (<= 64) */
    int regrid_int;      /* amr.regrid_int: rebuild every N coarse steps
                            (<= 0: never) */
    int force_full_refinement; /* axion_strings.tagging.force_full_refinement */
    double gradient_threshold;         /* axion_strings.tagging.* (DBL_MAX =  */
    double radial_gradient_threshold;  /* off), amplitude_threshold 0 = off  */
    double amplitude_threshold;
    int n_schedule; /* entries of level_add_log_mr_over_h; 0 = no gating */
    double level_add_log_mr_over_h[AXS_MAX_LEVELS]; /* entry l-1: level l may
                       exist from this log(m_r/H) on (the box plan's
                       axion_strings.derived_level_add_log_mr_over_h) */
    int use_graph; /* replay each level's step from a captured graph */
    int per_stage_launches; /* 0: every box level steps with fused stage
                               pairs (axs_boxlevel_set_stage_fusion mode 1);
                               1: one launch per stage */
} axs_amr_params;
/* The stage scalars' blend after the pre-evolution handoff
 * (AxionStringsLevel.cpp:889-907): lambda and the curvature coefficient go
 * from the *_start values to the background's over [tau_start, tau_start +
 * dtau] by smoothstep.  dtau = 0: no blend. */
typedef struct axs_coeff_blend
{
    double tau_start, dtau, lambda_start, curvature_start;
} axs_coeff_blend;
/* Builds the hierarchy at conformal time tau: level 0 takes the valid
 * cells of `base` (a uniform periodic level, cube cells) in boxes of
 * min(max_grid_size, n) cells per edge; each finer level is created from
 * the tags of the level below while the schedule allows it.  dt0 is the
 * base level's step (evolution.dt_multiplier * dx).  `base` is only read;
 * the caller keeps it. */
int axs_hierarchy_create(axs_ctx *ctx, const axs_level *base,
                         const axs_amr_params *p, const axs_background *bg,
                         const axs_coeff_blend *blend, double tau, double dt0,
                         axs_hierarchy **out);
int axs_hierarchy_destroy(axs_hierarchy *h);
/* One coarse step of the whole hierarchy, then a regrid when the step
 * count reaches a multiple of regrid_int. */
int axs_hierarchy_step(axs_hierarchy *h);
/* Rebuilds levels 1.. from fresh tags: a level whose boxes do not change
 * is kept as it is; a new level takes quartic interpolation from the level
 * below, then the old level's values where both cover a cell. */
int axs_hierarchy_regrid(axs_hierarchy *h);
/* nlevels (levels in use), tau, coarse steps taken, regrids done, and the
 * regrid wall seconds (device-synchronised): [0] total, [1] tagging and
 * boxes, [2] level creation, [3] interpolation, [4] copies from the old
 * level, [5] freeing it.  Any pointer may be NULL. */
int axs_hierarchy_info(const axs_hierarchy *h, int *nlevels, double *tau,
                       int64_t *nsteps, int64_t *nregrids,
                       double regrid_seconds[6]);
/* Level l (borrowed: valid until the next step, regrid or destroy), its
 * domain, cell size and box count.  Any output but lvl may be NULL. */
int axs_hierarchy_level(const axs_hierarchy *h, int l, axs_boxlevel **lvl,
                        int domain[3], double *dx, int *nboxes);
/* Box corners of level l (inclusive), 3 ints per box, in the level's order
 * (sorted by z, y, x of the low corner; level 0 tiles in that order). */
int axs_hierarchy_boxes(const axs_hierarchy *h, int l, int *lo, int *hi);
/* compute_composite_total_energy's raw sums (EnergyKernel.hpp:287-339):
 * every level's 14 energy sums over the cells no finer level covers,
 * weighted by dx_l^3 and added; out[13] is the comoving volume counted,
 * out[12] its unmasked part.  Fills every level's ghosts first. */
int axs_hierarchy_energy_sums(axs_hierarchy *h, double R, double lambda,
                              double R_prime_over_R, const axs_masking *m,
                              double out[14]);
/* Pierced-plaquette counts (plain, winding-weighted) of every level over
 * the cells no finer level covers (the composite xi of
 * AxionStringsLevel.cpp:1299-1337 weights level l's by dx_l):
 * counts[2 l], counts[2 l + 1].  Fills every level's ghosts first. */
int axs_hierarchy_plaquette_counts(axs_hierarchy *h, int64_t *counts);
/* collapse_to_level0 (AxionStringsLevel.cpp:110-134) into the uniform
 * level dst (level 0's domain and precision): level 0's values where
 * nothing refines it, the recursive 2x2x2 average of the finer levels
 * where something does.  The hierarchy is not changed. */
int axs_hierarchy_collapse(axs_hierarchy *h, axs_level *dst);
/* The host steps of box generation, exposed for testing:
 * tagged_blocks: blocks of `block` cells of a dense [z][y][x] tag array of
 * size dom that hold a tag after growing the tags by r cells (periodic,
 * r <= block) -> out [dom/block] bytes;
 * erode: a [z][y][x] block mask kept where its whole periodic
 * (2r+1)^3 neighbourhood is set;
 * boxes_from_blocks: boxes of the next finer level (ratio 2, fine_block
 * fine cells per block) over the set blocks: runs of up to max_blocks along
 * x, merged along y, then z (each up to max_blocks); *nboxes boxes written
 * to lo / hi (3 ints each) if at most max_boxes, else AXS_ERR_ARG. */
int axs_amr_tagged_blocks(const uint8_t *tags, const int dom[3], int block,
                          int r, uint8_t *out);
/* No reference counterpart: the same block mask from tags in device memory
 * into device memory (dom/block bytes), as the hierarchy's regrid forms it
 * on the device; synchronizes the context's stream. */
int axs_amr_tagged_blocks_device(axs_ctx *ctx, const uint8_t *tags_dev,
                                 const int dom[3], int block, int r,
                                 uint8_t *out_dev);
int axs_amr_erode(const uint8_t *mask, const int bdim[3], int r,
                  uint8_t *out);
int axs_amr_boxes_from_blocks(const uint8_t *blk, const int bdim[3],
                              int fine_block, int max_blocks, int max_boxes,
                              int *nboxes, int *lo, int *hi);

/* ------------------------------ adjoint ----------------------------- */
/* Gradient of J = sum_cells (|grad psi1|^2 + |grad psi2|^2) dx^3 / R_f^4
 * (the gradient term of rho_tot_pointwise, Energy.hpp:205-215, volume
 * summed as in EnergyKernel.hpp:301-316) at the end of nsteps RK4 steps,
 * with respect to the initial state (4n doubles, written to grad_state)
 * and to lambda0 (the overall scale of lambda(tau), *grad_lambda0).
 * Forward in the level's precision, adjoint accumulated in FP64,
 * checkpoints every `checkpoint_every` steps with recompute in between.
 * On return the level holds the final state; *J receives the objective. */
int axs_adjoint_grad_energy(axs_level *lvl, const axs_background *bg,
                            double lambda0, double tau, double dt,
                            int nsteps, int checkpoint_every, double *J,
                            double *grad_state, double *grad_lambda0);
/* Forward only: the objective J after nsteps (same definition). */
int axs_forward_grad_energy(axs_level *lvl, const axs_background *bg,
                            double lambda0, double tau, double dt,
                            int nsteps, double *J);

#ifdef __cplusplus
}
#endif

#endif /* AXS_H_ */
