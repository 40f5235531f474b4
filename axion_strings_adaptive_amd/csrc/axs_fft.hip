// axs_fft.hip -- the FFT-based pieces of the hot path on the device:
//   * compute_spectrum (SpectrumKernel.hpp:52-153): fill -> hipFFT R2C ->
//     shell binning.  The binning replaces the reference's host-serial loop
//     with a deterministic device reduction: along one (ky,kz) row the shell
//     index floor(|p|+0.5) is non-decreasing in kx, so a wavefront covering
//     64 consecutive kx holds contiguous runs of equal shells; a segmented
//     scan over DPP lane patterns sums each run and the run's last lane adds
//     it to the wavefront's private LDS histogram (no two lanes of one
//     instruction hit the same shell).  Shells come from integer |p|^2, the
//     row indices from wave-uniform counters.  Wave histograms are merged in fixed order into
//     per-block partials, summed in fixed block order by a second pass.
//   * FourierIC::generate (FourierIC.hpp:237-285): hash-seeded spectral
//     fill, Hermitian projection of the self-conjugate planes, hipFFT C2R,
//     deterministic sum of squares, rescale into the level's state.
// FFT memory is the level's RK stage scratch (idle between steps).
#include "axs_internal.hpp"

#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

namespace axs
{

#define AXS_FFT_TRY(expr)                                                     \
    do                                                                        \
    {                                                                         \
        hipfftResult r_ = (expr);                                             \
        if (r_ != HIPFFT_SUCCESS)                                             \
            return ::axs::set_error(AXS_ERR_HIP, std::string(#expr) +         \
                                                     " failed, code " +       \
                                                     std::to_string((int)r_)); \
    } while (0)

__device__ __forceinline__ int wrapped_freq_d(int idx, int n)
{
    return (idx <= n / 2) ? idx : idx - n;
}

// ------------------------------------------------------- shell binning --
constexpr int BIN_WAVES = 4;
constexpr int kBinBatch = 8; // units whose loads a wave keeps in flight
constexpr int kMaxShells = 1025; // up to 2048^3

struct BinParams
{
    int nx, ny, nz, nkx; // nkx = nx/2 + 1
    int n_shells, n_max;
    long chunks_per_row; // 64-lane chunks per (ky, kz) row
    int extra_kx;        // 1: kx = nx/2 is past the last chunk (nkx = 64 c + 1)
    long total_units;    // ny*nz*chunks_per_row
    long units_per_wave;
    double *partials; // [gridDim.x][2 n_shells + 2]
};

template <typename C>
__device__ __forceinline__ double mag2(const C *spec, size_t i)
{
    const double re = (double)spec[2 * i], im = (double)spec[2 * i + 1];
    return re * re + im * im;
}

// floor(|p| + 0.5) for |p|^2 = p2 (an integer below 2^22), in integers:
// the s with (2s - 1)^2 < 4 p2 < (2s + 1)^2 (4 p2 is even, the bounds
// odd, so neither can be equal).  This is the shell floor(sqrt(p2) + 0.5)
// of double arithmetic exactly: sqrt(p2) is never closer to a half-integer
// than 0.25 / (2 sqrt(p2) + 1) (> 7e-5 here), far more than double
// rounding moves it.  The float estimate (the hardware square root, within
// an ulp) is within one of s.
__device__ __forceinline__ int shell_of(int p2)
{
    int s = (int)(__builtin_amdgcn_sqrtf((float)p2) + 0.5f);
    const int q = 4 * p2;
    if ((2 * s + 1) * (2 * s + 1) < q)
        ++s;
    else if (s > 0 && (2 * s - 1) * (2 * s - 1) > q)
        --s;
    return s;
}

// One step of a segmented inclusive scan through a DPP lane pattern: each
// lane adds the value of the lane the pattern names if that lane is in the
// same run (equal shell); lanes the pattern gives no source keep the
// neighbour-shell sentinel -2 and add nothing.
template <int CTRL, int ROWS, bool ZERO>
__device__ __forceinline__ void seg_scan_step(int shell, double &v, int &cnt)
{
    // ZERO: lanes without a source read 0 (for the row shifts: v and cnt
    // add nothing, and shell 0 -- the DC mode alone -- matches only with a
    // zero value); else they keep the sentinel -2 and zeros
    const int os = ZERO ? 0 : -2;
    const long long b = __double_as_longlong(v);
    const int so = __builtin_amdgcn_update_dpp(os, shell, CTRL, ROWS, 0xf, ZERO);
    const int vlo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, ZERO);
    const int vhi =
        __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, ZERO);
    const int co = __builtin_amdgcn_update_dpp(0, cnt, CTRL, ROWS, 0xf, ZERO);
    if (so == shell)
    {
        v += __longlong_as_double(((long long)vhi << 32) | (unsigned)vlo);
        cnt += co;
    }
}

// DPP controls of the wave64 scan: shift right by 1, 2, 4, 8 lanes inside
// each row of 16 lanes, then lane 15 of rows 0 / 2 to rows 1 / 3 and lane
// 31 to rows 2 and 3.
constexpr int kDppRowShr = 0x110, kDppBcast15 = 0x142, kDppBcast31 = 0x143;

template <typename C>
__global__ void __launch_bounds__(64 * BIN_WAVES)
    k_bin(const C *spec, BinParams p)
{
    extern __shared__ double hist[]; // [BIN_WAVES][2][n_shells]
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double *h = hist + (size_t)wave * 2 * p.n_shells;
    double *hc = h + p.n_shells;
    for (int s = lane; s < 2 * p.n_shells; s += 64)
        h[s] = 0.0;
    __syncthreads();

    double full = 0.0, insc = 0.0; // per lane
    const long gw = (long)blockIdx.x * BIN_WAVES + wave;
    long u = gw * p.units_per_wave;
    const long u1 = min(u + p.units_per_wave, p.total_units);
    const double half = p.n_max / 2.0, hh = half * half;
    const int nk = p.extra_kx ? p.nkx - 1 : p.nkx; // kx of the chunks
    // the wave's units run through (k, j, chunk) with chunk fastest: one
    // division for the first unit, then counters (all wave-uniform).  Units
    // go in batches whose loads are all issued before the first is used, so
    // each wave keeps kBinBatch loads in flight.
    const long row0 = u / p.chunks_per_row;
    int chunk = (int)(u - row0 * p.chunks_per_row);
    int j = (int)(row0 % p.ny), k = (int)(row0 / p.ny);
    size_t rowbase = (size_t)row0 * p.nkx; // spectrum index of (0, j, k)
    auto kyz2_of = [&](int jj, int kk) {
        const int ky = wrapped_freq_d(jj, p.ny), kz = wrapped_freq_d(kk, p.nz);
        return ky * ky + kz * kz;
    };
    int kyz2 = kyz2_of(j, k);
    for (; u < u1; u += kBinBatch)
    {
        int bc[kBinBatch], bq[kBinBatch]; // chunk, ky^2 + kz^2
        size_t br[kBinBatch];             // row base
        C bre[kBinBatch], bim[kBinBatch];
#pragma unroll
        for (int g = 0; g < kBinBatch; ++g)
        {
            bc[g] = chunk;
            bq[g] = kyz2;
            br[g] = rowbase;
            // every lane loads (lanes past the row and units past the
            // wave's range read an element in bounds that is not used), so
            // no branch separates the batch's loads
            const int i = min(chunk * 64 + lane, nk - 1);
            const size_t o = u + g < u1 ? rowbase + i : 0;
            bre[g] = spec[2 * o];
            bim[g] = spec[2 * o + 1];
            if (++chunk == p.chunks_per_row)
            {
                chunk = 0;
                rowbase += p.nkx;
                if (++j == p.ny)
                {
                    j = 0;
                    ++k;
                }
                kyz2 = kyz2_of(j, k);
            }
        }
#pragma unroll
        for (int g = 0; g < kBinBatch; ++g)
        {
            if (u + g >= u1)
                break;
            const int kx = bc[g] * 64 + lane;
            const int kx0 = bc[g] * 64; // the chunk's smallest kx
            const double re = (double)bre[g], im = (double)bim[g];
            const double m2 = re * re + im * im;
            const int wgt = (kx == 0 || kx == p.nx / 2) ? 1 : 2;
            if ((double)(kx0 * kx0 + bq[g]) >= hh)
            {
                // the whole chunk lies outside the inscribed sphere: the
                // full-cube sum only (wave-uniform branch)
                if (kx < nk)
                    full += wgt * m2;
            }
            else
            {
                int shell = -1, cnt = 0;
                double v = 0.0;
                if (kx < nk)
                {
                    const int p2 = kx * kx + bq[g];
                    full += wgt * m2;
                    if ((double)p2 < hh) // |p| < n_max / 2
                    {
                        insc += wgt * m2;
                        const int sh = shell_of(p2);
                        if (sh < p.n_shells)
                        {
                            shell = sh;
                            v = wgt * 4.0 * 3.14159265358979323846 * (double)p2 *
                                m2;
                            cnt = wgt;
                        }
                    }
                }
                // segmented inclusive scan over runs of equal shell: along
                // a row the shell is non-decreasing in kx, so equal shells
                // are contiguous
                seg_scan_step<kDppRowShr + 1, 0xf, true>(shell, v, cnt);
                seg_scan_step<kDppRowShr + 2, 0xf, true>(shell, v, cnt);
                seg_scan_step<kDppRowShr + 4, 0xf, true>(shell, v, cnt);
                seg_scan_step<kDppRowShr + 8, 0xf, true>(shell, v, cnt);
                seg_scan_step<kDppBcast15, 0xa, false>(shell, v, cnt);
                seg_scan_step<kDppBcast31, 0xc, false>(shell, v, cnt);
                const int snext = __shfl_down(shell, 1, 64);
                const bool tail = (lane == 63) || (snext != shell);
                if (shell >= 0 && tail)
                {
                    h[shell] += v;
                    hc[shell] += (double)cnt;
                }
            }
            // kx = nx/2 past the last 64-lane chunk (nx a multiple of 128):
            // lane 63, whose run ended just above, bins it on its own
            if (p.extra_kx && bc[g] == p.chunks_per_row - 1 && lane == 63)
            {
                const int kxe = p.nkx - 1;
                const int p2 = kxe * kxe + bq[g];
                const double m2e = mag2(spec, br[g] + kxe);
                full += 1 * m2e;
                if ((double)p2 < hh)
                {
                    insc += 1 * m2e;
                    const int sh = shell_of(p2);
                    if (sh < p.n_shells)
                    {
                        h[sh] += 1 * 4.0 * 3.14159265358979323846 * (double)p2 * m2e;
                        hc[sh] += 1.0;
                    }
                }
            }
        }
    }
    // wave totals (fixed order)
    for (int o = 32; o > 0; o >>= 1)
    {
        full += __shfl_xor(full, o, 64);
        insc += __shfl_xor(insc, o, 64);
    }
    __shared__ double tot[BIN_WAVES][2];
    if (lane == 0)
    {
        tot[wave][0] = full;
        tot[wave][1] = insc;
    }
    __syncthreads();
    // merge wave histograms in wave order -> block partial row
    // [sums(n_shells) | counts(n_shells) | full | inscribed]
    double *out = p.partials + (size_t)blockIdx.x * (2 * p.n_shells + 2);
    for (int s = threadIdx.x; s < 2 * p.n_shells; s += 64 * BIN_WAVES)
    {
        double a = 0.0;
        for (int w = 0; w < BIN_WAVES; ++w)
            a += hist[(size_t)w * 2 * p.n_shells + s];
        out[s] = a;
    }
    if (threadIdx.x == 0)
    {
        double f = 0.0, g = 0.0;
        for (int w = 0; w < BIN_WAVES; ++w)
        {
            f += tot[w][0];
            g += tot[w][1];
        }
        out[2 * p.n_shells] = f;
        out[2 * p.n_shells + 1] = g;
    }
}

__global__ void k_sum_cols(const double *part, int nrow, int width,
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

// ---------------------------------------------------- sum of squares ----
template <typename B>
__global__ void k_sumsq(const B *x, size_t n, double *partials)
{
    __shared__ double red[256];
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256)
    {
        const double v = (double)x[i];
        s += v * v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int)threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partials[blockIdx.x] = red[0];
}

// ------------------------------------------------------------- Fourier IC
__device__ __forceinline__ uint64_t hash_u64_d(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__device__ __forceinline__ double uniform_from_hash_d(uint64_t h)
{
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Fills the half spectrum [nz][ny][nx/2+1] (FourierIC.hpp:253-271).
__global__ void k_ic_fill(double *spec, int nx, int ny, int nz,
                          double k_max_sq, uint64_t full_seed)
{
    const int nkx = nx / 2 + 1;
    const size_t total = (size_t)nkx * ny * nz;
    for (size_t t = blockIdx.x * (size_t)256 + threadIdx.x; t < total;
         t += (size_t)gridDim.x * 256)
    {
        const int i = (int)(t % nkx);
        const int j = (int)((t / nkx) % ny);
        const int k = (int)(t / ((size_t)nkx * ny));
        const int kx = i, ky = wrapped_freq_d(j, ny), kz = wrapped_freq_d(k, nz);
        const double k_sq = (double)(kx * kx + ky * ky + kz * kz);
        double re = 0.0, im = 0.0;
        if (!(k_sq > k_max_sq || k_sq == 0.0))
        {
            const uint64_t key = full_seed ^ ((uint64_t)0 << 60) ^
                                 ((uint64_t)(uint32_t)i << 40) ^
                                 ((uint64_t)(uint32_t)j << 20) ^
                                 (uint64_t)(uint32_t)k;
            const uint64_t h1 = hash_u64_d(key);
            const uint64_t h2 = hash_u64_d(key ^ 0x1234567890ABCDEFULL);
            const double u1 = fmax(uniform_from_hash_d(h1), 1.0e-300);
            const double u2 = uniform_from_hash_d(h2);
            const double r = sqrt(-2.0 * log(u1));
            const double theta = 2.0 * 3.14159265358979323846 * u2;
            re = r * cos(theta);
            im = r * sin(theta);
        }
        spec[2 * t] = re;
        spec[2 * t + 1] = im;
    }
}

// Hermitian projection (X(k) + conj X(-k))/2 of plane kx = i0.
__global__ void k_ic_hermitian(double *spec, int nx, int ny, int nz, int i0)
{
    const int nkx = nx / 2 + 1;
    const size_t total = (size_t)ny * nz;
    for (size_t t = blockIdx.x * (size_t)256 + threadIdx.x; t < total;
         t += (size_t)gridDim.x * 256)
    {
        const int j = (int)(t % ny), k = (int)(t / ny);
        const int jm = (ny - j) % ny, km = (nz - k) % nz;
        const size_t a = ((size_t)k * ny + j) * nkx + i0;
        const size_t b = ((size_t)km * ny + jm) * nkx + i0;
        if (b < a)
            continue;
        const double ar = spec[2 * a], ai = spec[2 * a + 1];
        const double br = spec[2 * b], bi = spec[2 * b + 1];
        const double nr = 0.5 * (ar + br), ni = 0.5 * (ai - bi);
        spec[2 * a] = nr;
        spec[2 * a + 1] = ni;
        spec[2 * b] = nr;
        spec[2 * b + 1] = -ni;
    }
}

template <typename T>
__global__ void k_ic_store(const double *field, double scale, T *dst,
                           float *lo, uint32_t n)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n;
         i += gridDim.x * 256)
    {
        const double v = field[i] * scale;
        const T h = (T)v;
        dst[i] = h;
        if (lo)
            lo[i] = (float)(v - (double)h);
    }
}

static int nblk(size_t n, int cap)
{
    size_t b = (n + 255) / 256;
    if (b > (size_t)cap)
        b = cap;
    return b < 1 ? 1 : (int)b;
}

static int sum_sq_dev(axs_level *L, const void *buf, bool dbl, size_t n,
                      double *out)
{
    hipStream_t s = L->ctx->stream;
    const int nb = nblk(n, 4096);
    if (dbl)
        hipLaunchKernelGGL(k_sumsq<double>, dim3(nb), dim3(256), 0, s,
                           (const double *)buf, n, L->ctx->d_partials);
    else
        hipLaunchKernelGGL(k_sumsq<float>, dim3(nb), dim3(256), 0, s,
                           (const float *)buf, n, L->ctx->d_partials);
    AXS_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sum_cols, dim3(1), dim3(256), 0, s,
                       (const double *)L->ctx->d_partials, nb, 1,
                       L->ctx->d_result);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipMemcpyAsync(L->ctx->h_result, L->ctx->d_result,
                               sizeof(double), hipMemcpyDeviceToHost, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    *out = L->ctx->h_result[0];
    return AXS_OK;
}

// Bytes of one real field and its half spectrum in type R.
static size_t real_bytes(const axs_level *L, size_t rs)
{
    return (size_t)L->g.ncell * rs;
}
static size_t spec_bytes(const axs_level *L, size_t rs)
{
    return (size_t)(L->g.nx / 2 + 1) * L->g.ny * L->g.nz * 2 * rs;
}

static int ensure_plan(axs_level *L, bool c2r, bool dbl, void **work,
                       size_t work_cap)
{
    void **slot = c2r ? &L->fft_c2r : &L->fft_r2c;
    hipfftHandle plan;
    if (*slot == nullptr)
    {
        AXS_FFT_TRY(hipfftCreate(&plan));
        AXS_FFT_TRY(hipfftSetAutoAllocation(plan, 0));
        size_t ws = 0;
        hipfftType t = c2r ? (dbl ? HIPFFT_Z2D : HIPFFT_C2R)
                           : (dbl ? HIPFFT_D2Z : HIPFFT_R2C);
        // x is the fastest index: dims given slowest first
        AXS_FFT_TRY(hipfftMakePlan3d(plan, L->g.nz, L->g.ny, L->g.nx, t, &ws));
        if (ws > work_cap)
        {
            hipfftDestroy(plan);
            return set_error(AXS_ERR_NOMEM,
                             "FFT work area larger than the stage scratch");
        }
        *slot = (void *)(intptr_t)plan;
    }
    plan = (hipfftHandle)(intptr_t)*slot;
    AXS_FFT_TRY(hipfftSetStream(plan, L->ctx->stream));
    AXS_FFT_TRY(hipfftSetWorkArea(plan, *work));
    return AXS_OK;
}

} // namespace axs

using namespace axs;

extern "C" {

int axs_level_spectrum(axs_level *L, int kind, const axs_masking *m, double R,
                       double *shell_sum, int64_t *shell_count, double tot[2],
                       double *sum_sq)
{
    AXS_CHECK_ARG(L && m && shell_sum && shell_count && tot,
                  "axs_level_spectrum: NULL argument");
    AXS_CHECK_ARG(L->g.nx % 2 == 0, "axs_level_spectrum: nx must be even");
    const int n_max = std::max(L->g.nx, std::max(L->g.ny, L->g.nz));
    const int n_shells = n_max / 2 + 1;
    AXS_CHECK_ARG(n_shells <= kMaxShells, "axs_level_spectrum: grid too large");
    const bool dbl = L->precision == AXS_FP64;
    const size_t rs = dbl ? 8 : 4;
    // scratch: real field in ya, half spectrum in acc, FFT work area in yb
    const size_t cap = (size_t)4 * L->g.ncell * L->word;
    AXS_CHECK_ARG(real_bytes(L, rs) <= cap && spec_bytes(L, rs) <= cap,
                  "axs_level_spectrum: scratch too small");
    void *field = L->ya, *spec = L->acc, *work = L->yb;
    int rc = axs_fill_field(L, kind, m, R, field);
    if (rc)
        return rc;
    if (sum_sq && (rc = sum_sq_dev(L, field, dbl, L->g.ncell, sum_sq)))
        return rc;
    if ((rc = ensure_plan(L, false, dbl, &work, cap)))
        return rc;
    hipfftHandle plan = (hipfftHandle)(intptr_t)L->fft_r2c;
    if (dbl)
        AXS_FFT_TRY(hipfftExecD2Z(plan, (hipfftDoubleReal *)field,
                                  (hipfftDoubleComplex *)spec));
    else
        AXS_FFT_TRY(
            hipfftExecR2C(plan, (hipfftReal *)field, (hipfftComplex *)spec));

    hipStream_t s = L->ctx->stream;
    BinParams p{};
    p.nx = L->g.nx;
    p.ny = L->g.ny;
    p.nz = L->g.nz;
    p.nkx = p.nx / 2 + 1;
    p.n_shells = n_shells;
    p.n_max = n_max;
    // nkx = 64 c + 1 (nx a multiple of 128): c chunks, and lane 63 of each
    // row's last chunk takes kx = nx/2, rather than a chunk of one lane
    p.extra_kx = (p.nkx % 64 == 1 && p.nkx > 1) ? 1 : 0;
    p.chunks_per_row = (p.nkx - p.extra_kx + 63) / 64;
    p.total_units = (long)p.ny * p.nz * p.chunks_per_row;
    int nb = 2048;
    const int width = 2 * n_shells + 2;
    while ((long)(nb + 1) * width > (long)kMaxPartials * 16)
        nb >>= 1;
    const long waves = (long)nb * BIN_WAVES;
    p.units_per_wave = (p.total_units + waves - 1) / waves;
    p.partials = L->ctx->d_partials;
    const size_t lds = (size_t)BIN_WAVES * 2 * n_shells * sizeof(double);
    if (dbl)
        hipLaunchKernelGGL(k_bin<double>, dim3(nb), dim3(64 * BIN_WAVES), lds, s,
                           (const double *)spec, p);
    else
        hipLaunchKernelGGL(k_bin<float>, dim3(nb), dim3(64 * BIN_WAVES), lds, s,
                           (const float *)spec, p);
    AXS_HIP_TRY(hipGetLastError());
    // column sums (fixed order) into the row after the partials
    double *colsum = L->ctx->d_partials + (size_t)nb * width;
    hipLaunchKernelGGL(k_sum_cols, dim3(width), dim3(256), 0, s,
                       (const double *)p.partials, nb, width, colsum);
    AXS_HIP_TRY(hipGetLastError());
    std::vector<double> hs(width);
    AXS_HIP_TRY(hipMemcpyAsync(hs.data(), colsum, width * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < n_shells; ++i)
    {
        shell_sum[i] += hs[i];
        shell_count[i] += (int64_t)llround(hs[n_shells + i]);
    }
    tot[0] += hs[2 * n_shells];
    tot[1] += hs[2 * n_shells + 1];
    return AXS_OK;
}

int axs_level_fourier_ic(axs_level *L, int comp, double k_max_cells,
                         double target_mean_square, uint64_t seed,
                         int comp_seed_offset)
{
    AXS_CHECK_ARG(L, "axs_level_fourier_ic: NULL level");
    AXS_CHECK_ARG(comp == 0 || comp == 1,
                  "axs_level_fourier_ic: comp must be 0 (psi1) or 1 (psi2)");
    AXS_CHECK_ARG(k_max_cells > 0.0, "axs_level_fourier_ic: k_max must be > 0");
    AXS_CHECK_ARG(target_mean_square >= 0.0,
                  "axs_level_fourier_ic: mean square must be >= 0");
    AXS_CHECK_ARG(L->g.nx % 2 == 0, "axs_level_fourier_ic: nx must be even");
    // FP64 spectrum in acc, FP64 real field in ya|yb (32 B/cell for FP32
    // levels, ya alone for FP64), work area in the remaining scratch.
    const size_t cap4 = (size_t)4 * L->g.ncell * L->word;
    void *spec = L->acc;
    void *field = L->ya;
    void *work = L->yb;
    size_t work_cap = cap4;
    if (L->precision != AXS_FP64)
    {
        // FP32 levels: ya|yb hold 32 B/cell = the real double field plus
        // room; the work area goes after the real field inside ya|yb.
        work = (char *)L->ya + real_bytes(L, 8);
        work_cap = 2 * cap4 - real_bytes(L, 8);
    }
    AXS_CHECK_ARG(spec_bytes(L, 8) <= cap4,
                  "axs_level_fourier_ic: scratch too small");
    hipStream_t s = L->ctx->stream;
    const uint64_t full_seed = seed ^ ((uint64_t)comp_seed_offset << 32);
    const size_t nspec = (size_t)(L->g.nx / 2 + 1) * L->g.ny * L->g.nz;
    hipLaunchKernelGGL(k_ic_fill, dim3(nblk(nspec, 8192)), dim3(256), 0, s,
                       (double *)spec, L->g.nx, L->g.ny, L->g.nz,
                       k_max_cells * k_max_cells, full_seed);
    AXS_HIP_TRY(hipGetLastError());
    const size_t nplane = (size_t)L->g.ny * L->g.nz;
    hipLaunchKernelGGL(k_ic_hermitian, dim3(nblk(nplane, 4096)), dim3(256), 0,
                       s, (double *)spec, L->g.nx, L->g.ny, L->g.nz, 0);
    hipLaunchKernelGGL(k_ic_hermitian, dim3(nblk(nplane, 4096)), dim3(256), 0,
                       s, (double *)spec, L->g.nx, L->g.ny, L->g.nz,
                       L->g.nx / 2);
    AXS_HIP_TRY(hipGetLastError());
    int rc = ensure_plan(L, true, true, &work, work_cap);
    if (rc)
        return rc;
    // the C2R plan is always FP64 (ensure_plan keyed by direction only)
    AXS_FFT_TRY(hipfftExecZ2D((hipfftHandle)(intptr_t)L->fft_c2r,
                              (hipfftDoubleComplex *)spec,
                              (hipfftDoubleReal *)field));
    double ssq = 0.0;
    if ((rc = sum_sq_dev(L, field, true, L->g.ncell, &ssq)))
        return rc;
    const double current = ssq / (double)L->g.ncell;
    const double scale =
        current > 0.0 ? std::sqrt(target_mean_square / current) : 1.0;
    const uint32_t n = L->g.ncell;
    if (L->precision == AXS_FP64)
        hipLaunchKernelGGL(k_ic_store<double>, dim3(nblk(n, 8192)), dim3(256),
                           0, s, (const double *)field, scale,
                           (double *)L->y + (size_t)comp * n, (float *)nullptr,
                           n);
    else
        hipLaunchKernelGGL(k_ic_store<float>, dim3(nblk(n, 8192)), dim3(256), 0,
                           s, (const double *)field, scale,
                           (float *)L->y + (size_t)comp * n,
                           L->comp ? L->comp + (size_t)comp * n : nullptr, n);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipStreamSynchronize(s));
    return AXS_OK;
}

} // extern "C"
