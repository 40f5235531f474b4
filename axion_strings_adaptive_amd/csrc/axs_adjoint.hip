// axs_adjoint.hip -- discrete adjoint of the RHS and the classic RK4 step.
//
// Objective: J = w * sum_cells (|D psi1|^2 + |D psi2|^2),  w = dx^3 / R_f^4,
// D the 4th-order first derivative of the energy diagnostic, evaluated on
// the final state -- the gradient term of rho_tot_pointwise
// (Energy.hpp:205-215) volume-summed as EnergyKernel.hpp:301-316 does.
//
// With periodic boundaries D is antisymmetric, so dJ/dpsi = -2 w D(D psi).
// The RHS f(psi, Pi) = (Pi, lap psi + c psi - (lam/2)(|psi|^2-R^2) psi) has
// the transposed Jacobian
//   (J^T kb)_psi_i = lap kb_Pi_i + c kb_Pi_i
//                    - (lam/2)[(|psi|^2-R^2) kb_Pi_i + 2 psi_i (psi . kb_Pi)]
//   (J^T kb)_Pi_i  = kb_psi_i
// (lap is symmetric).  For one RK4 step y' = y + h/6(k1 + 2k2 + 2k3 + k4),
// k_s = f(y_{s-1}), y_s = y + a_s k_s (a = h/2, h/2, h), the adjoint maps
// ybar' to ybar = ybar' + sum_s J_s^T kbar_s with
//   kbar_4 = h/6 ybar'                ybar_3 = J_4^T kbar_4
//   kbar_3 = h/3 ybar' + h   ybar_3   ybar_2 = J_3^T kbar_3
//   kbar_2 = h/3 ybar' + h/2 ybar_2   ybar_1 = J_2^T kbar_2
//   kbar_1 = h/6 ybar' + h/2 ybar_1   ybar_0 = J_1^T kbar_1
// and lambda(tau) = lambda0 * g(tau) contributes
//   dJ/dlambda0 += sum_s kbar_s,Pi . (-(g_s/2)(|psi|^2 - R^2) psi)(y_{s-1}).
// The forward runs in the level's precision (stage states recomputed from
// per-step states of the current checkpoint segment); every adjoint
// quantity is FP64.
#include "axs_internal.hpp"
#include "host/Background.hpp"

#include <cmath>
#include <vector>

namespace axs
{

__device__ __forceinline__ int wrp(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

template <typename T>
__device__ __forceinline__ double ldv(const T *p, const float *lo, uint32_t i)
{
    return lo ? (double)p[i] + (double)lo[i] : (double)p[i];
}

struct AdjGrid
{
    int nx, ny, nz;
    uint32_t n;
    double dx;
};

__device__ __forceinline__ uint32_t gid(const AdjGrid &g, int x, int y, int z)
{
    return ((uint32_t)wrp(z, g.nz) * g.ny + wrp(y, g.ny)) * g.nx + wrp(x, g.nx);
}

// g[c*3+d] = D_d psi_c (6 arrays of n doubles); partial sums of g^2
template <typename T>
__global__ void k_adj_grad1(const T *y, const float *lo, AdjGrid g,
                            double *gout, double *partials)
{
    __shared__ double red[256];
    const double i1 = 1.0 / g.dx;
    const double wf = 8.33333333333333333333e-2, wn = 6.66666666666666666667e-1;
    double s = 0.0;
    for (uint32_t id = blockIdx.x * 256 + threadIdx.x; id < g.n;
         id += gridDim.x * 256)
    {
        const int x = id % g.nx, yy = (id / g.nx) % g.ny,
                  z = id / ((uint32_t)g.nx * g.ny);
        for (int c = 0; c < 2; ++c)
        {
            const T *p = y + (size_t)c * g.n;
            const float *pl = lo ? lo + (size_t)c * g.n : nullptr;
            for (int d = 0; d < 3; ++d)
            {
                const int ex = d == 0, ey = d == 1, ez = d == 2;
                const double v =
                    (wf * ldv(p, pl, gid(g, x - 2 * ex, yy - 2 * ey, z - 2 * ez)) -
                     wn * ldv(p, pl, gid(g, x - ex, yy - ey, z - ez)) +
                     wn * ldv(p, pl, gid(g, x + ex, yy + ey, z + ez)) -
                     wf * ldv(p, pl, gid(g, x + 2 * ex, yy + 2 * ey, z + 2 * ez))) *
                    i1;
                gout[(size_t)(c * 3 + d) * g.n + id] = v;
                s += v * v;
            }
        }
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

// lam_N: psi_c = -2 w sum_d D_d g[c][d], Pi = 0
__global__ void k_adj_grad2(const double *gin, AdjGrid g, double w, double *out)
{
    const double i1 = 1.0 / g.dx;
    const double wf = 8.33333333333333333333e-2, wn = 6.66666666666666666667e-1;
    for (uint32_t id = blockIdx.x * 256 + threadIdx.x; id < g.n;
         id += gridDim.x * 256)
    {
        const int x = id % g.nx, yy = (id / g.nx) % g.ny,
                  z = id / ((uint32_t)g.nx * g.ny);
        for (int c = 0; c < 2; ++c)
        {
            double acc = 0.0;
            for (int d = 0; d < 3; ++d)
            {
                const double *p = gin + (size_t)(c * 3 + d) * g.n;
                const int ex = d == 0, ey = d == 1, ez = d == 2;
                acc += (wf * p[gid(g, x - 2 * ex, yy - 2 * ey, z - 2 * ez)] -
                        wn * p[gid(g, x - ex, yy - ey, z - ez)] +
                        wn * p[gid(g, x + ex, yy + ey, z + ez)] -
                        wf * p[gid(g, x + 2 * ex, yy + 2 * ey, z + 2 * ez)]) *
                       i1;
            }
            out[(size_t)c * g.n + id] = -2.0 * w * acc;
            out[(size_t)(c + 2) * g.n + id] = 0.0;
        }
    }
}

struct AdjStage
{
    const double *lp;    // ybar' (4n)
    double a;            // kbar = a*ybar' + b*yprev
    const double *yprev; // ybar of the previous adjoint stage or nullptr
    double b;
    double cc, lam, R2, gfac; // stage scalars, gfac = lambda / lambda0
    double *yout;        // ybar_s = J^T kbar (4n)
    double *acc;         // acc += ybar_s
    double *partials;    // dJ/dlambda0 block partials
};

// ybar_s = J_s^T kbar at the stage input state psi (level precision)
template <typename T>
__global__ void k_adj_stage(const T *y, const float *lo, AdjGrid g, AdjStage s)
{
    __shared__ double red[256];
    const double i2 = 1.0 / (g.dx * g.dx);
    const double w_far = 8.33333333333333333333e-2;
    const double w_near = 1.33333333333333333333e+0;
    const double w_local = 2.50000000000000000000e+0;
    double dl = 0.0;
    const uint32_t n = g.n;
    auto kb = [&](int comp, uint32_t i) -> double {
        const double v = s.a * s.lp[(size_t)comp * n + i];
        return s.yprev ? v + s.b * s.yprev[(size_t)comp * n + i] : v;
    };
    for (uint32_t id = blockIdx.x * 256 + threadIdx.x; id < n;
         id += gridDim.x * 256)
    {
        const int x = id % g.nx, yy = (id / g.nx) % g.ny,
                  z = id / ((uint32_t)g.nx * g.ny);
        const double psi1 = ldv(y, lo, id);
        const double psi2 = ldv(y + n, lo ? lo + n : nullptr, id);
        const double kP1 = kb(2, id), kP2 = kb(3, id);
        double lap[2];
        for (int c = 0; c < 2; ++c)
        {
            const int comp = 2 + c;
            double r = 0.0;
            for (int d = 0; d < 3; ++d)
            {
                const int ex = d == 0, ey = d == 1, ez = d == 2;
                r += (-w_far * kb(comp, gid(g, x - 2 * ex, yy - 2 * ey, z - 2 * ez)) +
                      w_near * kb(comp, gid(g, x - ex, yy - ey, z - ez)) -
                      w_local * kb(comp, id) +
                      w_near * kb(comp, gid(g, x + ex, yy + ey, z + ez)) -
                      w_far * kb(comp, gid(g, x + 2 * ex, yy + 2 * ey, z + 2 * ez))) *
                     i2;
            }
            lap[c] = r;
        }
        const double d = psi1 * psi1 + psi2 * psi2 - s.R2;
        const double dot = psi1 * kP1 + psi2 * kP2;
        const double o1 = lap[0] + s.cc * kP1 - 0.5 * s.lam * (d * kP1 + 2.0 * psi1 * dot);
        const double o2 = lap[1] + s.cc * kP2 - 0.5 * s.lam * (d * kP2 + 2.0 * psi2 * dot);
        const double o3 = kb(0, id), o4 = kb(1, id);
        s.yout[id] = o1;
        s.yout[n + id] = o2;
        s.yout[2 * (size_t)n + id] = o3;
        s.yout[3 * (size_t)n + id] = o4;
        s.acc[id] += o1;
        s.acc[n + id] += o2;
        s.acc[2 * (size_t)n + id] += o3;
        s.acc[3 * (size_t)n + id] += o4;
        // d f_Pi / d lambda0 = -(g/2) d psi
        dl += -0.5 * s.gfac * d * dot;
    }
    red[threadIdx.x] = dl;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int)threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        s.partials[blockIdx.x] = red[0];
}

__global__ void k_copy_d(const double *a, double *b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256)
        b[i] = a[i];
}

__global__ void k_sum_vec(const double *p, int n, double *out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256)
        s += p[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int)threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = red[0];
}

static int nblk(size_t n)
{
    size_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

static void make_bg(const axs_background *b, Background &bg)
{
    CTauSchedule s{};
    s.c0 = b->c0;
    s.c1 = b->c1;
    s.tau_switch = b->tau_switch;
    s.has_switch = b->has_switch != 0;
    bg = Background(b->a_inv, s);
}

// coefficient table [step][stage][3] with lambda scaled by lambda0, plus
// the per-stage lambda/lambda0 factors
static void build_table(const axs_background *b, double lambda0, double tau,
                        double dt, int nsteps, std::vector<double> &tab,
                        std::vector<double> &gfac)
{
    Background bg;
    if (!b->flat_space)
        make_bg(b, bg);
    tab.assign((size_t)nsteps * 12, 0.0);
    gfac.assign((size_t)nsteps * 4, 0.0);
    for (int i = 0; i < nsteps; ++i)
    {
        const double ts[4] = {tau, tau + 0.5 * dt, tau + 0.5 * dt, tau + dt};
        for (int s = 0; s < 4; ++s)
        {
            double cc, lam, R;
            if (b->flat_space)
            {
                cc = 0.0;
                lam = b->flat_m_r * b->flat_m_r;
                R = 1.0;
            }
            else
            {
                cc = bg.curvature_term_coeff(ts[s]);
                lam = bg.lambda(ts[s]);
                R = bg.R(ts[s]);
            }
            tab[(size_t)i * 12 + 3 * s] = cc;
            tab[(size_t)i * 12 + 3 * s + 1] = lambda0 * lam;
            tab[(size_t)i * 12 + 3 * s + 2] = R * R;
            gfac[(size_t)i * 4 + s] = lam;
        }
        tau += dt;
    }
}

struct Workspace
{
    double *d = nullptr;
    ~Workspace() { (void)hipFree(d); }
};

// One forward RK4 step with the table row `row` (explicit stage scalars).
static int fwd_step(axs_level *L, int row, double dt)
{
    const double *cf = L->d_coeff + (size_t)12 * row;
    int rc;
    if ((rc = launch_rk_stage(L, 1, L->y, L->ya, cf, nullptr, 0, dt)))
        return rc;
    if ((rc = launch_rk_stage(L, 2, L->ya, L->yb, cf, nullptr, 1, dt)))
        return rc;
    if ((rc = launch_rk_stage(L, 3, L->yb, L->ya, cf, nullptr, 2, dt)))
        return rc;
    return launch_rk_stage(L, 4, L->ya, L->y, cf, nullptr, 3, dt);
}

static int objective(axs_level *L, double R, double *J, double *gbuf,
                     double *lamN)
{
    hipStream_t s = L->ctx->stream;
    AdjGrid g{L->g.nx, L->g.ny, L->g.nz, L->g.ncell, L->dx};
    const int nb = nblk(g.n);
    const double w = L->dx * L->dx * L->dx / (R * R * R * R);
    if (L->precision == AXS_FP64)
        hipLaunchKernelGGL(k_adj_grad1<double>, dim3(nb), dim3(256), 0, s,
                           (const double *)L->y, (const float *)nullptr, g,
                           gbuf, L->ctx->d_partials);
    else
        hipLaunchKernelGGL(k_adj_grad1<float>, dim3(nb), dim3(256), 0, s,
                           (const float *)L->y, (const float *)L->comp, g, gbuf,
                           L->ctx->d_partials);
    hipLaunchKernelGGL(k_sum_vec, dim3(1), dim3(256), 0, s,
                       (const double *)L->ctx->d_partials, nb, L->ctx->d_result);
    if (lamN)
        hipLaunchKernelGGL(k_adj_grad2, dim3(nb), dim3(256), 0, s,
                           (const double *)gbuf, g, w, lamN);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipMemcpyAsync(L->ctx->h_result, L->ctx->d_result,
                               sizeof(double), hipMemcpyDeviceToHost, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));
    *J = w * L->ctx->h_result[0];
    return AXS_OK;
}

static double final_R(const axs_background *b, double tau_end)
{
    if (b->flat_space)
        return 1.0;
    Background bg;
    make_bg(b, bg);
    return bg.R(tau_end);
}

} // namespace axs

using namespace axs;

extern "C" {

int axs_forward_grad_energy(axs_level *L, const axs_background *bg,
                            double lambda0, double tau, double dt, int nsteps,
                            double *J)
{
    AXS_CHECK_ARG(L && bg && J, "axs_forward_grad_energy: NULL argument");
    AXS_CHECK_ARG(nsteps >= 0 && nsteps <= kCoeffTableSteps,
                  "axs_forward_grad_energy: 0 <= nsteps <= 4096");
    AXS_CHECK_ARG(bg->flat_space || bg->a_inv > 1.0,
                  "axs_forward_grad_energy: a_inv must be > 1");
    std::vector<double> tab, gf;
    build_table(bg, lambda0, tau, dt, nsteps, tab, gf);
    hipStream_t s = L->ctx->stream;
    if (nsteps > 0)
    {
        AXS_HIP_TRY(hipMemcpyAsync(L->d_coeff, tab.data(),
                                   tab.size() * sizeof(double),
                                   hipMemcpyHostToDevice, s));
        AXS_HIP_TRY(hipStreamSynchronize(s));
    }
    for (int i = 0; i < nsteps; ++i)
        if (int rc = fwd_step(L, i, dt))
            return rc;
    double tau_end = tau;
    for (int i = 0; i < nsteps; ++i)
        tau_end += dt;
    Workspace ws;
    AXS_HIP_TRY(hipMalloc(&ws.d, (size_t)6 * L->g.ncell * sizeof(double)));
    return objective(L, final_R(bg, tau_end), J, ws.d, nullptr);
}

int axs_adjoint_grad_energy(axs_level *L, const axs_background *bg,
                            double lambda0, double tau, double dt, int nsteps,
                            int K, double *J, double *grad_state,
                            double *grad_lambda0)
{
    AXS_CHECK_ARG(L && bg && J && grad_state && grad_lambda0,
                  "axs_adjoint_grad_energy: NULL argument");
    AXS_CHECK_ARG(nsteps >= 1 && nsteps <= kCoeffTableSteps,
                  "axs_adjoint_grad_energy: 1 <= nsteps <= 4096");
    AXS_CHECK_ARG(K >= 1, "axs_adjoint_grad_energy: checkpoint_every >= 1");
    AXS_CHECK_ARG(bg->flat_space || bg->a_inv > 1.0,
                  "axs_adjoint_grad_energy: a_inv must be > 1");
    hipStream_t s = L->ctx->stream;
    const uint32_t n = L->g.ncell;
    const size_t n4 = (size_t)4 * n;
    const size_t sbytes = n4 * L->word;          // one stored state
    const size_t cbytes = L->comp ? n4 * 4 : 0;   // its compensation
    const int ncp = (nsteps + K - 1) / K;
    const int seg = std::min(K, nsteps);

    std::vector<double> tab, gf;
    build_table(bg, lambda0, tau, dt, nsteps, tab, gf);
    AXS_HIP_TRY(hipMemcpyAsync(L->d_coeff, tab.data(), tab.size() * sizeof(double),
                               hipMemcpyHostToDevice, s));
    AXS_HIP_TRY(hipStreamSynchronize(s));

    // device workspace: checkpoints, segment states, final state, psi of
    // stage 1, adjoint vectors (lp, acc, yA, yB) and the objective scratch
    const size_t st_bytes = sbytes + cbytes;
    const size_t need = (size_t)(ncp + seg + 1) * st_bytes + n4 * L->word +
                        (4 * n4 + 6 * (size_t)n) * sizeof(double);
    char *base = nullptr;
    AXS_HIP_TRY(hipMalloc(&base, need));
    struct Free
    {
        char *p;
        ~Free() { (void)hipFree(p); }
    } guard{base};
    char *cp = base;
    char *segs = cp + (size_t)ncp * st_bytes;
    char *fin = segs + (size_t)seg * st_bytes;
    char *y1 = fin + st_bytes;
    double *lp = (double *)(y1 + n4 * L->word);
    double *acc = lp + n4, *yA = acc + n4, *yB = yA + n4, *gbuf = yB + n4;

    auto save = [&](char *dst) -> int {
        AXS_HIP_TRY(hipMemcpyAsync(dst, L->y, sbytes, hipMemcpyDeviceToDevice, s));
        if (cbytes)
            AXS_HIP_TRY(hipMemcpyAsync(dst + sbytes, L->comp, cbytes,
                                       hipMemcpyDeviceToDevice, s));
        return AXS_OK;
    };
    auto load = [&](const char *src) -> int {
        AXS_HIP_TRY(hipMemcpyAsync(L->y, src, sbytes, hipMemcpyDeviceToDevice, s));
        if (cbytes)
            AXS_HIP_TRY(hipMemcpyAsync(L->comp, src + sbytes, cbytes,
                                       hipMemcpyDeviceToDevice, s));
        return AXS_OK;
    };

    // ---- forward with checkpoints every K steps ----
    int rc;
    for (int i = 0; i < nsteps; ++i)
    {
        if (i % K == 0 && (rc = save(cp + (size_t)(i / K) * st_bytes)))
            return rc;
        if ((rc = fwd_step(L, i, dt)))
            return rc;
    }
    if ((rc = save(fin)))
        return rc;
    double tau_end = tau;
    for (int i = 0; i < nsteps; ++i)
        tau_end += dt;
    if ((rc = objective(L, final_R(bg, tau_end), J, gbuf, lp)))
        return rc;

    // ---- reverse sweep, one checkpoint segment at a time ----
    AdjGrid g{L->g.nx, L->g.ny, L->g.nz, n, L->dx};
    const int nb = nblk(n);
    const bool f64 = L->precision == AXS_FP64;
    double dlam = 0.0;
    std::vector<double> hp(nb);
    for (int m = ncp - 1; m >= 0; --m)
    {
        const int i0 = m * K, i1 = std::min(nsteps, i0 + K);
        if ((rc = load(cp + (size_t)m * st_bytes)))
            return rc;
        for (int i = i0; i < i1; ++i)
        {
            if ((rc = save(segs + (size_t)(i - i0) * st_bytes)))
                return rc;
            if (i + 1 < i1 && (rc = fwd_step(L, i, dt)))
                return rc;
        }
        for (int i = i1 - 1; i >= i0; --i)
        {
            // recompute the stage states of step i from y_i
            if ((rc = load(segs + (size_t)(i - i0) * st_bytes)))
                return rc;
            const double *cf = L->d_coeff + (size_t)12 * i;
            if ((rc = launch_rk_stage(L, 1, L->y, L->ya, cf, nullptr, 0, dt)))
                return rc;
            AXS_HIP_TRY(hipMemcpyAsync(y1, L->ya, n4 * L->word,
                                       hipMemcpyDeviceToDevice, s));
            if ((rc = launch_rk_stage(L, 2, L->ya, L->yb, cf, nullptr, 1, dt)))
                return rc;
            if ((rc = launch_rk_stage(L, 3, L->yb, L->ya, cf, nullptr, 2, dt)))
                return rc;
            // acc = ybar'
            hipLaunchKernelGGL(k_copy_d, dim3(nblk(n4)), dim3(256), 0, s,
                               (const double *)lp, acc, n4);
            const double h = dt;
            // (stage input state, a, yprev, b, stage index)
            const void *ins[4] = {L->ya, L->yb, y1, L->y};
            const float *los[4] = {nullptr, nullptr, nullptr, L->comp};
            const double as[4] = {h / 6.0, h / 3.0, h / 3.0, h / 6.0};
            const double bs[4] = {0.0, h, 0.5 * h, 0.5 * h};
            double *outs[4] = {yA, yB, yA, yB};
            const double *prevs[4] = {nullptr, yA, yB, yA};
            for (int k = 0; k < 4; ++k)
            {
                const int stg = 3 - k; // stage 4, 3, 2, 1
                AdjStage a{};
                a.lp = lp;
                a.a = as[k];
                a.yprev = prevs[k];
                a.b = bs[k];
                a.cc = tab[(size_t)i * 12 + 3 * stg];
                a.lam = tab[(size_t)i * 12 + 3 * stg + 1];
                a.R2 = tab[(size_t)i * 12 + 3 * stg + 2];
                a.gfac = gf[(size_t)i * 4 + stg];
                a.yout = outs[k];
                a.acc = acc;
                a.partials = L->ctx->d_partials + (size_t)k * nb;
                if (f64)
                    hipLaunchKernelGGL(k_adj_stage<double>, dim3(nb), dim3(256),
                                       0, s, (const double *)ins[k], los[k], g, a);
                else
                    hipLaunchKernelGGL(k_adj_stage<float>, dim3(nb), dim3(256),
                                       0, s, (const float *)ins[k], los[k], g, a);
                AXS_HIP_TRY(hipGetLastError());
            }
            hipLaunchKernelGGL(k_sum_vec, dim3(1), dim3(256), 0, s,
                               (const double *)L->ctx->d_partials, 4 * nb,
                               L->ctx->d_result);
            AXS_HIP_TRY(hipMemcpyAsync(L->ctx->h_result, L->ctx->d_result,
                                       sizeof(double), hipMemcpyDeviceToHost, s));
            // ybar_i = acc
            hipLaunchKernelGGL(k_copy_d, dim3(nblk(n4)), dim3(256), 0, s,
                               (const double *)acc, lp, n4);
            AXS_HIP_TRY(hipStreamSynchronize(s));
            dlam += L->ctx->h_result[0];
        }
    }
    AXS_HIP_TRY(hipMemcpyAsync(grad_state, lp, n4 * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    if ((rc = load(fin)))
        return rc;
    AXS_HIP_TRY(hipStreamSynchronize(s));
    *grad_lambda0 = dlam;
    (void)hp;
    return AXS_OK;
}

} // extern "C"
