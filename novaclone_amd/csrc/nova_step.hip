// nova_step.hip -- fused Runge-Kutta stage kernels and the step driver.
//
// Replaces the app-level Field-algebra chains of the reference time loops
// (app/TG_Piosson/RKLoop.H:15-38, app/OT_Poisson/UEqn.H:14-39,
// app/OT_Poisson/BEqn.H:15-41, app/shearLayer/UEqn.H:14-38), each of which
// makes one pass per operator and per algebraic term over pool
// temporaries, with a single kernel per stage: the advection, viscous (and
// Lorentz/induction) right-hand side and the stage combination are
// evaluated per node and written once.
#include "nova_internal.hpp"

struct nova_poisson;

namespace nova {
int poisson_run(nova_poisson *ps, const double *src, int src_is_vector, double scale, double *p);
int project(nova_poisson *ps, double *F, double s, double *p);
nova_ctx *poisson_ctx(nova_poisson *ps);
}  // namespace nova

struct nova_solver {
    nova_poisson *ps = nullptr;
    nova_ctx *ctx = nullptr;
    int model = 0, mode = 0;
    double nu = 0.0, eta = 0.0;
    double *wa = nullptr, *wb = nullptr, *wc = nullptr, *wd = nullptr;  // stage buffers (3 comps each)
    double *div_tmp = nullptr;
    int use_graph = 0;
    // replayed step graph and the arguments it was recorded with
    hipStream_t gstream = nullptr;
    hipGraphExec_t gexec = nullptr;
    double gdt = 0.0;
    double *gU = nullptr, *gB = nullptr, *gp = nullptr, *gpB = nullptr;
};

namespace nova {

namespace {

// Stage combination forms.
enum { FORM_SSP = 0, FORM_LIT4 = 1, FORM_SHEAR = 2 };

struct StageArgs {
    int form;
    double alpha, beta;  // FORM_SSP: alpha*U0 + beta*(U + dU); FORM_SHEAR: U0 + beta*dU
    double dt;
    double w[4];         // FORM_LIT4 accumulation weights
};

__device__ inline double combine(const StageArgs &a, double u, double u0, double du)
{
    if (a.form == FORM_SSP) {
        const double v = u + du;
        return a.alpha != 0.0 ? a.alpha * u0 + a.beta * v : v;
    }
    if (a.form == FORM_LIT4) {
        double s = u;
        for (int j = 0; j < 4; j++) s = s + a.w[j] * du;
        return s;
    }
    return u0 + a.beta * du;
}

// Neighbour table of one node: mapped indices and reflection flags for
// offsets -H..H along each direction.
template <int H>
struct Nbr {
    int64_t off[3][2 * H + 1];  // linear offset contribution of each axis
    unsigned char flip[3][2 * H + 1];
};

template <int H>
__device__ inline void make_nbr(const Grid &g, int i, int j, int k, Nbr<H> &nb)
{
    for (int o = -H; o <= H; o++) {
        int f;
        int ii = bmap(i + o, g.n[0], g.wall[0], f);
        nb.off[0][o + H] = (int64_t)ii * g.s[1] * g.s[2];
        nb.flip[0][o + H] = (unsigned char)f;
        int jj = bmap(j + o, g.n[1], g.wall[1], f);
        nb.off[1][o + H] = (int64_t)jj * g.s[2];
        nb.flip[1][o + H] = (unsigned char)f;
        int kk = bmap(k + o, g.n[2], g.wall[2], f);
        nb.off[2][o + H] = kk;
        nb.flip[2][o + H] = (unsigned char)f;
    }
}

// Value of component c of a 3-component field at the node displaced by o
// along direction d.
template <int H>
__device__ inline double at(const double *__restrict__ f, const Grid &g, const Nbr<H> &nb, int c, int d, int o)
{
    const int64_t q = (d == 0 ? nb.off[0][o + H] : nb.off[0][H]) + (d == 1 ? nb.off[1][o + H] : nb.off[1][H]) +
                      (d == 2 ? nb.off[2][o + H] : nb.off[2][H]);
    const double v = f[c * g.np + q];
    return (nb.flip[d][o + H] && c == d) ? -v : v;
}

// div(X, Y_c): 0.5 sum_d sum_l a [ (X_d Y_c)^+ - (X_d Y_c)^- + X_d (Y_c^+ - Y_c^-)
// + Y_c (X_d^+ - X_d^-) ]  (src/ex/div/div.hxx:101-169).
template <int H>
__device__ inline double div_xy(const Grid &g, const Nbr<H> &nb, const double *__restrict__ X,
                                const double *__restrict__ Y, int c, double zi)
{
    const double yc = at<H>(Y, g, nb, c, 0, 0);
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double xd = at<H>(X, g, nb, d, 0, 0);
        double ad = 0.0;
#pragma unroll
        for (int l = 0; l < H; l++) {
            const double xp = at<H>(X, g, nb, d, d, l + 1), xm = at<H>(X, g, nb, d, d, -l - 1);
            const double yp = at<H>(Y, g, nb, c, d, l + 1), ym = at<H>(Y, g, nb, c, d, -l - 1);
            ad += g.a[d][l] * ((yp * xp - ym * xm) + xd * (yp - ym) + yc * (xp - xm));
        }
        acc += d == 2 ? ad * zi : ad;
    }
    return 0.5 * acc;
}

// nu * laplacian(f_c) with the z-metric term
// (src/ex/laplacian/laplacian.hxx:140-188).
template <int H>
__device__ inline double lap_c(const Grid &g, const Nbr<H> &nb, const double *__restrict__ f, int c, double nu,
                               int k)
{
    const double f0 = at<H>(f, g, nb, c, 0, 0);
    double acc = 0.0, dz1 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double ad = 0.0;
#pragma unroll
        for (int l = 0; l < H; l++) {
            const double fp = at<H>(f, g, nb, c, d, l + 1), fm = at<H>(f, g, nb, c, d, -l - 1);
            ad += g.b[d][l] * ((fp + fm) - 2.0 * f0);
            if (d == 2) dz1 += g.a[2][l] * (fp - fm);
        }
        acc += d == 2 ? ad * g.zlap[k] : ad;
    }
    return nu * acc - nu * g.zmet[k] * dz1;
}

// d/dx_d of 0.5|B|^2 (app/OT_Poisson/UEqn.H:22).
template <int H>
__device__ inline double grad_pm(const Grid &g, const Nbr<H> &nb, const double *__restrict__ B, int d, double zi)
{
    double ad = 0.0;
#pragma unroll
    for (int l = 0; l < H; l++) {
        double bp = 0.0, bm = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double vp = at<H>(B, g, nb, c, d, l + 1), vm = at<H>(B, g, nb, c, d, -l - 1);
            bp += vp * vp;
            bm += vm * vm;
        }
        ad += g.a[d][l] * (0.5 * bp - 0.5 * bm);
    }
    return d == 2 ? ad * zi : ad;
}

// Incompressible NS stage: dU = dt*((-div(U,U)) + nu*lap(U)),
// out = combine(U, U0, dU).
template <int H>
__global__ void __launch_bounds__(256) k_ns_stage(Grid g, const double *__restrict__ U,
                                                  const double *U0, double *out, StageArgs sa,
                                                  double nu)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.s[2] || j >= g.s[1]) return;
    Nbr<H> nb;
    make_nbr<H>(g, i, j, k, nb);
    const double zi = g.zinv[k];
    const int64_t q = ((int64_t)i * g.s[1] + j) * g.s[2] + k;
    double res[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double du = sa.dt * ((-div_xy<H>(g, nb, U, U, c, zi)) + lap_c<H>(g, nb, U, c, nu, k));
        const double u = U[c * g.np + q];
        const double u0 = U0 ? U0[c * g.np + q] : u;
        res[c] = combine(sa, u, u0, du);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) out[c * g.np + q] = res[c];
}

// Incompressible MHD stage (app/OT_Poisson/UEqn.H:16-23, BEqn.H:15-23):
// dU = dt*((((-div(U,U)) + nu lap U) + div(B,B)) - grad(0.5 B.B)),
// dB = dt*(((-div(U,B)) + div(B,U)) + eta lap B).
template <int H>
__global__ void __launch_bounds__(256) k_mhd_stage(Grid g, const double *__restrict__ U,
                                                   const double *__restrict__ B, const double *U0,
                                                   const double *B0, double *outU, double *outB,
                                                   StageArgs sa, double nu, double eta)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.s[2] || j >= g.s[1]) return;
    Nbr<H> nb;
    make_nbr<H>(g, i, j, k, nb);
    const double zi = g.zinv[k];
    const int64_t q = ((int64_t)i * g.s[1] + j) * g.s[2] + k;
    double ru[3], rb[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double du = sa.dt * ((((-div_xy<H>(g, nb, U, U, c, zi)) + lap_c<H>(g, nb, U, c, nu, k)) +
                                    div_xy<H>(g, nb, B, B, c, zi)) - grad_pm<H>(g, nb, B, c, zi));
        const double db = sa.dt * (((-div_xy<H>(g, nb, U, B, c, zi)) + div_xy<H>(g, nb, B, U, c, zi)) +
                                   lap_c<H>(g, nb, B, c, eta, k));
        const double u = U[c * g.np + q], b = B[c * g.np + q];
        ru[c] = combine(sa, u, U0 ? U0[c * g.np + q] : u, du);
        rb[c] = combine(sa, b, B0 ? B0[c * g.np + q] : b, db);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        outU[c * g.np + q] = ru[c];
        outB[c * g.np + q] = rb[c];
    }
}

template <int H>
int launch_ns_stage_h(nova_ctx *ctx, const double *U, const double *U0, double *out, const StageArgs &sa, double nu)
{
    const Grid &g = ctx->g;
    dim3 grid((g.s[2] + 63) / 64, (g.s[1] + 3) / 4, g.s[0]);
    TimedScope ts(ctx, "ns_stage");
    hipLaunchKernelGGL(k_ns_stage<H>, grid, dim3(64, 4, 1), 0, ctx->stream, g, U, U0, out, sa, nu);
    return check_launch("ns_stage");
}

template <int H>
int launch_mhd_stage_h(nova_ctx *ctx, const double *U, const double *B, const double *U0, const double *B0,
                       double *oU, double *oB, const StageArgs &sa, double nu, double eta)
{
    const Grid &g = ctx->g;
    dim3 grid((g.s[2] + 63) / 64, (g.s[1] + 3) / 4, g.s[0]);
    TimedScope ts(ctx, "mhd_stage");
    hipLaunchKernelGGL(k_mhd_stage<H>, grid, dim3(64, 4, 1), 0, ctx->stream, g, U, B, U0, B0, oU, oB, sa, nu, eta);
    return check_launch("mhd_stage");
}

int ns_stage(nova_ctx *ctx, const double *U, const double *U0, double *out, const StageArgs &sa, double nu)
{
    switch (ctx->g.h) {
    case 1: return launch_ns_stage_h<1>(ctx, U, U0, out, sa, nu);
    case 2: return launch_ns_stage_h<2>(ctx, U, U0, out, sa, nu);
    case 3: return launch_ns_stage_h<3>(ctx, U, U0, out, sa, nu);
    case 4: return launch_ns_stage_h<4>(ctx, U, U0, out, sa, nu);
    }
    return fail(NOVA_ERR_ARG, "stencil order");
}

int mhd_stage(nova_ctx *ctx, const double *U, const double *B, const double *U0, const double *B0, double *oU,
              double *oB, const StageArgs &sa, double nu, double eta)
{
    switch (ctx->g.h) {
    case 1: return launch_mhd_stage_h<1>(ctx, U, B, U0, B0, oU, oB, sa, nu, eta);
    case 2: return launch_mhd_stage_h<2>(ctx, U, B, U0, B0, oU, oB, sa, nu, eta);
    case 3: return launch_mhd_stage_h<3>(ctx, U, B, U0, B0, oU, oB, sa, nu, eta);
    case 4: return launch_mhd_stage_h<4>(ctx, U, B, U0, B0, oU, oB, sa, nu, eta);
    }
    return fail(NOVA_ERR_ARG, "stencil order");
}

const double kRK4W[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};  // settings.cpp:49
const double kRK4C[4] = {0.5, 0.5, 1.0, 1.0};                          // settings.cpp:51

StageArgs ssp(int s, double dt)
{
    StageArgs a{};
    a.form = FORM_SSP;
    a.dt = dt;
    if (s == 0) { a.alpha = 0.0; a.beta = 1.0; }
    else if (s == 1) { a.alpha = 0.75; a.beta = 0.25; }
    else { a.alpha = 1.0 / 3.0; a.beta = 2.0 / 3.0; }
    return a;
}

int copy3(nova_ctx *ctx, double *dst, const double *src)
{
    TimedScope ts(ctx, "copy");
    NOVA_HIP(hipMemcpyAsync(dst, src, sizeof(double) * 3 * ctx->g.np, hipMemcpyDeviceToDevice, ctx->stream));
    return NOVA_OK;
}

int ns_step_once(nova_solver *sv, double dt, double *U, double *p)
{
    nova_ctx *ctx = sv->ctx;
    int rc;
    if (sv->mode == NOVA_RK_SSP3) {
        if ((rc = ns_stage(ctx, U, nullptr, sv->wa, ssp(0, dt), sv->nu))) return rc;
        if ((rc = ns_stage(ctx, sv->wa, U, sv->wb, ssp(1, dt), sv->nu))) return rc;
        if ((rc = ns_stage(ctx, sv->wb, U, U, ssp(2, dt), sv->nu))) return rc;
        return project(sv->ps, U, dt, p);
    }
    if (sv->mode == NOVA_RK_LITERAL4) {
        StageArgs a{};
        a.form = FORM_LIT4;
        a.dt = dt;
        for (int j = 0; j < 4; j++) a.w[j] = kRK4W[j];
        if ((rc = ns_stage(ctx, U, nullptr, sv->wa, a, sv->nu))) return rc;
        if ((rc = project(sv->ps, sv->wa, dt, p))) return rc;
        return copy3(ctx, U, sv->wa);
    }
    // NOVA_RK_SHEAR4
    if ((rc = copy3(ctx, sv->wb, U))) return rc;
    for (int rk = 0; rk < 4; rk++) {
        StageArgs a{};
        a.form = FORM_SHEAR;
        a.dt = dt;
        a.beta = kRK4W[rk];
        if ((rc = ns_stage(ctx, U, sv->wb, sv->wa, a, sv->nu))) return rc;
        if ((rc = project(sv->ps, sv->wa, dt * kRK4C[rk], p))) return rc;
        if ((rc = copy3(ctx, U, sv->wa))) return rc;
    }
    return NOVA_OK;
}

int mhd_step_once(nova_solver *sv, double dt, double *U, double *B, double *p, double *pB)
{
    nova_ctx *ctx = sv->ctx;
    int rc;
    if (sv->mode != NOVA_RK_SSP3) return fail(NOVA_ERR_UNSUPPORTED, "MHD supports the SSP-RK3 composition");
    // stage outputs: (wa, wb) after stage 1, (wc, wd) after stage 2, (U, B) after stage 3
    if ((rc = mhd_stage(ctx, U, B, nullptr, nullptr, sv->wa, sv->wb, ssp(0, dt), sv->nu, sv->eta))) return rc;
    if ((rc = project(sv->ps, sv->wa, dt, p)) || (rc = project(sv->ps, sv->wb, 1.0, pB))) return rc;
    if ((rc = mhd_stage(ctx, sv->wa, sv->wb, U, B, sv->wc, sv->wd, ssp(1, dt), sv->nu, sv->eta))) return rc;
    if ((rc = project(sv->ps, sv->wc, dt, p)) || (rc = project(sv->ps, sv->wd, 1.0, pB))) return rc;
    if ((rc = mhd_stage(ctx, sv->wc, sv->wd, U, B, U, B, ssp(2, dt), sv->nu, sv->eta))) return rc;
    if ((rc = project(sv->ps, U, dt, p)) || (rc = project(sv->ps, B, 1.0, pB))) return rc;
    return NOVA_OK;
}

}  // namespace

int solver_create(nova_poisson *ps, int model, int mode, double nu, double eta, nova_solver **out)
{
    nova_ctx *ctx = poisson_ctx(ps);
    if (model != NOVA_MODEL_NS && model != NOVA_MODEL_MHD) return fail(NOVA_ERR_ARG, "unknown model");
    if (mode < 0 || mode > 2) return fail(NOVA_ERR_ARG, "unknown RK composition");
    nova_solver *sv = new nova_solver();
    sv->ps = ps;
    sv->ctx = ctx;
    sv->model = model;
    sv->mode = mode;
    sv->nu = nu;
    sv->eta = eta;
    const size_t bytes = sizeof(double) * 3 * (size_t)ctx->g.np;
    hipError_t e = hipMalloc((void **)&sv->wa, bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&sv->wb, bytes);
    if (e == hipSuccess && model == NOVA_MODEL_MHD) e = hipMalloc((void **)&sv->wc, bytes);
    if (e == hipSuccess && model == NOVA_MODEL_MHD) e = hipMalloc((void **)&sv->wd, bytes);
    if (e != hipSuccess) {
        if (sv->wa) (void)hipFree(sv->wa);
        if (sv->wb) (void)hipFree(sv->wb);
        if (sv->wc) (void)hipFree(sv->wc);
        delete sv;
        return hip_fail(e, "solver work buffers");
    }
    *out = sv;
    return NOVA_OK;
}

int solver_destroy(nova_solver *sv)
{
    if (!sv) return NOVA_OK;
    if (sv->gexec) (void)hipGraphExecDestroy(sv->gexec);
    if (sv->gstream) (void)hipStreamDestroy(sv->gstream);
    if (sv->wa) (void)hipFree(sv->wa);
    if (sv->wb) (void)hipFree(sv->wb);
    if (sv->wc) (void)hipFree(sv->wc);
    if (sv->wd) (void)hipFree(sv->wd);
    if (sv->div_tmp) (void)hipFree(sv->div_tmp);
    delete sv;
    return NOVA_OK;
}

static int step_once(nova_solver *sv, double dt, double *U, double *B, double *p, double *pB)
{
    return sv->model == NOVA_MODEL_NS ? ns_step_once(sv, dt, U, p) : mhd_step_once(sv, dt, U, B, p, pB);
}

int solver_use_graph(nova_solver *sv, int enable)
{
    sv->use_graph = enable != 0;
    return NOVA_OK;
}

// Record one step on a capture stream and instantiate it; the graph is
// re-recorded when dt or a state pointer changes.
static int record_step(nova_solver *sv, double dt, double *U, double *B, double *p, double *pB)
{
    nova_ctx *ctx = sv->ctx;
    if (!sv->gstream) NOVA_HIP(hipStreamCreate(&sv->gstream));
    if (sv->gexec) {
        NOVA_HIP(hipGraphExecDestroy(sv->gexec));
        sv->gexec = nullptr;
    }
    hipStream_t saved = ctx->stream;
    ctx->stream = sv->gstream;
    NOVA_HIP(hipStreamBeginCapture(sv->gstream, hipStreamCaptureModeThreadLocal));
    int rc = step_once(sv, dt, U, B, p, pB);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(sv->gstream, &graph);
    ctx->stream = saved;
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&sv->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    sv->gdt = dt;
    sv->gU = U;
    sv->gB = B;
    sv->gp = p;
    sv->gpB = pB;
    return NOVA_OK;
}

int solver_step(nova_solver *sv, double dt, int nsteps, double *U, double *B, double *p, double *pB)
{
    if (sv->model == NOVA_MODEL_MHD && (!B || !pB)) return fail(NOVA_ERR_ARG, "MHD needs B and pB");
    if (!sv->use_graph) {
        for (int s = 0; s < nsteps; s++)
            if (int rc = step_once(sv, dt, U, B, p, pB)) return rc;
        return NOVA_OK;
    }
    if (nsteps == 0) return NOVA_OK;
    nova_ctx *ctx = sv->ctx;
    if (!sv->gexec || sv->gdt != dt || sv->gU != U || sv->gB != B || sv->gp != p || sv->gpB != pB)
        if (int rc = record_step(sv, dt, U, B, p, pB)) return rc;
    // order the replays after the caller's stream and the caller after them
    hipEvent_t ev;
    NOVA_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    NOVA_HIP(hipEventRecord(ev, ctx->stream));
    NOVA_HIP(hipStreamWaitEvent(sv->gstream, ev, 0));
    for (int s = 0; s < nsteps; s++) NOVA_HIP(hipGraphLaunch(sv->gexec, sv->gstream));
    NOVA_HIP(hipEventRecord(ev, sv->gstream));
    NOVA_HIP(hipStreamWaitEvent(ctx->stream, ev, 0));
    NOVA_HIP(hipEventDestroy(ev));
    return NOVA_OK;
}

}  // namespace nova
