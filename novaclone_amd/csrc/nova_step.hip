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
#include "nova_stencil.hpp"
#include "nova_stream.hpp"

struct nova_poisson;

namespace nova {
int poisson_run(nova_poisson *ps, const double *src, int src_is_vector, double scale, double *p);
int poisson_scheme(nova_poisson *ps);
int compact_ns_stage(nova_ctx *ctx, const double *U, const double *U0, double *out, const StageArgs &sa, double nu,
                     double *R);
int compact_mhd_stage(nova_ctx *ctx, const double *U, const double *B, const double *U0, const double *B0,
                      double *oU, double *oB, const StageArgs &sa, double nu, double eta, double *R);
int project(nova_poisson *ps, double *F, double s, double *p);
nova_ctx *poisson_ctx(nova_poisson *ps);
int mhd_u_stage(nova_ctx *ctx, const double *U, const double *B, double *out, const StageArgs &sa, double nu);
int mhd_b_stage(nova_ctx *ctx, const double *U, const double *B, const double *B0, double *acc, double *out,
                double beta, double w, int first, int last, double dt, double eta);
}  // namespace nova

struct nova_solver {
    nova_poisson *ps = nullptr;
    nova_ctx *ctx = nullptr;
    int model = 0, mode = 0;
    double nu = 0.0, eta = 0.0;
    double *wa = nullptr, *wb = nullptr, *wc = nullptr, *wd = nullptr;  // stage buffers (3 comps each)
    double *div_tmp = nullptr;
    const double *force = nullptr;  // constant right-hand-side term (nova_solver_set_force)
    int use_graph = 0;
    int scheme = NOVA_SCHEME_EXPLICIT;
    double *rwork = nullptr;  // compact right-hand-side accumulator (3 comps, 6 for MHD)
    // replayed step graph and the arguments it was recorded with
    hipStream_t gstream = nullptr;
    hipGraphExec_t gexec = nullptr;
    double gdt = 0.0;
    int gepoch = 0;  // launch-shape epoch of the context when recorded
    double *gU = nullptr, *gB = nullptr, *gp = nullptr, *gpB = nullptr;
};

namespace nova {

namespace {


template <int H, int NF, bool WALLS>
__global__ void __launch_bounds__(kTK *kTJ) k_stage_rows(Grid g, StreamArgs<NF> a)
{
    const int T = a.nkt * a.njt * a.nic;
    int t = blockIdx.x;
    if (T % 8 == 0) t = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
    const int kt = t % a.nkt;
    const int jt = (t / a.nkt) % a.njt;
    const int ic = t / (a.nkt * a.njt);
    const int k = kt * kTK + threadIdx.x, j = jt * kTJ + threadIdx.y;
    if (k >= g.s[2] || j >= g.s[1]) return;  // no barriers below
    const int i0 = ic * a.ichunk;
    int i1 = i0 + a.ichunk;
    if (i1 > g.s[0]) i1 = g.s[0];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    const int s2 = g.s[2];
    const int col = j * s2 + k;

    auto plane_off = [&](int ii, int &fi) -> int64_t {
        const int im = wrap<WALLS>(ii, g.n[0], g.wall[0], fi);
        return (int64_t)im * plane;
    };

    double q[NF][3][2 * H + 1];
#pragma unroll
    for (int o = 0; o < 2 * H + 1; o++) {
        int fi;
        const int64_t po = plane_off(i0 - H + o, fi);
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int c = 0; c < 3; c++) q[f][c][o] = rsign<WALLS>(g, c, 0, fi) * a.X[f][c * g.np + po + col];
    }
    const double zi = g.zinv[k], zl = g.zlap[k], zm = g.zmet[k];
    NsZ<H> nz{};
    if (NF == 1) nz = ns_z<H>(a.nk, zi, zl, zm);

    for (int i = i0; i < i1; i++) {
        double nq[NF][3];
        if (i + 1 < i1) {
            int fi;
            const int64_t po = plane_off(i + H + 1, fi);
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) nq[f][c] = rsign<WALLS>(g, c, 0, fi) * a.X[f][c * g.np + po + col];
        }
        const int64_t pi = (int64_t)i * plane;
        auto A = [&](int f, int c, int d, int o) -> double {
            if (d == 0 || o == 0) return q[f][c][H + (d == 0 ? o : 0)];
            int fl;
            int64_t off;
            if (d == 1) off = pi + (int64_t)wrap<WALLS>(j + o, g.n[1], g.wall[1], fl) * s2 + k;
            else off = pi + (int64_t)j * s2 + wrap<WALLS>(k + o, g.n[2], g.wall[2], fl);
            return rsign<WALLS>(g, c, d, fl) * a.X[f][c * g.np + off];
        };
        const int64_t qi = pi + col;
        const StageArgs &sa = a.sa;
        if (NF == 1) {
            double r[3];
            ns_du<H>(a.nk, nz, A, r);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double u = q[0][c][H];
                const double u0 = a.X0[0] ? a.X0[0][c * g.np + qi] : u;
                a.out[0][c * g.np + qi] = combine(sa, u, u0, r[c]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double du = sa.dt * ((((-div_xy<H>(g, A, 0, 0, c, zi)) + lap_c<H>(g, A, 0, c, a.nu, zl, zm)) +
                                            div_xy<H>(g, A, 1, 1, c, zi)) - grad_pm<H>(g, A, 1, c, zi));
                const double db = sa.dt * (((-div_xy<H>(g, A, 0, 1, c, zi)) + div_xy<H>(g, A, 1, 0, c, zi)) +
                                           lap_c<H>(g, A, 1, c, a.eta, zl, zm));
                const double u = q[0][c][H], b = q[NF - 1][c][H];
                const double u0 = a.X0[0] ? a.X0[0][c * g.np + qi] : u;
                const double b0 = a.X0[NF - 1] ? a.X0[NF - 1][c * g.np + qi] : b;
                a.out[0][c * g.np + qi] = combine(sa, u, u0, du);
                a.out[NF - 1][c * g.np + qi] = combine(sa, b, b0, db);
            }
        }
        if (i + 1 < i1) {
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
#pragma unroll
                    for (int o = 0; o < 2 * H; o++) q[f][c][o] = q[f][c][o + 1];
                    q[f][c][2 * H] = nq[f][c];
                }
        }
    }
}


template <int H, int NF>
int launch_stream_h(nova_ctx *ctx, StreamArgs<NF> a, const char *name)
{
    const Grid &g = ctx->g;
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    TimedScope ts(ctx, name);
    constexpr int MODEL = NF == 1 ? MODEL_NS : MODEL_MHD;
    // sixth-order NS stages compile the SSP combination in (the MHD stages,
    // at one wave per SIMD, spill when their branches go); the first stage
    // (combination 2) has a launch-shape slot of its own
    const int cmb = (H == 3 && MODEL == MODEL_NS && a.sa.form == FORM_SSP)
                        ? (a.sa.alpha == 0.0 ? 2 : (a.X0[0] ? 1 : 0))
                        : 0;
    const int slot = cmb == 2 ? 4 : MODEL;
    if (ctx->stage_variant[slot] == 1) {
        a.nkt = (g.s[2] + kTK - 1) / kTK;
        a.njt = (g.s[1] + kTJ - 1) / kTJ;
        const int ic = ctx->stage_ichunk[slot];
        a.ichunk = ic > 0 ? (ic < g.s[0] ? ic : g.s[0]) : pick_ichunk(g, a.nkt, a.njt);
        a.nic = (g.s[0] + a.ichunk - 1) / a.ichunk;
        const int T = a.nkt * a.njt * a.nic;
        if (walls)
            hipLaunchKernelGGL((k_stage_rows<H, NF, true>), dim3(T), dim3(kTK, kTJ, 1), 0, ctx->stream, g, a);
        else
            hipLaunchKernelGGL((k_stage_rows<H, NF, false>), dim3(T), dim3(kTK, kTJ, 1), 0, ctx->stream, g, a);
    } else if constexpr (H == 3 && MODEL == MODEL_NS) {
        if (walls) {
            if (cmb == 1) launch_stream_w<H, MODEL, true, 1>(ctx, a);
            else if (cmb == 2) launch_stream_w<H, MODEL, true, 2>(ctx, a);
            else launch_stream_w<H, MODEL, true>(ctx, a);
        } else {
            if (cmb == 1) launch_stream_w<H, MODEL, false, 1>(ctx, a);
            else if (cmb == 2) launch_stream_w<H, MODEL, false, 2>(ctx, a);
            else launch_stream_w<H, MODEL, false>(ctx, a);
        }
    } else if (walls) {
        launch_stream_w<H, MODEL, true>(ctx, a);
    } else {
        launch_stream_w<H, MODEL, false>(ctx, a);
    }
    return check_launch(name);
}

template <int NF>
int launch_stream(nova_ctx *ctx, const StreamArgs<NF> &a, const char *name)
{
    switch (ctx->g.h) {
    case 1: return launch_stream_h<1, NF>(ctx, a, name);
    case 2: return launch_stream_h<2, NF>(ctx, a, name);
    case 3: return launch_stream_h<3, NF>(ctx, a, name);
    case 4: return launch_stream_h<4, NF>(ctx, a, name);
    }
    return fail(NOVA_ERR_ARG, "stencil order");
}

int ns_stage(nova_ctx *ctx, const double *U, const double *U0, double *out, const StageArgs &sa, double nu)
{
    StreamArgs<1> a{};
    a.X[0] = U;
    a.X0[0] = U0;
    a.out[0] = out;
    a.sa = sa;
    a.nu = nu;
    a.nk = ns_k(ctx->g, nu, sa.dt);
    return launch_stream<1>(ctx, a, "ns_stage");
}

int mhd_stage(nova_ctx *ctx, const double *U, const double *B, const double *U0, const double *B0, double *oU,
              double *oB, const StageArgs &sa, double nu, double eta)
{
    StreamArgs<2> a{};
    a.X[0] = U;
    a.X[1] = B;
    a.X0[0] = U0;
    a.X0[1] = B0;
    a.out[0] = oU;
    a.out[1] = oB;
    a.sa = sa;
    a.nu = nu;
    a.eta = eta;
    return launch_stream<2>(ctx, a, "mhd_stage");
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
    if (sv->scheme == NOVA_SCHEME_COMPACT6) {
        if ((rc = compact_ns_stage(ctx, U, nullptr, sv->wa, ssp(0, dt), sv->nu, sv->rwork))) return rc;
        if ((rc = compact_ns_stage(ctx, sv->wa, U, sv->wb, ssp(1, dt), sv->nu, sv->rwork))) return rc;
        if ((rc = compact_ns_stage(ctx, sv->wb, U, U, ssp(2, dt), sv->nu, sv->rwork))) return rc;
        return project(sv->ps, U, dt, p);
    }
    // one stage: the stage kernel, the constant force with the weight the
    // stage gives its increment, the reference's wall fill of the result
    auto stage = [&](const double *X, const double *X0, double *out, const StageArgs &a, double fw) -> int {
        if (int r = ns_stage(ctx, X, X0, out, a, sv->nu)) return r;
        if (sv->force)
            if (int r = launch_axpby(ctx, 3 * (size_t)ctx->g.np, 1.0, out, fw * dt, sv->force, out)) return r;
        return wall_fill(ctx, out, 3);
    };
    auto proj = [&](double *F, double s) -> int {
        if (int r = project(sv->ps, F, s, p)) return r;
        return wall_fill(ctx, F, 3);  // U.correctBoundaryConditions()
    };
    if (sv->mode == NOVA_RK_SSP3) {
        if ((rc = stage(U, nullptr, sv->wa, ssp(0, dt), 1.0))) return rc;
        if ((rc = stage(sv->wa, U, sv->wb, ssp(1, dt), 0.25))) return rc;
        if ((rc = stage(sv->wb, U, U, ssp(2, dt), 2.0 / 3.0))) return rc;
        return proj(U, dt);
    }
    if (sv->mode == NOVA_RK_LITERAL4) {
        StageArgs a{};
        a.form = FORM_LIT4;
        a.dt = dt;
        for (int j = 0; j < 4; j++) a.w[j] = kRK4W[j];
        if ((rc = stage(U, nullptr, sv->wa, a, ((kRK4W[0] + kRK4W[1]) + kRK4W[2]) + kRK4W[3]))) return rc;
        if ((rc = proj(sv->wa, dt))) return rc;
        return copy3(ctx, U, sv->wa);
    }
    // NOVA_RK_SHEAR4
    if ((rc = copy3(ctx, sv->wb, U))) return rc;
    for (int rk = 0; rk < 4; rk++) {
        StageArgs a{};
        a.form = FORM_SHEAR;
        a.dt = dt;
        a.beta = kRK4W[rk];
        if ((rc = stage(U, sv->wb, sv->wa, a, kRK4W[rk]))) return rc;
        if ((rc = proj(sv->wa, dt * kRK4C[rk]))) return rc;
        if ((rc = copy3(ctx, U, sv->wa))) return rc;
    }
    return NOVA_OK;
}


int mhd_step_once(nova_solver *sv, double dt, double *U, double *B, double *p, double *pB)
{
    nova_ctx *ctx = sv->ctx;
    int rc;
    if (ctx->wall_fill) return fail(NOVA_ERR_UNSUPPORTED, "the reference's wall fills are set up for the NS model");
    if (sv->mode == NOVA_RK_LITERAL4) {
        // app/OT_Poisson/UEqn.H:14-40 then BEqn.H:14-41 (app/mhdjet alike)
        if (sv->scheme != NOVA_SCHEME_EXPLICIT)
            return fail(NOVA_ERR_UNSUPPORTED, "the reference's MHD composition uses the explicit stencils");
        StageArgs a{};
        a.form = FORM_LIT4;
        a.dt = dt;
        for (int j = 0; j < 4; j++) a.w[j] = kRK4W[j];
        if ((rc = mhd_u_stage(ctx, U, B, sv->wa, a, sv->nu))) return rc;
        if ((rc = project(sv->ps, sv->wa, dt, p))) return rc;
        if ((rc = copy3(ctx, U, sv->wa))) return rc;
        // classic RK4 of B with the new U frozen: stage rk writes
        // B0 + RK4a[rk][rk] dB_rk (the earlier entries of the row are zero)
        // and adds RK4a[3][rk] dB_rk to the running sum in wa; the last stage
        // writes the sum into B, then B is cleaned (s = 1)
        static const double kDiag[3] = {0.5, 0.5, 1.0};  // settings.cpp:44-48
        const double *in[4] = {B, sv->wb, sv->wc, sv->wd};
        double *out[4] = {sv->wb, sv->wc, sv->wd, B};
        for (int rk = 0; rk < 4; rk++)
            if ((rc = mhd_b_stage(ctx, U, in[rk], B, sv->wa, out[rk], rk < 3 ? kDiag[rk] : 0.0, kRK4W[rk], rk == 0,
                                  rk == 3, dt, sv->eta)))
                return rc;
        return project(sv->ps, B, 1.0, pB);
    }
    if (sv->mode != NOVA_RK_SSP3)
        return fail(NOVA_ERR_UNSUPPORTED, "MHD supports the SSP-RK3 and the reference's (LITERAL4) compositions");
    if (sv->scheme == NOVA_SCHEME_COMPACT6) {
        double *R = sv->rwork;
        if ((rc = compact_mhd_stage(ctx, U, B, nullptr, nullptr, sv->wa, sv->wb, ssp(0, dt), sv->nu, sv->eta, R)))
            return rc;
        if ((rc = project(sv->ps, sv->wa, dt, p)) || (rc = project(sv->ps, sv->wb, 1.0, pB))) return rc;
        if ((rc = compact_mhd_stage(ctx, sv->wa, sv->wb, U, B, sv->wc, sv->wd, ssp(1, dt), sv->nu, sv->eta, R)))
            return rc;
        if ((rc = project(sv->ps, sv->wc, dt, p)) || (rc = project(sv->ps, sv->wd, 1.0, pB))) return rc;
        if ((rc = compact_mhd_stage(ctx, sv->wc, sv->wd, U, B, U, B, ssp(2, dt), sv->nu, sv->eta, R))) return rc;
        if ((rc = project(sv->ps, U, dt, p)) || (rc = project(sv->ps, B, 1.0, pB))) return rc;
        return NOVA_OK;
    }
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

int stage_ns(nova_ctx *ctx, const double *U, const double *U0, double *out, const StageArgs &sa, double nu)
{
    return ns_stage(ctx, U, U0, out, sa, nu);
}

nova_poisson *solver_poisson(nova_solver *sv) { return sv->ps; }
int solver_scheme(nova_solver *sv) { return sv->scheme; }

int solver_model(nova_solver *sv, int *mode, double *nu)
{
    if (sv->model != NOVA_MODEL_NS) return fail(NOVA_ERR_UNSUPPORTED, "the adjoint covers the NS model");
    *mode = sv->mode;
    *nu = sv->nu;
    return NOVA_OK;
}

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
    if (sv->rwork) (void)hipFree(sv->rwork);
    delete sv;
    return NOVA_OK;
}

static int step_once(nova_solver *sv, double dt, double *U, double *B, double *p, double *pB)
{
    return sv->model == NOVA_MODEL_NS ? ns_step_once(sv, dt, U, p) : mhd_step_once(sv, dt, U, B, p, pB);
}

// Measured choice of the solver's stage-kernel launch shape and of the
// projection variant, on the caller's fields (which are not modified).
int solver_autotune(nova_solver *sv, const double *U, const double *B, int reps, char *report, size_t len)
{
    nova_ctx *ctx = sv->ctx;
    const bool mhd = sv->model == NOVA_MODEL_MHD;
    if (mhd && !B) return fail(NOVA_ERR_ARG, "MHD needs B");
    if (reps < 1) return fail(NOVA_ERR_ARG, "reps must be >= 1");
    std::string rep;
    // a middle stage (with the step-start operand, as two of the three
    // stages of a step have), its step-start operand a separate copy of the
    // input so that a trial moves the bytes a step's stage moves (an
    // operand read twice from one array would mostly hit the cache)
    const StageArgs sa = ssp(1, 1e-3);
    const size_t vb = sizeof(double) * 3 * ctx->g.np;
    NOVA_HIP(hipMemcpyAsync(sv->wb, U, vb, hipMemcpyDeviceToDevice, ctx->stream));
    if (mhd) NOVA_HIP(hipMemcpyAsync(sv->wd, B, vb, hipMemcpyDeviceToDevice, ctx->stream));
    // the staging variants 9-12 stay out of the search: in a full step they
    // measure level with 7 and 8, while a single stage timed on its own
    // (warm in the last-level cache) can favour them
    const int variants[] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    const int variants_mhd[] = {0, 1, 2, 3, 4, 11, 12};
    auto run = [&]() -> int {
        if (mhd) return mhd_stage(ctx, U, B, sv->wb, sv->wd, sv->wa, sv->wc, sa, sv->nu, sv->eta);
        return ns_stage(ctx, U, sv->wb, sv->wa, sa, sv->nu);
    };
    if (int rc = tune_stage_model(ctx, mhd ? MODEL_MHD : MODEL_NS, mhd ? variants_mhd : variants, mhd ? 7 : 9, reps, run,
                                  mhd ? "mhd_stage" : "ns_stage", rep))
        return rc;
    if (!mhd && ctx->g.h == 3) {
        // the first stage (no step-start operand) in its own slot; the
        // staging variants 11 and 12 join the search here
        const int first[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 11, 12};
        auto run_first = [&]() -> int { return ns_stage(ctx, U, nullptr, sv->wa, ssp(0, 1e-3), sv->nu); };
        if (int rc = tune_stage_model(ctx, 4, first, 11, reps, run_first, "ns_first_stage", rep)) return rc;
    }
    // projection variants, on a copy of U
    NOVA_HIP(hipMemcpyAsync(sv->wa, U, sizeof(double) * 3 * ctx->g.np, hipMemcpyDeviceToDevice, ctx->stream));
    double best = 1e300;
    int bp = ctx->proj_variant;
    for (int pv = 0; pv <= 2; pv++) {
        ctx->proj_variant = pv;
        double us = 0.0;
        if (int rc = time_avg_us(ctx, reps, [&]() { return project(sv->ps, sv->wa, 1.0, nullptr); }, &us)) {
            ctx->proj_variant = bp;
            return rc;
        }
        rep += "projection variant " + std::to_string(pv) + " " + std::to_string(us) + " us\n";
        if (us < best) {
            best = us;
            bp = pv;
        }
    }
    ctx->proj_variant = bp;
    ctx->tune_epoch++;
    rep += "projection chosen variant " + std::to_string(bp) + "\n";
    return write_report(rep, report, len);
}

int solver_set_force(nova_solver *sv, const double *F)
{
    if (F && sv->model != NOVA_MODEL_NS) return fail(NOVA_ERR_UNSUPPORTED, "a constant force is an NS model term");
    if (F && sv->scheme != NOVA_SCHEME_EXPLICIT) return fail(NOVA_ERR_UNSUPPORTED, "a constant force needs the explicit scheme");
    sv->force = F;
    sv->ctx->tune_epoch++;  // a recorded step graph lacks (or holds) the force
    return NOVA_OK;
}

int solver_set_scheme(nova_solver *sv, int scheme)
{
    if (scheme != NOVA_SCHEME_EXPLICIT && scheme != NOVA_SCHEME_COMPACT6) return fail(NOVA_ERR_ARG, "unknown scheme");
    if (scheme == NOVA_SCHEME_COMPACT6) {
        if (sv->mode != NOVA_RK_SSP3) return fail(NOVA_ERR_UNSUPPORTED, "the compact scheme covers SSP-RK3");
        if (sv->force) return fail(NOVA_ERR_UNSUPPORTED, "a constant force needs the explicit scheme");
        if (int rc = compact_supported(sv->ctx)) return rc;
        const int nc = sv->model == NOVA_MODEL_MHD ? 6 : 3;
        if (!sv->rwork) NOVA_HIP(hipMalloc((void **)&sv->rwork, sizeof(double) * nc * sv->ctx->g.np));
    }
    if (poisson_scheme(sv->ps) != scheme)
        return fail(NOVA_ERR_ARG, "set the Poisson solver to the same scheme first (nova_poisson_set_scheme)");
    sv->scheme = scheme;
    if (sv->gexec) {
        NOVA_HIP(hipGraphExecDestroy(sv->gexec));
        sv->gexec = nullptr;
    }
    return NOVA_OK;
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
    sv->gepoch = sv->ctx->tune_epoch;
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
    if (!sv->gexec || sv->gdt != dt || sv->gU != U || sv->gB != B || sv->gp != p || sv->gpB != pB ||
        sv->gepoch != ctx->tune_epoch)
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
