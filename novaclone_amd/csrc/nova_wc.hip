// nova_wc.hip -- the weakly-compressible step (app/TG/TG.cpp:71-76 with
// rhoEqn.H, pEqn.H and UEqn.H) on gfx950.
//
// The reference makes, per stage, one pass per operator and per algebraic
// term over pool temporaries: rho*U, ex::div, the negation, the dt scaling,
// the accumulator update and the stage value for rho (rhoEqn.H:19-33), and
// rho*U, ex::div(rho*U, U), ex::grad(p), ex::laplacian(mu, U), their sum,
// rhoU accumulator and U = rhoUOld/rho + c drhoU/rho for U (UEqn.H:16-33).
// Here each stage is one pass of the plane-streaming pipeline
// (nova_pipe.hpp): a block owns a 64 x 4 (k, j) tile and a chunk of i
// planes, stages the (j, k) planes of its inputs with their halos in LDS,
// keeps the i-neighbours of each thread's column in register queues, and
// has the loads of the next two planes in flight while a plane is computed;
// the products rho*U are formed at each node they are differenced at, as
// the reference's product field holds them.
//
// Stage buffers: rho ping-pongs user -> ra -> rb -> ra -> user with the
// accumulator in rc; U ping-pongs user -> ua -> ub -> ua -> user with the
// rho U accumulator in uc.  The user's U stays the step-start value until
// the last U stage, so rhoUOld = rho*U is formed from it at the node.
#include "nova_internal.hpp"
#include "nova_pipe.hpp"
#include "nova_stencil.hpp"

struct nova_wc {
    nova_ctx *ctx = nullptr;
    double mu = 0.0, pc = 50.0;
    double *ra = nullptr, *rb = nullptr, *rc = nullptr;  // one component each
    double *ua = nullptr, *ub = nullptr, *uc = nullptr;  // three components each
    // frozen magnetic field of app/OT (caller's array) and its Lorentz force
    const double *B = nullptr;
    double dbu = 0.0;
    double *lf = nullptr;  // three components
    // temperature equation of app/shearLayerWC: diffusivity, gravity, stage
    // buffers and the rho T accumulator (one component each)
    double ka = 0.0, grav[3] = {0.0, 0.0, 0.0};
    double *ta = nullptr, *tb = nullptr, *tc = nullptr;
};

namespace nova {
namespace {


struct WcRhoArgs {
    const double *rho_in;  // stage value (stencil)
    const double *U;       // frozen velocity (stencil)
    const double *rho0;    // step-start rho
    const double *rc_in;   // accumulator
    double *rho_out, *rc_out, *p;
    double dt, w, cs, pc;
    int last;
    int p_left;  // p = (pc rho) rho (app/shearLayerWC/pEqn.H:2) instead of pc (rho rho)
};

struct WcUArgs {
    const double *U_in;    // stage value (stencil)
    const double *U0;      // step-start U (null: U_in)
    const double *rho;     // frozen rho (stencil)
    const double *p;       // frozen p (stencil)
    const double *ruc_in;  // rho U accumulator (null: rho U0)
    const double *F;       // Lorentz force -div(2 dbu B, B) + grad(dbu B.B) (may be null)
    const double *T;       // temperature of the buoyancy term -(rho T) g (may be null)
    double g[3];
    double *U_out, *ruc_out;
    double dt, w, cs, mu;
    int last;
};

// rhoEqn.H:19-33: drho = dt * (-div(rho*U)) (src/ex/div/div.hxx:22-81),
// rc += w drho, rho = rho0 + cs drho (the last stage: rho = rc), and on
// the last stage p = pc rho^2 (pEqn.H:2).  Components rho, U_x, U_y, U_z:
// the flux rho U_d is formed at each node it is differenced at along d, so
// U_x is queued, U_y and U_z are staged and rho is both.
template <int H, bool WALLS, int KJ, int PD>
__global__ void __launch_bounds__(pipe::threads(KJ)) k_wc_rho(Grid g, pipe::Launch L, pipe::Src<4> S, WcRhoArgs a)
{
    pipe::run<H, 4, 0b0011u, 0b1101u, WALLS, KJ, PD>(g, L, S, [&](int, int, int, int64_t qi, const auto &nb, const pipe::ZK &z) {
        auto flux = [&](int d, int o) -> double { return nb(0, d, o) * nb(1 + d, d, o); };
        const double zi = z.zi;
        double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
        for (int l = 0; l < H; l++) {
            ax += g.a[0][l] * (flux(0, l + 1) - flux(0, -l - 1));
            ay += g.a[1][l] * (flux(1, l + 1) - flux(1, -l - 1));
            az += g.a[2][l] * (flux(2, l + 1) - flux(2, -l - 1));
        }
        const double dr = a.dt * (-((ax + ay) + az * zi));
        const double rc = a.rc_in[qi] + a.w * dr;
        if (a.last) {
            a.rho_out[qi] = rc;
            if (a.p) a.p[qi] = a.p_left ? (a.pc * rc) * rc : a.pc * (rc * rc);
        } else {
            a.rho_out[qi] = a.rho0[qi] + a.cs * dr;
            a.rc_out[qi] = rc;
        }
    });
}

// UEqn.H:16-33: drhoU = dt * ((-div(rho*U, U) - grad(p)) + laplacian(mu, U))
// (div.hxx:84-172, grad.cpp:25-62, laplacian.hxx:130-191), ruc += w drhoU,
// U = rhoUOld/rho + cs (drhoU/rho) (the last stage: U = ruc/rho), with
// rhoUOld = rho*U0.  Components rho, U_x, U_y, U_z, p, each queued and
// staged (every one is differenced along all three directions).
template <int H, bool WALLS, int KJ, int PD>
__global__ void __launch_bounds__(pipe::threads(KJ)) k_wc_u(Grid g, pipe::Launch L, pipe::Src<5> S, WcUArgs a)
{
    pipe::run<H, 5, 0b11111u, 0b11111u, WALLS, KJ, PD>(g, L, S, [&](int i, int j, int k, int64_t qi, const auto &nb, const pipe::ZK &z) {
        // field f at the node displaced by o along d: 0 = rho*U, 1 = U, 2 = p
        auto A = [&](int f, int c, int d, int o) -> double {
            if (f == 2) return nb(4, d, o);
            const double u = nb(1 + c, d, o);
            return f == 0 ? nb(0, d, o) * u : u;
        };
        const double zi = z.zi, zl = z.zl, zm = z.zm;
        const double r = nb(0, 0, 0);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double gp = 0.0;
#pragma unroll
            for (int l = 0; l < H; l++) gp += g.a[c][l] * (A(2, 0, c, l + 1) - A(2, 0, c, -l - 1));
            if (c == 2) gp *= zi;
            double t = ((-div_xy<H>(g, A, 0, 1, c, zi)) - gp) + lap_c<H>(g, A, 1, c, a.mu, zl, zm);
            if (a.F) t += a.F[c * g.np + qi];
            if (a.T) t = t - (r * a.T[qi]) * a.g[c];
            const double du = a.dt * t;
            const double u0 = a.U0 ? a.U0[c * g.np + qi] : nb(1 + c, 0, 0);
            const double ruold = r * u0;
            const double ruc = (a.ruc_in ? a.ruc_in[c * g.np + qi] : ruold) + a.w * du;
            // the wall-normal component is odd: zero on its walls (the
            // refill after each stage, UEqn.H:33, with a reflecting wall)
            bool on_wall = false;
            if constexpr (WALLS) {
                const int idx = c == 0 ? i : (c == 1 ? j : k);
                on_wall = g.wall[c] && (idx == 0 || idx == g.n[c]);
            }
            if (a.last) {
                a.U_out[c * g.np + qi] = on_wall ? 0.0 : ruc / r;
            } else {
                a.U_out[c * g.np + qi] = on_wall ? 0.0 : ruold / r + a.cs * (du / r);
                a.ruc_out[c * g.np + qi] = ruc;
            }
        }
    });
}

struct WcTArgs {
    const double *T_in;    // stage value (stencil)
    const double *T0;      // step-start T (null: T_in)
    const double *rho;     // frozen rho (stencil)
    const double *U;       // frozen U (stencil)
    const double *rtc_in;  // rho T accumulator (null: rho T0)
    double *T_out, *rtc_out;
    double dt, w, cs, ka;
    int last;
};

// TEqn.H (app/shearLayerWC:16-33): drhoT = dt * (-div(rho*U, T) +
// laplacian(ka, T)), rtc += w drhoT, T = rhoTOld/rho + cs (drhoT/rho) (the
// last stage: T = rtc/rho), with rhoTOld = rho*T0.  Components rho, U_x,
// U_y, U_z, T: rho U_d is differenced along d only (U_x queued, U_y and U_z
// staged), rho and T along all three.
template <int H, bool WALLS, int KJ, int PD>
__global__ void __launch_bounds__(pipe::threads(KJ)) k_wc_t(Grid g, pipe::Launch L, pipe::Src<5> S, WcTArgs a)
{
    pipe::run<H, 5, 0b10011u, 0b11101u, WALLS, KJ, PD>(g, L, S, [&](int, int, int, int64_t qi, const auto &nb, const pipe::ZK &z) {
        // field f at the node displaced by o along d: 0 = rho*U (component
        // c), 1 = T
        auto A = [&](int f, int c, int d, int o) -> double {
            return f == 0 ? nb(0, d, o) * nb(1 + c, d, o) : nb(4, d, o);
        };
        const double zi = z.zi, zl = z.zl, zm = z.zm;
        const double r = nb(0, 0, 0);
        const double dT = a.dt * ((-div_xy<H>(g, A, 0, 1, 0, zi)) + lap_c<H>(g, A, 1, 0, a.ka, zl, zm));
        const double t0 = a.T0 ? a.T0[qi] : nb(4, 0, 0);
        const double rtold = r * t0;
        const double rtc = (a.rtc_in ? a.rtc_in[qi] : rtold) + a.w * dT;
        if (a.last) {
            a.T_out[qi] = rtc / r;
        } else {
            a.T_out[qi] = rtold / r + a.cs * (dT / r);
            a.rtc_out[qi] = rtc;
        }
    });
}

struct WcLArgs {
    const double *B;
    double *F;
    double dbu;
};

// app/OT/UEqn.H:22-23 for a frozen B: F = -div((2 dbu) B, B) + grad(dbu (B & B))
// (div.hxx:84-172, grad.cpp:25-62); B's components queued and staged.
template <int H, bool WALLS, int KJ, int PD>
__global__ void __launch_bounds__(pipe::threads(KJ)) k_wc_lorentz(Grid g, pipe::Launch L, pipe::Src<3> S, WcLArgs a)
{
    pipe::run<H, 3, 0b111u, 0b111u, WALLS, KJ, PD>(g, L, S, [&](int, int, int, int64_t qi, const auto &nb, const pipe::ZK &z) {
        // field 0 = (2 dbu) B, 1 = B
        auto A = [&](int f, int c, int d, int o) -> double {
            const double b = nb(c, d, o);
            return f == 0 ? b * (2 * a.dbu) : b;
        };
        auto pm = [&](int d, int o) -> double {
            const double x = nb(0, d, o), y = nb(1, d, o), z = nb(2, d, o);
            return (x * x + y * y + z * z) * a.dbu;
        };
        const double zi = z.zi;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double gs = 0.0;
#pragma unroll
            for (int l = 0; l < H; l++) gs += g.a[c][l] * (pm(c, l + 1) - pm(c, -l - 1));
            if (c == 2) gs *= zi;
            a.F[c * g.np + qi] = (-div_xy<H>(g, A, 0, 1, c, zi)) + gs;
        }
    });
}

// tile rows and prefetch depth of each stage kernel and stencil half-width,
// the widest that keep it within two waves per SIMD without spills: 64 x 8
// tiles (one halo slot per thread instead of two; the rho stage 269 -> 236 us
// at 256^3 against 64 x 4) up to H = 3 (the momentum stage from H = 2), with
// one plane of loads in flight for the momentum stage and two for the
// others; at H = 4 64 x 4 tiles, the momentum stage with none in flight
enum { WC_RHO, WC_U, WC_T, WC_LOR };
template <int H, int KID>
struct WcCfg {
    static constexpr int KJ = (KID == WC_U && (H == 2 || H == 3)) || (KID != WC_U && H <= 3) ? 8 : 4;
    static constexpr int PD = KID == WC_U ? (H <= 1 ? 2 : (H <= 3 ? 1 : 0)) : (KID != WC_RHO && H >= 4 ? 1 : 2);
};

template <int H, int KID, int NS, class Args>
int launch_wc_kernel(nova_ctx *ctx, void (*kw)(Grid, pipe::Launch, pipe::Src<NS>, Args),
                     void (*kp)(Grid, pipe::Launch, pipe::Src<NS>, Args), const pipe::Src<NS> &S, const Args &a,
                     const char *name)
{
    constexpr int KJ = WcCfg<H, KID>::KJ;
    const Grid &g = ctx->g;
    TimedScope ts(ctx, name);
    const pipe::Launch L = pipe::tiling<H, KJ>(g);
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    hipLaunchKernelGGL(walls ? kw : kp, dim3(pipe::blocks(L)), dim3(pipe::threads(KJ)), 0, ctx->stream, g, L, S, a);
    return check_launch(name);
}

template <int H>
int launch_wc_h(nova_ctx *ctx, const WcRhoArgs *ra, const WcUArgs *ua, const WcLArgs *la, const WcTArgs *tt)
{
    const Grid &g = ctx->g;
    if (tt) {
        using C = WcCfg<H, WC_T>;
        pipe::Src<5> S;
        pipe::set_src(g, S, 0, tt->rho, 1, 0);
        for (int c = 0; c < 3; c++) pipe::set_src(g, S, 1 + c, tt->U, 3, c);
        pipe::set_src(g, S, 4, tt->T_in, 1, 0);
        return launch_wc_kernel<H, WC_T, 5>(ctx, k_wc_t<H, true, C::KJ, C::PD>, k_wc_t<H, false, C::KJ, C::PD>, S, *tt,
                                      "wc_t_stage");
    }
    if (la) {
        using C = WcCfg<H, WC_LOR>;
        pipe::Src<3> S;
        for (int c = 0; c < 3; c++) pipe::set_src(g, S, c, la->B, 3, c);
        return launch_wc_kernel<H, WC_LOR, 3>(ctx, k_wc_lorentz<H, true, C::KJ, C::PD>, k_wc_lorentz<H, false, C::KJ, C::PD>, S,
                                      *la, "wc_lorentz");
    }
    if (ra) {
        using C = WcCfg<H, WC_RHO>;
        pipe::Src<4> S;
        pipe::set_src(g, S, 0, ra->rho_in, 1, 0);
        for (int c = 0; c < 3; c++) pipe::set_src(g, S, 1 + c, ra->U, 3, c);
        return launch_wc_kernel<H, WC_RHO, 4>(ctx, k_wc_rho<H, true, C::KJ, C::PD>, k_wc_rho<H, false, C::KJ, C::PD>, S, *ra,
                                      "wc_rho_stage");
    }
    using C = WcCfg<H, WC_U>;
    pipe::Src<5> S;
    pipe::set_src(g, S, 0, ua->rho, 1, 0);
    for (int c = 0; c < 3; c++) pipe::set_src(g, S, 1 + c, ua->U_in, 3, c);
    pipe::set_src(g, S, 4, ua->p, 1, 0);
    return launch_wc_kernel<H, WC_U, 5>(ctx, k_wc_u<H, true, C::KJ, C::PD>, k_wc_u<H, false, C::KJ, C::PD>, S, *ua,
                                  "wc_u_stage");
}

int launch_wc(nova_ctx *ctx, const WcRhoArgs *ra, const WcUArgs *ua, const WcLArgs *la = nullptr,
              const WcTArgs *tt = nullptr)
{
    switch (ctx->g.h) {
    case 1: return launch_wc_h<1>(ctx, ra, ua, la, tt);
    case 2: return launch_wc_h<2>(ctx, ra, ua, la, tt);
    case 3: return launch_wc_h<3>(ctx, ra, ua, la, tt);
    case 4: return launch_wc_h<4>(ctx, ra, ua, la, tt);
    default: return fail(NOVA_ERR_ARG, "stencil order");
    }
}

// Classic RK4 of the reference apps (app/TG/TG.cpp:57): weights and the
// stage offsets (int(s/2)+1)/2.
constexpr double kRk4W[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
constexpr double kRk4C[3] = {0.5, 0.5, 1.0};

}  // namespace

int wc_create(nova_ctx *ctx, double mu, double pc, nova_wc **out)
{
    if (!ctx || !out) return fail(NOVA_ERR_ARG, "null argument");
    if (!(mu >= 0.0)) return fail(NOVA_ERR_ARG, "mu must be >= 0");
    if (ctx->wall_fill) return fail(NOVA_ERR_UNSUPPORTED, "the weakly-compressible steps take the reflection rule at walls");
    nova_wc *w = new nova_wc;
    w->ctx = ctx;
    w->mu = mu;
    w->pc = pc;
    const size_t b1 = sizeof(double) * ctx->g.np, b3 = 3 * b1;
    hipError_t e = hipSuccess;
    double **one[3] = {&w->ra, &w->rb, &w->rc};
    double **three[3] = {&w->ua, &w->ub, &w->uc};
    for (int q = 0; q < 3 && e == hipSuccess; q++) e = hipMalloc(one[q], b1);
    for (int q = 0; q < 3 && e == hipSuccess; q++) e = hipMalloc(three[q], b3);
    if (e != hipSuccess) {
        for (int q = 0; q < 3; q++) {
            if (*one[q]) (void)hipFree(*one[q]);
            if (*three[q]) (void)hipFree(*three[q]);
        }
        delete w;
        return hip_fail(e, "hipMalloc (weakly-compressible stage buffers)");
    }
    *out = w;
    return NOVA_OK;
}

int wc_destroy(nova_wc *w)
{
    if (!w) return NOVA_OK;
    for (double *b : {w->ra, w->rb, w->rc, w->ua, w->ub, w->uc, w->lf, w->ta, w->tb, w->tc})
        if (b) (void)hipFree(b);
    delete w;
    return NOVA_OK;
}

// The frozen magnetic field of app/OT (B is never reassigned there:
// BEqn.H:26-31 only forms Bstar): its Lorentz force enters every U stage.
// B stays the caller's array; NULL removes it.
int wc_set_magnetic(nova_wc *w, const double *B, double dbu)
{
    if (B && !w->lf) NOVA_HIP(hipMalloc(&w->lf, sizeof(double) * 3 * w->ctx->g.np));
    w->B = B;
    w->dbu = dbu;
    return NOVA_OK;
}

namespace {

// rhoEqn.H: four rho stages with U frozen; the last writes p (pEqn.H).
int wc_rho_eqn(nova_wc *w, double dt, const double *U, double *rho, double *p, int p_left)
{
    const double *rin[4] = {rho, w->ra, w->rb, w->ra};
    double *rout[4] = {w->ra, w->rb, w->ra, rho};
    for (int s = 0; s < 4; s++) {
        WcRhoArgs a{};
        a.rho_in = rin[s];
        a.U = U;
        a.rho0 = rho;
        a.rc_in = s == 0 ? rho : w->rc;
        a.rho_out = rout[s];
        a.rc_out = w->rc;
        a.p = p;
        a.dt = dt;
        a.w = kRk4W[s];
        a.cs = s < 3 ? kRk4C[s] : 0.0;
        a.pc = w->pc;
        a.last = s == 3;
        a.p_left = p_left;
        if (int rc = launch_wc(w->ctx, &a, nullptr)) return rc;
    }
    return NOVA_OK;
}

// UEqn.H: the Lorentz force of a frozen B (when set), then four U stages
// with rho, p (and T) frozen.
int wc_u_eqn(nova_wc *w, double dt, double *U, const double *rho, const double *p, const double *T)
{
    if (w->B) {
        WcLArgs l{};
        l.B = w->B;
        l.F = w->lf;
        l.dbu = w->dbu;
        if (int rc = launch_wc(w->ctx, nullptr, nullptr, &l)) return rc;
    }
    const double *uin[4] = {U, w->ua, w->ub, w->ua};
    double *uout[4] = {w->ua, w->ub, w->ua, U};
    for (int s = 0; s < 4; s++) {
        WcUArgs a{};
        a.U_in = uin[s];
        a.U0 = (s == 0 || s == 3) ? nullptr : U;
        a.rho = rho;
        a.p = p;
        a.ruc_in = s == 0 ? nullptr : w->uc;
        a.F = w->B ? w->lf : nullptr;
        a.T = T;
        for (int c = 0; c < 3; c++) a.g[c] = w->grav[c];
        a.U_out = uout[s];
        a.ruc_out = w->uc;
        a.dt = dt;
        a.w = kRk4W[s];
        a.cs = s < 3 ? kRk4C[s] : 0.0;
        a.mu = w->mu;
        a.last = s == 3;
        if (int rc = launch_wc(w->ctx, nullptr, &a)) return rc;
    }
    return NOVA_OK;
}

// TEqn.H: four T stages with rho and U frozen.
int wc_t_eqn(nova_wc *w, double dt, const double *U, const double *rho, double *T)
{
    const double *tin[4] = {T, w->ta, w->tb, w->ta};
    double *tout[4] = {w->ta, w->tb, w->ta, T};
    for (int s = 0; s < 4; s++) {
        WcTArgs a{};
        a.T_in = tin[s];
        a.T0 = (s == 0 || s == 3) ? nullptr : T;
        a.rho = rho;
        a.U = U;
        a.rtc_in = s == 0 ? nullptr : w->tc;
        a.T_out = tout[s];
        a.rtc_out = w->tc;
        a.dt = dt;
        a.w = kRk4W[s];
        a.cs = s < 3 ? kRk4C[s] : 0.0;
        a.ka = w->ka;
        a.last = s == 3;
        if (int rc = launch_wc(w->ctx, nullptr, nullptr, nullptr, &a)) return rc;
    }
    return NOVA_OK;
}

}  // namespace

// app/TG, app/OT order (TG.cpp:71-76): rhoEqn, pEqn, UEqn.
int wc_step(nova_wc *w, double dt, int nsteps, double *U, double *rho, double *p)
{
    for (int n = 0; n < nsteps; n++) {
        if (int rc = wc_rho_eqn(w, dt, U, rho, p, 0)) return rc;
        if (int rc = wc_u_eqn(w, dt, U, rho, p, nullptr)) return rc;
    }
    return NOVA_OK;
}

// The temperature equation and buoyancy of app/shearLayerWC.
int wc_set_temperature(nova_wc *w, double ka, const double g[3])
{
    if (!(ka >= 0.0)) return fail(NOVA_ERR_ARG, "ka must be >= 0");
    const size_t b1 = sizeof(double) * w->ctx->g.np;
    for (double **b : {&w->ta, &w->tb, &w->tc})
        if (!*b) NOVA_HIP(hipMalloc(b, b1));
    w->ka = ka;
    for (int c = 0; c < 3; c++) w->grav[c] = g[c];
    return NOVA_OK;
}

// app/shearLayerWC order (shearLayer.cpp:197-200): UEqn with buoyancy (p as
// given: the previous step's), rhoEqn, pEqn, TEqn.
int wc_step_t(nova_wc *w, double dt, int nsteps, double *U, double *rho, double *p, double *T)
{
    if (!w->ta) return fail(NOVA_ERR_ARG, "nova_wc_set_temperature first");
    for (int n = 0; n < nsteps; n++) {
        if (int rc = wc_u_eqn(w, dt, U, rho, p, T)) return rc;
        if (int rc = wc_rho_eqn(w, dt, U, rho, p, 1)) return rc;
        if (int rc = wc_t_eqn(w, dt, U, rho, T)) return rc;
    }
    return NOVA_OK;
}

// The initial hydrostatic pressure and density of app/shearLayerWC
// (hydrostaticP.H:1-37, shearLayer.cpp:139-151), run once before the time
// loop.  One thread per (i, j) column walks its contiguous k run (each
// cache line serves 16 consecutive k).  The partial sum is formed from the
// top in the reference's order, ((T dz) g_z) rho0 added to it, unfused.
__global__ void __launch_bounds__(256) k_wc_hydrostatic(Grid g, const double *__restrict__ T, double *__restrict__ p,
                                                        double *__restrict__ rho, double dz, double gz, double rho0,
                                                        double pc)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (int64_t)g.s[0] * g.s[1]) return;
    const int i = (int)(col / g.s[1]);
    const int64_t base = col * g.s[2];
    // the reference's i loop stops one node short of its last node (the
    // periodic image, absent here, or the wall node); its k sum starts one
    // node below the last (the image, or the top wall node)
    if (i < g.n[0]) {
        double acc = 0.0;
        for (int k = g.n[2] - 1; k >= 0; k--) {
#pragma clang fp contract(off)
            acc = acc + ((T[base + k] * dz) * gz) * rho0;
            p[base + k] = p[base + k] + acc;
        }
    }
    for (int k = 0; k < g.s[2]; k++) rho[base + k] = sqrt(p[base + k] / pc);
}

int wc_hydrostatic(nova_wc *w, const double *T, double rho0, double *p, double *rho)
{
    if (!w->ta) return fail(NOVA_ERR_ARG, "nova_wc_set_temperature first");
    nova_ctx *ctx = w->ctx;
    const Grid &g = ctx->g;
    const int64_t cols = (int64_t)g.s[0] * g.s[1];
    TimedScope ts(ctx, "wc_hydrostatic");
    hipLaunchKernelGGL(k_wc_hydrostatic, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, ctx->stream, g, T, p, rho,
                       ctx->d[2], w->grav[2], rho0, w->pc);
    return check_launch("wc_hydrostatic");
}

}  // namespace nova

