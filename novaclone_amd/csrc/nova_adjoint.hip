// nova_adjoint.hip -- tangent-linear model and discrete adjoint of the
// SSP-RK3 + projection step (no reference counterpart: NOVA has no adjoint).
//
// Forward step (nova_step.hip):  U1 = U + dt R(U),
//   U2 = 3/4 U + 1/4 (U1 + dt R(U1)),  U3 = 1/3 U + 2/3 (U2 + dt R(U2)),
//   U' = P U3,  with R(U) = -B(U, U) + nu L U, B(X, Y) = ex::div(X, Y)
//   (src/ex/div/div.hxx:84-172) and P the projection.
// Tangent-linear:  J_R(X) d = -B(d, X) - B(X, d) + nu L d.
// Adjoint (inner product sum over nodes, with weight 1/2 on wall nodes):
//   the central differences satisfy D^T = -D, so
//   J_R(X)^T m = B(X, m) - m div(X) - B(., X)^T m + nu L m,
//   (B(., X)^T m)_d = 1/2 sum_c [ -X_c D_d m_c + m_c D_d X_c - D_d(m_c X_c) ],
//   and P^T = P (the solve inverts D.G, which is symmetric).
// Reverse step, given m = dJ/dU':  m3 = P m;  m2 = 2/3 T(U2, m3);
//   m1 = 1/4 T(U1, m2);  dJ/dU = T(U, m1) + 3/4 m2 + 1/3 m3,
//   with T(X, m) = m + dt J_R(X)^T m.
// Checkpoints of U every C steps are kept; each segment is recomputed once
// in the reverse sweep, storing the stage inputs (U, U1, U2) of its steps.
#include "nova_internal.hpp"
#include "nova_stencil.hpp"

#include <vector>

struct nova_poisson;
struct nova_solver;

namespace nova {
int project(nova_poisson *ps, double *F, double s, double *p);
nova_ctx *poisson_ctx(nova_poisson *ps);
int stage_ns(nova_ctx *ctx, const double *U, const double *U0, double *out, const StageArgs &sa, double nu);
nova_poisson *solver_poisson(nova_solver *sv);
int solver_model(nova_solver *sv, int *mode, double *nu);
}  // namespace nova

struct nova_adjoint {
    nova_solver *sv = nullptr;
    nova_poisson *ps = nullptr;
    nova_ctx *ctx = nullptr;
    double nu = 0.0;
    int every = 1, max_steps = 0;
    int nsteps = 0;
    double dt = 0.0;
    std::vector<double *> ckpt;   // U at steps 0, C, 2C, ...
    std::vector<double *> seg;    // per step of a segment: U, U1, U2
    double *w[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // work vector fields
    double *ptmp = nullptr;
};

namespace nova {

namespace {

enum { LIN_TLM = 0, LIN_ADJ = 1 };

struct LinArgs {
    const double *X;    // forward stage input (stencil)
    const double *V;    // tangent or adjoint field (stencil)
    const double *V0;   // TLM: step-start tangent for the stage combination
    double *out;
    const double *Y;    // ADJ: pointwise term added to the output
    double *acc;        // ADJ: accumulator, acc = acc_a*acc + acc_b*V
    double beta, gamma, acc_a, acc_b;
    StageArgs sa;
    double nu;
    int ichunk, nkt, njt, nic;
};

constexpr int kLK = 64, kLJ = 4;

// (B(., X)^T m)_d for output component d: 1/2 sum_c sum_l a_l
// [ -X_c (m_c^+ - m_c^-) + m_c (X_c^+ - X_c^-) - ((m_c X_c)^+ - (m_c X_c)^-) ]
template <int H, class Acc>
__device__ inline double b_second_adj(const Grid &g, const Acc &A, int d, double zi)
{
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double x0 = A(0, c, 0, 0), m0 = A(1, c, 0, 0);
#pragma unroll
        for (int l = 0; l < H; l++) {
            const double xp = A(0, c, d, l + 1), xm = A(0, c, d, -l - 1);
            const double mp = A(1, c, d, l + 1), mm = A(1, c, d, -l - 1);
            acc += g.a[d][l] * (-x0 * (mp - mm) + m0 * (xp - xm) - (mp * xp - mm * xm));
        }
    }
    return 0.5 * (d == 2 ? acc * zi : acc);
}

template <int H, class Acc>
__device__ inline double div_of(const Grid &g, const Acc &A, int f, double zi)
{
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double ad = 0.0;
#pragma unroll
        for (int l = 0; l < H; l++) ad += g.a[d][l] * (A(f, d, d, l + 1) - A(f, d, d, -l - 1));
        acc += d == 2 ? ad * zi : ad;
    }
    return acc;
}

// One thread per (j,k) column streaming i, like the row variant of the
// forward stage kernel, with the forward input X (field 0) and the
// tangent/adjoint V (field 1) both queued.
template <int H, bool WALLS, int MODE>
__global__ void __launch_bounds__(kLK *kLJ) k_lin_rows(Grid g, LinArgs a)
{
    const int T = a.nkt * a.njt * a.nic;
    int t = blockIdx.x;
    if (T % 8 == 0) t = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
    const int kt = t % a.nkt;
    const int jt = (t / a.nkt) % a.njt;
    const int ic = t / (a.nkt * a.njt);
    const int k = kt * kLK + threadIdx.x, j = jt * kLJ + threadIdx.y;
    if (k >= g.s[2] || j >= g.s[1]) return;
    const int i0 = ic * a.ichunk;
    int i1 = i0 + a.ichunk;
    if (i1 > g.s[0]) i1 = g.s[0];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    const int s2 = g.s[2];
    const int col = j * s2 + k;
    const double *F[2] = {a.X, a.V};

    double q[2][3][2 * H + 1];
#pragma unroll
    for (int o = 0; o < 2 * H + 1; o++) {
        int fi;
        const int64_t po = (int64_t)wrap<WALLS>(i0 - H + o, g.n[0], g.wall[0], fi) * plane;
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int c = 0; c < 3; c++) q[f][c][o] = rsign<WALLS>(c, 0, fi) * F[f][c * g.np + po + col];
    }
    const double zi = g.zinv[k], zl = g.zlap[k], zm = g.zmet[k];
    for (int i = i0; i < i1; i++) {
        double nq[2][3];
        if (i + 1 < i1) {
            int fi;
            const int64_t po = (int64_t)wrap<WALLS>(i + H + 1, g.n[0], g.wall[0], fi) * plane;
#pragma unroll
            for (int f = 0; f < 2; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) nq[f][c] = rsign<WALLS>(c, 0, fi) * F[f][c * g.np + po + col];
        }
        const int64_t pi = (int64_t)i * plane;
        auto A = [&](int f, int c, int d, int o) -> double {
            if (d == 0 || o == 0) return q[f][c][H + (d == 0 ? o : 0)];
            int fl;
            int64_t off;
            if (d == 1) off = pi + (int64_t)wrap<WALLS>(j + o, g.n[1], g.wall[1], fl) * s2 + k;
            else off = pi + (int64_t)j * s2 + wrap<WALLS>(k + o, g.n[2], g.wall[2], fl);
            return rsign<WALLS>(c, d, fl) * F[f][c * g.np + off];
        };
        const int64_t qi = pi + col;
        if (MODE == LIN_TLM) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double r = (-div_xy<H>(g, A, 1, 0, c, zi)) - div_xy<H>(g, A, 0, 1, c, zi) +
                                 lap_c<H>(g, A, 1, c, a.nu, zl, zm);
                const double v = q[1][c][H];
                const double v0 = a.V0 ? a.V0[c * g.np + qi] : v;
                a.out[c * g.np + qi] = combine(a.sa, v, v0, a.sa.dt * r);
            }
        } else {
            const double divx = div_of<H>(g, A, 0, zi);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double m = q[1][c][H];
                const double jr = div_xy<H>(g, A, 0, 1, c, zi) - m * divx - b_second_adj<H>(g, A, c, zi) +
                                  lap_c<H>(g, A, 1, c, a.nu, zl, 0.0);
                const double tv = m + a.sa.dt * jr;
                double o = a.beta * tv;
                if (a.Y) o += a.gamma * a.Y[c * g.np + qi];
                if (a.acc) a.acc[c * g.np + qi] = (a.acc_a != 0.0 ? a.acc_a * a.acc[c * g.np + qi] : 0.0) + a.acc_b * m;
                a.out[c * g.np + qi] = o;
            }
        }
        if (i + 1 < i1) {
#pragma unroll
            for (int f = 0; f < 2; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
#pragma unroll
                    for (int o = 0; o < 2 * H; o++) q[f][c][o] = q[f][c][o + 1];
                    q[f][c][2 * H] = nq[f][c];
                }
        }
    }
}

template <int H, bool WALLS, int MODE>
void launch_lin_hw(nova_ctx *ctx, LinArgs a)
{
    const Grid &g = ctx->g;
    a.nkt = (g.s[2] + kLK - 1) / kLK;
    a.njt = (g.s[1] + kLJ - 1) / kLJ;
    int64_t per = (int64_t)g.s[0] * a.nkt * a.njt / 4096;
    a.ichunk = (int)(per < 8 ? 8 : (per > 32 ? 32 : per));
    if (a.ichunk > g.s[0]) a.ichunk = g.s[0];
    a.nic = (g.s[0] + a.ichunk - 1) / a.ichunk;
    const int T = a.nkt * a.njt * a.nic;
    hipLaunchKernelGGL((k_lin_rows<H, WALLS, MODE>), dim3(T), dim3(kLK, kLJ, 1), 0, ctx->stream, g, a);
}

template <int MODE>
int launch_lin(nova_ctx *ctx, const LinArgs &a, const char *name)
{
    const Grid &g = ctx->g;
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    TimedScope ts(ctx, name);
    switch (g.h * 2 + (walls ? 1 : 0)) {
    case 2: launch_lin_hw<1, false, MODE>(ctx, a); break;
    case 3: launch_lin_hw<1, true, MODE>(ctx, a); break;
    case 4: launch_lin_hw<2, false, MODE>(ctx, a); break;
    case 5: launch_lin_hw<2, true, MODE>(ctx, a); break;
    case 6: launch_lin_hw<3, false, MODE>(ctx, a); break;
    case 7: launch_lin_hw<3, true, MODE>(ctx, a); break;
    case 8: launch_lin_hw<4, false, MODE>(ctx, a); break;
    default: launch_lin_hw<4, true, MODE>(ctx, a); break;
    }
    return check_launch(name);
}

StageArgs ssp_args(int s, double dt)
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
    NOVA_HIP(hipMemcpyAsync(dst, src, sizeof(double) * 3 * ctx->g.np, hipMemcpyDeviceToDevice, ctx->stream));
    return NOVA_OK;
}

// Forward step from U storing the stage inputs U1, U2; the next state is
// written to Unext (which may be U itself).
int fwd_step_store(nova_adjoint *ad, double *U, double *U1, double *U2, double *Unext)
{
    nova_ctx *ctx = ad->ctx;
    int rc;
    if ((rc = stage_ns(ctx, U, nullptr, U1, ssp_args(0, ad->dt), ad->nu))) return rc;
    if ((rc = stage_ns(ctx, U1, U, U2, ssp_args(1, ad->dt), ad->nu))) return rc;
    if ((rc = stage_ns(ctx, U2, U, Unext, ssp_args(2, ad->dt), ad->nu))) return rc;
    return project(ad->ps, Unext, ad->dt, ad->ptmp);
}

// Recompute segment s (steps s*C .. min((s+1)*C, n) - 1), storing
// (U, U1, U2) per step.  Returns the number of steps in the segment.
int recompute_segment(nova_adjoint *ad, int s, int *count)
{
    const int C = ad->every;
    const int n0 = s * C;
    int n1 = n0 + C;
    if (n1 > ad->nsteps) n1 = ad->nsteps;
    int rc;
    if ((rc = copy3(ad->ctx, ad->seg[0], ad->ckpt[s]))) return rc;
    for (int n = n0; n < n1 - 1; n++) {  // the segment's last step is not needed forward
        const int q = n - n0;
        if ((rc = fwd_step_store(ad, ad->seg[3 * q], ad->seg[3 * q + 1], ad->seg[3 * q + 2], ad->seg[3 * q + 3])))
            return rc;
    }
    {
        // stage inputs of the last step only
        const int q = n1 - 1 - n0;
        if ((rc = stage_ns(ad->ctx, ad->seg[3 * q], nullptr, ad->seg[3 * q + 1], ssp_args(0, ad->dt), ad->nu)))
            return rc;
        if ((rc = stage_ns(ad->ctx, ad->seg[3 * q + 1], ad->seg[3 * q], ad->seg[3 * q + 2], ssp_args(1, ad->dt),
                           ad->nu)))
            return rc;
    }
    *count = n1 - n0;
    return NOVA_OK;
}

LinArgs lin_args(nova_adjoint *ad)
{
    LinArgs a{};
    a.nu = ad->nu;
    a.sa = ssp_args(0, ad->dt);
    return a;
}

}  // namespace

nova_poisson *solver_poisson(nova_solver *sv);

int adjoint_create(nova_solver *sv, int every, int max_steps, nova_adjoint **out)
{
    int mode = 0;
    double nu = 0.0;
    if (int rc = solver_model(sv, &mode, &nu)) return rc;
    if (mode != NOVA_RK_SSP3) return fail(NOVA_ERR_UNSUPPORTED, "the adjoint covers the NS SSP-RK3 composition");
    nova_poisson *ps = solver_poisson(sv);
    nova_ctx *ctx = poisson_ctx(ps);
    if (ctx->desc.hyp) return fail(NOVA_ERR_UNSUPPORTED, "the adjoint needs a uniform z grid");
    if (every < 1 || max_steps < 1) return fail(NOVA_ERR_ARG, "checkpoint interval and step budget must be >= 1");
    nova_adjoint *ad = new nova_adjoint();
    ad->sv = sv;
    ad->ps = ps;
    ad->ctx = ctx;
    ad->nu = nu;
    ad->every = every;
    ad->max_steps = max_steps;
    const size_t bytes = sizeof(double) * 3 * (size_t)ctx->g.np;
    const int nck = (max_steps + every - 1) / every;
    hipError_t e = hipSuccess;
    for (int q = 0; q < nck && e == hipSuccess; q++) {
        double *p = nullptr;
        e = hipMalloc((void **)&p, bytes);
        ad->ckpt.push_back(p);
    }
    for (int q = 0; q < 3 * every && e == hipSuccess; q++) {
        double *p = nullptr;
        e = hipMalloc((void **)&p, bytes);
        ad->seg.push_back(p);
    }
    for (int q = 0; q < 5 && e == hipSuccess; q++) e = hipMalloc((void **)&ad->w[q], bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&ad->ptmp, sizeof(double) * ctx->g.np);
    if (e != hipSuccess) {
        for (double *p : ad->ckpt) if (p) (void)hipFree(p);
        for (double *p : ad->seg) if (p) (void)hipFree(p);
        for (double *p : ad->w) if (p) (void)hipFree(p);
        if (ad->ptmp) (void)hipFree(ad->ptmp);
        delete ad;
        return hip_fail(e, "adjoint checkpoint memory");
    }
    *out = ad;
    return NOVA_OK;
}

int adjoint_destroy(nova_adjoint *ad)
{
    if (!ad) return NOVA_OK;
    (void)hipStreamSynchronize(ad->ctx->stream);
    for (double *p : ad->ckpt) if (p) (void)hipFree(p);
    for (double *p : ad->seg) if (p) (void)hipFree(p);
    for (double *p : ad->w) if (p) (void)hipFree(p);
    if (ad->ptmp) (void)hipFree(ad->ptmp);
    delete ad;
    return NOVA_OK;
}

int adjoint_forward(nova_adjoint *ad, double dt, int nsteps, const double *U0, double *U)
{
    if (nsteps < 1 || nsteps > ad->max_steps) return fail(NOVA_ERR_ARG, "nsteps outside the checkpoint budget");
    ad->dt = dt;
    ad->nsteps = nsteps;
    int rc;
    if (U != U0 && (rc = copy3(ad->ctx, U, U0))) return rc;
    for (int n = 0; n < nsteps; n++) {
        if (n % ad->every == 0 && (rc = copy3(ad->ctx, ad->ckpt[n / ad->every], U))) return rc;
        if ((rc = fwd_step_store(ad, U, ad->w[0], ad->w[1], U))) return rc;
    }
    return NOVA_OK;
}

int adjoint_backward(nova_adjoint *ad, const double *lamT, double *lam0)
{
    if (ad->nsteps < 1) return fail(NOVA_ERR_STATE, "run nova_adjoint_forward first");
    nova_ctx *ctx = ad->ctx;
    double *L = lam0, *A = ad->w[0], *B = ad->w[1], *ACC = ad->w[2];
    int rc;
    if (L != lamT && (rc = copy3(ctx, L, lamT))) return rc;
    const int nseg = (ad->nsteps + ad->every - 1) / ad->every;
    for (int s = nseg - 1; s >= 0; s--) {
        int cnt = 0;
        if ((rc = recompute_segment(ad, s, &cnt))) return rc;
        for (int q = cnt - 1; q >= 0; q--) {
            double *U = ad->seg[3 * q], *U1 = ad->seg[3 * q + 1], *U2 = ad->seg[3 * q + 2];
            if ((rc = project(ad->ps, L, 1.0, ad->ptmp))) return rc;  // m3 = P m
            LinArgs a = lin_args(ad);
            a.X = U2; a.V = L; a.out = A; a.beta = 2.0 / 3.0; a.acc = ACC; a.acc_a = 0.0; a.acc_b = 1.0 / 3.0;
            if ((rc = launch_lin<LIN_ADJ>(ctx, a, "adj_stage"))) return rc;
            a = lin_args(ad);
            a.X = U1; a.V = A; a.out = B; a.beta = 0.25; a.acc = ACC; a.acc_a = 1.0; a.acc_b = 0.75;
            if ((rc = launch_lin<LIN_ADJ>(ctx, a, "adj_stage"))) return rc;
            a = lin_args(ad);
            a.X = U; a.V = B; a.out = L; a.beta = 1.0; a.Y = ACC; a.gamma = 1.0;
            if ((rc = launch_lin<LIN_ADJ>(ctx, a, "adj_stage"))) return rc;
        }
    }
    return NOVA_OK;
}

int adjoint_tlm(nova_adjoint *ad, const double *d0, double *dT)
{
    if (ad->nsteps < 1) return fail(NOVA_ERR_STATE, "run nova_adjoint_forward first");
    nova_ctx *ctx = ad->ctx;
    double *D = dT, *D1 = ad->w[0], *D2 = ad->w[1];
    int rc;
    if (D != d0 && (rc = copy3(ctx, D, d0))) return rc;
    const int nseg = (ad->nsteps + ad->every - 1) / ad->every;
    for (int s = 0; s < nseg; s++) {
        int cnt = 0;
        if ((rc = recompute_segment(ad, s, &cnt))) return rc;
        for (int q = 0; q < cnt; q++) {
            double *U = ad->seg[3 * q], *U1 = ad->seg[3 * q + 1], *U2 = ad->seg[3 * q + 2];
            LinArgs a = lin_args(ad);
            a.X = U; a.V = D; a.V0 = nullptr; a.out = D1; a.sa = ssp_args(0, ad->dt);
            if ((rc = launch_lin<LIN_TLM>(ctx, a, "tlm_stage"))) return rc;
            a.X = U1; a.V = D1; a.V0 = D; a.out = D2; a.sa = ssp_args(1, ad->dt);
            if ((rc = launch_lin<LIN_TLM>(ctx, a, "tlm_stage"))) return rc;
            a.X = U2; a.V = D2; a.V0 = D; a.out = D; a.sa = ssp_args(2, ad->dt);
            if ((rc = launch_lin<LIN_TLM>(ctx, a, "tlm_stage"))) return rc;
            if ((rc = project(ad->ps, D, ad->dt, ad->ptmp))) return rc;
        }
    }
    return NOVA_OK;
}

}  // namespace nova
