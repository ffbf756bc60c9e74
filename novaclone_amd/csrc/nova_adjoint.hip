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
//   with T(X, m) = m + dt J_R(X)^T m: three adjoint stage launches, the last
//   reading m2 and m3 at its own nodes and writing over m3.
// Checkpoints of U every C steps are kept; each segment is recomputed once
// in the reverse sweep, storing the stage inputs (U, U1, U2) of its steps in
// a segment store.  The forward sweep writes the stage inputs of its last
// segments into the stores it has (one, plus those nova_adjoint_reserve
// added) as it goes, and the reverse sweep recomputes only the others.
#include "nova_internal.hpp"
#include "nova_stencil.hpp"
#include "nova_stream.hpp"

#include <vector>

struct nova_poisson;
struct nova_solver;

namespace nova {
int project(nova_poisson *ps, double *F, double s, double *p);
nova_ctx *poisson_ctx(nova_poisson *ps);
int stage_ns(nova_ctx *ctx, const double *U, const double *U0, double *out, const StageArgs &sa, double nu);
nova_poisson *solver_poisson(nova_solver *sv);
int solver_model(nova_solver *sv, int *mode, double *nu);
int solver_scheme(nova_solver *sv);
int compact_ns_stage(nova_ctx *ctx, const double *U, const double *U0, double *out, const StageArgs &sa, double nu,
                     double *R);
int compact_tlm_stage(nova_ctx *ctx, const double *X, const double *V, const double *V0, double *out,
                      const StageArgs &sa, double nu, double *R);
int compact_adj_stage(nova_ctx *ctx, const double *X, const double *M, double *out, double ob, const double *Y,
                      double yg, double *acc, double acc_a, double acc_b, double dt, double nu, double *R);
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
    // segment stores: per step of a segment U, U1, U2; store 0 receives the
    // recomputed segments, the others only segments of the forward sweep
    std::vector<std::vector<double *>> store;
    std::vector<int> held;        // the segment whose stage inputs each store holds (-1: none)
    double *w[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // work vector fields
    double *ptmp = nullptr;
    int scheme = NOVA_SCHEME_EXPLICIT;
    double *rwork = nullptr;  // compact right-hand-side accumulator (4 comps)
};

namespace nova {

namespace {

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

using LinArgs = StreamArgs<2>;

// forward stage out = combine(X, X0, dt R(X)) in the solver's scheme
int fwd_stage(nova_adjoint *ad, const double *X, const double *X0, double *out, const StageArgs &sa)
{
    if (ad->scheme == NOVA_SCHEME_COMPACT6) return compact_ns_stage(ad->ctx, X, X0, out, sa, ad->nu, ad->rwork);
    return stage_ns(ad->ctx, X, X0, out, sa, ad->nu);
}

// tangent-linear stage out = combine(V, V0, dt J_R(X) V)
int tlm_stage(nova_adjoint *ad, const double *X, const double *V, const double *V0, double *out, const StageArgs &sa)
{
    if (ad->scheme == NOVA_SCHEME_COMPACT6) return compact_tlm_stage(ad->ctx, X, V, V0, out, sa, ad->nu, ad->rwork);
    LinArgs a{};
    a.X[0] = X;
    a.X[1] = V;
    a.X0[1] = V0;
    a.out[0] = out;
    a.sa = sa;
    a.nu = ad->nu;
    return launch_stream_model<MODEL_TLM>(ad->ctx, a, "tlm_stage");
}

LinArgs adj_args(nova_adjoint *ad, const double *X, const double *M, double *out)
{
    LinArgs a{};
    a.X[0] = X;
    a.X[1] = M;
    a.out[0] = out;
    a.nu = ad->nu;
    a.sa.form = FORM_SSP;
    a.sa.dt = ad->dt;
    a.ak = adj_k(ad->ctx->g, ad->nu, ad->dt);
    return a;
}

// explicit adjoint stages: out = ob T(X, M), and out = T(X, M) + yg Y + zg Z
// (Z may be out)
int adj_stage_scaled(nova_adjoint *ad, const double *X, const double *M, double *out, double ob)
{
    LinArgs a = adj_args(ad, X, M, out);
    a.ob = ob;
    return launch_stream_model<MODEL_ADJ, 1>(ad->ctx, a, "adj_stage");
}

int adj_stage_sum(nova_adjoint *ad, const double *X, const double *M, double *out, const double *Y, double yg,
                  const double *Z, double zg)
{
    LinArgs a = adj_args(ad, X, M, out);
    a.Y = Y;
    a.yg = yg;
    a.Z = Z;
    a.zg = zg;
    return launch_stream_model<MODEL_ADJ, 2>(ad->ctx, a, "adj_stage");
}

// Forward step from U storing the stage inputs U1, U2; the next state is
// written to Unext (which may be U itself).
int fwd_step_store(nova_adjoint *ad, double *U, double *U1, double *U2, double *Unext)
{
    nova_ctx *ctx = ad->ctx;
    int rc;
    if ((rc = fwd_stage(ad, U, nullptr, U1, ssp_args(0, ad->dt)))) return rc;
    if ((rc = fwd_stage(ad, U1, U, U2, ssp_args(1, ad->dt)))) return rc;
    if ((rc = fwd_stage(ad, U2, U, Unext, ssp_args(2, ad->dt)))) return rc;
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
    std::vector<double *> &seg = ad->store[0];
    ad->held[0] = -1;
    if ((rc = copy3(ad->ctx, seg[0], ad->ckpt[s]))) return rc;
    for (int n = n0; n < n1 - 1; n++) {  // the segment's last step is not needed forward
        const int q = n - n0;
        if ((rc = fwd_step_store(ad, seg[3 * q], seg[3 * q + 1], seg[3 * q + 2], seg[3 * q + 3]))) return rc;
    }
    {
        // stage inputs of the last step only
        const int q = n1 - 1 - n0;
        if ((rc = fwd_stage(ad, seg[3 * q], nullptr, seg[3 * q + 1], ssp_args(0, ad->dt)))) return rc;
        if ((rc = fwd_stage(ad, seg[3 * q + 1], seg[3 * q], seg[3 * q + 2], ssp_args(1, ad->dt)))) return rc;
    }
    *count = n1 - n0;
    ad->held[0] = s;
    return NOVA_OK;
}

// The stage inputs of segment s: a store that holds them, else store 0
// after a recompute.
int segment_inputs(nova_adjoint *ad, int s, double *const **seg, int *count)
{
    for (size_t k = 0; k < ad->store.size(); k++)
        if (ad->held[k] == s) {
            *seg = ad->store[k].data();
            *count = (s + 1) * ad->every < ad->nsteps ? ad->every : ad->nsteps - s * ad->every;
            return NOVA_OK;
        }
    if (int rc = recompute_segment(ad, s, count)) return rc;
    *seg = ad->store[0].data();
    return NOVA_OK;
}

int alloc_store(size_t bytes, int nfields, std::vector<double *> &st)
{
    for (int q = 0; q < nfields; q++) {
        double *p = nullptr;
        const hipError_t e = hipMalloc((void **)&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();  // an allocation failure is not sticky
            for (double *x : st) (void)hipFree(x);
            st.clear();
            return NOVA_ERR_HIP;
        }
        st.push_back(p);
    }
    return NOVA_OK;
}

LinArgs lin_args(nova_adjoint *ad, const double *X, const double *V)
{
    LinArgs a{};
    a.X[0] = X;
    a.X[1] = V;
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
    if (ctx->wall_fill) return fail(NOVA_ERR_UNSUPPORTED, "the adjoint takes the reflection rule at walls");
    if (every < 1 || max_steps < 1) return fail(NOVA_ERR_ARG, "checkpoint interval and step budget must be >= 1");
    nova_adjoint *ad = new nova_adjoint();
    ad->sv = sv;
    ad->ps = ps;
    ad->ctx = ctx;
    ad->nu = nu;
    ad->every = every;
    ad->max_steps = max_steps;
    ad->scheme = solver_scheme(sv);
    const size_t bytes = sizeof(double) * 3 * (size_t)ctx->g.np;
    const int nck = (max_steps + every - 1) / every;
    hipError_t e = hipSuccess;
    for (int q = 0; q < nck && e == hipSuccess; q++) {
        double *p = nullptr;
        e = hipMalloc((void **)&p, bytes);
        ad->ckpt.push_back(p);
    }
    ad->store.resize(1);
    ad->held.assign(1, -1);
    for (int q = 0; q < 3 * every && e == hipSuccess; q++) {
        double *p = nullptr;
        e = hipMalloc((void **)&p, bytes);
        ad->store[0].push_back(p);
    }
    for (int q = 0; q < 5 && e == hipSuccess; q++) e = hipMalloc((void **)&ad->w[q], bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&ad->ptmp, sizeof(double) * ctx->g.np);
    if (e == hipSuccess && ad->scheme == NOVA_SCHEME_COMPACT6)
        e = hipMalloc((void **)&ad->rwork, sizeof(double) * 4 * ctx->g.np);
    if (e != hipSuccess) {
        for (double *p : ad->ckpt) if (p) (void)hipFree(p);
        for (double *p : ad->store[0]) if (p) (void)hipFree(p);
        for (double *p : ad->w) if (p) (void)hipFree(p);
        if (ad->ptmp) (void)hipFree(ad->ptmp);
        if (ad->rwork) (void)hipFree(ad->rwork);
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
    for (auto &st : ad->store)
        for (double *p : st) if (p) (void)hipFree(p);
    for (double *p : ad->w) if (p) (void)hipFree(p);
    if (ad->ptmp) (void)hipFree(ad->ptmp);
    if (ad->rwork) (void)hipFree(ad->rwork);
    delete ad;
    return NOVA_OK;
}

// Segment stores beyond the first, within max_bytes of device memory (at
// most one per checkpoint but the first); 0 releases them.
int adjoint_reserve(nova_adjoint *ad, uint64_t max_bytes, int *count)
{
    NOVA_HIP(hipStreamSynchronize(ad->ctx->stream));
    for (size_t k = 1; k < ad->store.size(); k++)
        for (double *p : ad->store[k]) (void)hipFree(p);
    ad->store.resize(1);
    ad->held.resize(1);
    const size_t bytes = sizeof(double) * 3 * (size_t)ad->ctx->g.np;
    const uint64_t per = (uint64_t)bytes * 3 * ad->every;
    const int nck = (int)ad->ckpt.size();
    int want = (int)(max_bytes / per);
    if (want > nck - 1) want = nck - 1;
    for (int k = 0; k < want; k++) {
        std::vector<double *> st;
        if (alloc_store(bytes, 3 * ad->every, st)) break;  // as many as the device holds
        ad->store.push_back(std::move(st));
        ad->held.push_back(-1);
    }
    if (count) *count = (int)ad->store.size() - 1;
    return NOVA_OK;
}

int adjoint_forward(nova_adjoint *ad, double dt, int nsteps, const double *U0, double *U)
{
    if (nsteps < 1 || nsteps > ad->max_steps) return fail(NOVA_ERR_ARG, "nsteps outside the checkpoint budget");
    ad->dt = dt;
    ad->nsteps = nsteps;
    for (int &h : ad->held) h = -1;
    int rc;
    if (U != U0 && (rc = copy3(ad->ctx, U, U0))) return rc;
    const int C = ad->every, nseg = (nsteps + C - 1) / C;
    const int nst = (int)ad->store.size();
    const int first_held = nseg > nst ? nseg - nst : 0;
    for (int s = 0; s < nseg; s++) {
        const int n0 = s * C, cnt = n0 + C < nsteps ? C : nsteps - n0;
        if ((rc = copy3(ad->ctx, ad->ckpt[s], U))) return rc;
        if (s < first_held) {
            for (int q = 0; q < cnt; q++)
                if ((rc = fwd_step_store(ad, U, ad->w[0], ad->w[1], U))) return rc;
            continue;
        }
        // the last segments run in the stores (the last one in store 0, which
        // the reverse sweep frees first), keeping their stage inputs
        const int k = nseg - 1 - s;
        std::vector<double *> &st = ad->store[k];
        if ((rc = copy3(ad->ctx, st[0], U))) return rc;
        for (int q = 0; q < cnt; q++)
            if ((rc = fwd_step_store(ad, st[3 * q], st[3 * q + 1], st[3 * q + 2], q + 1 < cnt ? st[3 * q + 3] : U)))
                return rc;
        ad->held[k] = s;
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
    const bool compact = ad->scheme == NOVA_SCHEME_COMPACT6;
    for (int s = nseg - 1; s >= 0; s--) {
        int cnt = 0;
        double *const *seg = nullptr;
        if ((rc = segment_inputs(ad, s, &seg, &cnt))) return rc;
        for (int q = cnt - 1; q >= 0; q--) {
            double *U = seg[3 * q], *U1 = seg[3 * q + 1], *U2 = seg[3 * q + 2];
            if ((rc = project(ad->ps, L, 1.0, ad->ptmp))) return rc;  // m3 = P m
            if (compact) {
                // the pencil passes carry one pointwise addend and an accumulator
                const double dt = ad->dt, nu = ad->nu;
                if ((rc = compact_adj_stage(ctx, U2, L, A, 2.0 / 3.0, nullptr, 0.0, ACC, 0.0, 1.0 / 3.0, dt, nu,
                                            ad->rwork)))
                    return rc;
                if ((rc = compact_adj_stage(ctx, U1, A, B, 0.25, nullptr, 0.0, ACC, 1.0, 0.75, dt, nu, ad->rwork)))
                    return rc;
                if ((rc = compact_adj_stage(ctx, U, B, L, 1.0, ACC, 1.0, nullptr, 0.0, 0.0, dt, nu, ad->rwork)))
                    return rc;
            } else {
                if ((rc = adj_stage_scaled(ad, U2, L, A, 2.0 / 3.0))) return rc;            // m2
                if ((rc = adj_stage_scaled(ad, U1, A, B, 0.25))) return rc;                 // m1
                if ((rc = adj_stage_sum(ad, U, B, L, A, 0.75, L, 1.0 / 3.0))) return rc;    // dJ/dU
            }
        }
    }
    return NOVA_OK;
}

// Measured choice of the launch shapes of the adjoint and tangent-linear
// stage kernels, on the first checkpoint of the last forward run.
int adjoint_autotune(nova_adjoint *ad, int reps, char *report, size_t len)
{
    if (ad->nsteps < 1) return fail(NOVA_ERR_STATE, "run nova_adjoint_forward first");
    if (reps < 1) return fail(NOVA_ERR_ARG, "reps must be >= 1");
    nova_ctx *ctx = ad->ctx;
    std::string rep;
    if (ad->scheme == NOVA_SCHEME_COMPACT6)  // pencil passes have no launch-shape choice
        return write_report("compact scheme: no stage launch shapes to tune\n", report, len);
    const int variants[] = {0, 2, 3, 4, 5, 6, 11, 12, 13, 14};
    // the forward state and a distinct second field (a copy), so that the
    // trials move the same bytes as the reverse sweep
    const double *X = ad->ckpt[0];
    double *V = ad->w[2];
    if (int rc = copy3(ctx, V, X)) return rc;
    if (int rc = copy3(ctx, ad->w[3], X)) return rc;
    auto adj = [&]() -> int { return adj_stage_scaled(ad, X, V, ad->w[0], 1.0); };
    if (int rc = tune_stage_model(ctx, MODEL_ADJ, variants, 10, reps, adj, "adj_stage", rep)) return rc;
    auto tlm = [&]() -> int {
        LinArgs a = lin_args(ad, X, V);
        a.out[0] = ad->w[0];
        a.X0[1] = ad->w[3];
        a.sa = ssp_args(1, ad->dt);
        return launch_stream_model<MODEL_TLM>(ctx, a, "tlm_stage");
    };
    if (int rc = tune_stage_model(ctx, MODEL_TLM, variants, 10, reps, tlm, "tlm_stage", rep)) return rc;
    return write_report(rep, report, len);
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
        double *const *seg = nullptr;
        if ((rc = segment_inputs(ad, s, &seg, &cnt))) return rc;
        for (int q = 0; q < cnt; q++) {
            double *U = seg[3 * q], *U1 = seg[3 * q + 1], *U2 = seg[3 * q + 2];
            if ((rc = tlm_stage(ad, U, D, nullptr, D1, ssp_args(0, ad->dt)))) return rc;
            if ((rc = tlm_stage(ad, U1, D1, D, D2, ssp_args(1, ad->dt)))) return rc;
            if ((rc = tlm_stage(ad, U2, D2, D, D, ssp_args(2, ad->dt)))) return rc;
            if ((rc = project(ad->ps, D, ad->dt, ad->ptmp))) return rc;
        }
    }
    return NOVA_OK;
}

}  // namespace nova
