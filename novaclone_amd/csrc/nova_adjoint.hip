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
// Checkpoints of U every C steps are kept.  A pool of vector fields (3C
// from nova_adjoint_create, more from nova_adjoint_reserve) holds the step
// start states U of the forward sweep's last S = min(pool, nsteps) steps,
// one field per step; the reverse sweep rebuilds U1 and U2 of such a step
// with two forward stage launches.  The earlier steps are recomputed in the
// reverse sweep a segment at a time from their checkpoint, storing (U, U1,
// U2) per step in the first 3C pool fields, which hold steps the reverse
// sweep has already passed by then.  The tangent-linear model walks forward
// with a running state instead.
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
    // the pool: pool[q] holds U of step res_first + q while slot_step[q]
    // says so; fields 0 .. 3C-1 also take the (U, U1, U2) of a recomputed
    // segment (rc_seg), whose step n0 + q reads U from its checkpoint for
    // q = 0 and from pool[3q] otherwise, U1 from pool[3q+1], U2 from pool[3q+2]
    std::vector<double *> pool;
    std::vector<int> slot_step;
    int n_create = 0;             // pool fields allocated by nova_adjoint_create
    int res_first = 0, nres = 0;  // resident steps res_first .. res_first + nres - 1
    int rc_seg = -1, rc_cnt = 0;  // recomputed segment and its number of steps
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

// Whether pool slot q holds U of step n (a resident step).
bool resident(const nova_adjoint *ad, int n, int *q)
{
    const int k = n - ad->res_first;
    if (k < 0 || k >= ad->nres || ad->slot_step[k] != n) return false;
    *q = k;
    return true;
}

// U1 and U2 of a step from its U, into the work fields w[3] and w[4].
int rebuild_stages(nova_adjoint *ad, double *U, double *st[3])
{
    int rc;
    st[0] = U;
    st[1] = ad->w[3];
    st[2] = ad->w[4];
    if ((rc = fwd_stage(ad, U, nullptr, st[1], ssp_args(0, ad->dt)))) return rc;
    return fwd_stage(ad, st[1], U, st[2], ssp_args(1, ad->dt));
}

// Recompute steps n0 .. n of segment s = n / C from its checkpoint, storing
// (U, U1, U2) per step in the pool's first 3C fields; the resident steps
// those fields held are dropped.
int recompute_segment(nova_adjoint *ad, int n)
{
    const int C = ad->every, s = n / C, n0 = s * C, cnt = n - n0 + 1;
    int rc;
    ad->rc_seg = -1;
    for (int q = 1; q < 3 * cnt; q++) ad->slot_step[q] = -1;
    std::vector<double *> &P = ad->pool;
    for (int q = 0; q < cnt; q++) {
        double *U = q == 0 ? ad->ckpt[s] : P[3 * q];
        if ((rc = fwd_stage(ad, U, nullptr, P[3 * q + 1], ssp_args(0, ad->dt)))) return rc;
        if ((rc = fwd_stage(ad, P[3 * q + 1], U, P[3 * q + 2], ssp_args(1, ad->dt)))) return rc;
        if (q + 1 < cnt) {  // the last step's next state is not needed
            if ((rc = fwd_stage(ad, P[3 * q + 2], U, P[3 * q + 3], ssp_args(2, ad->dt)))) return rc;
            if ((rc = project(ad->ps, P[3 * q + 3], ad->dt, ad->ptmp))) return rc;
        }
    }
    ad->rc_seg = s;
    ad->rc_cnt = cnt;
    return NOVA_OK;
}

// (U, U1, U2) of step n held by the recomputed segment, if it holds them.
bool recomputed(nova_adjoint *ad, int n, double *st[3])
{
    const int C = ad->every, q = n - ad->rc_seg * C;
    if (ad->rc_seg < 0 || n / C != ad->rc_seg || q >= ad->rc_cnt) return false;
    st[0] = q == 0 ? ad->ckpt[ad->rc_seg] : ad->pool[3 * q];
    st[1] = ad->pool[3 * q + 1];
    st[2] = ad->pool[3 * q + 2];
    return true;
}

// The stage inputs of step n for the reverse sweep.
int reverse_inputs(nova_adjoint *ad, int n, double *st[3])
{
    int q;
    if (resident(ad, n, &q)) return rebuild_stages(ad, ad->pool[q], st);
    if (!recomputed(ad, n, st)) {
        if (int rc = recompute_segment(ad, n)) return rc;
        recomputed(ad, n, st);
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
    for (int q = 0; q < 3 * every && e == hipSuccess; q++) {
        double *p = nullptr;
        e = hipMalloc((void **)&p, bytes);
        ad->pool.push_back(p);
    }
    for (int q = 0; q < 5 && e == hipSuccess; q++) e = hipMalloc((void **)&ad->w[q], bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&ad->ptmp, sizeof(double) * ctx->g.np);
    if (e == hipSuccess && ad->scheme == NOVA_SCHEME_COMPACT6)
        e = hipMalloc((void **)&ad->rwork, sizeof(double) * 4 * ctx->g.np);
    if (e != hipSuccess) {
        for (double *p : ad->ckpt) if (p) (void)hipFree(p);
        for (double *p : ad->pool) if (p) (void)hipFree(p);
        for (double *p : ad->w) if (p) (void)hipFree(p);
        if (ad->ptmp) (void)hipFree(ad->ptmp);
        if (ad->rwork) (void)hipFree(ad->rwork);
        delete ad;
        return hip_fail(e, "adjoint checkpoint memory");
    }
    ad->n_create = (int)ad->pool.size();
    ad->slot_step.assign(ad->pool.size(), -1);
    *out = ad;
    return NOVA_OK;
}

int adjoint_destroy(nova_adjoint *ad)
{
    if (!ad) return NOVA_OK;
    (void)hipStreamSynchronize(ad->ctx->stream);
    for (double *p : ad->ckpt) if (p) (void)hipFree(p);
    for (double *p : ad->pool) if (p) (void)hipFree(p);
    for (double *p : ad->w) if (p) (void)hipFree(p);
    if (ad->ptmp) (void)hipFree(ad->ptmp);
    if (ad->rwork) (void)hipFree(ad->rwork);
    delete ad;
    return NOVA_OK;
}

// Pool fields beyond the 3C of nova_adjoint_create within max_bytes of
// device memory (at most one per step of the budget beyond those); 0
// releases them.
int adjoint_reserve(nova_adjoint *ad, uint64_t max_bytes, int *count)
{
    NOVA_HIP(hipStreamSynchronize(ad->ctx->stream));
    for (size_t k = ad->n_create; k < ad->pool.size(); k++) (void)hipFree(ad->pool[k]);
    ad->pool.resize(ad->n_create);
    const size_t bytes = sizeof(double) * 3 * (size_t)ad->ctx->g.np;
    uint64_t want = max_bytes / bytes;
    const int room = ad->max_steps - ad->n_create;
    if (want > (uint64_t)(room > 0 ? room : 0)) want = room > 0 ? room : 0;
    for (uint64_t k = 0; k < want; k++) {
        double *p = nullptr;
        if (hipMalloc((void **)&p, bytes) != hipSuccess) {  // as many as the device holds
            (void)hipGetLastError();  // an allocation failure is not sticky
            break;
        }
        ad->pool.push_back(p);
    }
    // the resident steps of the last forward sweep are kept where their
    // slots survive
    ad->slot_step.resize(ad->pool.size(), -1);
    if (ad->nres > (int)ad->pool.size()) ad->nres = (int)ad->pool.size();
    if (count) *count = (int)ad->pool.size() - ad->n_create;
    return NOVA_OK;
}

int adjoint_forward(nova_adjoint *ad, double dt, int nsteps, const double *U0, double *U)
{
    if (nsteps < 1 || nsteps > ad->max_steps) return fail(NOVA_ERR_ARG, "nsteps outside the checkpoint budget");
    ad->dt = dt;
    ad->nsteps = nsteps;
    const int P = (int)ad->pool.size();
    ad->nres = P < nsteps ? P : nsteps;
    ad->res_first = nsteps - ad->nres;
    ad->rc_seg = -1;
    for (int &t : ad->slot_step) t = -1;
    int rc;
    const int C = ad->every;
    // the running state: U, until the resident steps start in the pool
    double *cur = U;
    if (ad->res_first == 0) cur = ad->pool[0];
    if (cur != U0 && (rc = copy3(ad->ctx, cur, U0))) return rc;
    for (int n = 0; n < nsteps; n++) {
        if (n % C == 0 && (rc = copy3(ad->ctx, ad->ckpt[n / C], cur))) return rc;
        double *next = U;
        if (n + 1 >= ad->res_first && n + 1 < nsteps) next = ad->pool[n + 1 - ad->res_first];
        if (n >= ad->res_first) ad->slot_step[n - ad->res_first] = n;
        if ((rc = fwd_step_store(ad, cur, ad->w[0], ad->w[1], next))) return rc;
        cur = next;
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
    const bool compact = ad->scheme == NOVA_SCHEME_COMPACT6;
    for (int n = ad->nsteps - 1; n >= 0; n--) {
        double *st[3];
        if ((rc = reverse_inputs(ad, n, st))) return rc;
        double *U = st[0], *U1 = st[1], *U2 = st[2];
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
    constexpr int nv = (int)(sizeof(variants) / sizeof(variants[0]));
    if (int rc = tune_stage_model(ctx, MODEL_ADJ, variants, nv, reps, adj, "adj_stage", rep)) return rc;
    auto tlm = [&]() -> int {
        LinArgs a = lin_args(ad, X, V);
        a.out[0] = ad->w[0];
        a.X0[1] = ad->w[3];
        a.sa = ssp_args(1, ad->dt);
        return launch_stream_model<MODEL_TLM>(ctx, a, "tlm_stage");
    };
    if (int rc = tune_stage_model(ctx, MODEL_TLM, variants, nv, reps, tlm, "tlm_stage", rep)) return rc;
    return write_report(rep, report, len);
}

int adjoint_tlm(nova_adjoint *ad, const double *d0, double *dT)
{
    if (ad->nsteps < 1) return fail(NOVA_ERR_STATE, "run nova_adjoint_forward first");
    nova_ctx *ctx = ad->ctx;
    double *D = dT, *D1 = ad->w[0], *D2 = ad->w[1];
    // steps neither resident nor recomputed walk forward from their
    // checkpoint in w[2] (cur_n: the step whose U it holds, -1: none)
    double *cur = ad->w[2];
    int cur_n = -1;
    int rc;
    if (D != d0 && (rc = copy3(ctx, D, d0))) return rc;
    for (int n = 0; n < ad->nsteps; n++) {
        double *st[3];
        int q;
        bool walk = false;
        if (resident(ad, n, &q)) {
            if ((rc = rebuild_stages(ad, ad->pool[q], st))) return rc;
        } else if (!recomputed(ad, n, st)) {
            if (cur_n != n) {
                const int s = n / ad->every;
                if ((rc = copy3(ctx, cur, ad->ckpt[s]))) return rc;
                for (int k = s * ad->every; k < n; k++)
                    if ((rc = fwd_step_store(ad, cur, ad->w[3], ad->w[4], cur))) return rc;
            }
            if ((rc = rebuild_stages(ad, cur, st))) return rc;
            walk = true;
        }
        if ((rc = tlm_stage(ad, st[0], D, nullptr, D1, ssp_args(0, ad->dt)))) return rc;
        if ((rc = tlm_stage(ad, st[1], D1, D, D2, ssp_args(1, ad->dt)))) return rc;
        if ((rc = tlm_stage(ad, st[2], D2, D, D, ssp_args(2, ad->dt)))) return rc;
        if ((rc = project(ad->ps, D, ad->dt, ad->ptmp))) return rc;
        cur_n = -1;
        if (walk && n + 1 < ad->nsteps) {  // advance the running state
            if ((rc = fwd_stage(ad, st[2], cur, cur, ssp_args(2, ad->dt)))) return rc;
            if ((rc = project(ad->ps, cur, ad->dt, ad->ptmp))) return rc;
            cur_n = n + 1;
        }
    }
    return NOVA_OK;
}

}  // namespace nova
