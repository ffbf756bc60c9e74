// nova_capi.hip -- the C ABI: context set-up (the reference Mesh on one
// device), memory, and thin wrappers over the kernel launchers.
#include "nova_internal.hpp"

#include <cmath>
#include <cstring>
#include <new>

struct nova_poisson;
struct nova_solver;
struct nova_adjoint;

namespace nova {

namespace {
thread_local std::string g_last_error;

// Central-difference weights, src/settings/settings.cpp:25-31.
const double kCoef[4][4] = {{1.0, 0.0, 0.0, 0.0},
                            {4.0 / 3.0, -1.0 / 3.0, 0.0, 0.0},
                            {6.0 / 4.0, -12.0 / 20.0, 1.0 / 10.0, 0.0},
                            {8.0 / 5.0, -4.0 / 5.0, 24.0 / 105.0, -8.0 / 280.0}};
}  // namespace

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return NOVA_ERR_HIP;
}

int write_report(const std::string &rep, char *buf, size_t len)
{
    if (!buf) return NOVA_OK;
    if (rep.size() + 1 > len) return fail(NOVA_ERR_ARG, "report buffer too small");
    std::memcpy(buf, rep.c_str(), rep.size() + 1);
    return NOVA_OK;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return NOVA_OK;
}

static hipEvent_t take_event(nova_ctx *ctx)
{
    if (!ctx->evpool.empty()) {
        hipEvent_t e = ctx->evpool.back();
        ctx->evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedScope::TimedScope(nova_ctx *ctx, const char *name) : ctx_(ctx), name_(name)
{
    if (!ctx_->timing) return;
    start_ = take_event(ctx_);
    if (start_) (void)hipEventRecord(start_, ctx_->stream);
}

TimedScope::~TimedScope()
{
    if (!ctx_->timing || !start_) return;
    hipEvent_t stop = take_event(ctx_);
    if (!stop) return;
    (void)hipEventRecord(stop, ctx_->stream);
    ctx_->tlog.push_back({name_, start_, stop});
}

int ensure_red(nova_ctx *ctx, size_t n)
{
    if (ctx->red_cap >= n) return NOVA_OK;
    if (ctx->d_red) NOVA_HIP(hipFree(ctx->d_red));
    ctx->d_red = nullptr;
    ctx->red_cap = 0;
    NOVA_HIP(hipMalloc((void **)&ctx->d_red, sizeof(double) * n));
    ctx->red_cap = n;
    return NOVA_OK;
}

int poisson_create(nova_ctx *ctx, nova_poisson **out, int batch_cap = 1);
int poisson_destroy(nova_poisson *ps);
int poisson_solve(nova_poisson *ps, const double *rhs, double *p);
int project(nova_poisson *ps, double *F, double s, double *p);
int solver_create(nova_poisson *ps, int model, int mode, double nu, double eta, nova_solver **out);
int solver_destroy(nova_solver *sv);
int solver_step(nova_solver *sv, double dt, int nsteps, double *U, double *B, double *p, double *pB);
int solver_use_graph(nova_solver *sv, int enable);
int solver_autotune(nova_solver *sv, const double *U, const double *B, int reps, char *report, size_t len);
int solver_set_scheme(nova_solver *sv, int scheme);
int solver_set_force(nova_solver *sv, const double *F);
int poisson_set_wall_solve(nova_poisson *ps, int mode);
int poisson_set_scheme(nova_poisson *ps, int scheme);
nova_ctx *poisson_ctx(nova_poisson *ps);
int adjoint_create(nova_solver *sv, int every, int max_steps, nova_adjoint **out);
int adjoint_destroy(nova_adjoint *ad);
int adjoint_forward(nova_adjoint *ad, double dt, int nsteps, const double *U0, double *U);
int adjoint_backward(nova_adjoint *ad, const double *lamT, double *lam0);
int adjoint_tlm(nova_adjoint *ad, const double *d0, double *dT);
int adjoint_autotune(nova_adjoint *ad, int reps, char *report, size_t len);
int adjoint_reserve(nova_adjoint *ad, uint64_t max_bytes, int *count);
int ensemble_create(const nova_instance_desc *inst, int n, int device, nova_ensemble **out);
int ensemble_destroy(nova_ensemble *e);
int ensemble_set_stream(nova_ensemble *e, hipStream_t st);
int ensemble_copy(nova_ensemble *e, int i, int store, double *U, double *B, double *p, double *pB);
int ensemble_step(nova_ensemble *e, double dt, int nsteps);
int ensemble_use_graph(nova_ensemble *e, int enable);
int ensemble_info(nova_ensemble *e, int64_t out[6]);
int wc_create(nova_ctx *ctx, double mu, double pc, nova_wc **out);
int wc_destroy(nova_wc *w);
int wc_step(nova_wc *w, double dt, int nsteps, double *U, double *rho, double *p);
int wc_set_magnetic(nova_wc *w, const double *B, double dbu);
int wc_set_temperature(nova_wc *w, double ka, const double g[3]);
int wc_step_t(nova_wc *w, double dt, int nsteps, double *U, double *rho, double *p, double *T);
int wc_hydrostatic(nova_wc *w, const double *T, double rho0, double *p, double *rho);
int fd_laplacian(nova_ctx *ctx, double mu, const double *phi, nova_fdmatrix **out);
int fd_destroy(nova_fdmatrix *m);
int fd_add_source(nova_fdmatrix *m, const double *f);
int fd_relax(nova_fdmatrix *m, const double *phi, double urf);
int fd_set_reference(nova_fdmatrix *m, int i, int j, int k, double val);
int fd_combine(nova_fdmatrix *m, const nova_fdmatrix *o, double s);
int fd_residual(nova_fdmatrix *m, const double *phi, double *res);
int fd_solve(nova_fdmatrix *m, double *phi, double tol, int max_its, int *its_out, double *res_out);
int fd_get(const nova_fdmatrix *m, double *dst, int what);
int fd_same_grid(const nova_fdmatrix *a, const nova_fdmatrix *b);
int fd_inside(const nova_fdmatrix *m, int i, int j, int k);
int time_create(double dt, double end, int write_every, double interval, bool timebased, nova_time **out);
int time_create_from_file(const char *file, nova_time **out);
int time_restart(nova_time *t, const char *file);
int time_advance(nova_time *t);
void time_destroy(nova_time *t);
void time_set_dt(nova_time *t, double dt);
void time_info(const nova_time *t, double dout[4], int64_t iout[5]);
void time_steps(const nova_time *t, double *out, int64_t cap);
int io_create(nova_ctx *ctx, nova_io **out);
int io_destroy(nova_io *io);
int io_wait(nova_io *io);
int io_write_field(nova_io *io, const nova_time *t, const char *path, const char *name, const double *F, int nc);
int io_write_mesh(nova_io *io, const nova_time *t, const char *path);
int io_read_field(nova_io *io, const char *path, const char *name, int file_num, double *F, int nc);
int io_checkpoint_write(nova_io *io, const char *file, int nfields, const double *const *F, const int *nc,
                        double time, int64_t step);
int io_checkpoint_read(nova_io *io, const char *file, int nfields, double *const *F, const int *nc, double *time,
                       int64_t *step);

// Mesh::z (src/Mesh/Mesh.cpp:373-407) for padded index kp on one rank.
static double mesh_z(const nova_grid_desc &d, double dz, double hb, int m, int kp)
{
    if (d.hyp) {
        const double A = std::pow(d.s1 / d.s2, 0.5);
        const int globK = kp - m / 2;
        const int globN = d.n[2] + 1;
        const double R = double(globK) / double(globN - 1) - 0.5;
        const double U = 1.0 + std::tanh(hb * R) / std::tanh(hb / 2.0);
        return U / (2.0 * A + (1 - A) * U) * d.L[2];
    }
    return (kp - m / 2) * dz + d.origin[2];
}

// z of padded index kp (Mesh::z) for the output coordinates.
double mesh_z_of(const nova_ctx *ctx, int kp) { return mesh_z(ctx->desc, ctx->d[2], ctx->hb, ctx->m, kp); }

static int ctx_init(nova_ctx *ctx, const nova_grid_desc *desc, int device)
{
    const nova_grid_desc &d = *desc;
    if (d.m != 2 && d.m != 4 && d.m != 6 && d.m != 8) return fail(NOVA_ERR_ARG, "m must be 2, 4, 6 or 8");
    const int h = d.m / 2;
    for (int a = 0; a < 3; a++) {
        if (d.n[a] < 2 * h || d.n[a] < 4) return fail(NOVA_ERR_ARG, "grid too small for the stencil");
        if (!(d.L[a] > 0.0)) return fail(NOVA_ERR_ARG, "lengths must be positive");
        if (d.bc[a] != NOVA_BC_PERIODIC && d.bc[a] != NOVA_BC_WALL) return fail(NOVA_ERR_ARG, "bad bc kind");
    }
    if (d.hyp && !(d.s1 > 0.0 && d.s2 > 0.0)) return fail(NOVA_ERR_ARG, "stretching needs s1, s2 > 0");
    // stage kernels address within an (j, k) plane by 32-bit byte offsets
    if ((int64_t)(d.n[1] + 1) * (d.n[2] + 1) >= ((int64_t)1 << 29))
        return fail(NOVA_ERR_UNSUPPORTED, "a (j, k) plane must hold fewer than 2^29 nodes");
    ctx->device = device;
    ctx->desc = d;
    ctx->m = d.m;
    for (int l = 0; l < 4; l++) ctx->coef[l] = kCoef[h - 1][l];
    Grid &g = ctx->g;
    g.h = h;
    for (int a = 0; a < 3; a++) {
        g.n[a] = d.n[a];
        g.wall[a] = d.bc[a] == NOVA_BC_WALL;
        g.odd[a] = 1u << a;  // the wall-normal velocity: dirichletBC 0; the others neumannBC 0
        g.s[a] = d.n[a] + (g.wall[a] ? 1 : 0);
        ctx->d[a] = d.L[a] / d.n[a];  // Mesh.cpp:68-70
        ctx->pad[a] = d.n[a] + 1 + d.m;  // Mesh.cpp:89-91,136-138
        for (int l = 0; l < kMaxH; l++) {
            const double c = l < h ? ctx->coef[l] : 0.0;
            g.a[a][l] = c / 2.0 / (l + 1) / ctx->d[a];
            g.b[a][l] = c / std::pow((l + 1) * ctx->d[a], 2);
            g.al[a][l] = c / (l + 1) / ctx->d[a];
        }
    }
    g.wb_sum2 = 0.0;
    for (int l = 0; l < kMaxH; l++) {
        const double c = l < h ? ctx->coef[l] : 0.0;
        g.wa[l] = c / 2.0 / (l + 1);
        g.wb[l] = c / ((l + 1) * (l + 1));
        g.wb_sum2 += 2.0 * g.wb[l];
    }
    for (int a = 0; a < 3; a++) {
        g.ih[a] = 1.0 / ctx->d[a];
        g.ih2[a] = 1.0 / (ctx->d[a] * ctx->d[a]);
    }
    g.np = (int64_t)g.s[0] * g.s[1] * g.s[2];
    // z metric (Mesh.cpp:147-179) on the stored k range
    double hb = 1.0;
    if (d.hyp) {
        const int globN = d.n[2] + 1;
        for (int it = 1; it < 10000; it++) hb = std::asinh(hb / ((globN - 1) * std::pow(d.s1 * d.s2, 0.5)));
    }
    ctx->hb = hb;
    const double dz = ctx->d[2];
    const int s2 = g.s[2];
    ctx->dzdxi.assign(s2, 0.0);
    ctx->d2zdxi2.assign(s2, 0.0);
    for (int k = 0; k < s2; k++) {
        const int kp = k + h;
        double a1 = 0.0, a2 = 0.0;
        for (int l = 0; l < h; l++) {
            const double c = ctx->coef[l];
            a1 += (c / 2.0 / (l + 1) / dz * mesh_z(d, dz, hb, d.m, kp + 1 + l) -
                   c / 2.0 / (l + 1) / dz * mesh_z(d, dz, hb, d.m, kp - 1 - l));
            a2 += (c / std::pow((l + 1) * dz, 2) * mesh_z(d, dz, hb, d.m, kp + 1 + l) +
                   c / std::pow((l + 1) * dz, 2) * mesh_z(d, dz, hb, d.m, kp - 1 - l) +
                   (-2.0 * c / std::pow((l + 1) * dz, 2)) * mesh_z(d, dz, hb, d.m, kp));
        }
        ctx->dzdxi[k] = a1;
        ctx->d2zdxi2[k] = a2;
    }
    std::vector<double> met(3 * (size_t)s2);
    for (int k = 0; k < s2; k++) {
        met[k] = 1.0 / ctx->dzdxi[k];
        met[s2 + k] = 1.0 / std::pow(ctx->dzdxi[k], 2);
        met[2 * s2 + k] = ctx->d2zdxi2[k] / std::pow(ctx->dzdxi[k], 3);
    }
    NOVA_HIP(hipSetDevice(device));
    NOVA_HIP(hipMalloc((void **)&ctx->d_metric, sizeof(double) * met.size()));
    NOVA_HIP(hipMemcpy(ctx->d_metric, met.data(), sizeof(double) * met.size(), hipMemcpyHostToDevice));
    g.zinv = ctx->d_metric;
    g.zlap = ctx->d_metric + s2;
    g.zmet = ctx->d_metric + 2 * s2;
    NOVA_HIP(hipHostMalloc((void **)&ctx->h_red, sizeof(double) * 16, hipHostMallocDefault));
    // the device's copy of the tdma workspace, resolved with the device current
    return tdma_scratch_init(device);
}

static void ctx_release(nova_ctx *ctx)
{
    for (auto &t : ctx->tlog) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto e : ctx->evpool) (void)hipEventDestroy(e);
    if (ctx->d_metric) (void)hipFree(ctx->d_metric);
    if (ctx->d_red) (void)hipFree(ctx->d_red);
    if (ctx->h_red) (void)hipHostFree(ctx->h_red);
}

}  // namespace nova

using namespace nova;

#define NOVA_GUARD(expr)                                                         \
    try {                                                                        \
        return (expr);                                                           \
    } catch (const std::bad_alloc &) {                                           \
        return fail(NOVA_ERR_NOMEM, "host allocation failed");                   \
    } catch (...) {                                                              \
        return fail(NOVA_ERR_STATE, "unexpected exception");                     \
    }

#define NOVA_REQUIRE(cond, msg) \
    do {                        \
        if (!(cond)) return fail(NOVA_ERR_ARG, msg); \
    } while (0)

extern "C" {

const char *nova_last_error(void) { return g_last_error.c_str(); }

int nova_ctx_create(const nova_grid_desc *desc, int device, nova_ctx **out)
{
    NOVA_REQUIRE(desc && out, "null argument");
    nova_ctx *ctx = new (std::nothrow) nova_ctx();
    if (!ctx) return fail(NOVA_ERR_NOMEM, "host allocation failed");
    int rc;
    try {
        rc = ctx_init(ctx, desc, device);
    } catch (...) {
        rc = fail(NOVA_ERR_NOMEM, "context set-up failed");
    }
    if (rc) {
        ctx_release(ctx);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return NOVA_OK;
}

int nova_ctx_destroy(nova_ctx *ctx)
{
    if (!ctx) return NOVA_OK;
    (void)hipStreamSynchronize(ctx->stream);
    ctx_release(ctx);
    delete ctx;
    return NOVA_OK;
}

int nova_ctx_set_stream(nova_ctx *ctx, void *stream)
{
    NOVA_REQUIRE(ctx, "null context");
    ctx->stream = (hipStream_t)stream;
    return NOVA_OK;
}

int nova_ctx_sync(nova_ctx *ctx)
{
    NOVA_REQUIRE(ctx, "null context");
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    return NOVA_OK;
}

int nova_ctx_info(const nova_ctx *ctx, int64_t out[8])
{
    NOVA_REQUIRE(ctx && out, "null argument");
    out[0] = ctx->g.s[0];
    out[1] = ctx->g.s[1];
    out[2] = ctx->g.s[2];
    out[3] = ctx->g.np;
    out[4] = ctx->pad[0];
    out[5] = ctx->pad[1];
    out[6] = ctx->pad[2];
    out[7] = ctx->m;
    return NOVA_OK;
}

int nova_ctx_metric(const nova_ctx *ctx, double *dzdxi, double *d2zdxi2)
{
    NOVA_REQUIRE(ctx, "null context");
    if (dzdxi) std::memcpy(dzdxi, ctx->dzdxi.data(), sizeof(double) * ctx->dzdxi.size());
    if (d2zdxi2) std::memcpy(d2zdxi2, ctx->d2zdxi2.data(), sizeof(double) * ctx->d2zdxi2.size());
    return NOVA_OK;
}

int nova_timing_enable(nova_ctx *ctx, int enable)
{
    NOVA_REQUIRE(ctx, "null context");
    ctx->timing = enable != 0;
    return NOVA_OK;
}

int nova_timing_collect(nova_ctx *ctx, char *buf, size_t buflen)
{
    NOVA_REQUIRE(ctx && buf && buflen > 0, "bad argument");
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<std::string> names;
    std::vector<double> tot;
    std::vector<long> cnt;
    for (auto &t : ctx->tlog) {
        float ms = 0.0f;
        NOVA_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        size_t q = 0;
        while (q < names.size() && names[q] != t.name) q++;
        if (q == names.size()) {
            names.push_back(t.name);
            tot.push_back(0.0);
            cnt.push_back(0);
        }
        tot[q] += ms;
        cnt[q] += 1;
        ctx->evpool.push_back(t.start);
        ctx->evpool.push_back(t.stop);
    }
    ctx->tlog.clear();
    std::string out;
    for (size_t q = 0; q < names.size(); q++)
        out += names[q] + " " + std::to_string(cnt[q]) + " " + std::to_string(tot[q]) + "\n";
    if (out.size() + 1 > buflen) return fail(NOVA_ERR_ARG, "timing buffer too small");
    std::memcpy(buf, out.c_str(), out.size() + 1);
    return NOVA_OK;
}

int nova_ctx_tune(nova_ctx *ctx, int stage_variant, int stage_ichunk, int proj_variant)
{
    NOVA_REQUIRE(ctx && stage_variant >= 0 && stage_variant <= 14 && stage_ichunk >= 0 && proj_variant >= 0 &&
                     proj_variant <= 2,
                 "bad argument");
    for (int k = 0; k < 5; k++) {
        ctx->stage_variant[k] = stage_variant;
        ctx->stage_ichunk[k] = stage_ichunk;
    }
    ctx->proj_variant = proj_variant;
    ctx->tune_epoch++;
    return NOVA_OK;
}

int nova_ctx_tune_kernel(nova_ctx *ctx, int kernel, int stage_variant, int stage_ichunk)
{
    NOVA_REQUIRE(ctx && kernel >= NOVA_KERNEL_NS_STAGE && kernel <= NOVA_KERNEL_NS_FIRST_STAGE && stage_variant >= 0 &&
                     stage_variant <= 14 && stage_ichunk >= 0,
                 "bad argument");
    ctx->stage_variant[kernel] = stage_variant;
    ctx->stage_ichunk[kernel] = stage_ichunk;
    ctx->tune_epoch++;
    return NOVA_OK;
}

int nova_alloc(nova_ctx *ctx, int ncomp, double **dptr)
{
    NOVA_REQUIRE(ctx && dptr && ncomp > 0, "bad argument");
    NOVA_HIP(hipMalloc((void **)dptr, sizeof(double) * (size_t)ncomp * ctx->g.np));
    return NOVA_OK;
}

int nova_free(nova_ctx *ctx, double *dptr)
{
    NOVA_REQUIRE(ctx, "null context");
    if (dptr) NOVA_HIP(hipFree(dptr));
    return NOVA_OK;
}

int nova_memcpy_h2d(nova_ctx *ctx, double *dst, const double *src, size_t count)
{
    NOVA_REQUIRE(ctx && dst && src, "null argument");
    NOVA_HIP(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyHostToDevice, ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    return NOVA_OK;
}

int nova_memcpy_d2h(nova_ctx *ctx, double *dst, const double *src, size_t count)
{
    NOVA_REQUIRE(ctx && dst && src, "null argument");
    NOVA_HIP(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    return NOVA_OK;
}

int nova_host_alloc(size_t bytes, void **ptr)
{
    NOVA_REQUIRE(ptr && bytes > 0, "bad argument");
    *ptr = nullptr;
    NOVA_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return NOVA_OK;
}

int nova_host_free(void *ptr)
{
    if (ptr) NOVA_HIP(hipHostFree(ptr));
    return NOVA_OK;
}

int nova_pack_padded(nova_ctx *ctx, const double *compact, int ncomp, double *padded)
{
    NOVA_REQUIRE(ctx && compact && padded && (ncomp == 1 || ncomp == 3), "bad argument");
    NOVA_GUARD(launch_pack(ctx, compact, ncomp, padded));
}

int nova_unpack_padded(nova_ctx *ctx, const double *padded, int ncomp, double *compact)
{
    NOVA_REQUIRE(ctx && compact && padded && ncomp > 0, "bad argument");
    NOVA_GUARD(launch_unpack(ctx, padded, ncomp, compact));
}

int nova_bc_fill_padded(nova_ctx *ctx, double *padded, int ncomp, int nent, const int *dir, const int *kind,
                        const int *comp, const double *val)
{
    NOVA_REQUIRE(ctx && padded && ncomp > 0 && nent >= 0 && (nent == 0 || (dir && kind && comp)), "bad argument");
    for (int e = 0; e < nent; e++) {
        NOVA_REQUIRE(comp[e] >= 0 && comp[e] < ncomp, "bc component out of range");
        if (int rc = launch_bc_fill(ctx, padded, ncomp, dir[e], kind[e], comp[e], val ? val[e] : 0.0)) return rc;
    }
    return NOVA_OK;
}

int nova_div_adv(nova_ctx *ctx, const double *U, const double *phi, int ncomp_phi, double *out)
{
    NOVA_REQUIRE(ctx && U && phi && out && (ncomp_phi == 1 || ncomp_phi == 3), "bad argument");
    NOVA_REQUIRE(out != U && out != phi, "out must not alias an input");
    NOVA_GUARD(launch_div_adv(ctx, U, phi, ncomp_phi, out));
}

int nova_div(nova_ctx *ctx, const double *U, double *out)
{
    NOVA_REQUIRE(ctx && U && out && out != U, "bad argument");
    NOVA_GUARD(launch_div(ctx, U, 1.0, out));
}

int nova_grad(nova_ctx *ctx, const double *p, double *out)
{
    NOVA_REQUIRE(ctx && p && out && out != p, "bad argument");
    NOVA_GUARD(launch_grad(ctx, p, out));
}

int nova_laplacian(nova_ctx *ctx, double nu, const double *f, int ncomp, double *out)
{
    NOVA_REQUIRE(ctx && f && out && out != f && (ncomp == 1 || ncomp == 3), "bad argument");
    NOVA_GUARD(launch_laplacian(ctx, nu, f, ncomp, out));
}

int nova_laplacian_var(nova_ctx *ctx, const double *nu, const double *f, int ncomp, double *out)
{
    NOVA_REQUIRE(ctx && nu && f && out && out != f && out != nu && (ncomp == 1 || ncomp == 3), "bad argument");
    NOVA_GUARD(launch_laplacian_var(ctx, nu, f, ncomp, out));
}

int nova_curl_literal(nova_ctx *ctx, const double *U, double *out)
{
    NOVA_REQUIRE(ctx && U && out && out != U, "bad argument");
    NOVA_GUARD(launch_curl(ctx, U, 1, out));
}

int nova_curl(nova_ctx *ctx, const double *U, double *out)
{
    NOVA_REQUIRE(ctx && U && out && out != U, "bad argument");
    NOVA_GUARD(launch_curl(ctx, U, 0, out));
}

int nova_axpby(nova_ctx *ctx, size_t count, double a, const double *x, double b, const double *y, double *out)
{
    NOVA_REQUIRE(ctx && x && out, "bad argument");
    NOVA_GUARD(launch_axpby(ctx, count, a, x, b, y, out));
}

int nova_field_op(nova_ctx *ctx, int op, const double *a, int nca, const double *b, int ncb, const double *cval,
                  double *out)
{
    NOVA_REQUIRE(ctx && out && (a || op == NOVA_OP_FILL), "bad argument");
    bool ok = false;
    switch (op) {
    case NOVA_OP_DOT: ok = b && nca == ncb && (nca == 1 || nca == 3); break;
    case NOVA_OP_CROSS: ok = b && nca == 3 && ncb == 3 && out != a && out != b; break;
    case NOVA_OP_SCALE: ok = b && nca == 1 && (ncb == 1 || ncb == 3 || ncb == 9) && out != a; break;
    case NOVA_OP_DIVIDE: ok = b && ncb == 1 && (nca == 1 || nca == 3 || nca == 9) && out != b; break;
    case NOVA_OP_DDOT: ok = b && nca == 9 && ncb == 9; break;
    case NOVA_OP_SQRT: ok = nca == 1; break;
    case NOVA_OP_OUTER: ok = b && nca == 3 && ncb == 3 && out != a && out != b; break;
    case NOVA_OP_ADD_CONST: ok = cval && (nca == 1 || nca == 3); break;
    case NOVA_OP_FILL: ok = cval && nca >= 1 && nca <= 9; break;
    default: break;
    }
    NOVA_REQUIRE(ok, "nova_field_op: operator and component counts do not match (see nova_capi.h)");
    NOVA_GUARD(field_op(ctx, op, a, nca, b, ncb, cval, out));
}

int nova_grad_vec_literal(nova_ctx *ctx, const double *U, double *out9)
{
    NOVA_REQUIRE(ctx && U && out9 && U != out9, "bad argument");
    NOVA_GUARD(grad_vec_literal(ctx, U, out9));
}

int nova_cfl(nova_ctx *ctx, const double *U, double dt, double out[6])
{
    NOVA_REQUIRE(ctx && U && out, "bad argument");
    NOVA_GUARD(reduce_cfl(ctx, U, dt, out));
}

int nova_hdf5_write(nova_ctx *ctx, const char *file, const double *F, int ncomp, int comp)
{
    NOVA_REQUIRE(ctx && file && F && comp >= 0 && comp < ncomp, "bad argument");
    NOVA_GUARD(hdf5_write_field(ctx, file, F, comp));
}

int nova_hdf5_write_host(const char *file, const int64_t dims[3], const int64_t stored[3], const double *host)
{
    NOVA_REQUIRE(file && dims && stored && host, "bad argument");
    NOVA_GUARD(hdf5_write_host(file, dims, stored, host));
}

int nova_hdf5_read_host(const char *file, int64_t dims[3], double *host, size_t cap)
{
    NOVA_REQUIRE(file && dims, "bad argument");
    NOVA_GUARD(hdf5_read_host(file, dims, host, cap));
}

int nova_poisson_create(nova_ctx *ctx, nova_poisson **out)
{
    NOVA_REQUIRE(ctx && out, "null argument");
    NOVA_GUARD(poisson_create(ctx, out));
}

int nova_poisson_destroy(nova_poisson *ps) { NOVA_GUARD(poisson_destroy(ps)); }

int nova_poisson_solve(nova_poisson *ps, const double *rhs, double *p)
{
    NOVA_REQUIRE(ps && rhs && p && rhs != p, "bad argument");
    NOVA_GUARD(poisson_solve(ps, rhs, p));
}

int nova_project(nova_poisson *ps, double *F, double s, double *p)
{
    NOVA_REQUIRE(ps && F && s != 0.0 && F != p, "bad argument");
    NOVA_GUARD(project(ps, F, s, p));
}

int nova_solver_create(nova_poisson *ps, int model, int mode, double nu, double eta, nova_solver **out)
{
    NOVA_REQUIRE(ps && out, "null argument");
    NOVA_GUARD(solver_create(ps, model, mode, nu, eta, out));
}

int nova_solver_destroy(nova_solver *sv) { NOVA_GUARD(solver_destroy(sv)); }

int nova_solver_step(nova_solver *sv, double dt, int nsteps, double *U, double *B, double *p, double *pB)
{
    NOVA_REQUIRE(sv && U && p && nsteps >= 0, "bad argument");
    NOVA_GUARD(solver_step(sv, dt, nsteps, U, B, p, pB));
}

int nova_solver_use_graph(nova_solver *sv, int enable)
{
    NOVA_REQUIRE(sv, "null solver");
    NOVA_GUARD(solver_use_graph(sv, enable));
}

int nova_solver_set_scheme(nova_solver *sv, int scheme)
{
    NOVA_REQUIRE(sv, "null solver");
    NOVA_GUARD(solver_set_scheme(sv, scheme));
}

int nova_solver_set_force(nova_solver *sv, const double *F)
{
    NOVA_REQUIRE(sv, "null solver");
    NOVA_GUARD(solver_set_force(sv, F));
}

int nova_poisson_set_wall_solve(nova_poisson *ps, int mode)
{
    NOVA_REQUIRE(ps, "null Poisson solver");
    NOVA_GUARD(poisson_set_wall_solve(ps, mode));
}

int nova_ctx_set_wall_bc(nova_ctx *ctx, int dir, const int kind[3])
{
    NOVA_REQUIRE(ctx && kind, "null argument");
    NOVA_GUARD(set_wall_bc(ctx, dir, kind));
}

int nova_wall_fill(nova_ctx *ctx, double *F, int ncomp)
{
    NOVA_REQUIRE(ctx && F && (ncomp == 1 || ncomp == 3), "bad argument");
    NOVA_GUARD(wall_fill(ctx, F, ncomp));
}

int nova_poisson_set_scheme(nova_poisson *ps, int scheme)
{
    NOVA_REQUIRE(ps, "null Poisson solver");
    NOVA_GUARD(poisson_set_scheme(ps, scheme));
}

int nova_compact_deriv(nova_ctx *ctx, const double *f, int ncomp, int dir, int order, double *out)
{
    NOVA_REQUIRE(ctx && f && out && ncomp > 0, "bad argument");
    NOVA_GUARD(compact_deriv(ctx, f, ncomp, dir, order, out));
}

int nova_solver_autotune(nova_solver *sv, const double *U, const double *B, int reps, char *report,
                         size_t report_len)
{
    NOVA_REQUIRE(sv && U && reps >= 1 && (!report || report_len > 0), "bad argument");
    NOVA_GUARD(solver_autotune(sv, U, B, reps, report, report_len));
}

int nova_reduce_ke(nova_ctx *ctx, const double *U, double *out)
{
    NOVA_REQUIRE(ctx && U && out, "bad argument");
    NOVA_GUARD(reduce_ke(ctx, U, out));
}

int nova_reduce_maxabs_div(nova_ctx *ctx, const double *U, double *out)
{
    NOVA_REQUIRE(ctx && U && out, "bad argument");
    double *d = nullptr;
    NOVA_HIP(hipMallocAsync((void **)&d, sizeof(double) * ctx->g.np, ctx->stream));
    int rc = launch_div(ctx, U, 1.0, d);
    if (!rc) rc = reduce_maxabs(ctx, d, out);
    (void)hipFreeAsync(d, ctx->stream);
    return rc;
}

int nova_reduce_dissipation(nova_ctx *ctx, const double *U, double nu, double *out)
{
    NOVA_REQUIRE(ctx && U && out, "bad argument");
    NOVA_GUARD(reduce_dissipation(ctx, U, nu, out));
}

int nova_mhd_diag(nova_ctx *ctx, const double *U, const double *B, double mu, double eta, double out[8])
{
    NOVA_REQUIRE(ctx && U && B && out, "bad argument");
    double *J = nullptr;
    NOVA_HIP(hipMallocAsync((void **)&J, sizeof(double) * 6 * ctx->g.np, ctx->stream));
    double *W = J + 3 * ctx->g.np;
    int rc = launch_curl(ctx, B, 1, J);
    if (!rc) rc = launch_curl(ctx, U, 1, W);
    if (!rc) rc = reduce_mhd(ctx, U, B, J, W, mu, eta, out);
    (void)hipFreeAsync(J, ctx->stream);
    return rc;
}

int nova_adjoint_create(nova_solver *sv, int checkpoint_every, int max_steps, nova_adjoint **out)
{
    NOVA_REQUIRE(sv && out, "null argument");
    NOVA_GUARD(adjoint_create(sv, checkpoint_every, max_steps, out));
}

int nova_adjoint_destroy(nova_adjoint *ad) { NOVA_GUARD(adjoint_destroy(ad)); }

int nova_adjoint_forward(nova_adjoint *ad, double dt, int nsteps, const double *U0, double *U)
{
    NOVA_REQUIRE(ad && U0 && U, "bad argument");
    NOVA_GUARD(adjoint_forward(ad, dt, nsteps, U0, U));
}

int nova_adjoint_backward(nova_adjoint *ad, const double *lam_final, double *lam0)
{
    NOVA_REQUIRE(ad && lam_final && lam0, "bad argument");
    NOVA_GUARD(adjoint_backward(ad, lam_final, lam0));
}

int nova_adjoint_tlm(nova_adjoint *ad, const double *dU0, double *dU_final)
{
    NOVA_REQUIRE(ad && dU0 && dU_final, "bad argument");
    NOVA_GUARD(adjoint_tlm(ad, dU0, dU_final));
}

int nova_adjoint_autotune(nova_adjoint *ad, int reps, char *report, size_t report_len)
{
    NOVA_REQUIRE(ad && reps >= 1 && (!report || report_len > 0), "bad argument");
    NOVA_GUARD(adjoint_autotune(ad, reps, report, report_len));
}

int nova_adjoint_reserve(nova_adjoint *ad, uint64_t max_bytes, int *fields)
{
    NOVA_REQUIRE(ad, "null adjoint");
    NOVA_GUARD(adjoint_reserve(ad, max_bytes, fields));
}

int nova_wc_create(nova_ctx *ctx, double mu, double pc, nova_wc **out)
{
    NOVA_REQUIRE(ctx && out, "null argument");
    NOVA_GUARD(wc_create(ctx, mu, pc, out));
}

int nova_wc_destroy(nova_wc *w) { NOVA_GUARD(wc_destroy(w)); }

int nova_wc_set_magnetic(nova_wc *w, const double *B, double dbu)
{
    NOVA_REQUIRE(w, "null solver");
    NOVA_GUARD(wc_set_magnetic(w, B, dbu));
}

int nova_wc_set_temperature(nova_wc *w, double ka, double gx, double gy, double gz)
{
    NOVA_REQUIRE(w, "null solver");
    const double g[3] = {gx, gy, gz};
    NOVA_GUARD(wc_set_temperature(w, ka, g));
}

int nova_wc_step_t(nova_wc *w, double dt, int nsteps, double *U, double *rho, double *p, double *T)
{
    NOVA_REQUIRE(w && U && rho && p && T && nsteps >= 0, "bad argument");
    NOVA_GUARD(wc_step_t(w, dt, nsteps, U, rho, p, T));
}

int nova_wc_hydrostatic(nova_wc *w, const double *T, double rho0, double *p, double *rho)
{
    NOVA_REQUIRE(w && T && p && rho, "bad argument");
    NOVA_GUARD(wc_hydrostatic(w, T, rho0, p, rho));
}

int nova_fd_laplacian(nova_ctx *ctx, double mu, const double *phi, nova_fdmatrix **out)
{
    NOVA_REQUIRE(ctx && phi && out, "bad argument");
    NOVA_GUARD(fd_laplacian(ctx, mu, phi, out));
}

int nova_fd_destroy(nova_fdmatrix *m)
{
    if (!m) return NOVA_OK;
    NOVA_GUARD(fd_destroy(m));
}

int nova_fd_add_source(nova_fdmatrix *m, const double *f)
{
    NOVA_REQUIRE(m && f, "bad argument");
    NOVA_GUARD(fd_add_source(m, f));
}

int nova_fd_relax(nova_fdmatrix *m, const double *phi, double urf)
{
    NOVA_REQUIRE(m && phi && urf != 0.0, "bad argument");
    NOVA_GUARD(fd_relax(m, phi, urf));
}

int nova_fd_set_reference(nova_fdmatrix *m, int i, int j, int k, double val)
{
    NOVA_REQUIRE(m && fd_inside(m, i, j, k), "bad argument");
    NOVA_GUARD(fd_set_reference(m, i, j, k, val));
}

int nova_fd_combine(nova_fdmatrix *m, const nova_fdmatrix *other, double s)
{
    NOVA_REQUIRE(m && (!other || fd_same_grid(m, other)), "bad argument");
    NOVA_GUARD(fd_combine(m, other, s));
}

int nova_fd_residual(nova_fdmatrix *m, const double *phi, double *res)
{
    NOVA_REQUIRE(m && phi && res, "bad argument");
    NOVA_GUARD(fd_residual(m, phi, res));
}

int nova_fd_solve(nova_fdmatrix *m, double *phi, double tol, int max_its, int *its, double *res)
{
    NOVA_REQUIRE(m && phi && max_its >= 0, "bad argument");
    NOVA_GUARD(fd_solve(m, phi, tol, max_its, its, res));
}

int nova_fd_get(const nova_fdmatrix *m, int what, double *padded_out)
{
    NOVA_REQUIRE(m && padded_out && what >= 0 && what <= 8, "bad argument");
    NOVA_GUARD(fd_get(m, padded_out, what));
}

int nova_time_create(double dt, double end_time, int write_every, nova_time **out)
{
    NOVA_REQUIRE(out && write_every > 0, "bad argument");
    NOVA_GUARD(time_create(dt, end_time, write_every, 0.0, false, out));
}

int nova_time_create_interval(double dt, double end_time, double write_interval, nova_time **out)
{
    NOVA_REQUIRE(out && write_interval > 0.0, "bad argument");
    NOVA_GUARD(time_create(dt, end_time, 0, write_interval, true, out));
}

int nova_time_create_from_file(const char *file, nova_time **out)
{
    NOVA_REQUIRE(file && out, "bad argument");
    NOVA_GUARD(time_create_from_file(file, out));
}

int nova_time_restart(nova_time *t, const char *restart_file)
{
    NOVA_REQUIRE(t && restart_file, "bad argument");
    NOVA_GUARD(time_restart(t, restart_file));
}

int nova_time_destroy(nova_time *t)
{
    time_destroy(t);
    return NOVA_OK;
}

int nova_time_advance(nova_time *t)
{
    NOVA_REQUIRE(t, "bad argument");
    NOVA_GUARD(time_advance(t));
}

int nova_time_set_dt(nova_time *t, double dt)
{
    NOVA_REQUIRE(t, "bad argument");
    time_set_dt(t, dt);
    return NOVA_OK;
}

int nova_time_info(const nova_time *t, double dout[4], int64_t iout[5])
{
    NOVA_REQUIRE(t && dout && iout, "bad argument");
    time_info(t, dout, iout);
    return NOVA_OK;
}

int nova_time_steps(const nova_time *t, double *out, int64_t cap)
{
    NOVA_REQUIRE(t && (out || cap == 0) && cap >= 0, "bad argument");
    time_steps(t, out, cap);
    return NOVA_OK;
}

int nova_io_create(nova_ctx *ctx, nova_io **out)
{
    NOVA_REQUIRE(ctx && out, "bad argument");
    NOVA_GUARD(io_create(ctx, out));
}

int nova_io_destroy(nova_io *io)
{
    if (!io) return NOVA_OK;
    NOVA_GUARD(io_destroy(io));
}

int nova_io_wait(nova_io *io)
{
    NOVA_REQUIRE(io, "bad argument");
    NOVA_GUARD(io_wait(io));
}

int nova_io_write_field(nova_io *io, const nova_time *t, const char *path, const char *name, const double *F,
                        int ncomp)
{
    NOVA_REQUIRE(io && t && path && name && F && (ncomp == 1 || ncomp == 3), "bad argument");
    NOVA_GUARD(io_write_field(io, t, path, name, F, ncomp));
}

int nova_io_write_mesh(nova_io *io, const nova_time *t, const char *path)
{
    NOVA_REQUIRE(io && t && path, "bad argument");
    NOVA_GUARD(io_write_mesh(io, t, path));
}

int nova_io_read_field(nova_io *io, const char *path, const char *name, int file_num, double *F, int ncomp)
{
    NOVA_REQUIRE(io && path && name && F && file_num >= 0 && (ncomp == 1 || ncomp == 3), "bad argument");
    NOVA_GUARD(io_read_field(io, path, name, file_num, F, ncomp));
}

int nova_io_checkpoint_write(nova_io *io, const char *file, int nfields, const double *const *fields,
                             const int *ncomp, double time, int64_t step)
{
    NOVA_REQUIRE(io && file && nfields > 0 && fields && ncomp, "bad argument");
    for (int f = 0; f < nfields; f++) NOVA_REQUIRE(fields[f] && ncomp[f] > 0, "bad field");
    NOVA_GUARD(io_checkpoint_write(io, file, nfields, fields, ncomp, time, step));
}

int nova_io_checkpoint_read(nova_io *io, const char *file, int nfields, double *const *fields, const int *ncomp,
                            double *time, int64_t *step)
{
    NOVA_REQUIRE(io && file && nfields > 0 && fields && ncomp, "bad argument");
    for (int f = 0; f < nfields; f++) NOVA_REQUIRE(fields[f] && ncomp[f] > 0, "bad field");
    NOVA_GUARD(io_checkpoint_read(io, file, nfields, fields, ncomp, time, step));
}

int nova_wc_step(nova_wc *w, double dt, int nsteps, double *U, double *rho, double *p)
{
    NOVA_REQUIRE(w && U && rho && p && nsteps >= 0, "bad argument");
    NOVA_GUARD(wc_step(w, dt, nsteps, U, rho, p));
}

int nova_ensemble_create(const nova_instance_desc *inst, int n, int device, nova_ensemble **out)
{
    NOVA_REQUIRE(inst && out && n > 0, "bad argument");
    NOVA_HIP(hipSetDevice(device));
    NOVA_GUARD(ensemble_create(inst, n, device, out));
}

int nova_ensemble_destroy(nova_ensemble *e)
{
    NOVA_GUARD(ensemble_destroy(e));
}

int nova_ensemble_set_stream(nova_ensemble *e, void *stream)
{
    NOVA_REQUIRE(e, "null ensemble");
    NOVA_GUARD(ensemble_set_stream(e, (hipStream_t)stream));
}

int nova_ensemble_copy(nova_ensemble *e, int i, int store, double *U, double *B, double *p, double *pB)
{
    NOVA_REQUIRE(e, "null ensemble");
    NOVA_GUARD(ensemble_copy(e, i, store, U, B, p, pB));
}

int nova_ensemble_step(nova_ensemble *e, double dt, int nsteps)
{
    NOVA_REQUIRE(e && nsteps >= 0, "bad argument");
    NOVA_GUARD(ensemble_step(e, dt, nsteps));
}

int nova_ensemble_use_graph(nova_ensemble *e, int enable)
{
    NOVA_REQUIRE(e, "null ensemble");
    NOVA_GUARD(ensemble_use_graph(e, enable));
}

int nova_ensemble_info(nova_ensemble *e, int64_t out[6])
{
    NOVA_REQUIRE(e && out, "bad argument");
    NOVA_GUARD(ensemble_info(e, out));
}

int nova_weighted_dot(nova_ctx *ctx, const double *a, const double *b, int ncomp, double *out)
{
    NOVA_REQUIRE(ctx && a && b && out && ncomp > 0, "bad argument");
    NOVA_GUARD(reduce_wdot(ctx, a, b, ncomp, out));
}

int nova_enstrophy(nova_ctx *ctx, const double *U, double *J, double *dJdU)
{
    NOVA_REQUIRE(ctx && U && J && dJdU != U, "bad argument");
    const double dV = ctx->d[0] * ctx->d[1] * ctx->d[2];
    double *W = nullptr;
    NOVA_HIP(hipMallocAsync((void **)&W, sizeof(double) * 3 * ctx->g.np, ctx->stream));
    int rc = launch_curl(ctx, U, 0, W);
    double s = 0.0;
    if (!rc) rc = reduce_wdot(ctx, W, W, 3, &s);
    if (!rc) *J = dV * s;
    if (!rc && dJdU) {
        rc = launch_curl(ctx, W, 0, dJdU, 1);  // W is a curl: pseudovector parity at walls
        if (!rc) rc = launch_axpby(ctx, 3 * (size_t)ctx->g.np, 2.0 * dV, dJdU, 0.0, nullptr, dJdU);
    }
    (void)hipFreeAsync(W, ctx->stream);
    return rc;
}

int nova_tdma_batched(nova_ctx *ctx, int n, int nbatch, const double *a, const double *b, const double *c,
                      const double *d, double *x, int cyclic)
{
    NOVA_REQUIRE(ctx && a && b && c && d && x && n >= 2 && nbatch >= 0, "bad argument");
    if (nbatch == 0) return NOVA_OK;
    NOVA_GUARD(launch_tdma(ctx, n, nbatch, a, b, c, d, x, cyclic));
}

int nova_tdma_batched_noalloc(nova_ctx *ctx, int n, int nbatch, const double *a, const double *b, const double *c,
                              const double *d, double *x, int cyclic)
{
    NOVA_REQUIRE(ctx && a && b && c && d && x && n >= 2 && nbatch >= 0, "bad argument");
    if (nbatch == 0) return NOVA_OK;
    NOVA_GUARD(launch_tdma_noalloc(ctx, n, nbatch, a, b, c, d, x, cyclic));
}

size_t nova_tdma_noalloc_max(void) { return tdma_scratch_max(); }

}  // extern "C"
