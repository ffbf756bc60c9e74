// nova_internal.hpp -- shared definitions of the gfx950 NOVA hot path.
//
// Device layout (compact SoA): component c of a field at base + c*np,
// node (i,j,k) at (i*s1 + j)*s2 + k.  Boundary conditions are applied at
// load time through index maps (bmap) instead of stored ghost layers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nova_capi.h"

namespace nova {

constexpr int kMaxH = 4;  // m/2 for m = 8

// Per-launch description of the grid, passed by value to kernels.
struct Grid {
    int s[3];          // stored extents
    int n[3];          // cells
    int wall[3];       // 1 = reflecting walls at nodes 0 and n
    unsigned odd[3];   // bit c: component c of a vector field is odd across walls normal to d
    int h;             // m/2
    int64_t np;        // s0*s1*s2
    double a[3][kMaxH];  // first derivative weights c_l / (2 (l+1) h_d)
    double b[3][kMaxH];  // second derivative weights c_l / ((l+1) h_d)^2
    double al[3][kMaxH]; // literal-gradient weights c_l / ((l+1) h_d)
    double wa[kMaxH];    // c_l / (2 (l+1))
    double wb[kMaxH];    // c_l / (l+1)^2
    double wb_sum2;      // 2 sum_l wb[l]
    double ih[3], ih2[3];  // 1/h_d, 1/h_d^2
    const double *zinv;  // per stored k: 1 / dzdxi(k)
    const double *zlap;  // per stored k: 1 / dzdxi(k)^2
    const double *zmet;  // per stored k: d2zdxi2(k) / dzdxi(k)^3
};

// Map a possibly out-of-range node index along one direction to a stored
// index.  Periodic: wrap modulo n.  Wall: reflect about node 0 and node n
// (flip = 1 when the load came through a reflection).  Indices beyond one
// reflection (threads of a partial tile) are clamped; they are never
// stored.
__host__ __device__ inline int bmap(int q, int n, int wall, int &flip)
{
    flip = 0;
    if (!wall) {
        q %= n;
        if (q < 0) q += n;
        return q;
    }
    if (q < 0) { q = -q; flip = 1; }
    if (q > n) { q = 2 * n - q; flip = 1; }
    if (q < 0) q = 0;
    if (q > n) q = n;
    return q;
}

// Sign of component c of an ncomp-component field reached through a
// reflection across a wall normal to direction d: odd components of a
// vector (g.odd[d], by default the wall-normal one) change sign, scalars
// and the other components are even.
__host__ __device__ inline double bsign(const Grid &g, int ncomp, int c, int d, int flip)
{
    return (flip && ncomp == 3 && ((g.odd[d] >> c) & 1u)) ? -1.0 : 1.0;
}

struct DevBuf {
    double *p = nullptr;
    size_t n = 0;
};

}  // namespace nova

namespace nova {
// One timed launch: events recorded on the context stream around it.
struct TimedLaunch {
    const char *name;
    hipEvent_t start, stop;
};
}  // namespace nova

struct nova_ctx {
    int device = 0;
    int timing = 0;                        // record events around launches
    // launch shape of the stage kernels per model (NS, MHD, tangent-linear,
    // adjoint): variant 0..4 and planes per streamed chunk (0: automatic)
    // (slot 4: the first SSP stage of the sixth-order NS step, which reads
    // no step-start operand and has its own best shape)
    int stage_variant[5] = {3, 3, 3, 3, 3};
    int stage_ichunk[5] = {0, 0, 0, 0, 0};
    int tune_epoch = 0;                    // bumped on every change of the launch shapes
    int proj_variant = 2;                  // 0: divergence fused into the r2c pass, 1: row kernels, 2: LDS tiles
    std::vector<nova::TimedLaunch> tlog;   // launches since the last collect
    std::vector<hipEvent_t> evpool;        // reusable events
    hipStream_t stream = nullptr;
    nova_grid_desc desc{};
    nova::Grid g{};
    int m = 6;
    int pad[3] = {0, 0, 0};          // reference padded extents n_d = N_d + 1 + m
    double d[3] = {0, 0, 0};         // spacings
    double hb = 1.0;                 // hyperbolic stretching constant (Mesh.cpp:57-60)
    double coef[4] = {0, 0, 0, 0};
    std::vector<double> dzdxi, d2zdxi2;  // stored-k metric
    double *d_metric = nullptr;      // zinv | zlap | zmet on the device
    double *d_red = nullptr;         // reduction partials
    size_t red_cap = 0;
    double *h_red = nullptr;         // pinned host result slot
    // the reference's wall fills (nova_ctx_set_wall_bc): wall nodes are
    // overwritten after every assignment, per direction and component
    int wall_fill = 0;
    int wall_kind[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
};

namespace nova {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
#define NOVA_HIP(call)                                   \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return nova::hip_fail(e_, #call); \
    } while (0)

int check_launch(const char *what);

// Brackets the launches issued in its scope with events on ctx->stream
// when the context's timing mode is on.
class TimedScope {
public:
    TimedScope(nova_ctx *ctx, const char *name);
    ~TimedScope();

private:
    nova_ctx *ctx_;
    const char *name_;
    hipEvent_t start_ = nullptr;
};
int ensure_red(nova_ctx *ctx, size_t n);

// Average duration in microseconds of reps calls of f() on the context
// stream, measured with events after one untimed call.
template <class F>
int time_avg_us(nova_ctx *ctx, int reps, F &&f, double *us)
{
    if (int rc = f()) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int rc = NOVA_OK;
    if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps && rc == NOVA_OK && e == hipSuccess; r++) rc = f();
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "tuning events");
    *us = 1e3 * ms / reps;
    return NOVA_OK;
}

// Measured search over the stage-kernel launch shapes of one model:
// run(v, c) launches the stage with variant v and chunk c already set in
// the context; the fastest pair is kept and every trial is appended to rep.
template <class F>
int tune_stage_model(nova_ctx *ctx, int model, const int *variants, int nvar, int reps, F &&run, const char *name,
                     std::string &rep)
{
    // planes per chunk (0: automatic); 128 gives a 256-plane grid two
    // chunks, so that one-block-per-CU shapes fill the GPU in one round with
    // half the queue fills of 64
    static const int kChunks[] = {0, 8, 16, 32, 64, 128};
    double best = 1e300;
    int bv = ctx->stage_variant[model], bc = ctx->stage_ichunk[model];
    for (int iv = 0; iv < nvar; iv++)
        for (int c : kChunks) {
            ctx->stage_variant[model] = variants[iv];
            ctx->stage_ichunk[model] = c;
            double us = 0.0;
            if (int rc = time_avg_us(ctx, reps, run, &us)) {
                ctx->stage_variant[model] = bv;
                ctx->stage_ichunk[model] = bc;
                return rc;
            }
            rep += std::string(name) + " variant " + std::to_string(variants[iv]) + " chunk " + std::to_string(c) +
                   " " + std::to_string(us) + " us\n";
            if (us < best) {
                best = us;
                bv = variants[iv];
                bc = c;
            }
        }
    ctx->stage_variant[model] = bv;
    ctx->stage_ichunk[model] = bc;
    ctx->tune_epoch++;
    rep += std::string(name) + " chosen variant " + std::to_string(bv) + " chunk " + std::to_string(bc) + "\n";
    return NOVA_OK;
}

int write_report(const std::string &rep, char *buf, size_t len);

// Kernel launchers (defined in the .hip files).
int launch_div_adv(nova_ctx *ctx, const double *U, const double *phi, int nc, double *out);
int launch_div(nova_ctx *ctx, const double *U, double scale, double *out);
int launch_grad(nova_ctx *ctx, const double *p, double *out);
int launch_laplacian(nova_ctx *ctx, double nu, const double *f, int nc, double *out);
int launch_laplacian_var(nova_ctx *ctx, const double *nu, const double *f, int nc, double *out);
int launch_curl(nova_ctx *ctx, const double *U, int literal, double *out, int pseudo = 0);
int launch_axpby(nova_ctx *ctx, size_t n, double a, const double *x, double b, const double *y,
                 double *out);
int launch_pack(nova_ctx *ctx, const double *compact, int nc, double *padded);
int launch_unpack(nova_ctx *ctx, const double *padded, int nc, double *compact);
int launch_bc_fill(nova_ctx *ctx, double *padded, int nc, int dir, int kind, int comp, double val);
int field_op(nova_ctx *ctx, int op, const double *a, int nca, const double *b, int ncb, const double *cval,
             double *out);
int grad_vec_literal(nova_ctx *ctx, const double *U, double *out);
int reduce_cfl(nova_ctx *ctx, const double *U, double dt, double out[6]);
int wall_fill(nova_ctx *ctx, double *F, int nc);
int set_wall_bc(nova_ctx *ctx, int dir, const int kind[3]);
int hdf5_write_host(const char *file, const int64_t dims[3], const int64_t stored[3], const double *host);
int hdf5_read_host(const char *file, int64_t dims[3], double *host, size_t cap);
int hdf5_write_field(nova_ctx *ctx, const char *file, const double *F, int comp);
int reduce_ke(nova_ctx *ctx, const double *U, double *out);
int reduce_maxabs(nova_ctx *ctx, const double *f, double *out);
int reduce_dissipation(nova_ctx *ctx, const double *U, double nu, double *out);
int reduce_wdot(nova_ctx *ctx, const double *a, const double *b, int nc, double *out);
int reduce_mhd(nova_ctx *ctx, const double *U, const double *B, const double *J,
               const double *W, double mu, double eta, double out[8]);
int launch_tdma(nova_ctx *ctx, int n, int nb, const double *a, const double *b, const double *c,
                const double *d, double *x, int cyclic);
// the same solve with its work array in the library's per-device workspace
int tdma_scratch_init(int device);
size_t tdma_scratch_max();
int launch_tdma_noalloc(nova_ctx *ctx, int n, int nb, const double *a, const double *b, const double *c,
                        const double *d, double *x, int cyclic);
// compact sixth-order scheme (nova_compact.hip)
int compact_supported(nova_ctx *ctx);
double compact_kmod(double kh, int order);
int compact_deriv(nova_ctx *ctx, const double *f, int nc, int dir, int order, double *out);
int compact_div(nova_ctx *ctx, const double *F, double scale, double *d);
int compact_grad_correct(nova_ctx *ctx, const double *p, double s, double *F);

}  // namespace nova
