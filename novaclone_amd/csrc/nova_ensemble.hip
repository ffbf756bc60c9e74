// nova_ensemble.hip -- batched ensembles of uneven instances (no reference
// counterpart: NOVA runs one case per process).
//
// Instances of different sizes and models (Navier-Stokes TG-type, MHD
// OT-type) advance together with the SSP-RK3 composition of nova_step.hip:
//   * each RK stage of every instance is two persistent launches, one per
//     model: a fixed grid of blocks walks the tile work list of that model's
//     instances, and each tile runs the same streamed stage body (stage_tile)
//     as the single-grid kernels, so an instance's result equals its
//     standalone run bit for bit;
//   * instances with the same grid form a size class that shares a context
//     (metric, wavenumbers) and one batched Poisson solver; a class's fields
//     sit in contiguous slots [NS U | MHD U | MHD B] so that each projection
//     of the class is one batched transform pipeline;
//   * the whole step is captured once as a HIP graph and replayed.
#include "nova_internal.hpp"
#include "nova_stencil.hpp"
#include "nova_stream.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

struct nova_poisson;

namespace nova {
int poisson_create(nova_ctx *ctx, nova_poisson **out, int batch_cap);
int poisson_destroy(nova_poisson *ps);
int project_batch(nova_poisson *ps, double *F, int64_t f_bs, int B, int split, double s_lo, double s_hi, double *p,
                  int64_t p_bs);
}  // namespace nova

extern "C" int nova_ctx_create(const nova_grid_desc *desc, int device, nova_ctx **out);
extern "C" int nova_ctx_destroy(nova_ctx *ctx);

namespace nova {
namespace {

constexpr int kEnsTK = 16;     // 16 x 16 (k x j) tiles: every class size is a multiple of 16
// planes per streamed chunk: 16 against 8 (2H = 6 warm-up planes per chunk)
// measured 2.607 against 2.622 ms per 48-instance step, 32 (fewer tiles
// than the persistent blocks can balance) 2.79
constexpr int kEnsChunk = 16;

// Device descriptor of one instance.
struct EnsInstance {
    Grid g;
    double *U[2], *W1[2], *W2[2];  // velocity / magnetic field: state, stage-1 and stage-2 outputs
    double nu, eta;
    int model;                     // MODEL_NS or MODEL_MHD
    int nkt, njt, nic, ichunk;
    int tile0;                     // first work-list tile
};

// One model's tiles of a stage: a persistent grid walks the work list of
// that model's instances.  The two models run as two launches, each with its
// own registers and LDS (a single launch over both held the NS tiles to the
// MHD stage's one wave per SIMD and its two-field LDS plane).
template <int H, int MODEL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODEL == MODEL_MHD || H >= 4 ? 1 : 2)))
k_stage_group(const EnsInstance *__restrict__ inst, int ninst, int ntiles, int stage, StageArgs sa)
{
    __shared__ double sh[stage_lds<H, MODEL, kEnsTK>()];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        // owner of tile t: the last instance whose first tile is <= t
        int lo = 0, hi = ninst - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (inst[mid].tile0 <= t) lo = mid;
            else hi = mid - 1;
        }
        const EnsInstance &I = inst[lo];
        const int tl = t - I.tile0;
        if constexpr (MODEL == MODEL_NS) {
            StreamArgs<1> a{};
            a.X[0] = stage == 0 ? I.U[0] : (stage == 1 ? I.W1[0] : I.W2[0]);
            a.X0[0] = stage == 0 ? nullptr : I.U[0];
            a.out[0] = stage == 0 ? I.W1[0] : (stage == 1 ? I.W2[0] : I.U[0]);
            a.sa = sa;
            a.nu = I.nu;
            a.nk = ns_k(I.g, I.nu, sa.dt);  // the products the host forms for a standalone stage
            a.ichunk = I.ichunk;
            a.nkt = I.nkt;
            a.njt = I.njt;
            a.nic = I.nic;
            stage_tile<H, MODEL_NS, false, kEnsTK>(I.g, a, tl, sh);
        } else {
            StreamArgs<2> a{};
            for (int f = 0; f < 2; f++) {
                a.X[f] = stage == 0 ? I.U[f] : (stage == 1 ? I.W1[f] : I.W2[f]);
                a.X0[f] = stage == 0 ? nullptr : I.U[f];
                a.out[f] = stage == 0 ? I.W1[f] : (stage == 1 ? I.W2[f] : I.U[f]);
            }
            a.sa = sa;
            a.nu = I.nu;
            a.eta = I.eta;
            a.ichunk = I.ichunk;
            a.nkt = I.nkt;
            a.njt = I.njt;
            a.nic = I.nic;
            stage_tile<H, MODEL_MHD, false, kEnsTK>(I.g, a, tl, sh);
        }
        __syncthreads();  // the next tile restages the LDS plane
    }
}

StageArgs ssp_stage(int s, double dt)
{
    StageArgs a{};
    a.form = FORM_SSP;
    a.dt = dt;
    if (s == 0) { a.alpha = 0.0; a.beta = 1.0; }
    else if (s == 1) { a.alpha = 0.75; a.beta = 0.25; }
    else { a.alpha = 1.0 / 3.0; a.beta = 2.0 / 3.0; }
    return a;
}

bool same_grid(const nova_grid_desc &a, const nova_grid_desc &b)
{
    for (int d = 0; d < 3; d++)
        if (a.n[d] != b.n[d] || a.L[d] != b.L[d] || a.origin[d] != b.origin[d] || a.bc[d] != b.bc[d]) return false;
    return a.m == b.m && a.hyp == b.hyp && a.s1 == b.s1 && a.s2 == b.s2;
}

}  // namespace
}  // namespace nova

struct nova_ensemble {
    struct Class {
        nova_grid_desc desc{};
        nova_ctx *ctx = nullptr;
        nova_poisson *ps = nullptr;
        int nT = 0, nM = 0;              // NS and MHD members
        double *U = nullptr, *W1 = nullptr, *W2 = nullptr, *P = nullptr;
        int64_t np = 0;
        int slots() const { return nT + 2 * nM; }
    };
    struct Member {
        int cls = 0, model = 0;
        int slotU = 0, slotB = -1;       // field slots in the class regions
    };
    int device = 0;
    hipStream_t stream = nullptr;        // caller's stream (0 = default)
    std::vector<Class> classes;
    std::vector<Member> members;
    // work lists: the NS instances (hinst[0 .. nns)) then the MHD ones, each
    // list's tiles numbered from 0; one persistent launch per model
    std::vector<nova::EnsInstance> hinst;
    nova::EnsInstance *dinst = nullptr;
    int nns = 0;
    int ntiles_m[2] = {0, 0}, nblocks_m[2] = {0, 0};  // [0] NS, [1] MHD
    int ntiles = 0, nblocks = 0, h = 3;
    int use_graph = 0;
    hipStream_t gstream = nullptr;
    hipGraphExec_t gexec = nullptr;
    double gdt = 0.0;
    // the size classes' projections run side by side: class i on side[i],
    // forked from the step's stream after each stage (fork) and joined back
    // (done[i]) before the next stage; default-priority streams
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> done;
    hipEvent_t fork = nullptr;
};

namespace nova {
namespace {

void release(nova_ensemble *e)
{
    for (auto &c : e->classes) {
        if (c.ps) (void)poisson_destroy(c.ps);
        if (c.U) (void)hipFree(c.U);
        if (c.W1) (void)hipFree(c.W1);
        if (c.W2) (void)hipFree(c.W2);
        if (c.P) (void)hipFree(c.P);
        if (c.ctx) (void)nova_ctx_destroy(c.ctx);
    }
    if (e->dinst) (void)hipFree(e->dinst);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->gstream) (void)hipStreamDestroy(e->gstream);
    for (hipStream_t s : e->side) (void)hipStreamDestroy(s);
    for (hipEvent_t v : e->done) (void)hipEventDestroy(v);
    if (e->fork) (void)hipEventDestroy(e->fork);
}

template <int H>
int launch_group(nova_ensemble *e, hipStream_t st, int stage, double dt)
{
    const int nm = (int)e->hinst.size() - e->nns;
    if (e->nns > 0)
        hipLaunchKernelGGL((k_stage_group<H, MODEL_NS>), dim3(e->nblocks_m[0]), dim3(kEnsTK, 256 / kEnsTK, 1), 0, st,
                           e->dinst, e->nns, e->ntiles_m[0], stage, ssp_stage(stage, dt));
    if (nm > 0)
        hipLaunchKernelGGL((k_stage_group<H, MODEL_MHD>), dim3(e->nblocks_m[1]), dim3(kEnsTK, 256 / kEnsTK, 1), 0, st,
                           e->dinst + e->nns, nm, e->ntiles_m[1], stage, ssp_stage(stage, dt));
    return check_launch("ensemble stage");
}

// blocks of one model's launch that are resident at once
template <int H, int MODEL>
hipError_t group_residency(int *per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_stage_group<H, MODEL>, 256, 0);
}

int step_once(nova_ensemble *e, hipStream_t st, double dt)
{
    for (auto &c : e->classes) c.ctx->stream = st;
    for (int s = 0; s < 3; s++) {
        int rc;
        {
            TimedScope ts(e->classes[0].ctx, "ensemble_stage");
            switch (e->h) {
            case 1: rc = launch_group<1>(e, st, s, dt); break;
            case 2: rc = launch_group<2>(e, st, s, dt); break;
            case 3: rc = launch_group<3>(e, st, s, dt); break;
            default: rc = launch_group<4>(e, st, s, dt); break;
            }
        }
        if (rc) return rc;
        const bool fan = !e->side.empty();
        if (fan) NOVA_HIP(hipEventRecord(e->fork, st));
        for (size_t ci = 0; ci < e->classes.size(); ci++) {
            auto &c = e->classes[ci];
            // after stages 1 and 2 only the MHD members project (velocity with
            // dt, field with 1); after stage 3 every member does
            double *region = s == 0 ? c.W1 : (s == 1 ? c.W2 : c.U);
            const int first = s < 2 ? c.nT : 0;
            const int B = c.slots() - first;
            if (B == 0) continue;
            const int split = c.nT + c.nM - first;
            if (fan) {
                NOVA_HIP(hipStreamWaitEvent(e->side[ci], e->fork, 0));
                c.ctx->stream = e->side[ci];
            }
            if ((rc = project_batch(c.ps, region + first * 3 * c.np, 3 * c.np, B, split, dt, 1.0, c.P + first * c.np,
                                    c.np)))
                return rc;
            if (fan) {
                NOVA_HIP(hipEventRecord(e->done[ci], e->side[ci]));
                NOVA_HIP(hipStreamWaitEvent(st, e->done[ci], 0));
                c.ctx->stream = st;
            }
        }
    }
    return NOVA_OK;
}

int record(nova_ensemble *e, double dt)
{
    if (!e->gstream) NOVA_HIP(hipStreamCreate(&e->gstream));
    if (e->gexec) {
        NOVA_HIP(hipGraphExecDestroy(e->gexec));
        e->gexec = nullptr;
    }
    NOVA_HIP(hipStreamBeginCapture(e->gstream, hipStreamCaptureModeThreadLocal));
    int rc = step_once(e, e->gstream, dt);
    hipGraph_t graph = nullptr;
    hipError_t err = hipStreamEndCapture(e->gstream, &graph);
    for (auto &c : e->classes) c.ctx->stream = e->stream;
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (err != hipSuccess) return hip_fail(err, "hipStreamEndCapture");
    err = hipGraphInstantiate(&e->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (err != hipSuccess) return hip_fail(err, "hipGraphInstantiate");
    e->gdt = dt;
    return NOVA_OK;
}

}  // namespace

int ensemble_create(const nova_instance_desc *inst, int n, int device, nova_ensemble **out)
{
    if (n < 1) return fail(NOVA_ERR_ARG, "an ensemble needs at least one instance");
    const int m = inst[0].grid.m;
    for (int i = 0; i < n; i++) {
        const nova_instance_desc &d = inst[i];
        if (d.model != NOVA_MODEL_NS && d.model != NOVA_MODEL_MHD) return fail(NOVA_ERR_ARG, "unknown model");
        if (d.grid.m != m) return fail(NOVA_ERR_UNSUPPORTED, "ensemble instances must share the stencil order m");
        for (int a = 0; a < 3; a++)
            if (d.grid.bc[a] != NOVA_BC_PERIODIC) return fail(NOVA_ERR_UNSUPPORTED, "ensemble instances are periodic");
        if (d.grid.hyp) return fail(NOVA_ERR_UNSUPPORTED, "ensemble instances use uniform grids");
    }
    nova_ensemble *e = new nova_ensemble();
    e->device = device;
    e->h = m / 2;
    // size classes
    e->members.resize(n);
    for (int i = 0; i < n; i++) {
        int c = -1;
        for (size_t q = 0; q < e->classes.size(); q++)
            if (same_grid(e->classes[q].desc, inst[i].grid)) c = (int)q;
        if (c < 0) {
            nova_ensemble::Class cl;
            cl.desc = inst[i].grid;
            e->classes.push_back(cl);
            c = (int)e->classes.size() - 1;
        }
        e->members[i].cls = c;
        e->members[i].model = inst[i].model;
        if (inst[i].model == NOVA_MODEL_NS) e->classes[c].nT++;
        else e->classes[c].nM++;
    }
    // slots: [NS U | MHD U | MHD B] in member order
    std::vector<int> tcount(e->classes.size(), 0), mcount(e->classes.size(), 0);
    for (int i = 0; i < n; i++) {
        auto &mb = e->members[i];
        auto &cl = e->classes[mb.cls];
        if (mb.model == NOVA_MODEL_NS) {
            mb.slotU = tcount[mb.cls]++;
        } else {
            mb.slotU = cl.nT + mcount[mb.cls];
            mb.slotB = cl.nT + cl.nM + mcount[mb.cls];
            mcount[mb.cls]++;
        }
    }
    int rc = NOVA_OK;
    for (auto &cl : e->classes) {
        if ((rc = nova_ctx_create(&cl.desc, device, &cl.ctx))) break;
        if ((rc = poisson_create(cl.ctx, &cl.ps, cl.slots()))) break;
        cl.np = cl.ctx->g.np;
        const size_t vb = sizeof(double) * 3 * (size_t)cl.np * cl.slots();
        hipError_t err = hipMalloc((void **)&cl.U, vb);
        if (err == hipSuccess) err = hipMalloc((void **)&cl.W1, vb);
        if (err == hipSuccess) err = hipMalloc((void **)&cl.W2, vb);
        if (err == hipSuccess) err = hipMalloc((void **)&cl.P, sizeof(double) * (size_t)cl.np * cl.slots());
        if (err == hipSuccess) err = hipMemsetAsync(cl.U, 0, vb, cl.ctx->stream);
        if (err == hipSuccess) err = hipMemsetAsync(cl.P, 0, sizeof(double) * (size_t)cl.np * cl.slots(), cl.ctx->stream);
        if (err != hipSuccess) {
            rc = hip_fail(err, "ensemble fields");
            break;
        }
    }
    if (rc) {
        release(e);
        delete e;
        return rc;
    }
    // work list: per instance its tiles, in member order
    // NS instances first, then MHD: each model's work list is contiguous
    for (int pass = 0; pass < 2; pass++) {
        int tiles = 0;
        for (int i = 0; i < n; i++) {
            const auto &mb = e->members[i];
            const bool mhd = mb.model == NOVA_MODEL_MHD;
            if (mhd != (pass == 1)) continue;
            const auto &cl = e->classes[mb.cls];
            const Grid &g = cl.ctx->g;
            EnsInstance I{};
            I.g = g;
            const int64_t f3 = 3 * cl.np;
            I.U[0] = cl.U + mb.slotU * f3;
            I.W1[0] = cl.W1 + mb.slotU * f3;
            I.W2[0] = cl.W2 + mb.slotU * f3;
            if (mb.slotB >= 0) {
                I.U[1] = cl.U + mb.slotB * f3;
                I.W1[1] = cl.W1 + mb.slotB * f3;
                I.W2[1] = cl.W2 + mb.slotB * f3;
            }
            I.nu = inst[i].nu;
            I.eta = inst[i].eta;
            I.model = mhd ? MODEL_MHD : MODEL_NS;
            I.nkt = (g.s[2] + kEnsTK - 1) / kEnsTK;
            I.njt = (g.s[1] + 256 / kEnsTK - 1) / (256 / kEnsTK);
            I.ichunk = kEnsChunk < g.s[0] ? kEnsChunk : g.s[0];
            I.nic = (g.s[0] + I.ichunk - 1) / I.ichunk;
            I.tile0 = tiles;
            tiles += I.nkt * I.njt * I.nic;
            e->hinst.push_back(I);
        }
        e->ntiles_m[pass] = tiles;
        if (pass == 0) e->nns = (int)e->hinst.size();
    }
    e->ntiles = e->ntiles_m[0] + e->ntiles_m[1];
    int cus = 0, per_cu[2] = {0, 0};
    hipError_t err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    for (int m = 0; m < 2 && err == hipSuccess; m++) {
        switch (e->h) {
        case 1: err = m ? group_residency<1, MODEL_MHD>(&per_cu[m]) : group_residency<1, MODEL_NS>(&per_cu[m]); break;
        case 2: err = m ? group_residency<2, MODEL_MHD>(&per_cu[m]) : group_residency<2, MODEL_NS>(&per_cu[m]); break;
        case 3: err = m ? group_residency<3, MODEL_MHD>(&per_cu[m]) : group_residency<3, MODEL_NS>(&per_cu[m]); break;
        default: err = m ? group_residency<4, MODEL_MHD>(&per_cu[m]) : group_residency<4, MODEL_NS>(&per_cu[m]); break;
        }
    }
    if (err == hipSuccess) err = hipMalloc((void **)&e->dinst, sizeof(EnsInstance) * e->hinst.size());
    if (err == hipSuccess)
        err = hipMemcpy(e->dinst, e->hinst.data(), sizeof(EnsInstance) * e->hinst.size(), hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        release(e);
        delete e;
        return hip_fail(err, "ensemble set-up");
    }
    for (int m = 0; m < 2; m++)
        e->nblocks_m[m] = std::min(std::max(1, cus) * std::max(1, per_cu[m]), std::max(1, e->ntiles_m[m]));
    e->nblocks = (e->ntiles_m[0] ? e->nblocks_m[0] : 0) + (e->ntiles_m[1] ? e->nblocks_m[1] : 0);
    for (auto &cl : e->classes) cl.ctx->stream = e->stream;
    if (e->classes.size() > 1) {
        err = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
        for (size_t ci = 0; ci < e->classes.size() && err == hipSuccess; ci++) {
            hipStream_t st = nullptr;
            hipEvent_t ev = nullptr;
            err = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
            if (err == hipSuccess) e->side.push_back(st);
            if (err == hipSuccess) err = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (err == hipSuccess) e->done.push_back(ev);
        }
        if (err != hipSuccess) {
            release(e);
            delete e;
            return hip_fail(err, "ensemble side streams");
        }
    }
    *out = e;
    return NOVA_OK;
}

int ensemble_destroy(nova_ensemble *e)
{
    if (!e) return NOVA_OK;
    (void)hipStreamSynchronize(e->stream);
    release(e);
    delete e;
    return NOVA_OK;
}

int ensemble_set_stream(nova_ensemble *e, hipStream_t st)
{
    e->stream = st;
    for (auto &c : e->classes) c.ctx->stream = st;
    return NOVA_OK;
}

int ensemble_copy(nova_ensemble *e, int i, int store, double *U, double *B, double *p, double *pB)
{
    if (i < 0 || i >= (int)e->members.size()) return fail(NOVA_ERR_ARG, "instance index out of range");
    const auto &mb = e->members[i];
    const auto &cl = e->classes[mb.cls];
    const size_t vb = sizeof(double) * 3 * (size_t)cl.np, sb = sizeof(double) * (size_t)cl.np;
    double *dU = cl.U + mb.slotU * 3 * cl.np;
    double *dB = mb.slotB >= 0 ? cl.U + mb.slotB * 3 * cl.np : nullptr;
    double *dp = cl.P + mb.slotU * cl.np;
    double *dpB = mb.slotB >= 0 ? cl.P + mb.slotB * cl.np : nullptr;
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    if (!store) {
        if (U) NOVA_HIP(hipMemcpyAsync(dU, U, vb, k, e->stream));
        if (B && dB) NOVA_HIP(hipMemcpyAsync(dB, B, vb, k, e->stream));
        return NOVA_OK;
    }
    if (U) NOVA_HIP(hipMemcpyAsync(U, dU, vb, k, e->stream));
    if (B && dB) NOVA_HIP(hipMemcpyAsync(B, dB, vb, k, e->stream));
    if (p) NOVA_HIP(hipMemcpyAsync(p, dp, sb, k, e->stream));
    if (pB && dpB) NOVA_HIP(hipMemcpyAsync(pB, dpB, sb, k, e->stream));
    return NOVA_OK;
}

int ensemble_step(nova_ensemble *e, double dt, int nsteps)
{
    if (!e->use_graph) {
        for (int s = 0; s < nsteps; s++) {
            int rc = step_once(e, e->stream, dt);
            for (auto &c : e->classes) c.ctx->stream = e->stream;
            if (rc) return rc;
        }
        return NOVA_OK;
    }
    if (nsteps == 0) return NOVA_OK;
    if (!e->gexec || e->gdt != dt)
        if (int rc = record(e, dt)) return rc;
    hipEvent_t ev;
    NOVA_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    NOVA_HIP(hipEventRecord(ev, e->stream));
    NOVA_HIP(hipStreamWaitEvent(e->gstream, ev, 0));
    for (int s = 0; s < nsteps; s++) NOVA_HIP(hipGraphLaunch(e->gexec, e->gstream));
    NOVA_HIP(hipEventRecord(ev, e->gstream));
    NOVA_HIP(hipStreamWaitEvent(e->stream, ev, 0));
    NOVA_HIP(hipEventDestroy(ev));
    return NOVA_OK;
}

int ensemble_use_graph(nova_ensemble *e, int enable)
{
    e->use_graph = enable ? 1 : 0;
    return NOVA_OK;
}

int ensemble_info(nova_ensemble *e, int64_t out[6])
{
    int64_t pts = 0;
    for (const auto &mb : e->members) {
        const auto &d = e->classes[mb.cls].desc;
        pts += (int64_t)d.n[0] * d.n[1] * d.n[2];
    }
    out[0] = (int64_t)e->members.size();
    out[1] = (int64_t)e->classes.size();
    out[2] = e->ntiles;
    out[3] = e->nblocks;
    out[4] = pts;
    out[5] = 0;
    for (const auto &c : e->classes) out[5] += c.slots();
    return NOVA_OK;
}

nova_ctx *ensemble_ctx(nova_ensemble *e, int cls)
{
    return (cls >= 0 && cls < (int)e->classes.size()) ? e->classes[cls].ctx : nullptr;
}

}  // namespace nova
