// nova_stream.hpp -- the streamed (LDS plane + register queue) stage
// kernel shared by the forward stages (Navier-Stokes, MHD) and the
// tangent-linear and adjoint stages.
#pragma once
#include "nova_internal.hpp"
#include "nova_stencil.hpp"

namespace nova {
namespace {

enum { MODEL_NS = 0, MODEL_MHD = 1, MODEL_TLM = 2, MODEL_ADJ = 3 };
template <int MODEL>
constexpr int model_nf() { return MODEL == MODEL_NS ? 1 : 2; }
// whether the stage combination of the model reads a step-start field f
template <int MODEL>
constexpr bool model_x0(int f) { return MODEL == MODEL_MHD || (MODEL == MODEL_NS && f == 0) || (MODEL == MODEL_TLM && f == 1); }

// ---- the streamed stage kernel ----
//
// A 64 x 4 block owns a (k, j) tile and walks i over a chunk of planes.
// The i-neighbours of its own column live in a register queue q (planes
// i-H .. i+H); the current plane, with its j and k halos, is staged in
// LDS.  Every input value is loaded from global memory once per block
// (plus the halo ring and the 2H warm-up planes of the chunk), and the
// loads for plane i+1 are issued before plane i is computed.
constexpr int kTK = 64, kTJ = 4;

template <int NF>
struct StreamArgs {
    const double *X[NF];   // stage inputs (stencil)
    const double *X0[NF];  // step-start fields for the stage combination (may be null)
    double *out[NF];
    StageArgs sa;
    double nu, eta;
    // adjoint stage only: out = ob * T + yg * Y, acc = acc_a * acc + acc_b * m
    const double *Y;
    double *acc;
    double ob, yg, acc_a, acc_b;
    int ichunk;
    int nkt, njt, nic;
};

// MODEL_NS: X = {U}; MODEL_MHD: X = {U, B};  out = stage combination.
// MODEL_TLM: X = {U, d}, X0[1] = step-start tangent; out[0] = tangent stage.
// MODEL_ADJ: X = {U, m}; out[0] = ob (m + dt J_R(U)^T m) + yg Y.
// LDS doubles of one tile's staged plane
template <int H, int MODEL, int kTK, int kTJ = 256 / kTK>
constexpr int stage_lds()
{
    return 3 * model_nf<MODEL>() * (kTJ + 2 * H) * (kTK + 2 * H);
}

// One tile t = (kt, jt, ic) of a stage: the block streams planes i of its
// chunk for the (k, j) tile.  shbuf holds stage_lds() doubles of LDS.
template <int H, int MODEL, bool WALLS, int kTK, int PF = 1, int kTJ = 256 / kTK>
__device__ __attribute__((always_inline)) inline void stage_tile(const Grid &g, const StreamArgs<model_nf<MODEL>()> &a,
                                                                 int t, double *shbuf)
{
    constexpr int NF = model_nf<MODEL>();
    constexpr int QW = 2 * H + 1;
    constexpr int NC = 3 * NF;
    constexpr int LK = kTK + 2 * H, LJ = kTJ + 2 * H;
    constexpr int NHALO = 2 * H * (kTK + kTJ);
    constexpr int NSLOT = (NHALO + kTK * kTJ - 1) / (kTK * kTJ);
    auto sh = [&](int n, int r, int c) -> double & { return shbuf[(n * LJ + r) * LK + c]; };
    const int kt = t % a.nkt;
    const int jt = (t / a.nkt) % a.njt;
    const int ic = t / (a.nkt * a.njt);
    const int tk = threadIdx.x, tj = threadIdx.y;
    const int tid = tj * kTK + tk;
    const int k = kt * kTK + tk, j = jt * kTJ + tj;
    const bool valid = k < g.s[2] && j < g.s[1];
    const int i0 = ic * a.ichunk;
    int i1 = i0 + a.ichunk;
    if (i1 > g.s[0]) i1 = g.s[0];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    // element (component c, plane offset po, in-plane offset o): the plane
    // base is block-uniform and 64-bit, the lane offset a 32-bit byte offset
    // (a plane holds far fewer than 2^29 nodes), so loads address from a
    // scalar base plus one 32-bit register
    auto ldp = [&](const double *b, int c, int64_t po, uint32_t o) -> double {
        return *(const double *)((const char *)(b + c * g.np + po) + (uint32_t)(o * 8u));
    };
    auto stp = [&](double *b, int c, int64_t po, uint32_t o, double v) {
        *(double *)((char *)(b + c * g.np + po) + (uint32_t)(o * 8u)) = v;
    };

    // own column
    int fj, fk;
    const int jm = bmap(j, g.n[1], g.wall[1], fj);
    const int km = bmap(k, g.n[2], g.wall[2], fk);
    const uint32_t col = (uint32_t)jm * g.s[2] + km;
    double sgn_col[3];
#pragma unroll
    for (int c = 0; c < 3; c++) sgn_col[c] = rsign<WALLS>(c, 1, fj) * rsign<WALLS>(c, 2, fk);

    // halo slots of this thread: LDS position, source column, signs
    int hrow[NSLOT], hcolk[NSLOT];
    uint32_t hsrc[NSLOT];
    double hsg[NSLOT][3];
    bool hon[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; s++) {
        const int e = tid + s * kTK * kTJ;
        hon[s] = e < NHALO;
        int r = 0, cc = 0, jj = 0, kk = 0;
        if (e < 2 * H * kTK) {  // j halo rows
            const int rr = e / kTK;
            cc = e % kTK + H;
            r = rr < H ? rr : kTJ + rr;
            jj = jt * kTJ + (rr < H ? rr - H : kTJ + rr - H);
            kk = kt * kTK + (cc - H);
        } else if (e < NHALO) {  // k halo columns
            const int e2 = e - 2 * H * kTK;
            const int rr = e2 / (2 * H), c2 = e2 % (2 * H);
            r = rr + H;
            cc = c2 < H ? c2 : kTK + c2;
            jj = jt * kTJ + rr;
            kk = kt * kTK + (c2 < H ? c2 - H : kTK + c2 - H);
        }
        int f1, f2;
        const int jm2 = bmap(jj, g.n[1], g.wall[1], f1);
        const int km2 = bmap(kk, g.n[2], g.wall[2], f2);
        hrow[s] = r;
        hcolk[s] = cc;
        hsrc[s] = (uint32_t)jm2 * g.s[2] + km2;
#pragma unroll
        for (int c = 0; c < 3; c++) hsg[s][c] = rsign<WALLS>(c, 1, f1) * rsign<WALLS>(c, 2, f2);
    }

    // loaders
    auto plane_off = [&](int ii, int &fi) -> int64_t {
        const int im = bmap(ii, g.n[0], g.wall[0], fi);
        return (int64_t)im * plane;
    };

    double q[NF][3][QW];
#pragma unroll
    for (int o = 0; o < QW; o++) {
        int fi;
        const int64_t po = plane_off(i0 - H + o, fi);
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                q[f][c][o] = rsign<WALLS>(c, 0, fi) * sgn_col[c] * ldp(a.X[f], c, po, col);
            }
    }
    // prefetched data of plane p: the queue entry of plane p + H, the halo of
    // plane p and the step-start values of plane p
    // Values are kept as loaded; the reflection signs (walls) are applied
    // when they are consumed, so that no arithmetic waits on a load in flight.
    struct Pf {
        double nq[NF][3];
        double hv[NSLOT][NC];
        double nx0[NF][3];
        double ny[3], nacc[3];  // adjoint stage: pointwise Y and accumulator
        int fq, fh;             // reflection flags of the queue plane and the halo plane
    };
    constexpr bool ADJ = MODEL == MODEL_ADJ;
    // Every lane issues the same loads for every plane: slots past the halo
    // read column (0, 0), lanes outside the grid read their mapped column,
    // and the prefetch past the end of a chunk re-reads its last plane.  With
    // a fixed count of memory operations per plane the wait for one plane's
    // data leaves the loads of the next plane in flight.
    auto load_pf = [&](Pf &P, int p, bool queue) {
        if (queue) {
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++)
                    P.nx0[f][c] = (model_x0<MODEL>(f) && a.X0[f]) ? ldp(a.X0[f], c, (int64_t)p * plane, col) : 0.0;
            if constexpr (ADJ) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    P.ny[c] = a.Y ? ldp(a.Y, c, (int64_t)p * plane, col) : 0.0;
                    P.nacc[c] = (a.acc && a.acc_a != 0.0) ? ldp(a.acc, c, (int64_t)p * plane, col) : 0.0;
                }
            }
            const int64_t po = plane_off(p + H, P.fq);
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) P.nq[f][c] = ldp(a.X[f], c, po, col);
        }
        const int64_t ph = plane_off(p, P.fh);
#pragma unroll
        for (int s = 0; s < NSLOT; s++)
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) P.hv[s][f * 3 + c] = ldp(a.X[f], c, ph, hsrc[s]);
    };

    const double zi = g.zinv[km], zl = g.zlap[km], zm = g.zmet[km];
    // step-start values of the own column, one plane ahead like the queue
    double x0c[NF][3];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            x0c[f][c] = (model_x0<MODEL>(f) && a.X0[f]) ? ldp(a.X0[f], c, (int64_t)i0 * plane, col) : 0.0;
    double yc[3] = {0.0, 0.0, 0.0}, accc[3] = {0.0, 0.0, 0.0};
    if constexpr (ADJ) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            yc[c] = a.Y ? ldp(a.Y, c, (int64_t)i0 * plane, col) : 0.0;
            accc[c] = (a.acc && a.acc_a != 0.0) ? ldp(a.acc, c, (int64_t)i0 * plane, col) : 0.0;
        }
    }

    // One plane: stage plane i (halo from Cur) in LDS, issue the loads of
    // plane i + PF into the set they go to, compute plane i, then advance
    // the queue with Nxt (the data of plane i + 1).  PF = 1: Cur and Nxt
    // are one set; PF = 2: two sets alternate, so two planes of loads are
    // in flight while a plane is computed.
    auto body = [&](int i, Pf &Cur, Pf &Nxt) {
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int c = 0; c < 3; c++) sh(f * 3 + c, tj + H, tk + H) = q[f][c][H];
#pragma unroll
        for (int s = 0; s < NSLOT; s++)
            if (hon[s])
#pragma unroll
                for (int n = 0; n < NC; n++)
                    sh(n, hrow[s], hcolk[s]) = rsign<WALLS>(n % 3, 0, Cur.fh) * hsg[s][n % 3] * Cur.hv[s][n];
        __syncthreads();
        load_pf(Cur, i + PF < i1 ? i + PF : i1 - 1, true);

        if (valid) {
            auto A = [&](int f, int c, int d, int o) -> double {
                if (d == 0) return q[f][c][H + o];
                if (d == 1) return sh(f * 3 + c, tj + H + o, tk + H);
                return sh(f * 3 + c, tj + H, tk + H + o);
            };
            const int64_t pi = (int64_t)i * plane;
            const StageArgs &sa = a.sa;
            if constexpr (MODEL == MODEL_NS) {
                double r[3];
                ns_rhs<H>(g, A, a.nu, zi, zl, zm, r);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double du = sa.dt * r[c];
                    const double u = A(0, c, 0, 0);
                    const double u0 = a.X0[0] ? x0c[0][c] : u;
                    stp(a.out[0], c, pi, col, combine(sa, u, u0, du));
                }
            } else if constexpr (MODEL == MODEL_TLM) {
                double r[3];
                tlm_rhs<H>(g, A, a.nu, zi, zl, zm, r);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double v = q[1][c][H];
                    const double v0 = a.X0[1] ? x0c[1][c] : v;
                    stp(a.out[0], c, pi, col, combine(sa, v, v0, sa.dt * r[c]));
                }
            } else if constexpr (MODEL == MODEL_ADJ) {
                double r[3];
                adj_rhs<H>(g, A, a.nu, zi, zl, r);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double m = q[1][c][H];
                    double o = a.ob * (m + sa.dt * r[c]);
                    if (a.Y) o += a.yg * yc[c];
                    if (a.acc) stp(a.acc, c, pi, col, (a.acc_a != 0.0 ? a.acc_a * accc[c] : 0.0) + a.acc_b * m);
                    stp(a.out[0], c, pi, col, o);
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double du =
                        sa.dt * ((((-div_xy<H>(g, A, 0, 0, c, zi)) + lap_c<H>(g, A, 0, c, a.nu, zl, zm)) +
                                  div_xy<H>(g, A, 1, 1, c, zi)) - grad_pm<H>(g, A, 1, c, zi));
                    const double db = sa.dt * (((-div_xy<H>(g, A, 0, 1, c, zi)) + div_xy<H>(g, A, 1, 0, c, zi)) +
                                               lap_c<H>(g, A, 1, c, a.eta, zl, zm));
                    const double u = A(0, c, 0, 0), b = q[NF - 1][c][H];
                    const double u0 = a.X0[0] ? x0c[0][c] : u;
                    const double b0 = a.X0[NF - 1] ? x0c[NF - 1][c] : b;
                    stp(a.out[0], c, pi, col, combine(sa, u, u0, du));
                    stp(a.out[NF - 1], c, pi, col, combine(sa, b, b0, db));
                }
            }
        }
        __syncthreads();
        {
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
#pragma unroll
                    for (int o = 0; o < 2 * H; o++) q[f][c][o] = q[f][c][o + 1];
                    q[f][c][2 * H] = rsign<WALLS>(c, 0, Nxt.fq) * sgn_col[c] * Nxt.nq[f][c];
                    x0c[f][c] = Nxt.nx0[f][c];
                }
            if constexpr (ADJ) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    yc[c] = Nxt.ny[c];
                    accc[c] = Nxt.nacc[c];
                }
            }
        }
    };

    if constexpr (PF == 1) {
        Pf S;
        load_pf(S, i0, false);
        for (int i = i0; i < i1; i++) body(i, S, S);
    } else {
        Pf A, B;
        load_pf(A, i0, false);
        load_pf(B, i0 + 1 < i1 ? i0 + 1 : i0, true);
        int i = i0;
        for (; i + 1 < i1; i += 2) {
            body(i, A, B);
            body(i + 1, B, A);
        }
        if (i < i1) body(i, A, B);
    }
}

template <int H, int MODEL, bool WALLS, int MINW, int kTK, int kTJ>
__global__ void __launch_bounds__(kTK *kTJ) __attribute__((amdgpu_waves_per_eu(MINW)))
k_stage_stream(Grid g, StreamArgs<model_nf<MODEL>()> a)
{
    __shared__ double sh[stage_lds<H, MODEL, kTK, kTJ>()];
    // consecutive tiles share halos, so each XCD (blocks are dealt
    // round-robin over 8) takes a contiguous run of tiles
    const int T = a.nkt * a.njt * a.nic;
    int t = blockIdx.x;
    if (T % 8 == 0) t = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
    // shapes held to one or two waves per SIMD have the registers for a
    // second prefetch set
    stage_tile<H, MODEL, WALLS, kTK, MINW <= 2 ? 2 : 1, kTJ>(g, a, t, sh);
}

int pick_ichunk(const Grid &g, int nkt, int njt)
{
    // about 2048 tiles (8 blocks per CU on 256 CUs), chunks of 8..64 planes
    int64_t per = (int64_t)g.s[0] * nkt * njt / 2048;
    int c = (int)(per < 8 ? 8 : (per > 64 ? 64 : per));
    if (c > g.s[0]) c = g.s[0];
    return c;
}

// ---- the DMA-staged Navier-Stokes stage (periodic one-field) ----
//
// The block's haloed (j, k) planes travel from global memory straight into
// a ring of R = H + 1 + D LDS buffers (global_load_lds: no registers hold
// data in flight), D planes ahead of the plane that first needs them.  Each
// buffer holds one haloed plane of the three components; the plane i + H
// of the ring feeds the register queue of the i-neighbours, the plane i
// the j and k stencils.  The buffers are distinct LDS variables addressed
// with compile-time slot numbers (the loop is unrolled over the ring), so
// the wait for one buffer's data leaves the loads into the others in
// flight.
template <int N>
struct IC {
    static constexpr int value = N;
};

template <int B, int E, class F>
__device__ __attribute__((always_inline)) inline void static_for(F &&f)
{
    if constexpr (B < E) {
        f(IC<B>{});
        static_for<B + 1, E>(f);
    }
}

template <int H, int kTK, int kTJ>
struct DmaGeom {
    static constexpr int LK = kTK + 2 * H, LJ = kTJ + 2 * H;
    static constexpr int NPL = LK * LJ;                // nodes of one haloed plane
    static constexpr int PADC = (NPL + 31) / 32 * 32;  // one component, whole 32-double loads
    static constexpr int NI = 3 * PADC / 32;           // loads (of 32 doubles) per plane
    static constexpr int NW = kTK * kTJ / 64;          // waves per block
    static constexpr int NIW = (NI + NW - 1) / NW;     // loads per wave per plane
    static constexpr int PLANE = NIW * NW * 32;        // doubles per ring buffer
};

template <int H, int D>
constexpr int dma_ring() { return H + 1 + D; }

// s_waitcnt with only the vector-memory count set (the others at their
// maximum): vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8
template <int N>
__device__ __attribute__((always_inline)) inline void wait_vm()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

typedef double dbl2 __attribute__((ext_vector_type(2)));

// LDS byte address of a __shared__ location
__device__ __attribute__((always_inline)) inline uint32_t lds_addr(const double *p)
{
    return (uint32_t)(size_t)(const __attribute__((address_space(3))) double *)p;
}

// The stencil, queue and step-start reads of the ring are written as LDS
// reads that wait for themselves (s_waitcnt lgkmcnt(0) inside the block),
// the step-start values travel through LDS as well, and the block barrier
// is a bare s_barrier: the compiler waits for every outstanding load
// (vmcnt(0)) before any LDS read, fence or use of a loaded register that
// follows a global_load_lds, which would drain the ring each plane.  The
// explicit vmcnt wait before the barrier is what orders the ring: the
// loads of the buffers read next are complete in every wave when the
// barrier opens.
template <int H, int kTK, int kTJ, int D>
__global__ void __launch_bounds__(kTK *kTJ) k_stage_dma(Grid g, StreamArgs<1> a)
{
    using G = DmaGeom<H, kTK, kTJ>;
    constexpr int R = dma_ring<H, D>();
    constexpr int NT = kTK * kTJ;
    constexpr int NXW = 3 * NT / 32 / G::NW;  // step-start loads per wave per plane
    static_assert(H == 3 && kTK == 32 && kTJ == 16 && D == 2,
                  "read offsets are written for a 32 x 16 tile with a 3-wide halo");
    static_assert(G::LK == 38 && G::PADC == 864 && R % 2 == 0, "layout of the read offsets");
    static_assert(3 * NT % (32 * G::NW) == 0, "whole step-start loads per wave");
    __shared__ double r0[G::PLANE], r1[G::PLANE], r2[G::PLANE], r3[G::PLANE], r4[G::PLANE], r5[G::PLANE];
    __shared__ double x0s0[3 * NT], x0s1[3 * NT];
    auto slot = [&](auto S) -> double * {
        constexpr int s = decltype(S)::value;
        if constexpr (s == 0) return r0;
        else if constexpr (s == 1) return r1;
        else if constexpr (s == 2) return r2;
        else if constexpr (s == 3) return r3;
        else if constexpr (s == 4) return r4;
        else return r5;
    };
    auto xslot = [&](auto S) -> double * {
        if constexpr (decltype(S)::value % 2 == 0) return x0s0;
        else return x0s1;
    };

    const int T = a.nkt * a.njt * a.nic;
    int t = blockIdx.x;
    if (T % 8 == 0) t = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
    const int kt = t % a.nkt;
    const int jt = (t / a.nkt) % a.njt;
    const int ic = t / (a.nkt * a.njt);
    const int tk = threadIdx.x, tj = threadIdx.y;
    const int tid = tj * kTK + tk;
    const int k = kt * kTK + tk, j = jt * kTJ + tj;
    const bool valid = k < g.s[2] && j < g.s[1];
    const int i0 = ic * a.ichunk;
    int i1 = i0 + a.ichunk;
    if (i1 > g.s[0]) i1 = g.s[0];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    int f;
    const int km = bmap(k, g.n[2], g.wall[2], f);
    const uint32_t col = (uint32_t)bmap(j, g.n[1], g.wall[1], f) * g.s[2] + km;
    auto ldp = [&](const double *b, int c, int64_t po, uint32_t o) -> double {
        return *(const double *)((const char *)(b + c * g.np + po) + (uint32_t)(o * 8u));
    };
    auto stp = [&](double *b, int c, int64_t po, uint32_t o, double v) {
        *(double *)((char *)(b + c * g.np + po) + (uint32_t)(o * 8u)) = v;
    };

    // this lane's part of the plane loads: load m = w + NW u of the block
    // fills doubles 32 m .. 32 m + 31 of a buffer (lane L the low or high
    // word of double 32 m + L/2); per load the byte offset from the plane
    // base of the field, component included (three components of a field
    // span less than 4 GiB).  Ring buffers hold the haloed plane (component,
    // row, column); step-start buffers the tile interior (component, tj,
    // tk), one row of 32 per load.
    const int lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t cstride = (uint32_t)(g.np * 8);
    uint32_t loff[G::NIW];
#pragma unroll
    for (int u = 0; u < G::NIW; u++) {
        const int m = w + G::NW * u;
        const int cm = (32 * m) / G::PADC;
        const int e2 = 32 * m + (lane >> 1) - cm * G::PADC;
        uint32_t off = 0;
        if (cm < 3 && e2 < G::NPL) {
            const int r = e2 / G::LK, cc = e2 - r * G::LK;
            int f1, f2;
            off = (uint32_t)bmap(jt * kTJ - H + r, g.n[1], g.wall[1], f1) * g.s[2] +
                  bmap(kt * kTK - H + cc, g.n[2], g.wall[2], f2);
        }
        loff[u] = (uint32_t)(cm < 3 ? cm : 2) * cstride + off * 8u + 4u * (uint32_t)(lane & 1);
    }
    uint32_t xoff[NXW];
#pragma unroll
    for (int u = 0; u < NXW; u++) {
        const int m = w + G::NW * u;
        const int cm = m / kTJ, row = m - cm * kTJ;
        int f1, f2;
        const uint32_t off = (uint32_t)bmap(jt * kTJ + row, g.n[1], g.wall[1], f1) * g.s[2] +
                             bmap(kt * kTK + (lane >> 1), g.n[2], g.wall[2], f2);
        xoff[u] = (uint32_t)cm * cstride + off * 8u + 4u * (uint32_t)(lane & 1);
    }
    auto dma = [&](auto S, int p) {
        int fi;
        const char *pb = (const char *)(a.X[0] + (int64_t)bmap(p, g.n[0], g.wall[0], fi) * plane);
        double *ls = slot(S);
#pragma unroll
        for (int u = 0; u < G::NIW; u++)
            __builtin_amdgcn_global_load_lds((const void *)(pb + loff[u]),
                                             (__attribute__((address_space(3))) void *)(ls + 32 * (w + G::NW * u)),
                                             4, 0, 0);
    };
    auto dma_x0 = [&](auto S, int p) {
        if (!a.X0[0]) return;
        const char *pb = (const char *)(a.X0[0] + (int64_t)(p < g.s[0] ? p : g.s[0] - 1) * plane);
        double *ls = xslot(S);
#pragma unroll
        for (int u = 0; u < NXW; u++)
            __builtin_amdgcn_global_load_lds((const void *)(pb + xoff[u]),
                                             (__attribute__((address_space(3))) void *)(ls + 32 * (w + G::NW * u)),
                                             4, 0, 0);
    };

    // queue of the own column: planes i - H .. i + H (entry 2H arrives from
    // the ring at plane i)
    double q[3][2 * H + 1];
#pragma unroll
    for (int o = 0; o < 2 * H; o++) {
        int fi;
        const int64_t po = (int64_t)bmap(i0 - H + o, g.n[0], g.wall[0], fi) * plane;
#pragma unroll
        for (int c = 0; c < 3; c++) q[c][o] = ldp(a.X[0], c, po, col);
    }
    const double zi = g.zinv[km], zl = g.zlap[km], zm = g.zmet[km];
    // the ring starts with planes i0 .. i0 + H + D - 1 in buffers 0 .. R - 2,
    // the step-start buffer 0 with plane i0 (one ring plane after it, so
    // that the wait of the first plane covers it)
    static_for<0, R - 2>([&](auto S) { dma(S, i0 + decltype(S)::value); });
    dma_x0(IC<0>{}, i0);
    dma(IC<R - 2>{}, i0 + R - 2);

    // byte offsets in a buffer: the 7 x 7 window corner (tj, tk) of this
    // thread and its own node (tj + H, tk + H); in a step-start buffer the
    // own node
    const uint32_t win = (uint32_t)(tj * G::LK + tk) * 8u;
    const uint32_t own = (uint32_t)((tj + H) * G::LK + tk + H) * 8u;
    const uint32_t xown = (uint32_t)tid * 8u;
    const StageArgs &sa = a.sa;

    // plane i from ring buffer S.  Every plane issues the same loads (planes
    // past the chunk end are computed and not stored).  At the barrier of
    // plane i the ring plane i + H (issued D planes back) and the
    // step-start plane i (issued one plane back, before that plane's ring
    // load) must be complete: both have at least the NIW ring loads of
    // the previous plane after them.
    auto body = [&](auto S, int i) {
        constexpr int s = decltype(S)::value;
        wait_vm<G::NIW>();
        asm volatile("s_barrier" ::: "memory");
        dma_x0(IC<(s + 1) % 2>{}, i + 1);
        dma(IC<(s + R - 1) % R>{}, i + H + D);
        // j and k neighbours (offsets in doubles from the window corner:
        // column k = tk + 3 at rows 0..6 -> 38 r + 3, row j = tj + 3 at
        // columns 0..6 -> 114 + c; the centre is the queue's), the queue
        // entry of plane i + H and the step-start values of plane i
        // (components 864 and 512 doubles apart)
        const uint32_t sb = lds_addr(slot(S)) + win;
        const uint32_t b0 = sb, b1 = sb + G::PADC * 8u, b2 = sb + 2u * G::PADC * 8u;
        const uint32_t qb = lds_addr(slot(IC<(s + H) % R>{})) + own;
        const uint32_t xb = lds_addr(xslot(S)) + xown;
        dbl2 v[3][6];
        double qn[3], xv[3];
        asm volatile(
            "ds_read2_b64 %0, %24 offset0:3 offset1:41\n\t"
            "ds_read2_b64 %1, %24 offset0:79 offset1:155\n\t"
            "ds_read2_b64 %2, %24 offset0:193 offset1:231\n\t"
            "ds_read2_b64 %3, %24 offset0:114 offset1:115\n\t"
            "ds_read2_b64 %4, %24 offset0:116 offset1:118\n\t"
            "ds_read2_b64 %5, %24 offset0:119 offset1:120\n\t"
            "ds_read2_b64 %6, %25 offset0:3 offset1:41\n\t"
            "ds_read2_b64 %7, %25 offset0:79 offset1:155\n\t"
            "ds_read2_b64 %8, %25 offset0:193 offset1:231\n\t"
            "ds_read2_b64 %9, %25 offset0:114 offset1:115\n\t"
            "ds_read2_b64 %10, %25 offset0:116 offset1:118\n\t"
            "ds_read2_b64 %11, %25 offset0:119 offset1:120\n\t"
            "ds_read2_b64 %12, %26 offset0:3 offset1:41\n\t"
            "ds_read2_b64 %13, %26 offset0:79 offset1:155\n\t"
            "ds_read2_b64 %14, %26 offset0:193 offset1:231\n\t"
            "ds_read2_b64 %15, %26 offset0:114 offset1:115\n\t"
            "ds_read2_b64 %16, %26 offset0:116 offset1:118\n\t"
            "ds_read2_b64 %17, %26 offset0:119 offset1:120\n\t"
            "ds_read_b64 %18, %27\n\t"
            "ds_read_b64 %19, %27 offset:6912\n\t"
            "ds_read_b64 %20, %27 offset:13824\n\t"
            "ds_read_b64 %21, %28\n\t"
            "ds_read_b64 %22, %28 offset:4096\n\t"
            "ds_read_b64 %23, %28 offset:8192\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(v[0][0]), "=&v"(v[0][1]), "=&v"(v[0][2]), "=&v"(v[0][3]), "=&v"(v[0][4]), "=&v"(v[0][5]),
              "=&v"(v[1][0]), "=&v"(v[1][1]), "=&v"(v[1][2]), "=&v"(v[1][3]), "=&v"(v[1][4]), "=&v"(v[1][5]),
              "=&v"(v[2][0]), "=&v"(v[2][1]), "=&v"(v[2][2]), "=&v"(v[2][3]), "=&v"(v[2][4]), "=&v"(v[2][5]),
              "=&v"(qn[0]), "=&v"(qn[1]), "=&v"(qn[2]), "=&v"(xv[0]), "=&v"(xv[1]), "=&v"(xv[2])
            : "v"(b0), "v"(b1), "v"(b2), "v"(qb), "v"(xb)
            : "memory");
#pragma unroll
        for (int c = 0; c < 3; c++) q[c][2 * H] = qn[c];
        // neighbours by offset -3..3 (index o + 3; the centre entries are
        // not read)
        double nj[3][7], nk[3][7];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            nj[c][0] = v[c][0].x; nj[c][1] = v[c][0].y; nj[c][2] = v[c][1].x; nj[c][3] = 0.0;
            nj[c][4] = v[c][1].y; nj[c][5] = v[c][2].x; nj[c][6] = v[c][2].y;
            nk[c][0] = v[c][3].x; nk[c][1] = v[c][3].y; nk[c][2] = v[c][4].x; nk[c][3] = 0.0;
            nk[c][4] = v[c][4].y; nk[c][5] = v[c][5].x; nk[c][6] = v[c][5].y;
        }
        auto A = [&](int, int c, int d, int o) -> double {
            if (d == 0) return q[c][H + o];
            if (d == 1) return nj[c][H + o];
            return nk[c][H + o];
        };
        double r[3];
        ns_rhs<H>(g, A, a.nu, zi, zl, zm, r);
        if (valid && i < i1) {
            const int64_t pi = (int64_t)i * plane;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double du = sa.dt * r[c];
                const double u = q[c][H];
                stp(a.out[0], c, pi, col, combine(sa, u, a.X0[0] ? xv[c] : u, du));
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int o = 0; o < 2 * H; o++) q[c][o] = q[c][o + 1];
    };
    for (int b = i0; b < i1; b += R) static_for<0, R>([&](auto S) { body(S, b + decltype(S)::value); });
    wait_vm<0>();
}

template <int H, int TK, int TJ, int D>
void launch_stage_dma(nova_ctx *ctx, StreamArgs<1> a)
{
    const Grid &g = ctx->g;
    constexpr int R = dma_ring<H, D>();
    a.nkt = (g.s[2] + TK - 1) / TK;
    a.njt = (g.s[1] + TJ - 1) / TJ;
    const int ic = ctx->stage_ichunk[MODEL_NS];
    int c = ic > 0 ? ic : pick_ichunk(g, a.nkt, a.njt);
    // whole turns of the ring per chunk
    c = (c + R - 1) / R * R;
    a.ichunk = c < g.s[0] ? c : g.s[0];
    a.nic = (g.s[0] + a.ichunk - 1) / a.ichunk;
    const int T = a.nkt * a.njt * a.nic;
    hipLaunchKernelGGL((k_stage_dma<H, TK, TJ, D>), dim3(T), dim3(TK, TJ, 1), 0, ctx->stream, g, a);
}


template <int H, int MODEL, bool WALLS, int MINW, int TK, int TJ = 256 / TK>
void launch_stream_tile(nova_ctx *ctx, StreamArgs<model_nf<MODEL>()> a)
{
    const Grid &g = ctx->g;
    a.nkt = (g.s[2] + TK - 1) / TK;
    a.njt = (g.s[1] + TJ - 1) / TJ;
    const int ic = ctx->stage_ichunk[MODEL];
    a.ichunk = ic > 0 ? (ic < g.s[0] ? ic : g.s[0]) : pick_ichunk(g, a.nkt, a.njt);
    a.nic = (g.s[0] + a.ichunk - 1) / a.ichunk;
    const int T = a.nkt * a.njt * a.nic;
    hipLaunchKernelGGL((k_stage_stream<H, MODEL, WALLS, MINW, TK, TJ>), dim3(T), dim3(TK, TJ, 1), 0, ctx->stream, g,
                       a);
}

template <int H, int MODEL, bool WALLS>
void launch_stream_w(nova_ctx *ctx, const StreamArgs<model_nf<MODEL>()> &a)
{
    // variants 2-4 hold one-field models to 3 waves per SIMD; two queued
    // fields need more than 256 registers and run unspilled at one wave per
    // SIMD (spilled shapes measure 3-4x slower)
    constexpr int W = model_nf<MODEL>() == 1 ? 3 : 1;
    switch (ctx->stage_variant[MODEL]) {
    case 0: launch_stream_tile<H, MODEL, WALLS, 1, 64>(ctx, a); break;
    case 2: launch_stream_tile<H, MODEL, WALLS, W, 64>(ctx, a); break;
    case 4: launch_stream_tile<H, MODEL, WALLS, W, 16>(ctx, a); break;
    // two waves per SIMD with two planes of loads in flight (one-field models)
    case 5: launch_stream_tile<H, MODEL, WALLS, W == 3 ? 2 : 1, 32>(ctx, a); break;
    case 6: launch_stream_tile<H, MODEL, WALLS, W == 3 ? 2 : 1, 16>(ctx, a); break;
    // 512-thread tiles (32 x 16, 64 x 8): a smaller halo ring per node.  A
    // 512-thread block is held to 256 registers per lane, which two queued
    // fields exceed, so two-field models keep their 256-thread shapes here
    case 7:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, 2, 32, 16>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, 1, 32>(ctx, a);
        break;
    case 8:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, 2, 64, 8>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, 1, 64>(ctx, a);
        break;
    // DMA-staged ring (periodic sixth-order Navier-Stokes stage, 32 x 16
    // tiles, two planes of loads ahead; the others use the default shape)
    case 9:
        // (its per-lane load offsets span the three components in 32 bits)
        if constexpr (MODEL == MODEL_NS && !WALLS && H == 3) {
            if (3 * ctx->g.np * 8 < ((int64_t)1 << 32)) launch_stage_dma<H, 32, 16, 2>(ctx, a);
            else launch_stream_tile<H, MODEL, WALLS, W, 32>(ctx, a);
        } else {
            launch_stream_tile<H, MODEL, WALLS, W, 32>(ctx, a);
        }
        break;
    default: launch_stream_tile<H, MODEL, WALLS, W, 32>(ctx, a); break;
    }
}

// Streamed launch for any model and stencil half-width (the row-stream
// variant of the forward stages is dispatched by their own launcher).
template <int MODEL>
int launch_stream_model(nova_ctx *ctx, const StreamArgs<model_nf<MODEL>()> &a, const char *name)
{
    const Grid &g = ctx->g;
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    TimedScope ts(ctx, name);
    switch (g.h * 2 + (walls ? 1 : 0)) {
    case 2: launch_stream_w<1, MODEL, false>(ctx, a); break;
    case 3: launch_stream_w<1, MODEL, true>(ctx, a); break;
    case 4: launch_stream_w<2, MODEL, false>(ctx, a); break;
    case 5: launch_stream_w<2, MODEL, true>(ctx, a); break;
    case 6: launch_stream_w<3, MODEL, false>(ctx, a); break;
    case 7: launch_stream_w<3, MODEL, true>(ctx, a); break;
    case 8: launch_stream_w<4, MODEL, false>(ctx, a); break;
    case 9: launch_stream_w<4, MODEL, true>(ctx, a); break;
    default: return fail(NOVA_ERR_ARG, "stencil order");
    }
    return check_launch(name);
}

}  // namespace
}  // namespace nova
