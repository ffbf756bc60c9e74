// nova_wc.hip -- the weakly-compressible step (app/TG/TG.cpp:71-76 with
// rhoEqn.H, pEqn.H and UEqn.H) on gfx950.
//
// The reference makes, per stage, one pass per operator and per algebraic
// term over pool temporaries: rho*U, ex::div, the negation, the dt scaling,
// the accumulator update and the stage value for rho (rhoEqn.H:19-33), and
// rho*U, ex::div(rho*U, U), ex::grad(p), ex::laplacian(mu, U), their sum,
// rhoU accumulator and U = rhoUOld/rho + c drhoU/rho for U (UEqn.H:16-33).
// Here each stage is one kernel: a block owns a 32 x 8 (k, j) tile and a
// chunk of i planes, stages each haloed (j, k) plane of its inputs in LDS
// and keeps the i-neighbours of each thread's column in register queues;
// the products rho*U are formed per node, as the reference's product field
// holds them.
//
// Stage buffers: rho ping-pongs user -> ra -> rb -> ra -> user with the
// accumulator in rc; U ping-pongs user -> ua -> ub -> ua -> user with the
// rho U accumulator in uc.  The user's U stays the step-start value until
// the last U stage, so rhoUOld = rho*U is formed from it at the node.
#include "nova_internal.hpp"
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

constexpr int kWcTK = 32, kWcTJ = 8;  // threads of a block along k and j (one node each)

struct WcRhoArgs {
    const double *rho_in;  // stage value (stencil)
    const double *U;       // frozen velocity (stencil)
    const double *rho0;    // step-start rho
    const double *rc_in;   // accumulator
    double *rho_out, *rc_out, *p;
    double dt, w, cs, pc;
    int last;
    int p_left;  // p = (pc rho) rho (app/shearLayerWC/pEqn.H:2) instead of pc (rho rho)
    int ichunk, nkt, njt, nic;
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
    int ichunk, nkt, njt, nic;
};

// The (k, j) tile and i chunk of this block (consecutive tiles on one
// XCD), the thread's node and whether it lies in the grid.
struct WcTile {
    int k0, j0, k, j, i0, i1;
    bool valid;
};

__device__ inline WcTile wc_tile(const Grid &g, int ichunk, int nkt, int njt, int nic)
{
    const int T = nkt * njt * nic;
    int t = blockIdx.x;
    if (T % 8 == 0) t = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
    const int kt = t % nkt, jt = (t / nkt) % njt, ic = t / (nkt * njt);
    WcTile w;
    w.k0 = kt * kWcTK;
    w.j0 = jt * kWcTJ;
    w.k = w.k0 + threadIdx.x;
    w.j = w.j0 + threadIdx.y;
    w.valid = w.k < g.s[2] && w.j < g.s[1];
    w.i0 = ic * ichunk;
    w.i1 = w.i0 + ichunk < g.s[0] ? w.i0 + ichunk : g.s[0];
    return w;
}

// Node offset and reflection flags of (i, j, k) (indices at most H outside
// the grid, or inside a partial tile).
template <bool WALLS>
__device__ inline int64_t wc_node(const Grid &g, int i, int j, int k, int (&fl)[3])
{
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    return (int64_t)bmap(i, g.n[0], g.wall[0], fl[0]) * plane + (int64_t)bmap(j, g.n[1], g.wall[1], fl[1]) * g.s[2] +
           bmap(k, g.n[2], g.wall[2], fl[2]);
}

// Component c of a vector field at a node reached through reflections.
template <bool WALLS>
__device__ inline double wc_vec(const Grid &g, const double *F, int c, int64_t off, const int (&fl)[3])
{
    return rsign<WALLS>(g, c, 0, fl[0]) * rsign<WALLS>(g, c, 1, fl[1]) * rsign<WALLS>(g, c, 2, fl[2]) * F[c * g.np + off];
}

// The thread's share of the haloed (j, k) plane of the tile: slot s holds
// position e = tid + s * threads (row r = e / LK, column cc = e % LK; rows
// j0 - H + r, columns k0 - H + cc) with its in-plane offset and reflection
// flags, fixed for every plane.
template <int H>
struct WcSlots {
    static constexpr int LK = kWcTK + 2 * H, LJ = kWcTJ + 2 * H;
    static constexpr int NS = (LJ * LK + kWcTK * kWcTJ - 1) / (kWcTK * kWcTJ);
    int64_t off[NS];
    int pos[NS];  // e, or -1 past the plane
    int fj[NS], fk[NS];
};

template <int H>
__device__ inline void wc_slots(const Grid &g, const WcTile &tl, WcSlots<H> &S)
{
    constexpr int LK = WcSlots<H>::LK, LJ = WcSlots<H>::LJ;
    const int tid = threadIdx.y * kWcTK + threadIdx.x;
#pragma unroll
    for (int s = 0; s < WcSlots<H>::NS; s++) {
        const int e = tid + s * kWcTK * kWcTJ;
        const int r = e / LK, cc = e - r * LK;
        S.pos[s] = e < LJ * LK ? e : -1;
        S.off[s] = (int64_t)bmap(tl.j0 - H + r, g.n[1], g.wall[1], S.fj[s]) * g.s[2] +
                   bmap(tl.k0 - H + cc, g.n[2], g.wall[2], S.fk[s]);
    }
}

// Stages NC values per node of the haloed plane i of the tile in LDS
// (sh[c][e]); val(off, fl, c) gives component c at a node.
// Every slot's values are loaded (a slot past the plane reads the mapped
// node wc_slots gave it, a valid one) before any is written, so that the
// plane's loads are in flight together instead of one slot at a time.
template <int H, int NC, class Val>
__device__ inline void wc_stage_plane(const Grid &g, const WcSlots<H> &S, int i, double *sh, Val val)
{
    constexpr int NP = WcSlots<H>::LK * WcSlots<H>::LJ;
    constexpr int NS = WcSlots<H>::NS;
    int fl[3];
    const int64_t po = (int64_t)bmap(i, g.n[0], g.wall[0], fl[0]) * g.s[1] * g.s[2];
    double v[NS][NC];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        fl[1] = S.fj[s];
        fl[2] = S.fk[s];
#pragma unroll
        for (int c = 0; c < NC; c++) v[s][c] = val(po + S.off[s], fl, c);
    }
#pragma unroll
    for (int s = 0; s < NS; s++)
        if (S.pos[s] >= 0)
#pragma unroll
            for (int c = 0; c < NC; c++) sh[c * NP + S.pos[s]] = v[s][c];
}

// rhoEqn.H:19-33: drho = dt * (-div(rho*U)) (src/ex/div/div.hxx:22-81),
// rc += w drho, rho = rho0 + cs drho (the last stage: rho = rc), and on
// the last stage p = pc rho^2 (pEqn.H:2).  The j- and k-fluxes rho U_y and
// rho U_z of the plane are staged in LDS, the x-flux of the own column
// walks a register queue.
template <int H, bool WALLS>
__global__ void __launch_bounds__(kWcTK *kWcTJ) k_wc_rho(Grid g, WcRhoArgs a)
{
    constexpr int LK = kWcTK + 2 * H, LJ = kWcTJ + 2 * H;
    __shared__ double sh[2 * LJ * LK];
    const WcTile tl = wc_tile(g, a.ichunk, a.nkt, a.njt, a.nic);
    const int tk = threadIdx.x, tj = threadIdx.y;
    WcSlots<H> slots;
    wc_slots<H>(g, tl, slots);
    int f0;
    const int km = bmap(tl.k, g.n[2], g.wall[2], f0);
    auto flux = [&](int64_t off, const int (&fl)[3], int d) -> double {
        return a.rho_in[off] * wc_vec<WALLS>(g, a.U, d, off, fl);
    };
    double q[2 * H + 1];  // x-flux of the own column, planes i - H .. i + H
#pragma unroll
    for (int o = 0; o < 2 * H; o++) {
        int fl[3];
        q[o] = flux(wc_node<WALLS>(g, tl.i0 - H + o, tl.j, tl.k, fl), fl, 0);
    }
    const double zi = g.zinv[km];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    auto S = [&](int c, int r, int cc) -> double { return sh[(c * LJ + r) * LK + cc]; };
    for (int i = tl.i0; i < tl.i1; i++) {
        {
            int fl[3];
            q[2 * H] = flux(wc_node<WALLS>(g, i + H, tl.j, tl.k, fl), fl, 0);
        }
        wc_stage_plane<H, 2>(g, slots, i, sh,
                             [&](int64_t off, const int (&fl)[3], int c) { return flux(off, fl, c + 1); });
        __syncthreads();
        if (tl.valid) {
            double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
            for (int l = 0; l < H; l++) {
                ax += g.a[0][l] * (q[H + 1 + l] - q[H - 1 - l]);
                ay += g.a[1][l] * (S(0, tj + H + 1 + l, tk + H) - S(0, tj + H - 1 - l, tk + H));
                az += g.a[2][l] * (S(1, tj + H, tk + H + 1 + l) - S(1, tj + H, tk + H - 1 - l));
            }
            const double dr = a.dt * (-((ax + ay) + az * zi));
            const int64_t qi = (int64_t)i * plane + (int64_t)tl.j * g.s[2] + tl.k;
            const double rc = a.rc_in[qi] + a.w * dr;
            if (a.last) {
                a.rho_out[qi] = rc;
                if (a.p) a.p[qi] = a.p_left ? (a.pc * rc) * rc : a.pc * (rc * rc);
            } else {
                a.rho_out[qi] = a.rho0[qi] + a.cs * dr;
                a.rc_out[qi] = rc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < 2 * H; o++) q[o] = q[o + 1];
    }
}

// UEqn.H:16-33: drhoU = dt * ((-div(rho*U, U) - grad(p)) + laplacian(mu, U))
// (div.hxx:84-172, grad.cpp:25-62, laplacian.hxx:130-191), ruc += w drhoU,
// U = rhoUOld/rho + cs (drhoU/rho) (the last stage: U = ruc/rho), with
// rhoUOld = rho*U0.  rho, U and p of the plane are staged in LDS; the own
// column's i-neighbours walk register queues.
template <int H, bool WALLS>
__global__ void __launch_bounds__(kWcTK *kWcTJ) k_wc_u(Grid g, WcUArgs a)
{
    constexpr int LK = kWcTK + 2 * H, LJ = kWcTJ + 2 * H;
    constexpr int QW = 2 * H + 1;
    __shared__ double sh[5 * LJ * LK];  // rho, U_x, U_y, U_z, p
    const WcTile tl = wc_tile(g, a.ichunk, a.nkt, a.njt, a.nic);
    const int tk = threadIdx.x, tj = threadIdx.y;
    WcSlots<H> slots;
    wc_slots<H>(g, tl, slots);
    int f0;
    const int km = bmap(tl.k, g.n[2], g.wall[2], f0);
    // queues of the own column, planes i - H .. i + H: rho, U, p
    double qr[QW], qu[3][QW], qp[QW];
    auto load_plane = [&](int o, int ii) {
        int fl[3];
        const int64_t off = wc_node<WALLS>(g, ii, tl.j, tl.k, fl);
        qr[o] = a.rho[off];
        qp[o] = a.p[off];
#pragma unroll
        for (int c = 0; c < 3; c++) qu[c][o] = wc_vec<WALLS>(g, a.U_in, c, off, fl);
    };
#pragma unroll
    for (int o = 0; o < 2 * H; o++) load_plane(o, tl.i0 - H + o);
    const double zi = g.zinv[km], zl = g.zlap[km], zm = g.zmet[km];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    auto S = [&](int c, int r, int cc) -> double { return sh[(c * LJ + r) * LK + cc]; };
    for (int i = tl.i0; i < tl.i1; i++) {
        load_plane(2 * H, i + H);
        wc_stage_plane<H, 5>(g, slots, i, sh, [&](int64_t off, const int (&fl)[3], int c) -> double {
            if (c == 0) return a.rho[off];
            if (c == 4) return a.p[off];
            return wc_vec<WALLS>(g, a.U_in, c - 1, off, fl);
        });
        __syncthreads();
        if (tl.valid) {
            // field f at the node displaced by o along d: 0 = rho*U, 1 = U, 2 = p
            auto A = [&](int f, int c, int d, int o) -> double {
                if (d == 0 || o == 0) {
                    const int e = H + (d == 0 ? o : 0);
                    return f == 0 ? qr[e] * qu[c][e] : (f == 1 ? qu[c][e] : qp[e]);
                }
                const int r = tj + H + (d == 1 ? o : 0), cc = tk + H + (d == 2 ? o : 0);
                if (f == 2) return S(4, r, cc);
                const double u = S(1 + c, r, cc);
                return f == 0 ? S(0, r, cc) * u : u;
            };
            const int64_t qi = (int64_t)i * plane + (int64_t)tl.j * g.s[2] + tl.k;
            const double r = qr[H];
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
                const double u0 = a.U0 ? a.U0[c * g.np + qi] : qu[c][H];
                const double ruold = r * u0;
                const double ruc = (a.ruc_in ? a.ruc_in[c * g.np + qi] : ruold) + a.w * du;
                // the wall-normal component is odd: zero on its walls (the
                // refill after each stage, UEqn.H:33, with a reflecting wall)
                bool on_wall = false;
                if constexpr (WALLS) {
                    const int idx = c == 0 ? i : (c == 1 ? tl.j : tl.k);
                    on_wall = g.wall[c] && (idx == 0 || idx == g.n[c]);
                }
                if (a.last) {
                    a.U_out[c * g.np + qi] = on_wall ? 0.0 : ruc / r;
                } else {
                    a.U_out[c * g.np + qi] = on_wall ? 0.0 : ruold / r + a.cs * (du / r);
                    a.ruc_out[c * g.np + qi] = ruc;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < 2 * H; o++) {
            qr[o] = qr[o + 1];
            qp[o] = qp[o + 1];
#pragma unroll
            for (int c = 0; c < 3; c++) qu[c][o] = qu[c][o + 1];
        }
    }
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
    int ichunk, nkt, njt, nic;
};

// TEqn.H (app/shearLayerWC:16-33): drhoT = dt * (-div(rho*U, T) +
// laplacian(ka, T)), rtc += w drhoT, T = rhoTOld/rho + cs (drhoT/rho) (the
// last stage: T = rtc/rho), with rhoTOld = rho*T0.  rho, U and T of the
// plane are staged in LDS; the own column's i-neighbours walk register
// queues.
template <int H, bool WALLS>
__global__ void __launch_bounds__(kWcTK *kWcTJ) k_wc_t(Grid g, WcTArgs a)
{
    constexpr int LK = kWcTK + 2 * H, LJ = kWcTJ + 2 * H;
    constexpr int QW = 2 * H + 1;
    __shared__ double sh[5 * LJ * LK];  // rho, U_x, U_y, U_z, T
    const WcTile tl = wc_tile(g, a.ichunk, a.nkt, a.njt, a.nic);
    const int tk = threadIdx.x, tj = threadIdx.y;
    WcSlots<H> slots;
    wc_slots<H>(g, tl, slots);
    int f0;
    const int km = bmap(tl.k, g.n[2], g.wall[2], f0);
    double qr[QW], qu[3][QW], qt[QW];
    auto load_plane = [&](int o, int ii) {
        int fl[3];
        const int64_t off = wc_node<WALLS>(g, ii, tl.j, tl.k, fl);
        qr[o] = a.rho[off];
        qt[o] = a.T_in[off];
#pragma unroll
        for (int c = 0; c < 3; c++) qu[c][o] = wc_vec<WALLS>(g, a.U, c, off, fl);
    };
#pragma unroll
    for (int o = 0; o < 2 * H; o++) load_plane(o, tl.i0 - H + o);
    const double zi = g.zinv[km], zl = g.zlap[km], zm = g.zmet[km];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    auto S = [&](int c, int r, int cc) -> double { return sh[(c * LJ + r) * LK + cc]; };
    for (int i = tl.i0; i < tl.i1; i++) {
        load_plane(2 * H, i + H);
        wc_stage_plane<H, 5>(g, slots, i, sh, [&](int64_t off, const int (&fl)[3], int c) -> double {
            if (c == 0) return a.rho[off];
            if (c == 4) return a.T_in[off];
            return wc_vec<WALLS>(g, a.U, c - 1, off, fl);
        });
        __syncthreads();
        if (tl.valid) {
            // field f at the node displaced by o along d: 0 = rho*U (component
            // c), 1 = T
            auto A = [&](int f, int c, int d, int o) -> double {
                if (d == 0 || o == 0) {
                    const int e = H + (d == 0 ? o : 0);
                    return f == 0 ? qr[e] * qu[c][e] : qt[e];
                }
                const int r = tj + H + (d == 1 ? o : 0), cc = tk + H + (d == 2 ? o : 0);
                return f == 0 ? S(0, r, cc) * S(1 + c, r, cc) : S(4, r, cc);
            };
            const int64_t qi = (int64_t)i * plane + (int64_t)tl.j * g.s[2] + tl.k;
            const double r = qr[H];
            const double dT = a.dt * ((-div_xy<H>(g, A, 0, 1, 0, zi)) + lap_c<H>(g, A, 1, 0, a.ka, zl, zm));
            const double t0 = a.T0 ? a.T0[qi] : qt[H];
            const double rtold = r * t0;
            const double rtc = (a.rtc_in ? a.rtc_in[qi] : rtold) + a.w * dT;
            if (a.last) {
                a.T_out[qi] = rtc / r;
            } else {
                a.T_out[qi] = rtold / r + a.cs * (dT / r);
                a.rtc_out[qi] = rtc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < 2 * H; o++) {
            qr[o] = qr[o + 1];
            qt[o] = qt[o + 1];
#pragma unroll
            for (int c = 0; c < 3; c++) qu[c][o] = qu[c][o + 1];
        }
    }
}

struct WcLArgs {
    const double *B;
    double *F;
    double dbu;
    int ichunk, nkt, njt, nic;
};

// app/OT/UEqn.H:22-23 for a frozen B: F = -div((2 dbu) B, B) + grad(dbu (B & B))
// (div.hxx:84-172, grad.cpp:25-62), with B staged in LDS and queued along i.
template <int H, bool WALLS>
__global__ void __launch_bounds__(kWcTK *kWcTJ) k_wc_lorentz(Grid g, WcLArgs a)
{
    constexpr int LK = kWcTK + 2 * H, LJ = kWcTJ + 2 * H;
    constexpr int QW = 2 * H + 1;
    __shared__ double sh[3 * LJ * LK];
    const WcTile tl = wc_tile(g, a.ichunk, a.nkt, a.njt, a.nic);
    const int tk = threadIdx.x, tj = threadIdx.y;
    WcSlots<H> slots;
    wc_slots<H>(g, tl, slots);
    int f0;
    const int km = bmap(tl.k, g.n[2], g.wall[2], f0);
    double qb[3][QW];
    auto load_plane = [&](int o, int ii) {
        int fl[3];
        const int64_t off = wc_node<WALLS>(g, ii, tl.j, tl.k, fl);
#pragma unroll
        for (int c = 0; c < 3; c++) qb[c][o] = wc_vec<WALLS>(g, a.B, c, off, fl);
    };
#pragma unroll
    for (int o = 0; o < 2 * H; o++) load_plane(o, tl.i0 - H + o);
    const double zi = g.zinv[km];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    auto S = [&](int c, int r, int cc) -> double { return sh[(c * LJ + r) * LK + cc]; };
    for (int i = tl.i0; i < tl.i1; i++) {
        load_plane(2 * H, i + H);
        wc_stage_plane<H, 3>(g, slots, i, sh,
                             [&](int64_t off, const int (&fl)[3], int c) { return wc_vec<WALLS>(g, a.B, c, off, fl); });
        __syncthreads();
        if (tl.valid) {
            auto Bv = [&](int c, int d, int o) -> double {
                if (d == 0 || o == 0) return qb[c][H + (d == 0 ? o : 0)];
                return S(c, tj + H + (d == 1 ? o : 0), tk + H + (d == 2 ? o : 0));
            };
            // field 0 = (2 dbu) B, 1 = B
            auto A = [&](int f, int c, int d, int o) -> double {
                const double b = Bv(c, d, o);
                return f == 0 ? b * (2 * a.dbu) : b;
            };
            auto pm = [&](int d, int o) -> double {
                const double x = Bv(0, d, o), y = Bv(1, d, o), z = Bv(2, d, o);
                return (x * x + y * y + z * z) * a.dbu;
            };
            const int64_t qi = (int64_t)i * plane + (int64_t)tl.j * g.s[2] + tl.k;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double gs = 0.0;
#pragma unroll
                for (int l = 0; l < H; l++) gs += g.a[c][l] * (pm(c, l + 1) - pm(c, -l - 1));
                if (c == 2) gs *= zi;
                a.F[c * g.np + qi] = (-div_xy<H>(g, A, 0, 1, c, zi)) + gs;
            }
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < 2 * H; o++)
#pragma unroll
            for (int c = 0; c < 3; c++) qb[c][o] = qb[c][o + 1];
    }
}

template <class Args>
void wc_shape(const Grid &g, Args &a)
{
    a.nkt = (g.s[2] + kWcTK - 1) / kWcTK;
    a.njt = (g.s[1] + kWcTJ - 1) / kWcTJ;
    // about 2048 blocks, chunks of 8..64 planes
    int64_t per = (int64_t)g.s[0] * a.nkt * a.njt / 2048;
    int c = (int)(per < 8 ? 8 : (per > 64 ? 64 : per));
    a.ichunk = c < g.s[0] ? c : g.s[0];
    a.nic = (g.s[0] + a.ichunk - 1) / a.ichunk;
}

template <int H>
int launch_wc_h(nova_ctx *ctx, const WcRhoArgs *ra, const WcUArgs *ua, const WcLArgs *la, const WcTArgs *tt)
{
    const Grid &g = ctx->g;
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    if (tt) {
        TimedScope ts(ctx, "wc_t_stage");
        WcTArgs a = *tt;
        wc_shape(g, a);
        const dim3 grid(a.nkt * a.njt * a.nic), block(kWcTK, kWcTJ, 1);
        if (walls) hipLaunchKernelGGL((k_wc_t<H, true>), grid, block, 0, ctx->stream, g, a);
        else hipLaunchKernelGGL((k_wc_t<H, false>), grid, block, 0, ctx->stream, g, a);
        return check_launch("wc_t_stage");
    }
    if (la) {
        TimedScope ts(ctx, "wc_lorentz");
        WcLArgs a = *la;
        wc_shape(g, a);
        const dim3 grid(a.nkt * a.njt * a.nic), block(kWcTK, kWcTJ, 1);
        if (walls) hipLaunchKernelGGL((k_wc_lorentz<H, true>), grid, block, 0, ctx->stream, g, a);
        else hipLaunchKernelGGL((k_wc_lorentz<H, false>), grid, block, 0, ctx->stream, g, a);
        return check_launch("wc_lorentz");
    }
    if (ra) {
        TimedScope ts(ctx, "wc_rho_stage");
        WcRhoArgs a = *ra;
        wc_shape(g, a);
        const dim3 grid(a.nkt * a.njt * a.nic), block(kWcTK, kWcTJ, 1);
        if (walls) hipLaunchKernelGGL((k_wc_rho<H, true>), grid, block, 0, ctx->stream, g, a);
        else hipLaunchKernelGGL((k_wc_rho<H, false>), grid, block, 0, ctx->stream, g, a);
        return check_launch("wc_rho_stage");
    }
    TimedScope ts(ctx, "wc_u_stage");
    WcUArgs a = *ua;
    wc_shape(g, a);
    const dim3 grid(a.nkt * a.njt * a.nic), block(kWcTK, kWcTJ, 1);
    if (walls) hipLaunchKernelGGL((k_wc_u<H, true>), grid, block, 0, ctx->stream, g, a);
    else hipLaunchKernelGGL((k_wc_u<H, false>), grid, block, 0, ctx->stream, g, a);
    return check_launch("wc_u_stage");
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
