// nova_stream.hpp -- the streamed (LDS plane + register queue) stage
// kernel shared by the forward stages (Navier-Stokes, MHD) and the
// tangent-linear and adjoint stages.
#pragma once
#include "nova_internal.hpp"
#include "nova_stencil.hpp"

namespace nova {
namespace {

// MODEL_MHDU / MODEL_MHDB: the two halves of the reference's MHD step
// (app/OT_Poisson/UEqn.H and BEqn.H), each reading both fields' stencils
// and writing one field
enum { MODEL_NS = 0, MODEL_MHD = 1, MODEL_TLM = 2, MODEL_ADJ = 3, MODEL_MHDU = 4, MODEL_MHDB = 5 };
template <int MODEL>
constexpr int model_nf() { return MODEL == MODEL_NS ? 1 : 2; }
// whether the stage combination of the model reads a step-start field f
template <int MODEL>
constexpr bool model_x0(int f)
{
    return MODEL == MODEL_MHD || (MODEL == MODEL_NS && f == 0) || ((MODEL == MODEL_TLM || MODEL == MODEL_MHDB) && f == 1);
}

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
    NsK nk;   // NS stages: ns_k(g, nu, sa.dt)
    MhdK mk;  // MHD stages: mhd_k(g, nu, eta, sa.dt)
    AdjK ak;  // adjoint stages: adj_k(g, nu, sa.dt)
    // adjoint stage: out = ob T (combination 1) or T + yg Y + zg Z
    // (combination 2), T = m + dt J_R(X)^T m; MHD B stages: the accumulator
    const double *Y, *Z;
    double *acc;
    double ob, yg, zg, acc_a, acc_b;
    int ichunk;
    int nkt, njt, nic;
};

// MODEL_NS: X = {U}; MODEL_MHD: X = {U, B};  out = stage combination.
// MODEL_TLM: X = {U, d}, X0[1] = step-start tangent; out[0] = tangent stage.
// MODEL_ADJ: X = {U, m}; T = m + dt J_R(U)^T m; out[0] = ob T (CMB 1) or
//   T + yg Y + zg Z (CMB 2, Y and Z read at the node itself; Z may be
//   out[0]: each node is read and then written by the same lane).
// MODEL_MHDU: X = {U, B}; out[0] = U + sum_j w_j dU, dU = dt R_U(U, B)
//   (sa.form FORM_LIT4).
// MODEL_MHDB: X = {U, B}, X0[1] = B at the step start, acc = the running
//   RK4 sum; dB = dt R_B(U, B); s = (acc_a != 0 ? acc : X0[1]) + w_0 dB;
//   alpha != 0 (last stage): out[0] = s; else acc = s and
//   out[0] = X0[1] + beta dB.
// Staging of the (j,k) planes in LDS: STG_ONE holds one plane (two barriers
// per plane), STG_TWO two alternating planes (one barrier), STG_RING a ring
// of H + 2 planes whose halos are loaded H planes ahead (one barrier).
constexpr int STG_ONE = 0, STG_TWO = 1, STG_RING = 2;
template <int H, int STG>
constexpr int stage_slots()
{
    return STG == STG_ONE ? 1 : (STG == STG_TWO ? 2 : H + 2);
}
// LDS doubles of one tile's staged planes
template <int H, int MODEL, int kTK, int kTJ = 256 / kTK, int STG = STG_ONE>
constexpr int stage_lds()
{
    return stage_slots<H, STG>() * 3 * model_nf<MODEL>() * (kTJ + 2 * H) * (kTK + 2 * H);
}

// A copy of x in a register of its own.  With two prefetch sets the value
// taken into the queue must not stay in the register its load wrote: the
// set is loaded again one plane later, and a queue entry sharing that
// register makes the compiler move the new load's result at the loop
// latch, which waits for the load to land.
template <int PF>
__device__ __attribute__((always_inline)) inline double fresh(double x)
{
    if constexpr (PF == 1) {
        return x;
    } else {
        double r;
        asm("v_mov_b64 %0, %1" : "=v"(r) : "v"(x));
        return r;
    }
}

// One tile t = (kt, jt, ic) of a stage: the block streams planes i of its
// chunk for the (k, j) tile.  shbuf holds stage_lds() doubles of LDS.
// Stage combination of a launch: CMB 0 reads the form at run time; the
// SSP forms are compiled in where the launch knows them (1: alpha X0 +
// beta (X + dX), step-start fields present; 2: X + dX, none read).
// STG_RING: the halo of plane p is loaded together with the queue entry
// (the interior) of plane p, H planes before plane p is computed, and kept
// in the ring of staged planes: a neighbouring tile loads those nodes as its
// interior at about the same time, so the halo comes from L2 instead of
// being fetched again H planes later, after the streaming traffic has
// evicted it.  With two or more slots, the slot a plane is written to was
// last read two planes earlier, before the barrier of the previous plane,
// so one barrier per plane suffices.
template <int H, int MODEL, bool WALLS, int kTK, int PF = 1, int kTJ = 256 / kTK, int CMB = 0, int STG = STG_ONE>
__device__ __attribute__((always_inline)) inline void stage_tile(const Grid &g, const StreamArgs<model_nf<MODEL>()> &a,
                                                                 int t, double *shbuf)
{
    constexpr int NF = model_nf<MODEL>();
    constexpr int QW = 2 * H + 1;
    constexpr int NC = 3 * NF;
    constexpr int LK = kTK + 2 * H, LJ = kTJ + 2 * H;
    constexpr int NHALO = 2 * H * (kTK + kTJ);
    constexpr int NSLOT = (NHALO + kTK * kTJ - 1) / (kTK * kTJ);
    // staged plane slots (plane i in slot (i - i0) mod NS)
    constexpr bool RING = STG == STG_RING;
    constexpr int NS = stage_slots<H, STG>();
    constexpr int RS = NC * LJ * LK;
    auto shr = [&](int slot, int n, int r, int c) -> double & { return shbuf[slot * RS + (n * LJ + r) * LK + c]; };
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
    // results are stored non-temporally: a stage writes its whole output
    // once (hundreds of MB, more than the caches hold), and streaming it past
    // L2 leaves the cache to the halo planes that neighbouring tiles re-read
    auto stp = [&](double *b, int c, int64_t po, uint32_t o, double v) {
        __builtin_nontemporal_store(v, (double *)((char *)(b + c * g.np + po) + (uint32_t)(o * 8u)));
    };

    // own column
    int fj, fk;
    const int jm = bmap(j, g.n[1], g.wall[1], fj);
    const int km = bmap(k, g.n[2], g.wall[2], fk);
    const uint32_t col = (uint32_t)jm * g.s[2] + km;
    // reflection masks: bit c set when component c changes sign on the way
    // the value was loaded (the i plane's reflection is added per plane)
    const unsigned mcol = (fj ? g.odd[1] : 0u) ^ (fk ? g.odd[2] : 0u);
    const unsigned modd_i = g.odd[0];

    // halo slots of this thread: LDS position, source column, signs
    int hrow[NSLOT], hcolk[NSLOT];
    uint32_t hsrc[NSLOT];
    unsigned hmk[NSLOT];
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
        hmk[s] = (f1 ? g.odd[1] : 0u) ^ (f2 ? g.odd[2] : 0u);
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
                q[f][c][o] = rneg<WALLS>(ldp(a.X[f], c, po, col), mcol ^ (fi ? modd_i : 0u), c);
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
        double ny[3], nacc[3];  // adjoint combination 2: Y and Z; MHD B stages: the accumulator
        int fq, fh;             // reflection flags of the queue plane and the halo plane
    };
    constexpr bool ADJ = MODEL == MODEL_ADJ;
    static_assert(!ADJ || CMB == 1 || CMB == 2, "adjoint stages take a compiled-in combination");
    // reads pointwise addends (adjoint combination 2) or the accumulator (MHD B stages)
    constexpr bool ACC = (ADJ && CMB == 2) || MODEL == MODEL_MHDB;
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
                    P.nx0[f][c] = (model_x0<MODEL>(f) && (CMB == 1 || (CMB == 0 && a.X0[f])))
                                      ? ldp(a.X0[f], c, (int64_t)p * plane, col)
                                      : 0.0;
            if constexpr (ACC) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    if constexpr (ADJ) {
                        P.ny[c] = ldp(a.Y, c, (int64_t)p * plane, col);
                        P.nacc[c] = ldp(a.Z, c, (int64_t)p * plane, col);
                    } else {
                        P.nacc[c] = a.acc_a != 0.0 ? ldp(a.acc, c, (int64_t)p * plane, col) : 0.0;
                    }
                }
            }
            const int64_t po = plane_off(p + H, P.fq);
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) P.nq[f][c] = ldp(a.X[f], c, po, col);
        }
        const int64_t ph = plane_off(RING ? p + H : p, P.fh);
#pragma unroll
        for (int s = 0; s < NSLOT; s++)
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) P.hv[s][f * 3 + c] = ldp(a.X[f], c, ph, hsrc[s]);
    };

    const double zi = g.zinv[km], zl = g.zlap[km], zm = g.zmet[km];
    NsZ<H> nz{};
    MhdZ mz{};
    AdjZ<H> az{};
    if constexpr (MODEL == MODEL_NS) nz = ns_z<H>(a.nk, zi, zl, zm);
    if constexpr (ADJ) az = adj_z<H>(a.ak, zi, zl);
    if constexpr (MODEL == MODEL_MHDU || MODEL == MODEL_MHDB) mz = mhd_z(a.mk, zi, zl, zm);
    // step-start values of the own column, one plane ahead like the queue
    double x0c[NF][3];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            x0c[f][c] = (model_x0<MODEL>(f) && (CMB == 1 || (CMB == 0 && a.X0[f])))
                            ? ldp(a.X0[f], c, (int64_t)i0 * plane, col)
                            : 0.0;
    double yc[3] = {0.0, 0.0, 0.0}, accc[3] = {0.0, 0.0, 0.0};
    if constexpr (ACC) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if constexpr (ADJ) {
                yc[c] = ldp(a.Y, c, (int64_t)i0 * plane, col);
                accc[c] = ldp(a.Z, c, (int64_t)i0 * plane, col);
            } else {
                accc[c] = a.acc_a != 0.0 ? ldp(a.acc, c, (int64_t)i0 * plane, col) : 0.0;
            }
        }
    }

    // One plane: stage plane i (halo from Cur) in LDS, issue the loads of
    // plane i + PF into the set they go to, compute plane i, then advance
    // the queue with Nxt (the data of plane i + 1).  PF = 1: Cur and Nxt
    // are one set; PF = 2: two sets alternate, so two planes of loads are
    // in flight while a plane is computed.
    auto body = [&](int i, Pf &Cur, Pf &Nxt) {
        if constexpr (RING) {
            // plane i + H into its slot (it held plane i - 2)
            const int sw = (i + H - i0) % NS;
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) shr(sw, f * 3 + c, tj + H, tk + H) = q[f][c][2 * H];
#pragma unroll
            for (int s = 0; s < NSLOT; s++)
                if (hon[s])
#pragma unroll
                    for (int n = 0; n < NC; n++)
                        shr(sw, n, hrow[s], hcolk[s]) = rneg<WALLS>(Cur.hv[s][n], hmk[s] ^ (Cur.fh ? modd_i : 0u), n % 3);
        } else {
            const int sw = NS == 1 ? 0 : (i - i0) & 1;
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) shr(sw, f * 3 + c, tj + H, tk + H) = q[f][c][H];
#pragma unroll
            for (int s = 0; s < NSLOT; s++)
                if (hon[s])
#pragma unroll
                    for (int n = 0; n < NC; n++)
                        shr(sw, n, hrow[s], hcolk[s]) = rneg<WALLS>(Cur.hv[s][n], hmk[s] ^ (Cur.fh ? modd_i : 0u), n % 3);
        }
        __syncthreads();
        load_pf(Cur, i + PF < i1 ? i + PF : i1 - 1, true);

        if (valid) {
            const int sr = NS == 1 ? 0 : (RING ? (i - i0) % NS : (i - i0) & 1);
            auto A = [&](int f, int c, int d, int o) -> double {
                if (d == 0) return q[f][c][H + o];
                if (d == 1) return shr(sr, f * 3 + c, tj + H + o, tk + H);
                return shr(sr, f * 3 + c, tj + H, tk + H + o);
            };
            const int64_t pi = (int64_t)i * plane;
            const StageArgs &sa = a.sa;
            if constexpr (MODEL == MODEL_NS) {
                double r[3];
                ns_du<H>(a.nk, nz, A, r);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double du = r[c];
                    const double u = A(0, c, 0, 0);
                    if constexpr (CMB == 1) {
                        const double v = u + du;
                        stp(a.out[0], c, pi, col, sa.alpha * x0c[0][c] + sa.beta * v);
                    } else if constexpr (CMB == 2) {
                        stp(a.out[0], c, pi, col, u + du);
                    } else {
                        const double u0 = a.X0[0] ? x0c[0][c] : u;
                        stp(a.out[0], c, pi, col, combine(sa, u, u0, du));
                    }
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
                adj_du<H>(a.ak, az, A, r);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double t = q[1][c][H] + r[c];
                    if constexpr (CMB == 1) stp(a.out[0], c, pi, col, a.ob * t);
                    else stp(a.out[0], c, pi, col, (t + a.yg * yc[c]) + a.zg * accc[c]);
                }
            } else if constexpr (MODEL == MODEL_MHDU) {
                double du[3], db[3];
                mhd_dudb<H, true, false>(g, a.mk, mz, A, du, db);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double u = A(0, c, 0, 0);
                    stp(a.out[0], c, pi, col, combine(sa, u, u, du[c]));
                }
            } else if constexpr (MODEL == MODEL_MHDB) {
                double du[3], dbv[3];
                mhd_dudb<H, false, true>(g, a.mk, mz, A, du, dbv);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double db = dbv[c];
                    const double s = (a.acc_a != 0.0 ? accc[c] : x0c[1][c]) + sa.w[0] * db;
                    if (sa.alpha != 0.0) {
                        stp(a.out[0], c, pi, col, s);
                    } else {
                        stp(a.acc, c, pi, col, s);
                        stp(a.out[0], c, pi, col, x0c[1][c] + sa.beta * db);
                    }
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
        if constexpr (NS == 1) __syncthreads();
        {
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
#pragma unroll
                    for (int o = 0; o < 2 * H; o++) q[f][c][o] = q[f][c][o + 1];
                    q[f][c][2 * H] = fresh<PF>(rneg<WALLS>(Nxt.nq[f][c], mcol ^ (Nxt.fq ? modd_i : 0u), c));
                    x0c[f][c] = fresh<PF>(Nxt.nx0[f][c]);
                }
            if constexpr (ACC) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    if constexpr (ADJ) yc[c] = fresh<PF>(Nxt.ny[c]);
                    accc[c] = fresh<PF>(Nxt.nacc[c]);
                }
            }
        }
    };

    if constexpr (RING) {
        // planes i0 .. i0 + H - 1 into slots 0 .. H - 1: interiors from the
        // queue, halos loaded here (body i writes plane i + H)
#pragma unroll
        for (int pp = 0; pp < H; pp++) {
            int fh;
            const int64_t ph = plane_off(i0 + pp, fh);
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int c = 0; c < 3; c++) shr(pp, f * 3 + c, tj + H, tk + H) = q[f][c][H + pp];
#pragma unroll
            for (int s = 0; s < NSLOT; s++)
                if (hon[s])
#pragma unroll
                    for (int f = 0; f < NF; f++)
#pragma unroll
                        for (int c = 0; c < 3; c++)
                            shr(pp, f * 3 + c, hrow[s], hcolk[s]) =
                                rneg<WALLS>(ldp(a.X[f], c, ph, hsrc[s]), hmk[s] ^ (fh ? modd_i : 0u), c);
        }
    }
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

template <int H, int MODEL, bool WALLS, int MINW, int kTK, int kTJ, int CMB, int STG = STG_ONE>
__global__ void __launch_bounds__(kTK *kTJ) __attribute__((amdgpu_waves_per_eu(MINW)))
k_stage_stream(Grid g, StreamArgs<model_nf<MODEL>()> a)
{
    __shared__ double sh[stage_lds<H, MODEL, kTK, kTJ, STG>()];
    // consecutive tiles share halos, so each XCD (blocks are dealt
    // round-robin over 8) takes a contiguous run of tiles
    const int T = a.nkt * a.njt * a.nic;
    int t = blockIdx.x;
    if (T % 8 == 0) t = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
    // shapes held to one wave per SIMD, and one-field shapes held to two,
    // have the registers for a second prefetch set
    constexpr int PF = (MINW == 1 || (MINW == 2 && model_nf<MODEL>() == 1)) ? 2 : 1;
    stage_tile<H, MODEL, WALLS, kTK, PF, kTJ, CMB, STG>(g, a, t, sh);
}

// Launch-shape slot of a stage: the model, or slot 4 for the first SSP
// stage of the NS model (compiled-in combination 2)
template <int MODEL, int CMB>
constexpr int stage_slot()
{
    return MODEL == MODEL_NS && CMB == 2 ? 4 : (MODEL >= MODEL_MHDU ? MODEL_MHD : MODEL);
}

int pick_ichunk(const Grid &g, int nkt, int njt)
{
    // about 2048 tiles (8 blocks per CU on 256 CUs), chunks of 8..64 planes
    int64_t per = (int64_t)g.s[0] * nkt * njt / 2048;
    int c = (int)(per < 8 ? 8 : (per > 64 ? 64 : per));
    if (c > g.s[0]) c = g.s[0];
    return c;
}

template <int H, int MODEL, bool WALLS, int CMB, int MINW, int TK, int TJ = 256 / TK, int STG = STG_ONE>
void launch_stream_tile(nova_ctx *ctx, StreamArgs<model_nf<MODEL>()> a)
{
    const Grid &g = ctx->g;
    a.nkt = (g.s[2] + TK - 1) / TK;
    a.njt = (g.s[1] + TJ - 1) / TJ;
    const int ic = ctx->stage_ichunk[stage_slot<MODEL, CMB>()];
    a.ichunk = ic > 0 ? (ic < g.s[0] ? ic : g.s[0]) : pick_ichunk(g, a.nkt, a.njt);
    a.nic = (g.s[0] + a.ichunk - 1) / a.ichunk;
    const int T = a.nkt * a.njt * a.nic;
    hipLaunchKernelGGL((k_stage_stream<H, MODEL, WALLS, MINW, TK, TJ, CMB, STG>), dim3(T), dim3(TK, TJ, 1), 0, ctx->stream, g,
                       a);
}

template <int H, int MODEL, bool WALLS, int CMB = 0>
void launch_stream_w(nova_ctx *ctx, const StreamArgs<model_nf<MODEL>()> &a)
{
    // variants 2-4 hold one-field models to 3 waves per SIMD; two queued
    // fields need more than 256 registers and run unspilled at one wave per
    // SIMD (spilled shapes measure 3-4x slower)
    constexpr int W = model_nf<MODEL>() == 1 ? 3 : 1;
    switch (ctx->stage_variant[stage_slot<MODEL, CMB>()]) {
    case 0: launch_stream_tile<H, MODEL, WALLS, CMB, 1, 64>(ctx, a); break;
    case 2: launch_stream_tile<H, MODEL, WALLS, CMB, W, 64>(ctx, a); break;
    case 4: launch_stream_tile<H, MODEL, WALLS, CMB, W, 16>(ctx, a); break;
    // two waves per SIMD: one-field models with two planes of loads in
    // flight; two-field models with one prefetch set on 32 x 8 tiles (one
    // staged plane, or two alternating), which spills a few registers
    case 5:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, CMB, 2, 32>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, CMB, 2, 32, 8>(ctx, a);
        break;
    case 6:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, CMB, 2, 16>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, CMB, 2, 32, 8, STG_TWO>(ctx, a);
        break;
    // 512-thread tiles (32 x 16, 64 x 8): a smaller halo ring per node.  A
    // 512-thread block is held to 256 registers per lane, which two queued
    // fields exceed, so two-field models keep their 256-thread shapes here
    case 7:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, CMB, 2, 32, 16>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, CMB, 1, 32>(ctx, a);
        break;
    case 8:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, CMB, 2, 64, 8>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, CMB, 1, 64>(ctx, a);
        break;
    // the halo ring (the halo of a plane loaded with its interior, H planes
    // ahead): 512-thread 64 x 8 and 32 x 16 tiles (one-field models; a ring
    // that does not fit the 160 KiB of LDS alternates two planes instead)
    case 9:
        if constexpr (W == 3) {
            constexpr int S = stage_lds<H, MODEL, 64, 8, STG_RING>() * 8 <= 160 * 1024 ? STG_RING : STG_TWO;
            launch_stream_tile<H, MODEL, WALLS, CMB, 2, 64, 8, S>(ctx, a);
        } else launch_stream_tile<H, MODEL, WALLS, CMB, 1, 64>(ctx, a);
        break;
    case 10:
        if constexpr (W == 3) {
            constexpr int S = stage_lds<H, MODEL, 32, 16, STG_RING>() * 8 <= 160 * 1024 ? STG_RING : STG_TWO;
            launch_stream_tile<H, MODEL, WALLS, CMB, 2, 32, 16, S>(ctx, a);
        } else launch_stream_tile<H, MODEL, WALLS, CMB, 1, 32>(ctx, a);
        break;
    // two alternating staged planes (one barrier per plane): 512-thread
    // 64 x 8 and 32 x 16 tiles, 256-thread 64 x 4 tiles for two-field models
    case 11:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, CMB, 2, 64, 8, STG_TWO>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, CMB, 1, 64, 4, STG_TWO>(ctx, a);
        break;
    case 12:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, CMB, 2, 32, 16, STG_TWO>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, CMB, 1, 32, 8, STG_TWO>(ctx, a);
        break;
    // two-field models: blocks of one wave (16 x 4) and two waves (32 x 4),
    // two alternating staged planes, so that a CU holds several independent
    // blocks at one wave per SIMD instead of one block whose four waves meet
    // at every plane's barrier
    case 13:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, CMB, W, 32>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, CMB, 1, 16, 4, STG_TWO>(ctx, a);
        break;
    case 14:
        if constexpr (W == 3) launch_stream_tile<H, MODEL, WALLS, CMB, W, 32>(ctx, a);
        else launch_stream_tile<H, MODEL, WALLS, CMB, 1, 32, 4, STG_TWO>(ctx, a);
        break;
    default: launch_stream_tile<H, MODEL, WALLS, CMB, W, 32>(ctx, a); break;
    }
}

// Streamed launch for any model and stencil half-width (the row-stream
// variant of the forward stages is dispatched by their own launcher).
template <int MODEL, int CMB = 0>
int launch_stream_model(nova_ctx *ctx, const StreamArgs<model_nf<MODEL>()> &a, const char *name)
{
    const Grid &g = ctx->g;
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    TimedScope ts(ctx, name);
    switch (g.h * 2 + (walls ? 1 : 0)) {
    case 2: launch_stream_w<1, MODEL, false, CMB>(ctx, a); break;
    case 3: launch_stream_w<1, MODEL, true, CMB>(ctx, a); break;
    case 4: launch_stream_w<2, MODEL, false, CMB>(ctx, a); break;
    case 5: launch_stream_w<2, MODEL, true, CMB>(ctx, a); break;
    case 6: launch_stream_w<3, MODEL, false, CMB>(ctx, a); break;
    case 7: launch_stream_w<3, MODEL, true, CMB>(ctx, a); break;
    case 8: launch_stream_w<4, MODEL, false, CMB>(ctx, a); break;
    case 9: launch_stream_w<4, MODEL, true, CMB>(ctx, a); break;
    default: return fail(NOVA_ERR_ARG, "stencil order");
    }
    return check_launch(name);
}

}  // namespace
}  // namespace nova
