// nova_pipe.hpp -- the plane-streaming pipeline of the per-operator kernels
// (nova_ops.hip) and the weakly-compressible stages (nova_wc.hip).
//
// A 64 x KJ block owns a (k, j) tile and walks i over a chunk of planes.
// The inputs are NS components, each a plain stored field read through the
// boundary index maps (with its reflection signs at walls).  Components in
// the queue mask QM keep the i-neighbours of the block's own column in a
// register queue of planes i-H .. i+H; components in the LDS mask LM are
// staged plane by plane in LDS with their j and k halos (H nodes each side,
// no corners).  The loads of the next planes are in flight while a plane is
// computed: each plane's queue entry and halo are loaded up to two planes
// ahead, kept as loaded, and given their signs when they are staged.  Each
// input value is read from global memory about once (plus the halo ring and
// the 2H warm-up planes of a chunk), and every load addresses a column
// computed once per thread plus a plane offset computed once per plane.
#pragma once
#include "nova_internal.hpp"

namespace nova {
namespace pipe {

constexpr int kK = 64;
// threads of a block with KJ tile rows
constexpr int threads(int KJ) { return kK * KJ; }

constexpr int popc(unsigned m) { return m ? (int)(m & 1u) + popc(m >> 1) : 0; }
// position of component s among the components of mask m
constexpr int rank(unsigned m, int s) { return popc(m & ((1u << s) - 1u)); }

template <int NS>
struct Src {
    const double *f[NS];  // stored field of component s
    // bit d of flip[s]: component s changes sign when loaded through a
    // reflection across the walls normal to d (a vector component odd there)
    unsigned flip[NS];
};

struct Launch {
    int ichunk, nkt, njt, nic;
};

// the z metric of a thread's column (per stored k): 1/dzdxi, 1/dzdxi^2,
// d2zdxi2/dzdxi^3, loaded once per thread (per plane, the compiler could
// not hoist them past the result stores, which it cannot tell apart)
struct ZK {
    double zi, zl, zm;
};

// i chunks enough for about 2048 blocks of 256 threads (eight per CU), at
// least 8 planes (and 4H) long so that the 2H warm-up planes stay a small
// part of a chunk; KJ: rows of a tile (64 x KJ threads)
template <int H, int KJ = 4>
inline Launch tiling(const Grid &g)
{
    Launch L;
    L.nkt = (g.s[2] + kK - 1) / kK;
    L.njt = (g.s[1] + KJ - 1) / KJ;
    const int tiles = L.nkt * L.njt;
    const int nic = (2048 * 4 / KJ + tiles - 1) / tiles;
    const int minchunk = 4 * H > 8 ? 4 * H : 8;
    int ichunk = (g.s[0] + nic - 1) / nic;
    if (ichunk < minchunk) ichunk = minchunk;
    if (ichunk > g.s[0]) ichunk = g.s[0];
    L.ichunk = ichunk;
    L.nic = (g.s[0] + ichunk - 1) / ichunk;
    return L;
}
inline unsigned blocks(const Launch &L) { return (unsigned)(L.nkt * L.njt * L.nic); }

// Source entry s = component c of an nc-component field at f (signs as the
// boundary maps give them: a three-component field's component c is odd
// across the walls normal to d when bit c of g.odd[d] is set).
template <int NS>
inline void set_src(const Grid &g, Src<NS> &S, int s, const double *f, int nc, int c)
{
    S.f[s] = f + c * g.np;
    S.flip[s] = 0;
    if (nc == 3)
        for (int d = 0; d < 3; d++)
            if ((g.odd[d] >> c) & 1u) S.flip[s] |= 1u << d;
}

// The pipeline.  compute(i, j, k, q0, nb, zk) runs for every node of the
// block's chunk inside the grid (q0: the node's stored offset; zk: the z
// metric of the node's k); nb(s, d, o) is
// component s at the node displaced by o (-H .. H) along d: d = 0 needs s in
// QM, d = 1 or 2 with o != 0 needs s in LM, o = 0 either.  Blocks map to
// tiles so that each of the eight XCDs (given blocks in turn) holds a
// contiguous eighth of them: tiles that share halo rows share an L2.
// KJ: rows of the tile (64 x KJ threads).  PD: planes whose loads are in
// flight while a plane is computed (2: the queue entries and halos of the
// next two planes; 1: of the next plane; 0: none, a plane's loads are issued
// when it is staged), which trades registers against latency hiding.
template <int H, int NS, unsigned QM, unsigned LM, bool WALLS, int KJ = 4, int PD = 2, class Compute>
__device__ __attribute__((always_inline)) inline void run(const Grid &g, const Launch &L, const Src<NS> &S,
                                                          Compute compute)
{
    constexpr int kJ = KJ, kThreads = kK * KJ;
    static_assert(PD >= 0 && PD <= 2, "prefetch depth 0, 1 or 2");
    constexpr int NQ = popc(QM), NL = popc(LM);
    constexpr int QW = 2 * H + 1, LJ = kJ + 2 * H, LK = kK + 2 * H;
    constexpr int NHALO = 2 * H * (kK + kJ);
    constexpr int NSLOT = (NHALO + kThreads - 1) / kThreads;
    static_assert((QM | LM) == (1u << NS) - 1u, "every component is queued or staged");
    __shared__ double sh[NL > 0 ? NL : 1][LJ][LK];

    const int T = L.nkt * L.njt * L.nic;
    int bt = blockIdx.x;
    if (T % 8 == 0) bt = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
    const int kt = bt % L.nkt, jt = (bt / L.nkt) % L.njt, ic = bt / (L.nkt * L.njt);
    const int tk = threadIdx.x % kK, tj = threadIdx.x / kK;
    const int k = kt * kK + tk, j = jt * kJ + tj;
    const bool valid = k < g.s[2] && j < g.s[1];
    const int i0 = ic * L.ichunk;
    const int i1 = i0 + L.ichunk < g.s[0] ? i0 + L.ichunk : g.s[0];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];

    // own column and its reflection flags (bit 1: j, bit 2: k)
    int fj, fk;
    const int jm = bmap(j, g.n[1], g.wall[1], fj);
    const int km = bmap(k, g.n[2], g.wall[2], fk);
    const int64_t col = (int64_t)jm * g.s[2] + km;
    const unsigned cfl = (fj ? 2u : 0u) | (fk ? 4u : 0u);
    const ZK zk{g.zinv[km], g.zlap[km], g.zmet[km]};
    // halo slots of this thread: LDS row and column, source column, flags
    int hr[NSLOT], hc[NSLOT];
    int64_t hsrc[NSLOT];
    unsigned hfl[NSLOT];
    bool hon[NSLOT];
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++) {
        const int e = threadIdx.x + sl * kThreads;
        int jj = 0, kk = 0;
        hon[sl] = e < NHALO;
        hr[sl] = hc[sl] = 0;
        if (e < 2 * H * kK) {  // j halo rows
            const int rr = e / kK, cc = e % kK;
            hr[sl] = rr < H ? rr : kJ + rr;
            hc[sl] = cc + H;
            jj = jt * kJ + (rr < H ? rr - H : kJ + rr - H);
            kk = kt * kK + cc;
        } else if (e < NHALO) {  // k halo columns
            const int e2 = e - 2 * H * kK;
            const int rr = e2 / (2 * H), c2 = e2 % (2 * H);
            hr[sl] = rr + H;
            hc[sl] = c2 < H ? c2 : kK + c2;
            jj = jt * kJ + rr;
            kk = kt * kK + (c2 < H ? c2 - H : kK + c2 - H);
        }
        int f1, f2;
        const int jm2 = bmap(jj, g.n[1], g.wall[1], f1);
        const int km2 = bmap(kk, g.n[2], g.wall[2], f2);
        hsrc[sl] = (int64_t)jm2 * g.s[2] + km2;
        hfl[sl] = (f1 ? 2u : 0u) | (f2 ? 4u : 0u);
    }
    // value of component s loaded through the reflections in flags fl
    auto sgn = [&](int s, unsigned fl, double v) -> double {
        if constexpr (WALLS) return (__popc(fl & S.flip[s]) & 1) ? -v : v;
        else return v;
    };
    auto plane_of = [&](int ii, unsigned &fl) -> int64_t {
        int fi;
        const int im = bmap(ii, g.n[0], g.wall[0], fi);
        fl = fi ? 1u : 0u;
        return (int64_t)im * plane;
    };

    // loads of plane p: the queue entries of plane p + H and the halo of
    // plane p
    struct Pf {
        double nq[NQ > 0 ? NQ : 1];
        double hv[NSLOT][NL > 0 ? NL : 1];
        unsigned fq, fh;
    };
    auto load_q = [&](Pf &P, int p) {
        const int64_t po = plane_of(p + H, P.fq);
#pragma unroll
        for (int s = 0; s < NS; s++)
            if ((QM >> s) & 1u) P.nq[rank(QM, s)] = S.f[s][po + col];
    };
    auto load_pf = [&](Pf &P, int p, bool queue) {
        if (queue) load_q(P, p);
        const int64_t pp = plane_of(p, P.fh);
#pragma unroll
        for (int sl = 0; sl < NSLOT; sl++)
#pragma unroll
            for (int s = 0; s < NS; s++)
                if ((LM >> s) & 1u) P.hv[sl][rank(LM, s)] = hon[sl] ? S.f[s][pp + hsrc[sl]] : 0.0;
    };

    // queue, and the centre of the components kept only in LDS
    double q[NQ > 0 ? NQ : 1][QW];
    double cen[NS];
    {
#pragma unroll
        for (int o = 0; o < QW; o++) {
            unsigned fl;
            const int64_t po = plane_of(i0 - H + o, fl);
#pragma unroll
            for (int s = 0; s < NS; s++)
                if ((QM >> s) & 1u) q[rank(QM, s)][o] = sgn(s, fl | cfl, S.f[s][po + col]);
        }
    }
    auto nb = [&](int s, int d, int o) -> double {
        if (o == 0) return cen[s];
        if (d == 0) return q[rank(QM, s)][H + o];
        if (d == 1) return sh[rank(LM, s)][tj + H + o][tk + H];
        return sh[rank(LM, s)][tj + H][tk + H + o];
    };

    // the centre of an LDS-only component is loaded with the halo, at the
    // own column: one more value per plane in the prefetch
    struct Pc {
        double v[NS];
        unsigned f;
    };
    constexpr unsigned CM = LM & ~QM;  // components whose centre is not queued
    auto load_c = [&](Pc &C, int p) {
        if constexpr (CM != 0) {
            const int64_t pp = plane_of(p, C.f);
#pragma unroll
            for (int s = 0; s < NS; s++)
                if ((CM >> s) & 1u) C.v[s] = S.f[s][pp + col];
        }
    };

    auto stage = [&](Pf &Cur, Pc &Ccur) {
#pragma unroll
        for (int s = 0; s < NS; s++) {
            if ((QM >> s) & 1u) cen[s] = q[rank(QM, s)][H];
            else cen[s] = sgn(s, Ccur.f | cfl, Ccur.v[s]);
            if ((LM >> s) & 1u) sh[rank(LM, s)][tj + H][tk + H] = cen[s];
        }
#pragma unroll
        for (int sl = 0; sl < NSLOT; sl++)
            if (hon[sl])
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if ((LM >> s) & 1u)
                        sh[rank(LM, s)][hr[sl]][hc[sl]] = sgn(s, Cur.fh | hfl[sl], Cur.hv[sl][rank(LM, s)]);
    };
    auto shift = [&](const Pf &Nxt) {
#pragma unroll
        for (int s = 0; s < NQ; s++) {
#pragma unroll
            for (int o = 0; o < 2 * H; o++) q[s][o] = q[s][o + 1];
        }
#pragma unroll
        for (int s = 0; s < NS; s++)
            if ((QM >> s) & 1u) q[rank(QM, s)][2 * H] = sgn(s, Nxt.fq | cfl, Nxt.nq[rank(QM, s)]);
    };
    if constexpr (PD == 2) {
        auto body = [&](int i, Pf &Cur, Pf &Nxt, Pc &Ccur) {
            stage(Cur, Ccur);
            __syncthreads();
            if (i + 2 < i1) {
                load_pf(Cur, i + 2, true);
                load_c(Ccur, i + 2);
            }
            if (valid) compute(i, j, k, (int64_t)i * plane + col, nb, zk);
            __syncthreads();
            if (i + 1 < i1) shift(Nxt);
        };
        Pf A, B;
        Pc CA, CB;
        load_pf(A, i0, false);
        load_c(CA, i0);
        A.fq = 0;
        B.fq = B.fh = 0;
        CB.f = 0;
        if (i0 + 1 < i1) {
            load_pf(B, i0 + 1, true);
            load_c(CB, i0 + 1);
        }
        for (int i = i0; i < i1; i += 2) {
            body(i, A, B, CA);
            if (i + 1 < i1) body(i + 1, B, A, CB);
        }
    } else {
        // one set of loads: plane i + 1's issued after plane i is staged
        // (PD 1) or plane i's just before it is staged (PD 0)
        Pf A;
        Pc CA;
        A.fq = 0;
        if constexpr (PD == 1) {
            load_pf(A, i0, false);
            load_c(CA, i0);
        }
        for (int i = i0; i < i1; i++) {
            if constexpr (PD == 0) {
                // the queue entry of plane i + 1 + H (shifted in at the top of
                // the next plane), then plane i's halo and centres
                if (i > i0) shift(A);
                if (i + 1 < i1) load_q(A, i + 1);
                load_pf(A, i, false);
                load_c(CA, i);
            }
            stage(A, CA);
            __syncthreads();
            if constexpr (PD == 1) {
                if (i + 1 < i1) {
                    load_pf(A, i + 1, true);
                    load_c(CA, i + 1);
                }
            }
            if (valid) compute(i, j, k, (int64_t)i * plane + col, nb, zk);
            __syncthreads();
            if constexpr (PD == 1) {
                if (i + 1 < i1) shift(A);
            }
        }
    }
}

}  // namespace pipe
}  // namespace nova
