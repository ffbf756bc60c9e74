// nova_stencil.hpp -- device helpers shared by the stage, tangent-linear
// and adjoint kernels: stage combinations, the skew-symmetric advection
// form of ex::div(U, phi), the Laplacian and boundary index maps.
#pragma once

#include "nova_internal.hpp"

namespace nova {

// Stage combination forms.
enum { FORM_SSP = 0, FORM_LIT4 = 1, FORM_SHEAR = 2 };

struct StageArgs {
    int form;
    double alpha, beta;  // FORM_SSP: alpha*U0 + beta*(U + dU); FORM_SHEAR: U0 + beta*dU
    double dt;
    double w[4];         // FORM_LIT4 accumulation weights
};

__device__ inline double combine(const StageArgs &a, double u, double u0, double du)
{
    if (a.form == FORM_SSP) {
        const double v = u + du;
        return a.alpha != 0.0 ? a.alpha * u0 + a.beta * v : v;
    }
    if (a.form == FORM_LIT4) {
        double s = u;
        for (int j = 0; j < 4; j++) s = s + a.w[j] * du;
        return s;
    }
    return u0 + a.beta * du;
}

// ---- right-hand sides, written against an accessor A(f, c, d, o) that
// returns component c of field f at the node displaced by o along
// direction d (d = 0: i, 1: j, 2: k; o = 0 is the node itself) ----

// div(X, Y_c): 0.5 sum_d sum_l a [ (X_d Y_c)^+ - (X_d Y_c)^- + X_d (Y_c^+ - Y_c^-)
// + Y_c (X_d^+ - X_d^-) ]  (src/ex/div/div.hxx:101-169).
template <int H, class Acc>
__device__ inline double div_xy(const Grid &g, const Acc &A, int fx, int fy, int c, double zi)
{
    const double yc = A(fy, c, 0, 0);
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double xd = A(fx, d, 0, 0);
        double ad = 0.0;
#pragma unroll
        for (int l = 0; l < H; l++) {
            const double xp = A(fx, d, d, l + 1), xm = A(fx, d, d, -l - 1);
            const double yp = A(fy, c, d, l + 1), ym = A(fy, c, d, -l - 1);
            ad += g.a[d][l] * ((yp * xp - ym * xm) + xd * (yp - ym) + yc * (xp - xm));
        }
        acc += d == 2 ? ad * zi : ad;
    }
    return 0.5 * acc;
}

// nu * laplacian(f_c) with the z-metric term
// (src/ex/laplacian/laplacian.hxx:140-188).
template <int H, class Acc>
__device__ inline double lap_c(const Grid &g, const Acc &A, int f, int c, double nu, double zl, double zm)
{
    const double f0 = A(f, c, 0, 0);
    double acc = 0.0, dz1 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double ad = 0.0;
#pragma unroll
        for (int l = 0; l < H; l++) {
            const double fp = A(f, c, d, l + 1), fm = A(f, c, d, -l - 1);
            ad += g.b[d][l] * ((fp + fm) - 2.0 * f0);
            if (d == 2) dz1 += g.a[2][l] * (fp - fm);
        }
        acc += d == 2 ? ad * zl : ad;
    }
    return nu * acc - nu * zm * dz1;
}

// d/dx_d of 0.5|B|^2 (app/OT_Poisson/UEqn.H:22).
template <int H, class Acc>
__device__ inline double grad_pm(const Grid &g, const Acc &A, int fb, int d, double zi)
{
    double ad = 0.0;
#pragma unroll
    for (int l = 0; l < H; l++) {
        double bp = 0.0, bm = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double vp = A(fb, c, d, l + 1), vm = A(fb, c, d, -l - 1);
            bp += vp * vp;
            bm += vm * vm;
        }
        ad += g.a[d][l] * (0.5 * bp - 0.5 * bm);
    }
    return d == 2 ? ad * zi : ad;
}

// NS stage increment du = dt (-div(U, U) + nu lap(U)) of all three
// components, with the weights folded per direction d and offset l + 1
// (the same terms as div_xy and lap_c, regrouped):
//   du_c = sum_d sum_l [ bp U_c(+) + bm U_c(-) ] + U_c(0) (dv - lam)
//   bp = k1 (U_d(+) + U_d(0)) + kp,   bm = km - k1 (U_d(-) + U_d(0))
//   dv = sum_d sum_l k1 (U_d(+) - U_d(-))
// with k1 = -dt/2 wa_l s_d (s_d = 1/h_d, times 1/dzdxi on z),
// kp = km = dt nu wb_l s2_d (s2_d = 1/h_d^2, times 1/dzdxi^2 on z; on z
// minus / plus the metric term dt nu (d2zdxi2/dzdxi^3) wa_l / h_z) and
// lam = dt nu wb_sum2 sum_d s2_d.  The advective half of bp, bm carries
// the conservative flux U_c U_d and the U_d D U_c part of the skew form,
// dv the U_c div U part (U(+), U(-), U(0): the values at the offset node
// +(l+1), -(l+1) along d and at the node itself).  About 120 FP64
// operations per node for m = 6.

// Launch-uniform coefficients of the x and y directions and the z factors
// before the per-k metric (products only, so host and device round them
// alike).
struct NsK {
    double k1[2][kMaxH], kb[2][kMaxH];   // x, y: -dt/2 wa_l / h_d, dt nu wb_l / h_d^2
    double k1z[kMaxH], kbz[kMaxH], kmz[kMaxH];  // z: -dt/2 wa_l / h_z, dt nu wb_l / h_z^2, dt nu wa_l / h_z
    double lxy, lz;                      // dt nu wb_sum2 (1/h_x^2 + 1/h_y^2), dt nu wb_sum2 / h_z^2
};

__host__ __device__ inline NsK ns_k(const Grid &g, double nu, double dt)
{
    NsK k{};
    const double hdt = -0.5 * dt, ndt = nu * dt;
    for (int l = 0; l < g.h; l++) {
        for (int d = 0; d < 2; d++) {
            k.k1[d][l] = hdt * (g.wa[l] * g.ih[d]);
            k.kb[d][l] = ndt * (g.wb[l] * g.ih2[d]);
        }
        k.k1z[l] = hdt * (g.wa[l] * g.ih[2]);
        k.kbz[l] = ndt * (g.wb[l] * g.ih2[2]);
        k.kmz[l] = ndt * (g.wa[l] * g.ih[2]);
    }
    const double sxy = g.ih2[0] + g.ih2[1];
    k.lxy = ndt * (g.wb_sum2 * sxy);
    k.lz = ndt * (g.wb_sum2 * g.ih2[2]);
    return k;
}

// The z coefficients of one thread (its k fixes the metric factors).
template <int H>
struct NsZ {
    double k1[H], kp[H], km[H];
    double lam;
};

template <int H>
__device__ inline NsZ<H> ns_z(const NsK &k, double zi, double zl, double zm)
{
    NsZ<H> z;
#pragma unroll
    for (int l = 0; l < H; l++) {
        z.k1[l] = k.k1z[l] * zi;
        const double b = k.kbz[l] * zl, mt = k.kmz[l] * zm;
        z.kp[l] = b - mt;
        z.km[l] = b + mt;
    }
    z.lam = k.lxy + k.lz * zl;
    return z;
}

template <int H, class Acc>
__device__ inline void ns_du(const NsK &k, const NsZ<H> &z, const Acc &A, double (&du)[3])
{
    double u0[3];
#pragma unroll
    for (int c = 0; c < 3; c++) u0[c] = A(0, c, 0, 0);
    double acc[3] = {0.0, 0.0, 0.0}, dv = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int l = 0; l < H; l++) {
            const double k1 = d == 2 ? z.k1[l] : k.k1[d][l];
            const double kp = d == 2 ? z.kp[l] : k.kb[d][l];
            const double km = d == 2 ? z.km[l] : k.kb[d][l];
            const double up = A(0, d, d, l + 1), um = A(0, d, d, -l - 1);
            const double bp = __builtin_fma(k1, up + u0[d], kp);
            const double bm = __builtin_fma(-k1, um + u0[d], km);
            dv = __builtin_fma(k1, up - um, dv);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double cp = c == d ? up : A(0, c, d, l + 1);
                const double cm = c == d ? um : A(0, c, d, -l - 1);
                acc[c] = __builtin_fma(bp, cp, acc[c]);
                acc[c] = __builtin_fma(bm, cm, acc[c]);
            }
        }
    }
    const double g0 = dv - z.lam;
#pragma unroll
    for (int c = 0; c < 3; c++) du[c] = __builtin_fma(u0[c], g0, acc[c]);
}

// MHD stage increments of app/OT_Poisson/UEqn.H:16-23 and BEqn.H:15-23,
//   du = dt (-div(U, U) + nu lap(U) + div(B, B) - grad(|B|^2 / 2))
//   db = dt (-div(U, B) + div(B, U) + eta lap(B)),
// regrouped per direction d and offset l + 1 as ns_du, with
// k1 = -dt/2 wa_l s_d and the Laplacian weights kn = dt nu wb_l s2_d,
// ke = dt eta wb_l s2_d (on z with the metric term, as in ns_du):
//   du_c += (k1 aU+ + kn+) U_c(+) + (kn- - k1 aU-) U_c(-) + b+ B_c(+) + b- B_c(-)
//   db_c += (k1 aU+ + ke+) B_c(+) + (ke- - k1 aU-) B_c(-) + b+ U_c(+) + b- U_c(-)
//   du_d += k1 (|B(+)|^2 - |B(-)|^2)
// with aU+- = U_d(+-) + U_d(0), b+ = -k1 (B_d(+) + B_d(0)),
// b- = k1 (B_d(-) + B_d(0)), and at the end du_c += U_c(0) (dvU - lam_n)
// - B_c(0) dvB, db_c += B_c(0) (dvU - lam_e) - U_c(0) dvB, with
// dvX = sum k1 (X_d(+) - X_d(-)).  The per-direction factors are kept and the
// offset weights wa_l, wb_l applied where they are used: two fields leave
// too few registers to hold every coefficient.  Used by the U block and the
// B stages of the reference's composition (one increment each); the SSP
// stage that forms both increments at once keeps the per-component div_xy /
// lap_c form, whose working set fits one wave's registers (the regrouped
// form of both increments together spills).
struct MhdK {
    double K[3], N[3], E[3];  // -dt/2 / h_d, dt nu / h_d^2, dt eta / h_d^2
    double mn, me;            // dt nu / h_z, dt eta / h_z (times wa_l d2zdxi2/dzdxi^3: the z-metric term)
    double lnxy, lnz, lexy, lez;
};

__host__ __device__ inline MhdK mhd_k(const Grid &g, double nu, double eta, double dt)
{
    MhdK k{};
    const double hdt = -0.5 * dt, ndt = nu * dt, edt = eta * dt;
    for (int d = 0; d < 3; d++) {
        k.K[d] = hdt * g.ih[d];
        k.N[d] = ndt * g.ih2[d];
        k.E[d] = edt * g.ih2[d];
    }
    k.mn = ndt * g.ih[2];
    k.me = edt * g.ih[2];
    const double sxy = g.ih2[0] + g.ih2[1];
    k.lnxy = ndt * (g.wb_sum2 * sxy);
    k.lnz = ndt * (g.wb_sum2 * g.ih2[2]);
    k.lexy = edt * (g.wb_sum2 * sxy);
    k.lez = edt * (g.wb_sum2 * g.ih2[2]);
    return k;
}

// The z factors of one thread.
struct MhdZ {
    double K, N, E, Mn, Me, lamn, lame;
};

__device__ inline MhdZ mhd_z(const MhdK &k, double zi, double zl, double zm)
{
    MhdZ z;
    z.K = k.K[2] * zi;
    z.N = k.N[2] * zl;
    z.E = k.E[2] * zl;
    z.Mn = k.mn * zm;
    z.Me = k.me * zm;
    z.lamn = k.lnxy + k.lnz * zl;
    z.lame = k.lexy + k.lez * zl;
    return z;
}

// WU / WB: which increments to form (the reference's composition advances U
// and B in separate blocks); fields 0 = U, 1 = B of the accessor.
template <int H, bool WU, bool WB, class Acc>
__device__ inline void mhd_dudb(const Grid &g, const MhdK &k, const MhdZ &z, const Acc &A, double (&du)[3],
                                double (&db)[3])
{
    double u0[3], b0[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        u0[c] = A(0, c, 0, 0);
        b0[c] = A(1, c, 0, 0);
        du[c] = 0.0;
        db[c] = 0.0;
    }
    double dvu = 0.0, dvb = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int l = 0; l < H; l++) {
            const double wa = g.wa[l], wb = g.wb[l];
            const double k1 = (d == 2 ? z.K : k.K[d]) * wa;
            const double up = A(0, d, d, l + 1), um = A(0, d, d, -l - 1);
            const double bp = A(1, d, d, l + 1), bm = A(1, d, d, -l - 1);
            const double aup = up + u0[d], aum = um + u0[d];
            const double bcp = -k1 * (bp + b0[d]), bcm = k1 * (bm + b0[d]);
            dvu = __builtin_fma(k1, up - um, dvu);
            dvb = __builtin_fma(k1, bp - bm, dvb);
            double ap = 0.0, am = 0.0, ep = 0.0, em = 0.0;
            if constexpr (WU) {
                const double b = (d == 2 ? z.N : k.N[d]) * wb;
                const double mt = d == 2 ? z.Mn * wa : 0.0;
                ap = __builtin_fma(k1, aup, b - mt);
                am = __builtin_fma(-k1, aum, b + mt);
                double q = 0.0;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double vp = c == d ? bp : A(1, c, d, l + 1), vm = c == d ? bm : A(1, c, d, -l - 1);
                    q = __builtin_fma(vp, vp, q);
                    q = __builtin_fma(-vm, vm, q);
                }
                du[d] = __builtin_fma(k1, q, du[d]);
            }
            if constexpr (WB) {
                const double b = (d == 2 ? z.E : k.E[d]) * wb;
                const double mt = d == 2 ? z.Me * wa : 0.0;
                ep = __builtin_fma(k1, aup, b - mt);
                em = __builtin_fma(-k1, aum, b + mt);
            }
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double cup = c == d ? up : A(0, c, d, l + 1), cum = c == d ? um : A(0, c, d, -l - 1);
                const double cbp = c == d ? bp : A(1, c, d, l + 1), cbm = c == d ? bm : A(1, c, d, -l - 1);
                if constexpr (WU) {
                    du[c] = __builtin_fma(ap, cup, du[c]);
                    du[c] = __builtin_fma(am, cum, du[c]);
                    du[c] = __builtin_fma(bcp, cbp, du[c]);
                    du[c] = __builtin_fma(bcm, cbm, du[c]);
                }
                if constexpr (WB) {
                    db[c] = __builtin_fma(ep, cbp, db[c]);
                    db[c] = __builtin_fma(em, cbm, db[c]);
                    db[c] = __builtin_fma(bcp, cup, db[c]);
                    db[c] = __builtin_fma(bcm, cum, db[c]);
                }
            }
        }
    }
    const double gu = dvu - z.lamn, gb = dvu - z.lame;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if constexpr (WU) du[c] = __builtin_fma(u0[c], gu, __builtin_fma(-b0[c], dvb, du[c]));
        if constexpr (WB) db[c] = __builtin_fma(b0[c], gb, __builtin_fma(-u0[c], dvb, db[c]));
    }
}

// Tangent-linear right-hand side J_R(X) d = -B(d, X) - B(X, d) + nu L d for
// all three components (X = field 0, d = field 1), direction by direction:
//   B(d,X)_c + B(X,d)_c = 1/2 sum_d [ D_d(d_d X_c + X_d d_c) + d_d D_d X_c
//                          + X_c D_d d_d + X_d D_d d_c + d_c D_d X_d ].
template <int H, class Acc>
__device__ inline void tlm_rhs(const Grid &g, const Acc &A, double nu, double zi, double zl, double zm,
                               double (&out)[3])
{
    double x0[3], v0[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        x0[c] = A(0, c, 0, 0);
        v0[c] = A(1, c, 0, 0);
    }
    double adv[3] = {0.0, 0.0, 0.0}, lap[3] = {0.0, 0.0, 0.0}, dz1[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double dq[3] = {0.0, 0.0, 0.0}, dx[3] = {0.0, 0.0, 0.0}, dv[3] = {0.0, 0.0, 0.0}, sl[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int l = 0; l < H; l++) {
            const double xdp = A(0, d, d, l + 1), xdm = A(0, d, d, -l - 1);
            const double vdp = A(1, d, d, l + 1), vdm = A(1, d, d, -l - 1);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double xp = c == d ? xdp : A(0, c, d, l + 1), xm = c == d ? xdm : A(0, c, d, -l - 1);
                const double vp = c == d ? vdp : A(1, c, d, l + 1), vm = c == d ? vdm : A(1, c, d, -l - 1);
                dq[c] += g.wa[l] * ((vdp * xp + xdp * vp) - (vdm * xm + xdm * vm));
                dx[c] += g.wa[l] * (xp - xm);
                dv[c] += g.wa[l] * (vp - vm);
                sl[c] += g.wb[l] * (vp + vm);
            }
        }
        const double sa = d == 2 ? g.ih[2] * zi : g.ih[d];
        const double sb = d == 2 ? g.ih2[2] * zl : g.ih2[d];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            adv[c] += sa * (dq[c] + v0[d] * dx[c] + x0[c] * dv[d] + x0[d] * dv[c] + v0[c] * dx[d]);
            lap[c] += sb * (sl[c] - g.wb_sum2 * v0[c]);
            if (d == 2) dz1[c] = g.ih[2] * dv[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) out[c] = (-(0.5 * adv[c])) + nu * (lap[c] - zm * dz1[c]);
}

// The adjoint increment dt J_R(X)^T m, J_R(X)^T m = B(X, m) - m div X
// - B(., X)^T m + nu L m (X = field 0, m = field 1), with
//   B(X,m)_c = 1/2 sum_d [ D_d(X_d m_c) + X_d D_d m_c + m_c D_d X_d ]
//   (B(.,X)^T m)_c = 1/2 [ sum_c' ( -X_c' D_c m_c' + m_c' D_c X_c' ) - D_c s ],
//   s = sum_c m_c X_c,
// and the weights folded per direction d and offset l + 1 as in ns_du:
//   du_c = sum_d sum_l [ bp m_c(+) + bm m_c(-) ] - m_c(0) (dv + lam) + T_c
//   bp = h (X_d(+) + X_d(0)) + kb,   bm = kb - h (X_d(-) + X_d(0))
//   dv = sum_d sum_l h (X_d(+) - X_d(-))
//   T_c = sum_l h_c,l sum_c' [ (m_c'(+) - m_c'(0)) (X_c'(+) + X_c'(0))
//                             - (m_c'(-) - m_c'(0)) (X_c'(-) + X_c'(0)) ]
//         (offsets along direction c only: the -(B(., X)^T m)_c term)
// with h = dt/2 wa_l s_d, kb = dt nu wb_l s2_d, lam = dt nu wb_sum2 sum_d
// s2_d (s_d, s2_d as in ns_du; the adjoint needs a uniform z, so there is no
// metric term).  About 270 FP64 operations per node for m = 6, against about
// 400 for the direction-outer form of the six derivative sums.
struct AdjK {
    double h[3][kMaxH], kb[3][kMaxH];  // dt/2 wa_l / h_d, dt nu wb_l / h_d^2 (z: before the per-k factor)
    double lxy, lz;                    // dt nu wb_sum2 (1/h_x^2 + 1/h_y^2), dt nu wb_sum2 / h_z^2
};

__host__ __device__ inline AdjK adj_k(const Grid &g, double nu, double dt)
{
    AdjK k{};
    const double hdt = 0.5 * dt, ndt = nu * dt;
    for (int l = 0; l < g.h; l++)
        for (int d = 0; d < 3; d++) {
            k.h[d][l] = hdt * (g.wa[l] * g.ih[d]);
            k.kb[d][l] = ndt * (g.wb[l] * g.ih2[d]);
        }
    k.lxy = ndt * (g.wb_sum2 * (g.ih2[0] + g.ih2[1]));
    k.lz = ndt * (g.wb_sum2 * g.ih2[2]);
    return k;
}

template <int H>
struct AdjZ {
    double h[H], kb[H];
    double lam;
};

template <int H>
__device__ inline AdjZ<H> adj_z(const AdjK &k, double zi, double zl)
{
    AdjZ<H> z;
#pragma unroll
    for (int l = 0; l < H; l++) {
        z.h[l] = k.h[2][l] * zi;
        z.kb[l] = k.kb[2][l] * zl;
    }
    z.lam = k.lxy + k.lz * zl;
    return z;
}

// The terms of the adjoint increment that come from the offsets along one
// direction d (weights h[l], kb[l] of that direction): A_c = sum_l [bp m_c(+)
// + bm m_c(-)], dv = sum_l h (X_d(+) - X_d(-)), T = sum_l h t_l (the T_d term
// above), each summed over l from zero.  x0, m0: the values at the node.
template <int H, class Acc>
__device__ inline void adj_dir(const double *h, const double *kb, const Acc &A, int d, const double (&x0)[3],
                               const double (&m0)[3], double (&Ad)[3], double &dv, double &T)
{
    Ad[0] = Ad[1] = Ad[2] = 0.0;
    dv = 0.0;
    T = 0.0;
#pragma unroll
    for (int l = 0; l < H; l++) {
        double P[3], M[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            P[c] = A(0, c, d, l + 1) + x0[c];
            M[c] = A(0, c, d, -l - 1) + x0[c];
        }
        const double bp = __builtin_fma(h[l], P[d], kb[l]);
        const double bm = __builtin_fma(-h[l], M[d], kb[l]);
        dv = __builtin_fma(h[l], A(0, d, d, l + 1) - A(0, d, d, -l - 1), dv);
        double t = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double mp = A(1, c, d, l + 1), mm = A(1, c, d, -l - 1);
            Ad[c] = __builtin_fma(bp, mp, Ad[c]);
            Ad[c] = __builtin_fma(bm, mm, Ad[c]);
            t = __builtin_fma(mp - m0[c], P[c], t);
            t = __builtin_fma(m0[c] - mm, M[c], t);
        }
        T = __builtin_fma(h[l], t, T);
    }
}

// The increment from the three directions' terms: the sums over directions
// are taken as (x + y) + z, so that a kernel computing the directions in
// different waves (the direction-split adjoint stage) gives the same bits.
__device__ inline void adj_combine(const double (&Ad)[3][3], const double (&dv)[3], const double (&T)[3],
                                   const double (&m0)[3], double lam, double (&du)[3])
{
    const double g0 = ((dv[0] + dv[1]) + dv[2]) + lam;
#pragma unroll
    for (int c = 0; c < 3; c++) du[c] = __builtin_fma(-m0[c], g0, (Ad[0][c] + Ad[1][c]) + Ad[2][c]) + T[c];
}

template <int H, class Acc>
__device__ inline void adj_du(const AdjK &k, const AdjZ<H> &z, const Acc &A, double (&du)[3])
{
    double x0[3], m0[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        x0[c] = A(0, c, 0, 0);
        m0[c] = A(1, c, 0, 0);
    }
    double Ad[3][3], dv[3], T[3];
    adj_dir<H>(k.h[0], k.kb[0], A, 0, x0, m0, Ad[0], dv[0], T[0]);
    adj_dir<H>(k.h[1], k.kb[1], A, 1, x0, m0, Ad[1], dv[1], T[1]);
    adj_dir<H>(z.h, z.kb, A, 2, x0, m0, Ad[2], dv[2], T[2]);
    adj_combine(Ad, dv, T, m0, z.lam, du);
}

// Sign of a value reached through a reflection (walls only; a periodic
// grid has no reflections and carries no signs).
template <bool WALLS>
__device__ inline double rsign(const Grid &g, int c, int d, int flip)
{
    return WALLS ? bsign(g, 3, c, d, flip) : 1.0;
}

// v, negated when bit c of m is set: m is the exclusive or of the odd
// masks (Grid::odd) of the directions the value was reflected across (walls
// only).  The sign bit is flipped, which equals a product with -1 exactly.
template <bool WALLS>
__device__ inline double rneg(double v, unsigned m, int c)
{
    if constexpr (!WALLS) return v;
    const long long b = (long long)((m >> c) & 1u) << 63;
    return __longlong_as_double(__double_as_longlong(v) ^ b);
}

// One wrap suffices on a periodic axis (|offset| <= 4 < n); walls reflect.
template <bool WALLS>
__device__ inline int wrap(int q, int n, int wall, int &flip)
{
    if (!WALLS) {
        flip = 0;
        q += q < 0 ? n : 0;
        q -= q >= n ? n : 0;
        return q;
    }
    return bmap(q, n, wall, flip);
}


}  // namespace nova
