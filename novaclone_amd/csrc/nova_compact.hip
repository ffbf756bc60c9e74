// nova_compact.hip -- compact sixth-order derivatives on periodic lines (no
// reference counterpart: the reference computes only the explicit central
// stencils of src/ex; this is the tridiagonal scheme of Lele (1992),
// J. Comput. Phys. 103:16-42, first derivative alpha = 1/3, a = 14/9,
// b = 1/9; second derivative alpha = 2/11, a = 12/11, b = 3/11).
//
// Line solver.  A = tridiag(alpha, 1, alpha) (cyclic) factors exactly as
// A = kappa (1 + r E^-1)(1 + r E) with alpha r^2 - r + alpha = 0, |r| < 1,
// kappa = alpha / r (E = shift).  A x = rhs is then two first-order
// recursions, w_i = rhs_i / kappa - r w_{i-1} and x_i = w_i - r x_{i+1}, whose
// periodic solutions are geometric sums.  One wavefront owns a line: lane p
// holds the MP consecutive points [p MP, (p+1) MP) in registers, runs each
// recursion over its chunk (Thomas), and obtains the carry into its chunk
// from a cyclic Kogge-Stone scan across the P = N / MP lanes with
// multiplier rho = (-r)^MP (partition method, PCR-like across lanes); the
// scan stops once rho^(2^K) is below 1e-20, so the result is the exact
// periodic solution to rounding.
//
// Pencil passes.  A pass computes, for every line along one direction, the
// compact derivatives a model needs and combines them pointwise.  Lines
// along k are contiguous and read straight into registers; lines along i and
// j are staged through an LDS tile of TL consecutive k values, stored so
// that each lane's chunk element j sits at j P + lane (conflict-free reads).
#include "nova_internal.hpp"
#include "nova_stencil.hpp"

#include <cmath>

struct nova_poisson;

namespace nova {

namespace {

constexpr int kScanMax = 6;

// One compact scheme along one direction.
struct CmpScheme {
    double r;               // factor root
    double ca, cb;          // right-hand-side weights, divided by kappa
    double rp[9];           // (-r)^q, q = 0..8
    double rho2[kScanMax];  // rho^(2^s), rho = (-r)^MP
    int dsh[kScanMax];      // lane distance of scan step s: 2^s mod P
    int K;                  // scan steps
};

struct CmpDir {
    CmpScheme s[2];  // first and second derivative
    int N, P;        // line length and lanes per line
};

constexpr double kAl[2] = {1.0 / 3.0, 2.0 / 11.0};
constexpr double kA[2] = {14.0 / 9.0, 12.0 / 11.0};
constexpr double kB[2] = {1.0 / 9.0, 3.0 / 11.0};

int pick_mp(int n)
{
    if (n <= 64) return 1;
    if (n <= 128) return 2;
    if (n <= 256) return 4;
    if (n <= 512) return 8;
    return 0;
}

CmpDir make_dir(int n, double h, int mp)
{
    CmpDir D{};
    D.N = n;
    D.P = n / mp;
    for (int o = 0; o < 2; o++) {
        CmpScheme &S = D.s[o];
        const double al = kAl[o];
        S.r = (1.0 - std::sqrt(1.0 - 4.0 * al * al)) / (2.0 * al);
        const double kappa = al / S.r;
        if (o == 0) {
            S.ca = kA[o] / (2.0 * h) / kappa;
            S.cb = kB[o] / (4.0 * h) / kappa;
        } else {
            S.ca = kA[o] / (h * h) / kappa;
            S.cb = kB[o] / (4.0 * h * h) / kappa;
        }
        double q = 1.0;
        for (int e = 0; e <= 8; e++) {
            S.rp[e] = q;
            q *= -S.r;
        }
        const double rho = S.rp[mp];
        double rr = rho;
        S.K = 0;
        for (int s = 0; s < kScanMax; s++) {
            S.rho2[s] = rr;
            S.dsh[s] = (1 << s) % D.P;
            if (std::fabs(rr) > 1e-20) S.K = s + 1;
            rr *= rr;
        }
    }
    return D;
}

// Unsigned division by a run-time divisor d >= 2 as a multiply-high by
// m = ceil(2^32 / d): x m / 2^32 = x / d + x e / 2^32 with 0 <= e < 1, whose
// floor is floor(x / d) while x < 2^32 / d, which the tile indices (x below
// 2^16, d at most 2^13) satisfy.  A division by a run-time value otherwise
// costs a few dozen instructions, on every staged element.
struct FastDiv {
    uint32_t m, d;
    __device__ explicit FastDiv(uint32_t dd) : m(0xFFFFFFFFu / dd + 1u), d(dd) {}
    __device__ uint32_t div(uint32_t x) const { return __umulhi(x, m); }
    __device__ uint32_t mod(uint32_t x) const { return x - div(x) * d; }
};

// ---- wavefront line algebra ----

// UR: scan steps unrolled to kScanMax, each taken when the scheme needs it
// (a uniform branch); fewer scalar instructions, more registers (the k
// passes and the MHD passes take it, whose occupancy the registers do not
// set lower)
template <int MP, bool UR = false>
__device__ inline void cmp_solve(const CmpScheme &S, int P, int lane, double (&x)[MP])
{
    const double r = S.r;
    double w[MP];
    w[0] = x[0];
#pragma unroll
    for (int j = 1; j < MP; j++) w[j] = x[j] - r * w[j - 1];
    double acc = w[MP - 1];
    // the scan's lane distances come precomputed (a modulo by the lane
    // count on every step cost tens of scalar instructions)
    auto scan_dn = [&](double &v) {
        if constexpr (UR) {
#pragma unroll
            for (int s = 0; s < kScanMax; s++) {
                if (s >= S.K) break;
                const int d = S.dsh[s];
                const int src = lane - d < 0 ? lane - d + P : lane - d;
                v = v + S.rho2[s] * __shfl(v, src, 64);
            }
        } else {
            for (int s = 0; s < S.K; s++) {
                const int d = S.dsh[s];
                const int src = lane - d < 0 ? lane - d + P : lane - d;
                v = v + S.rho2[s] * __shfl(v, src, 64);
            }
        }
    };
    auto scan_up = [&](double &v) {
        if constexpr (UR) {
#pragma unroll
            for (int s = 0; s < kScanMax; s++) {
                if (s >= S.K) break;
                const int d = S.dsh[s];
                const int src = lane + d >= P ? lane + d - P : lane + d;
                v = v + S.rho2[s] * __shfl(v, src, 64);
            }
        } else {
            for (int s = 0; s < S.K; s++) {
                const int d = S.dsh[s];
                const int src = lane + d >= P ? lane + d - P : lane + d;
                v = v + S.rho2[s] * __shfl(v, src, 64);
            }
        }
    };
    scan_dn(acc);
    const double cin = __shfl(acc, lane == 0 ? P - 1 : lane - 1, 64);
#pragma unroll
    for (int j = 0; j < MP; j++) w[j] = w[j] + S.rp[j + 1] * cin;
    x[MP - 1] = w[MP - 1];
#pragma unroll
    for (int j = MP - 2; j >= 0; j--) x[j] = w[j] - r * x[j + 1];
    acc = x[0];
    scan_up(acc);
    const double cin2 = __shfl(acc, lane + 1 >= P ? lane + 1 - P : lane + 1, 64);
#pragma unroll
    for (int j = 0; j < MP; j++) x[j] = x[j] + S.rp[MP - j] * cin2;
}

// right-hand side of the compact scheme (ORDER 1 or 2) from the chunk f
template <int MP, int ORDER>
__device__ inline void cmp_rhs(const CmpScheme &S, int P, int lane, const double (&f)[MP], double (&e)[MP])
{
    double ext[MP + 4];
#pragma unroll
    for (int j = 0; j < MP; j++) ext[j + 2] = f[j];
    const int lm1 = lane == 0 ? P - 1 : lane - 1;
    const int lp1 = lane + 1 >= P ? lane + 1 - P : lane + 1;
    if constexpr (MP >= 2) {
        ext[1] = __shfl(f[MP - 1], lm1, 64);
        ext[0] = __shfl(f[MP - 2], lm1, 64);
        ext[MP + 2] = __shfl(f[0], lp1, 64);
        ext[MP + 3] = __shfl(f[1], lp1, 64);
    } else {
        const int lm2 = lane - 2 < 0 ? lane - 2 + P : lane - 2;
        const int lp2 = lane + 2 >= P ? lane + 2 - P : lane + 2;
        ext[1] = __shfl(f[0], lm1, 64);
        ext[0] = __shfl(f[0], lm2, 64);
        ext[3] = __shfl(f[0], lp1, 64);
        ext[4] = __shfl(f[0], lp2, 64);
    }
#pragma unroll
    for (int j = 0; j < MP; j++) {
        if constexpr (ORDER == 1)
            e[j] = S.ca * (ext[j + 3] - ext[j + 1]) + S.cb * (ext[j + 4] - ext[j]);
        else
            e[j] = S.ca * ((ext[j + 3] - 2.0 * ext[j + 2]) + ext[j + 1]) +
                   S.cb * ((ext[j + 4] - 2.0 * ext[j + 2]) + ext[j]);
    }
}

template <int MP, int ORDER, bool UR = false>
__device__ inline void cmp_deriv(const CmpDir &D, int lane, const double (&f)[MP], double (&out)[MP])
{
    cmp_rhs<MP, ORDER>(D.s[ORDER - 1], D.P, lane, f, out);
    cmp_solve<MP, UR>(D.s[ORDER - 1], D.P, lane, out);
}

// NS systems of one scheme solved together: the same operations per system
// as cmp_solve (so the same results bit for bit), with every step of the
// cross-lane scans issued for all systems before any of them is used, so
// that NS shuffles are in flight at once instead of one dependent chain per
// system.
template <int MP, int NS>
__device__ inline void cmp_solve_n(const CmpScheme &S, int P, int lane, double (&x)[NS][MP])
{
    const double r = S.r;
    double w[NS][MP], acc[NS], o[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        w[s][0] = x[s][0];
#pragma unroll
        for (int j = 1; j < MP; j++) w[s][j] = x[s][j] - r * w[s][j - 1];
        acc[s] = w[s][MP - 1];
    }
    for (int st = 0; st < S.K; st++) {
        const int d = S.dsh[st];
        const int src = lane - d < 0 ? lane - d + P : lane - d;
#pragma unroll
        for (int s = 0; s < NS; s++) o[s] = __shfl(acc[s], src, 64);
#pragma unroll
        for (int s = 0; s < NS; s++) acc[s] = acc[s] + S.rho2[st] * o[s];
    }
    const int lm1 = lane == 0 ? P - 1 : lane - 1;
#pragma unroll
    for (int s = 0; s < NS; s++) o[s] = __shfl(acc[s], lm1, 64);
#pragma unroll
    for (int s = 0; s < NS; s++) {
#pragma unroll
        for (int j = 0; j < MP; j++) w[s][j] = w[s][j] + S.rp[j + 1] * o[s];
        x[s][MP - 1] = w[s][MP - 1];
#pragma unroll
        for (int j = MP - 2; j >= 0; j--) x[s][j] = w[s][j] - r * x[s][j + 1];
        acc[s] = x[s][0];
    }
    for (int st = 0; st < S.K; st++) {
        const int d = S.dsh[st];
        const int src = lane + d >= P ? lane + d - P : lane + d;
#pragma unroll
        for (int s = 0; s < NS; s++) o[s] = __shfl(acc[s], src, 64);
#pragma unroll
        for (int s = 0; s < NS; s++) acc[s] = acc[s] + S.rho2[st] * o[s];
    }
    const int lp1 = lane + 1 >= P ? lane + 1 - P : lane + 1;
#pragma unroll
    for (int s = 0; s < NS; s++) o[s] = __shfl(acc[s], lp1, 64);
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int j = 0; j < MP; j++) x[s][j] = x[s][j] + S.rp[MP - j] * o[s];
}

// NS compact derivatives of one order, f[s] -> out[s]
template <int MP, int ORDER, int NS>
__device__ inline void cmp_deriv_n(const CmpDir &D, int lane, const double (&f)[NS][MP], double (&out)[NS][MP])
{
#pragma unroll
    for (int s = 0; s < NS; s++) cmp_rhs<MP, ORDER>(D.s[ORDER - 1], D.P, lane, f[s], out[s]);
    cmp_solve_n<MP, NS>(D.s[ORDER - 1], D.P, lane, out);
}

// ---- pass policies: inputs differentiated along the pass direction and
// the pointwise use of the results ----

// out = D^order f (one component)
struct DerivW {
    static constexpr int NI = 1, NO = 1;
    static constexpr bool PAIR = false;  // the k pass stores point by point
    struct Args {
        const double *f;
        double *out;
        int order;
    };
    __device__ static const double *input(const Args &a, const Grid &, int) { return a.f; }
    template <int MP, int DIR>
    __device__ static void compute(const Args &a, const CmpDir &D, int lane, double (&in)[NI][MP],
                                   double (&out)[NO][MP])
    {
        if (a.order == 1) cmp_deriv<MP, 1>(D, lane, in[0], out[0]);
        else cmp_deriv<MP, 2>(D, lane, in[0], out[0]);
    }
    __device__ static void store(const Args &a, const Grid &, int, int64_t q, double v, double) { a.out[q] = v; }
    // the i / j passes store in two phases: every operand a store reads is
    // loaded first, then every result combined and written
    __device__ static double operand(const Args &, const Grid &, int, int64_t) { return 0.0; }
    __device__ static void store_op(const Args &a, const Grid &, int, int64_t q, double v, double) { a.out[q] = v; }
    __device__ static void store_point(const Args &a, const Grid &g, int64_t q, const double *v, const double *x)
    {
        store(a, g, 0, q, v[0], x[0]);
    }
};

// Navier-Stokes right-hand side along one direction d:
//   c_d = -1/2 [D_d(U_d U_c) + U_d D_d U_c + U_c D_d U_d] + nu D2_d U_c,
// mode 0: R = c_d; 1: R += c_d; 2: out = stage combination with dt (R + c_d).
struct NsW {
    static constexpr int NI = 3, NO = 3;
    static constexpr bool PAIR = true;  // the k pass stores two points at a time (store_pair)
    struct Args {
        const double *U;
        const double *U0;
        double *R;
        double *out;
        StageArgs sa;
        double nu;
        int mode;
    };
    __device__ static const double *input(const Args &a, const Grid &g, int c) { return a.U + c * g.np; }
    template <int MP, int DIR>
    __device__ static void compute(const Args &a, const CmpDir &D, int lane, double (&u)[NI][MP],
                                   double (&out)[NO][MP])
    {
        if constexpr (DIR == 2) {
            // k lines (no LDS tile, occupancy set by registers): one system
            // at a time, fewer registers live
            double dud[MP];
            cmp_deriv<MP, 1, true>(D, lane, u[DIR], dud);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double duc[MP], fl[MP], dfl[MP], l2[MP];
                if (c == DIR) {
#pragma unroll
                    for (int j = 0; j < MP; j++) duc[j] = dud[j];
                } else {
                    cmp_deriv<MP, 1, true>(D, lane, u[c], duc);
                }
#pragma unroll
                for (int j = 0; j < MP; j++) fl[j] = u[DIR][j] * u[c][j];
                cmp_deriv<MP, 1, true>(D, lane, fl, dfl);
                cmp_deriv<MP, 2, true>(D, lane, u[c], l2);
#pragma unroll
                for (int j = 0; j < MP; j++)
                    out[c][j] = -0.5 * ((dfl[j] + u[DIR][j] * duc[j]) + u[c][j] * dud[j]) + a.nu * l2[j];
            }
            return;
        }
        // i and j lines (occupancy set by the LDS tile): the six
        // first-derivative systems (D U_c, D (U_d U_c)) together, then the
        // three second-derivative ones, the same operations in the same order
        double f1[6][MP], d1[6][MP], d2[3][MP];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int j = 0; j < MP; j++) {
                f1[c][j] = u[c][j];
                f1[3 + c][j] = u[DIR][j] * u[c][j];
            }
        cmp_deriv_n<MP, 1, 6>(D, lane, f1, d1);
        cmp_deriv_n<MP, 2, 3>(D, lane, u, d2);
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int j = 0; j < MP; j++)
                out[c][j] = -0.5 * ((d1[3 + c][j] + u[DIR][j] * d1[c][j]) + u[c][j] * d1[DIR][j]) + a.nu * d2[c][j];
    }
    __device__ static void store(const Args &a, const Grid &g, int o, int64_t q, double v, double x)
    {
        const int64_t e = o * g.np + q;
        if (a.mode == 0) {
            a.R[e] = v;
        } else if (a.mode == 1) {
            a.R[e] = a.R[e] + v;
        } else {
            const double du = a.sa.dt * (a.R[e] + v);
            const double x0 = a.U0 ? a.U0[e] : x;
            a.out[e] = combine(a.sa, x, x0, du);
        }
    }
    __device__ static double operand(const Args &a, const Grid &g, int o, int64_t q)
    {
        return a.mode == 1 ? a.R[o * g.np + q] : 0.0;
    }
    __device__ static void store_op(const Args &a, const Grid &g, int o, int64_t q, double v, double r)
    {
        a.R[o * g.np + q] = a.mode == 1 ? r + v : v;  // the i and j passes: modes 0 and 1
    }
    // store() for the points q and q + 1 (q even on a line of even length,
    // so 16-byte aligned): the same expressions, 16-byte loads and stores
    __device__ static void store_pair(const Args &a, const Grid &g, int64_t q, const double (&v)[2][NO],
                                      const double (&x)[2][NI])
    {
#pragma unroll
        for (int o = 0; o < 3; o++) {
            const int64_t e = o * g.np + q;
            double2 *R2 = (double2 *)(a.R + e);
            if (a.mode == 0) {
                *R2 = double2{v[0][o], v[1][o]};
            } else if (a.mode == 1) {
                const double2 r = *R2;
                *R2 = double2{r.x + v[0][o], r.y + v[1][o]};
            } else {
                const double2 r = *R2;
                const double2 x0 = a.U0 ? *(const double2 *)(a.U0 + e) : double2{x[0][o], x[1][o]};
                const double du0 = a.sa.dt * (r.x + v[0][o]), du1 = a.sa.dt * (r.y + v[1][o]);
                *(double2 *)(a.out + e) = double2{combine(a.sa, x[0][o], x0.x, du0), combine(a.sa, x[1][o], x0.y, du1)};
            }
        }
    }    __device__ static void store_point(const Args &a, const Grid &g, int64_t q, const double *v, const double *x)
    {
#pragma unroll
        for (int o = 0; o < 3; o++) store(a, g, o, q, v[o], x[o]);
    }
};

// divergence along one direction: mode 0: d = D F_d; 1: d += D F_d;
// 2: d = (d + D F_d) * scale
struct DivW {
    static constexpr int NI = 1, NO = 1;
    static constexpr bool PAIR = false;  // the k pass stores point by point
    struct Args {
        const double *F;  // the pass direction's component
        double *d;
        double scale;
        int mode;
    };
    __device__ static const double *input(const Args &a, const Grid &, int) { return a.F; }
    template <int MP, int DIR>
    __device__ static void compute(const Args &, const CmpDir &D, int lane, double (&in)[NI][MP],
                                   double (&out)[NO][MP])
    {
        cmp_deriv<MP, 1>(D, lane, in[0], out[0]);
    }
    __device__ static void store(const Args &a, const Grid &, int, int64_t q, double v, double)
    {
        if (a.mode == 0) a.d[q] = v;
        else if (a.mode == 1) a.d[q] = a.d[q] + v;
        else a.d[q] = (a.d[q] + v) * a.scale;
    }
    __device__ static double operand(const Args &a, const Grid &, int, int64_t q) { return a.mode ? a.d[q] : 0.0; }
    __device__ static void store_op(const Args &a, const Grid &, int, int64_t q, double v, double d)
    {
        if (a.mode == 0) a.d[q] = v;
        else if (a.mode == 1) a.d[q] = d + v;
        else a.d[q] = (d + v) * a.scale;
    }
    __device__ static void store_point(const Args &a, const Grid &g, int64_t q, const double *v, const double *x)
    {
        store(a, g, 0, q, v[0], x[0]);
    }
};

// gradient correction along one direction: F_d -= s D_d p
struct GradW {
    static constexpr int NI = 1, NO = 1;
    static constexpr bool PAIR = false;  // the k pass stores point by point
    struct Args {
        const double *p;
        double *F;  // the pass direction's component
        double s;
    };
    __device__ static const double *input(const Args &a, const Grid &, int) { return a.p; }
    template <int MP, int DIR>
    __device__ static void compute(const Args &, const CmpDir &D, int lane, double (&in)[NI][MP],
                                   double (&out)[NO][MP])
    {
        cmp_deriv<MP, 1>(D, lane, in[0], out[0]);
    }
    __device__ static void store(const Args &a, const Grid &, int, int64_t q, double v, double)
    {
        a.F[q] = a.F[q] - a.s * v;
    }
    __device__ static double operand(const Args &a, const Grid &, int, int64_t q) { return a.F[q]; }
    __device__ static void store_op(const Args &a, const Grid &, int, int64_t q, double v, double f)
    {
        a.F[q] = f - a.s * v;
    }
    __device__ static void store_point(const Args &a, const Grid &g, int64_t q, const double *v, const double *x)
    {
        store(a, g, 0, q, v[0], x[0]);
    }
};

// Tangent-linear right-hand side along direction d (X = forward state, v =
// tangent): -1/2 [D_d(v_d X_c + X_d v_c) + v_d D_d X_c + X_c D_d v_d
// + X_d D_d v_c + v_c D_d X_d] + nu D2_d v_c; modes as NsW with v the
// stage input and v0 the step-start tangent.
struct TlmW {
    static constexpr int NI = 6, NO = 3;
    static constexpr bool PAIR = false;  // the k pass stores point by point
    struct Args {
        const double *X, *V, *V0;
        double *R, *out;
        StageArgs sa;
        double nu;
        int mode;
    };
    __device__ static const double *input(const Args &a, const Grid &g, int c)
    {
        return c < 3 ? a.X + c * g.np : a.V + (c - 3) * g.np;
    }
    template <int MP, int DIR>
    __device__ static void compute(const Args &a, const CmpDir &D, int lane, double (&in)[NI][MP],
                                   double (&out)[NO][MP])
    {
        double (&x)[3][MP] = *reinterpret_cast<double(*)[3][MP]>(&in[0][0]);
        double (&v)[3][MP] = *reinterpret_cast<double(*)[3][MP]>(&in[3][0]);
        double dxd[MP], dvd[MP];
        cmp_deriv<MP, 1>(D, lane, x[DIR], dxd);
        cmp_deriv<MP, 1>(D, lane, v[DIR], dvd);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double dxc[MP], dvc[MP], fl[MP], dfl[MP], l2[MP];
            if (c == DIR) {
#pragma unroll
                for (int j = 0; j < MP; j++) {
                    dxc[j] = dxd[j];
                    dvc[j] = dvd[j];
                }
            } else {
                cmp_deriv<MP, 1>(D, lane, x[c], dxc);
                cmp_deriv<MP, 1>(D, lane, v[c], dvc);
            }
#pragma unroll
            for (int j = 0; j < MP; j++) fl[j] = v[DIR][j] * x[c][j] + x[DIR][j] * v[c][j];
            cmp_deriv<MP, 1>(D, lane, fl, dfl);
            cmp_deriv<MP, 2>(D, lane, v[c], l2);
#pragma unroll
            for (int j = 0; j < MP; j++)
                out[c][j] = -0.5 * ((((dfl[j] + v[DIR][j] * dxc[j]) + x[c][j] * dvd[j]) + x[DIR][j] * dvc[j]) +
                                    v[c][j] * dxd[j]) +
                            a.nu * l2[j];
        }
    }
    __device__ static void store(const Args &a, const Grid &g, int o, int64_t q, double v, double x)
    {
        const int64_t e = o * g.np + q;
        if (a.mode == 0) {
            a.R[e] = v;
        } else if (a.mode == 1) {
            a.R[e] = a.R[e] + v;
        } else {
            const double du = a.sa.dt * (a.R[e] + v);
            const double x0 = a.V0 ? a.V0[e] : x;
            a.out[e] = combine(a.sa, x, x0, du);
        }
    }
    __device__ static double operand(const Args &a, const Grid &g, int o, int64_t q)
    {
        return a.mode == 1 ? a.R[o * g.np + q] : 0.0;
    }
    __device__ static void store_op(const Args &a, const Grid &g, int o, int64_t q, double v, double r)
    {
        a.R[o * g.np + q] = a.mode == 1 ? r + v : v;
    }    __device__ static void store_point(const Args &a, const Grid &g, int64_t q, const double *v, const double *x)
    {
#pragma unroll
        for (int o = 0; o < 3; o++) store(a, g, o, q, v[o], x[3 + o]);
    }
};

// Adjoint right-hand side along direction d (X = forward state, m =
// adjoint): B(X,m)_c = 1/2 [D_d(X_d m_c) + X_d D_d m_c + m_c D_d X_d],
// nu D2_d m_c, and for output component d the transposed advection term
// (B(., X)^T m)_d = 1/2 [sum_c (-X_c D_d m_c + m_c D_d X_c) - D_d(sum_c m_c X_c)];
// the fourth accumulator collects div X.  Final (k) pass:
// out = ob (m + dt (R_c - m_c div X)) + yg Y, acc = acc_a acc + acc_b m.
struct AdjW {
    static constexpr int NI = 6, NO = 4;
    static constexpr bool PAIR = false;  // the k pass stores point by point
    struct Args {
        const double *X, *M;
        const double *Y;
        double *acc;
        double *R;  // 4 components
        double *out;
        double dt, nu, ob, yg, acc_a, acc_b;
        int mode;
    };
    __device__ static const double *input(const Args &a, const Grid &g, int c)
    {
        return c < 3 ? a.X + c * g.np : a.M + (c - 3) * g.np;
    }
    template <int MP, int DIR>
    __device__ static void compute(const Args &a, const CmpDir &D, int lane, double (&in)[NI][MP],
                                   double (&out)[NO][MP])
    {
        double (&x)[3][MP] = *reinterpret_cast<double(*)[3][MP]>(&in[0][0]);
        double (&m)[3][MP] = *reinterpret_cast<double(*)[3][MP]>(&in[3][0]);
        double dx[3][MP];
#pragma unroll
        for (int c = 0; c < 3; c++) cmp_deriv<MP, 1>(D, lane, x[c], dx[c]);
        double bt[MP], sp[MP], dsp[MP];
#pragma unroll
        for (int j = 0; j < MP; j++) {
            bt[j] = 0.0;
            sp[j] = (m[0][j] * x[0][j] + m[1][j] * x[1][j]) + m[2][j] * x[2][j];
        }
        cmp_deriv<MP, 1>(D, lane, sp, dsp);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double dm[MP], fl[MP], dfl[MP], l2[MP];
            cmp_deriv<MP, 1>(D, lane, m[c], dm);
#pragma unroll
            for (int j = 0; j < MP; j++) {
                fl[j] = x[DIR][j] * m[c][j];
                bt[j] = bt[j] + ((-x[c][j] * dm[j]) + m[c][j] * dx[c][j]);
            }
            cmp_deriv<MP, 1>(D, lane, fl, dfl);
            cmp_deriv<MP, 2>(D, lane, m[c], l2);
#pragma unroll
            for (int j = 0; j < MP; j++)
                out[c][j] = 0.5 * ((dfl[j] + x[DIR][j] * dm[j]) + m[c][j] * dx[DIR][j]) + a.nu * l2[j];
        }
#pragma unroll
        for (int j = 0; j < MP; j++) {
            out[DIR][j] = out[DIR][j] - 0.5 * (bt[j] - dsp[j]);
            out[3][j] = dx[DIR][j];
        }
    }
    __device__ static void store(const Args &a, const Grid &g, int o, int64_t q, double v, double)
    {
        const int64_t e = o * g.np + q;
        if (a.mode == 0) a.R[e] = v;
        else a.R[e] = a.R[e] + v;
    }
    __device__ static double operand(const Args &a, const Grid &g, int o, int64_t q)
    {
        return a.mode ? a.R[o * g.np + q] : 0.0;
    }
    __device__ static void store_op(const Args &a, const Grid &g, int o, int64_t q, double v, double r)
    {
        a.R[o * g.np + q] = a.mode ? r + v : v;
    }
    __device__ static void store_point(const Args &a, const Grid &g, int64_t q, const double *v, const double *x)
    {
        if (a.mode < 2) {
#pragma unroll
            for (int o = 0; o < 4; o++) store(a, g, o, q, v[o], 0.0);
            return;
        }
        const double divx = a.R[3 * g.np + q] + v[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int64_t e = c * g.np + q;
            const double m = x[3 + c];
            const double jr = (a.R[e] + v[c]) - m * divx;
            double o = a.ob * (m + a.dt * jr);
            if (a.Y) o += a.yg * a.Y[e];
            if (a.acc) a.acc[e] = (a.acc_a != 0.0 ? a.acc_a * a.acc[e] : 0.0) + a.acc_b * m;
            a.out[e] = o;
        }
    }
};

// MHD right-hand sides along direction d (app/OT_Poisson/UEqn.H:16-23,
// BEqn.H:15-23 with compact operators), with Bf(X,Y)_c =
// 1/2 [D_d(X_d Y_c) + X_d D_d Y_c + Y_c D_d X_d]:
//   dU_c = -Bf(U,U)_c + Bf(B,B)_c - [c = d] D_d(|B|^2/2) + nu D2_d U_c,
//   dB_c = -Bf(U,B)_c + Bf(B,U)_c + eta D2_d B_c;  modes as NsW (6 comps).
struct MhdW {
    static constexpr int NI = 6, NO = 6;
    static constexpr bool PAIR = false;  // the k pass stores point by point
    struct Args {
        const double *U, *B, *U0, *B0;
        double *R;  // 6 components
        double *oU, *oB;
        StageArgs sa;
        double nu, eta;
        int mode;
    };
    __device__ static const double *input(const Args &a, const Grid &g, int c)
    {
        return c < 3 ? a.U + c * g.np : a.B + (c - 3) * g.np;
    }
    template <int MP, int DIR>
    __device__ static void compute(const Args &a, const CmpDir &D, int lane, double (&in)[NI][MP],
                                   double (&out)[NO][MP])
    {
        double (&u)[3][MP] = *reinterpret_cast<double(*)[3][MP]>(&in[0][0]);
        double (&b)[3][MP] = *reinterpret_cast<double(*)[3][MP]>(&in[3][0]);
        // The systems are solved in batches (cmp_deriv_n: the same operations
        // per system as one solve at a time, with the batch's cross-lane scan
        // steps interleaved): first u_d, b_d and the magnetic pressure
        // 0.5 |B|^2, then the six second derivatives, then per component c
        // the first derivatives of u_c, b_c and the four products together.
        double dd[3][MP];
        {
            double f[3][MP];
#pragma unroll
            for (int j = 0; j < MP; j++) {
                f[0][j] = u[DIR][j];
                f[1][j] = b[DIR][j];
                f[2][j] = 0.5 * ((b[0][j] * b[0][j] + b[1][j] * b[1][j]) + b[2][j] * b[2][j]);
            }
            cmp_deriv_n<MP, 1, 3>(D, lane, f, dd);
        }
        double l2a[6][MP];
        cmp_deriv_n<MP, 2, 6>(D, lane, in, l2a);
        const double(&dud)[MP] = dd[0];
        const double(&dbd)[MP] = dd[1];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            // first derivatives: u_c, b_c, u_d u_c, b_d b_c, u_d b_c, b_d u_c
            double r[6][MP];
            if (c == DIR) {
                double f[3][MP], r3[3][MP];
#pragma unroll
                for (int j = 0; j < MP; j++) {
                    f[0][j] = u[DIR][j] * u[c][j];
                    f[1][j] = b[DIR][j] * b[c][j];
                    f[2][j] = u[DIR][j] * b[c][j];
                }
                cmp_deriv_n<MP, 1, 3>(D, lane, f, r3);
#pragma unroll
                for (int j = 0; j < MP; j++) {
                    r[0][j] = dud[j];
                    r[1][j] = dbd[j];
                    r[2][j] = r3[0][j];
                    r[3][j] = r3[1][j];
                    r[4][j] = r3[2][j];
                    r[5][j] = r3[2][j];  // b_d u_d = u_d b_d
                }
            } else {
                double f[6][MP];
#pragma unroll
                for (int j = 0; j < MP; j++) {
                    f[0][j] = u[c][j];
                    f[1][j] = b[c][j];
                    f[2][j] = u[DIR][j] * u[c][j];
                    f[3][j] = b[DIR][j] * b[c][j];
                    f[4][j] = u[DIR][j] * b[c][j];
                    f[5][j] = b[DIR][j] * u[c][j];
                }
                cmp_deriv_n<MP, 1, 6>(D, lane, f, r);
            }
            const double(&duc)[MP] = r[0];
            const double(&dbc)[MP] = r[1];
            const double(&d1)[MP] = r[2];
            const double(&d2)[MP] = r[3];
            const double(&d3)[MP] = r[4];
            const double(&d4)[MP] = r[5];
            const double(&l2u)[MP] = l2a[c];
            const double(&l2b)[MP] = l2a[3 + c];
#pragma unroll
            for (int j = 0; j < MP; j++)
                out[c][j] = ((-0.5 * ((d1[j] + u[DIR][j] * duc[j]) + u[c][j] * dud[j])) +
                             0.5 * ((d2[j] + b[DIR][j] * dbc[j]) + b[c][j] * dbd[j])) +
                            a.nu * l2u[j];
#pragma unroll
            for (int j = 0; j < MP; j++)
                out[3 + c][j] = ((-0.5 * ((d3[j] + u[DIR][j] * dbc[j]) + b[c][j] * dud[j])) +
                                 0.5 * ((d4[j] + b[DIR][j] * duc[j]) + u[c][j] * dbd[j])) +
                                a.eta * l2b[j];
        }
        const double(&dpm)[MP] = dd[2];
#pragma unroll
        for (int j = 0; j < MP; j++) out[DIR][j] = out[DIR][j] - dpm[j];
    }
    __device__ static void store(const Args &a, const Grid &g, int o, int64_t q, double v, double)
    {
        const int64_t e = o * g.np + q;
        if (a.mode == 0) a.R[e] = v;
        else a.R[e] = a.R[e] + v;
    }
    __device__ static double operand(const Args &a, const Grid &g, int o, int64_t q)
    {
        return a.mode ? a.R[o * g.np + q] : 0.0;
    }
    __device__ static void store_op(const Args &a, const Grid &g, int o, int64_t q, double v, double r)
    {
        a.R[o * g.np + q] = a.mode ? r + v : v;
    }
    __device__ static void store_point(const Args &a, const Grid &g, int64_t q, const double *v, const double *x)
    {
        if (a.mode < 2) {
#pragma unroll
            for (int o = 0; o < 6; o++) store(a, g, o, q, v[o], 0.0);
            return;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int64_t e = c * g.np + q;
            const double du = a.sa.dt * (a.R[e] + v[c]);
            const double db = a.sa.dt * (a.R[(3 + c) * g.np + q] + v[3 + c]);
            const double u0 = a.U0 ? a.U0[e] : x[c];
            const double b0 = a.B0 ? a.B0[e] : x[3 + c];
            a.oU[e] = combine(a.sa, x[c], u0, du);
            a.oB[e] = combine(a.sa, x[3 + c], b0, db);
        }
    }
};

// k lines per LDS tile and threads per block: single-field passes
// (derivative, divergence, gradient correction; few registers) at up to
// 256 points a line take 16 consecutive k (128-byte rows, 32 KB) on 512
// threads; the model passes 48 KB at most on 256 threads, since their
// registers leave three waves per SIMD (measured at 256^3: the NS passes
// ran 454 / 523 us with 8 k on 256 threads and 589 / 742 us with 16 on 512,
// the divergence and correction 115 / 115 and 94 / 88 us)
template <int MP, class W>
constexpr int tile_lines()
{
    constexpr int nb = W::NI > W::NO ? W::NI : W::NO;
    return nb == 1 && MP <= 4 ? 16 : (nb <= 3 ? 8 : 4) / (MP <= 4 ? 1 : 2);
}
template <int MP, class W>
constexpr int tile_threads()
{
    return tile_lines<MP, W>() == 16 ? 512 : 256;
}

// Lines along k: one wavefront per (i, j) line, no LDS.
template <int MP, class W>
__global__ void __launch_bounds__(256) k_cmp_pass_k(Grid g, CmpDir D, typename W::Args a)
{
    const int lane = threadIdx.x & 63;
    const int64_t line = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (line >= (int64_t)g.s[0] * g.s[1]) return;  // whole wavefronts leave; no block barriers below
    const int64_t base = line * g.s[2];
    const bool on = lane < D.P;
    double in[W::NI][MP], out[W::NO][MP];
#pragma unroll
    for (int c = 0; c < W::NI; c++) {
        const double *f = W::input(a, g, c) + base;
        if constexpr (MP % 2 == 0) {
            // a lane's MP points are consecutive and start at an even node
            // of a line of even length: 16-byte loads, two points each
#pragma unroll
            for (int j = 0; j < MP; j += 2) {
                const double2 v = on ? *(const double2 *)(f + lane * MP + j) : double2{0.0, 0.0};
                in[c][j] = v.x;
                in[c][j + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < MP; j++) in[c][j] = on ? f[lane * MP + j] : 0.0;
        }
    }
    W::template compute<MP, 2>(a, D, lane, in, out);
    if (!on) return;
    if constexpr (W::PAIR && MP % 2 == 0) {
#pragma unroll
        for (int j = 0; j < MP; j += 2) {
            double v[2][W::NO], x[2][W::NI];
#pragma unroll
            for (int h = 0; h < 2; h++) {
#pragma unroll
                for (int o = 0; o < W::NO; o++) v[h][o] = out[o][j + h];
#pragma unroll
                for (int c = 0; c < W::NI; c++) x[h][c] = in[c][j + h];
            }
            W::store_pair(a, g, base + lane * MP + j, v, x);
        }
    } else {
#pragma unroll
        for (int j = 0; j < MP; j++) {
            double v[W::NO], x[W::NI];
#pragma unroll
            for (int o = 0; o < W::NO; o++) v[o] = out[o][j];
#pragma unroll
            for (int c = 0; c < W::NI; c++) x[c] = in[c][j];
            W::store_point(a, g, base + lane * MP + j, v, x);
        }
    }
}

// Lines along i (DIR 0) or j (DIR 1): a block stages TL consecutive k
// lines through LDS; each wavefront solves TL / 4 of them.
template <int DIR, int MP, class W>
__global__ void __launch_bounds__((tile_threads<MP, W>())) k_cmp_pass_t(Grid g, CmpDir D, typename W::Args a)
{
    constexpr int TL = tile_lines<MP, W>();
    constexpr int NT = tile_threads<MP, W>(), NW = NT / 64;
    constexpr int NB = W::NI > W::NO ? W::NI : W::NO;
    constexpr int NMAX = 64 * MP;
    __shared__ double sh[NB * TL * NMAX];
    const int N = D.N, P = D.P;
    const int nkt = (g.s[2] + TL - 1) / TL;
    const int fixed = blockIdx.x / nkt;  // i for DIR 1, j for DIR 0
    const int k0 = (blockIdx.x % nkt) * TL;
    const int64_t lstride = DIR == 0 ? (int64_t)g.s[1] * g.s[2] : g.s[2];  // step along the line
    const int64_t base = (DIR == 0 ? (int64_t)fixed * g.s[2] : (int64_t)fixed * g.s[1] * g.s[2]) + k0;
    // element (c, kk, q) sits in row (c TL + kk) MP + q % MP of P doubles at
    // column q / MP rotated by rot(row): the solve reads whole rows (any
    // rotation keeps that free of bank conflicts), and the staging copies,
    // whose lanes run over kk and q, hit distinct banks
    const FastDiv dP((uint32_t)P), dNT((uint32_t)(N * TL));
    auto rot = [&](int row) { return (int)dP.mod((uint32_t)(2 * row + ((row >> 5) & 1))); };
    auto col = [&](int row, int c0) {
        const int v = c0 + rot(row);
        return v >= P ? v - P : v;
    };
    auto slot = [&](int c, int kk, int q) {
        const int row = (c * TL + kk) * MP + q % MP;
        return (c * TL + kk) * NMAX + (q % MP) * P + col(row, q / MP);
    };
    // one-field passes: a lane's loads go out in groups of GS, all in
    // flight before the group's LDS writes (a loop that writes each value as
    // it lands keeps one load in flight); the two-field models, whose
    // registers are spoken for, keep the plain loop
    constexpr int GS = 8;
    constexpr bool GROUPED = W::NI == 1;
    if constexpr (GROUPED) {
        constexpr int PIN = (W::NI * NMAX * TL + NT - 1) / NT;
#pragma unroll
        for (int r0 = 0; r0 < PIN; r0 += GS) {
            double v[GS];
#pragma unroll
            for (int u = 0; u < GS; u++) {
                const int e = threadIdx.x + (r0 + u) * NT;
                const int rem = (int)dNT.mod(e), q = rem / TL, kk = rem % TL;
                v[u] = 0.0;
                if (r0 + u < PIN && e < W::NI * N * TL && k0 + kk < g.s[2])
                    v[u] = W::input(a, g, (int)dNT.div(e))[base + q * lstride + kk];
            }
#pragma unroll
            for (int u = 0; u < GS; u++) {
                const int e = threadIdx.x + (r0 + u) * NT;
                const int rem = (int)dNT.mod(e);
                if (r0 + u < PIN && e < W::NI * N * TL) sh[slot((int)dNT.div(e), rem % TL, rem / TL)] = v[u];
            }
        }
    } else {
        // the same elements as e = tid + r NT (c = e / (N TL), q = e % (N TL)
        // / TL, kk = e % TL), walked incrementally: kk stays tid % TL, q steps
        // by NT / TL (a multiple of MP, so q % MP and the rotation of the
        // element's LDS row change only with c) and wraps past N into the
        // next component
        constexpr int QS = NT / TL;
        static_assert(QS % MP == 0, "the walk keeps q % MP");
        const int kk = threadIdx.x % TL;
        const bool kin = k0 + kk < g.s[2];
        int q = threadIdx.x / TL, c = 0;
        while (c < W::NI && q >= N) {
            q -= N;
            c++;
        }
        while (c < W::NI) {
            const int row = (c * TL + kk) * MP + q % MP, r = rot(row);
            double *const srow = sh + (c * TL + kk) * NMAX + (q % MP) * P;
            const double *src = W::input(a, g, c) + base + kk + q * lstride;
            // GS elements of the line at a time: their loads all in flight
            // before the group's LDS writes; the walk then goes on from the
            // first element past the line (into the next component)
            const int qe = q < N ? q + (N - q + QS - 1) / QS * QS : q;
            for (; q < N; q += GS * QS, src += GS * QS * lstride) {
                double v[GS];
#pragma unroll
                for (int u = 0; u < GS; u++) v[u] = kin && q + u * QS < N ? src[u * QS * lstride] : 0.0;
#pragma unroll
                for (int u = 0; u < GS; u++) {
                    int cc = (q + u * QS) / MP + r;
                    cc -= cc >= P ? P : 0;
                    if (q + u * QS < N) srow[cc] = v[u];
                }
            }
            q = qe - N;
            c++;
            while (c < W::NI && q >= N) {
                q -= N;
                c++;
            }
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool on = lane < P;
    for (int kk = wave; kk < TL; kk += NW) {
        double in[W::NI][MP], out[W::NO][MP];
        // unconditional reads (lanes past P read lane 0's place), then a select
        const int ln = on ? lane : 0;
#pragma unroll
        for (int c = 0; c < W::NI; c++)
#pragma unroll
            for (int j = 0; j < MP; j++) {
                const double v = sh[(c * TL + kk) * NMAX + j * P + col((c * TL + kk) * MP + j, ln)];
                in[c][j] = on ? v : 0.0;
            }
        W::template compute<MP, DIR>(a, D, lane, in, out);
        if (on)
#pragma unroll
            for (int o = 0; o < W::NO; o++)
#pragma unroll
                for (int j = 0; j < MP; j++) sh[(o * TL + kk) * NMAX + j * P + col((o * TL + kk) * MP + j, lane)] = out[o][j];
    }
    __syncthreads();
    // the operands the stores read, a group loaded, then its results written
    constexpr int POUT = (W::NO * NMAX * TL + NT - 1) / NT;
    if constexpr (!GROUPED) {
        // the staging loop's walk over (o, q) with kk fixed
        constexpr int QS = NT / TL;
        const int kk = threadIdx.x % TL;
        if (k0 + kk >= g.s[2]) return;  // no barriers below
        int q = threadIdx.x / TL, o = 0;
        while (o < W::NO && q >= N) {
            q -= N;
            o++;
        }
        while (o < W::NO) {
            const int row = (o * TL + kk) * MP + q % MP, r = rot(row);
            const double *const srow = sh + (o * TL + kk) * NMAX + (q % MP) * P;
            int64_t off = base + kk + q * lstride;
            // the operands of GS stores loaded first, then their results
            const int qe = q < N ? q + (N - q + QS - 1) / QS * QS : q;
            for (; q < N; q += GS * QS, off += GS * QS * lstride) {
                double opnd[GS];
#pragma unroll
                for (int u = 0; u < GS; u++)
                    opnd[u] = q + u * QS < N ? W::operand(a, g, o, off + u * QS * lstride) : 0.0;
#pragma unroll
                for (int u = 0; u < GS; u++) {
                    int cc = (q + u * QS) / MP + r;
                    cc -= cc >= P ? P : 0;
                    if (q + u * QS < N) W::store_op(a, g, o, off + u * QS * lstride, srow[cc], opnd[u]);
                }
            }
            q = qe - N;
            o++;
            while (o < W::NO && q >= N) {
                q -= N;
                o++;
            }
        }
        return;
    }
#pragma unroll
    for (int r0 = 0; r0 < POUT; r0 += GS) {
        double opnd[GS];
#pragma unroll
        for (int u = 0; u < GS; u++) {
            const int e = threadIdx.x + (r0 + u) * NT;
            const int rem = (int)dNT.mod(e), q = rem / TL, kk = rem % TL;
            opnd[u] = (r0 + u < POUT && e < W::NO * N * TL && k0 + kk < g.s[2])
                          ? W::operand(a, g, (int)dNT.div(e), base + q * lstride + kk)
                          : 0.0;
        }
#pragma unroll
        for (int u = 0; u < GS; u++) {
            const int e = threadIdx.x + (r0 + u) * NT;
            const int o = (int)dNT.div(e), rem = (int)dNT.mod(e);
            const int q = rem / TL, kk = rem % TL;
            if (r0 + u < POUT && e < W::NO * N * TL && k0 + kk < g.s[2])
                W::store_op(a, g, o, base + q * lstride + kk, sh[slot(o, kk, q)], opnd[u]);
        }
    }
}

template <int MP, class W>
int launch_pass_mp(nova_ctx *ctx, int dir, const CmpDir &D, const typename W::Args &a)
{
    const Grid &g = ctx->g;
    if (dir == 2) {
        const int64_t lines = (int64_t)g.s[0] * g.s[1];
        hipLaunchKernelGGL((k_cmp_pass_k<MP, W>), dim3((unsigned)((lines + 3) / 4)), dim3(256), 0, ctx->stream, g, D,
                           a);
    } else {
        constexpr int TL = tile_lines<MP, W>();
        const int nkt = (g.s[2] + TL - 1) / TL;
        const int fixed = dir == 0 ? g.s[1] : g.s[0];
        constexpr int NT = tile_threads<MP, W>();
        if (dir == 0)
            hipLaunchKernelGGL((k_cmp_pass_t<0, MP, W>), dim3((unsigned)(fixed * nkt)), dim3(NT), 0, ctx->stream, g,
                               D, a);
        else
            hipLaunchKernelGGL((k_cmp_pass_t<1, MP, W>), dim3((unsigned)(fixed * nkt)), dim3(NT), 0, ctx->stream, g,
                               D, a);
    }
    return check_launch("compact pass");
}

template <class W>
int launch_pass(nova_ctx *ctx, int dir, const typename W::Args &a, const char *name)
{
    const Grid &g = ctx->g;
    const int n = g.s[dir];
    const int mp = pick_mp(n);
    if (!mp || n % mp || n / mp > 64 || n < 4) return fail(NOVA_ERR_UNSUPPORTED, "compact line length");
    const CmpDir D = make_dir(n, ctx->d[dir], mp);
    TimedScope ts(ctx, name);
    switch (mp) {
    case 1: return launch_pass_mp<1, W>(ctx, dir, D, a);
    case 2: return launch_pass_mp<2, W>(ctx, dir, D, a);
    case 4: return launch_pass_mp<4, W>(ctx, dir, D, a);
    default: return launch_pass_mp<8, W>(ctx, dir, D, a);
    }
}

}  // namespace

// Host side: the compact scheme is defined on uniform periodic grids.
int compact_supported(nova_ctx *ctx)
{
    const Grid &g = ctx->g;
    if (g.wall[0] || g.wall[1] || g.wall[2] || ctx->desc.hyp)
        return fail(NOVA_ERR_UNSUPPORTED, "the compact scheme needs a uniform periodic grid");
    for (int d = 0; d < 3; d++) {
        const int n = g.s[d], mp = pick_mp(n);
        if (!mp || n % mp || n / mp > 64 || n < 4)
            return fail(NOVA_ERR_UNSUPPORTED, "compact lines need N = MP * P with MP in {1,2,4,8}, P <= 64");
    }
    return NOVA_OK;
}

double compact_kmod(double kh, int order)
{
    const int o = order - 1;
    if (order == 1) return (kA[o] * std::sin(kh) + 0.5 * kB[o] * std::sin(2.0 * kh)) / (1.0 + 2.0 * kAl[o] * std::cos(kh));
    return (2.0 * kA[o] * (1.0 - std::cos(kh)) + 0.5 * kB[o] * (1.0 - std::cos(2.0 * kh))) /
           (1.0 + 2.0 * kAl[o] * std::cos(kh));
}

int compact_deriv(nova_ctx *ctx, const double *f, int nc, int dir, int order, double *out)
{
    if (int rc = compact_supported(ctx)) return rc;
    if (dir < 0 || dir > 2 || (order != 1 && order != 2)) return fail(NOVA_ERR_ARG, "direction 0..2, order 1 or 2");
    for (int c = 0; c < nc; c++) {
        DerivW::Args a{f + c * ctx->g.np, out + c * ctx->g.np, order};
        if (int rc = launch_pass<DerivW>(ctx, dir, a, "compact_deriv")) return rc;
    }
    return NOVA_OK;
}

// One SSP-type stage out = combine(U, U0, dt R(U)) with the compact NS
// right-hand side, accumulated in R over passes i, j and finished in the k
// pass (which has U in registers).
int compact_ns_stage(nova_ctx *ctx, const double *U, const double *U0, double *out, const StageArgs &sa, double nu,
                     double *R)
{
    NsW::Args a{U, U0, R, out, sa, nu, 0};
    if (int rc = launch_pass<NsW>(ctx, 0, a, "compact_ns_i")) return rc;
    a.mode = 1;
    if (int rc = launch_pass<NsW>(ctx, 1, a, "compact_ns_j")) return rc;
    a.mode = 2;
    return launch_pass<NsW>(ctx, 2, a, "compact_ns_k");
}

// Tangent-linear stage: out = combine(V, V0, dt J_R(X) V) (passes i, j, k).
int compact_tlm_stage(nova_ctx *ctx, const double *X, const double *V, const double *V0, double *out,
                      const StageArgs &sa, double nu, double *R)
{
    TlmW::Args a{X, V, V0, R, out, sa, nu, 0};
    if (int rc = launch_pass<TlmW>(ctx, 0, a, "compact_tlm_i")) return rc;
    a.mode = 1;
    if (int rc = launch_pass<TlmW>(ctx, 1, a, "compact_tlm_j")) return rc;
    a.mode = 2;
    return launch_pass<TlmW>(ctx, 2, a, "compact_tlm_k");
}

// Adjoint stage: out = ob (M + dt J_R(X)^T M) + yg Y, acc = acc_a acc + acc_b M
// (R holds four components).
int compact_adj_stage(nova_ctx *ctx, const double *X, const double *M, double *out, double ob, const double *Y,
                      double yg, double *acc, double acc_a, double acc_b, double dt, double nu, double *R)
{
    AdjW::Args a{X, M, Y, acc, R, out, dt, nu, ob, yg, acc_a, acc_b, 0};
    if (int rc = launch_pass<AdjW>(ctx, 0, a, "compact_adj_i")) return rc;
    a.mode = 1;
    if (int rc = launch_pass<AdjW>(ctx, 1, a, "compact_adj_j")) return rc;
    a.mode = 2;
    return launch_pass<AdjW>(ctx, 2, a, "compact_adj_k");
}

// MHD stage: (oU, oB) = combine((U, B), (U0, B0), dt (dU, dB)); R holds six
// components.
int compact_mhd_stage(nova_ctx *ctx, const double *U, const double *B, const double *U0, const double *B0,
                      double *oU, double *oB, const StageArgs &sa, double nu, double eta, double *R)
{
    MhdW::Args a{U, B, U0, B0, R, oU, oB, sa, nu, eta, 0};
    if (int rc = launch_pass<MhdW>(ctx, 0, a, "compact_mhd_i")) return rc;
    a.mode = 1;
    if (int rc = launch_pass<MhdW>(ctx, 1, a, "compact_mhd_j")) return rc;
    a.mode = 2;
    return launch_pass<MhdW>(ctx, 2, a, "compact_mhd_k");
}

// d = (compact div F) * scale
int compact_div(nova_ctx *ctx, const double *F, double scale, double *d)
{
    for (int dir = 0; dir < 3; dir++) {
        DivW::Args a{F + dir * ctx->g.np, d, scale, dir};
        if (int rc = launch_pass<DivW>(ctx, dir, a, "compact_div")) return rc;
    }
    return NOVA_OK;
}

// F -= s (compact grad p)
int compact_grad_correct(nova_ctx *ctx, const double *p, double s, double *F)
{
    for (int dir = 0; dir < 3; dir++) {
        GradW::Args a{p, F + dir * ctx->g.np, s};
        if (int rc = launch_pass<GradW>(ctx, dir, a, "compact_grad_correct")) return rc;
    }
    return NOVA_OK;
}

}  // namespace nova
