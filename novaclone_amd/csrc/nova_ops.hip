// nova_ops.hip -- operator-level kernels of the NOVA hot path on gfx950.
//
// These are the drop-in counterparts of the reference's src/ex operators,
// its BC fills and its diagnostics, on the compact layout.  The operators
// the reference apps call every stage (ex::div of a product, ex::div,
// ex::grad, ex::laplacian) are streamed kernels (k_sop_stream: LDS plane
// tiles and an i register queue); the others run one thread per output
// node, with stencil neighbours loaded through the boundary index maps, so
// the k-neighbours of a 64-wide wave share cache lines and the j/i
// neighbours hit L2.  The fused time-step kernels (nova_step.hip) carry the
// bandwidth-critical work of the library's own time loops.
#include <mutex>
#include <string>

#include "nova_internal.hpp"
#include "nova_pipe.hpp"

namespace nova {

namespace {

__device__ inline int64_t lin(const Grid &g, int i, int j, int k)
{
    return ((int64_t)i * g.s[1] + j) * g.s[2] + k;
}

// Load component c of an nc-component field at a possibly out-of-range
// node, applying the boundary maps.
__device__ inline double ld(const double *__restrict__ f, const Grid &g, int nc, int c, int i,
                            int j, int k)
{
    int fi, fj, fk;
    int ii = bmap(i, g.n[0], g.wall[0], fi);
    int jj = bmap(j, g.n[1], g.wall[1], fj);
    int kk = bmap(k, g.n[2], g.wall[2], fk);
    double s = bsign(g, nc, c, 0, fi) * bsign(g, nc, c, 1, fj) * bsign(g, nc, c, 2, fk);
    return s * f[c * g.np + lin(g, ii, jj, kk)];
}

__device__ inline void offs(int d, int q, int &di, int &dj, int &dk)
{
    di = d == 0 ? q : 0;
    dj = d == 1 ? q : 0;
    dk = d == 2 ? q : 0;
}

// ---- streamed operators: ex::div(U, phi), ex::div(U), ex::grad(p),
// ex::laplacian(nu, f) ----
//
// Each is one pass of the plane-streaming pipeline (nova_pipe.hpp): LDS
// plane tiles for the j and k neighbours, an i register queue, the loads of
// the next two planes in flight, the boundary maps evaluated once per column
// and per plane.  A component is queued or staged only where the operator
// reads its neighbours: div(U, phi) queues U_x and stages U_y, U_z (each
// velocity component is differenced along its own direction only).
// Each operator's arithmetic is written term for term as the reference's
// (cited at each case), in the same order for every node.
enum { SOP_DIV_ADV = 0, SOP_LAPLACIAN = 1, SOP_DIV = 2, SOP_GRAD = 3, SOP_CURL_LIT = 4, SOP_CURL = 5 };

// components: DIV_ADV U_x, U_y, U_z then phi's (none when phi is U);
// LAPLACIAN f's; DIV and CURL U's; GRAD p
template <int OP, int NCP, bool SHARED>
constexpr int sop_staged()
{
    return OP == SOP_DIV_ADV ? (SHARED ? 3 : 3 + NCP)
                             : (OP == SOP_LAPLACIAN ? NCP : (OP == SOP_GRAD ? 1 : 3));
}
// queued (i-neighbours) and staged (j, k neighbours) components: the curl
// differences U_y and U_z along i, each component along the two directions
// other than its own
template <int OP, int NCP, bool SHARED>
constexpr unsigned sop_qm()
{
    constexpr unsigned all = (1u << sop_staged<OP, NCP, SHARED>()) - 1u;
    return (OP == SOP_DIV_ADV && !SHARED) ? (1u | (((1u << NCP) - 1u) << 3))
                                          : (OP == SOP_DIV ? 1u : ((OP == SOP_CURL_LIT || OP == SOP_CURL) ? 6u : all));
}
template <int OP, int NCP, bool SHARED>
constexpr unsigned sop_lm()
{
    constexpr unsigned all = (1u << sop_staged<OP, NCP, SHARED>()) - 1u;
    return (OP == SOP_DIV_ADV && !SHARED) ? (6u | (((1u << NCP) - 1u) << 3)) : (OP == SOP_DIV ? 6u : all);
}

struct SopArgs {
    double *out;
    double nu;     // LAPLACIAN
    double scale;  // DIV
};

// tile rows: 64 x 8 tiles (one halo slot per thread: 1.84 loads per node
// and component against 2.59 on 64 x 4) measured 4-8% faster for div and
// grad at 256^3 and the same for the Laplacian; the curl, 1% slower there,
// and div(U, phi) at H = 4, which spills within 256 registers, keep 64 x 4
template <int H, int OP, int NCP, bool SHARED>
constexpr int sop_kj()
{
    return (OP == SOP_CURL_LIT || OP == SOP_CURL || (H >= 4 && OP == SOP_DIV_ADV && !SHARED && NCP == 3)) ? 4 : 8;
}
template <int H, int OP, int NCP, bool SHARED, bool WALLS>
__global__ void __launch_bounds__(pipe::threads(sop_kj<H, OP, NCP, SHARED>()))
    __attribute__((amdgpu_waves_per_eu(OP == SOP_DIV_ADV && !SHARED && NCP == 3 && H <= 3 ? 2 : 1)))
    k_sop_stream(Grid g, pipe::Launch L, pipe::Src<sop_staged<OP, NCP, SHARED>()> S, SopArgs a)
{
    constexpr int NS = sop_staged<OP, NCP, SHARED>();
    // planes of loads in flight: the Laplacian with one (353 -> 330 us for a
    // vector at 256^3, its registers allowing more waves), the others two
    constexpr int PD = OP == SOP_LAPLACIAN ? 1 : 2;
    pipe::run<H, NS, sop_qm<OP, NCP, SHARED>(), sop_lm<OP, NCP, SHARED>(), WALLS, sop_kj<H, OP, NCP, SHARED>(), PD>(
        g, L, S, [&](int, int, int, int64_t q0, const auto &nb, const pipe::ZK &z) {
            if constexpr (OP == SOP_DIV_ADV) {
                // src/ex/div/div.hxx:101-169: the mean of the conservative
                // and the expanded forms, per direction and stencil offset
                // (U_d phi)^+ - (U_d phi)^- + U_d (phi^+ - phi^-) + phi (U_d^+ - U_d^-)
                const double zi = z.zi;
#pragma unroll
                for (int c = 0; c < NCP; c++) {
                    const int sp = SHARED ? c : 3 + c;
                    const double pc = nb(sp, 0, 0);
                    double acc = 0.0;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double u0 = nb(d, d, 0);
                        double ad = 0.0;
#pragma unroll
                        for (int l = 0; l < H; l++) {
                            const double up = nb(d, d, l + 1), um = nb(d, d, -l - 1);
                            const double pp = nb(sp, d, l + 1), pm = nb(sp, d, -l - 1);
                            ad += g.a[d][l] * ((pp * up - pm * um) + u0 * (pp - pm) + pc * (up - um));
                        }
                        acc += d == 2 ? ad * zi : ad;
                    }
                    __builtin_nontemporal_store(0.5 * acc, a.out + c * g.np + q0);
                }
            } else if constexpr (OP == SOP_LAPLACIAN) {
                // src/ex/laplacian/laplacian.hxx:140-188: the second
                // derivatives, z through 1/dzdxi^2, minus nu d2zdxi2/dzdxi^3
                // times the z first derivative
#pragma unroll
                for (int c = 0; c < NCP; c++) {
                    const double f0 = nb(c, 0, 0);
                    double acc = 0.0, dz1 = 0.0;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        double ad = 0.0;
#pragma unroll
                        for (int l = 0; l < H; l++) {
                            const double fp = nb(c, d, l + 1), fm = nb(c, d, -l - 1);
                            ad += g.b[d][l] * ((fp + fm) - 2.0 * f0);
                            if (d == 2) dz1 += g.a[2][l] * (fp - fm);
                        }
                        acc += d == 2 ? ad * z.zl : ad;
                    }
                    __builtin_nontemporal_store(a.nu * acc - a.nu * z.zm * dz1, a.out + c * g.np + q0);
                }
            } else if constexpr (OP == SOP_DIV) {
                // src/ex/div/div.hxx:37-77
                double acc = 0.0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    double ad = 0.0;
#pragma unroll
                    for (int l = 0; l < H; l++) ad += g.a[d][l] * (nb(d, d, l + 1) - nb(d, d, -l - 1));
                    acc += d == 2 ? ad * z.zi : ad;
                }
                __builtin_nontemporal_store(acc * a.scale, a.out + q0);
            } else if constexpr (OP == SOP_GRAD) {
                // src/ex/grad/grad.cpp:33-58
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    double ad = 0.0;
#pragma unroll
                    for (int l = 0; l < H; l++) ad += g.a[d][l] * (nb(0, d, l + 1) - nb(0, d, -l - 1));
                    __builtin_nontemporal_store(d == 2 ? ad * z.zi : ad, a.out + d * g.np + q0);
                }
            } else {
                // curl: literal = antisymmetric part of the half-shifted
                // gradient (src/ex/grad/grad.cpp:73-98, src/ex/curl/curl.cpp:38-53,
                // offsets +l and -(l + 1)); otherwise the central stencil
                double gr[3][3];
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        if (c == d) continue;
                        double ad = 0.0;
#pragma unroll
                        for (int l = 0; l < H; l++) {
                            if constexpr (OP == SOP_CURL_LIT) ad += g.al[d][l] * (nb(c, d, l) - nb(c, d, -l - 1));
                            else ad += g.a[d][l] * (nb(c, d, l + 1) - nb(c, d, -l - 1));
                        }
                        gr[c][d] = d == 2 ? ad * z.zi : ad;
                    }
                __builtin_nontemporal_store(gr[2][1] - gr[1][2], a.out + q0);
                __builtin_nontemporal_store(gr[0][2] - gr[2][0], a.out + g.np + q0);
                __builtin_nontemporal_store(gr[1][0] - gr[0][1], a.out + 2 * g.np + q0);
            }
        });
}

// ex::laplacian(Field<scalar> nu, f) (src/ex/laplacian/laplacian.hxx:64-127):
// per l each neighbour weighted by nu on its near side (nu(j+l) with
// f(j+1+l), nu(j-1-l) with f(j-1-l)), the centre by minus the pair's sum,
// the z terms through 1/dzdxi^2; then, l by l, minus the z first
// derivative times nu d2zdxi2/dzdxi^3.  nu is an even scalar at walls.
__global__ void k_laplacian_var(Grid g, const double *__restrict__ nu, const double *__restrict__ f, int nc,
                                double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.s[2] || j >= g.s[1]) return;
    const double n0 = ld(nu, g, 1, 0, i, j, k);
    for (int c = 0; c < nc; c++) {
        const double f0 = ld(f, g, nc, c, i, j, k);
        double acc = 0.0;
        for (int l = 0; l < g.h; l++) {
            double t = 0.0, cP = 0.0;
            for (int d = 0; d < 3; d++) {
                int di, dj, dk;
                offs(d, 1, di, dj, dk);
                const double np = ld(nu, g, 1, 0, i + l * di, j + l * dj, k + l * dk);
                const double nm = ld(nu, g, 1, 0, i - (1 + l) * di, j - (1 + l) * dj, k - (1 + l) * dk);
                const double fp = ld(f, g, nc, c, i + (1 + l) * di, j + (1 + l) * dj, k + (1 + l) * dk);
                const double fm = ld(f, g, nc, c, i - (1 + l) * di, j - (1 + l) * dj, k - (1 + l) * dk);
                const double w = d == 2 ? g.b[2][l] * g.zlap[k] : g.b[d][l];
                t += np * w * fp + nm * w * fm;
                cP -= (np + nm) * w;
            }
            acc += t + cP * f0;
        }
        for (int l = 0; l < g.h; l++) {
            const double fT = ld(f, g, nc, c, i, j, k + 1 + l), fB = ld(f, g, nc, c, i, j, k - 1 - l);
            acc -= (g.a[2][l] * fT - g.a[2][l] * fB) * n0 * g.zmet[k];
        }
        out[c * g.np + lin(g, i, j, k)] = acc;
    }
}

__global__ void k_axpby(size_t n, double a, const double *__restrict__ x, double b,
                        const double *__restrict__ y, double *__restrict__ out)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        out[q] = a * x[q] + (y ? b * y[q] : 0.0);
}

// compact SoA -> padded AoS with every ghost and the duplicated endpoint
// filled through the boundary maps.
__global__ void k_pack(Grid g, int h, int p0, int p1, int p2, const double *__restrict__ cf, int nc,
                       double *__restrict__ pad)
{
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int I = blockIdx.z;
    if (K >= p2 || J >= p1 || I >= p0) return;
    const size_t o = (((size_t)I * p1 + J) * p2 + K) * nc;
    for (int c = 0; c < nc; c++) pad[o + c] = ld(cf, g, nc, c, I - h, J - h, K - h);
}

__global__ void k_unpack(Grid g, int h, int p1, int p2, const double *__restrict__ pad, int nc,
                         double *__restrict__ cf)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.s[2] || j >= g.s[1]) return;
    const size_t o = (((size_t)(i + h) * p1 + (j + h)) * p2 + (k + h)) * nc;
    for (int c = 0; c < nc; c++) cf[c * g.np + lin(g, i, j, k)] = pad[o + c];
}

// One entry of the reference's BC list on a padded AoS field: periodicBC
// (src/BC/periodicBC.hxx:1-66), dirichletBC (src/BC/dirichletBC.hxx),
// neumannBC (src/BC/neumannBC.hxx) and the pure reflections.
__global__ void k_bc_entry(int p0, int p1, int p2, int h, double hd, double *__restrict__ f, int nc,
                           int axis, int low, int kind, int comp, double val)
{
#pragma clang fp contract(off)
    const int a = blockIdx.y * blockDim.y + threadIdx.y;  // first transverse index
    const int b = blockIdx.x * blockDim.x + threadIdx.x;  // second transverse index
    const int na = axis == 0 ? p1 : p0, nb = axis == 2 ? p1 : p2;
    if (a >= na || b >= nb) return;
    size_t base, stride;
    int n;
    if (axis == 0) { base = ((size_t)0 * p1 + a) * p2 + b; stride = (size_t)p1 * p2; n = p0; }
    else if (axis == 1) { base = ((size_t)a * p1) * p2 + b; stride = (size_t)p2; n = p1; }
    else { base = ((size_t)a * p1 + b) * p2; stride = 1; n = p2; }
    const int m = 2 * h;
#define V(q, c) f[(base + (size_t)(q) * stride) * nc + (c)]
    if (kind == NOVA_PBC_PERIODIC) {
        if (!low) return;  // periodicBC acts from the east/south/bottom entries only
        for (int q = 0; q < h; q++)
            for (int c = 0; c < nc; c++) V(q, c) = V(q + n - m - 1, c);
        for (int q = n - h - 1; q < n; q++)
            for (int c = 0; c < nc; c++) V(q, c) = V(q - n + m + 1, c);
        return;
    }
    const int w = low ? h : n - h - 1;
    const int s = low ? -1 : 1;
    if (kind == NOVA_PBC_DIRICHLET) {
        V(w, comp) = val;
        for (int l = 0; l < h; l++) V(w + s * (l + 1), comp) = 2.0 * (l + 1) * val - V(w - s * (l + 1), comp);
    } else if (kind == NOVA_PBC_NEUMANN) {
        if (low) V(w, comp) = V(w + 1, comp) - val * hd;
        else V(w, comp) = V(w - 1, comp) + val * hd;
        for (int l = 0; l < h; l++) V(w + s * (l + 1), comp) = V(w - s * (l + 1), comp) - 2.0 * (l + 1) * val * hd;
    } else if (kind == NOVA_PBC_EVEN) {
        for (int l = 0; l < h; l++) V(w + s * (l + 1), comp) = V(w - s * (l + 1), comp);
    } else if (kind == NOVA_PBC_ODD) {
        V(w, comp) = 0.0;
        for (int l = 0; l < h; l++) V(w + s * (l + 1), comp) = -V(w - s * (l + 1), comp);
    }
#undef V
}

// ---- deterministic reductions: per-block partials, then one block ----
constexpr int kRedThreads = 256;

template <int NQ>
__device__ inline void block_reduce(double (&v)[NQ], const int *is_max, double *__restrict__ part)
{
    __shared__ double sh[NQ][kRedThreads];
    for (int q = 0; q < NQ; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int st = kRedThreads / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int q = 0; q < NQ; q++) {
                double o = sh[q][threadIdx.x + st];
                sh[q][threadIdx.x] = is_max[q] ? fmax(sh[q][threadIdx.x], o) : sh[q][threadIdx.x] + o;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < NQ; q++) part[q * gridDim.x + blockIdx.x] = sh[q][0];
}

template <int NQ>
__global__ void k_red_final(const double *__restrict__ part, int nb, double *__restrict__ out,
                            int m0, int m1, int m2, int m3, int m4, int m5, int m6, int m7)
{
    const int is_max[8] = {m0, m1, m2, m3, m4, m5, m6, m7};
    double v[NQ];
    for (int q = 0; q < NQ; q++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nb; b += kRedThreads)
            a = is_max[q] ? fmax(a, part[q * nb + b]) : a + part[q * nb + b];
        v[q] = a;
    }
    __shared__ double sh[NQ][kRedThreads];
    for (int q = 0; q < NQ; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int st = kRedThreads / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int q = 0; q < NQ; q++) {
                double o = sh[q][threadIdx.x + st];
                sh[q][threadIdx.x] = is_max[q] ? fmax(sh[q][threadIdx.x], o) : sh[q][threadIdx.x] + o;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < NQ; q++) out[q] = sh[q][0];
}

// Sum of 0.5|U|^2 over nodes [0, n_d) (app/TG_Piosson/TG.cpp:151-161).
__global__ void k_ke_part(Grid g, const double *__restrict__ U, double *__restrict__ part)
{
    const int64_t tot = (int64_t)g.n[0] * g.n[1] * g.n[2];
    double a = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(q % g.n[2]);
        const int64_t r = q / g.n[2];
        const int j = (int)(r % g.n[1]);
        const int i = (int)(r / g.n[1]);
        const int64_t o = lin(g, i, j, k);
        const double u = U[o], v = U[g.np + o], w = U[2 * g.np + o];
        a += 0.5 * (u * u + v * v + w * w);
    }
    double v1[1] = {a};
    const int mx[1] = {0};
    block_reduce<1>(v1, mx, part);
}

// nu * mean over nodes [0, n_d) of sum_{c,d} (D_d U_c)^2 with the central
// stencil of src/ex/grad/grad.cpp:33-58 applied to each component: the
// viscous dissipation rate of the velocity field.
__global__ void k_diss_part(Grid g, const double *__restrict__ U, double *__restrict__ part)
{
    const int64_t tot = (int64_t)g.n[0] * g.n[1] * g.n[2];
    double a = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(q % g.n[2]);
        const int64_t r = q / g.n[2];
        const int j = (int)(r % g.n[1]);
        const int i = (int)(r / g.n[1]);
        for (int c = 0; c < 3; c++)
            for (int d = 0; d < 3; d++) {
                double ad = 0.0;
                for (int l = 0; l < g.h; l++) {
                    int di, dj, dk;
                    offs(d, l + 1, di, dj, dk);
                    ad += g.a[d][l] * (ld(U, g, 3, c, i + di, j + dj, k + dk) - ld(U, g, 3, c, i - di, j - dj, k - dk));
                }
                if (d == 2) ad *= g.zinv[k];
                a += ad * ad;
            }
    }
    double v1[1] = {a};
    const int mx[1] = {0};
    block_reduce<1>(v1, mx, part);
}

// Weighted inner product sum_nodes w sum_c a_c b_c, w = 1/2 on wall nodes
// (the trapezoidal weight under which the central differences with
// reflecting walls satisfy D^T = -D).
__global__ void k_wdot_part(Grid g, const double *__restrict__ a, const double *__restrict__ b, int nc,
                            double *__restrict__ part)
{
    double acc = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < g.np; q += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(q % g.s[2]);
        const int64_t r = q / g.s[2];
        const int j = (int)(r % g.s[1]);
        const int i = (int)(r / g.s[1]);
        double w = 1.0;
        if (g.wall[0] && (i == 0 || i == g.n[0])) w *= 0.5;
        if (g.wall[1] && (j == 0 || j == g.n[1])) w *= 0.5;
        if (g.wall[2] && (k == 0 || k == g.n[2])) w *= 0.5;
        double v = 0.0;
        for (int c = 0; c < nc; c++) v += a[c * g.np + q] * b[c * g.np + q];
        acc += w * v;
    }
    double v1[1] = {acc};
    const int mx[1] = {0};
    block_reduce<1>(v1, mx, part);
}

__global__ void k_maxabs_part(int64_t n, const double *__restrict__ f, double *__restrict__ part)
{
    double a = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        a = fmax(a, fabs(f[q]));
    double v1[1] = {a};
    const int mx[1] = {1};
    block_reduce<1>(v1, mx, part);
}

// app/OT_Poisson/OT.cpp:151-162 sums and maxima.
__global__ void k_mhd_part(Grid g, const double *__restrict__ U, const double *__restrict__ B,
                           const double *__restrict__ J, const double *__restrict__ W,
                           double *__restrict__ part)
{
    const int64_t tot = (int64_t)g.n[0] * g.n[1] * g.n[2];
    double ek = 0, em = 0, jmax = 0, umax = 0, om = 0, jav = 0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(q % g.n[2]);
        const int64_t r = q / g.n[2];
        const int j = (int)(r % g.n[1]);
        const int i = (int)(r / g.n[1]);
        const int64_t o = lin(g, i, j, k);
        const double u0 = U[o], u1 = U[g.np + o], u2 = U[2 * g.np + o];
        const double b0 = B[o], b1 = B[g.np + o], b2 = B[2 * g.np + o];
        const double j0 = J[o], j1 = J[g.np + o], j2 = J[2 * g.np + o];
        const double w0 = W[o], w1 = W[g.np + o], w2 = W[2 * g.np + o];
        ek += 0.5 * (u0 * u0 + u1 * u1 + u2 * u2);
        em += 0.5 * (b0 * b0 + b1 * b1 + b2 * b2);
        const double jm = sqrt(j0 * j0 + j1 * j1 + j2 * j2);
        umax = fmax(umax, sqrt(u0 * u0 + u1 * u1 + u2 * u2));
        jmax = fmax(jmax, jm);
        om += sqrt(w0 * w0 + w1 * w1 + w2 * w2);
        jav += jm;
    }
    double v[6] = {ek, em, jav, jmax, om, umax};
    const int mx[6] = {0, 0, 0, 1, 0, 1};
    block_reduce<6>(v, mx, part);
}

// Batched Thomas solve, one system per thread (src/Tools/Tools.cpp:43-62);
// the cyclic form uses the Sherman-Morrison correction with the same
// forward/backward sweeps.
__global__ void k_tdma(int n, int nb, const double *__restrict__ a, const double *__restrict__ b,
                       const double *__restrict__ c, const double *__restrict__ d,
                       double *__restrict__ x, double *__restrict__ work, int cyclic)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nb) return;
    const double *A = a + (size_t)s * n, *Bd = b + (size_t)s * n, *C = c + (size_t)s * n, *D = d + (size_t)s * n;
    double *X = x + (size_t)s * n;
    double *cp = work + (size_t)s * 3 * n, *dp = cp + n, *zz = dp + n;
    if (!cyclic) {
        cp[0] = C[0] / Bd[0];
        dp[0] = D[0] / Bd[0];
        for (int i = 1; i < n; i++) {
            const double den = Bd[i] - A[i] * cp[i - 1];
            cp[i] = C[i] / den;
            dp[i] = (D[i] - A[i] * dp[i - 1]) / den;
        }
        X[n - 1] = dp[n - 1];
        for (int i = n - 2; i >= 0; i--) X[i] = dp[i] - cp[i] * X[i + 1];
        return;
    }
    // cyclic: A x = d with corners alpha = c[n-1] (row n-1, col 0) and
    // beta = a[0] (row 0, col n-1).  gamma = -b[0].
    const double alpha = C[n - 1], beta = A[0], gamma = -Bd[0];
    // modified diagonal: b0' = b0 - gamma, b_{n-1}' = b_{n-1} - alpha*beta/gamma
    // solve T y = d and T z = u, u = (gamma, 0, ..., 0, alpha)
    for (int pass = 0; pass < 2; pass++) {
        double *out = pass == 0 ? X : zz;
        for (int i = 0; i < n; i++) {
            double bi = Bd[i];
            if (i == 0) bi -= gamma;
            if (i == n - 1) bi -= alpha * beta / gamma;
            const double ai = i == 0 ? 0.0 : A[i];
            const double ci = i == n - 1 ? 0.0 : C[i];
            double ri;
            if (pass == 0) ri = D[i];
            else ri = (i == 0) ? gamma : (i == n - 1 ? alpha : 0.0);
            if (i == 0) {
                cp[0] = ci / bi;
                dp[0] = ri / bi;
            } else {
                const double den = bi - ai * cp[i - 1];
                cp[i] = ci / den;
                dp[i] = (ri - ai * dp[i - 1]) / den;
            }
        }
        out[n - 1] = dp[n - 1];
        for (int i = n - 2; i >= 0; i--) out[i] = dp[i] - cp[i] * out[i + 1];
    }
    const double fact = (X[0] + beta * X[n - 1] / gamma) / (1.0 + zz[0] + beta * zz[n - 1] / gamma);
    for (int i = 0; i < n; i++) X[i] -= fact * zz[i];
}

dim3 pt_block() { return dim3(64, 4, 1); }
dim3 pt_grid(const Grid &g) { return dim3((g.s[2] + 63) / 64, (g.s[1] + 3) / 4, g.s[0]); }

}  // namespace

namespace {

template <int H, int OP, int NCP, bool SHARED>
int launch_sop_h(nova_ctx *ctx, const pipe::Src<sop_staged<OP, NCP, SHARED>()> &S, const SopArgs &a, const char *name)
{
    const Grid &g = ctx->g;
    constexpr int KJ = sop_kj<H, OP, NCP, SHARED>();
    const pipe::Launch L = pipe::tiling<H, KJ>(g);
    if (g.wall[0] || g.wall[1] || g.wall[2])
        hipLaunchKernelGGL((k_sop_stream<H, OP, NCP, SHARED, true>), dim3(pipe::blocks(L)), dim3(pipe::threads(KJ)), 0,
                           ctx->stream, g, L, S, a);
    else
        hipLaunchKernelGGL((k_sop_stream<H, OP, NCP, SHARED, false>), dim3(pipe::blocks(L)), dim3(pipe::threads(KJ)), 0,
                           ctx->stream, g, L, S, a);
    return check_launch(name);
}

template <int OP, int NCP, bool SHARED = false>
int launch_sop(nova_ctx *ctx, const pipe::Src<sop_staged<OP, NCP, SHARED>()> &S, const SopArgs &a, const char *name)
{
    switch (ctx->g.h) {
    case 1: return launch_sop_h<1, OP, NCP, SHARED>(ctx, S, a, name);
    case 2: return launch_sop_h<2, OP, NCP, SHARED>(ctx, S, a, name);
    case 3: return launch_sop_h<3, OP, NCP, SHARED>(ctx, S, a, name);
    case 4: return launch_sop_h<4, OP, NCP, SHARED>(ctx, S, a, name);
    }
    return fail(NOVA_ERR_ARG, "stencil half-width outside 1..4");
}

template <int NCP>
int launch_div_adv_nc(nova_ctx *ctx, const double *U, const double *phi, const SopArgs &a)
{
    pipe::Src<3 + NCP> S;
    for (int d = 0; d < 3; d++) pipe::set_src(ctx->g, S, d, U, 3, d);
    for (int c = 0; c < NCP; c++) pipe::set_src(ctx->g, S, 3 + c, phi, NCP, c);
    return launch_sop<SOP_DIV_ADV, NCP>(ctx, S, a, "div_adv");
}

template <int NCP>
int launch_laplacian_nc(nova_ctx *ctx, const double *f, const SopArgs &a)
{
    pipe::Src<NCP> S;
    for (int c = 0; c < NCP; c++) pipe::set_src(ctx->g, S, c, f, NCP, c);
    return launch_sop<SOP_LAPLACIAN, NCP>(ctx, S, a, "laplacian");
}

}  // namespace

int launch_div_adv(nova_ctx *ctx, const double *U, const double *phi, int nc, double *out)
{
    SopArgs a{};
    a.out = out;
    if (phi == U && nc == 3) {
        pipe::Src<3> S;
        for (int d = 0; d < 3; d++) pipe::set_src(ctx->g, S, d, U, 3, d);
        return launch_sop<SOP_DIV_ADV, 3, true>(ctx, S, a, "div_adv");
    }
    return nc == 3 ? launch_div_adv_nc<3>(ctx, U, phi, a) : launch_div_adv_nc<1>(ctx, U, phi, a);
}

int launch_div(nova_ctx *ctx, const double *U, double scale, double *out)
{
    SopArgs a{};
    a.out = out;
    a.scale = scale;
    pipe::Src<3> S;
    for (int d = 0; d < 3; d++) pipe::set_src(ctx->g, S, d, U, 3, d);
    return launch_sop<SOP_DIV, 3>(ctx, S, a, "div");
}

int launch_grad(nova_ctx *ctx, const double *p, double *out)
{
    SopArgs a{};
    a.out = out;
    pipe::Src<1> S;
    pipe::set_src(ctx->g, S, 0, p, 1, 0);
    return launch_sop<SOP_GRAD, 1>(ctx, S, a, "grad");
}

int launch_laplacian_var(nova_ctx *ctx, const double *nu, const double *f, int nc, double *out)
{
    hipLaunchKernelGGL(k_laplacian_var, pt_grid(ctx->g), pt_block(), 0, ctx->stream, ctx->g, nu, f, nc, out);
    return check_launch("laplacian_var");
}

int launch_laplacian(nova_ctx *ctx, double nu, const double *f, int nc, double *out)
{
    SopArgs a{};
    a.out = out;
    a.nu = nu;
    return nc == 3 ? launch_laplacian_nc<3>(ctx, f, a) : launch_laplacian_nc<1>(ctx, f, a);
}

int launch_curl(nova_ctx *ctx, const double *U, int literal, double *out, int pseudo)
{
    SopArgs a{};
    a.out = out;
    pipe::Src<3> S;
    for (int c = 0; c < 3; c++) {
        pipe::set_src(ctx->g, S, c, U, 3, c);
        // a pseudovector (the input is itself a curl, central stencil only):
        // the wall-normal component even, the others odd
        if (pseudo && !literal) S.flip[c] = 7u & ~(1u << c);
    }
    return literal ? launch_sop<SOP_CURL_LIT, 3>(ctx, S, a, "curl") : launch_sop<SOP_CURL, 3>(ctx, S, a, "curl");
}

int launch_axpby(nova_ctx *ctx, size_t n, double a, const double *x, double b, const double *y, double *out)
{
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return NOVA_OK;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, n, a, x, b, y, out);
    return check_launch("axpby");
}

int launch_pack(nova_ctx *ctx, const double *cf, int nc, double *pad)
{
    const int *p = ctx->pad;
    dim3 grid((p[2] + 63) / 64, (p[1] + 3) / 4, p[0]);
    hipLaunchKernelGGL(k_pack, grid, pt_block(), 0, ctx->stream, ctx->g, ctx->g.h, p[0], p[1], p[2], cf, nc, pad);
    return check_launch("pack");
}

int launch_unpack(nova_ctx *ctx, const double *pad, int nc, double *cf)
{
    hipLaunchKernelGGL(k_unpack, pt_grid(ctx->g), pt_block(), 0, ctx->stream, ctx->g, ctx->g.h, ctx->pad[1],
                       ctx->pad[2], pad, nc, cf);
    return check_launch("unpack");
}

int launch_bc_fill(nova_ctx *ctx, double *pad, int nc, int dir, int kind, int comp, double val)
{
    int axis, low;
    switch (dir) {
    case 2: axis = 0; low = 1; break;  // east
    case 3: axis = 0; low = 0; break;  // west
    case 1: axis = 1; low = 1; break;  // south
    case 0: axis = 1; low = 0; break;  // north
    case 5: axis = 2; low = 1; break;  // bottom
    case 4: axis = 2; low = 0; break;  // top
    default: return fail(NOVA_ERR_ARG, "bc dir out of range");
    }
    if (kind == NOVA_PBC_PERIODIC && !low) return NOVA_OK;
    const int *p = ctx->pad;
    const int na = axis == 0 ? p[1] : p[0], nb = axis == 2 ? p[1] : p[2];
    dim3 grid((nb + 63) / 64, (na + 3) / 4, 1);
    hipLaunchKernelGGL(k_bc_entry, grid, pt_block(), 0, ctx->stream, p[0], p[1], p[2], ctx->g.h, ctx->d[axis],
                       pad, nc, axis, low, kind, comp, val);
    return check_launch("bc_fill");
}

static int red_blocks(int64_t n)
{
    int64_t b = (n + kRedThreads * 4 - 1) / (kRedThreads * 4);
    if (b < 1) b = 1;
    if (b > 2048) b = 2048;
    return (int)b;
}

int reduce_ke(nova_ctx *ctx, const double *U, double *out)
{
    const Grid &g = ctx->g;
    const int64_t tot = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int nb = red_blocks(tot);
    if (int rc = ensure_red(ctx, (size_t)nb * 8 + 8)) return rc;
    hipLaunchKernelGGL(k_ke_part, dim3(nb), dim3(kRedThreads), 0, ctx->stream, g, U, ctx->d_red + 8);
    hipLaunchKernelGGL(k_red_final<1>, dim3(1), dim3(kRedThreads), 0, ctx->stream, ctx->d_red + 8, nb, ctx->d_red,
                       0, 0, 0, 0, 0, 0, 0, 0);
    if (int rc = check_launch("reduce_ke")) return rc;
    NOVA_HIP(hipMemcpyAsync(ctx->h_red, ctx->d_red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->h_red[0] / (double)tot;
    return NOVA_OK;
}

int reduce_dissipation(nova_ctx *ctx, const double *U, double nu, double *out)
{
    const Grid &g = ctx->g;
    const int64_t tot = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int nb = red_blocks(tot);
    if (int rc = ensure_red(ctx, (size_t)nb * 8 + 8)) return rc;
    hipLaunchKernelGGL(k_diss_part, dim3(nb), dim3(kRedThreads), 0, ctx->stream, g, U, ctx->d_red + 8);
    hipLaunchKernelGGL(k_red_final<1>, dim3(1), dim3(kRedThreads), 0, ctx->stream, ctx->d_red + 8, nb, ctx->d_red,
                       0, 0, 0, 0, 0, 0, 0, 0);
    if (int rc = check_launch("reduce_dissipation")) return rc;
    NOVA_HIP(hipMemcpyAsync(ctx->h_red, ctx->d_red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    *out = nu * ctx->h_red[0] / (double)tot;
    return NOVA_OK;
}

int reduce_wdot(nova_ctx *ctx, const double *a, const double *b, int nc, double *out)
{
    const int nb = red_blocks(ctx->g.np);
    if (int rc = ensure_red(ctx, (size_t)nb * 8 + 8)) return rc;
    hipLaunchKernelGGL(k_wdot_part, dim3(nb), dim3(kRedThreads), 0, ctx->stream, ctx->g, a, b, nc, ctx->d_red + 8);
    hipLaunchKernelGGL(k_red_final<1>, dim3(1), dim3(kRedThreads), 0, ctx->stream, ctx->d_red + 8, nb, ctx->d_red,
                       0, 0, 0, 0, 0, 0, 0, 0);
    if (int rc = check_launch("reduce_wdot")) return rc;
    NOVA_HIP(hipMemcpyAsync(ctx->h_red, ctx->d_red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->h_red[0];
    return NOVA_OK;
}

int reduce_maxabs(nova_ctx *ctx, const double *f, double *out)
{
    const int64_t n = ctx->g.np;
    const int nb = red_blocks(n);
    if (int rc = ensure_red(ctx, (size_t)nb * 8 + 8)) return rc;
    hipLaunchKernelGGL(k_maxabs_part, dim3(nb), dim3(kRedThreads), 0, ctx->stream, n, f, ctx->d_red + 8);
    hipLaunchKernelGGL(k_red_final<1>, dim3(1), dim3(kRedThreads), 0, ctx->stream, ctx->d_red + 8, nb, ctx->d_red,
                       1, 0, 0, 0, 0, 0, 0, 0);
    if (int rc = check_launch("reduce_maxabs")) return rc;
    NOVA_HIP(hipMemcpyAsync(ctx->h_red, ctx->d_red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->h_red[0];
    return NOVA_OK;
}

int reduce_mhd(nova_ctx *ctx, const double *U, const double *B, const double *J, const double *W,
               double mu, double eta, double out[8])
{
    const Grid &g = ctx->g;
    const int64_t tot = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int nb = red_blocks(tot);
    if (int rc = ensure_red(ctx, (size_t)nb * 6 + 8)) return rc;
    hipLaunchKernelGGL(k_mhd_part, dim3(nb), dim3(kRedThreads), 0, ctx->stream, g, U, B, J, W, ctx->d_red + 8);
    hipLaunchKernelGGL(k_red_final<6>, dim3(1), dim3(kRedThreads), 0, ctx->stream, ctx->d_red + 8, nb, ctx->d_red,
                       0, 0, 0, 1, 0, 1, 0, 0);
    if (int rc = check_launch("reduce_mhd")) return rc;
    NOVA_HIP(hipMemcpyAsync(ctx->h_red, ctx->d_red, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    const double n = (double)tot;
    const double ek = ctx->h_red[0] / n, em = ctx->h_red[1] / n, jav = ctx->h_red[2] / n;
    const double om = ctx->h_red[4] / n;
    out[0] = ek;
    out[1] = em;
    out[2] = jav;
    out[3] = ctx->h_red[3];
    out[4] = om;
    out[5] = -mu * (om * om) - eta * (jav * jav);
    out[6] = ctx->h_red[5];
    out[7] = 0.0;
    return NOVA_OK;
}

// Work array of nova_tdma_batched_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends, so that the entry point
// allocates nothing.  Host code never uses the name `scratch` as an
// address: tdma_scratch_init resolves each device's copy with
// hipGetSymbolAddress once, with that device current, and every launch
// passes the resolved pointer.  Calls that overlap on one GPU share it
// and race.
constexpr size_t kTdmaScratchDoubles = size_t(1) << 21;  // 16 MiB: 3 n nbatch doubles at most
static __device__ __align__(256) double scratch[kTdmaScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int tdma_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(NOVA_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return NOVA_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = static_cast<double *>(ptr);
    return NOVA_OK;
}

size_t tdma_scratch_max() { return kTdmaScratchDoubles; }

int launch_tdma_noalloc(nova_ctx *ctx, int n, int nb, const double *a, const double *b, const double *c,
                        const double *d, double *x, int cyclic)
{
    if ((size_t)3 * (size_t)n * (size_t)nb > kTdmaScratchDoubles)
        return fail(NOVA_ERR_ARG, "nova_tdma_batched_noalloc: 3 n nbatch = " + std::to_string(3 * (size_t)n * nb) +
                                      " doubles exceed its workspace of " + std::to_string(kTdmaScratchDoubles));
    double *work = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        work = g_scratch_ptr[ctx->device];
    }
    if (!work) return fail(NOVA_ERR_STATE, "the device's tdma workspace was not resolved at context creation");
    hipLaunchKernelGGL(k_tdma, dim3((nb + 127) / 128), dim3(128), 0, ctx->stream, n, nb, a, b, c, d, x, work, cyclic);
    return check_launch("tdma_noalloc");
}

int launch_tdma(nova_ctx *ctx, int n, int nb, const double *a, const double *b, const double *c,
                const double *d, double *x, int cyclic)
{
    double *work = nullptr;
    NOVA_HIP(hipMallocAsync((void **)&work, sizeof(double) * 3 * (size_t)n * nb, ctx->stream));
    hipLaunchKernelGGL(k_tdma, dim3((nb + 127) / 128), dim3(128), 0, ctx->stream, n, nb, a, b, c, d, x, work, cyclic);
    int rc = check_launch("tdma");
    NOVA_HIP(hipFreeAsync(work, ctx->stream));
    return rc;
}

}  // namespace nova
