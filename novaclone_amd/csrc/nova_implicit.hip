// nova_implicit.hip -- the implicit finite-difference matrix im::fdMatrix
// (src/im/fdMatrix/fdMatrix.hxx) and its Laplacian im::laplacian
// (src/im/laplacian/laplacian.hxx:23-58) on gfx950, for scalar fields.
//
// The matrix is a 7-point operator stored as six neighbour coefficient
// fields al(0..5) (north, south, east, west, top, bottom), the centre ap and
// the right-hand side, all in the reference's padded layout (n_d = N_d + 1 +
// m nodes a direction, k fastest), as is the transported field while the
// matrix solves for it: the solver's line sweeps read the ghost layers and
// the periodic image node as stale copies until the refill that ends each
// iteration, and only the padded layout has them.
//
// fdMatrix::solve is a line Gauss-Seidel: in each of three sweeps (lines
// along j, then i, then k) every line is a Thomas solve whose right-hand
// side reads the four off-line neighbours, two of them (the lower indices of
// the two loop directions) already updated in this sweep.  Line (a, b) of a
// sweep therefore depends on lines (a-1, b) and (a, b-1) only: the lines of
// one anti-diagonal a + b = t are independent, and one launch per diagonal
// reproduces the reference's order exactly.  The Thomas recurrences run in
// place on the line (forward values, then the back substitution), with the
// same operations as tools::tdma (src/Tools/Tools.cpp:43-62), unfused.
#include "nova_internal.hpp"

#include <algorithm>
#include <cmath>

struct nova_fdmatrix {
    nova_ctx *ctx = nullptr;
    int n[3] = {0, 0, 0};  // padded extents
    int64_t np = 0;        // padded points
    double *al = nullptr;  // six coefficient fields, al + q np
    double *ap = nullptr, *rhs = nullptr, *phi = nullptr;
    double *cp = nullptr;   // Thomas scratch: one padded line per concurrent line
    double *part = nullptr; // residual partial sums
    double *hres = nullptr; // pinned host slot
    int nb = 0;
};

namespace nova {
namespace {

constexpr int kFdThreads = 256;

__device__ inline int64_t pix(const int n[3], int i, int j, int k) { return ((int64_t)i * n[1] + j) * n[2] + k; }

struct FdDims {
    int n[3];
    int h;
    int64_t np;
};

FdDims fd_dims(const nova_fdmatrix *m)
{
    FdDims d;
    for (int a = 0; a < 3; a++) d.n[a] = m->n[a];
    d.h = m->ctx->g.h;
    d.np = m->np;
    return d;
}

// im::laplacian: al = mu/d^2 on the interior [h, n-h) (zero elsewhere), ap
// = -2 mu/dx^2 - 2 mu/dy^2 - 2 mu/dz^2; then deferredCorrection
// (fdMatrix.hxx:213-245): rhs = -hoa + sum al phi_nb + ap phi on the
// interior, zero elsewhere.
__global__ void k_fd_build(FdDims d, double cy, double cx, double cz, double cp, const double *__restrict__ hoa,
                           const double *__restrict__ phi, double *__restrict__ al, double *__restrict__ ap,
                           double *__restrict__ rhs)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.np) return;
    const int k = (int)(q % d.n[2]);
    const int j = (int)((q / d.n[2]) % d.n[1]);
    const int i = (int)(q / ((int64_t)d.n[2] * d.n[1]));
    const bool in = i >= d.h && i < d.n[0] - d.h && j >= d.h && j < d.n[1] - d.h && k >= d.h && k < d.n[2] - d.h;
    const double c[6] = {cy, cy, cx, cx, cz, cz};
    for (int l = 0; l < 6; l++) al[l * d.np + q] = in ? c[l] : 0.0;
    ap[q] = in ? cp : 0.0;
    double r = 0.0;
    if (in) {
#pragma clang fp contract(off)
        r = 0.0 + (((((((-hoa[q] + cy * phi[pix(d.n, i, j + 1, k)]) + cy * phi[pix(d.n, i, j - 1, k)]) +
                       cx * phi[pix(d.n, i + 1, j, k)]) +
                      cx * phi[pix(d.n, i - 1, j, k)]) +
                     cz * phi[pix(d.n, i, j, k + 1)]) +
                    cz * phi[pix(d.n, i, j, k - 1)]) +
                   cp * phi[q]);
    }
    rhs[q] = r;
}

// Squared residual rhs - sum al phi_nb - ap phi (fdMatrix.hxx:278-289) over
// the interior, per-block partial sums.
__global__ void __launch_bounds__(kFdThreads) k_fd_res_part(FdDims d, const double *__restrict__ al,
                                                             const double *__restrict__ ap,
                                                             const double *__restrict__ rhs,
                                                             const double *__restrict__ phi, double *__restrict__ part)
{
    const int I = d.n[0] - 2 * d.h, J = d.n[1] - 2 * d.h, K = d.n[2] - 2 * d.h;
    const int64_t tot = (int64_t)I * J * K;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % K) + d.h;
        const int j = (int)((e / K) % J) + d.h;
        const int i = (int)(e / ((int64_t)K * J)) + d.h;
        const int64_t q = pix(d.n, i, j, k);
        double r;
        {
#pragma clang fp contract(off)
            r = rhs[q] - al[q] * phi[pix(d.n, i, j + 1, k)] - al[d.np + q] * phi[pix(d.n, i, j - 1, k)] -
                al[2 * d.np + q] * phi[pix(d.n, i + 1, j, k)] - al[3 * d.np + q] * phi[pix(d.n, i - 1, j, k)] -
                al[4 * d.np + q] * phi[pix(d.n, i, j, k + 1)] - al[5 * d.np + q] * phi[pix(d.n, i, j, k - 1)] -
                ap[q] * phi[q];
            acc += r * r;
        }
    }
    __shared__ double sh[kFdThreads];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = kFdThreads / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kFdThreads) k_fd_res_final(const double *__restrict__ part, int nb,
                                                              double *__restrict__ out)
{
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += kFdThreads) a += part[b];
    __shared__ double sh[kFdThreads];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int st = kFdThreads / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// One anti-diagonal of a sweep.  DIR = line direction (1: j, 0: i, 2: k);
// the two loop directions (A outer, B inner in the reference) are the
// other two, in the reference's order: j sweep (i, k), i sweep (j, k), k
// sweep (i, j).  Line (a, b) with a + b = t, a, b counted from h.
template <int DIR>
__global__ void __launch_bounds__(kFdThreads) k_fd_sweep(FdDims d, int t, const double *__restrict__ al,
                                                          const double *__restrict__ ap,
                                                          const double *__restrict__ rhs, double *__restrict__ phi,
                                                          double *__restrict__ cpbuf)
{
    constexpr int DA = DIR == 0 ? 1 : 0;  // outer loop direction
    constexpr int DB = DIR == 2 ? 1 : 2;  // inner loop direction
    const int NA = d.n[DA] - 2 * d.h, NB = d.n[DB] - 2 * d.h;
    const int a0 = t - (NB - 1) > 0 ? t - (NB - 1) : 0;
    const int a = a0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int amax = t < NA - 1 ? t : NA - 1;
    if (a > amax) return;
    const int b = t - a;
    int idx[3];
    idx[DA] = a + d.h;
    idx[DB] = b + d.h;
    idx[DIR] = 0;
    // coefficient slots: lower / upper neighbour along the line, and the two
    // off-line pairs subtracted in the reference's order
    constexpr int LO = DIR == 1 ? 1 : (DIR == 0 ? 3 : 5);
    constexpr int UP = DIR == 1 ? 0 : (DIR == 0 ? 2 : 4);
    // j sweep: E, W, T, B; i sweep: N, S, T, B; k sweep: E, W, N, S
    constexpr int O1 = DIR == 0 ? 0 : 2, O2 = O1 + 1;
    constexpr int O3 = DIR == 2 ? 0 : 4, O4 = O3 + 1;
    constexpr int D1 = DIR == 0 ? 1 : 0;  // direction of the O1/O2 pair
    constexpr int D3 = DIR == 2 ? 1 : 2;  // direction of the O3/O4 pair
    int st[3] = {d.n[1] * d.n[2], d.n[2], 1};
    const int64_t base = pix(d.n, idx[0], idx[1], idx[2]);
    const int64_t sl = st[DIR];
    const int len = d.n[DIR];
    double *cp = cpbuf + (int64_t)(a - a0) * len;
    double cprev = 0.0, dprev = 0.0;
    for (int r = 0; r < len; r++) {
        const int64_t q = base + r * sl;
        const bool bnd = r < d.h || r >= len - d.h - 1;
        const double av = bnd ? 0.0 : al[LO * d.np + q];
        const double bv = bnd ? 1.0 : ap[q];
        const double cv = bnd ? 0.0 : al[UP * d.np + q];
        double dv;
        {
#pragma clang fp contract(off)
            dv = bnd ? phi[q]
                     : rhs[q] - al[O1 * d.np + q] * phi[q + st[D1]] - al[O2 * d.np + q] * phi[q - st[D1]] -
                           al[O3 * d.np + q] * phi[q + st[D3]] - al[O4 * d.np + q] * phi[q - st[D3]];
        }
        double c, dd;
        {
#pragma clang fp contract(off)
            if (r == 0) {
                c = cv / bv;
                dd = dv / bv;
            } else {
                c = cv / (bv - av * cprev);
                dd = (dv - av * dprev) / (bv - av * cprev);
            }
        }
        cp[r] = c;
        phi[q] = dd;  // the line's own value is read only by its own row, above
        cprev = c;
        dprev = dd;
    }
    double x = dprev;
    for (int r = len - 2; r >= 0; r--) {
        const int64_t q = base + r * sl;
        {
#pragma clang fp contract(off)
            x = phi[q] - cp[r] * x;
        }
        phi[q] = x;
    }
}

// relax (fdMatrix.hxx:248-253): rhs += (1-urf)/urf * (ap phi); ap /= urf,
// on every padded node.
__global__ void k_fd_relax(int64_t np, double s, double urf, double *__restrict__ ap, double *__restrict__ rhs,
                           const double *__restrict__ phi)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    {
#pragma clang fp contract(off)
        rhs[q] = rhs[q] + s * (ap[q] * phi[q]);
        ap[q] = ap[q] / urf;
    }
}

// rhs += f on every padded node (operator==, fdMatrix.hxx:372-382)
__global__ void k_fd_add(int64_t np, const double *__restrict__ f, double *__restrict__ rhs)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < np) rhs[q] = rhs[q] + f[q];
}

// y = y + s x over the matrix's nine fields (operator+= / -=, :327-356) or
// y = s y (unary minus, :394-413)
__global__ void k_fd_axpy(int64_t nt, double s, const double *__restrict__ x, double *__restrict__ y)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nt) return;
    if (x) {
#pragma clang fp contract(off)
        y[q] = y[q] + s * x[q];
    } else {
        y[q] = y[q] * s;
    }
}

dim3 fd_grid(int64_t n) { return dim3((unsigned)((n + kFdThreads - 1) / kFdThreads)); }

// The transported field: stored layout -> padded with the grid's fill.
int fd_load_phi(nova_fdmatrix *m, const double *phi)
{
    return launch_pack(m->ctx, phi, 1, m->phi);
}

int fd_residual_dev(nova_fdmatrix *m, double *out)
{
    const FdDims d = fd_dims(m);
    nova_ctx *ctx = m->ctx;
    hipLaunchKernelGGL(k_fd_res_part, dim3(m->nb), dim3(kFdThreads), 0, ctx->stream, d, m->al, m->ap, m->rhs, m->phi,
                       m->part);
    hipLaunchKernelGGL(k_fd_res_final, dim3(1), dim3(kFdThreads), 0, ctx->stream, m->part, m->nb, m->part + m->nb);
    if (int rc = check_launch("fd_residual")) return rc;
    NOVA_HIP(hipMemcpyAsync(m->hres, m->part + m->nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t n = (int64_t)(m->n[0] - ctx->m) * (m->n[1] - ctx->m) * (m->n[2] - ctx->m);
    *out = std::pow(m->hres[0] / (double)(int)n, 0.5);
    return NOVA_OK;
}

template <int DIR>
int fd_sweep(nova_fdmatrix *m)
{
    const FdDims d = fd_dims(m);
    constexpr int DA = DIR == 0 ? 1 : 0;
    constexpr int DB = DIR == 2 ? 1 : 2;
    const int NA = d.n[DA] - 2 * d.h, NB = d.n[DB] - 2 * d.h;
    for (int t = 0; t <= NA + NB - 2; t++) {
        const int a0 = t - (NB - 1) > 0 ? t - (NB - 1) : 0;
        const int a1 = t < NA - 1 ? t : NA - 1;
        const int cnt = a1 - a0 + 1;
        hipLaunchKernelGGL(k_fd_sweep<DIR>, fd_grid(cnt), dim3(kFdThreads), 0, m->ctx->stream, d, t, m->al, m->ap,
                           m->rhs, m->phi, m->cp);
    }
    return check_launch("fd_sweep");
}

}  // namespace

int fd_laplacian(nova_ctx *ctx, double mu, const double *phi, nova_fdmatrix **out)
{
    auto *m = new nova_fdmatrix();
    m->ctx = ctx;
    for (int a = 0; a < 3; a++) m->n[a] = ctx->pad[a];
    m->np = (int64_t)m->n[0] * m->n[1] * m->n[2];
    int lmax = 0;
    for (int a = 0; a < 3; a++) lmax = m->n[a] > lmax ? m->n[a] : lmax;
    const int64_t I = m->n[0] - 2 * ctx->g.h, J = m->n[1] - 2 * ctx->g.h, K = m->n[2] - 2 * ctx->g.h;
    const int64_t tot = I * J * K;
    m->nb = (int)std::min<int64_t>(1024, (tot + kFdThreads * 8 - 1) / (kFdThreads * 8));
    int64_t maxlines = std::max(std::max(I, J), K);
    auto cleanup = [&](int rc) {
        (void)hipFree(m->al);
        (void)hipFree(m->ap);
        (void)hipFree(m->rhs);
        (void)hipFree(m->phi);
        (void)hipFree(m->cp);
        (void)hipFree(m->part);
        if (m->hres) (void)hipHostFree(m->hres);
        delete m;
        return rc;
    };
    if (hipMalloc(&m->al, sizeof(double) * 6 * m->np) != hipSuccess ||
        hipMalloc(&m->ap, sizeof(double) * m->np) != hipSuccess ||
        hipMalloc(&m->rhs, sizeof(double) * m->np) != hipSuccess ||
        hipMalloc(&m->phi, sizeof(double) * m->np) != hipSuccess ||
        hipMalloc(&m->cp, sizeof(double) * maxlines * lmax) != hipSuccess ||
        hipMalloc(&m->part, sizeof(double) * (m->nb + 1)) != hipSuccess ||
        hipHostMalloc(&m->hres, sizeof(double), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(NOVA_ERR_NOMEM, "fdMatrix buffers"));
    // the explicit Laplacian of the stored field (the deferred correction's
    // high-order part), then both in the padded layout
    double *hoa_c = nullptr, *hoa = nullptr;
    if (hipMalloc(&hoa_c, sizeof(double) * ctx->g.np) != hipSuccess ||
        hipMalloc(&hoa, sizeof(double) * m->np) != hipSuccess) {
        (void)hipFree(hoa_c);
        return cleanup(fail(NOVA_ERR_NOMEM, "fdMatrix work"));
    }
    int rc = launch_laplacian(ctx, mu, phi, 1, hoa_c);
    if (!rc) rc = launch_pack(ctx, hoa_c, 1, hoa);
    if (!rc) rc = fd_load_phi(m, phi);
    if (!rc) {
        // laplacian.hxx:35-44, in its operation order
        const double cy = mu / std::pow(ctx->d[1], 2), cx = mu / std::pow(ctx->d[0], 2),
                     cz = mu / std::pow(ctx->d[2], 2);
        const double cp = -2.0 * mu / std::pow(ctx->d[0], 2) - 2.0 * mu / std::pow(ctx->d[1], 2) -
                          2.0 * mu / std::pow(ctx->d[2], 2);
        hipLaunchKernelGGL(k_fd_build, fd_grid(m->np), dim3(kFdThreads), 0, ctx->stream, fd_dims(m), cy, cx, cz, cp,
                           hoa, m->phi, m->al, m->ap, m->rhs);
        rc = check_launch("fd_build");
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(NOVA_ERR_HIP, "fd_build sync");
    (void)hipFree(hoa_c);
    (void)hipFree(hoa);
    if (rc) return cleanup(rc);
    *out = m;
    return NOVA_OK;
}

int fd_destroy(nova_fdmatrix *m)
{
    (void)hipFree(m->al);
    (void)hipFree(m->ap);
    (void)hipFree(m->rhs);
    (void)hipFree(m->phi);
    (void)hipFree(m->cp);
    (void)hipFree(m->part);
    if (m->hres) (void)hipHostFree(m->hres);
    delete m;
    return NOVA_OK;
}

// operator== (fdMatrix.hxx:372-382): rhs += f (stored layout, padded with
// the grid's fill first, as a reference field carries its ghosts)
int fd_add_source(nova_fdmatrix *m, const double *f)
{
    double *fp = nullptr;
    NOVA_HIP(hipMalloc(&fp, sizeof(double) * m->np));
    int rc = launch_pack(m->ctx, f, 1, fp);
    if (!rc) {
        hipLaunchKernelGGL(k_fd_add, fd_grid(m->np), dim3(kFdThreads), 0, m->ctx->stream, m->np, fp, m->rhs);
        rc = check_launch("fd_add_source");
    }
    if (!rc && hipStreamSynchronize(m->ctx->stream) != hipSuccess) rc = fail(NOVA_ERR_HIP, "fd sync");
    (void)hipFree(fp);
    return rc;
}

int fd_relax(nova_fdmatrix *m, const double *phi, double urf)
{
    if (int rc = fd_load_phi(m, phi)) return rc;
    hipLaunchKernelGGL(k_fd_relax, fd_grid(m->np), dim3(kFdThreads), 0, m->ctx->stream, m->np, (1.0 - urf) / urf, urf,
                       m->ap, m->rhs, m->phi);
    return check_launch("fd_relax");
}

// setReference (fdMatrix.hxx:256-268), padded node indices
int fd_set_reference(nova_fdmatrix *m, int i, int j, int k, double val)
{
    const int64_t q = ((int64_t)i * m->n[1] + j) * m->n[2] + k;
    double v[2];
    NOVA_HIP(hipMemcpyAsync(&v[0], m->rhs + q, sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
    NOVA_HIP(hipMemcpyAsync(&v[1], m->ap + q, sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
    NOVA_HIP(hipStreamSynchronize(m->ctx->stream));
    v[0] += 1e15 * val;
    v[1] *= 1e15;
    NOVA_HIP(hipMemcpyAsync(m->rhs + q, &v[0], sizeof(double), hipMemcpyHostToDevice, m->ctx->stream));
    NOVA_HIP(hipMemcpyAsync(m->ap + q, &v[1], sizeof(double), hipMemcpyHostToDevice, m->ctx->stream));
    NOVA_HIP(hipStreamSynchronize(m->ctx->stream));
    return NOVA_OK;
}

// operator+= (s = 1), operator-= (s = -1) of another matrix on the same
// grid, or unary minus (other = null, s = -1)
int fd_combine(nova_fdmatrix *m, const nova_fdmatrix *o, double s)
{
    struct Pair {
        const double *x;
        double *y;
        int64_t n;
    };
    const Pair ps[3] = {{o ? o->al : nullptr, m->al, 6 * m->np}, {o ? o->ap : nullptr, m->ap, m->np},
                        {o ? o->rhs : nullptr, m->rhs, m->np}};
    for (const Pair &p : ps)
        hipLaunchKernelGGL(k_fd_axpy, fd_grid(p.n), dim3(kFdThreads), 0, m->ctx->stream, p.n, s, p.x, p.y);
    return check_launch("fd_combine");
}

int fd_residual(nova_fdmatrix *m, const double *phi, double *res)
{
    if (int rc = fd_load_phi(m, phi)) return rc;
    return fd_residual_dev(m, res);
}

// solve (fdMatrix.hxx:59-209): until the residual is at most tol or maxIts
// iterations: j sweeps, i sweeps, k sweeps, then the ghost and image refill.
int fd_solve(nova_fdmatrix *m, double *phi, double tol, int max_its, int *its_out, double *res_out)
{
    if (int rc = fd_load_phi(m, phi)) return rc;
    int its = 0;
    double res = 0.0;
    while (its < max_its) {
        if (int rc = fd_residual_dev(m, &res)) return rc;
        if (!(res > tol)) break;
        its++;
        if (int rc = fd_sweep<1>(m)) return rc;
        if (int rc = fd_sweep<0>(m)) return rc;
        if (int rc = fd_sweep<2>(m)) return rc;
        // correctBoundaryConditions: stored nodes back, then the fill
        if (int rc = launch_unpack(m->ctx, m->phi, 1, phi)) return rc;
        if (int rc = launch_pack(m->ctx, phi, 1, m->phi)) return rc;
    }
    if (int rc = launch_unpack(m->ctx, m->phi, 1, phi)) return rc;
    NOVA_HIP(hipStreamSynchronize(m->ctx->stream));
    if (its_out) *its_out = its;
    if (res_out) *res_out = res;
    return NOVA_OK;
}

int fd_same_grid(const nova_fdmatrix *a, const nova_fdmatrix *b)
{
    return a->ctx == b->ctx;
}

int fd_inside(const nova_fdmatrix *m, int i, int j, int k)
{
    return i >= 0 && i < m->n[0] && j >= 0 && j < m->n[1] && k >= 0 && k < m->n[2];
}

int fd_get(const nova_fdmatrix *m, double *dst, int what)
{
    const double *src = what < 6 ? m->al + what * m->np : (what == 6 ? m->ap : (what == 7 ? m->rhs : m->phi));
    NOVA_HIP(hipMemcpyAsync(dst, src, sizeof(double) * m->np, hipMemcpyDeviceToDevice, m->ctx->stream));
    return NOVA_OK;
}

}  // namespace nova
