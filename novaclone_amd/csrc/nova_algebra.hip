// nova_algebra.hip -- pointwise Field algebra and the CFL reduction.
//
// The reference's out-of-place Field operators (src/Field/Field.hxx:617-1150,
// src/Field/Field.cpp:88-150) loop over every node of the padded array and
// apply the node type's operator (src/Types/vector/vector.cpp:67-160,
// src/Types/tensor/tensor.cpp:237-251).  Here each is one pass over the
// stored nodes (ghosts are never stored: they follow from the boundary
// conditions on load).  Expressions are evaluated with the reference's
// operand order and without fused multiply-adds, so the results equal the
// reference's node for node.
#include <cmath>

#include "nova_internal.hpp"

namespace nova {
namespace {

#pragma clang fp contract(off)

// out = op(a, b) per node; n nodes, component stride np
__global__ void k_field_op(int op, int64_t n, int64_t np, const double *a, int nca, const double *b, int ncb,
                           double c0, double c1, double c2, double *out)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        auto A = [&](int c) { return a[c * np + q]; };
        auto B = [&](int c) { return b[c * np + q]; };
        switch (op) {
        case NOVA_OP_DOT: {  // vector.cpp:120-127 (x x + y y + z z); scalar & scalar = product
            double s = A(0) * B(0);
            for (int c = 1; c < nca; c++) s = s + A(c) * B(c);
            out[q] = s;
            break;
        }
        case NOVA_OP_CROSS:  // vector.cpp:144-156
            out[q] = A(1) * B(2) - A(2) * B(1);
            out[np + q] = A(2) * B(0) - A(0) * B(2);
            out[2 * np + q] = A(0) * B(1) - A(1) * B(0);
            break;
        case NOVA_OP_SCALE: {  // scalar field * T field: s * v per component (vector.cpp:67-74)
            const double s = A(0);
            for (int c = 0; c < ncb; c++) out[c * np + q] = s * B(c);
            break;
        }
        case NOVA_OP_DIVIDE: {  // T field / scalar field: v / s per component (vector.cpp:85-92)
            const double s = B(0);
            for (int c = 0; c < nca; c++) out[c * np + q] = A(c) / s;
            break;
        }
        case NOVA_OP_DDOT: {  // tensor && tensor as tensor.cpp:237-251 writes it: every pair (i, j)
            double r = 0.0;
            for (int i = 0; i < 9; i++)
                for (int j = 0; j < 9; j++) r += A(i) * B(j);
            out[q] = r;
            break;
        }
        case NOVA_OP_SQRT:  // Field.cpp:88-105
            out[q] = sqrt(A(0));
            break;
        case NOVA_OP_OUTER:  // vector * vector -> tensor, row-major (xx, xy, xz, yx, ...)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) out[(3 * i + j) * np + q] = A(i) * B(j);
            break;
        case NOVA_OP_ADD_CONST: {  // Field<T> + T (Field - T with the constant negated)
            const double cv[3] = {c0, c1, c2};
            for (int c = 0; c < nca; c++) out[c * np + q] = A(c) + cv[c];
            break;
        }
        }
    }
}

// out = the constant vals over nc components (Field = T)
struct FillVals {
    double v[9];
};
__global__ void k_fill(int64_t n, int64_t np, int nc, FillVals vals, double *out)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < nc; c++) out[c * np + q] = vals.v[c];
}

// The literal vector gradient of src/ex/grad/grad.cpp:65-102 (the
// half-shifted stencil f(i + l) - f(i - 1 - l)) as a 9-component field,
// row = component, column = direction.
__global__ void k_grad_vec(Grid g, const double *__restrict__ U, double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.s[2] || j >= g.s[1]) return;
    const int64_t o = ((int64_t)i * g.s[1] + j) * g.s[2] + k;
    const int idx[3] = {i, j, k};
    for (int q = 0; q < 3; q++)
        for (int d = 0; d < 3; d++) {
            double acc = 0.0;
            for (int l = 0; l < g.h; l++) {
                int fp, fm, ip[3] = {i, j, k}, im[3] = {i, j, k};
                ip[d] = bmap(idx[d] + l, g.n[d], g.wall[d], fp);
                im[d] = bmap(idx[d] - 1 - l, g.n[d], g.wall[d], fm);
                const double vp = bsign(g, 3, q, d, fp) * U[q * g.np + ((int64_t)ip[0] * g.s[1] + ip[1]) * g.s[2] + ip[2]];
                const double vm = bsign(g, 3, q, d, fm) * U[q * g.np + ((int64_t)im[0] * g.s[1] + im[1]) * g.s[2] + im[2]];
                const double w = d == 2 ? g.al[2][l] * g.zinv[k] : g.al[d][l];
                acc += w * (vp - vm);
            }
            out[(3 * q + d) * g.np + o] = acc;
        }
}

constexpr int kThreads = 256;

// tools::CFL (src/Tools/Tools.cpp:9-41) over the reference's loop range
// [m/2, n - m/2) (every stored node; a periodic image node repeats node 0):
// part = per-block maxima of U_d dt / h_d (from 0, signed, as the reference)
// and of |U_d| dt / h_d.
__global__ void k_cfl_part(Grid g, const double *__restrict__ U, double dt, double hx, double hy, double hz,
                           double *__restrict__ part)
{
    double m[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double h[3] = {hx, hy, hz};
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < g.np; q += (int64_t)gridDim.x * blockDim.x)
        for (int d = 0; d < 3; d++) {
            const double u = U[d * g.np + q];
            m[d] = fmax(m[d], u * dt / h[d]);
            m[3 + d] = fmax(m[3 + d], fabs(u) * dt / h[d]);
        }
    __shared__ double sh[6][kThreads];
    for (int c = 0; c < 6; c++) sh[c][threadIdx.x] = m[c];
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int c = 0; c < 6; c++) sh[c][threadIdx.x] = fmax(sh[c][threadIdx.x], sh[c][threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; c++) part[c * gridDim.x + blockIdx.x] = sh[c][0];
}

__global__ void k_cfl_final(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    __shared__ double sh[6][kThreads];
    for (int c = 0; c < 6; c++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nb; b += kThreads) a = fmax(a, part[c * nb + b]);
        sh[c][threadIdx.x] = a;
    }
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int c = 0; c < 6; c++) sh[c][threadIdx.x] = fmax(sh[c][threadIdx.x], sh[c][threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; c++) out[c] = sh[c][0];
}

int grid_blocks(int64_t n)
{
    int64_t b = (n + kThreads * 4 - 1) / (kThreads * 4);
    return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

}  // namespace

int field_op(nova_ctx *ctx, int op, const double *a, int nca, const double *b, int ncb, const double *cval,
             double *out)
{
    const int64_t np = ctx->g.np;
    if (op == NOVA_OP_FILL) {
        FillVals v{};
        for (int c = 0; c < nca; c++) v.v[c] = cval[c];
        TimedScope ts(ctx, "field_fill");
        hipLaunchKernelGGL(k_fill, dim3(grid_blocks(np)), dim3(kThreads), 0, ctx->stream, np, np, nca, v, out);
        return check_launch("field_fill");
    }
    const double c0 = cval ? cval[0] : 0.0, c1 = cval && nca > 1 ? cval[1] : 0.0, c2 = cval && nca > 2 ? cval[2] : 0.0;
    TimedScope ts(ctx, "field_op");
    hipLaunchKernelGGL(k_field_op, dim3(grid_blocks(np)), dim3(kThreads), 0, ctx->stream, op, np, np, a, nca, b, ncb,
                       c0, c1, c2, out);
    return check_launch("field_op");
}

int grad_vec_literal(nova_ctx *ctx, const double *U, double *out)
{
    const Grid &g = ctx->g;
    TimedScope ts(ctx, "grad_vec_literal");
    hipLaunchKernelGGL(k_grad_vec, dim3((g.s[2] + 63) / 64, (g.s[1] + 3) / 4, g.s[0]), dim3(64, 4, 1), 0, ctx->stream,
                       g, U, out);
    return check_launch("grad_vec_literal");
}

int reduce_cfl(nova_ctx *ctx, const double *U, double dt, double out[6])
{
    const Grid &g = ctx->g;
    const int nb = grid_blocks(g.np) > 1024 ? 1024 : grid_blocks(g.np);
    if (int rc = ensure_red(ctx, (size_t)nb * 6 + 8)) return rc;
    hipLaunchKernelGGL(k_cfl_part, dim3(nb), dim3(kThreads), 0, ctx->stream, g, U, dt, ctx->d[0], ctx->d[1], ctx->d[2],
                       ctx->d_red + 8);
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(kThreads), 0, ctx->stream, ctx->d_red + 8, nb, ctx->d_red);
    if (int rc = check_launch("reduce_cfl")) return rc;
    NOVA_HIP(hipMemcpyAsync(ctx->h_red, ctx->d_red, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    // the reference's last loop node (m/2 before the padded end in every
    // direction): the image of node 0 on a periodic axis, node N on a walled
    // one, where its cfl_y and cfl_z take max(cfl_x, U_d dt / h_d)
    int64_t last = 0;
    {
        const int li = g.wall[0] ? g.n[0] : 0, lj = g.wall[1] ? g.n[1] : 0, lk = g.wall[2] ? g.n[2] : 0;
        last = ((int64_t)li * g.s[1] + lj) * g.s[2] + lk;
    }
    double ul[3];
    for (int d = 1; d < 3; d++)
        NOVA_HIP(hipMemcpyAsync(&ul[d], U + d * g.np + last, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    const double cx = ctx->h_red[0];
    out[0] = cx;
    out[1] = std::fmax(cx, ul[1] * dt / ctx->d[1]);
    out[2] = std::fmax(cx, ul[2] * dt / ctx->d[2]);
    for (int d = 0; d < 3; d++) out[3 + d] = ctx->h_red[3 + d];
    return NOVA_OK;
}

}  // namespace nova
