// nova_walls.hip -- the wall-node part of the reference's boundary fills.
//
// dirichletBC (src/BC/dirichletBC.hxx:90-125) and neumannBC
// (src/BC/neumannBC.hxx:100-135) with value 0 do two things on a wall:
// they fill the ghost layers by an odd resp. even reflection about the wall
// node, which every stencil here applies when it loads (Grid::odd), and they
// overwrite the wall node itself: Dirichlet with the value, Neumann with the
// neighbouring interior node (a first-order one-sided condition).  This file
// is the second part, applied to a stored field after each assignment when a
// context follows the reference's fills (nova_ctx_set_wall_bc).  Vector
// components take the kinds set per wall direction; scalar fields
// (the pressure) take Neumann 0, as the reference's pBC files state.
#include "nova_internal.hpp"

namespace nova {
namespace {

constexpr int kThreads = 256;

// Both wall planes of direction d, every component: node 0 and node N along
// d (stored index s[d] - 1 = N) overwritten.
__global__ void k_wall_fill(Grid g, int d, int nc, int k0, int k1, int k2, double *__restrict__ F)
{
    const int kind[3] = {k0, k1, k2};
    // the two other directions, slow then fast
    const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
    const int64_t nab = (int64_t)g.s[a] * g.s[b];
    const int64_t total = nab * 2 * nc;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = q % nab;
        const int side = (int)((q / nab) % 2), c = (int)(q / (2 * nab));
        const int ia = (int)(e / g.s[b]), ib = (int)(e % g.s[b]);
        const int w = side ? g.s[d] - 1 : 0;     // wall node
        const int nb = side ? w - 1 : 1;         // its interior neighbour
        int idx[3], jdx[3];
        idx[a] = jdx[a] = ia;
        idx[b] = jdx[b] = ib;
        idx[d] = w;
        jdx[d] = nb;
        const int64_t ow = ((int64_t)idx[0] * g.s[1] + idx[1]) * g.s[2] + idx[2];
        const int64_t on = ((int64_t)jdx[0] * g.s[1] + jdx[1]) * g.s[2] + jdx[2];
        const int kd = nc == 3 ? kind[c] : NOVA_WALL_NEUMANN0;
        if (kd == NOVA_WALL_DIRICHLET0) F[c * g.np + ow] = 0.0;
        else if (kd == NOVA_WALL_NEUMANN0) F[c * g.np + ow] = F[c * g.np + on];
    }
}

}  // namespace

int wall_fill(nova_ctx *ctx, double *F, int nc)
{
    if (!ctx->wall_fill) return NOVA_OK;
    const Grid &g = ctx->g;
    TimedScope ts(ctx, "wall_fill");
    for (int d = 0; d < 3; d++) {
        if (!g.wall[d]) continue;
        const int64_t n = (int64_t)g.np / g.s[d] * 2 * nc;
        int blocks = (int)((n + kThreads - 1) / kThreads);
        if (blocks > 4096) blocks = 4096;
        const int *k = ctx->wall_kind[d];
        hipLaunchKernelGGL(k_wall_fill, dim3(blocks), dim3(kThreads), 0, ctx->stream, g, d, nc, k[0], k[1], k[2], F);
        if (int rc = check_launch("wall_fill")) return rc;
    }
    return NOVA_OK;
}

int set_wall_bc(nova_ctx *ctx, int dir, const int kind[3])
{
    if (dir < 0 || dir > 2) return fail(NOVA_ERR_ARG, "wall direction must be 0, 1 or 2");
    if (!ctx->g.wall[dir]) return fail(NOVA_ERR_ARG, "the direction has no walls");
    if (ctx->g.s[dir] < 2) return fail(NOVA_ERR_ARG, "a walled direction needs two nodes");
    unsigned odd = 0;
    for (int c = 0; c < 3; c++) {
        if (kind[c] != NOVA_WALL_NEUMANN0 && kind[c] != NOVA_WALL_DIRICHLET0)
            return fail(NOVA_ERR_ARG, "wall kinds are NOVA_WALL_NEUMANN0 or NOVA_WALL_DIRICHLET0");
        if (kind[c] == NOVA_WALL_DIRICHLET0) odd |= 1u << c;
    }
    ctx->g.odd[dir] = odd;
    for (int c = 0; c < 3; c++) ctx->wall_kind[dir][c] = kind[c];
    ctx->wall_fill = 1;
    ctx->tune_epoch++;  // recorded step graphs carry the old conditions
    return NOVA_OK;
}

}  // namespace nova
