// nova_mhd_lit.hip -- stage kernels of the reference's own MHD step.
//
// app/OT_Poisson (and app/mhdjet) advance U and B in two blocks per step:
// UEqn.H:14-40 forms dU = dt(-div(U,U) + laplacian(mu,U) + div(B,B)
// - grad(0.5*(B&B))) four times at the same U and B (the stages are
// frozen) and accumulates U* = UOld + sum_j RK4a[3][j] dU, which the
// projection then makes solenoidal; BEqn.H:14-45 runs a classic RK4 of
// dB = dt(-div(U,B) + div(B,U) + laplacian(eta,B)) with that U frozen,
// B assigned and its boundaries filled after every stage, then cleans B
// with the s = 1 projection.  The U block is one launch (MODEL_MHDU); each
// of the four B stages is one launch (MODEL_MHDB) that writes the next
// stage input and carries the running RK4a[3] sum in an accumulator, so the
// four increments are never stored.  Both read U and B through the same
// staged-plane stream as the other stage kernels; the boundary fill is the
// load-time index map.
#include "nova_internal.hpp"
#include "nova_stencil.hpp"
#include "nova_stream.hpp"

namespace nova {

int mhd_u_stage(nova_ctx *ctx, const double *U, const double *B, double *out, const StageArgs &sa, double nu)
{
    if (sa.form != FORM_LIT4) return fail(NOVA_ERR_ARG, "the frozen-stage U block takes the accumulated form");
    StreamArgs<2> a{};
    a.X[0] = U;
    a.X[1] = B;
    a.out[0] = out;
    a.sa = sa;
    a.nu = nu;
    a.mk = mhd_k(ctx->g, nu, 0.0, sa.dt);
    return launch_stream_model<MODEL_MHDU>(ctx, a, "mhd_u_block");
}

int mhd_b_stage(nova_ctx *ctx, const double *U, const double *B, const double *B0, double *acc, double *out,
                double beta, double w, int first, int last, double dt, double eta)
{
    StreamArgs<2> a{};
    a.X[0] = U;
    a.X[1] = B;
    a.X0[1] = last ? nullptr : B0;  // the last stage adds to the accumulator only
    a.out[0] = out;
    a.acc = acc;
    a.acc_a = first ? 0.0 : 1.0;
    a.sa.form = FORM_LIT4;
    a.sa.dt = dt;
    a.sa.beta = beta;
    a.sa.w[0] = w;
    a.sa.alpha = last ? 1.0 : 0.0;
    a.eta = eta;
    a.mk = mhd_k(ctx->g, 0.0, eta, dt);
    return launch_stream_model<MODEL_MHDB>(ctx, a, "mhd_b_stage");
}

}  // namespace nova
