/*
 * nova_capi.h -- C ABI of the MI355X (gfx950) NOVA hot path.
 *
 * The reference (NOVA) has no FFI: its boundary is the C++ header API that
 * the app sources include (Field/Mesh/ex::/poisson/Time).  Each entry point
 * below names the reference interface it replaces (paths relative to the
 * reference repository root).  The C++ drop-in layer
 * (novaclone_amd/host/nova/) and the Python harness (novaclone_amd/) both
 * call only these functions.
 *
 * Conventions
 *  - Every function returns 0 (NOVA_OK) or a negative NOVA_ERR_* code and
 *    never throws; nova_last_error() returns the message of the calling
 *    thread's most recent failure.
 *  - Fields live on the device in the compact layout: structure of arrays
 *    over components, each component a dense s0*s1*s2 block, index
 *    (i*s1 + j)*s2 + k (k fastest, like the reference's Field, minus the
 *    ghost layers and the duplicated periodic endpoint).  s_d = n_d cells
 *    for a periodic direction, n_d + 1 nodes for a walled one.  Ghosts are
 *    not stored: every stencil applies the boundary condition when it
 *    loads (periodic wrap, or reflection about the wall node; the wall-
 *    normal component 1 of a 3-component field is odd, everything else
 *    even).
 *  - Device pointers are caller-owned (nova_alloc or any HIP allocation).
 *  - Work is asynchronous on the context's stream (nova_ctx_set_stream;
 *    default: the null stream).  One host thread per context at a time.
 */
#ifndef NOVA_CAPI_H
#define NOVA_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    NOVA_OK = 0,
    NOVA_ERR_ARG = -1,
    NOVA_ERR_HIP = -2,
    NOVA_ERR_UNSUPPORTED = -3,
    NOVA_ERR_STATE = -4,
    NOVA_ERR_NOMEM = -5,
    NOVA_ERR_IO = -6
};

enum { NOVA_BC_PERIODIC = 0, NOVA_BC_WALL = 1 };

/* Time-integration compositions (nova_solver_create). */
enum {
    NOVA_RK_LITERAL4 = 0, /* app/TG_Piosson/RKLoop.H:15-38: frozen-stage RK4, one projection;
                           * MHD: app/OT_Poisson/UEqn.H:14-40 then BEqn.H:14-41 (frozen-stage
                           * U and its projection, classic RK4 of B with that U frozen,
                           * cleaning with s = 1) */
    NOVA_RK_SSP3 = 1,     /* SSP-RK3 (Shu-Osher), one projection per step */
    NOVA_RK_SHEAR4 = 2    /* app/shearLayer/UEqn.H:14-38: projection at every stage */
};

enum { NOVA_MODEL_NS = 0, NOVA_MODEL_MHD = 1 };
/* Derivative schemes: the reference's explicit central stencils of order m
 * (src/ex), or the compact sixth-order tridiagonal scheme (no reference
 * counterpart; uniform periodic grids). */
enum { NOVA_SCHEME_EXPLICIT = 0, NOVA_SCHEME_COMPACT6 = 1 };

/* Reference BC kinds for the padded-layout fill (src/BC/BCenums.h:8). */
enum { NOVA_PBC_PERIODIC = 0, NOVA_PBC_NEUMANN = 1, NOVA_PBC_DIRICHLET = 2,
       NOVA_PBC_EVEN = 10, NOVA_PBC_ODD = 11 };

typedef struct nova_grid_desc {
    int n[3];          /* cells per direction (src/Mesh/Mesh.cpp:49-52 "N:") */
    double L[3];       /* lengths ("L:") */
    double origin[3];  /* ("origin:") */
    int m;             /* explicit stencil order 2/4/6/8 (src/settings/settings.cpp:24) */
    int bc[3];         /* NOVA_BC_* per direction */
    int hyp;           /* hyperbolic z stretching ("hypz:", Mesh.cpp:57-60) */
    double s1, s2;
} nova_grid_desc;

typedef struct nova_ctx nova_ctx;
typedef struct nova_poisson nova_poisson;
typedef struct nova_solver nova_solver;
typedef struct nova_adjoint nova_adjoint;
typedef struct nova_ensemble nova_ensemble;
typedef struct nova_wc nova_wc;
typedef struct nova_time nova_time;
typedef struct nova_io nova_io;
typedef struct nova_fdmatrix nova_fdmatrix;

/* ---- context / memory (replaces Mesh construction, Mesh.cpp:24-179, and
 * Field storage, src/Field/Field.hxx:21-134) ---- */
int nova_ctx_create(const nova_grid_desc *desc, int device, nova_ctx **out);
int nova_ctx_destroy(nova_ctx *ctx);
int nova_ctx_set_stream(nova_ctx *ctx, void *hip_stream);
int nova_ctx_sync(nova_ctx *ctx);
/* out: s0, s1, s2, points per component, padded n_i, n_j, n_k, m */
int nova_ctx_info(const nova_ctx *ctx, int64_t out[8]);
/* Per-k metric on the stored k range (Mesh.cpp:147-179): dzdxi, d2zdxi2. */
int nova_ctx_metric(const nova_ctx *ctx, double *dzdxi, double *d2zdxi2);
const char *nova_last_error(void);
/* Per-kernel timing (no reference counterpart; the reference prints whole-
 * step wall time, app/TG_Piosson/TG.cpp:142): when enabled, every kernel
 * the library launches is bracketed by HIP events on the context stream.
 * collect synchronises and writes "name launches total_ms" lines. */
int nova_timing_enable(nova_ctx *ctx, int enable);
int nova_timing_collect(nova_ctx *ctx, char *buf, size_t buflen);
/* Launch-shape choice of the fused stage kernels (no reference
 * counterpart), for every stage model at once: variant 0 stages each (j,k)
 * plane of a 64 x 4 (k x j) tile in LDS at one wave per SIMD with two planes
 * of loads in flight, variant 1 streams rows without LDS (forward models
 * only; the others use variant 3), variants 2, 3 and 4 stage 64 x 4, 32 x 8
 * and 16 x 16 tiles with one-field models held to three waves per SIMD
 * (default: 3), variants 5 and 6 stage 32 x 8 and 16 x 16 tiles at two waves
 * per SIMD with two planes in flight (one-field models; two-field models:
 * 32 x 8 tiles at two waves with one plane in flight, one staged plane or two
 * alternating), variants 7 and 8 the
 * same with 512-thread 32 x 16 and 64 x 8 tiles (two-field models keep
 * 32 x 8 and 64 x 4 at one wave), variants 9 and 10 the 64 x 8 and 32 x 16
 * tiles with a ring of staged planes (the halo of a plane is loaded with its
 * interior, H planes ahead; two-field models as for 8 and 7), variants 11
 * and 12 the 64 x 8 and 32 x 16 tiles with two alternating staged planes
 * (two-field models: 64 x 4 and 32 x 8 at one wave), variants 13 and 14
 * blocks of one and two waves (16 x 4, 32 x 4) with two alternating planes
 * (two-field models; one-field models run variant 3); stage_ichunk = planes per
 * streamed chunk (0 = auto); proj_variant 0 fuses the divergence into the
 * first FFT pass, 1 computes it (and the gradient correction) with streamed
 * row kernels, 2 (default) with LDS-staged streamed tiles.  Results are
 * identical up to rounding; only speed differs. */
int nova_ctx_tune(nova_ctx *ctx, int stage_variant, int stage_ichunk, int proj_variant);
/* Stage kernels that carry their own launch shape. */
enum {
    NOVA_KERNEL_NS_STAGE = 0,
    NOVA_KERNEL_MHD_STAGE = 1,
    NOVA_KERNEL_TLM_STAGE = 2,
    NOVA_KERNEL_ADJ_STAGE = 3,
    NOVA_KERNEL_NS_FIRST_STAGE = 4  /* the first SSP-RK3 stage of the sixth-order NS step */
};
/* Launch shape of one stage kernel (variants as for nova_ctx_tune). */
int nova_ctx_tune_kernel(nova_ctx *ctx, int kernel, int stage_variant, int stage_ichunk);
/* The reference's boundary fills on the walls normal to direction dir,
 * src/BC/dirichletBC.hxx:90-125 and neumannBC.hxx:100-135 with value 0, as
 * Field::setBC (Field.hxx:519-547) installs them from a BC file: kind[c]
 * for component c of vector fields (scalar fields take neumann 0, as the
 * reference's pBC files state).  Stencils then reflect component c oddly
 * (Dirichlet) or evenly (Neumann) about the wall node, and the solver steps
 * and projections overwrite the wall nodes after every assignment as the
 * fills do (Dirichlet: 0; Neumann: the neighbouring interior node), the
 * potential included.  Without this call a walled direction reflects the
 * wall-normal component oddly and the others evenly and keeps the computed
 * wall nodes.  The NS model only (MHD, weakly-compressible and adjoint
 * steps refuse a context with these fills). */
enum { NOVA_WALL_NEUMANN0 = 1, NOVA_WALL_DIRICHLET0 = 2 };
int nova_ctx_set_wall_bc(nova_ctx *ctx, int dir, const int kind[3]);
/* The wall-node overwrite of those fills on one stored field (ncomp 1 or
 * 3), for callers composing their own steps (a no-op without them). */
int nova_wall_fill(nova_ctx *ctx, double *F, int ncomp);

int nova_alloc(nova_ctx *ctx, int ncomp, double **dptr);
int nova_free(nova_ctx *ctx, double *dptr);
int nova_memcpy_h2d(nova_ctx *ctx, double *dst, const double *src, size_t count);
int nova_memcpy_d2h(nova_ctx *ctx, double *dst, const double *src, size_t count);
/* Page-locked host memory for buffers that cross to and from the device
 * (the C++ layer's host mirrors, which stand in for the reference's host
 * arrays, src/Field/Field.h:62): nova_memcpy_* from and to it runs at
 * the link's rate without a staging copy.  Freed with nova_host_free. */
int nova_host_alloc(size_t bytes, void **ptr);
int nova_host_free(void *ptr);

/* ---- reference-layout boundary (Field::correctBoundaryConditions,
 * src/Field/Field.hxx:310-401, with periodicBC/dirichletBC/neumannBC) ---- */
/* compact SoA -> padded AoS (the reference's Field layout) with the ghost
 * layers and the duplicated endpoint filled by the grid's BCs. */
int nova_pack_padded(nova_ctx *ctx, const double *compact, int ncomp, double *padded);
/* padded AoS -> compact SoA (interior nodes only). */
int nova_unpack_padded(nova_ctx *ctx, const double *padded, int ncomp, double *compact);
/* In-place ghost fill of a padded AoS field, entries applied in order with
 * the reference's semantics: src/BC/periodicBC.hxx:1-66,
 * src/BC/dirichletBC.hxx, src/BC/neumannBC.hxx (dir: 0 north .. 5 bottom,
 * src/BC/BCenums.h:9). */
int nova_bc_fill_padded(nova_ctx *ctx, double *padded, int ncomp, int nent,
                        const int *dir, const int *kind, const int *comp, const double *val);

/* ---- explicit operators, compact layout (src/ex) ---- */
/* ex::div(Field<vector>, Field<T>), src/ex/div/div.hxx:84-172 */
int nova_div_adv(nova_ctx *ctx, const double *U, const double *phi, int ncomp_phi, double *out);
/* ex::div(Field<vector>), src/ex/div/div.hxx:22-81 */
int nova_div(nova_ctx *ctx, const double *U, double *out);
/* ex::grad(Field<scalar>), src/ex/grad/grad.cpp:25-62 */
int nova_grad(nova_ctx *ctx, const double *p, double *out);
/* ex::laplacian(scalar, Field<T>), src/ex/laplacian/laplacian.hxx:130-191 */
int nova_laplacian(nova_ctx *ctx, double nu, const double *f, int ncomp, double *out);
/* ex::laplacian(Field<scalar> nu, Field<T>), src/ex/laplacian/laplacian.hxx:64-127:
 * the variable-coefficient form, nu a one-component field (even at walls) */
int nova_laplacian_var(nova_ctx *ctx, const double *nu, const double *f, int ncomp, double *out);
/* ex::curl, src/ex/curl/curl.cpp:25-57 (half-shifted literal stencil) */
int nova_curl_literal(nova_ctx *ctx, const double *U, double *out);
/* consistent curl from the central stencil (no reference counterpart) */
int nova_curl(nova_ctx *ctx, const double *U, double *out);
/* out = a*x + b*y over count doubles (Field algebra, Field.hxx:617-1150) */
int nova_axpby(nova_ctx *ctx, size_t count, double a, const double *x, double b,
               const double *y, double *out);
/* Pointwise Field products (src/Field/Field.h:227-363, Field.cpp:88-150 over
 * the node operators of src/Types), each node computed as the reference's
 * node type computes it (same operand order, no fused multiply-adds):
 *   NOVA_OP_DOT       a & b: nca = ncb = 1 or 3 -> 1 component
 *   NOVA_OP_CROSS     a ^ b: 3, 3 -> 3
 *   NOVA_OP_SCALE     a * b with a a scalar field: 1, ncb -> ncb
 *   NOVA_OP_DIVIDE    a / b with b a scalar field: nca, 1 -> nca
 *   NOVA_OP_DDOT      a && b of tensors (9, 9 -> 1), as tensor.cpp:237-251
 *                     writes it: the sum over every component pair (i, j)
 *   NOVA_OP_SQRT      sqrt(a): 1 -> 1 (b unused)
 *   NOVA_OP_OUTER     a * b of vectors: 3, 3 -> 9 (tensor, row-major)
 *   NOVA_OP_ADD_CONST a + cval (nca components; Field - T with cval = -T)
 *   NOVA_OP_FILL      out = cval: nca (at most 9) components, every stored
 *                     node (Field = T, Field.hxx:141-305); a and b unused
 * out may alias a or b only for DOT, SQRT, SCALE, DIVIDE and ADD_CONST when
 * it has the components of the aliased operand. */
enum { NOVA_OP_DOT = 0, NOVA_OP_CROSS = 1, NOVA_OP_SCALE = 2, NOVA_OP_DIVIDE = 3, NOVA_OP_DDOT = 4,
       NOVA_OP_SQRT = 5, NOVA_OP_OUTER = 6, NOVA_OP_ADD_CONST = 7, NOVA_OP_FILL = 8 };
int nova_field_op(nova_ctx *ctx, int op, const double *a, int nca, const double *b, int ncb,
                  const double *cval, double *out);
/* ex::grad(Field<vector>) -> Field<tensor>, src/ex/grad/grad.cpp:65-102: the
 * half-shifted literal stencil, 9 components, row = component, column =
 * direction (the tensor that ex::curl takes the antisymmetric part of). */
int nova_grad_vec_literal(nova_ctx *ctx, const double *U, double *out9);
/* tools::CFL(U, mesh), src/Tools/Tools.cpp:9-41: out[0..2] the values it
 * prints, with its cfl_y = max(cfl_x, U_y dt / dy at the last node of the
 * loop) and cfl_z alike, as written; out[3..5] the per-direction maxima of
 * |U_d| dt / h_d (the Courant numbers). */
int nova_cfl(nova_ctx *ctx, const double *U, double dt, double out[6]);

/* ---- Poisson (src/poisson/poisson.cpp) ---- */
/* poisson::poisson + wavenumbersFFT_, poisson.cpp:6-164,337-396.  Periodic
 * directions use FFTs; a walled y direction uses the even (DCT-I) solve. */
int nova_poisson_create(nova_ctx *ctx, nova_poisson **out);
int nova_poisson_destroy(nova_poisson *ps);
/* poisson::rhs + poisson::solve, poisson.cpp:167-181,223-265: p = lap^-1 rhs
 * with the modified wavenumbers, null modes (k'^2 < 1e-16) set to zero. */
int nova_poisson_solve(nova_poisson *ps, const double *rhs, double *p);
/* Projection F <- F - s*grad(p), lap p = div(F)/s
 * (app/TG_Piosson/RKLoop.H:30-38 with s = dt; app/OT_Poisson/BEqn.H:37-41
 * with s = 1).  p (scalar, may be NULL) receives the potential. */
int nova_project(nova_poisson *ps, double *F, double s, double *p);
/* Derivative scheme of the projection (NOVA_SCHEME_*): compact divergence
 * and gradient passes and the compact modified wavenumbers in the solve. */
int nova_poisson_set_scheme(nova_poisson *ps, int scheme);
/* Solve across a walled y: NOVA_WALL_SOLVE_EVEN (default) the even extension
 * (a DCT-I through an FFT of length 2 N_y); NOVA_WALL_SOLVE_REFERENCE the
 * reference's own FFT solve (poisson.cpp:167-265, which its walled apps
 * call), y taken as periodic over the nodes 0 .. N_y - 1 and the wall node
 * N_y left to the boundary fill. */
enum { NOVA_WALL_SOLVE_EVEN = 0, NOVA_WALL_SOLVE_REFERENCE = 1 };
int nova_poisson_set_wall_solve(nova_poisson *ps, int mode);
/* Compact sixth-order derivative (order 1 or 2) along direction dir of each
 * of ncomp components (no reference counterpart; Lele's tridiagonal scheme
 * with alpha = 1/3 resp. 2/11, solved per line by a wavefront). */
int nova_compact_deriv(nova_ctx *ctx, const double *f, int ncomp, int dir, int order, double *out);

/* ---- time stepping (the app-level RK loops) ---- */
int nova_solver_create(nova_poisson *ps, int model, int mode, double nu, double eta,
                       nova_solver **out);
int nova_solver_destroy(nova_solver *sv);
/* Advance nsteps of size dt.  U (and B for MHD) in/out, p (and pB) out.
 * NS: app/TG_Piosson/RKLoop.H; MHD: app/OT_Poisson/{UEqn,BEqn}.H (the
 * reference's composition with NOVA_RK_LITERAL4, both fields projected
 * every stage with NOVA_RK_SSP3).  The current density J = ex::curl(B)
 * that BEqn.H:43 forms after the cleaning is output only and is left to
 * nova_curl_literal / nova_mhd_diag. */
int nova_solver_step(nova_solver *sv, double dt, int nsteps, double *U, double *B,
                     double *p, double *pB);
/* Record the step sequence as a HIP graph and replay it (1) or launch
 * kernels one by one (0). */
int nova_solver_use_graph(nova_solver *sv, int enable);
/* Derivative scheme of the solver's right-hand side (NOVA_SCHEME_*); the
 * compact scheme needs SSP-RK3 (NS or MHD) and a Poisson solver set to the
 * same scheme. */
int nova_solver_set_scheme(nova_solver *sv, int scheme);
/* A constant vector field F added to the NS right-hand side of every stage,
 * dU = dt (R(U) + F): the Lorentz force div(B, B) - grad(B & B / 2) of a
 * magnetic field that does not evolve (app/shearLayer/UEqn.H:16-26 with
 * BEqn.H switched off, shearLayer.cpp:178-179).  F stays the caller's and
 * must outlive the solver's use of it; NULL removes it. */
int nova_solver_set_force(nova_solver *sv, const double *F);
/* Autotuner (no reference counterpart): times every launch shape of the
 * solver's stage kernel (variants x chunks of 0, 8, 16, 32, 64, 128 planes) and
 * the three projection variants on U (and B), reps launches each after one
 * untimed launch, and keeps the fastest.  U and B are not modified.  When
 * report is non-null it receives one line per trial and the choices. */
int nova_solver_autotune(nova_solver *sv, const double *U, const double *B, int reps, char *report,
                         size_t report_len);

/* ---- diagnostics (app/TG_Piosson/TG.cpp:148-166, src/Tools/Tools.h:19-36,
 * app/OT_Poisson/OT.cpp:134-176) ---- */
int nova_reduce_ke(nova_ctx *ctx, const double *U, double *out);
int nova_reduce_maxabs_div(nova_ctx *ctx, const double *U, double *out);
/* nu * mean |grad U|^2 (central stencil per component): the viscous
 * dissipation rate, the well-conditioned counterpart of the -dEk/dt column
 * of app/TG_Piosson/TG.cpp:170. */
int nova_reduce_dissipation(nova_ctx *ctx, const double *U, double nu, double *out);
/* out = {Ek, Em, Jav, Jmax, omega, epsilon, Umax, 0} */
int nova_mhd_diag(nova_ctx *ctx, const double *U, const double *B, double mu, double eta,
                  double out[8]);

/* ---- discrete adjoint of the SSP-RK3 + projection step (no reference
 * counterpart: NOVA has no adjoint).  Inner product: sum over stored nodes,
 * weight 1/2 on wall nodes (nova_weighted_dot); gradients are with respect
 * to that inner product. ---- */
/* Reserve checkpoint memory for up to max_steps steps, one checkpoint of U
 * every checkpoint_every steps, plus a pool of 3*checkpoint_every vector
 * fields that keeps the step-start states of the forward sweep's last steps
 * (one field per step). */
int nova_adjoint_create(nova_solver *sv, int checkpoint_every, int max_steps, nova_adjoint **out);
int nova_adjoint_destroy(nova_adjoint *ad);
/* U = forward(U0) over nsteps of dt, storing the checkpoints. */
int nova_adjoint_forward(nova_adjoint *ad, double dt, int nsteps, const double *U0, double *U);
/* lam0 = (dU_final/dU0)^T lam_final along the last forward trajectory. */
int nova_adjoint_backward(nova_adjoint *ad, const double *lam_final, double *lam0);
/* dU_final = (dU_final/dU0) dU0 (tangent-linear model). */
int nova_adjoint_tlm(nova_adjoint *ad, const double *dU0, double *dU_final);
/* Autotuner for the adjoint and tangent-linear stage kernels, run on the
 * first checkpoint of the last forward run (see nova_solver_autotune). */
int nova_adjoint_autotune(nova_adjoint *ad, int reps, char *report, size_t report_len);
/* Further pool fields (one vector field each) within max_bytes of device
 * memory, at most max_steps - 3*checkpoint_every of them; as many as the
 * device can allocate are kept and *fields (may be null) receives their
 * count.  The forward sweep keeps the step-start state U of its last
 * min(pool, nsteps) steps; the reverse sweep rebuilds the stage inputs U1, U2
 * of those steps with two stage launches each and recomputes the earlier
 * steps from their checkpoints a segment at a time.  max_bytes = 0 releases
 * them.  Takes effect from the next nova_adjoint_forward. */
int nova_adjoint_reserve(nova_adjoint *ad, uint64_t max_bytes, int *fields);
/* out = sum_nodes w a.b over ncomp components. */
int nova_weighted_dot(nova_ctx *ctx, const double *a, const double *b, int ncomp, double *out);
/* J = dV sum_nodes w |curl U|^2 (central stencil curl); dJdU (may be NULL)
 * receives its gradient 2 dV curl(curl U). */
int nova_enstrophy(nova_ctx *ctx, const double *U, double *J, double *dJdU);

/* ---- weakly-compressible step (app/TG/TG.cpp:71-76: rhoEqn.H, pEqn.H,
 * UEqn.H) ---- */
/* A solver on ctx's grid with viscosity mu and p = pc rho^2 (pEqn.H:2 has
 * pc = 50, i.e. c = 10).  Allocates six stage buffers (12 fields). */
int nova_wc_create(nova_ctx *ctx, double mu, double pc, nova_wc **out);
int nova_wc_destroy(nova_wc *w);
/* Advance nsteps of size dt.  Per step: a classic RK4 of rho with U frozen
 * (drho = dt (-div(rho U)), rhoEqn.H:19-33), p = pc rho^2, and a classic
 * RK4 of rho U with rho and p frozen (drhoU = dt (-div(rho U, U) - grad p +
 * laplacian(mu, U)), U = rhoUOld/rho + c drhoU/rho, UEqn.H:16-33).  U
 * (3 components) and rho in/out, p out; device pointers, stored layout. */
int nova_wc_step(nova_wc *w, double dt, int nsteps, double *U, double *rho, double *p);
/* The magnetic field of app/OT (OT.cpp:92-95 runs rhoEqn.H, pEqn.H,
 * UEqn.H, BEqn.H; BEqn.H never assigns B, so B stays as set): each U stage
 * adds the Lorentz terms -div((2 dbu) B, B) + grad(dbu (B & B)) of
 * app/OT/UEqn.H:22-23, formed once per step.  B (3 components, device,
 * stored layout) is read, not copied; NULL removes it. */
int nova_wc_set_magnetic(nova_wc *w, const double *B, double dbu);
/* The temperature equation of app/shearLayerWC (TEqn.H: classic RK4 of
 * rho T with rho and U frozen, drhoT = dt (-div(rho U, T) + laplacian(ka, T)))
 * and the buoyancy -(rho T) g in UEqn.H; allocates three more buffers. */
int nova_wc_set_temperature(nova_wc *w, double ka, double gx, double gy, double gz);
/* Advance nsteps in the order of app/shearLayerWC (shearLayer.cpp:197-200):
 * UEqn (with buoyancy; p as given, i.e. the previous step's), rhoEqn, pEqn
 * (p = (pc rho) rho, pc = c^2/(2 rho0)), TEqn.  U, rho, T in/out; p in/out. */
int nova_wc_step_t(nova_wc *w, double dt, int nsteps, double *U, double *rho, double *p, double *T);
/* The initial state of app/shearLayerWC after its initial conditions
 * (hydrostaticP.H, shearLayer.cpp:139-151): per (i, j) column, from the
 * node below the top (k = N_z - 1) down to k = 0, p += sum_{k' >= k}
 * ((T dz) g_z) rho0, accumulated from the top in that order (g_z from
 * nova_wc_set_temperature, which must come first); the top node k = N_z of a
 * walled z and the plane i = N_x of a walled x keep p, as the reference's
 * loop bounds leave them.  Then rho = (p / pc)^(1/2) on every node.  T in,
 * p in/out, rho out; device pointers, stored layout. */
int nova_wc_hydrostatic(nova_wc *w, const double *T, double rho0, double *p, double *rho);

/* ---- implicit matrix im::fdMatrix<scalar> (src/im/fdMatrix/fdMatrix.hxx)
 * and im::laplacian (src/im/laplacian/laplacian.hxx:23-58) ----
 * The matrix keeps its six neighbour coefficient fields (north, south,
 * east, west, top, bottom), ap and the right-hand side in the reference's
 * padded layout (nova_ctx_info's padded extents, k fastest).  Fields passed
 * in or out (phi, f) are stored-layout scalars; the matrix pads them with
 * the grid's boundary fill, as a reference Field carries its ghosts. */
/* im::laplacian(mu, phi): al = mu/d^2 and ap = -2 mu (1/dx^2 + 1/dy^2 +
 * 1/dz^2) on the interior, then deferredCorrection(ex::laplacian(mu, phi))
 * (fdMatrix.hxx:213-245): rhs = -lap + sum al phi_nb + ap phi. */
int nova_fd_laplacian(nova_ctx *ctx, double mu, const double *phi, nova_fdmatrix **out);
int nova_fd_destroy(nova_fdmatrix *m);
/* operator==(matrix, f) (fdMatrix.hxx:372-382): rhs += f */
int nova_fd_add_source(nova_fdmatrix *m, const double *f);
/* relax(urf) (fdMatrix.hxx:248-253) with phi the transported field */
int nova_fd_relax(nova_fdmatrix *m, const double *phi, double urf);
/* setReference(i, j, k, val) (fdMatrix.hxx:256-268), padded indices */
int nova_fd_set_reference(nova_fdmatrix *m, int i, int j, int k, double val);
/* m += s * other (operator+= with s = 1, operator-= with s = -1,
 * fdMatrix.hxx:327-356), or m *= s with other = NULL (unary minus, s = -1,
 * :394-413); both matrices on one context */
int nova_fd_combine(nova_fdmatrix *m, const nova_fdmatrix *other, double s);
/* residual() (fdMatrix.hxx:271-310): RMS of rhs - A phi over the interior */
int nova_fd_residual(nova_fdmatrix *m, const double *phi, double *res);
/* solve(tol, maxIts) (fdMatrix.hxx:59-209) on phi in place: until the
 * residual is at most tol, sweeps of Thomas line solves along j, i and k
 * in the reference's line order (each line sees the lines before it in its
 * sweep already updated), then the boundary refill; its = iterations run,
 * res = the last residual computed. */
int nova_fd_solve(nova_fdmatrix *m, double *phi, double tol, int max_its, int *its, double *res);
/* A padded array of the matrix: what = 0..5 al(what), 6 ap, 7 rhs, 8 the
 * padded transported field of the last call. */
int nova_fd_get(const nova_fdmatrix *m, int what, double *padded_out);

/* ---- output and restart (src/Time/Time.cpp, src/Mesh/Mesh.cpp:224-369,
 * src/Field/Field.hxx:403-522) ---- */
/* The run clock.  Time(dt, endTime, writeEvery) (Time.cpp:23-62): output
 * every write_every steps; Time(dt, endTime, writeInterval) (:64-74): output
 * when the time reaches the next multiple of the interval; Time(file)
 * (:77-109): the "dt:", "writeT:", "writeN:", "end:" lines of a settings
 * file.  nova_time_restart is the restart branch of :40-61: it reads the
 * output count and times of a restart file (Mesh::writeRestartInfo's
 * format), resumes the clock at the last time and numbers files on from
 * the count. */
int nova_time_create(double dt, double end_time, int write_every, nova_time **out);
int nova_time_create_interval(double dt, double end_time, double write_interval, nova_time **out);
int nova_time_create_from_file(const char *file, nova_time **out);
int nova_time_restart(nova_time *t, const char *restart_file);
int nova_time_destroy(nova_time *t);
/* Time::operator++ (Time.cpp:112-139): decide whether this step writes (and
 * if so count the file and record the current time), then advance the time
 * by dt and the step count by one.  A settings file without writeN/writeT
 * leaves no output rule: advancing it is an error. */
int nova_time_advance(nova_time *t);
int nova_time_set_dt(nova_time *t, double dt);  /* Time::dt() is writable */
/* dout: dt, current time, end time, write interval; iout: time step, file
 * number, write (0/1), log (0/1), number of output times.  run() is
 * current time < end time. */
int nova_time_info(const nova_time *t, double dout[4], int64_t iout[5]);
int nova_time_steps(const nova_time *t, double *out, int64_t cap);

/* Output of one grid's fields.  Writes are asynchronous: the layout change
 * runs on the context's stream (so the field may be overwritten by work
 * queued after the call), the copy to host and the file write run behind
 * it; nova_io_wait blocks until every queued file is written and returns
 * the first error.  One rank: the processor number in the file names is
 * 0000. */
int nova_io_create(nova_ctx *ctx, nova_io **out);
int nova_io_destroy(nova_io *io);  /* waits for queued writes first */
int nova_io_wait(nova_io *io);
/* Field::write (Field.hxx:403-451): nothing unless the clock writes this
 * step; else path/name.0000.FFFFF.dat (FFFFF = the clock's file number):
 * 80-byte "description", "part", int 1, "point", then for each component
 * the N_x+1 x N_y+1 x N_z+1 nodes (the periodic image node N is node 0) as
 * float32, i fastest, then j, then k.  The reference's 80-byte strings
 * carry undefined bytes after their NUL; these are NUL-padded. */
int nova_io_write_field(nova_io *io, const nova_time *t, const char *path, const char *name, const double *F,
                        int ncomp);
/* Mesh::write (Mesh.cpp:224-303, geom.p0000.geo: EnSight Gold binary
 * "block rectilinear", node coordinates as float32, z through the
 * stretching), writeCase (:306-347, fd.p0000.case) and writeRestartInfo
 * (:349-365, restart.p.0000.case), when the clock writes this step. */
int nova_io_write_mesh(nova_io *io, const nova_time *t, const char *path);
/* Field::read (Field.hxx:455-522): path/name.0000.FFFFF.dat, the header
 * skipped, float32 nodes into the stored FP64 field (image entries
 * ignored).  Synchronous; waits for queued writes first. */
int nova_io_read_field(nova_io *io, const char *path, const char *name, int file_num, double *F, int ncomp);
/* FP64 checkpoint (no reference counterpart: the reference restarts from
 * its float32 output): a 64-byte header ("NOVAFP64", version 1, N, wall
 * flags, m, field count, time, step), each field's component count, then
 * the fields in the stored layout, bit for bit.  Writing is asynchronous
 * like nova_io_write_field; reading checks the header against the grid and
 * the given component counts. */
int nova_io_checkpoint_write(nova_io *io, const char *file, int nfields, const double *const *fields,
                             const int *ncomp, double time, int64_t step);
int nova_io_checkpoint_read(nova_io *io, const char *file, int nfields, double *const *fields, const int *ncomp,
                            double *time, int64_t *step);

/* HDF5 dumps of app/mhdjet (mhdjet.cpp:25-89, 277-400): component comp of
 * an ncomp-component stored field to its own file, one dataset "/3DArray"
 * of native doubles with dimensions (N_x, N_y, N_z) (the global cell
 * counts), element [i][j][k] = node (i, j, k), k fastest (the first N_d
 * stored nodes of a walled direction).  Synchronous.  The HDF5 C library
 * (1.10 ABI) is opened at the first call: NOVA_HDF5_LIB, else libhdf5.so,
 * else /opt/conda/lib/libhdf5.so; without it the call returns
 * NOVA_ERR_UNSUPPORTED. */
int nova_hdf5_write(nova_ctx *ctx, const char *file, const double *F, int ncomp, int comp);
/* The same file from a host array of extents stored[] >= dims[] (k
 * fastest), and a reader of "/3DArray" (dims out; host may be NULL to query
 * the extents, else cap doubles at least). */
int nova_hdf5_write_host(const char *file, const int64_t dims[3], const int64_t stored[3], const double *host);
int nova_hdf5_read_host(const char *file, int64_t dims[3], double *host, size_t cap);

/* ---- ensembles of uneven instances (no reference counterpart: NOVA runs
 * one case per process; the step of each instance is that of
 * nova_solver_step with NOVA_RK_SSP3, app/TG_Piosson/RKLoop.H for
 * NOVA_MODEL_NS and app/OT_Poisson/{UEqn,BEqn}.H for NOVA_MODEL_MHD) ----
 * Instances share the stencil order m and are periodic; instances with the
 * same grid form a size class solved by one batched Poisson pipeline.  Each
 * RK stage of all instances is one persistent kernel launch, and with the
 * graph enabled a whole step is one replayed HIP graph.  Every instance's
 * fields equal those of its standalone nova_solver_step run bit for bit. */
typedef struct nova_instance_desc {
    nova_grid_desc grid;
    int model;      /* NOVA_MODEL_NS or NOVA_MODEL_MHD */
    double nu, eta; /* viscosity and (MHD) magnetic diffusivity */
} nova_instance_desc;
int nova_ensemble_create(const nova_instance_desc *inst, int n, int device, nova_ensemble **out);
int nova_ensemble_destroy(nova_ensemble *e);
int nova_ensemble_set_stream(nova_ensemble *e, void *stream);
/* Copy instance i's fields in (store = 0: U, B from the caller) or out
 * (store = 1: U, B, p, pB to the caller); device pointers in the stored
 * layout of the instance's grid, NULL skips a field. */
int nova_ensemble_copy(nova_ensemble *e, int i, int store, double *U, double *B, double *p, double *pB);
/* Advance every instance nsteps of dt (SSP-RK3). */
int nova_ensemble_step(nova_ensemble *e, double dt, int nsteps);
/* Record a step as a HIP graph and replay it (1) or launch directly (0). */
int nova_ensemble_use_graph(nova_ensemble *e, int enable);
/* out: instances, size classes, work-list tiles per stage, persistent
 * blocks (both models' launches together), total nodes, field slots. */
int nova_ensemble_info(nova_ensemble *e, int64_t out[6]);

/* ---- batched tridiagonal (src/Tools/Tools.cpp:43-62, tools::tdma) ----
 * nbatch systems of size n, system s at a[s*n ...]; cyclic != 0 adds the
 * periodic corner entries a[0] (row 0, col n-1) and c[n-1] (row n-1, col 0). */
int nova_tdma_batched(nova_ctx *ctx, int n, int nbatch, const double *a, const double *b,
                      const double *c, const double *d, double *x, int cyclic);
/* No counterpart in the reference: nova_tdma_batched for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  It
 * calls no hipMalloc/hipFree: the forward-sweep work array (3 n nbatch
 * doubles) lives in a per-GPU workspace of nova_tdma_noalloc_max()
 * doubles (16 MiB) that the library holds until the process ends; larger
 * calls fail with NOVA_ERR_ARG.  The workspace needs no clearing.  Every
 * call on one GPU shares that GPU's workspace: calls that overlap on one
 * GPU (other streams, other host threads) race. */
int nova_tdma_batched_noalloc(nova_ctx *ctx, int n, int nbatch, const double *a, const double *b,
                              const double *c, const double *d, double *x, int cyclic);
size_t nova_tdma_noalloc_max(void);

#ifdef __cplusplus
}
#endif
#endif
