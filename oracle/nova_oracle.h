/*
 * nova_oracle.h -- CPU restatement of NOVA's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the bench's CPU-baseline leg call it.  The product
 * (libnova_hip.so) never links, loads or calls it.
 *
 * Every array is in the reference's padded layout: a node-based grid with
 * n_d = N_d + 1 + m nodes per direction (duplicated periodic endpoint plus
 * m/2 ghost layers per side), array-of-structures over the components,
 * index ((i*nj + j)*nk + k)*nc + c, i.e. k fastest
 * (reference src/Field/Field.h:82-85, src/Mesh/Mesh.cpp:89-91,136-138).
 */
#ifndef NOVA_ORACLE_H
#define NOVA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* BC directions, same numbering as the reference (src/BC/BCenums.h:9):
 * north = j high, south = j low, east = i low, west = i high,
 * top = k high, bottom = k low. */
enum { ORC_NORTH = 0, ORC_SOUTH, ORC_EAST, ORC_WEST, ORC_TOP, ORC_BOTTOM };

/* BC kinds.  PERIODIC/NEUMANN/DIRICHLET follow the reference's literal
 * ghost fills; EVEN/ODD are pure reflections about the wall node (EVEN keeps
 * the wall node, ODD sets it to zero), used by the free-slip wall case. */
enum { ORC_BC_PERIODIC = 0, ORC_BC_NEUMANN = 1, ORC_BC_DIRICHLET = 2,
       ORC_BC_EVEN = 10, ORC_BC_ODD = 11 };

/* Time-integration compositions. */
enum {
    ORC_RK_LITERAL4 = 0,   /* frozen-stage RK4 + one projection per step */
    ORC_RK_SSP3     = 1,   /* SSP-RK3 (Shu-Osher) + one projection per step */
    ORC_RK_SHEAR4   = 2    /* RK4 weights, projection at every stage */
};

typedef struct orc_mesh orc_mesh;
typedef struct orc_poisson orc_poisson;

/* Mesh (src/Mesh/Mesh.cpp:24-179, z(k) :373-407).  N = cells per direction,
 * L = lengths, o = origin, m = stencil order, hyp/s1/s2 = optional
 * hyperbolic z stretching.  wall_y != 0 declares reflecting walls at the two
 * y ends (the y direction then has no periodic images). */
orc_mesh *orc_mesh_create(const int N[3], const double L[3], const double o[3],
                          int m, int hyp, double s1, double s2, int wall_y);
void orc_mesh_free(orc_mesh *);
void orc_mesh_dims(const orc_mesh *, int n_out[3]);
void orc_mesh_metric(const orc_mesh *, double *dzdxi, double *d2zdxi2);
double orc_mesh_z(const orc_mesh *, int k);
void orc_coef(int m, double out[4]);

/* Ghost fill.  Entries are applied in the given order, as
 * Field::correctBoundaryConditions walks its BC list (Field.hxx:316-322). */
void orc_bc_apply(const orc_mesh *, double *f, int nc, int nent,
                  const int *dir, const int *kind, const int *comp,
                  const double *val);
/* The standard fill of the mesh: periodic in every direction without walls,
 * EVEN/ODD reflections in y when wall_y is set (component `odd_comp` is odd;
 * -1 means none). */
void orc_fill(const orc_mesh *, double *f, int nc, int odd_comp);

/* Explicit operators (src/ex).  `out` is overwritten with zero first, then
 * accumulated over the interior [m/2, n-m/2), as the reference's pool
 * temporaries are (src/reuseTmp/reuseTmp.hxx:47). */
void orc_div_adv(const orc_mesh *, const double *U, const double *phi, int ncphi, double *out);
void orc_div(const orc_mesh *, const double *U, double *out);
void orc_grad(const orc_mesh *, const double *p, double *out);
void orc_laplacian(const orc_mesh *, double nu, const double *f, int nc, double *out);
/* Variable-coefficient form (laplacian.hxx:64-127); nu is a padded scalar. */
void orc_laplacian_var(const orc_mesh *, const double *nu, const double *f, int nc, double *out);
void orc_grad_vec_literal(const orc_mesh *, const double *U, double *out9);
void orc_curl_literal(const orc_mesh *, const double *U, double *out);
void orc_curl_central(const orc_mesh *, const double *U, double *out);

/* FFT (unnormalised; sign -1 forward, +1 backward), complex interleaved. */
void orc_fft1d(int n, double *data, int sign);
void orc_fft3d(const int n[3], double *data, int sign); /* data index i + n0*(j + n1*k) */

/* Poisson (src/poisson/poisson.cpp:167-181,223-265,337-396). */
orc_poisson *orc_poisson_create(const orc_mesh *);
void orc_poisson_free(orc_poisson *);
void orc_poisson_k2(const orc_poisson *, double *out); /* N0*N1*N2, i fastest */
/* rhs and p are padded scalars; p's interior [m/2, n-1-m/2) is overwritten. */
void orc_poisson_solve(orc_poisson *, const double *rhs, double *p);

/* One time step of the incompressible TG/NS case (app/TG_Piosson/RKLoop.H
 * and the declared compositions).  Ustar and p are caller-owned work/output
 * fields (padded vector and scalar). */
void orc_ns_step(const orc_mesh *, orc_poisson *, double *U, double *Ustar,
                 double *p, double nu, double dt, int mode);
/* One MHD step.  mode 0 = app/OT_Poisson/{UEqn,BEqn}.H literally;
 * mode 1 = SSP-RK3 with u and B projected at every stage. */
void orc_ns_step_fused(const orc_mesh *, orc_poisson *, double *U, double *W1, double *W2, double *p, double nu,
                       double dt);
void orc_set_threads(int n);
void orc_mhd_rhs(const orc_mesh *, const double *U, const double *B, double nu, double eta,
                 double *dU, double *dB);
void orc_mhd_step(const orc_mesh *, orc_poisson *, double *U, double *B,
                  double *p, double *pB, double nu, double eta, double dt, int mode);
/* Weakly-compressible step (app/TG/TG.cpp with rhoEqn.H, pEqn.H, UEqn.H):
 * rho and U in/out, p = pc rho^2 out.  B (may be NULL) adds the Lorentz
 * terms of app/OT/UEqn.H with the frozen B of app/OT (coefficient dbu). */
void orc_wc_step(const orc_mesh *, double *U, double *rho, double *p, const double *B, double mu, double pc,
                 double dbu, double dt);
/* Weakly-compressible step of app/shearLayerWC (UEqn with buoyancy, rhoEqn,
 * pEqn, TEqn): U, rho, T in/out; p in (used by UEqn) and out. */
void orc_wcs_step(const orc_mesh *, double *U, double *rho, double *p, double *T, double mu, double ka,
                  const double g[3], double pc, double dt);
/* Initial hydrostatic pressure and density of app/shearLayerWC
 * (hydrostaticP.H, shearLayer.cpp:139-151): per (i, j) column of
 * i in [h, ni-h-1), j in [h, nj-h), from k = nk-2-h down to h,
 * pHyd(k) = sum_{k' >= k} T dz g[2] rho0 accumulated in that order;
 * p += pHyd; then rho = pow(p / pc, 0.5) on [h, n-h) in every direction. */
void orc_wc_hydrostatic(const orc_mesh *, const double *T, double *p, double *rho, const double g[3],
                        double rho0, double pc);
/* EnSight Gold output of the reference's Field::write / Field::read
 * (src/Field/Field.hxx:403-491) and Mesh::write + writeCase +
 * writeRestartInfo (src/Mesh/Mesh.cpp:224-365) for one rank (procNo) and
 * the given output times.  f is an nc-component padded field.  Return 0 or
 * -1 (file error). */
int orc_field_write(const orc_mesh *, const double *f, int nc, const char *path, const char *name, int procNo,
                    int fileNum);
int orc_field_read(const orc_mesh *, double *f, int nc, const char *path, const char *name, int procNo,
                   int fileNum);
int orc_mesh_write(const orc_mesh *, const char *path, int procNo, const double *steps, int nsteps);
/* Implicit matrix (src/im/fdMatrix/fdMatrix.hxx, src/im/laplacian/laplacian.hxx:23-58)
 * for scalars: al = six padded coefficient fields back to back (N, S, E, W,
 * T, B), ap, rhs and phi padded.  orc_fd_laplacian builds im::laplacian(mu,
 * phi) with its deferred correction; orc_fd_solve runs fdMatrix::solve on
 * phi in place and returns the iteration count. */
void orc_fd_laplacian(const orc_mesh *, double mu, const double *phi, double *al, double *ap, double *rhs);
double orc_fd_residual(const orc_mesh *, const double *al, const double *ap, const double *rhs, const double *phi);
int orc_fd_solve(const orc_mesh *, const double *al, const double *ap, const double *rhs, double *phi, double tol,
                 int max_its, double *res_out);
void orc_fd_relax(const orc_mesh *, double *ap, double *rhs, const double *phi, double urf);
/* Projection: F <- F - s*grad(lap^-1(div(F)/s)), p receives the potential. */
void orc_project(const orc_mesh *, orc_poisson *, double *F, double *p, double s);

/* Diagnostics (app/TG_Piosson/TG.cpp:148-166, src/Tools/Tools.h:19-36,
 * app/OT_Poisson/OT.cpp:134-181). */
double orc_ke(const orc_mesh *, const double *U);
double orc_maxabs_div(const orc_mesh *, const double *U);
double orc_dissipation(const orc_mesh *, const double *U, double nu);
void orc_cfl(const orc_mesh *, const double *U, double dt, double out[6]);
void orc_mhd_diag(const orc_mesh *, const double *U, const double *B,
                  double mu, double eta, double out[8]);

/* Batched tridiagonal Thomas solve, src/Tools/Tools.cpp:43-62 semantics. */
void orc_tdma(int n, const double *a, const double *b, const double *c,
              const double *d, double *x);

/* ---- compact 6th-order scheme (no reference counterpart: the reference
 * computes only the explicit stencils above).  Lele (1992), J. Comput.
 * Phys. 103:16-42, tridiagonal schemes with alpha = 1/3, a = 14/9, b = 1/9
 * (first derivative) and alpha = 2/11, a = 12/11, b = 3/11 (second
 * derivative), on periodic lines. ---- */
/* Cyclic tridiagonal solve with constant sub-diagonal a, diagonal b,
 * super-diagonal c (corners: row 0 col n-1 = a, row n-1 col 0 = c). */
void orc_cyclic_tridiag(int n, double a, double b, double c, const double *d, double *x);
/* order 1: k'h = (a sin kh + b/2 sin 2kh) / (1 + 2 alpha cos kh);
 * order 2: k''^2 h^2 = (2a (1 - cos kh) + b/2 (1 - cos 2kh)) / (1 + 2 alpha cos kh). */
double orc_compact_kmod(double kh, int order);
/* out = compact derivative (order 1 or 2) of every component of the padded
 * periodic field f along direction dir (0, 1, 2); ghosts of out are filled. */
void orc_compact_deriv(const orc_mesh *, const double *f, int nc, int dir, int order, double *out);
/* Poisson solver whose k'^2 is the compact first-derivative wavenumber
 * squared, the symbol of (compact div)(compact grad). */
orc_poisson *orc_poisson_create_compact(const orc_mesh *);
/* R_c = -1/2 sum_d [D_d(U_d U_c) + U_d D_d U_c + U_c D_d U_d] + nu sum_d D2_d U_c */
void orc_compact_ns_rhs(const orc_mesh *, const double *U, double nu, double *out);
/* F <- F - s D p with (compact div)(D p) = (compact div F) / s. */
void orc_compact_project(const orc_mesh *, orc_poisson *, double *F, double *p, double s);
/* SSP-RK3 step with the compact right-hand side and one compact projection. */
void orc_compact_ns_step(const orc_mesh *, orc_poisson *, double *U, double *Ustar, double *p,
                         double nu, double dt);
/* MHD right-hand sides with compact operators (the terms of
 * app/OT_Poisson/UEqn.H:16-23 and BEqn.H:15-23):
 *   dU = -B(U,U) + B(B,B) - D(|B|^2/2) + nu D2 U,
 *   dB = -B(U,B) + B(B,U) + eta D2 B,
 * B(X,Y)_c = 1/2 sum_d [D_d(X_d Y_c) + X_d D_d Y_c + Y_c D_d X_d]. */
void orc_compact_mhd_rhs(const orc_mesh *, const double *U, const double *B, double nu, double eta,
                         double *dU, double *dB);
/* SSP-RK3 MHD step with compact operators, u and B projected at every stage. */
void orc_compact_mhd_step(const orc_mesh *, orc_poisson *, double *U, double *B, double *p, double *pB,
                          double nu, double eta, double dt);
/* max |compact div U| over the unique nodes. */
double orc_compact_maxabs_div(const orc_mesh *, const double *U);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
