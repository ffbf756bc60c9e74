/*
 * nova_oracle.c -- CPU restatement of NOVA's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product.  Each function cites the reference file:line it restates
 * (paths relative to the reference repository root).  Arithmetic follows
 * the reference's loop order and expression grouping (l-outer sweeps,
 * left-to-right sums, pool temporaries zero-filled), so that this code and
 * the reference round the same way wherever the C library calls (sin, pow)
 * agree.  OpenMP parallelises only over independent output points; the
 * accumulation order of every point is unchanged.
 *
 * Layout: padded AoS, index ((i*nj + j)*nk + k)*nc + c (k fastest).
 *
 * Pins (every function here is checked against something other than its
 * own formula): discrete eigenvalues of the stencils on Fourier modes,
 * brute-force and closed-form Poisson solves, dense solves of the
 * tridiagonal and implicit systems, stability polynomials of the RK
 * compositions, closed forms of the MHD, weakly-compressible, dissipation
 * and diagnostic functions, the stretched mesh's inverse map, the EnSight
 * byte layout (tests/test_oracle*.py); planted mistakes (ORC_PLANT) fail
 * those pins (tests/test_oracle_planted.py).
 */
#include "nova_oracle.h"

/* ORC_PLANT selects one deliberate mistake (test builds only, see
 * tests/test_oracle_planted.py): 0, the default, is the faithful oracle.
 * A non-zero value exists to show that the oracle's pins fail on it. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_mesh {
    int N[3];
    int m, h;
    int n[3];          /* padded extents ni, nj, nk */
    double d[3];       /* dx, dy, dz */
    double L[3];
    double o[3];
    int hyp;
    double s1, s2, hb; /* stretching parameters and the solved b */
    int wall_y;
    double *dzdxi, *d2zdxi2;
    double coef[4];
};

struct orc_poisson {
    const orc_mesh *mesh;
    int N[3];
    int ext;           /* 1: y is solved on the even extension (walls) */
    int M[3];          /* transform extents (M[1] = 2*N[1] with walls) */
    double *k2;        /* M0*M1*M2, i fastest */
    double complex *buf;
};

/* ------------------------------------------------------------------ */
/* Constants: src/settings/settings.cpp:25-31 (central-difference weights
 * for orders 2, 4, 6, 8), src/Tools/Tools.cpp:6-7 (pi, eps). */
static const double COEF[4][4] = {
    {1.0, 0.0, 0.0, 0.0},
    {4.0 / 3.0, -1.0 / 3.0, 0.0, 0.0},
    {6.0 / 4.0, -12.0 / 20.0, 1.0 / 10.0, 0.0},
    {8.0 / 5.0, -4.0 / 5.0, 24.0 / 105.0, -8.0 / 280.0}};
static const double ORC_PI = 3.14159265358979323846;
static const double ORC_EPS = 1.0e-16;
/* settings.cpp:44-49, last row: the RK4 accumulation weights. */
static const double RK4W[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
/* settings.cpp:51 */
static const double RK4C[4] = {0.5, 0.5, 1.0, 1.0};

void orc_coef(int m, double out[4])
{
    int r = m / 2 - 1;
    if (r < 0) r = 0;
    if (r > 3) r = 3;
    for (int l = 0; l < 4; l++) out[l] = COEF[r][l];
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

#define IX(M, i, j, k) ((((size_t)(i)) * (size_t)(M)->n[1] + (size_t)(j)) * (size_t)(M)->n[2] + (size_t)(k))
#define NPTS(M) ((size_t)(M)->n[0] * (size_t)(M)->n[1] * (size_t)(M)->n[2])

/* Mesh::z, src/Mesh/Mesh.cpp:373-407 (single rank: globK = k - m/2). */
double orc_mesh_z(const orc_mesh *M, int k)
{
    if (M->hyp) {
        double A = ORC_PLANT == 10 ? pow(M->s2 / M->s1, 0.5) : pow(M->s1 / M->s2, 0.5);
        int globK = k - M->m / 2;
        int globN = M->N[2] + 1;
        double R = (double)globK / (double)(globN - 1) - (ORC_PLANT == 11 ? 0.45 : 0.5);
        double b = M->hb;
        double U = 1.0 + tanh(b * R) / tanh(b / 2.0);
        return U / (2.0 * A + (1 - A) * U) * M->L[2];
    }
    return (k - M->m / 2) * M->d[2] + M->o[2];
}

/* Mesh::Mesh, src/Mesh/Mesh.cpp:64-179 on one rank. */
orc_mesh *orc_mesh_create(const int N[3], const double L[3], const double o[3],
                          int m, int hyp, double s1, double s2, int wall_y)
{
    orc_mesh *M = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    M->m = m;
    M->h = m / 2;
    orc_coef(m, M->coef);
    for (int d = 0; d < 3; d++) {
        M->N[d] = N[d];
        M->L[d] = L[d];
        M->o[d] = o[d];
        M->d[d] = L[d] / N[d];       /* Mesh.cpp:68-70 */
        M->n[d] = N[d] + 1 + m;      /* Mesh.cpp:89-91,136-138 */
    }
    M->hyp = hyp;
    M->s1 = s1;
    M->s2 = s2;
    M->wall_y = wall_y;
    if (hyp) { /* fixed point of Mesh.cpp:386-396 */
        double b = 1.0;
        int globN = N[2] + 1;
        for (int i = 1; i < 10000; i++) b = asinh(b / ((globN - 1) * pow(s1 * s2, 0.5)));
        M->hb = b;
    }
    int nk = M->n[2];
    M->dzdxi = (double *)calloc((size_t)nk, sizeof(double));
    M->d2zdxi2 = (double *)calloc((size_t)nk, sizeof(double));
    const double dz = M->d[2];
    for (int k = M->h; k < nk - M->h; k++) { /* Mesh.cpp:149-160 */
        M->dzdxi[k] = 0;
        for (int l = 0; l < M->h; l++) {
            M->dzdxi[k] += (M->coef[l] / 2.0 / (l + 1) / dz * orc_mesh_z(M, k + 1 + l)
                            - M->coef[l] / 2.0 / (l + 1) / dz * orc_mesh_z(M, k - 1 - l));
        }
    }
    for (int k = M->h; k < nk - M->h; k++) { /* Mesh.cpp:165-179 */
        M->d2zdxi2[k] = 0;
        for (int l = 0; l < M->h; l++) {
            M->d2zdxi2[k] += (M->coef[l] / pow((l + 1) * dz, 2) * orc_mesh_z(M, k + 1 + l)
                              + M->coef[l] / pow((l + 1) * dz, 2) * orc_mesh_z(M, k - 1 - l)
                              + (-2.0 * M->coef[l] / pow((l + 1) * dz, 2)) * orc_mesh_z(M, k));
        }
    }
    return M;
}

void orc_mesh_free(orc_mesh *M)
{
    if (!M) return;
    free(M->dzdxi);
    free(M->d2zdxi2);
    free(M);
}

void orc_mesh_dims(const orc_mesh *M, int n_out[3])
{
    for (int d = 0; d < 3; d++) n_out[d] = M->n[d];
}

void orc_mesh_metric(const orc_mesh *M, double *dzdxi, double *d2zdxi2)
{
    memcpy(dzdxi, M->dzdxi, sizeof(double) * (size_t)M->n[2]);
    memcpy(d2zdxi2, M->d2zdxi2, sizeof(double) * (size_t)M->n[2]);
}

/* ------------------------------------------------------------------ */
/* Ghost fills. */

/* periodicBC<T>::update, src/BC/periodicBC.hxx:1-66: copies the whole node
 * value (all components) for the direction it belongs to. */
static void bc_periodic(const orc_mesh *M, double *f, int nc, int dir)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], m = M->m, h = M->h;
    if (dir == ORC_EAST) {
        #pragma omp parallel for collapse(2)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++) {
                for (int i = 0; i < h; i++)
                    memcpy(&f[IX(M, i, j, k) * nc], &f[IX(M, i + ni - m - 1, j, k) * nc], sizeof(double) * nc);
                for (int i = ni - h - 1; i < ni; i++)
                    memcpy(&f[IX(M, i, j, k) * nc], &f[IX(M, i - ni + m + 1, j, k) * nc], sizeof(double) * nc);
            }
    }
    if (dir == ORC_SOUTH) {
        #pragma omp parallel for collapse(2)
        for (int k = 0; k < nk; k++)
            for (int i = 0; i < ni; i++) {
                for (int j = 0; j < h; j++)
                    memcpy(&f[IX(M, i, j, k) * nc], &f[IX(M, i, j + nj - m - 1, k) * nc], sizeof(double) * nc);
                for (int j = nj - h - 1; j < nj; j++)
                    memcpy(&f[IX(M, i, j, k) * nc], &f[IX(M, i, j - nj + m + 1, k) * nc], sizeof(double) * nc);
            }
    }
    if (dir == ORC_BOTTOM) {
        #pragma omp parallel for collapse(2)
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++) {
                for (int k = 0; k < h; k++)
                    memcpy(&f[IX(M, i, j, k) * nc], &f[IX(M, i, j, k + nk - m - 1) * nc], sizeof(double) * nc);
                for (int k = nk - h - 1; k < nk; k++)
                    memcpy(&f[IX(M, i, j, k) * nc], &f[IX(M, i, j, k - nk + m + 1) * nc], sizeof(double) * nc);
            }
    }
}

/* One face of a wall-type fill for component c.  `w` is the wall node,
 * `s` = -1 for a low face (ghosts at w-(l+1)) and +1 for a high face.
 * Kinds: DIRICHLET (src/BC/dirichletBC.hxx), NEUMANN
 * (src/BC/neumannBC.hxx), EVEN/ODD (pure reflections, no reference row). */
static void bc_wall(const orc_mesh *M, double *f, int nc, int dir, int kind, int c, double val)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    int axis, low;
    switch (dir) {
    case ORC_EAST: axis = 0; low = 1; break;
    case ORC_WEST: axis = 0; low = 0; break;
    case ORC_SOUTH: axis = 1; low = 1; break;
    case ORC_NORTH: axis = 1; low = 0; break;
    case ORC_BOTTOM: axis = 2; low = 1; break;
    default: axis = 2; low = 0; break;
    }
    const int n = M->n[axis];
    const int w = low ? h : n - h - 1;
    const int s = low ? -1 : 1;
    const double hd = M->d[axis];
    const int na = axis == 0 ? nj : ni, nb = axis == 2 ? nj : nk;
    #pragma omp parallel for collapse(2)
    for (int a = 0; a < na; a++)
        for (int b = 0; b < nb; b++) {
            size_t base, stride;
            if (axis == 0) { base = IX(M, 0, a, b); stride = (size_t)nj * nk; }
            else if (axis == 1) { base = IX(M, a, 0, b); stride = (size_t)nk; }
            else { base = IX(M, a, b, 0); stride = 1; }
#define V(q) f[(base + (size_t)(q) * stride) * nc + c]
            if (kind == ORC_BC_DIRICHLET) {
                V(w) = val;
                for (int l = 0; l < h; l++)
                    V(w + s * (l + 1)) = 2.0 * (l + 1) * val - V(w - s * (l + 1));
            } else if (kind == ORC_BC_NEUMANN) {
                if (low) V(w) = V(w + 1) - val * hd;
                else V(w) = V(w - 1) + val * hd;
                for (int l = 0; l < h; l++)
                    V(w + s * (l + 1)) = V(w - s * (l + 1)) - 2.0 * (l + 1) * val * hd;
            } else if (kind == ORC_BC_EVEN) {
                for (int l = 0; l < h; l++) V(w + s * (l + 1)) = V(w - s * (l + 1));
            } else if (kind == ORC_BC_ODD) {
                V(w) = 0.0;
                for (int l = 0; l < h; l++) V(w + s * (l + 1)) = -V(w - s * (l + 1));
            }
#undef V
        }
    (void)ni;
}

void orc_bc_apply(const orc_mesh *M, double *f, int nc, int nent, const int *dir,
                  const int *kind, const int *comp, const double *val)
{
    for (int e = 0; e < nent; e++) {
        if (kind[e] == ORC_BC_PERIODIC) bc_periodic(M, f, nc, dir[e]);
        else bc_wall(M, f, nc, dir[e], kind[e], comp[e], val ? val[e] : 0.0);
    }
}

/* The fill used by the time steppers: the reference's periodic dispatch
 * order south, east, bottom (entries north..bottom per component,
 * Field.hxx:316-322 with periodicBC acting on south/east/bottom only); with
 * y walls the north/south entries are reflections. */
void orc_fill(const orc_mesh *M, double *f, int nc, int odd_comp)
{
    for (int c = 0; c < nc; c++) {
        if (M->wall_y) {
            int kind = (c == odd_comp) ? ORC_BC_ODD : ORC_BC_EVEN;
            bc_wall(M, f, nc, ORC_NORTH, kind, c, 0.0);
            bc_wall(M, f, nc, ORC_SOUTH, kind, c, 0.0);
        } else {
            bc_periodic(M, f, nc, ORC_SOUTH);
        }
        bc_periodic(M, f, nc, ORC_EAST);
        bc_periodic(M, f, nc, ORC_BOTTOM);
    }
}

/* ------------------------------------------------------------------ */
/* Explicit operators. */

/* ex::div(Field<vector>, Field<T>), src/ex/div/div.hxx:84-172. */
void orc_div_adv(const orc_mesh *M, const double *U, const double *phi, int nc, double *out)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    const double dx = M->d[0], dy = M->d[1], dz = M->d[2];
    const double *dzdxi = M->dzdxi;
    memset(out, 0, sizeof(double) * NPTS(M) * nc);
    for (int l = 0; l < h; l++) {
        const double c = M->coef[l];
        #pragma omp parallel for
        for (int i = h; i < ni - h; i++) {
            for (int j = h; j < nj - h; j++) {
                for (int k = h; k < nk - h; k++) {
                    const double Vp = U[IX(M, i, j + 1 + l, k) * 3 + 1];
                    const double Vm = U[IX(M, i, j - 1 - l, k) * 3 + 1];
                    const double Up = U[IX(M, i + 1 + l, j, k) * 3 + 0];
                    const double Um = U[IX(M, i - 1 - l, j, k) * 3 + 0];
                    const double Wp = U[IX(M, i, j, k + 1 + l) * 3 + 2];
                    const double Wm = U[IX(M, i, j, k - 1 - l) * 3 + 2];
                    const double *Uc = &U[IX(M, i, j, k) * 3];
                    const double *pN = &phi[IX(M, i, j + 1 + l, k) * nc];
                    const double *pS = &phi[IX(M, i, j - 1 - l, k) * nc];
                    const double *pE = &phi[IX(M, i + 1 + l, j, k) * nc];
                    const double *pW = &phi[IX(M, i - 1 - l, j, k) * nc];
                    const double *pT = &phi[IX(M, i, j, k + 1 + l) * nc];
                    const double *pB = &phi[IX(M, i, j, k - 1 - l) * nc];
                    const double *pP = &phi[IX(M, i, j, k) * nc];
                    double *o = &out[IX(M, i, j, k) * nc];
                    for (int q = 0; q < nc; q++) { /* div.hxx:120-128 */
                        o[q] += (0.5 * c / 2.0 / (l + 1) / dy * pN[q] * Vp
                                 - 0.5 * c / 2.0 / (l + 1) / dy * pS[q] * Vm
                                 + 0.5 * c / 2.0 / (l + 1) / dx * pE[q] * Up);
                    }
                    for (int q = 0; q < nc; q++) { /* div.hxx:129-138 */
                        o[q] += (-0.5 * c / 2.0 / (l + 1) / dx * pW[q] * Um
                                 + 0.5 * c / 2.0 / (l + 1) / dz * pT[q] * Wp / dzdxi[k]
                                 - 0.5 * c / 2.0 / (l + 1) / dz * pB[q] * Wm / dzdxi[k]);
                    }
                    for (int q = 0; q < nc; q++) { /* div.hxx:141-152 */
                        o[q] += (0.5 * Uc[1] * c / 2.0 / (l + 1) / dy * pN[q]
                                 - 0.5 * Uc[1] * c / 2.0 / (l + 1) / dy * pS[q]
                                 + 0.5 * Uc[0] * c / 2.0 / (l + 1) / dx * pE[q]
                                 - 0.5 * Uc[0] * c / 2.0 / (l + 1) / dx * pW[q]
                                 + 0.5 * Uc[2] * c / 2.0 / (l + 1) / dz * pT[q] / dzdxi[k]
                                 - 0.5 * Uc[2] * c / 2.0 / (l + 1) / dz * pB[q] / dzdxi[k]);
                    }
                    for (int q = 0; q < nc; q++) { /* div.hxx:154-165 */
                        o[q] += (0.5 * pP[q] * (c / 2.0 / (l + 1) / dy * Vp)
                                 - 0.5 * pP[q] * (c / 2.0 / (l + 1) / dy * Vm)
                                 + 0.5 * pP[q] * (c / 2.0 / (l + 1) / dx * Up)
                                 - 0.5 * pP[q] * (c / 2.0 / (l + 1) / dx * Um)
                                 + 0.5 * pP[q] * (c / 2.0 / (l + 1) / dz * Wp) / dzdxi[k]
                                 - 0.5 * pP[q] * (c / 2.0 / (l + 1) / dz * Wm) / dzdxi[k]);
                    }
                }
            }
        }
    }
}

/* ex::div(Field<vector>), src/ex/div/div.hxx:22-81. */
void orc_div(const orc_mesh *M, const double *U, double *out)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    const double dx = M->d[0], dy = M->d[1], dz = M->d[2];
    const double *dzdxi = M->dzdxi;
    memset(out, 0, sizeof(double) * NPTS(M));
    for (int l = 0; l < h; l++) {
        const double c = M->coef[l];
        #pragma omp parallel for
        for (int i = h; i < ni - h; i++)
            for (int j = h; j < nj - h; j++)
                for (int k = h; k < nk - h; k++) {
                    const double Vp = U[IX(M, i, j + 1 + l, k) * 3 + 1];
                    const double Vm = U[IX(M, i, j - 1 - l, k) * 3 + 1];
                    const double Up = U[IX(M, i + 1 + l, j, k) * 3 + 0];
                    const double Um = U[IX(M, i - 1 - l, j, k) * 3 + 0];
                    const double Wp = U[IX(M, i, j, k + 1 + l) * 3 + 2];
                    const double Wm = U[IX(M, i, j, k - 1 - l) * 3 + 2];
                    double *o = &out[IX(M, i, j, k)];
                    *o += (c / 2.0 / (l + 1) / dy * Vp - c / 2.0 / (l + 1) / dy * Vm
                           + c / 2.0 / (l + 1) / dx * Up);
                    *o += (-c / 2.0 / (l + 1) / dx * Um + c / 2.0 / (l + 1) / dz * Wp / dzdxi[k]
                           - c / 2.0 / (l + 1) / dz * Wm / dzdxi[k]);
                }
    }
}

/* ex::grad(Field<scalar>), src/ex/grad/grad.cpp:25-62. */
void orc_grad(const orc_mesh *M, const double *p, double *out)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    const double dx = M->d[0], dy = M->d[1], dz = M->d[2];
    const double *dzdxi = M->dzdxi;
    memset(out, 0, sizeof(double) * NPTS(M) * 3);
    for (int l = 0; l < h; l++) {
        const double c = M->coef[l];
        #pragma omp parallel for
        for (int i = h; i < ni - h; i++)
            for (int j = h; j < nj - h; j++)
                for (int k = h; k < nk - h; k++) {
                    double *o = &out[IX(M, i, j, k) * 3];
                    o[0] += c / 2.0 / (l + 1) / dx * p[IX(M, i + 1 + l, j, k)]
                            - c / 2.0 / (l + 1) / dx * p[IX(M, i - 1 - l, j, k)];
                    o[1] += c / 2.0 / (l + 1) / dy * p[IX(M, i, j + 1 + l, k)]
                            - c / 2.0 / (l + 1) / dy * p[IX(M, i, j - 1 - l, k)];
                    o[2] += c / 2.0 / (l + 1) / dz * p[IX(M, i, j, k + 1 + l)] / dzdxi[k]
                            - c / 2.0 / (l + 1) / dz * p[IX(M, i, j, k - 1 - l)] / dzdxi[k];
                }
    }
}

/* ex::laplacian(scalar nu, Field<T>), src/ex/laplacian/laplacian.hxx:130-191. */
void orc_laplacian(const orc_mesh *M, double nu, const double *f, int nc, double *out)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    const double dx = M->d[0], dy = M->d[1], dz = M->d[2];
    const double *dzdxi = M->dzdxi, *d2 = M->d2zdxi2;
    memset(out, 0, sizeof(double) * NPTS(M) * nc);
    for (int l = 0; l < h; l++) {
        const double c = M->coef[l];
        #pragma omp parallel for
        for (int i = h; i < ni - h; i++)
            for (int j = h; j < nj - h; j++)
                for (int k = h; k < nk - h; k++) {
                    const double *fN = &f[IX(M, i, j + 1 + l, k) * nc];
                    const double *fS = &f[IX(M, i, j - 1 - l, k) * nc];
                    const double *fE = &f[IX(M, i + 1 + l, j, k) * nc];
                    const double *fW = &f[IX(M, i - 1 - l, j, k) * nc];
                    const double *fT = &f[IX(M, i, j, k + 1 + l) * nc];
                    const double *fB = &f[IX(M, i, j, k - 1 - l) * nc];
                    const double *fP = &f[IX(M, i, j, k) * nc];
                    double *o = &out[IX(M, i, j, k) * nc];
                    const double cP = (-(nu + nu) * c / pow((l + 1) * dx, 2)
                                       - (nu + nu) * c / pow((l + 1) * dy, 2)
                                       - (nu + nu) * c / pow((l + 1) * dz, 2) / pow(dzdxi[k], 2));
                    for (int q = 0; q < nc; q++) {
                        o[q] += (nu * c / pow((l + 1) * dy, 2) * fN[q]
                                 + nu * c / pow((l + 1) * dy, 2) * fS[q]
                                 + nu * c / pow((l + 1) * dx, 2) * fE[q]
                                 + nu * c / pow((l + 1) * dx, 2) * fW[q]
                                 + nu * c / pow((l + 1) * dz, 2) * fT[q] / pow(dzdxi[k], 2)
                                 + nu * c / pow((l + 1) * dz, 2) * fB[q] / pow(dzdxi[k], 2)
                                 + cP * fP[q]);
                    }
                }
    }
    for (int l = 0; l < h; l++) { /* laplacian.hxx:170-188 (z-metric term) */
        const double c = M->coef[l];
        #pragma omp parallel for
        for (int i = h; i < ni - h; i++)
            for (int j = h; j < nj - h; j++)
                for (int k = h; k < nk - h; k++) {
                    const double *fT = &f[IX(M, i, j, k + 1 + l) * nc];
                    const double *fB = &f[IX(M, i, j, k - 1 - l) * nc];
                    double *o = &out[IX(M, i, j, k) * nc];
                    for (int q = 0; q < nc; q++)
                        o[q] -= (c / 2.0 / (l + 1) / dz * fT[q] - c / 2.0 / (l + 1) / dz * fB[q])
                                * nu * d2[k] / pow(dzdxi[k], 3);
                }
    }
}

/* ex::laplacian(Field<scalar> nu, Field<T>), src/ex/laplacian/laplacian.hxx:64-127:
 * per l, each neighbour weighted by nu on its near side (nu(j+l) with
 * f(j+1+l), nu(j-1-l) with f(j-1-l)), the centre by minus the sum of the
 * pair's nu; the z terms divided by dzdxi^2; then per l minus the z first
 * derivative times nu(i,j,k) d2zdxi2 / dzdxi^3.  nu is a padded scalar. */
void orc_laplacian_var(const orc_mesh *M, const double *nu, const double *f, int nc, double *out)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    const double dx = M->d[0], dy = M->d[1], dz = M->d[2];
    const double *dzdxi = M->dzdxi, *d2 = M->d2zdxi2;
    memset(out, 0, sizeof(double) * NPTS(M) * nc);
    for (int l = 0; l < h; l++) {
        const double c = M->coef[l];
        #pragma omp parallel for
        for (int i = h; i < ni - h; i++)
            for (int j = h; j < nj - h; j++)
                for (int k = h; k < nk - h; k++) {
                    const double nN = nu[IX(M, i, j + l, k)], nS = nu[IX(M, i, j - 1 - l, k)];
                    const double nE = nu[IX(M, i + l, j, k)], nW = nu[IX(M, i - 1 - l, j, k)];
                    const double nT = nu[IX(M, i, j, k + l)], nB = nu[IX(M, i, j, k - 1 - l)];
                    const double *fN = &f[IX(M, i, j + 1 + l, k) * nc];
                    const double *fS = &f[IX(M, i, j - 1 - l, k) * nc];
                    const double *fE = &f[IX(M, i + 1 + l, j, k) * nc];
                    const double *fW = &f[IX(M, i - 1 - l, j, k) * nc];
                    const double *fT = &f[IX(M, i, j, k + 1 + l) * nc];
                    const double *fB = &f[IX(M, i, j, k - 1 - l) * nc];
                    const double *fP = &f[IX(M, i, j, k) * nc];
                    double *o = &out[IX(M, i, j, k) * nc];
                    const double cP = (-(nE + nW) * c / pow((l + 1) * dx, 2)
                                       - (nN + nS) * c / pow((l + 1) * dy, 2)
                                       - (nT + nB) * c / pow((l + 1) * dz, 2) / pow(dzdxi[k], 2));
                    for (int q = 0; q < nc; q++) {
                        o[q] += (nN * c / pow((l + 1) * dy, 2) * fN[q]
                                 + nS * c / pow((l + 1) * dy, 2) * fS[q]
                                 + nE * c / pow((l + 1) * dx, 2) * fE[q]
                                 + nW * c / pow((l + 1) * dx, 2) * fW[q]
                                 + nT * c / pow((l + 1) * dz, 2) * fT[q] / pow(dzdxi[k], 2)
                                 + nB * c / pow((l + 1) * dz, 2) * fB[q] / pow(dzdxi[k], 2)
                                 + cP * fP[q]);
                    }
                }
    }
    for (int l = 0; l < h; l++) { /* laplacian.hxx:106-124 (z-metric term) */
        const double c = M->coef[l];
        #pragma omp parallel for
        for (int i = h; i < ni - h; i++)
            for (int j = h; j < nj - h; j++)
                for (int k = h; k < nk - h; k++) {
                    const double *fT = &f[IX(M, i, j, k + 1 + l) * nc];
                    const double *fB = &f[IX(M, i, j, k - 1 - l) * nc];
                    double *o = &out[IX(M, i, j, k) * nc];
                    for (int q = 0; q < nc; q++)
                        o[q] -= (c / 2.0 / (l + 1) / dz * fT[q] - c / 2.0 / (l + 1) / dz * fB[q])
                                * nu[IX(M, i, j, k)] * d2[k] / pow(dzdxi[k], 3);
                }
    }
}

/* ------------------------------------------------------------------ */
/* Implicit matrix im::fdMatrix<scalar> (src/im/fdMatrix/fdMatrix.hxx) and
 * im::laplacian (src/im/laplacian/laplacian.hxx:23-58).  al holds the six
 * padded coefficient fields (north, south, east, west, top, bottom) one
 * after the other; ap, rhs and phi are padded scalars. */

static double *falloc(const orc_mesh *M, int nc);

#define AL(q, n) al[(size_t)(q) * NPTS(M) + (n)]

void orc_fd_laplacian(const orc_mesh *M, double mu, const double *phi, double *al, double *ap, double *rhs)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    const size_t np = NPTS(M);
    memset(al, 0, sizeof(double) * 6 * np);
    memset(ap, 0, sizeof(double) * np);
    memset(rhs, 0, sizeof(double) * np);
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                const size_t n = IX(M, i, j, k);
                AL(0, n) = mu / pow(M->d[1], 2);
                AL(1, n) = mu / pow(M->d[1], 2);
                AL(2, n) = mu / pow(M->d[0], 2);
                AL(3, n) = mu / pow(M->d[0], 2);
                AL(4, n) = mu / pow(M->d[2], 2);
                AL(5, n) = mu / pow(M->d[2], 2);
                ap[n] = -2.0 * mu / pow(M->d[0], 2) - 2.0 * mu / pow(M->d[1], 2) - 2.0 * mu / pow(M->d[2], 2);
            }
    /* deferredCorrection(ex::laplacian(mu, phi)), fdMatrix.hxx:213-245 */
    double *hoa = falloc(M, 1);
    orc_laplacian(M, mu, phi, 1, hoa);
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                const size_t n = IX(M, i, j, k);
                rhs[n] += -hoa[n] + AL(0, n) * phi[IX(M, i, j + 1, k)] + AL(1, n) * phi[IX(M, i, j - 1, k)]
                          + AL(2, n) * phi[IX(M, i + 1, j, k)] + AL(3, n) * phi[IX(M, i - 1, j, k)]
                          + AL(4, n) * phi[IX(M, i, j, k + 1)] + AL(5, n) * phi[IX(M, i, j, k - 1)]
                          + ap[n] * phi[n];
            }
    free(hoa);
}

/* fdMatrix.hxx:271-310 */
double orc_fd_residual(const orc_mesh *M, const double *al, const double *ap, const double *rhs, const double *phi)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    double res = 0.0;
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                const size_t n = IX(M, i, j, k);
                const double r = rhs[n] - AL(0, n) * phi[IX(M, i, j + 1, k)] - AL(1, n) * phi[IX(M, i, j - 1, k)]
                                 - AL(2, n) * phi[IX(M, i + 1, j, k)] - AL(3, n) * phi[IX(M, i - 1, j, k)]
                                 - AL(4, n) * phi[IX(M, i, j, k + 1)] - AL(5, n) * phi[IX(M, i, j, k - 1)]
                                 - ap[n] * phi[n];
                res += pow(r, 2);
            }
    const int cnt = (ni - M->m) * (nj - M->m) * (nk - M->m);
    return pow(res / cnt, 0.5);
}

/* tools::tdma, src/Tools/Tools.cpp:43-62 */
static void tdma(int n, const double *a, const double *b, const double *c, const double *d, double *x, double *cp,
                 double *dp)
{
    cp[0] = c[0] / b[0];
    dp[0] = d[0] / b[0];
    for (int i = 1; i < n; i++) {
        cp[i] = c[i] / (b[i] - a[i] * cp[i - 1]);
        dp[i] = (d[i] - a[i] * dp[i - 1]) / (b[i] - a[i] * cp[i - 1]);
    }
    x[n - 1] = dp[n - 1];
    for (int i = n - 2; i >= 0; i--) x[i] = dp[i] - cp[i] * x[i + 1];
}

/* fdMatrix.hxx:59-209: returns the iteration count; *res_out the last
 * residual computed. */
int orc_fd_solve(const orc_mesh *M, const double *al, const double *ap, const double *rhs, double *phi, double tol,
                 int max_its, double *res_out)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    int nmax = ni > nj ? ni : nj;
    nmax = nmax > nk ? nmax : nk;
    double *a = malloc(sizeof(double) * nmax), *b = malloc(sizeof(double) * nmax), *c = malloc(sizeof(double) * nmax);
    double *d = malloc(sizeof(double) * nmax), *x = malloc(sizeof(double) * nmax);
    double *cp = malloc(sizeof(double) * nmax), *dp = malloc(sizeof(double) * nmax);
    int its = 0;
    double res = 0.0;
    while (its < max_its) {
        res = orc_fd_residual(M, al, ap, rhs, phi);
        if (!(res > tol)) break;
        its++;
        /* j sweeps */
        for (int i = h; i < ni - h; i++)
            for (int k = h; k < nk - h; k++) {
                for (int j = 0; j < nj; j++) {
                    const int bnd = j < h || j >= nj - h - 1;
                    const size_t n = IX(M, i, j, k);
                    a[j] = bnd ? 0.0 : AL(1, n);
                    b[j] = bnd ? 1.0 : ap[n];
                    c[j] = bnd ? 0.0 : AL(0, n);
                    d[j] = bnd ? phi[n]
                               : rhs[n] - AL(2, n) * phi[IX(M, i + 1, j, k)] - AL(3, n) * phi[IX(M, i - 1, j, k)]
                                     - AL(4, n) * phi[IX(M, i, j, k + 1)] - AL(5, n) * phi[IX(M, i, j, k - 1)];
                }
                tdma(nj, a, b, c, d, x, cp, dp);
                for (int j = h; j < nj - h; j++) phi[IX(M, i, j, k)] = x[j];
            }
        /* i sweeps */
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                for (int i = 0; i < ni; i++) {
                    const int bnd = i < h || i >= ni - h - 1;
                    const size_t n = IX(M, i, j, k);
                    a[i] = bnd ? 0.0 : AL(3, n);
                    b[i] = bnd ? 1.0 : ap[n];
                    c[i] = bnd ? 0.0 : AL(2, n);
                    d[i] = bnd ? phi[n]
                               : rhs[n] - AL(0, n) * phi[IX(M, i, j + 1, k)] - AL(1, n) * phi[IX(M, i, j - 1, k)]
                                     - AL(4, n) * phi[IX(M, i, j, k + 1)] - AL(5, n) * phi[IX(M, i, j, k - 1)];
                }
                tdma(ni, a, b, c, d, x, cp, dp);
                for (int i = h; i < ni - h; i++) phi[IX(M, i, j, k)] = x[i];
            }
        /* k sweeps */
        for (int i = h; i < ni - h; i++)
            for (int j = h; j < nj - h; j++) {
                for (int k = 0; k < nk; k++) {
                    const int bnd = k < h || k >= nk - h - 1;
                    const size_t n = IX(M, i, j, k);
                    a[k] = bnd ? 0.0 : AL(5, n);
                    b[k] = bnd ? 1.0 : ap[n];
                    c[k] = bnd ? 0.0 : AL(4, n);
                    d[k] = bnd ? phi[n]
                               : rhs[n] - AL(2, n) * phi[IX(M, i + 1, j, k)] - AL(3, n) * phi[IX(M, i - 1, j, k)]
                                     - AL(0, n) * phi[IX(M, i, j + 1, k)] - AL(1, n) * phi[IX(M, i, j - 1, k)];
                }
                tdma(nk, a, b, c, d, x, cp, dp);
                for (int k = h; k < nk - h; k++) phi[IX(M, i, j, k)] = x[k];
            }
        orc_fill(M, phi, 1, -1); /* transportedFld_->correctBoundaryConditions() */
    }
    free(a); free(b); free(c); free(d); free(x); free(cp); free(dp);
    if (res_out) *res_out = res;
    return its;
}

/* relax (fdMatrix.hxx:248-253) over every padded node */
void orc_fd_relax(const orc_mesh *M, double *ap, double *rhs, const double *phi, double urf)
{
    const size_t np = NPTS(M);
    for (size_t n = 0; n < np; n++) rhs[n] += (1.0 - urf) / urf * (ap[n] * phi[n]);
    for (size_t n = 0; n < np; n++) ap[n] /= urf;
}

#undef AL

/* ex::grad(Field<vector>) -> tensor, src/ex/grad/grad.cpp:65-102 (the
 * half-shifted stencil f(i+l) - f(i-1-l)); out9 is row-major
 * (xx, xy, xz, yx, ...), row = component, column = direction. */
void orc_grad_vec_literal(const orc_mesh *M, const double *U, double *out)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    memset(out, 0, sizeof(double) * NPTS(M) * 9);
    #pragma omp parallel for
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                double *o = &out[IX(M, i, j, k) * 9];
                for (int q = 0; q < 9; q++) o[q] = 0.0;
                for (int l = 0; l < h; l++) {
                    const double c = M->coef[l];
                    const double dx = c / (l + 1) / M->d[0];
                    const double dy = c / (l + 1) / M->d[1];
                    const double dz = c / (l + 1) / M->d[2] / M->dzdxi[k];
                    for (int q = 0; q < 3; q++) {
                        o[3 * q + 0] += dx * (U[IX(M, i + l, j, k) * 3 + q] - U[IX(M, i - 1 - l, j, k) * 3 + q]);
                        o[3 * q + 1] += dy * (U[IX(M, i, j + l, k) * 3 + q] - U[IX(M, i, j - 1 - l, k) * 3 + q]);
                        o[3 * q + 2] += dz * (U[IX(M, i, j, k + l) * 3 + q] - U[IX(M, i, j, k - 1 - l) * 3 + q]);
                    }
                }
            }
}

/* ex::curl, src/ex/curl/curl.cpp:25-57: antisymmetric part of the literal
 * vector gradient. */
void orc_curl_literal(const orc_mesh *M, const double *U, double *out)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    double *g = (double *)malloc(sizeof(double) * NPTS(M) * 9);
    orc_grad_vec_literal(M, U, g);
    memset(out, 0, sizeof(double) * NPTS(M) * 3);
    #pragma omp parallel for
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                const double *t = &g[IX(M, i, j, k) * 9];
                double *o = &out[IX(M, i, j, k) * 3];
                o[0] = t[7] - t[5]; /* zy - yz */
                o[1] = t[2] - t[6]; /* xz - zx */
                o[2] = t[3] - t[1]; /* yx - xy */
            }
    free(g);
}

/* Consistent curl: central gradient of each component (the stencil of
 * grad.cpp:25-62), used for the enstrophy objective.  No reference row. */
void orc_curl_central(const orc_mesh *M, const double *U, double *out)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    memset(out, 0, sizeof(double) * NPTS(M) * 3);
    #pragma omp parallel for
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                double g[3][3] = {{0}};
                for (int l = 0; l < h; l++) {
                    const double c = M->coef[l];
                    const double ax = c / 2.0 / (l + 1) / M->d[0];
                    const double ay = c / 2.0 / (l + 1) / M->d[1];
                    const double az = c / 2.0 / (l + 1) / M->d[2] / M->dzdxi[k];
                    for (int q = 0; q < 3; q++) {
                        g[q][0] += ax * (U[IX(M, i + 1 + l, j, k) * 3 + q] - U[IX(M, i - 1 - l, j, k) * 3 + q]);
                        g[q][1] += ay * (U[IX(M, i, j + 1 + l, k) * 3 + q] - U[IX(M, i, j - 1 - l, k) * 3 + q]);
                        g[q][2] += az * (U[IX(M, i, j, k + 1 + l) * 3 + q] - U[IX(M, i, j, k - 1 - l) * 3 + q]);
                    }
                }
                double *o = &out[IX(M, i, j, k) * 3];
                o[0] = g[2][1] - g[1][2];
                o[1] = g[0][2] - g[2][0];
                o[2] = g[1][0] - g[0][1];
            }
}

/* ------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time, any length (radix 2, 3,
 * 4, 5 and a direct DFT for other prime factors).  Stands in for
 * heFFTe/FFTW (src/poisson/poisson.cpp:109,225,250): forward uses e^{-i},
 * backward e^{+i}, neither normalises. */
static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

static void fft_rec(int n, int stride, const double complex *in, double complex *out,
                    int sign, double complex *tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n);
    int m = n / p;
    for (int r = 0; r < p; r++) fft_rec(m, stride * p, in + (size_t)r * stride, out + (size_t)r * m, sign, tmp);
    /* combine: X[k + q m] = sum_r W_n^{r (k + q m)} Y_r[k] */
    const double ang = sign * 2.0 * ORC_PI / n;
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            long e = (long)r * k;
            double complex w = (e == 0) ? 1.0 : (cos(ang * e) + I * sin(ang * e));
            tmp[r] = out[(size_t)r * m + k] * w;
        }
        for (int q = 0; q < p; q++) {
            double complex s = 0;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * q) % p;
                double complex w = (e == 0) ? 1.0 : (cos(2.0 * ORC_PI * sign * e / p) + I * sin(2.0 * ORC_PI * sign * e / p));
                s += tmp[r] * w;
            }
            out[(size_t)k + (size_t)q * m] = s;
        }
    }
}

void orc_fft1d(int n, double *data, int sign)
{
    double complex *x = (double complex *)data;
    double complex *y = (double complex *)malloc(sizeof(double complex) * (size_t)n);
    double complex tmp[64];
    double complex *t = n > 64 ? (double complex *)malloc(sizeof(double complex) * (size_t)n) : tmp;
    fft_rec(n, 1, x, y, sign, t);
    memcpy(x, y, sizeof(double complex) * (size_t)n);
    if (t != tmp) free(t);
    free(y);
}

void orc_fft3d(const int n[3], double *data, int sign)
{
    double complex *x = (double complex *)data;
    const size_t n0 = n[0], n1 = n[1], n2 = n[2];
    for (int ax = 0; ax < 3; ax++) {
        const size_t len = (size_t)n[ax];
        const size_t stride = ax == 0 ? 1 : (ax == 1 ? n0 : n0 * n1);
        const size_t nlines = n0 * n1 * n2 / len;
        #pragma omp parallel
        {
            double complex *a = (double complex *)malloc(sizeof(double complex) * len);
            double complex *b = (double complex *)malloc(sizeof(double complex) * len);
            double complex *t = (double complex *)malloc(sizeof(double complex) * len);
            #pragma omp for
            for (size_t ln = 0; ln < nlines; ln++) {
                size_t base;
                if (ax == 0) base = ln * n0;
                else if (ax == 1) base = (ln % n0) + (ln / n0) * n0 * n1;
                else base = ln;
                for (size_t q = 0; q < len; q++) a[q] = x[base + q * stride];
                fft_rec((int)len, 1, a, b, sign, t);
                for (size_t q = 0; q < len; q++) x[base + q * stride] = b[q];
            }
            free(a); free(b); free(t);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Poisson.  poisson::wavenumbersFFT_ (poisson.cpp:337-396) with
 * localStart = 0; with y walls the solve runs on the even extension of
 * the y direction (period 2 L_y, 2 N_y points), which is the DCT-I solve
 * for even-reflected data (no reference row: solveDCT, poisson.cpp:268-335,
 * is not a Poisson inverse). */
static double kmod(double kk, double h, const double *coef, int hm)
{
    double kp = 0.0;
    if (kk > ORC_EPS)
        for (int ll = 0; ll < hm; ll++)
            kp += kk * coef[ll] / (ll + 1) * sin((ll + 1) * kk * h) / (kk * h);
    return kp;
}

orc_poisson *orc_poisson_create(const orc_mesh *M)
{
    orc_poisson *P = (orc_poisson *)calloc(1, sizeof(orc_poisson));
    P->mesh = M;
    for (int d = 0; d < 3; d++) { P->N[d] = M->N[d]; P->M[d] = M->N[d]; }
    P->ext = M->wall_y;
    double Ly = M->L[1];
    if (P->ext) { P->M[1] = 2 * M->N[1]; Ly = 2.0 * M->L[1]; }
    const int M0 = P->M[0], M1 = P->M[1], M2 = P->M[2];
    P->k2 = (double *)malloc(sizeof(double) * (size_t)M0 * M1 * M2);
    P->buf = (double complex *)malloc(sizeof(double complex) * (size_t)M0 * M1 * M2);
    const double pi = ORC_PI;
    size_t l = 0;
    for (int k = 0; k < M2; k++)
        for (int j = 0; j < M1; j++)
            for (int i = 0; i < M0; i++) {
                double kx, ky, kz;
                if (2 * i < M0) kx = 2.0 * pi * i / M->L[0];
                else kx = 2.0 * pi * (M0 - i) / M->L[0];
                if (2 * j < M1) ky = 2.0 * pi * j / Ly;
                else ky = 2.0 * pi * (M1 - j) / Ly;
                kz = 2.0 * pi * k / M->L[2];
                double kpx = kmod(kx, M->d[0], M->coef, M->h);
                double kpy = kmod(ky, M->d[1], M->coef, M->h);
                double kpz = kmod(kz, M->d[2], M->coef, M->h);
                P->k2[l++] = kpx * kpx + kpy * kpy + kpz * kpz;
            }
    return P;
}

void orc_poisson_free(orc_poisson *P)
{
    if (!P) return;
    free(P->k2);
    free(P->buf);
    free(P);
}

void orc_poisson_k2(const orc_poisson *P, double *out)
{
    memcpy(out, P->k2, sizeof(double) * (size_t)P->M[0] * P->M[1] * P->M[2]);
}

/* poisson::rhs + poisson::solve, poisson.cpp:167-181,223-265. */
void orc_poisson_solve(orc_poisson *P, const double *rhs, double *p)
{
    const orc_mesh *M = P->mesh;
    const int h = M->h;
    const int M0 = P->M[0], M1 = P->M[1], M2 = P->M[2];
    const int Ny = M->N[1];
    double complex *phi = P->buf;
    /* gather (rhs over [h, n-1-h), i fastest); with walls, the y line is
     * extended evenly: index j in [0, 2Ny) maps to node min(j, 2Ny - j). */
    #pragma omp parallel for
    for (int k = 0; k < M2; k++)
        for (int j = 0; j < M1; j++) {
            int jj = j;
            if (P->ext && j > Ny) jj = 2 * Ny - j;
            for (int i = 0; i < M0; i++)
                phi[(size_t)i + (size_t)M0 * ((size_t)j + (size_t)M1 * k)] = rhs[IX(M, i + h, jj + h, k + h)];
        }
    orc_fft3d(P->M, (double *)phi, -1);
    size_t total = (size_t)M0 * M1 * M2;
    #pragma omp parallel for
    for (size_t l = 0; l < total; l++) { /* poisson.cpp:231-249 */
        if (P->k2[l] < ORC_EPS) phi[l] = 0.0;
        else phi[l] = (-creal(phi[l]) / P->k2[l]) + I * (-cimag(phi[l]) / P->k2[l]);
    }
    orc_fft3d(P->M, (double *)phi, +1);
    const int jmax = P->ext ? Ny + 1 : M1;
    const double norm = (double)((long)M0 * (long)M1 * (long)M2);
    #pragma omp parallel for
    for (int k = 0; k < M2; k++) /* poisson.cpp:252-264 */
        for (int j = 0; j < jmax; j++)
            for (int i = 0; i < M0; i++)
                p[IX(M, i + h, j + h, k + h)] = creal(phi[(size_t)i + (size_t)M0 * ((size_t)j + (size_t)M1 * k)]) / norm;
}

/* ------------------------------------------------------------------ */
/* Field algebra helpers (src/Field/Field.hxx:617-1150: elementwise over
 * the whole padded array). */
static double *falloc(const orc_mesh *M, int nc)
{
    return (double *)calloc(NPTS(M) * (size_t)nc, sizeof(double));
}

/* dU = dt * ((-div(U,U)) + laplacian(nu,U)) [+ div(B,B) - grad(0.5*(B&B))]
 * (app/TG_Piosson/RKLoop.H:17-22, app/OT_Poisson/UEqn.H:16-23). */
static void rhs_u(const orc_mesh *M, const double *U, const double *B, double nu, double dt,
                  double *dU, double *t1, double *t2, double *s1, double *s2)
{
    const size_t n3 = NPTS(M) * 3;
    orc_div_adv(M, U, U, 3, t1);
    orc_laplacian(M, nu, U, 3, t2);
    #pragma omp parallel for
    for (size_t q = 0; q < n3; q++) dU[q] = (-t1[q]) + t2[q];
    if (B) {
        orc_div_adv(M, B, B, 3, t1);
        const size_t n1 = NPTS(M);
        #pragma omp parallel for
        for (size_t q = 0; q < n1; q++) /* vector.h inner product, then 0.5* */
            s1[q] = 0.5 * (B[3 * q] * B[3 * q] + B[3 * q + 1] * B[3 * q + 1] + B[3 * q + 2] * B[3 * q + 2]);
        orc_grad(M, s1, t2);
        #pragma omp parallel for
        for (size_t q = 0; q < n3; q++) {
            if (ORC_PLANT == 1) dU[q] = (dU[q] - t1[q]) - t2[q];      /* Lorentz div(B,B) sign */
            else if (ORC_PLANT == 2) dU[q] = dU[q] - t2[q];          /* div(B,B) dropped */
            else if (ORC_PLANT == 3) dU[q] = (dU[q] + t1[q]) + t2[q]; /* magnetic pressure sign */
            else dU[q] = (dU[q] + t1[q]) - t2[q];
        }
    }
    (void)s2;
    #pragma omp parallel for
    for (size_t q = 0; q < n3; q++) dU[q] = dt * dU[q];
}

/* dB = dt * ((-div(U,B) + div(B,U)) + laplacian(eta,B)),
 * app/OT_Poisson/BEqn.H:15-23. */
static void rhs_b(const orc_mesh *M, const double *U, const double *B, double eta, double dt,
                  double *dB, double *t1, double *t2)
{
    const size_t n3 = NPTS(M) * 3;
    orc_div_adv(M, U, B, 3, t1);
    orc_div_adv(M, B, U, 3, t2);
    #pragma omp parallel for
    for (size_t q = 0; q < n3; q++) dB[q] = ORC_PLANT == 4 ? t1[q] - t2[q] : (-t1[q]) + t2[q]; /* 4: induction sign */
    orc_laplacian(M, eta, B, 3, t1);
    #pragma omp parallel for
    for (size_t q = 0; q < n3; q++) dB[q] = dt * (dB[q] + t1[q]);
}

/* The two right-hand sides of app/OT_Poisson/UEqn.H:16-23 and BEqn.H:15-23
 * per unit dt, for the pins: dU = -div(U,U) + laplacian(nu,U) + div(B,B)
 * - grad(0.5*(B&B)), dB = -div(U,B) + div(B,U) + laplacian(eta,B). */
void orc_mhd_rhs(const orc_mesh *M, const double *U, const double *B, double nu, double eta, double *dU,
                 double *dB)
{
    double *t1 = falloc(M, 3), *t2 = falloc(M, 3), *s1 = falloc(M, 1);
    rhs_u(M, U, B, nu, 1.0, dU, t1, t2, s1, NULL);
    rhs_b(M, U, B, eta, 1.0, dB, t1, t2);
    free(t1); free(t2); free(s1);
}

/* Projection F <- F - s*grad(p), lap p = div(F)/s (RKLoop.H:30-38 with
 * s = dt; BEqn.H:37-41 with s = 1). */
void orc_project(const orc_mesh *M, orc_poisson *P, double *F, double *p, double s)
{
    const size_t n1 = NPTS(M), n3 = n1 * 3;
    double *d = falloc(M, 1);
    double *g = falloc(M, 3);
    orc_fill(M, F, 3, 1);
    orc_div(M, F, d);
    #pragma omp parallel for
    for (size_t q = 0; q < n1; q++) d[q] = d[q] / s;
    orc_poisson_solve(P, d, p);
    orc_fill(M, p, 1, -1);
    orc_grad(M, p, g);
    #pragma omp parallel for
    for (size_t q = 0; q < n3; q++) F[q] = F[q] - s * g[q];
    orc_fill(M, F, 3, 1);
    free(d);
    free(g);
}

void orc_ns_step(const orc_mesh *M, orc_poisson *P, double *U, double *Ustar, double *p,
                 double nu, double dt, int mode)
{
    const size_t n3 = NPTS(M) * 3;
    double *U0 = falloc(M, 3), *dU = falloc(M, 3), *t1 = falloc(M, 3), *t2 = falloc(M, 3);
    memcpy(U0, U, sizeof(double) * n3);
    if (mode == ORC_RK_LITERAL4) {
        /* RKLoop.H:15-29: every stage evaluates the RHS at the same U, so
         * the four dU are equal; the last stage's accumulation is kept. */
        rhs_u(M, U, NULL, nu, dt, dU, t1, t2, NULL, NULL);
        #pragma omp parallel for
        for (size_t q = 0; q < n3; q++) {
            double s = U0[q];
            for (int j = 0; j < 4; j++) s = s + RK4W[j] * dU[q];
            Ustar[q] = s;
        }
        orc_project(M, P, Ustar, p, dt);
        memcpy(U, Ustar, sizeof(double) * n3);
    } else if (mode == ORC_RK_SSP3) {
        /* U1 = U0 + L(U0); U2 = 3/4 U0 + 1/4 (U1 + L(U1));
         * U3 = 1/3 U0 + 2/3 (U2 + L(U2)); then one projection. */
        double *Us = Ustar;
        memcpy(Us, U0, sizeof(double) * n3);
        for (int s = 0; s < 3; s++) {
            rhs_u(M, Us, NULL, nu, dt, dU, t1, t2, NULL, NULL);
            #pragma omp parallel for
            for (size_t q = 0; q < n3; q++) {
                double v = Us[q] + dU[q];
                if (s == 0) Us[q] = v;
                else if (s == 1) Us[q] = 0.75 * U0[q] + 0.25 * v;
                else Us[q] = (1.0 / 3.0) * U0[q] + (2.0 / 3.0) * v;
            }
            orc_fill(M, Us, 3, 1);
        }
        orc_project(M, P, Us, p, dt);
        memcpy(U, Us, sizeof(double) * n3);
    } else { /* ORC_RK_SHEAR4: app/shearLayer/UEqn.H:14-38 */
        for (int rk = 0; rk < 4; rk++) {
            rhs_u(M, U, NULL, nu, dt, dU, t1, t2, NULL, NULL);
            #pragma omp parallel for
            for (size_t q = 0; q < n3; q++) Ustar[q] = U0[q] + RK4W[rk] * dU[q];
            orc_project(M, P, Ustar, p, dt * RK4C[rk]);
            memcpy(U, Ustar, sizeof(double) * n3);
        }
    }
    free(U0); free(dU); free(t1); free(t2);
}

/* CPU baseline variant (ii), "fused": the SSP-RK3 step of orc_ns_step with
 * the same operators evaluated one node at a time in a single OpenMP pass
 * per stage (weights precomputed, every term of -div(U,U) + laplacian(nu,U)
 * and the stage combination accumulated in registers, the output written
 * once), the projection with one fused divergence pass, the same FFT solve
 * and one fused correction pass.  No reference row: it times the
 * reference's arithmetic without its operator-at-a-time memory traffic.
 * Agrees with orc_ns_step(..., ORC_RK_SSP3) to rounding. */
static void fused_ns_stage(const orc_mesh *M, const double *X, const double *U0, double *out, double nu,
                           double dt, double alpha, double beta)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    double a[3][4], b[3][4];
    for (int l = 0; l < h; l++)
        for (int d = 0; d < 3; d++) {
            a[d][l] = M->coef[l] / 2.0 / (l + 1) / M->d[d];
            b[d][l] = M->coef[l] / (((l + 1) * M->d[d]) * ((l + 1) * M->d[d]));
        }
    const long sj = (long)nk, si = (long)nj * nk, sk = 1;
    const long st[3] = {si, sj, sk};
    #pragma omp parallel for collapse(2)
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                const long o = (long)IX(M, i, j, k);
                const double zi = 1.0 / M->dzdxi[k], zl = zi * zi;
                const double zm = M->d2zdxi2[k] * zi * zi * zi;
                const double *x0 = &X[o * 3];
                double adv[3] = {0, 0, 0}, lap[3] = {0, 0, 0}, dz1[3] = {0, 0, 0};
                for (int d = 0; d < 3; d++) {
                    const double sa = d == 2 ? zi : 1.0, sb = d == 2 ? zl : 1.0;
                    for (int l = 0; l < h; l++) {
                        const double *xp = &X[(o + (l + 1) * st[d]) * 3], *xm = &X[(o - (l + 1) * st[d]) * 3];
                        for (int q = 0; q < 3; q++) {
                            adv[q] += sa * a[d][l] * ((xp[q] * xp[d] - xm[q] * xm[d]) + x0[d] * (xp[q] - xm[q])
                                                      + x0[q] * (xp[d] - xm[d]));
                            lap[q] += sb * b[d][l] * ((xp[q] + xm[q]) - 2.0 * x0[q]);
                            if (d == 2) dz1[q] += a[2][l] * (xp[q] - xm[q]);
                        }
                    }
                }
                for (int q = 0; q < 3; q++) {
                    const double du = dt * ((-(0.5 * adv[q])) + nu * (lap[q] - zm * dz1[q]));
                    const double v = x0[q] + du;
                    out[o * 3 + q] = alpha != 0.0 ? alpha * U0[o * 3 + q] + beta * v : v;
                }
            }
}

void orc_ns_step_fused(const orc_mesh *M, orc_poisson *P, double *U, double *W1, double *W2, double *p, double nu,
                       double dt)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    fused_ns_stage(M, U, U, W1, nu, dt, 0.0, 1.0);
    orc_fill(M, W1, 3, 1);
    fused_ns_stage(M, W1, U, W2, nu, dt, 0.75, 0.25);
    orc_fill(M, W2, 3, 1);
    fused_ns_stage(M, W2, U, W1, nu, dt, 1.0 / 3.0, 2.0 / 3.0);
    orc_fill(M, W1, 3, 1);
    /* projection: divergence / dt in one pass, the FFT solve, one correction pass */
    double *dv = falloc(M, 1);
    const long st[3] = {(long)nj * nk, (long)nk, 1};
    double a[3][4];
    for (int l = 0; l < h; l++)
        for (int d = 0; d < 3; d++) a[d][l] = M->coef[l] / 2.0 / (l + 1) / M->d[d];
    #pragma omp parallel for collapse(2)
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                const long o = (long)IX(M, i, j, k);
                double s = 0.0;
                for (int d = 0; d < 3; d++) {
                    double t = 0.0;
                    for (int l = 0; l < h; l++) t += a[d][l] * (W1[(o + (l + 1) * st[d]) * 3 + d] - W1[(o - (l + 1) * st[d]) * 3 + d]);
                    s += d == 2 ? t / M->dzdxi[k] : t;
                }
                dv[o] = s / dt;
            }
    orc_poisson_solve(P, dv, p);
    orc_fill(M, p, 1, -1);
    #pragma omp parallel for collapse(2)
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                const long o = (long)IX(M, i, j, k);
                for (int d = 0; d < 3; d++) {
                    double t = 0.0;
                    for (int l = 0; l < h; l++) t += a[d][l] * (p[o + (l + 1) * st[d]] - p[o - (l + 1) * st[d]]);
                    if (d == 2) t /= M->dzdxi[k];
                    U[o * 3 + d] = W1[o * 3 + d] - dt * t;
                }
            }
    orc_fill(M, U, 3, 1);
    free(dv);
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void orc_mhd_step(const orc_mesh *M, orc_poisson *P, double *U, double *B, double *p,
                  double *pB, double nu, double eta, double dt, int mode)
{
    const size_t n1 = NPTS(M), n3 = n1 * 3;
    double *U0 = falloc(M, 3), *B0 = falloc(M, 3), *dU = falloc(M, 3), *dB[4];
    double *t1 = falloc(M, 3), *t2 = falloc(M, 3), *s1 = falloc(M, 1);
    for (int r = 0; r < 4; r++) dB[r] = falloc(M, 3);
    memcpy(U0, U, sizeof(double) * n3);
    memcpy(B0, B, sizeof(double) * n3);
    if (mode == 0) {
        /* UEqn.H:14-39 (frozen-stage RK4 + projection) */
        rhs_u(M, U, B, nu, dt, dU, t1, t2, s1, NULL);
        #pragma omp parallel for
        for (size_t q = 0; q < n3; q++) {
            double s = U0[q];
            for (int j = 0; j < 4; j++) s = s + RK4W[j] * dU[q];
            U[q] = s;
        }
        orc_project(M, P, U, p, dt);
        /* BEqn.H:15-41 (staged RK4 of B with U frozen, then cleaning) */
        static const double A[4][4] = {{0.5, 0, 0, 0}, {0, 0.5, 0, 0}, {0, 0, 1.0, 0},
                                       {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0}};
        static const double Aw[4][4] = {{0.5, 0, 0, 0}, {0, 0.5, 0, 0}, {0, 0, 1.0, 0},
                                        {0.25, 0.25, 0.25, 0.25}};
        for (int rk = 0; rk < 4; rk++) {
            /* 5: the B stages see the step-start U instead of the projected one;
             * 6: the last row of RK4a replaced */
            rhs_b(M, ORC_PLANT == 5 ? U0 : U, B, eta, dt, dB[rk], t1, t2);
            #pragma omp parallel for
            for (size_t q = 0; q < n3; q++) {
                double s = B0[q];
                for (int j = 0; j <= rk; j++) s = s + (ORC_PLANT == 6 ? Aw[rk][j] : A[rk][j]) * dB[j][q];
                B[q] = s;
            }
            orc_fill(M, B, 3, 1);
        }
        orc_project(M, P, B, pB, 1.0);
    } else {
        double *Bs = dB[1], *Ut = dB[2], *Bt = dB[3];
        for (int s = 0; s < 3; s++) {
            rhs_u(M, U, B, nu, dt, dU, t1, t2, s1, NULL);
            rhs_b(M, U, B, eta, dt, dB[0], t1, t2);
            #pragma omp parallel for
            for (size_t q = 0; q < n3; q++) {
                double vu = U[q] + dU[q], vb = B[q] + dB[0][q];
                if (s == 0) { Ut[q] = vu; Bt[q] = vb; }
                else if (s == 1 && ORC_PLANT == 7) { Ut[q] = 0.25 * U0[q] + 0.75 * vu; Bt[q] = 0.25 * B0[q] + 0.75 * vb; }
                else if (s == 1) { Ut[q] = 0.75 * U0[q] + 0.25 * vu; Bt[q] = 0.75 * B0[q] + 0.25 * vb; }
                else { Ut[q] = (1.0 / 3.0) * U0[q] + (2.0 / 3.0) * vu; Bt[q] = (1.0 / 3.0) * B0[q] + (2.0 / 3.0) * vb; }
            }
            memcpy(U, Ut, sizeof(double) * n3);
            memcpy(B, Bt, sizeof(double) * n3);
            orc_project(M, P, U, p, dt);
            orc_project(M, P, B, pB, 1.0);
        }
        (void)Bs;
    }
    free(U0); free(B0); free(dU); free(t1); free(t2); free(s1);
    for (int r = 0; r < 4; r++) free(dB[r]);
}

/* The weakly-compressible equation blocks.  Field algebra acts on every
 * node of the padded array (Field.hxx:617-1150), the operators on the
 * interior with zero ghosts, and the advanced field is refilled after every
 * stage.  Stage s starts from the step-start value plus (int(s/2)+1)/2 of
 * the previous increment; the accumulator adds RK4W[s] of each increment. */

/* rhoEqn.H:10-35: classic RK4 of rho with U frozen, drho = dt (-div(rho U)). */
static void wc_rho_eqn(const orc_mesh *M, const double *U, double *rho, double dt)
{
    const size_t n1 = NPTS(M);
    double *rU = falloc(M, 3), *dv = falloc(M, 1), *dr = falloc(M, 1), *rc = falloc(M, 1), *r0 = falloc(M, 1);
    memcpy(r0, rho, sizeof(double) * n1);
    memcpy(rc, rho, sizeof(double) * n1);
    for (int rk = 0; rk < 4; rk++) {
        #pragma omp parallel for
        for (size_t q = 0; q < n1; q++)
            for (int c = 0; c < 3; c++) rU[3 * q + c] = rho[q] * U[3 * q + c];
        orc_div(M, rU, dv);
        #pragma omp parallel for
        for (size_t q = 0; q < n1; q++) {
            dr[q] = dt * (-dv[q]);
            rc[q] = rc[q] + RK4W[rk] * dr[q];
        }
        if (rk < 3) {
            const double cs = (rk / 2 + 1) / 2.0;
            #pragma omp parallel for
            for (size_t q = 0; q < n1; q++) rho[q] = r0[q] + cs * dr[q];
        } else {
            memcpy(rho, rc, sizeof(double) * n1);
        }
        orc_fill(M, rho, 1, -1);
    }
    free(rU); free(dv); free(dr); free(rc); free(r0);
}

/* UEqn.H (app/TG:12-35, app/OT:12-35, app/shearLayerWC:12-40): classic RK4
 * of rho U with rho, p (and B, T) frozen:
 * drhoU = dt (((-div(rho U, U) - grad p) + laplacian(mu, U)) [- div((2 DBU)
 * B, B) + grad(DBU (B & B))] [- (rho T) g]). */
static void wc_u_eqn(const orc_mesh *M, double *U, const double *rho, const double *p, const double *B, double dbu,
                     const double *T, const double g[3], double mu, double dt)
{
    const size_t n1 = NPTS(M), n3 = n1 * 3;
    double *rU = falloc(M, 3), *t1 = falloc(M, 3), *t2 = falloc(M, 3), *gp = falloc(M, 3), *dU = falloc(M, 3);
    double *rUc = falloc(M, 3), *rU0 = falloc(M, 3);
    double *t3 = NULL, *t4 = NULL;
    if (B) {
        double *X = falloc(M, 3), *s1 = falloc(M, 1);
        t3 = falloc(M, 3);
        t4 = falloc(M, 3);
        #pragma omp parallel for
        for (size_t q = 0; q < n1; q++) {
            for (int c = 0; c < 3; c++) X[3 * q + c] = B[3 * q + c] * (2 * dbu);
            s1[q] = (B[3 * q] * B[3 * q] + B[3 * q + 1] * B[3 * q + 1] + B[3 * q + 2] * B[3 * q + 2]) * dbu;
        }
        orc_div_adv(M, X, B, 3, t3);
        orc_grad(M, s1, t4);
        free(X); free(s1);
    }
    #pragma omp parallel for
    for (size_t q = 0; q < n1; q++)
        for (int c = 0; c < 3; c++) rUc[3 * q + c] = rU0[3 * q + c] = rho[q] * U[3 * q + c];
    orc_grad(M, p, gp);
    for (int rk = 0; rk < 4; rk++) {
        #pragma omp parallel for
        for (size_t q = 0; q < n1; q++)
            for (int c = 0; c < 3; c++) rU[3 * q + c] = rho[q] * U[3 * q + c];
        orc_div_adv(M, rU, U, 3, t1);
        orc_laplacian(M, mu, U, 3, t2);
        #pragma omp parallel for
        for (size_t q = 0; q < n3; q++) {
            double t = ((-t1[q]) - gp[q]) + t2[q];
            if (B) t = (t - t3[q]) + t4[q];
            if (T) t = t - (rho[q / 3] * T[q / 3]) * g[q % 3];
            dU[q] = dt * t;
            rUc[q] = rUc[q] + RK4W[rk] * dU[q];
        }
        const double cs = (rk / 2 + 1) / 2.0;
        #pragma omp parallel for
        for (size_t q = 0; q < n1; q++)
            for (int c = 0; c < 3; c++)
                U[3 * q + c] = rk < 3 ? rU0[3 * q + c] / rho[q] + cs * (dU[3 * q + c] / rho[q])
                                      : rUc[3 * q + c] / rho[q];
        orc_fill(M, U, 3, 1);
    }
    free(rU); free(t1); free(t2); free(gp); free(dU); free(rUc); free(rU0); free(t3); free(t4);
}

/* TEqn.H (app/shearLayerWC:12-36): classic RK4 of rho T with rho and U
 * frozen, drhoT = dt (-div(rho U, T) + laplacian(ka, T)). */
static void wc_t_eqn(const orc_mesh *M, const double *U, const double *rho, double *T, double ka, double dt)
{
    const size_t n1 = NPTS(M);
    double *rU = falloc(M, 3), *t1 = falloc(M, 1), *t2 = falloc(M, 1), *dT = falloc(M, 1);
    double *rTc = falloc(M, 1), *rT0 = falloc(M, 1);
    #pragma omp parallel for
    for (size_t q = 0; q < n1; q++) {
        rTc[q] = rT0[q] = rho[q] * T[q];
        for (int c = 0; c < 3; c++) rU[3 * q + c] = rho[q] * U[3 * q + c];
    }
    for (int rk = 0; rk < 4; rk++) {
        orc_div_adv(M, rU, T, 1, t1);
        orc_laplacian(M, ka, T, 1, t2);
        #pragma omp parallel for
        for (size_t q = 0; q < n1; q++) {
            dT[q] = dt * ((-t1[q]) + t2[q]);
            rTc[q] = rTc[q] + RK4W[rk] * dT[q];
        }
        const double cs = (rk / 2 + 1) / 2.0;
        #pragma omp parallel for
        for (size_t q = 0; q < n1; q++) T[q] = rk < 3 ? rT0[q] / rho[q] + cs * (dT[q] / rho[q]) : rTc[q] / rho[q];
        orc_fill(M, T, 1, -1);
    }
    free(rU); free(t1); free(t2); free(dT); free(rTc); free(rT0);
}

/* Weakly-compressible step of app/TG (TG.cpp:71-76) and app/OT (OT.cpp:92-95,
 * B non-null; BEqn.H never assigns B, so it stays frozen): rhoEqn.H, pEqn.H
 * (p = pc rho^2; pc = 50 for c = 10), UEqn.H. */
void orc_wc_step(const orc_mesh *M, double *U, double *rho, double *p, const double *B, double mu, double pc,
                 double dbu, double dt)
{
    const size_t n1 = NPTS(M);
    const double g0[3] = {0.0, 0.0, 0.0};
    wc_rho_eqn(M, U, rho, dt);
    #pragma omp parallel for
    for (size_t q = 0; q < n1; q++) p[q] = pc * (rho[q] * rho[q]);
    wc_u_eqn(M, U, rho, p, B, dbu, NULL, g0, mu, dt);
}

/* Weakly-compressible step of app/shearLayerWC (shearLayer.cpp:197-200):
 * UEqn.H with the buoyancy -(rho T) g, rhoEqn.H, pEqn.H (p = pc rho^2 with
 * pc = c^2/(2 rho0)), TEqn.H.  p enters UEqn as given (the previous
 * step's). */
void orc_wcs_step(const orc_mesh *M, double *U, double *rho, double *p, double *T, double mu, double ka,
                  const double g[3], double pc, double dt)
{
    const size_t n1 = NPTS(M);
    wc_u_eqn(M, U, rho, p, NULL, 0.0, T, g, mu, dt);
    wc_rho_eqn(M, U, rho, dt);
    #pragma omp parallel for
    for (size_t q = 0; q < n1; q++) p[q] = (pc * rho[q]) * rho[q]; /* pEqn.H:2, (c*c/2.0/rho0)*rho*rho */
    wc_t_eqn(M, U, rho, T, ka, dt);
}

/* app/shearLayerWC/hydrostaticP.H:1-37 and shearLayer.cpp:141-151.  The
 * column total of hydrostaticP.H:7-10 only feeds the other ranks' offset,
 * which the reference leaves out of pHyd (:28), so it is not formed. */
void orc_wc_hydrostatic(const orc_mesh *M, const double *T, double *p, double *rho, const double g[3],
                        double rho0, double pc)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    const double dz = M->d[2];
    #pragma omp parallel for
    for (int i = h; i < ni - h - 1; i++)
        for (int j = h; j < nj - h; j++) {
            double acc = 0.0;
            for (int k = nk - 2 - h; k >= h; k--) {
                acc += T[IX(M, i, j, k)] * dz * g[2] * rho0;
                p[IX(M, i, j, k)] += acc;
            }
        }
    #pragma omp parallel for
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) rho[IX(M, i, j, k)] = pow(p[IX(M, i, j, k)] / pc, 0.5);
}

/* ------------------------------------------------------------------ */
/* EnSight Gold output (src/Field/Field.hxx:403-522, src/Mesh/Mesh.cpp:224-365). */

/* An 80-byte string as the reference's buf.str().resize(80) writes it:
 * the text and its NUL are defined there, the rest is unspecified; zeros
 * here. */
static int put80(FILE *f, const char *txt)
{
    char b[80];
    memset(b, 0, sizeof(b));
    strncpy(b, txt, 79);
    return fwrite(b, 1, 80, f) == 80 ? 0 : -1;
}

/* Field.hxx:403-451: path/name.%04d.%05d.dat (processor, file number);
 * description, part, int 1, point; then every component's interior nodes
 * [h, n-h) as float, i fastest, then j, then k. */
int orc_field_write(const orc_mesh *M, const double *f, int nc, const char *path, const char *name, int procNo,
                    int fileNum)
{
    char fn[4096];
    snprintf(fn, sizeof(fn), "%s/%s.%04d.%05d.dat", path, name, procNo, fileNum);
    FILE *fo = fopen(fn, "wb");
    if (!fo) return -1;
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    int partNum = 1, rc = 0;
    rc |= put80(fo, "description");
    rc |= put80(fo, "part");
    rc |= fwrite(&partNum, sizeof(int), 1, fo) == 1 ? 0 : -1;
    rc |= put80(fo, "point");
    for (int c = 0; c < nc; c++)
        for (int k = h; k < nk - h; k++)
            for (int j = h; j < nj - h; j++)
                for (int i = h; i < ni - h; i++) {
                    float v = (float)f[IX(M, i, j, k) * nc + c];
                    rc |= fwrite(&v, sizeof(float), 1, fo) == 1 ? 0 : -1;
                }
    if (fclose(fo) != 0) rc = -1;
    return rc;
}

/* Field.hxx:455-491: skip 240 + sizeof(int) bytes, read the same nodes. */
int orc_field_read(const orc_mesh *M, double *f, int nc, const char *path, const char *name, int procNo,
                   int fileNum)
{
    char fn[4096];
    snprintf(fn, sizeof(fn), "%s/%s.%04d.%05d.dat", path, name, procNo, fileNum);
    FILE *fi = fopen(fn, "rb");
    if (!fi) return -1;
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    int rc = fseek(fi, 240 + sizeof(int), SEEK_SET) == 0 ? 0 : -1;
    for (int c = 0; c < nc && !rc; c++)
        for (int k = h; k < nk - h && !rc; k++)
            for (int j = h; j < nj - h && !rc; j++)
                for (int i = h; i < ni - h; i++) {
                    float v;
                    if (fread(&v, sizeof(float), 1, fi) != 1) {
                        rc = -1;
                        break;
                    }
                    f[IX(M, i, j, k) * nc + c] = (double)v;
                }
    fclose(fi);
    return rc;
}

/* Mesh::write (Mesh.cpp:224-303), writeCase (:306-347), writeRestartInfo
 * (:349-365).  The output times print as the reference's ofstream prints a
 * double by default, i.e. as printf's %g. */
int orc_mesh_write(const orc_mesh *M, const char *path, int procNo, const double *steps, int nsteps)
{
    char fn[4096];
    snprintf(fn, sizeof(fn), "%s/geom.p%04d.geo", path, procNo);
    FILE *fo = fopen(fn, "wb");
    if (!fo) return -1;
    int rc = 0, partNum = 1;
    rc |= put80(fo, "C Binary");
    rc |= put80(fo, "desc1");
    rc |= put80(fo, "desc2");
    rc |= put80(fo, "node id assign");
    rc |= put80(fo, "element id assign");
    rc |= put80(fo, "part");
    rc |= fwrite(&partNum, sizeof(int), 1, fo) == 1 ? 0 : -1;
    rc |= put80(fo, "part desc");
    rc |= put80(fo, "block rectilinear");
    int np[3];
    for (int a = 0; a < 3; a++) np[a] = M->n[a] - M->m;
    rc |= fwrite(np, sizeof(int), 3, fo) == 3 ? 0 : -1;
    for (int i = 0; i < np[0]; i++) {
        float x = M->o[0] + M->d[0] * i;
        rc |= fwrite(&x, sizeof(float), 1, fo) == 1 ? 0 : -1;
    }
    for (int j = 0; j < np[1]; j++) {
        float y = M->o[1] + M->d[1] * j;
        rc |= fwrite(&y, sizeof(float), 1, fo) == 1 ? 0 : -1;
    }
    for (int k = 0; k < np[2]; k++) {
        float z = orc_mesh_z(M, k + M->m / 2);
        rc |= fwrite(&z, sizeof(float), 1, fo) == 1 ? 0 : -1;
    }
    if (fclose(fo) != 0) rc = -1;

    snprintf(fn, sizeof(fn), "%s/fd.p%04d.case", path, procNo);
    fo = fopen(fn, "w");
    if (!fo) return -1;
    fprintf(fo, "FORMAT\n");
    fprintf(fo, "type: ensight gold\n");
    fprintf(fo, "GEOMETRY\n");
    fprintf(fo, "model: geom.p%04d.geo\n", procNo);
    fprintf(fo, "VARIABLE\n");
    fprintf(fo, "vector per node: velocity U.%04d.*****.dat\n", procNo);
    fprintf(fo, "vector per node: Uprime Uprime.%04d.*****.dat\n", procNo);
    fprintf(fo, "vector per node: Udash Udash.%04d.*****.dat\n", procNo);
    fprintf(fo, "scalar per node: p  p.%04d.*****.dat\n", procNo);
    fprintf(fo, "scalar per node: T  T.%04d.*****.dat\n", procNo);
    fprintf(fo, "vector per node: B B.%04d.*****.dat\n", procNo);
    fprintf(fo, "vector per node: J J.%04d.*****.dat\n", procNo);
    fprintf(fo, "TIME\n");
    fprintf(fo, "time set: 1\n");
    fprintf(fo, "number of steps: %d\n", nsteps);
    fprintf(fo, "filename start number: 00001\n");
    fprintf(fo, "filename increment: 1\n");
    fprintf(fo, "time values: ");
    for (int i = 0; i < nsteps; i++) fprintf(fo, "%g\n", steps[i]);
    fprintf(fo, "\n");
    if (fclose(fo) != 0) rc = -1;

    snprintf(fn, sizeof(fn), "%s/restart.p.%04d.case", path, procNo);
    fo = fopen(fn, "w");
    if (!fo) return -1;
    fprintf(fo, "%d\n", nsteps);
    for (int i = 0; i < nsteps; i++) fprintf(fo, "%g\n", steps[i]);
    if (fclose(fo) != 0) rc = -1;
    return rc;
}

/* ------------------------------------------------------------------ */
/* Diagnostics. */

/* app/TG_Piosson/TG.cpp:148-166: mean of 0.5|U|^2 over [h, n-1-h). */
double orc_ke(const orc_mesh *M, const double *U)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    double Ek = 0.0;
    long n = 0;
    for (int i = h; i < ni - h - 1; i++)
        for (int j = h; j < nj - h - 1; j++)
            for (int k = h; k < nk - h - 1; k++) {
                const double *u = &U[IX(M, i, j, k) * 3];
                Ek += 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                n++;
            }
    return Ek / n;
}

/* nu * mean over [h, n-1-h) of sum_{c,d} (D_d U_c)^2, D = the central
 * stencil of src/ex/grad/grad.cpp:33-58 per component (no reference row:
 * the viscous dissipation rate). */
double orc_dissipation(const orc_mesh *M, const double *U, double nu)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    double *g = falloc(M, 3), *s = falloc(M, 1);
    double acc = 0.0;
    long n = 0;
    for (int c = 0; c < 3; c++) {
        const size_t np = NPTS(M);
        for (size_t q = 0; q < np; q++) s[q] = U[3 * q + c];
        orc_grad(M, s, g);
        for (int i = h; i < ni - h - 1; i++)
            for (int j = h; j < nj - h - 1; j++)
                for (int k = h; k < nk - h - 1; k++) {
                    const double *v = &g[IX(M, i, j, k) * 3];
                    acc += ORC_PLANT == 9 ? v[0] * v[0] + v[1] * v[1] : v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                    if (c == 0) n++;
                }
    }
    free(g);
    free(s);
    return ORC_PLANT == 8 ? acc / n : nu * acc / n; /* 8: nu missing */
}

/* tools::CFL(U, mesh), src/Tools/Tools.cpp:9-41, as written: over
 * [h, n-h) in every direction, cfl_x = max(cfl_x, U_x dt / dx) from 0 and
 * cfl_y = max(cfl_x, U_y dt / dy), cfl_z = max(cfl_x, U_z dt / dz) at every
 * node (so they keep only the last node's terms).  out[3..5]: the maxima
 * of |U_d| dt / h_d over the same nodes. */
void orc_cfl(const orc_mesh *M, const double *U, double dt, double out[6])
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    double cx = 0.0, cy = 0.0, cz = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                const double *u = &U[IX(M, i, j, k) * 3];
                cx = fmax(cx, u[0] * dt / M->d[0]);
                cy = fmax(cx, u[1] * dt / M->d[1]);
                cz = fmax(cx, u[2] * dt / M->d[2]);
                ax = fmax(ax, fabs(u[0]) * dt / M->d[0]);
                ay = fmax(ay, fabs(u[1]) * dt / M->d[1]);
                az = fmax(az, fabs(u[2]) * dt / M->d[2]);
            }
    out[0] = cx; out[1] = cy; out[2] = cz; out[3] = ax; out[4] = ay; out[5] = az;
}

/* tools::maximum(ex::div(U)), src/Tools/Tools.h:19-36 over [h, n-h). */
double orc_maxabs_div(const orc_mesh *M, const double *U)
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    double *d = falloc(M, 1);
    orc_div(M, U, d);
    double mx = 0.0;
    for (int i = h; i < ni - h; i++)
        for (int j = h; j < nj - h; j++)
            for (int k = h; k < nk - h; k++) {
                double a = fabs(d[IX(M, i, j, k)]);
                mx = a > mx ? a : mx;
            }
    free(d);
    return mx;
}

/* app/OT_Poisson/OT.cpp:134-176: out = {Ek, Em, Jav, Jmax, omega,
 * epsilon, Umax, 0} with J = curl(B), omega = curl(U) (literal curl). */
void orc_mhd_diag(const orc_mesh *M, const double *U, const double *B, double mu, double eta,
                  double out[8])
{
    const int ni = M->n[0], nj = M->n[1], nk = M->n[2], h = M->h;
    double *J = falloc(M, 3), *W = falloc(M, 3);
    if (ORC_PLANT == 12) orc_curl_central(M, B, J); /* 12: J from the consistent curl */
    else orc_curl_literal(M, B, J);
    orc_curl_literal(M, U, W);
    double Ek = 0, Em = 0, Jmax = 0, Umax = 0, Jav = 0, om = 0;
    long n = 0;
    for (int i = h; i < ni - h - 1; i++)
        for (int j = h; j < nj - h - 1; j++)
            for (int k = h; k < nk - h - 1; k++) {
                const size_t q = IX(M, i, j, k) * 3;
                const double *u = &U[q], *b = &B[q], *jj = &J[q], *w = &W[q];
                Ek += 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                Em += (ORC_PLANT == 13 ? 1.0 : 0.5) * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
                double jm = sqrt(jj[0] * jj[0] + jj[1] * jj[1] + jj[2] * jj[2]);
                double um = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                Jmax = Jmax > jm ? Jmax : jm;
                Umax = Umax > um ? Umax : um;
                om += sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                Jav += jm;
                n++;
            }
    Ek /= n; Em /= n; Jav /= n; om /= n;
    out[0] = Ek; out[1] = Em; out[2] = Jav; out[3] = Jmax; out[4] = om;
    out[5] = ORC_PLANT == 14 ? -mu * (om * om) + eta * (Jav * Jav) : -mu * (om * om) - eta * (Jav * Jav);
    out[6] = Umax; out[7] = 0.0;
    free(J); free(W);
}

/* tools::tdma, src/Tools/Tools.cpp:43-62 (Thomas, no pivoting). */
void orc_tdma(int n, const double *a, const double *b, const double *c, const double *d, double *x)
{
    double *cp = (double *)malloc(sizeof(double) * (size_t)n);
    double *dp = (double *)malloc(sizeof(double) * (size_t)n);
    cp[0] = c[0] / b[0];
    dp[0] = d[0] / b[0];
    for (int i = 1; i < n; i++) {
        cp[i] = c[i] / (b[i] - a[i] * cp[i - 1]);
        dp[i] = (d[i] - a[i] * dp[i - 1]) / (b[i] - a[i] * cp[i - 1]);
    }
    x[n - 1] = dp[n - 1];
    for (int i = n - 2; i >= 0; i--) x[i] = dp[i] - cp[i] * x[i + 1];
    free(cp);
    free(dp);
}

/* ------------------------------------------------------------------ */
/* Compact 6th-order scheme (no reference counterpart).  Lele (1992),
 * J. Comput. Phys. 103:16-42: first derivative (2.1.7) with alpha = 1/3,
 * a = 14/9, b = 1/9; second derivative (2.2.7) with alpha = 2/11,
 * a = 12/11, b = 3/11; periodic lines through the N unique nodes. */
static const double CMP_AL[2] = {1.0 / 3.0, 2.0 / 11.0};
static const double CMP_A[2] = {14.0 / 9.0, 12.0 / 11.0};
static const double CMP_B[2] = {1.0 / 9.0, 3.0 / 11.0};

/* Sherman-Morrison reduction of the cyclic system to two Thomas solves
 * (Press et al., Numerical Recipes, 2nd ed., sec. 2.7, "cyclic"). */
void orc_cyclic_tridiag(int n, double a, double b, double c, const double *d, double *x)
{
    const double alpha = c, beta = a;  /* corners (n-1, 0) and (0, n-1) */
    const double gamma = -b;
    double *bb = (double *)malloc(sizeof(double) * n);
    double *u = (double *)calloc(n, sizeof(double));
    double *z = (double *)malloc(sizeof(double) * n);
    double *cp = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) bb[i] = b;
    bb[0] = b - gamma;
    bb[n - 1] = b - alpha * beta / gamma;
    /* Thomas on (a, bb, c) for two right-hand sides */
    const double *rhs[2] = {d, u};
    double *sol[2] = {x, z};
    u[0] = gamma;
    u[n - 1] = alpha;
    for (int r = 0; r < 2; r++) {
        const double *dd = rhs[r];
        double *xx = sol[r];
        cp[0] = c / bb[0];
        xx[0] = dd[0] / bb[0];
        for (int i = 1; i < n; i++) {
            const double den = bb[i] - a * cp[i - 1];
            cp[i] = c / den;
            xx[i] = (dd[i] - a * xx[i - 1]) / den;
        }
        for (int i = n - 2; i >= 0; i--) xx[i] = xx[i] - cp[i] * xx[i + 1];
    }
    const double fact = (x[0] + beta * x[n - 1] / gamma) / (1.0 + z[0] + beta * z[n - 1] / gamma);
    for (int i = 0; i < n; i++) x[i] = x[i] - fact * z[i];
    free(bb); free(u); free(z); free(cp);
}

double orc_compact_kmod(double kh, int order)
{
    const int o = order - 1;
    if (order == 1)
        return (CMP_A[o] * sin(kh) + 0.5 * CMP_B[o] * sin(2.0 * kh)) / (1.0 + 2.0 * CMP_AL[o] * cos(kh));
    return (2.0 * CMP_A[o] * (1.0 - cos(kh)) + 0.5 * CMP_B[o] * (1.0 - cos(2.0 * kh))) /
           (1.0 + 2.0 * CMP_AL[o] * cos(kh));
}

/* one periodic line of n values -> its compact derivative */
static void compact_line(int n, double h, int order, const double *f, double *out, double *rhs)
{
    const int o = order - 1;
    for (int i = 0; i < n; i++) {
        const double fm2 = f[(i - 2 + 2 * n) % n], fm1 = f[(i - 1 + n) % n];
        const double fp1 = f[(i + 1) % n], fp2 = f[(i + 2) % n];
        if (order == 1)
            rhs[i] = CMP_A[o] / (2.0 * h) * (fp1 - fm1) + CMP_B[o] / (4.0 * h) * (fp2 - fm2);
        else
            rhs[i] = CMP_A[o] / (h * h) * (fp1 - 2.0 * f[i] + fm1) +
                     CMP_B[o] / (4.0 * h * h) * (fp2 - 2.0 * f[i] + fm2);
    }
    orc_cyclic_tridiag(n, CMP_AL[o], 1.0, CMP_AL[o], rhs, out);
}

void orc_compact_deriv(const orc_mesh *M, const double *f, int nc, int dir, int order, double *out)
{
    const int h = M->h;
    const int N[3] = {M->N[0], M->N[1], M->N[2]};
    const int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
    const double hd = M->d[dir];
    #pragma omp parallel
    {
        const int n = N[dir];
        double *line = (double *)malloc(sizeof(double) * n);
        double *res = (double *)malloc(sizeof(double) * n);
        double *rhs = (double *)malloc(sizeof(double) * n);
        #pragma omp for collapse(2)
        for (int p = 0; p < N[a1]; p++)
            for (int q = 0; q < N[a2]; q++)
                for (int c = 0; c < nc; c++) {
                    int ix[3];
                    ix[a1] = p + h;
                    ix[a2] = q + h;
                    for (int t = 0; t < n; t++) {
                        ix[dir] = t + h;
                        line[t] = f[IX(M, ix[0], ix[1], ix[2]) * nc + c];
                    }
                    compact_line(n, hd, order, line, res, rhs);
                    for (int t = 0; t < n; t++) {
                        ix[dir] = t + h;
                        out[IX(M, ix[0], ix[1], ix[2]) * nc + c] = res[t];
                    }
                }
        free(line); free(res); free(rhs);
    }
    orc_fill(M, out, nc, nc == 3 ? 1 : -1);
}

orc_poisson *orc_poisson_create_compact(const orc_mesh *M)
{
    orc_poisson *P = orc_poisson_create(M);
    const int M0 = P->M[0], M1 = P->M[1], M2 = P->M[2];
    const double pi = ORC_PI;
    size_t l = 0;
    for (int k = 0; k < M2; k++)
        for (int j = 0; j < M1; j++)
            for (int i = 0; i < M0; i++) {
                const double kx = (2 * i < M0) ? 2.0 * pi * i / M->L[0] : 2.0 * pi * (M0 - i) / M->L[0];
                const double ky = (2 * j < M1) ? 2.0 * pi * j / M->L[1] : 2.0 * pi * (M1 - j) / M->L[1];
                const double kz = 2.0 * pi * k / M->L[2];
                const double kpx = orc_compact_kmod(kx * M->d[0], 1) / M->d[0];
                const double kpy = orc_compact_kmod(ky * M->d[1], 1) / M->d[1];
                const double kpz = orc_compact_kmod(kz * M->d[2], 1) / M->d[2];
                P->k2[l++] = kpx * kpx + kpy * kpy + kpz * kpz;
            }
    return P;
}

static void compact_div(const orc_mesh *M, const double *F, double *out)
{
    const size_t n1 = NPTS(M);
    double *comp = falloc(M, 1), *dd = falloc(M, 1);
    memset(out, 0, sizeof(double) * n1);
    for (int d = 0; d < 3; d++) {
        for (size_t q = 0; q < n1; q++) comp[q] = F[3 * q + d];
        orc_fill(M, comp, 1, -1);
        orc_compact_deriv(M, comp, 1, d, 1, dd);
        for (size_t q = 0; q < n1; q++) out[q] += dd[q];
    }
    free(comp); free(dd);
}

void orc_compact_ns_rhs(const orc_mesh *M, const double *U, double nu, double *out)
{
    const size_t n1 = NPTS(M), n3 = n1 * 3;
    double *flux = falloc(M, 3), *dflux = falloc(M, 3), *dU = falloc(M, 3), *d2U = falloc(M, 3);
    double *adv = falloc(M, 3), *lap = falloc(M, 3);
    for (int d = 0; d < 3; d++) {
        for (size_t q = 0; q < n1; q++)
            for (int c = 0; c < 3; c++) flux[3 * q + c] = U[3 * q + d] * U[3 * q + c];
        orc_fill(M, flux, 3, -1);
        orc_compact_deriv(M, flux, 3, d, 1, dflux);
        orc_compact_deriv(M, U, 3, d, 1, dU);
        orc_compact_deriv(M, U, 3, d, 2, d2U);
        for (size_t q = 0; q < n1; q++)
            for (int c = 0; c < 3; c++) {
                adv[3 * q + c] += dflux[3 * q + c] + U[3 * q + d] * dU[3 * q + c] + U[3 * q + c] * dU[3 * q + d];
                lap[3 * q + c] += d2U[3 * q + c];
            }
    }
    for (size_t q = 0; q < n3; q++) out[q] = -0.5 * adv[q] + nu * lap[q];
    orc_fill(M, out, 3, 1);
    free(flux); free(dflux); free(dU); free(d2U); free(adv); free(lap);
}

void orc_compact_project(const orc_mesh *M, orc_poisson *P, double *F, double *p, double s)
{
    const size_t n1 = NPTS(M);
    double *d = falloc(M, 1), *g = falloc(M, 1);
    orc_fill(M, F, 3, 1);
    compact_div(M, F, d);
    for (size_t q = 0; q < n1; q++) d[q] = d[q] / s;
    orc_poisson_solve(P, d, p);
    orc_fill(M, p, 1, -1);
    for (int c = 0; c < 3; c++) {
        orc_compact_deriv(M, p, 1, c, 1, g);
        for (size_t q = 0; q < n1; q++) F[3 * q + c] = F[3 * q + c] - s * g[q];
    }
    orc_fill(M, F, 3, 1);
    free(d); free(g);
}

void orc_compact_ns_step(const orc_mesh *M, orc_poisson *P, double *U, double *Ustar, double *p,
                         double nu, double dt)
{
    const size_t n3 = NPTS(M) * 3;
    double *U0 = falloc(M, 3), *R = falloc(M, 3);
    memcpy(U0, U, sizeof(double) * n3);
    memcpy(Ustar, U0, sizeof(double) * n3);
    for (int s = 0; s < 3; s++) {
        orc_compact_ns_rhs(M, Ustar, nu, R);
        for (size_t q = 0; q < n3; q++) {
            const double v = Ustar[q] + dt * R[q];
            if (s == 0) Ustar[q] = v;
            else if (s == 1) Ustar[q] = 0.75 * U0[q] + 0.25 * v;
            else Ustar[q] = (1.0 / 3.0) * U0[q] + (2.0 / 3.0) * v;
        }
        orc_fill(M, Ustar, 3, 1);
    }
    orc_compact_project(M, P, Ustar, p, dt);
    memcpy(U, Ustar, sizeof(double) * n3);
    free(U0); free(R);
}

double orc_compact_maxabs_div(const orc_mesh *M, const double *U)
{
    double *d = falloc(M, 1);
    compact_div(M, U, d);
    double mx = 0.0;
    const int h = M->h;
    for (int i = h; i < h + M->N[0]; i++)
        for (int j = h; j < h + M->N[1]; j++)
            for (int k = h; k < h + M->N[2]; k++) {
                const double v = fabs(d[IX(M, i, j, k)]);
                if (v > mx) mx = v;
            }
    free(d);
    return mx;
}

/* out_c += s * 1/2 sum_d [D_d(X_d Y_c) + X_d D_d Y_c + Y_c D_d X_d] */
static void compact_bform(const orc_mesh *M, const double *X, const double *Y, double s, double *out)
{
    const size_t n1 = NPTS(M);
    double *flux = falloc(M, 3), *dflux = falloc(M, 3), *dY = falloc(M, 3), *dX = falloc(M, 3);
    double *acc = falloc(M, 3);
    for (int d = 0; d < 3; d++) {
        for (size_t q = 0; q < n1; q++)
            for (int c = 0; c < 3; c++) flux[3 * q + c] = X[3 * q + d] * Y[3 * q + c];
        orc_fill(M, flux, 3, -1);
        orc_compact_deriv(M, flux, 3, d, 1, dflux);
        orc_compact_deriv(M, Y, 3, d, 1, dY);
        orc_compact_deriv(M, X, 3, d, 1, dX);
        for (size_t q = 0; q < n1; q++)
            for (int c = 0; c < 3; c++)
                acc[3 * q + c] += dflux[3 * q + c] + X[3 * q + d] * dY[3 * q + c] + Y[3 * q + c] * dX[3 * q + d];
    }
    for (size_t q = 0; q < 3 * n1; q++) out[q] += s * 0.5 * acc[q];
    free(flux); free(dflux); free(dY); free(dX); free(acc);
}

void orc_compact_mhd_rhs(const orc_mesh *M, const double *U, const double *B, double nu, double eta,
                         double *dU, double *dB)
{
    const size_t n1 = NPTS(M), n3 = 3 * n1;
    memset(dU, 0, sizeof(double) * n3);
    memset(dB, 0, sizeof(double) * n3);
    compact_bform(M, U, U, -1.0, dU);
    compact_bform(M, B, B, 1.0, dU);
    double *pm = falloc(M, 1), *g = falloc(M, 1), *l = falloc(M, 3);
    for (size_t q = 0; q < n1; q++)
        pm[q] = 0.5 * (B[3 * q] * B[3 * q] + B[3 * q + 1] * B[3 * q + 1] + B[3 * q + 2] * B[3 * q + 2]);
    orc_fill(M, pm, 1, -1);
    for (int c = 0; c < 3; c++) {
        orc_compact_deriv(M, pm, 1, c, 1, g);
        for (size_t q = 0; q < n1; q++) dU[3 * q + c] -= g[q];
    }
    for (int d = 0; d < 3; d++) {
        orc_compact_deriv(M, U, 3, d, 2, l);
        for (size_t q = 0; q < n3; q++) dU[q] += nu * l[q];
        orc_compact_deriv(M, B, 3, d, 2, l);
        for (size_t q = 0; q < n3; q++) dB[q] += eta * l[q];
    }
    compact_bform(M, U, B, -1.0, dB);
    compact_bform(M, B, U, 1.0, dB);
    orc_fill(M, dU, 3, 1);
    orc_fill(M, dB, 3, 1);
    free(pm); free(g); free(l);
}

void orc_compact_mhd_step(const orc_mesh *M, orc_poisson *P, double *U, double *B, double *p, double *pB,
                          double nu, double eta, double dt)
{
    const size_t n3 = NPTS(M) * 3;
    double *U0 = falloc(M, 3), *B0 = falloc(M, 3), *dU = falloc(M, 3), *dB = falloc(M, 3);
    memcpy(U0, U, sizeof(double) * n3);
    memcpy(B0, B, sizeof(double) * n3);
    for (int s = 0; s < 3; s++) {
        orc_compact_mhd_rhs(M, U, B, nu, eta, dU, dB);
        for (size_t q = 0; q < n3; q++) {
            const double vu = U[q] + dt * dU[q], vb = B[q] + dt * dB[q];
            if (s == 0) { U[q] = vu; B[q] = vb; }
            else if (s == 1) { U[q] = 0.75 * U0[q] + 0.25 * vu; B[q] = 0.75 * B0[q] + 0.25 * vb; }
            else { U[q] = (1.0 / 3.0) * U0[q] + (2.0 / 3.0) * vu; B[q] = (1.0 / 3.0) * B0[q] + (2.0 / 3.0) * vb; }
        }
        orc_compact_project(M, P, U, p, dt);
        orc_compact_project(M, P, B, pB, 1.0);
    }
    free(U0); free(B0); free(dU); free(dB);
}
