"""Pins the oracle's implicit matrix im::fdMatrix<scalar> and
im::laplacian (src/im/fdMatrix/fdMatrix.hxx, src/im/laplacian/laplacian.hxx).
The reference has no test of them; they are pinned by what the formulas
fix: the coefficients in closed form (zero off the interior), the deferred
correction's right-hand side minus the low-order operator equal to minus the
explicit Laplacian (itself pinned by its Fourier eigenvalues), the residual
against an independent stencil evaluation, and the solver's converged field
against a dense solve of the same linear system (the null space of the
periodic Laplacian removed by setReference's 1e15 weighting)."""
import numpy as np
import pytest


def interior(M, a):
    h = M.h
    return a[h:M.n[0] - h, h:M.n[1] - h, h:M.n[2] - h]


def stencil(M, A, phi):
    """rhs - sum al phi_nb - ap phi on the interior, evaluated with numpy."""
    h = M.h
    p = phi[..., 0] if phi.ndim == 4 else phi
    s = (slice(h, M.n[0] - h), slice(h, M.n[1] - h), slice(h, M.n[2] - h))

    def sh(di, dj, dk):
        return p[h + di:M.n[0] - h + di, h + dj:M.n[1] - h + dj, h + dk:M.n[2] - h + dk]
    al = A["al"]
    return (A["rhs"][s] - al[0][s] * sh(0, 1, 0) - al[1][s] * sh(0, -1, 0) - al[2][s] * sh(1, 0, 0)
            - al[3][s] * sh(-1, 0, 0) - al[4][s] * sh(0, 0, 1) - al[5][s] * sh(0, 0, -1) - A["ap"][s] * p[s])


@pytest.mark.parametrize("wall", [False, True])
def test_matrix_coefficients_and_deferred_correction(orc, wall):
    N, L = (8, 6, 10), (1.0, 2.0, 1.5)
    M = orc.OMesh(N, L, wall_y=wall)
    rng = np.random.default_rng(1)
    phi = M.from_compact(rng.standard_normal((1, *M.stored)))
    mu = 0.7
    A = M.fd_laplacian(mu, phi)
    d = [L[a] / N[a] for a in range(3)]
    want = [mu / d[1] ** 2, mu / d[1] ** 2, mu / d[0] ** 2, mu / d[0] ** 2, mu / d[2] ** 2, mu / d[2] ** 2]
    for q in range(6):
        assert np.all(interior(M, A["al"][q]) == want[q])
        assert np.count_nonzero(A["al"][q]) == interior(M, A["al"][q]).size
    assert np.allclose(interior(M, A["ap"]), -2 * mu * sum(1 / x ** 2 for x in d), rtol=1e-15)
    # rhs - A_low phi = -laplacian(mu, phi) on the interior
    lap = M.laplacian(mu, phi)[..., 0]
    r = stencil(M, A, phi)
    assert np.abs(r + interior(M, lap)).max() < 1e-12 * np.abs(interior(M, lap)).max()
    assert abs(M.fd_residual(A, phi) - np.sqrt(np.mean(r ** 2))) < 1e-13 * np.sqrt(np.mean(r ** 2))


def periodic_dense(M, A):
    """The system on the stored nodes of a periodic grid, neighbours
    wrapped: rows = interior nodes 0..N-1 in each direction."""
    N = M.N
    h = M.h
    n = N[0] * N[1] * N[2]
    K = np.zeros((n, n))
    b = np.zeros(n)

    def idx(i, j, k):
        return ((i % N[0]) * N[1] + (j % N[1])) * N[2] + (k % N[2])
    for i in range(N[0]):
        for j in range(N[1]):
            for k in range(N[2]):
                r = idx(i, j, k)
                P = (i + h, j + h, k + h)
                K[r, r] += A["ap"][P]
                for q, (di, dj, dk) in enumerate([(0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0), (0, 0, 1),
                                                  (0, 0, -1)]):
                    K[r, idx(i + di, j + dj, k + dk)] += A["al"][q][P]
                b[r] = A["rhs"][P]
    return K, b


def test_solve_converges_to_the_dense_solution(orc):
    N = (6, 5, 7)
    M = orc.OMesh(N, (1.0, 1.0, 1.0))
    rng = np.random.default_rng(2)
    phi0 = M.from_compact(rng.standard_normal((1, *M.stored)))
    A = M.fd_laplacian(1.0, phi0)
    src = M.from_compact(rng.standard_normal((1, *M.stored)))
    A["rhs"] += src[..., 0]                       # matrix == source
    h = M.h
    A["rhs"][h + 2, h + 1, h + 3] += 1e15 * 0.25  # setReference(., 0.25)
    A["ap"][h + 2, h + 1, h + 3] *= 1e15
    K, b = periodic_dense(M, A)
    x = np.linalg.solve(K, b).reshape(N)
    phi = phi0.copy()
    its, res = M.fd_solve(A, phi, 1e-10, 2000)
    assert 1 < its < 2000 and res <= 1e-10
    got = phi[h:h + N[0], h:h + N[1], h:h + N[2], 0]
    assert np.abs(got - x).max() < 1e-7 * np.abs(x).max()
    # the image node repeats node 0 after the refill
    assert np.array_equal(phi[h + N[0], h:h + N[1], h:h + N[2]], phi[h, h:h + N[1], h:h + N[2]])
    # the residual the solver reports is the stencil's RMS
    r = stencil(M, A, phi)
    assert M.fd_residual(A, phi) == pytest.approx(np.sqrt(np.mean(r ** 2)), rel=1e-12)


def test_iterations_stop_at_the_tolerance_and_at_max_its(orc):
    M = orc.OMesh((6, 6, 6), (1.0, 1.0, 1.0))
    rng = np.random.default_rng(3)
    phi = M.from_compact(rng.standard_normal((1, *M.stored)))
    A = M.fd_laplacian(1.0, phi)
    p1 = phi.copy()
    its, res = M.fd_solve(A, p1, 0.0, 3)
    assert its == 3 and res > 0
    p2 = phi.copy()
    its0, res0 = M.fd_solve(A, p2, 1e300, 3)   # converged before any sweep
    assert its0 == 0 and np.array_equal(p2, phi)
    assert res0 == pytest.approx(M.fd_residual(A, phi), rel=0, abs=0)


def test_relax_scales_the_centre_and_compensates_the_rhs(orc):
    M = orc.OMesh((6, 6, 6), (1.0, 1.0, 1.0))
    rng = np.random.default_rng(4)
    phi = M.from_compact(rng.standard_normal((1, *M.stored)))
    A = M.fd_laplacian(1.0, phi)
    ap0, rhs0 = A["ap"].copy(), A["rhs"].copy()
    M.fd_relax(A, phi, 0.7)
    assert np.allclose(A["ap"], ap0 / 0.7, rtol=1e-15)
    # at phi the relaxed equation has the same residual: (ap/urf) phi vs
    # rhs + (1-urf)/urf ap phi
    r0 = stencil(M, {"al": A["al"], "ap": ap0, "rhs": rhs0}, phi)
    r1 = stencil(M, A, phi)
    assert np.abs(r1 - r0).max() < 1e-11 * np.abs(r0).max()
