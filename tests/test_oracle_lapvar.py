"""Pins the oracle's variable-coefficient Laplacian
ex::laplacian(Field<scalar> nu, Field<T>) (src/ex/laplacian/laplacian.hxx:64-127).
The reference has no test of it; it is pinned by what the formula fixes: a
uniform nu gives nu times the scalar-coefficient Laplacian (itself pinned by
its Fourier eigenvalues), a constant field is annihilated for any nu, and a
unit impulse of nu at one node has the closed-form response
w_l (f(q+1) - f(q-l)) at q - l and w_l (f(q) - f(q+1+l)) at q + 1 + l along
each axis (w_l = c_l / ((l+1) h)^2), zero elsewhere, which fixes which node's
nu multiplies which neighbour."""
import ctypes as C
import math

import numpy as np
import pytest

TWO_PI = 2.0 * math.pi


def inner(M):
    h = M.h
    return tuple(slice(h, h + s) for s in M.stored)


@pytest.mark.parametrize("m", [2, 4, 6, 8])
@pytest.mark.parametrize("wall", [False, True])
def test_uniform_nu_is_nu_times_the_laplacian(orc, m, wall):
    N = (12, 10, 14)
    M = orc.OMesh(N, (TWO_PI, 3.0, 2.0), m=m, wall_y=wall)
    rng = np.random.default_rng(m)
    U = rng.standard_normal((3, *M.stored))
    if wall:
        U[1, :, 0, :] = U[1, :, -1, :] = 0.0
    f = M.from_compact(U)
    nu = M.from_compact(np.full((1, *M.stored), 0.37))
    a = M.laplacian_var(nu, f)[inner(M)]
    b = M.laplacian(0.37, f)[inner(M)]
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_constant_field_is_annihilated(orc):
    M = orc.OMesh((10, 12, 8), (1.0, 2.0, 3.0), hyp=(0.02, 0.03))
    rng = np.random.default_rng(1)
    nu = M.from_compact(1.0 + rng.uniform(0, 1, (1, *M.stored)))
    f = M.field(3)
    f[...] = (2.0, -1.0, 0.5)
    out = M.laplacian_var(nu, f)[inner(M)]
    scale = max(1.0 / d ** 2 for d in (0.1, 1 / 6, 3 / 8)) * 4.0
    assert np.abs(out).max() < 1e-13 * scale


@pytest.mark.parametrize("m", [4, 6])
def test_impulse_of_nu_has_the_closed_form_response(orc, m):
    N = (16, 12, 14)
    L = (2.0, 1.5, 1.75)
    M = orc.OMesh(N, L, m=m)
    h = M.h
    buf = (C.c_double * 4)()
    orc.lib().orc_coef(m, buf)
    coef = list(buf)
    rng = np.random.default_rng(7)
    fc = rng.standard_normal((1, *M.stored))
    f = M.from_compact(fc)
    q = (7, 5, 6)
    nuc = np.zeros((1, *M.stored))
    nuc[0, q[0], q[1], q[2]] = 1.0
    nu = M.from_compact(nuc)
    out = M.laplacian_var(nu, f)[inner(M)][..., 0]
    want = np.zeros(M.stored)
    F = fc[0]
    for d in range(3):
        hd = L[d] / N[d]
        e = np.zeros(3, dtype=int)
        e[d] = 1
        for l in range(h):
            w = coef[l] / ((l + 1) * hd) ** 2
            qq = np.array(q)
            a = tuple(qq - l * e)            # node q - l: nu(i + l) = nu(q) with f(q + 1)
            want[a] += w * (F[tuple(qq + e)] - F[a])
            b = tuple(qq + (1 + l) * e)      # node q + 1 + l: nu(i - 1 - l) = nu(q) with f(q)
            want[b] += w * (F[tuple(qq)] - F[b])
    # the z metric of a uniform grid is 1 and 0 to rounding
    assert np.abs(out - want).max() < 1e-10 * np.abs(want).max()
    # node q itself answers on all three axes (l = 0)
    assert np.count_nonzero(np.abs(out) > 1e-9 * np.abs(want).max()) == 6 * h - 2
