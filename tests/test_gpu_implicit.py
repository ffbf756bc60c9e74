"""The implicit matrix im::fdMatrix<scalar> / im::laplacian
(src/im/fdMatrix/fdMatrix.hxx, src/im/laplacian/laplacian.hxx) through the
C ABI against the oracle: coefficients bit for bit, the deferred-correction
right-hand side and the residual to rounding, and the line Gauss-Seidel
solve after 1, 2 and 3 iterations and at convergence (same iteration count)
to 1e-12 relative, which only the reference's line order gives; relax,
operator==, setReference and the matrix algebra."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TWO_PI = 2.0 * math.pi


@pytest.fixture(scope="module")
def nv():
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available(), "GPU tests need a device"
    nv.lib()
    return nv


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


CASES = [
    dict(N=(10, 8, 12), L=(1.0, 2.0, 1.5), m=6, wall=False),
    dict(N=(12, 10, 8), L=(TWO_PI, 3.0, 2.0), m=4, wall=True),
    dict(N=(9, 11, 10), L=(1.0, 1.0, 1.0), m=8, wall=False),
]


def pair(nv, orc, case, seed):
    bc = ("periodic", "wall" if case["wall"] else "periodic", "periodic")
    G = nv.Mesh(case["N"], case["L"], m=case["m"], bc=bc)
    O = orc.OMesh(case["N"], case["L"], m=case["m"], wall_y=case["wall"])
    phi = np.random.default_rng(seed).standard_normal((1, *O.stored))
    return G, O, phi


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['N']}-m{c['m']}-wall{int(c['wall'])}")
def test_matrix_and_solve_match_oracle(nv, orc, case):
    print("case", case, flush=True)
    G, O, phi = pair(nv, orc, case, 1)
    mu = 0.6
    pd = dev(phi)
    A = nv.FdMatrix(G, mu, pd)
    po = O.from_compact(phi)
    Ao = O.fd_laplacian(mu, po)
    for q in range(6):
        assert np.array_equal(host(A.padded(q)), Ao["al"][q]), q
    assert np.array_equal(host(A.padded(6)), Ao["ap"])
    assert rel(host(A.padded(7)), Ao["rhs"]) < 1e-13
    r_g, r_o = A.residual(pd), O.fd_residual(Ao, po)
    assert abs(r_g - r_o) <= 1e-12 * r_o
    # 1, 2, 3 sweeps from the same start, then to convergence
    for its in (1, 2, 3):
        g = dev(phi)
        o = O.from_compact(phi)
        ig, _ = A.solve(g, 0.0, its)
        io, _ = O.fd_solve(Ao, o, 0.0, its)
        e = rel(host(g), O.interior(o))
        print("its", its, "rel", e, flush=True)
        assert ig == io == its and e < 1e-12
    g = dev(phi)
    o = O.from_compact(phi)
    ig, rg = A.solve(g, 1e-8, 500)
    io, ro = O.fd_solve(Ao, o, 1e-8, 500)
    print("converged in", ig, io, rg, ro, flush=True)
    assert ig == io and 1 < ig < 500 and rg <= 1e-8
    assert rel(host(g), O.interior(o)) < 1e-12
    A.close()


def test_relax_source_reference_and_algebra_match_oracle(nv, orc):
    case = CASES[0]
    G, O, phi = pair(nv, orc, case, 2)
    src = np.random.default_rng(3).standard_normal((1, *O.stored))
    pd, sd = dev(phi), dev(src)
    po, so = O.from_compact(phi), O.from_compact(src)
    A = nv.FdMatrix(G, 1.0, pd)
    Ao = O.fd_laplacian(1.0, po)
    A.add_source(sd)
    Ao["rhs"] += so[..., 0]
    A.relax(pd, 0.8)
    O.fd_relax(Ao, po, 0.8)
    h = O.h
    A.set_reference(h + 1, h + 2, h + 3, 0.5)
    Ao["rhs"][h + 1, h + 2, h + 3] += 1e15 * 0.5
    Ao["ap"][h + 1, h + 2, h + 3] *= 1e15
    assert np.array_equal(host(A.padded(6)), Ao["ap"])
    assert rel(host(A.padded(7)), Ao["rhs"]) < 1e-13
    g, o = dev(phi), O.from_compact(phi)
    ig, _ = A.solve(g, 1e-7, 400)
    io, _ = O.fd_solve(Ao, o, 1e-7, 400)
    assert ig == io and rel(host(g), O.interior(o)) < 1e-11
    # operator+=, -= and unary minus are elementwise on al, ap, rhs
    B = nv.FdMatrix(G, 0.3, sd)
    Bo = O.fd_laplacian(0.3, so)
    ref = [host(A.padded(q)) for q in range(8)]
    A += B
    for q in range(8):
        want = (Ao["al"][q] + Bo["al"][q]) if q < 6 else ((Ao["ap"] + Bo["ap"]) if q == 6 else None)
        if want is not None:
            assert np.array_equal(host(A.padded(q)), want), q
    A -= B
    for q in range(7):
        assert np.array_equal(host(A.padded(q)), (ref[q] + host(B.padded(q))) - host(B.padded(q))), q
    A.negate()
    assert np.array_equal(host(A.padded(6)), -((ref[6] + host(B.padded(6))) - host(B.padded(6))))
    with pytest.raises(nv.NovaError):
        A.set_reference(-1, 0, 0, 1.0)
    H = nv.Mesh(case["N"], case["L"], m=case["m"])
    C2 = nv.FdMatrix(H, 1.0, H.field(1))
    with pytest.raises(nv.NovaError):  # another context
        A += C2
