"""Pointwise Field algebra (src/Field/Field.h:227-363 over the node types of
src/Types), the literal vector gradient (src/ex/grad/grad.cpp:65-102) and
tools::CFL (src/Tools/Tools.cpp:9-41) on the GPU.  The products are pure
IEEE arithmetic in the reference's operand order without fused
multiply-adds, so they must equal the same numpy expressions bit for bit;
the gradient and the CFL numbers are checked against the oracle."""
import math

import numpy as np
import pytest

from test_gpu_parity import dev, host, pair, rel_l2, wall_clean

pytestmark = pytest.mark.gpu
TWO_PI = 2.0 * math.pi


@pytest.fixture(scope="module")
def nv():
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available()
    nv.lib()
    return nv


def test_pointwise_products_bitwise(nv):
    G = nv.Mesh((12, 16, 8), (1.0, 2.0, 3.0))
    rng = np.random.default_rng(8)
    a, b = rng.standard_normal((3, *G.shape)), rng.standard_normal((3, *G.shape))
    s = 0.5 + rng.random((1, *G.shape))
    A9, B9 = rng.standard_normal((9, *G.shape)), rng.standard_normal((9, *G.shape))
    da, db, ds = dev(a), dev(b), dev(s)
    got = host(nv.field_op(G, nv.OP_DOT, da, db))[0]
    assert np.array_equal(got, a[0] * b[0] + a[1] * b[1] + a[2] * b[2])
    got = host(nv.field_op(G, nv.OP_CROSS, da, db))
    exp = np.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])
    assert np.array_equal(got, exp)
    assert np.array_equal(host(nv.field_op(G, nv.OP_SCALE, ds, da)), s * a)
    assert np.array_equal(host(nv.field_op(G, nv.OP_SCALE, ds, ds)), s * s)
    assert np.array_equal(host(nv.field_op(G, nv.OP_DIVIDE, da, ds)), a / s)
    assert np.array_equal(host(nv.field_op(G, nv.OP_SQRT, ds)), np.sqrt(s))
    got = host(nv.field_op(G, nv.OP_OUTER, da, db))
    assert np.array_equal(got, np.stack([a[i] * b[j] for i in range(3) for j in range(3)]))
    r = np.zeros(G.shape)
    for i in range(9):  # tensor.cpp:237-251: every pair (i, j)
        for j in range(9):
            r = r + A9[i] * B9[j]
    assert np.array_equal(host(nv.field_op(G, nv.OP_DDOT, dev(A9), dev(B9)))[0], r)
    got = host(nv.field_op(G, nv.OP_ADD_CONST, da, None, (-1.5, 2.0, 0.25)))
    assert np.array_equal(got, a + np.array([-1.5, 2.0, 0.25])[:, None, None, None])
    with pytest.raises(nv.NovaError):
        nv.field_op(G, nv.OP_CROSS, ds, da)


@pytest.mark.parametrize("wall_y", [False, True])
def test_literal_vector_gradient_matches_oracle(nv, orc, wall_y):
    G, O = pair(nv, orc, (12, 16, 16), (TWO_PI, 3.0 if wall_y else TWO_PI, 2.0), wall_y=wall_y)
    c = wall_clean(O, np.random.default_rng(2).standard_normal((3, *O.stored)))
    Uo = O.from_compact(c)
    got = host(nv.grad_vec_literal(G, dev(c)))
    ref = O.interior(O.grad_vec_literal(Uo))
    assert rel_l2(got, ref) <= 1e-13


@pytest.mark.parametrize("wall_y", [False, True])
def test_cfl_matches_oracle(nv, orc, wall_y):
    G, O = pair(nv, orc, (16, 12, 8), (1.0, 2.0 if wall_y else 1.5, 0.5), wall_y=wall_y)
    rng = np.random.default_rng(6)
    c = wall_clean(O, rng.standard_normal((3, *O.stored)))
    c[1] *= 4.0
    got = nv.cfl(G, dev(c), 0.01)
    ref = O.cfl(O.from_compact(c), 0.01)
    assert np.array_equal(np.array(got), ref), (got, ref)
