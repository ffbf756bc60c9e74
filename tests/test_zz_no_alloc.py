"""nova_tdma_batched_noalloc (no reference counterpart): the batched
tridiagonal solve with its work array in the library's per-GPU workspace
instead of a per-call allocation, for callers that cannot allocate.  Called
through the C ABI and checked bit for bit against nova_tdma_batched on two
calls in a row, the second with a smaller input, plain and cyclic, and
against a dense solve; a call beyond the workspace is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr())


def _systems(rng, nb, n):
    a, c = rng.standard_normal((nb, n)), rng.standard_normal((nb, n))
    b = 4.0 + np.abs(rng.standard_normal((nb, n)))
    d = rng.standard_normal((nb, n))
    return a, b, c, d


@pytest.mark.parametrize("cyclic", [0, 1])
def test_noalloc_tdma_equals_allocating_entry_point(cyclic):
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available()
    lib = nv.lib()
    G = nv.Mesh((8, 8, 8), (1.0, 1.0, 1.0))
    rng = np.random.default_rng(11 + cyclic)
    for nb, n in ((3000, 129), (700, 64)):  # the second call smaller than the first
        print("case", "cyclic" if cyclic else "plain", nb, n, flush=True)
        hs = _systems(rng, nb, n)
        a, b, c, d = (torch.from_numpy(v).cuda() for v in hs)
        x_ref = torch.empty_like(d)
        x = torch.full_like(d, float("nan"))
        torch.cuda.synchronize()
        assert lib.nova_tdma_batched(G.handle, n, nb, _p(a), _p(b), _p(c), _p(d), _p(x_ref), cyclic) == 0
        assert lib.nova_tdma_batched_noalloc(G.handle, n, nb, _p(a), _p(b), _p(c), _p(d), _p(x), cyclic) == 0, \
            lib.nova_last_error()
        torch.cuda.synchronize()
        assert torch.equal(x, x_ref)
        xs = x.cpu().numpy()
        for s in (0, nb // 2, nb - 1):
            A = np.diag(hs[1][s]) + np.diag(hs[0][s, 1:], -1) + np.diag(hs[2][s, :-1], 1)
            if cyclic:
                A[0, n - 1] = hs[0][s, 0]
                A[n - 1, 0] = hs[2][s, n - 1]
            assert np.allclose(xs[s], np.linalg.solve(A, hs[3][s]), rtol=0, atol=1e-12)


def test_noalloc_tdma_refuses_sizes_beyond_its_workspace():
    import torch
    import novaclone_amd as nv
    lib = nv.lib()
    G = nv.Mesh((8, 8, 8), (1.0, 1.0, 1.0))
    cap = lib.nova_tdma_noalloc_max()
    assert cap == 1 << 21
    n = 64
    nb = cap // (3 * n) + 1
    a = torch.zeros((1,), dtype=torch.float64, device="cuda")  # never read: the size check comes first
    rc = lib.nova_tdma_batched_noalloc(G.handle, n, nb, _p(a), _p(a), _p(a), _p(a), _p(a), 0)
    assert rc == -1  # NOVA_ERR_ARG
    assert b"workspace" in lib.nova_last_error()
