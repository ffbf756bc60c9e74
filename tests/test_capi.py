"""C ABI checks that need no GPU: the library loads, exports every entry
point that include/nova_capi.h declares, and fails loudly (error code +
message, no fallback) when no device is present or arguments are bad."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "nova_capi.h").read_text()
    body = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nova_[a-z0-9_]+)\s*\(", body)))


def test_library_exports_every_declared_symbol():
    import novaclone_amd as nv
    lib = nv.load_library()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the python mirror binds exactly the declared set
    assert set(nv._SIGS) | {"nova_last_error", "nova_tdma_noalloc_max"} == set(syms)


def test_bad_arguments_are_rejected_with_a_message():
    import novaclone_amd as nv
    lib = nv.load_library()
    d = nv.GridDesc()
    d.n[:] = [16, 16, 16]
    d.L[:] = [1.0, 1.0, 1.0]
    d.m = 5
    h = C.c_void_p()
    assert lib.nova_ctx_create(C.byref(d), 0, C.byref(h)) == -1
    assert b"m must be" in lib.nova_last_error()
    d.m = 6
    d.n[0] = 2
    assert lib.nova_ctx_create(C.byref(d), 0, C.byref(h)) == -1
    assert lib.nova_ctx_create(None, 0, C.byref(h)) == -1


def test_host_allocation_arguments_are_checked():
    """nova_host_alloc (the drop-in's page-locked host mirrors) refuses a
    null result pointer and a zero size before any runtime call;
    nova_host_free(NULL) is a no-op."""
    import novaclone_amd as nv
    lib = nv.load_library()
    p = C.c_void_p()
    assert lib.nova_host_alloc(0, C.byref(p)) == -1 and not p.value
    assert b"bad argument" in lib.nova_last_error()
    assert lib.nova_host_alloc(64, None) == -1
    assert lib.nova_host_free(None) == 0


def test_no_device_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import novaclone_amd as nv
    lib = nv.load_library()
    d = nv.GridDesc()
    d.n[:] = [16, 16, 16]
    d.L[:] = [1.0, 1.0, 1.0]
    d.m = 6
    h = C.c_void_p()
    assert lib.nova_ctx_create(C.byref(d), 0, C.byref(h)) == -2
    assert b"hipSetDevice" in lib.nova_last_error()
    with pytest.raises(nv.NovaError):
        nv.Mesh((16, 16, 16), (1.0, 1.0, 1.0))


class _StubMesh:
    """Stands in for a Mesh of 4 x 4 x 4 stored nodes (no device, no handle):
    the shape checks must raise before anything reaches the library."""
    np = 64
    padded = (10, 10, 10)
    handle = None
    _h = None

    def __init__(self):
        import torch
        self.device = torch.device("cuda", 0)

    def field(self, nc=3):
        import torch
        return torch.zeros((nc, 4, 4, 4), dtype=torch.float64)


def test_field_sizes_are_checked_before_any_launch():
    import torch
    import novaclone_amd as nv
    m = _StubMesh()
    short = torch.zeros((3, 4, 4, 3), dtype=torch.float64)   # one plane short
    scalar = torch.zeros((1, 4, 4, 4), dtype=torch.float64)
    vector = torch.zeros((3, 4, 4, 4), dtype=torch.float64)
    with pytest.raises(nv.NovaError, match="elements"):
        nv.ex.div(m, short)
    with pytest.raises(nv.NovaError, match="elements"):
        nv.ex.grad(m, vector)  # three components where one is read
    with pytest.raises(nv.NovaError, match="elements"):
        nv.ke(m, scalar)
    sv = object.__new__(nv.Solver)
    sv._h, sv.mesh = None, m
    with pytest.raises(nv.NovaError, match="elements"):
        sv.step(1e-3, 1, short, scalar)
    with pytest.raises(nv.NovaError, match="elements"):
        sv.step(1e-3, 1, None, vector)  # p must be one component
    ps = object.__new__(nv.Poisson)
    ps._h, ps.mesh = None, m
    with pytest.raises(nv.NovaError, match="elements"):
        ps.project(short, 1.0)
    # the right size on the wrong device is refused too
    with pytest.raises(nv.NovaError, match="is on cpu"):
        nv.ex.div(m, vector)
