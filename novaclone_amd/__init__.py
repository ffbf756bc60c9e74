"""MI355X-native NOVA hot path: Python host mirror over the C ABI.

The names follow the reference's host API (``Mesh``, ``ex.div``/``grad``/
``laplacian``/``curl``, ``poisson.rhs``/``solve``, the app-level RK loop) so
that a NOVA user finds the same operators.  Every call goes to the gfx950
library ``libnova_hip.so`` through the C ABI declared in
``include/nova_capi.h``; there is no CPU fallback: importing this package
without the built library, or calling it without a GPU, raises.

Device memory and streams come from PyTorch (plumbing only): a field is a
float64 CUDA tensor of shape ``(ncomp, s0, s1, s2)`` in the compact layout
(k fastest, no ghost layers).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libnova_hip.so"

NOVA_OK = 0
BC_PERIODIC, BC_WALL = 0, 1
RK_LITERAL4, RK_SSP3, RK_SHEAR4 = 0, 1, 2
MODEL_NS, MODEL_MHD = 0, 1
KERNEL_NS_STAGE, KERNEL_MHD_STAGE, KERNEL_TLM_STAGE, KERNEL_ADJ_STAGE, KERNEL_NS_FIRST_STAGE = 0, 1, 2, 3, 4
SCHEME_EXPLICIT, SCHEME_COMPACT6 = 0, 1
WALL_NEUMANN0, WALL_DIRICHLET0 = 1, 2          # nova_ctx_set_wall_bc kinds
WALL_SOLVE_EVEN, WALL_SOLVE_REFERENCE = 0, 1   # nova_poisson_set_wall_solve
# reference BC kinds for the padded-layout fill (src/BC/BCenums.h:8)
PBC_PERIODIC, PBC_NEUMANN, PBC_DIRICHLET, PBC_EVEN, PBC_ODD = 0, 1, 2, 10, 11
# pointwise Field products (nova_field_op)
OP_DOT, OP_CROSS, OP_SCALE, OP_DIVIDE, OP_DDOT, OP_SQRT, OP_OUTER, OP_ADD_CONST = range(8)
_OP_OUT = {OP_DOT: lambda a, b: 1, OP_CROSS: lambda a, b: 3, OP_SCALE: lambda a, b: b, OP_DIVIDE: lambda a, b: a,
           OP_DDOT: lambda a, b: 1, OP_SQRT: lambda a, b: 1, OP_OUTER: lambda a, b: 9, OP_ADD_CONST: lambda a, b: a}
DIRS = {"north": 0, "south": 1, "east": 2, "west": 3, "top": 4, "bottom": 5}


class NovaError(RuntimeError):
    pass


class GridDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3),
        ("L", C.c_double * 3),
        ("origin", C.c_double * 3),
        ("m", C.c_int),
        ("bc", C.c_int * 3),
        ("hyp", C.c_int),
        ("s1", C.c_double),
        ("s2", C.c_double),
    ]


class InstanceDesc(C.Structure):
    _fields_ = [("grid", GridDesc), ("model", C.c_int), ("nu", C.c_double), ("eta", C.c_double)]


_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_DP = C.POINTER(C.c_double)

# name -> (argtypes); every function returns int status unless noted
_SIGS = {
    "nova_ctx_create": [C.POINTER(GridDesc), _I, C.POINTER(_P)],
    "nova_ctx_destroy": [_P],
    "nova_ctx_set_stream": [_P, _P],
    "nova_ctx_sync": [_P],
    "nova_ctx_info": [_P, C.POINTER(C.c_int64)],
    "nova_ctx_metric": [_P, _DP, _DP],
    "nova_timing_enable": [_P, _I],
    "nova_ctx_tune": [_P, _I, _I, _I],
    "nova_ctx_tune_kernel": [_P, _I, _I, _I],
    "nova_timing_collect": [_P, C.c_char_p, C.c_size_t],
    "nova_alloc": [_P, _I, C.POINTER(_P)],
    "nova_free": [_P, _P],
    "nova_memcpy_h2d": [_P, _P, _P, C.c_size_t],
    "nova_memcpy_d2h": [_P, _P, _P, C.c_size_t],
    "nova_host_alloc": [C.c_size_t, C.POINTER(_P)],
    "nova_host_free": [_P],
    "nova_pack_padded": [_P, _P, _I, _P],
    "nova_unpack_padded": [_P, _P, _I, _P],
    "nova_bc_fill_padded": [_P, _P, _I, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), _DP],
    "nova_div_adv": [_P, _P, _P, _I, _P],
    "nova_div": [_P, _P, _P],
    "nova_grad": [_P, _P, _P],
    "nova_laplacian": [_P, _D, _P, _I, _P],
    "nova_laplacian_var": [_P, _P, _P, _I, _P],
    "nova_curl_literal": [_P, _P, _P],
    "nova_curl": [_P, _P, _P],
    "nova_axpby": [_P, C.c_size_t, _D, _P, _D, _P, _P],
    "nova_field_op": [_P, _I, _P, _I, _P, _I, _DP, _P],
    "nova_grad_vec_literal": [_P, _P, _P],
    "nova_cfl": [_P, _P, _D, _DP],
    "nova_hdf5_write": [_P, C.c_char_p, _P, _I, _I],
    "nova_hdf5_write_host": [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _DP],
    "nova_hdf5_read_host": [C.c_char_p, C.POINTER(C.c_int64), _DP, C.c_size_t],
    "nova_poisson_create": [_P, C.POINTER(_P)],
    "nova_poisson_destroy": [_P],
    "nova_poisson_solve": [_P, _P, _P],
    "nova_project": [_P, _P, _D, _P],
    "nova_solver_create": [_P, _I, _I, _D, _D, C.POINTER(_P)],
    "nova_solver_destroy": [_P],
    "nova_solver_step": [_P, _D, _I, _P, _P, _P, _P],
    "nova_solver_use_graph": [_P, _I],
    "nova_solver_set_scheme": [_P, _I],
    "nova_poisson_set_scheme": [_P, _I],
    "nova_compact_deriv": [_P, _P, _I, _I, _I, _P],
    "nova_solver_autotune": [_P, _P, _P, _I, C.c_char_p, C.c_size_t],
    "nova_reduce_ke": [_P, _P, _DP],
    "nova_reduce_maxabs_div": [_P, _P, _DP],
    "nova_reduce_dissipation": [_P, _P, _D, _DP],
    "nova_mhd_diag": [_P, _P, _P, _D, _D, _DP],
    "nova_tdma_batched": [_P, _I, _I, _P, _P, _P, _P, _P, _I],
    "nova_tdma_batched_noalloc": [_P, _I, _I, _P, _P, _P, _P, _P, _I],
    "nova_wc_create": [_P, _D, _D, C.POINTER(_P)],
    "nova_wc_destroy": [_P],
    "nova_wc_step": [_P, _D, _I, _P, _P, _P],
    "nova_wc_set_magnetic": [_P, _P, _D],
    "nova_wc_set_temperature": [_P, _D, _D, _D, _D],
    "nova_wc_step_t": [_P, _D, _I, _P, _P, _P, _P],
    "nova_wc_hydrostatic": [_P, _P, _D, _P, _P],
    "nova_fd_laplacian": [_P, _D, _P, C.POINTER(_P)],
    "nova_fd_destroy": [_P],
    "nova_fd_add_source": [_P, _P],
    "nova_fd_relax": [_P, _P, _D],
    "nova_fd_set_reference": [_P, _I, _I, _I, _D],
    "nova_fd_combine": [_P, _P, _D],
    "nova_fd_residual": [_P, _P, _DP],
    "nova_fd_solve": [_P, _P, _D, _I, C.POINTER(C.c_int), _DP],
    "nova_fd_get": [_P, _I, _P],
    "nova_time_create": [_D, _D, _I, C.POINTER(_P)],
    "nova_time_create_interval": [_D, _D, _D, C.POINTER(_P)],
    "nova_time_create_from_file": [C.c_char_p, C.POINTER(_P)],
    "nova_time_restart": [_P, C.c_char_p],
    "nova_time_destroy": [_P],
    "nova_time_advance": [_P],
    "nova_time_set_dt": [_P, _D],
    "nova_time_info": [_P, _DP, C.POINTER(C.c_int64)],
    "nova_time_steps": [_P, _DP, C.c_int64],
    "nova_io_create": [_P, C.POINTER(_P)],
    "nova_io_destroy": [_P],
    "nova_io_wait": [_P],
    "nova_io_write_field": [_P, _P, C.c_char_p, C.c_char_p, _P, _I],
    "nova_io_write_mesh": [_P, _P, C.c_char_p],
    "nova_io_read_field": [_P, C.c_char_p, C.c_char_p, _I, _P, _I],
    "nova_io_checkpoint_write": [_P, C.c_char_p, _I, C.POINTER(_P), C.POINTER(C.c_int), _D, C.c_int64],
    "nova_io_checkpoint_read": [_P, C.c_char_p, _I, C.POINTER(_P), C.POINTER(C.c_int), _DP,
                                C.POINTER(C.c_int64)],
    "nova_ensemble_create": [C.POINTER(InstanceDesc), _I, _I, C.POINTER(_P)],
    "nova_ensemble_destroy": [_P],
    "nova_ensemble_set_stream": [_P, _P],
    "nova_ensemble_copy": [_P, _I, _I, _P, _P, _P, _P],
    "nova_ensemble_step": [_P, _D, _I],
    "nova_ensemble_use_graph": [_P, _I],
    "nova_ensemble_info": [_P, C.POINTER(C.c_int64)],
    "nova_adjoint_create": [_P, _I, _I, C.POINTER(_P)],
    "nova_adjoint_destroy": [_P],
    "nova_adjoint_forward": [_P, _D, _I, _P, _P],
    "nova_adjoint_backward": [_P, _P, _P],
    "nova_adjoint_tlm": [_P, _P, _P],
    "nova_adjoint_autotune": [_P, _I, C.c_char_p, C.c_size_t],
    "nova_ctx_set_wall_bc": [_P, _I, _P],
    "nova_wall_fill": [_P, _P, _I],
    "nova_poisson_set_wall_solve": [_P, _I],
    "nova_solver_set_force": [_P, _P],
    "nova_adjoint_reserve": [_P, C.c_uint64, _P],
    "nova_weighted_dot": [_P, _P, _P, _I, _DP],
    "nova_enstrophy": [_P, _P, _DP, _P],
}


def load_library(path: os.PathLike | str = LIB_PATH) -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    p = Path(path)
    if not p.exists():
        raise NovaError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.nova_last_error.argtypes = []
    lib.nova_last_error.restype = C.c_char_p
    lib.nova_tdma_noalloc_max.argtypes = []
    lib.nova_tdma_noalloc_max.restype = C.c_size_t
    return lib


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != NOVA_OK:
        msg = lib().nova_last_error().decode(errors="replace")
        raise NovaError(f"{what} failed ({rc}): {msg}")


def _ptr(t: torch.Tensor | None):
    if t is None:
        return None
    if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
        raise NovaError("fields must be contiguous float64 CUDA tensors")
    return C.c_void_p(t.data_ptr())


def _fld(t: torch.Tensor | None, mesh, ncomp: int, what: str = "field", npts: int | None = None):
    """_ptr of a field that the kernels read or write as ncomp components of
    the mesh's stored nodes (npts, default mesh.np): anything else would be
    an out-of-bounds device access, so the element count and the device are
    checked before the call."""
    if t is None:
        return None
    n = mesh.np if npts is None else npts
    if t.numel() != ncomp * n:
        raise NovaError(f"{what}: {t.numel()} elements, the mesh needs {ncomp} x {n} = {ncomp * n} "
                        f"(shape {tuple(t.shape)})")
    if t.device != mesh.device:
        raise NovaError(f"{what} is on {t.device}, the mesh on {mesh.device}")
    return _ptr(t)


class Mesh:
    """Grid + device context (reference ``Mesh``, src/Mesh/Mesh.cpp:24-179).

    ``N`` cells, ``L`` lengths, ``origin``, stencil order ``m`` and a BC kind
    per direction (``"periodic"`` or ``"wall"``: reflecting walls at nodes 0
    and N, stored as N+1 nodes)."""

    def __init__(self, N, L, origin=(0.0, 0.0, 0.0), m: int = 6, bc=("periodic",) * 3,
                 device: int = 0, hyp: tuple[float, float] | None = None):
        if not torch.cuda.is_available():
            raise NovaError("no GPU visible: the NOVA HIP path has no CPU fallback")
        d = GridDesc()
        for a in range(3):
            d.n[a] = int(N[a])
            d.L[a] = float(L[a])
            d.origin[a] = float(origin[a])
            d.bc[a] = BC_WALL if bc[a] == "wall" else BC_PERIODIC
        d.m = int(m)
        d.hyp = 1 if hyp else 0
        d.s1, d.s2 = (float(hyp[0]), float(hyp[1])) if hyp else (0.0, 0.0)
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        h = C.c_void_p()
        _check(lib().nova_ctx_create(C.byref(d), int(device), C.byref(h)), "nova_ctx_create")
        self._h = h
        self.desc = d
        self.N = tuple(int(x) for x in N)
        self.L = tuple(float(x) for x in L)
        self.m = int(m)
        self.bc = tuple(bc)
        info = (C.c_int64 * 8)()
        _check(lib().nova_ctx_info(h, info), "nova_ctx_info")
        self.shape = (info[0], info[1], info[2])
        self.np = info[3]
        self.padded = (info[4], info[5], info[6])
        self.sync_stream()

    @property
    def handle(self):
        return self._h

    def sync_stream(self) -> None:
        """Issue work on torch's current stream."""
        s = torch.cuda.current_stream(self.device).cuda_stream
        _check(lib().nova_ctx_set_stream(self._h, C.c_void_p(s)), "nova_ctx_set_stream")

    def tune(self, stage_variant: int = 3, stage_ichunk: int = 0, proj_variant: int = 2) -> None:
        """Pick the stage-kernel variant of every stage model, its i-chunk
        (0 = automatic) and the projection variant."""
        _check(lib().nova_ctx_tune(self._h, int(stage_variant), int(stage_ichunk), int(proj_variant)),
               "nova_ctx_tune")

    def set_wall_bc(self, direction: int, kinds) -> None:
        """The reference's boundary fills (dirichletBC / neumannBC with value
        0) on the walls normal to ``direction``: ``kinds[c]`` per vector
        component (WALL_NEUMANN0 or WALL_DIRICHLET0); wall nodes are then
        overwritten after every assignment, as the fills do."""
        k = (C.c_int * 3)(*[int(x) for x in kinds])
        _check(lib().nova_ctx_set_wall_bc(self._h, int(direction), k), "nova_ctx_set_wall_bc")

    def wall_fill(self, F: torch.Tensor) -> None:
        """The wall-node overwrite of those fills on one field, in place."""
        nc = F.shape[0] if F.dim() == 4 else 1
        _check(lib().nova_wall_fill(self._h, _fld(F, self, nc), int(nc)), "nova_wall_fill")

    def tune_kernel(self, kernel: int, stage_variant: int, stage_ichunk: int = 0) -> None:
        """Launch shape of one stage kernel (KERNEL_NS_STAGE .. KERNEL_ADJ_STAGE)."""
        _check(lib().nova_ctx_tune_kernel(self._h, int(kernel), int(stage_variant), int(stage_ichunk)),
               "nova_ctx_tune_kernel")

    def timing(self, enable: bool) -> None:
        """Bracket every library kernel with HIP events (per-kernel times)."""
        _check(lib().nova_timing_enable(self._h, int(enable)), "nova_timing_enable")

    def timing_collect(self) -> dict:
        """{kernel: (launches, total_ms)} since the last collect."""
        buf = C.create_string_buffer(1 << 16)
        _check(lib().nova_timing_collect(self._h, buf, len(buf)), "nova_timing_collect")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def field(self, ncomp: int = 3) -> torch.Tensor:
        return torch.zeros((ncomp, *self.shape), dtype=torch.float64, device=self.device)

    def dzdxi(self):
        import numpy as np
        a = np.zeros(self.shape[2])
        b = np.zeros(self.shape[2])
        _check(lib().nova_ctx_metric(self._h, a.ctypes.data_as(_DP), b.ctypes.data_as(_DP)), "nova_ctx_metric")
        return a, b

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nova_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # reference-layout boundary
    def pack_padded(self, f: torch.Tensor) -> torch.Tensor:
        nc = f.shape[0]
        out = torch.empty((*self.padded, nc), dtype=torch.float64, device=self.device)
        _check(lib().nova_pack_padded(self._h, _fld(f, self, nc), nc, _ptr(out)), "nova_pack_padded")
        return out

    def unpack_padded(self, padded: torch.Tensor) -> torch.Tensor:
        nc = padded.shape[-1]
        out = self.field(nc)
        npad = self.padded[0] * self.padded[1] * self.padded[2]
        _check(lib().nova_unpack_padded(self._h, _fld(padded.contiguous(), self, nc, "padded field", npad), nc,
                                        _ptr(out)), "nova_unpack_padded")
        return out

    def bc_fill_padded(self, padded: torch.Tensor, entries) -> None:
        """Apply reference BC entries (dir, kind, comp, val) in order to a
        padded AoS field in place."""
        n = len(entries)
        dirs = (C.c_int * max(n, 1))(*[DIRS[e[0]] if isinstance(e[0], str) else e[0] for e in entries])
        kinds = (C.c_int * max(n, 1))(*[e[1] for e in entries])
        comps = (C.c_int * max(n, 1))(*[e[2] for e in entries])
        vals = (C.c_double * max(n, 1))(*[float(e[3]) for e in entries])
        npad = self.padded[0] * self.padded[1] * self.padded[2]
        pp = _fld(padded, self, padded.shape[-1], "padded field", npad)
        _check(lib().nova_bc_fill_padded(self._h, pp, padded.shape[-1], n, dirs, kinds, comps, vals),
               "nova_bc_fill_padded")


class ex:
    """Explicit operators (reference namespace ``ex``, src/ex)."""

    @staticmethod
    def div(mesh: Mesh, U: torch.Tensor, phi: torch.Tensor | None = None) -> torch.Tensor:
        if phi is None:
            out = mesh.field(1)
            _check(lib().nova_div(mesh.handle, _fld(U, mesh, 3, "U"), _ptr(out)), "nova_div")
            return out
        out = mesh.field(phi.shape[0])
        _check(lib().nova_div_adv(mesh.handle, _fld(U, mesh, 3, "U"), _fld(phi, mesh, phi.shape[0], "phi"),
                                  phi.shape[0], _ptr(out)), "nova_div_adv")
        return out

    @staticmethod
    def grad(mesh: Mesh, p: torch.Tensor) -> torch.Tensor:
        out = mesh.field(3)
        _check(lib().nova_grad(mesh.handle, _fld(p, mesh, 1, "p"), _ptr(out)), "nova_grad")
        return out

    @staticmethod
    def laplacian(mesh: Mesh, nu, f: torch.Tensor) -> torch.Tensor:
        """ex::laplacian(nu, f): nu a number, or a one-component field (the
        variable-coefficient form, laplacian.hxx:64-127)."""
        out = mesh.field(f.shape[0])
        if isinstance(nu, torch.Tensor):
            _check(lib().nova_laplacian_var(mesh.handle, _fld(nu, mesh, 1, "nu"), _fld(f, mesh, f.shape[0]),
                                            f.shape[0], _ptr(out)),
                   "nova_laplacian_var")
        else:
            _check(lib().nova_laplacian(mesh.handle, float(nu), _fld(f, mesh, f.shape[0]), f.shape[0], _ptr(out)),
                   "nova_laplacian")
        return out

    @staticmethod
    def curl(mesh: Mesh, U: torch.Tensor, literal: bool = True) -> torch.Tensor:
        out = mesh.field(3)
        fn = lib().nova_curl_literal if literal else lib().nova_curl
        _check(fn(mesh.handle, _fld(U, mesh, 3, "U"), _ptr(out)), "nova_curl")
        return out


def compact_deriv(mesh: "Mesh", f: torch.Tensor, direction: int, order: int = 1) -> torch.Tensor:
    """Compact sixth-order derivative (order 1 or 2) of every component of f
    along a direction (nova_compact_deriv)."""
    out = torch.empty_like(f)
    _check(lib().nova_compact_deriv(mesh.handle, _fld(f, mesh, f.shape[0]), f.shape[0], int(direction), int(order),
                                    _ptr(out)),
           "nova_compact_deriv")
    return out


class Poisson:
    """FFT (and wall-y DCT) Poisson solver (reference ``poisson``,
    src/poisson/poisson.cpp)."""

    def __init__(self, mesh: Mesh, scheme: int = SCHEME_EXPLICIT):
        self.mesh = mesh
        h = C.c_void_p()
        _check(lib().nova_poisson_create(mesh.handle, C.byref(h)), "nova_poisson_create")
        self._h = h
        self.scheme = scheme
        if scheme != SCHEME_EXPLICIT:
            _check(lib().nova_poisson_set_scheme(h, int(scheme)), "nova_poisson_set_scheme")

    @property
    def handle(self):
        return self._h

    def solve(self, rhs: torch.Tensor, p: torch.Tensor | None = None) -> torch.Tensor:
        if p is None:
            p = self.mesh.field(1)
        _check(lib().nova_poisson_solve(self._h, _fld(rhs, self.mesh, 1, "rhs"), _fld(p, self.mesh, 1, "p")),
               "nova_poisson_solve")
        return p

    def project(self, F: torch.Tensor, s: float = 1.0, p: torch.Tensor | None = None) -> None:
        _check(lib().nova_project(self._h, _fld(F, self.mesh, 3, "F"), float(s), _fld(p, self.mesh, 1, "p")),
               "nova_project")

    def set_wall_solve(self, mode: int) -> None:
        """WALL_SOLVE_EVEN (the even extension, default) or
        WALL_SOLVE_REFERENCE (the reference's FFT, y periodic over the nodes
        0 .. N_y - 1) across a walled y."""
        _check(lib().nova_poisson_set_wall_solve(self._h, int(mode)), "nova_poisson_set_wall_solve")

    def close(self):
        if getattr(self, "_h", None):
            lib().nova_poisson_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """The app-level time loop (app/TG_Piosson/RKLoop.H, app/OT_Poisson/
    {UEqn,BEqn}.H, app/shearLayer/UEqn.H) as fused stage kernels."""

    def __init__(self, poisson: Poisson, model: int = MODEL_NS, mode: int = RK_SSP3, nu: float = 0.0,
                 eta: float = 0.0, graph: bool = False):
        self.poisson = poisson
        self.mesh = poisson.mesh
        h = C.c_void_p()
        _check(lib().nova_solver_create(poisson.handle, model, mode, float(nu), float(eta), C.byref(h)),
               "nova_solver_create")
        self._h = h
        self.model = model
        if poisson.scheme != SCHEME_EXPLICIT:
            _check(lib().nova_solver_set_scheme(h, int(poisson.scheme)), "nova_solver_set_scheme")
        if graph:
            _check(lib().nova_solver_use_graph(h, 1), "nova_solver_use_graph")

    def step(self, dt: float, nsteps: int, U, p, B=None, pB=None) -> None:
        m = self.mesh
        _check(lib().nova_solver_step(self._h, float(dt), int(nsteps), _fld(U, m, 3, "U"), _fld(B, m, 3, "B"),
                                      _fld(p, m, 1, "p"), _fld(pB, m, 1, "pB")), "nova_solver_step")

    def set_force(self, F: torch.Tensor | None) -> None:
        """A constant vector field added to the NS right-hand side of every
        stage (kept referenced here while it is in use)."""
        self._force = F
        _check(lib().nova_solver_set_force(self._h, _fld(F, self.mesh, 3, "F")), "nova_solver_set_force")

    def autotune(self, U, B=None, reps: int = 3) -> str:
        """Time every stage launch shape and both projection variants on U
        (and B), keep the fastest, and return the trial report."""
        buf = C.create_string_buffer(1 << 16)
        _check(lib().nova_solver_autotune(self._h, _fld(U, self.mesh, 3, "U"), _fld(B, self.mesh, 3, "B"),
                                          int(reps), buf, len(buf)),
               "nova_solver_autotune")
        return buf.value.decode()

    def close(self):
        if getattr(self, "_h", None):
            lib().nova_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FdMatrix:
    """im::fdMatrix<scalar> built by im::laplacian(mu, phi)
    (src/im/laplacian/laplacian.hxx:23-58): a 7-point matrix with the
    explicit Laplacian's deferred correction, solved by line Gauss-Seidel
    sweeps (fdMatrix::solve).  Fields are stored-layout scalars of shape
    (1, s0, s1, s2); the matrix's own arrays are padded."""

    def __init__(self, mesh: Mesh, mu: float, phi: torch.Tensor, _handle=None):
        self.mesh = mesh
        h = C.c_void_p()
        if _handle is not None:
            h = _handle
        else:
            _check(lib().nova_fd_laplacian(mesh.handle, float(mu), _fld(phi, mesh, 1, "phi"), C.byref(h)),
                   "nova_fd_laplacian")
        self._h = h

    def add_source(self, f: torch.Tensor) -> "FdMatrix":
        """matrix == f: the right-hand side gains f."""
        _check(lib().nova_fd_add_source(self._h, _fld(f, self.mesh, 1)), "nova_fd_add_source")
        return self

    def relax(self, phi: torch.Tensor, urf: float) -> None:
        _check(lib().nova_fd_relax(self._h, _fld(phi, self.mesh, 1, "phi"), float(urf)), "nova_fd_relax")

    def set_reference(self, i: int, j: int, k: int, val: float) -> None:
        _check(lib().nova_fd_set_reference(self._h, int(i), int(j), int(k), float(val)), "nova_fd_set_reference")

    def __iadd__(self, other: "FdMatrix"):
        _check(lib().nova_fd_combine(self._h, other._h, 1.0), "nova_fd_combine")
        return self

    def __isub__(self, other: "FdMatrix"):
        _check(lib().nova_fd_combine(self._h, other._h, -1.0), "nova_fd_combine")
        return self

    def negate(self) -> "FdMatrix":
        _check(lib().nova_fd_combine(self._h, None, -1.0), "nova_fd_combine")
        return self

    def residual(self, phi: torch.Tensor) -> float:
        r = C.c_double()
        _check(lib().nova_fd_residual(self._h, _fld(phi, self.mesh, 1, "phi"), C.byref(r)), "nova_fd_residual")
        return r.value

    def solve(self, phi: torch.Tensor, tol: float, max_its: int):
        """Solve for phi in place; returns (iterations, last residual)."""
        its = C.c_int()
        r = C.c_double()
        _check(lib().nova_fd_solve(self._h, _fld(phi, self.mesh, 1, "phi"), float(tol), int(max_its), C.byref(its),
                                   C.byref(r)),
               "nova_fd_solve")
        return its.value, r.value

    def padded(self, what: int) -> torch.Tensor:
        """al(0..5), 6 ap, 7 rhs, 8 the padded transported field."""
        out = torch.empty(self.mesh.padded, dtype=torch.float64, device=self.mesh.device)
        _check(lib().nova_fd_get(self._h, int(what), _ptr(out)), "nova_fd_get")
        torch.cuda.synchronize()
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().nova_fd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Time:
    """The run clock (src/Time/Time.cpp): ``Time(dt, end, writeEvery)`` with
    an int writes every writeEvery steps, with a float at multiples of that
    time interval; ``Time.from_file(settings)`` reads "dt:", "writeT:",
    "writeN:", "end:"; ``restart(file)`` resumes from a restart file.
    ``advance()`` is ``operator++``."""

    def __init__(self, dt: float, end_time: float, write_every, _handle=None):
        h = C.c_void_p()
        if _handle is not None:
            h = _handle
        elif isinstance(write_every, float):
            _check(lib().nova_time_create_interval(float(dt), float(end_time), write_every, C.byref(h)),
                   "nova_time_create_interval")
        else:
            _check(lib().nova_time_create(float(dt), float(end_time), int(write_every), C.byref(h)),
                   "nova_time_create")
        self._h = h

    @classmethod
    def from_file(cls, path) -> "Time":
        h = C.c_void_p()
        _check(lib().nova_time_create_from_file(str(path).encode(), C.byref(h)), "nova_time_create_from_file")
        return cls(0.0, 0.0, 0, _handle=h)

    def restart(self, restart_file) -> None:
        _check(lib().nova_time_restart(self._h, str(restart_file).encode()), "nova_time_restart")

    def advance(self) -> None:
        _check(lib().nova_time_advance(self._h), "nova_time_advance")

    def _info(self):
        d = (C.c_double * 4)()
        i = (C.c_int64 * 5)()
        _check(lib().nova_time_info(self._h, d, i), "nova_time_info")
        return list(d), list(i)

    def dt(self) -> float:
        return self._info()[0][0]

    def set_dt(self, dt: float) -> None:
        _check(lib().nova_time_set_dt(self._h, float(dt)), "nova_time_set_dt")

    def curTime(self) -> float:
        return self._info()[0][1]

    def run(self) -> bool:
        d, _ = self._info()
        return d[1] < d[2]

    def write(self) -> bool:
        return bool(self._info()[1][2])

    def writelog(self) -> bool:
        return bool(self._info()[1][3])

    def timeStep(self) -> int:
        return self._info()[1][0]

    def fileNum(self) -> int:
        return self._info()[1][1]

    def steps(self) -> list:
        n = self._info()[1][4]
        out = (C.c_double * max(n, 1))()
        _check(lib().nova_time_steps(self._h, out, n), "nova_time_steps")
        return list(out)[:n]

    def close(self):
        if getattr(self, "_h", None):
            lib().nova_time_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IO:
    """EnSight Gold output and restart of one mesh's fields
    (Mesh::write, Field::write/read) and FP64 checkpoints.  Writes are
    asynchronous; ``wait()`` blocks until the files are written."""

    def __init__(self, mesh: "Mesh"):
        self.mesh = mesh
        h = C.c_void_p()
        _check(lib().nova_io_create(mesh.handle, C.byref(h)), "nova_io_create")
        self._h = h

    def write_mesh(self, time: Time, path) -> None:
        _check(lib().nova_io_write_mesh(self._h, time._h, str(path).encode()), "nova_io_write_mesh")

    def write_field(self, time: Time, path, name: str, F) -> None:
        _check(lib().nova_io_write_field(self._h, time._h, str(path).encode(), name.encode(),
                                         _fld(F, self.mesh, F.shape[0], name), F.shape[0]),
               "nova_io_write_field")

    def read_field(self, path, name: str, file_num: int, F) -> None:
        _check(lib().nova_io_read_field(self._h, str(path).encode(), name.encode(), int(file_num),
                                        _fld(F, self.mesh, F.shape[0], name), F.shape[0]), "nova_io_read_field")

    def checkpoint_write(self, file, fields, time: float = 0.0, step: int = 0) -> None:
        ptrs = (C.c_void_p * len(fields))(*[_fld(f, self.mesh, f.shape[0]).value for f in fields])
        nc = (C.c_int * len(fields))(*[f.shape[0] for f in fields])
        _check(lib().nova_io_checkpoint_write(self._h, str(file).encode(), len(fields), ptrs, nc, float(time),
                                              int(step)), "nova_io_checkpoint_write")

    def checkpoint_read(self, file, fields):
        ptrs = (C.c_void_p * len(fields))(*[_fld(f, self.mesh, f.shape[0]).value for f in fields])
        nc = (C.c_int * len(fields))(*[f.shape[0] for f in fields])
        t = C.c_double()
        st = C.c_int64()
        _check(lib().nova_io_checkpoint_read(self._h, str(file).encode(), len(fields), ptrs, nc, C.byref(t),
                                             C.byref(st)), "nova_io_checkpoint_read")
        return t.value, st.value

    def wait(self) -> None:
        _check(lib().nova_io_wait(self._h), "nova_io_wait")

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            _check(lib().nova_io_destroy(h), "nova_io_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WC:
    """The weakly-compressible step (app/TG/TG.cpp:71-76 with rhoEqn.H,
    pEqn.H, UEqn.H): a classic RK4 of rho with U frozen, p = pc rho^2
    (pc = 50 in pEqn.H), and a classic RK4 of rho U with rho and p frozen."""

    def __init__(self, mesh: Mesh, mu: float, pc: float = 50.0):
        self.mesh = mesh
        h = C.c_void_p()
        _check(lib().nova_wc_create(mesh.handle, float(mu), float(pc), C.byref(h)), "nova_wc_create")
        self._h = h

    def set_magnetic(self, B, dbu: float) -> None:
        """The frozen magnetic field of app/OT and its coefficient DBU; the
        tensor must stay alive while it is set (None removes it)."""
        self._B = B
        _check(lib().nova_wc_set_magnetic(self._h, _fld(B, self.mesh, 3, "B"), float(dbu)), "nova_wc_set_magnetic")

    def step(self, dt: float, nsteps: int, U, rho, p) -> None:
        """Advance nsteps: U (3 components) and rho in/out, p out."""
        m = self.mesh
        _check(lib().nova_wc_step(self._h, float(dt), int(nsteps), _fld(U, m, 3, "U"), _fld(rho, m, 1, "rho"),
                                  _fld(p, m, 1, "p")), "nova_wc_step")

    def set_temperature(self, ka: float, g) -> None:
        """Enable the temperature equation and buoyancy of app/shearLayerWC."""
        _check(lib().nova_wc_set_temperature(self._h, float(ka), *map(float, g)), "nova_wc_set_temperature")

    def step_t(self, dt: float, nsteps: int, U, rho, p, T) -> None:
        """Advance nsteps in app/shearLayerWC order: U, rho, T in/out, p in/out."""
        m = self.mesh
        _check(lib().nova_wc_step_t(self._h, float(dt), int(nsteps), _fld(U, m, 3, "U"), _fld(rho, m, 1, "rho"),
                                    _fld(p, m, 1, "p"), _fld(T, m, 1, "T")), "nova_wc_step_t")

    def hydrostatic(self, T, rho0: float, p, rho) -> None:
        """app/shearLayerWC's initial state (hydrostaticP.H): p += the
        hydrostatic pressure of T under g_z, then rho = (p/pc)^(1/2)."""
        m = self.mesh
        _check(lib().nova_wc_hydrostatic(self._h, _fld(T, m, 1, "T"), float(rho0), _fld(p, m, 1, "p"),
                                         _fld(rho, m, 1, "rho")), "nova_wc_hydrostatic")

    def close(self):
        if getattr(self, "_h", None):
            lib().nova_wc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def field_op(mesh: Mesh, op: int, a: torch.Tensor, b: torch.Tensor | None = None, cval=None) -> torch.Tensor:
    """The reference's pointwise Field products (Field.h:227-363): OP_DOT
    (a & b), OP_CROSS (a ^ b), OP_SCALE (scalar field a times b), OP_DIVIDE
    (a over scalar field b), OP_DDOT (tensors a && b), OP_SQRT, OP_OUTER
    (vectors a * b), OP_ADD_CONST (a + cval)."""
    nca = a.shape[0]
    ncb = b.shape[0] if b is not None else 0
    out = mesh.field(_OP_OUT[op](nca, ncb))
    cv = (C.c_double * 3)(*(list(cval) + [0.0] * 3)[:3]) if cval is not None else None
    _check(lib().nova_field_op(mesh.handle, int(op), _fld(a, mesh, nca, "a"), nca, _fld(b, mesh, ncb, "b"), ncb,
                               cv, _ptr(out)), "nova_field_op")
    return out


def grad_vec_literal(mesh: Mesh, U: torch.Tensor) -> torch.Tensor:
    """ex::grad(Field<vector>): the literal half-shifted gradient tensor
    (9 components, row = component, column = direction)."""
    out = mesh.field(9)
    _check(lib().nova_grad_vec_literal(mesh.handle, _fld(U, mesh, 3, "U"), _ptr(out)), "nova_grad_vec_literal")
    return out


def cfl(mesh: Mesh, U: torch.Tensor, dt: float):
    """tools::CFL: (cfl_x, cfl_y, cfl_z) as the reference prints them, then
    the Courant numbers max |U_d| dt / h_d."""
    out = (C.c_double * 6)()
    _check(lib().nova_cfl(mesh.handle, _fld(U, mesh, 3, "U"), float(dt), out), "nova_cfl")
    return list(out)


def hdf5_write(mesh: Mesh, file, F: torch.Tensor, comp: int = 0) -> None:
    """app/mhdjet's HDF5 dump of one component of F (dataset /3DArray)."""
    _check(lib().nova_hdf5_write(mesh.handle, str(file).encode(), _fld(F, mesh, F.shape[0], "F"), F.shape[0],
                                 int(comp)), "nova_hdf5_write")


def hdf5_write_host(file, data, dims=None) -> None:
    """Write a host float64 array (s0, s1, s2) as /3DArray of extents dims
    (default: the array's)."""
    import numpy as np
    a = np.ascontiguousarray(data, dtype=np.float64)
    st = (C.c_int64 * 3)(*a.shape)
    dm = (C.c_int64 * 3)(*(dims if dims is not None else a.shape))
    _check(lib().nova_hdf5_write_host(str(file).encode(), dm, st, a.ctypes.data_as(_DP)), "nova_hdf5_write_host")


def hdf5_read_host(file):
    """/3DArray of an HDF5 file as a float64 numpy array."""
    import numpy as np
    dm = (C.c_int64 * 3)()
    _check(lib().nova_hdf5_read_host(str(file).encode(), dm, None, 0), "nova_hdf5_read_host")
    out = np.empty(tuple(dm), dtype=np.float64)
    _check(lib().nova_hdf5_read_host(str(file).encode(), dm, out.ctypes.data_as(_DP), out.size),
           "nova_hdf5_read_host")
    return out


def ke(mesh: Mesh, U: torch.Tensor) -> float:
    out = C.c_double()
    _check(lib().nova_reduce_ke(mesh.handle, _fld(U, mesh, 3, "U"), C.byref(out)), "nova_reduce_ke")
    return out.value


def maxabs_div(mesh: Mesh, U: torch.Tensor) -> float:
    out = C.c_double()
    _check(lib().nova_reduce_maxabs_div(mesh.handle, _fld(U, mesh, 3, "U"), C.byref(out)), "nova_reduce_maxabs_div")
    return out.value


def dissipation(mesh: Mesh, U: torch.Tensor, nu: float) -> float:
    out = C.c_double()
    _check(lib().nova_reduce_dissipation(mesh.handle, _fld(U, mesh, 3, "U"), float(nu), C.byref(out)),
           "nova_reduce_dissipation")
    return out.value


def mhd_diag(mesh: Mesh, U, B, mu: float, eta: float):
    out = (C.c_double * 8)()
    _check(lib().nova_mhd_diag(mesh.handle, _fld(U, mesh, 3, "U"), _fld(B, mesh, 3, "B"), float(mu), float(eta), out),
           "nova_mhd_diag")
    return list(out)


def tdma_batched(mesh: Mesh, a, b, c, d, cyclic: bool = False) -> torch.Tensor:
    nb, n = a.shape
    x = torch.empty_like(d)
    a, b, c, d = (_fld(v, mesh, 1, nm, nb * n) for v, nm in zip((a, b, c, d), "abcd"))
    _check(lib().nova_tdma_batched(mesh.handle, n, nb, a, b, c, d, _ptr(x), int(cyclic)),
           "nova_tdma_batched")
    return x


def weighted_dot(mesh: Mesh, a: torch.Tensor, b: torch.Tensor) -> float:
    """sum over nodes of w a.b, w = 1/2 on wall nodes."""
    out = C.c_double()
    _check(lib().nova_weighted_dot(mesh.handle, _fld(a, mesh, a.shape[0], "a"), _fld(b, mesh, a.shape[0], "b"),
                                   a.shape[0], C.byref(out)), "nova_weighted_dot")
    return out.value


def enstrophy(mesh: Mesh, U: torch.Tensor, grad: bool = False):
    """J = dV sum w |curl U|^2; with grad=True also dJ/dU = 2 dV curl curl U."""
    J = C.c_double()
    g = mesh.field(3) if grad else None
    _check(lib().nova_enstrophy(mesh.handle, _fld(U, mesh, 3, "U"), C.byref(J), _ptr(g)), "nova_enstrophy")
    return (J.value, g) if grad else J.value


class Adjoint:
    """Checkpointed discrete adjoint and tangent-linear model of the SSP-RK3
    + projection step (nova_adjoint_*)."""

    def __init__(self, solver: Solver, checkpoint_every: int, max_steps: int):
        self.solver = solver
        self.mesh = solver.mesh
        h = C.c_void_p()
        _check(lib().nova_adjoint_create(solver._h, int(checkpoint_every), int(max_steps), C.byref(h)),
               "nova_adjoint_create")
        self._h = h

    def forward(self, dt: float, nsteps: int, U0: torch.Tensor) -> torch.Tensor:
        U = torch.empty_like(U0)
        _check(lib().nova_adjoint_forward(self._h, float(dt), int(nsteps), _fld(U0, self.mesh, 3, "U0"), _ptr(U)),
               "nova_adjoint_forward")
        return U

    def backward(self, lam: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(lam)
        _check(lib().nova_adjoint_backward(self._h, _fld(lam.contiguous(), self.mesh, 3, "lambda"), _ptr(out)),
               "nova_adjoint_backward")
        return out

    def tlm(self, d0: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(d0)
        _check(lib().nova_adjoint_tlm(self._h, _fld(d0.contiguous(), self.mesh, 3, "dU0"), _ptr(out)),
               "nova_adjoint_tlm")
        return out

    def autotune(self, reps: int = 3) -> str:
        """Tune the adjoint and tangent-linear stage kernels on the first
        checkpoint of the last forward run; returns the trial report."""
        buf = C.create_string_buffer(1 << 16)
        _check(lib().nova_adjoint_autotune(self._h, int(reps), buf, len(buf)), "nova_adjoint_autotune")
        return buf.value.decode()

    def reserve(self, max_bytes: int) -> int:
        """Further pool fields (one vector field each) within max_bytes of
        device memory, so that the forward sweep keeps the step-start states
        of more of its last steps and the reverse sweep recomputes fewer
        segments; returns how many were allocated."""
        n = C.c_int(0)
        _check(lib().nova_adjoint_reserve(self._h, C.c_uint64(int(max_bytes)), C.byref(n)), "nova_adjoint_reserve")
        return n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().nova_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Batched ensemble of uneven periodic instances (nova_ensemble_*).

    ``instances`` is a list of ``(N, L, model, nu, eta)``; every stage of all
    instances is one persistent launch and, with ``graph``, a step is one
    replayed HIP graph.  Fields are copied in and out with ``load``/``store``."""

    def __init__(self, instances, m: int = 6, device: int = 0, graph: bool = True):
        if not torch.cuda.is_available():
            raise NovaError("no GPU visible: the NOVA HIP path has no CPU fallback")
        n = len(instances)
        arr = (InstanceDesc * n)()
        self.shapes, self.models = [], []
        for q, (N, L, model, nu, eta) in enumerate(instances):
            d = arr[q]
            for a in range(3):
                d.grid.n[a] = int(N[a])
                d.grid.L[a] = float(L[a])
                d.grid.origin[a] = 0.0
                d.grid.bc[a] = BC_PERIODIC
            d.grid.m = int(m)
            d.model = int(model)
            d.nu = float(nu)
            d.eta = float(eta)
            self.shapes.append(tuple(int(x) for x in N))
            self.models.append(int(model))
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        h = C.c_void_p()
        _check(lib().nova_ensemble_create(arr, n, int(device), C.byref(h)), "nova_ensemble_create")
        self._h = h
        s = torch.cuda.current_stream(self.device).cuda_stream
        _check(lib().nova_ensemble_set_stream(h, C.c_void_p(s)), "nova_ensemble_set_stream")
        if graph:
            _check(lib().nova_ensemble_use_graph(h, 1), "nova_ensemble_use_graph")

    def __len__(self) -> int:
        return len(self.shapes)

    def info(self) -> dict:
        out = (C.c_int64 * 6)()
        _check(lib().nova_ensemble_info(self._h, out), "nova_ensemble_info")
        return {"instances": out[0], "classes": out[1], "tiles": out[2], "blocks": out[3], "nodes": out[4],
                "slots": out[5]}

    def load(self, i: int, U: torch.Tensor, B: torch.Tensor | None = None) -> None:
        shp = self.shapes[i]
        npts = shp[0] * shp[1] * shp[2]
        if self.models[i] == MODEL_MHD and B is None:
            raise NovaError(f"instance {i} is MHD: load needs B")
        _check(lib().nova_ensemble_copy(self._h, int(i), 0, _fld(U.contiguous(), self, 3, "U", npts),
                                        _fld(B, self, 3, "B", npts), None, None), "nova_ensemble_copy")

    def store(self, i: int):
        """(U, B, p, pB) of instance i (B and pB are None for NS)."""
        shp = self.shapes[i]
        U = torch.empty((3, *shp), dtype=torch.float64, device=self.device)
        p = torch.empty((1, *shp), dtype=torch.float64, device=self.device)
        mhd = self.models[i] == MODEL_MHD
        B = torch.empty_like(U) if mhd else None
        pB = torch.empty_like(p) if mhd else None
        _check(lib().nova_ensemble_copy(self._h, int(i), 1, _ptr(U), _ptr(B), _ptr(p), _ptr(pB)),
               "nova_ensemble_copy")
        return U, B, p, pB

    def step(self, dt: float, nsteps: int) -> None:
        _check(lib().nova_ensemble_step(self._h, float(dt), int(nsteps)), "nova_ensemble_step")

    def close(self):
        if getattr(self, "_h", None):
            lib().nova_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EnstrophyAfterSteps(torch.autograd.Function):
    """J(U0) = enstrophy of the flow advanced nsteps from U0; backward runs the
    discrete adjoint.  Gradients are with respect to the node-sum inner
    product (weight 1/2 on wall nodes)."""

    @staticmethod
    def forward(ctx, U0: torch.Tensor, adjoint: "Adjoint", dt: float, nsteps: int):
        U = adjoint.forward(dt, nsteps, U0.detach().contiguous())
        J, g = enstrophy(adjoint.mesh, U, grad=True)
        ctx.adjoint = adjoint
        ctx.save_for_backward(g)
        return torch.tensor(J, dtype=torch.float64, device=U0.device)

    @staticmethod
    def backward(ctx, gJ: torch.Tensor):
        (g,) = ctx.saved_tensors
        lam0 = ctx.adjoint.backward(g)
        return gJ * lam0, None, None, None
