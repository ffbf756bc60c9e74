"""ctypes binding of the oracle library (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg
import this module; it is the checker, never the product.  Fields are numpy
arrays in the reference's padded AoS layout, shape (n_i, n_j, n_k, ncomp).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# NOVA_ORACLE_LIB points the binding at another build of the same source
# (the planted-mistake builds of tests/test_oracle_planted.py)
LIB = Path(os.environ.get("NOVA_ORACLE_LIB", HERE / "_build" / "libnova_oracle.so"))

NORTH, SOUTH, EAST, WEST, TOP, BOTTOM = range(6)
BC_PERIODIC, BC_NEUMANN, BC_DIRICHLET, BC_EVEN, BC_ODD = 0, 1, 2, 10, 11
RK_LITERAL4, RK_SSP3, RK_SHEAR4 = 0, 1, 2

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_P = C.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _load() -> C.CDLL:
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    sig = {
        "orc_mesh_create": ([_I, _D, _D, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int], _P),
        "orc_mesh_free": ([_P], None),
        "orc_mesh_dims": ([_P, _I], None),
        "orc_mesh_metric": ([_P, _D, _D], None),
        "orc_mesh_z": ([_P, C.c_int], C.c_double),
        "orc_coef": ([C.c_int, _D], None),
        "orc_bc_apply": ([_P, _D, C.c_int, C.c_int, _I, _I, _I, _D], None),
        "orc_fill": ([_P, _D, C.c_int, C.c_int], None),
        "orc_div_adv": ([_P, _D, _D, C.c_int, _D], None),
        "orc_div": ([_P, _D, _D], None),
        "orc_grad": ([_P, _D, _D], None),
        "orc_laplacian": ([_P, C.c_double, _D, C.c_int, _D], None),
        "orc_laplacian_var": ([_P, _D, _D, C.c_int, _D], None),
        "orc_grad_vec_literal": ([_P, _D, _D], None),
        "orc_curl_literal": ([_P, _D, _D], None),
        "orc_curl_central": ([_P, _D, _D], None),
        "orc_fft1d": ([C.c_int, _D, C.c_int], None),
        "orc_fft3d": ([_I, _D, C.c_int], None),
        "orc_poisson_create": ([_P], _P),
        "orc_poisson_free": ([_P], None),
        "orc_poisson_k2": ([_P, _D], None),
        "orc_poisson_solve": ([_P, _D, _D], None),
        "orc_ns_step": ([_P, _P, _D, _D, _D, C.c_double, C.c_double, C.c_int], None),
        "orc_ns_step_fused": ([_P, _P, _D, _D, _D, _D, C.c_double, C.c_double], None),
        "orc_set_threads": ([C.c_int], None),
        "orc_mhd_rhs": ([_P, _D, _D, C.c_double, C.c_double, _D, _D], None),
        "orc_mhd_step": ([_P, _P, _D, _D, _D, _D, C.c_double, C.c_double, C.c_double, C.c_int], None),
        "orc_wc_step": ([_P, _D, _D, _D, _D, C.c_double, C.c_double, C.c_double, C.c_double], None),
        "orc_wcs_step": ([_P, _D, _D, _D, _D, C.c_double, C.c_double, _D, C.c_double, C.c_double], None),
        "orc_wc_hydrostatic": ([_P, _D, _D, _D, _D, C.c_double, C.c_double], None),
        "orc_fd_laplacian": ([_P, C.c_double, _D, _D, _D, _D], None),
        "orc_fd_residual": ([_P, _D, _D, _D, _D], C.c_double),
        "orc_fd_solve": ([_P, _D, _D, _D, _D, C.c_double, C.c_int, _D], C.c_int),
        "orc_fd_relax": ([_P, _D, _D, _D, C.c_double], None),
        "orc_field_write": ([_P, _D, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.c_int], C.c_int),
        "orc_field_read": ([_P, _D, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.c_int], C.c_int),
        "orc_mesh_write": ([_P, C.c_char_p, C.c_int, _D, C.c_int], C.c_int),
        "orc_project": ([_P, _P, _D, _D, C.c_double], None),
        "orc_ke": ([_P, _D], C.c_double),
        "orc_maxabs_div": ([_P, _D], C.c_double),
        "orc_dissipation": ([_P, _D, C.c_double], C.c_double),
        "orc_cfl": ([_P, _D, C.c_double, _D], None),
        "orc_mhd_diag": ([_P, _D, _D, C.c_double, C.c_double, _D], None),
        "orc_tdma": ([C.c_int, _D, _D, _D, _D, _D], None),
        "orc_num_threads": ([], C.c_int),
        "orc_cyclic_tridiag": ([C.c_int, C.c_double, C.c_double, C.c_double, _D, _D], None),
        "orc_compact_kmod": ([C.c_double, C.c_int], C.c_double),
        "orc_compact_deriv": ([_P, _D, C.c_int, C.c_int, C.c_int, _D], None),
        "orc_poisson_create_compact": ([_P], _P),
        "orc_compact_ns_rhs": ([_P, _D, C.c_double, _D], None),
        "orc_compact_project": ([_P, _P, _D, _D, C.c_double], None),
        "orc_compact_ns_step": ([_P, _P, _D, _D, _D, C.c_double, C.c_double], None),
        "orc_compact_maxabs_div": ([_P, _D], C.c_double),
        "orc_compact_mhd_rhs": ([_P, _D, _D, C.c_double, C.c_double, _D, _D], None),
        "orc_compact_mhd_step": ([_P, _P, _D, _D, _D, _D, C.c_double, C.c_double, C.c_double], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


class OMesh:
    def __init__(self, N, L, origin=(0.0, 0.0, 0.0), m=6, hyp=None, wall_y=False):
        n = (C.c_int * 3)(*N)
        l = (C.c_double * 3)(*L)
        o = (C.c_double * 3)(*origin)
        s1, s2 = hyp if hyp else (0.0, 0.0)
        self._h = lib().orc_mesh_create(n, l, o, m, 1 if hyp else 0, s1, s2, 1 if wall_y else 0)
        dims = (C.c_int * 3)()
        lib().orc_mesh_dims(self._h, dims)
        self.n = tuple(dims)
        self.N = tuple(N)
        self.L = tuple(L)
        self.m = m
        self.h = m // 2
        self.wall_y = wall_y
        self.stored = (N[0], N[1] + (1 if wall_y else 0), N[2])

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_mesh_free(self._h)
            self._h = None

    def field(self, nc=3):
        return np.zeros((*self.n, nc))

    def metric(self):
        a = np.zeros(self.n[2])
        b = np.zeros(self.n[2])
        lib().orc_mesh_metric(self._h, _d(a), _d(b))
        return a, b

    # layout conversions
    def interior(self, f: np.ndarray) -> np.ndarray:
        """padded AoS -> compact SoA (ncomp, s0, s1, s2)."""
        h = self.h
        s0, s1, s2 = self.stored
        return np.ascontiguousarray(np.moveaxis(f[h:h + s0, h:h + s1, h:h + s2, :], -1, 0))

    def from_compact(self, c: np.ndarray, odd_comp=1) -> np.ndarray:
        """compact SoA -> padded AoS with the standard ghost fill."""
        nc = c.shape[0]
        f = self.field(nc)
        h = self.h
        s0, s1, s2 = self.stored
        f[h:h + s0, h:h + s1, h:h + s2, :] = np.moveaxis(c, 0, -1)
        self.fill(f, odd_comp if nc == 3 else -1)
        return f

    def fill(self, f, odd_comp=1):
        lib().orc_fill(self._h, _d(f), f.shape[-1], odd_comp)

    def bc_apply(self, f, entries):
        n = len(entries)
        dirs = (C.c_int * max(n, 1))(*[e[0] for e in entries])
        kinds = (C.c_int * max(n, 1))(*[e[1] for e in entries])
        comps = (C.c_int * max(n, 1))(*[e[2] for e in entries])
        vals = (C.c_double * max(n, 1))(*[float(e[3]) for e in entries])
        lib().orc_bc_apply(self._h, _d(f), f.shape[-1], n, dirs, kinds, comps, vals)

    def div_adv(self, U, phi):
        out = self.field(phi.shape[-1])
        lib().orc_div_adv(self._h, _d(U), _d(phi), phi.shape[-1], _d(out))
        return out

    def div(self, U):
        out = self.field(1)
        lib().orc_div(self._h, _d(U), _d(out))
        return out

    def grad(self, p):
        out = self.field(3)
        lib().orc_grad(self._h, _d(p), _d(out))
        return out

    def laplacian(self, nu, f):
        out = self.field(f.shape[-1])
        lib().orc_laplacian(self._h, nu, _d(f), f.shape[-1], _d(out))
        return out

    def laplacian_var(self, nu, f):
        """ex::laplacian(Field<scalar> nu, f); nu padded with nc = 1."""
        out = self.field(f.shape[-1])
        lib().orc_laplacian_var(self._h, _d(nu), _d(f), f.shape[-1], _d(out))
        return out

    def grad_vec_literal(self, U):
        out = self.field(9)
        lib().orc_grad_vec_literal(self._h, _d(U), _d(out))
        return out

    def curl_literal(self, U):
        out = self.field(3)
        lib().orc_curl_literal(self._h, _d(U), _d(out))
        return out

    def curl_central(self, U):
        out = self.field(3)
        lib().orc_curl_central(self._h, _d(U), _d(out))
        return out

    def ke(self, U):
        return lib().orc_ke(self._h, _d(U))

    def maxabs_div(self, U):
        return lib().orc_maxabs_div(self._h, _d(U))

    def cfl(self, U, dt):
        out = np.zeros(6)
        lib().orc_cfl(self._h, _d(U), dt, _d(out))
        return out

    def dissipation(self, U, nu):
        return lib().orc_dissipation(self._h, _d(U), nu)

    def mhd_rhs(self, U, B, nu, eta):
        """(dU, dB) per unit dt of the MHD step (UEqn.H:16-23, BEqn.H:15-23)."""
        dU, dB = self.field(3), self.field(3)
        lib().orc_mhd_rhs(self._h, _d(U), _d(B), nu, eta, _d(dU), _d(dB))
        return dU, dB

    def z(self, k):
        """Mesh::z(k) at padded index k."""
        return lib().orc_mesh_z(self._h, int(k))

    def mhd_diag(self, U, B, mu, eta):
        out = np.zeros(8)
        lib().orc_mhd_diag(self._h, _d(U), _d(B), mu, eta, _d(out))
        return out

    # weakly-compressible path
    def wc_step(self, U, rho, p, mu, pc, dt, B=None, dbu=0.0):
        """One weakly-compressible step in place (U: nc=3, rho and p: nc=1);
        B (nc=3, frozen) adds the Lorentz terms of app/OT/UEqn.H."""
        lib().orc_wc_step(self._h, _d(U), _d(rho), _d(p), _d(B) if B is not None else None, mu, pc, dbu, dt)

    def wcs_step(self, U, rho, p, T, mu, ka, g, pc, dt):
        """One app/shearLayerWC step in place (U nc=3; rho, p, T nc=1)."""
        gv = (C.c_double * 3)(*g)
        lib().orc_wcs_step(self._h, _d(U), _d(rho), _d(p), _d(T), mu, ka, gv, pc, dt)

    def wc_hydrostatic(self, T, p, rho, g, rho0, pc):
        """Initial hydrostatic pressure (added to p) and rho = (p/pc)^0.5 of
        app/shearLayerWC, in place."""
        gv = (C.c_double * 3)(*g)
        lib().orc_wc_hydrostatic(self._h, _d(T), _d(p), _d(rho), gv, rho0, pc)

    # implicit matrix im::fdMatrix<scalar> / im::laplacian
    def fd_laplacian(self, mu, phi):
        """-> dict(al (6, *n), ap, rhs) for padded scalar phi (shape (*n, 1))."""
        al = np.zeros((6, *self.n))
        ap = np.zeros(self.n)
        rhs = np.zeros(self.n)
        lib().orc_fd_laplacian(self._h, mu, _d(np.ascontiguousarray(phi)), _d(al), _d(ap), _d(rhs))
        return {"al": al, "ap": ap, "rhs": rhs}

    def fd_residual(self, A, phi):
        return lib().orc_fd_residual(self._h, _d(A["al"]), _d(A["ap"]), _d(A["rhs"]), _d(phi))

    def fd_solve(self, A, phi, tol, max_its):
        """Solve in place on padded phi; returns (iterations, last residual)."""
        r = np.zeros(1)
        its = lib().orc_fd_solve(self._h, _d(A["al"]), _d(A["ap"]), _d(A["rhs"]), _d(phi), tol, max_its, _d(r))
        return its, r[0]

    def fd_relax(self, A, phi, urf):
        lib().orc_fd_relax(self._h, _d(A["ap"]), _d(A["rhs"]), _d(phi), urf)

    # EnSight Gold output
    def field_write(self, f, path, name, file_num, proc=0):
        """Field::write of padded field f (shape (*n, nc))."""
        rc = lib().orc_field_write(self._h, _d(f), f.shape[-1], str(path).encode(), name.encode(), proc, file_num)
        assert rc == 0, "orc_field_write"

    def field_read(self, f, path, name, file_num, proc=0):
        rc = lib().orc_field_read(self._h, _d(f), f.shape[-1], str(path).encode(), name.encode(), proc, file_num)
        assert rc == 0, "orc_field_read"

    def mesh_write(self, path, steps, proc=0):
        st = np.ascontiguousarray(steps, dtype=np.float64)
        if st.size == 0:
            st = np.zeros(1)
            n = 0
        else:
            n = len(steps)
        rc = lib().orc_mesh_write(self._h, str(path).encode(), proc, _d(st), n)
        assert rc == 0, "orc_mesh_write"

    # compact 6th-order scheme

    def compact_deriv(self, f, direction, order=1):
        out = self.field(f.shape[-1])
        lib().orc_compact_deriv(self._h, _d(f), f.shape[-1], direction, order, _d(out))
        return out

    def compact_ns_rhs(self, U, nu):
        out = self.field(3)
        lib().orc_compact_ns_rhs(self._h, _d(U), nu, _d(out))
        return out

    def compact_mhd_rhs(self, U, B, nu, eta):
        dU, dB = self.field(3), self.field(3)
        lib().orc_compact_mhd_rhs(self._h, _d(U), _d(B), nu, eta, _d(dU), _d(dB))
        return dU, dB

    def compact_maxabs_div(self, U):
        return lib().orc_compact_maxabs_div(self._h, _d(U))


class OPoisson:
    def __init__(self, mesh: OMesh, compact: bool = False):
        self.mesh = mesh
        self.compact = compact
        self._h = (lib().orc_poisson_create_compact if compact else lib().orc_poisson_create)(mesh._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_poisson_free(self._h)
            self._h = None

    def k2(self):
        N = self.mesh.N
        M1 = 2 * N[1] if self.mesh.wall_y else N[1]
        out = np.zeros(N[0] * M1 * N[2])
        lib().orc_poisson_k2(self._h, _d(out))
        return out.reshape(N[2], M1, N[0])  # [k][j][i]

    def solve(self, rhs, p=None):
        if p is None:
            p = self.mesh.field(1)
        lib().orc_poisson_solve(self._h, _d(rhs), _d(p))
        return p

    def project(self, F, s, p=None):
        if p is None:
            p = self.mesh.field(1)
        fn = lib().orc_compact_project if self.compact else lib().orc_project
        fn(self.mesh._h, self._h, _d(F), _d(p), s)
        return p

    def compact_mhd_step(self, U, B, p, pB, nu, eta, dt):
        lib().orc_compact_mhd_step(self.mesh._h, self._h, _d(U), _d(B), _d(p), _d(pB), nu, eta, dt)

    def compact_ns_step(self, U, p, nu, dt, Ustar=None):
        if Ustar is None:
            Ustar = self.mesh.field(3)
        lib().orc_compact_ns_step(self.mesh._h, self._h, _d(U), _d(Ustar), _d(p), nu, dt)

    def ns_step(self, U, p, nu, dt, mode, Ustar=None):
        if Ustar is None:
            Ustar = self.mesh.field(3)
        lib().orc_ns_step(self.mesh._h, self._h, _d(U), _d(Ustar), _d(p), nu, dt, mode)

    def ns_step_fused(self, U, p, nu, dt, W1=None, W2=None):
        """SSP-RK3 step with one fused pass per stage (CPU baseline variant)."""
        W1 = self.mesh.field(3) if W1 is None else W1
        W2 = self.mesh.field(3) if W2 is None else W2
        lib().orc_ns_step_fused(self.mesh._h, self._h, _d(U), _d(W1), _d(W2), _d(p), nu, dt)

    def mhd_step(self, U, B, p, pB, nu, eta, dt, mode):
        lib().orc_mhd_step(self.mesh._h, self._h, _d(U), _d(B), _d(p), _d(pB), nu, eta, dt, mode)


BC_NAMES = {"north": NORTH, "south": SOUTH, "east": EAST, "west": WEST, "top": TOP, "bottom": BOTTOM}
BC_KINDS = {"periodic": BC_PERIODIC, "neumann": BC_NEUMANN, "dirichlet": BC_DIRICHLET}


def bc_entries(text: str, ncomp: int):
    """The "dir type val comp" lines of a BC file (src/Field/Field.hxx:519-547)
    as bc_apply entries in the order Field::correctBoundaryConditions
    applies them (Field.hxx:310-322: bc_[dir + 6 comp], index order)."""
    table = {}
    for line in text.splitlines():
        f = line.split()
        if len(f) < 4:
            continue
        d, kind, val, comp = BC_NAMES[f[0]], BC_KINDS[f[1]], float(f[2]), int(f[3])
        table[d + 6 * comp] = (d, kind, comp, val)
    return [table[i] for i in sorted(table) if table[i][2] < ncomp]


SHEAR_RK4 = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)  # app/shearLayer/shearLayer.cpp:145 (RK4)
RK4C = (0.5, 0.5, 1.0, 1.0)                               # src/settings/settings.cpp:51 (RK4c)


def shear_layer_step(M: "OMesh", P: "OPoisson", U, B, p, mu, dt, ubc, pbc):
    """One time step of app/shearLayer (shearLayer.cpp:159-185; TEqn and
    BEqn switched off, lines 178-179) with UEqn.H:1-54 as written, on the
    padded fields of the walled mesh M, in place: per stage rk
        dU = dt (((-div(U, U) + lap(mu, U)) + div(B, B)) - grad(0.5 (B & B)))
        U* = UOld + RK4[rk] dU;  U*.correctBoundaryConditions()
        p = solve(div(U*) / dt / RK4c[rk]);  p.correctBoundaryConditions()
        U = U* - (RK4c[rk] dt) grad(p);  U.correctBoundaryConditions()
    with the literal fills of the BC entries ubc, pbc (bc_entries) and the
    reference's FFT solve (poisson.cpp:167-265), which takes nodes
    0 .. N - 1 of every direction: P is an OPoisson on the periodic mesh of
    the same cells (the padded layouts agree), p's wall node is left to
    its fill.  B keeps its values (BEqn off)."""
    UOld = U.copy()
    BB = (B[..., 0] * B[..., 0] + B[..., 1] * B[..., 1]) + B[..., 2] * B[..., 2]  # vector.cpp:120-127
    half = np.ascontiguousarray((0.5 * BB)[..., None])
    for rk in range(4):
        dU = dt * ((((-M.div_adv(U, U)) + M.laplacian(mu, U)) + M.div_adv(B, B)) - M.grad(half))
        Ustar = UOld + SHEAR_RK4[rk] * dU
        M.bc_apply(Ustar, ubc)
        rhs = np.ascontiguousarray(M.div(Ustar) / dt / RK4C[rk])
        P.solve(rhs, p)
        M.bc_apply(p, pbc)
        U[...] = Ustar - (RK4C[rk] * dt) * M.grad(p)
        M.bc_apply(U, ubc)


def coef(m):
    out = np.zeros(4)
    lib().orc_coef(m, _d(out))
    return out


def fft1d(x: np.ndarray, sign=-1) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft1d(a.size, a.view(np.float64).ctypes.data_as(_D), sign)
    return a


def fft3d(x: np.ndarray, sign=-1) -> np.ndarray:
    """x has numpy shape (n2, n1, n0) (i fastest)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = (C.c_int * 3)(a.shape[2], a.shape[1], a.shape[0])
    lib().orc_fft3d(n, a.view(np.float64).ctypes.data_as(_D), sign)
    return a


def tdma(a, b, c, d):
    n = len(b)
    x = np.zeros(n)
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d)]
    lib().orc_tdma(n, *[_d(v) for v in arrs], _d(x))
    return x


def cyclic_tridiag(a, b, c, d):
    d = np.ascontiguousarray(d, dtype=np.float64)
    x = np.zeros_like(d)
    lib().orc_cyclic_tridiag(len(d), a, b, c, _d(d), _d(x))
    return x


def compact_kmod(kh, order=1):
    return lib().orc_compact_kmod(float(kh), int(order))


def num_threads() -> int:
    return lib().orc_num_threads()


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


class OTime:
    """The run clock of src/Time/Time.cpp, restated: Time(dt, end, writeEvery)
    (:23-38), with a float the interval form (:64-74), from_file (:77-109),
    restart (:40-61) and advance = operator++ (:112-139)."""

    def __init__(self, dt, end, write_every):
        self.steps = []
        self.dt, self.cur, self.end = dt, 0.0, end
        self.time_step, self.file_num = 0, 0
        self.write, self.log = False, True
        self.timebased = isinstance(write_every, float)
        self.write_every = 0 if self.timebased else write_every
        self.interval = write_every if self.timebased else 0.0

    @classmethod
    def from_file(cls, path):
        t = cls(0.0, 0.0, 0)
        for line in open(path):
            w = line.split()
            if not w:
                continue
            if w[0] == "dt:":
                t.dt = float(w[1])
            elif w[0] == "writeT:":
                t.timebased, t.interval = True, float(w[1])
            elif w[0] == "writeN:":
                t.timebased, t.write_every = False, int(w[1])
            elif w[0] == "end:":
                t.end = float(w[1])
        return t

    def restart(self, path):
        w = open(path).read().split()
        n = int(w[0])
        vals = [float(x) for x in w[1:1 + n]]
        self.steps += vals
        if n:
            self.cur = vals[-1]
        self.file_num = len(self.steps)

    def advance(self):
        if self.timebased:
            self.write = (self.cur + self.dt) >= (self.file_num + 1) * self.interval
        else:
            self.write = self.time_step % self.write_every == 0
        if self.write:
            self.file_num += 1
            self.steps.append(self.cur)
        self.log = self.time_step % 100 == 0
        self.cur += self.dt
        self.time_step += 1

    def run(self):
        return self.cur < self.end
