"""ctypes wrapper of the CPU oracle (libfrf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py, always as the checker.  The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfrf_oracle.so")

BC_PERIODIC, BC_TRANSMISSIVE, BC_DIRICHLET = 0, 1, 3
FLUX = {"rusanov": 0, "hllc": 1}
CAPT = {"null": 0, "persson_br0": 1, "persson_dd": 2}
STATUS = {0: "ok", 1: "diverged", 2: "max_steps"}


class OrcProblem(C.Structure):
    _fields_ = [("p", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
                ("bc_x", C.c_int32), ("bc_y", C.c_int32), ("flux", C.c_int32),
                ("capturing", C.c_int32), ("gamma", C.c_double), ("kappa", C.c_double),
                ("s0_factor", C.c_double), ("eps_floor", C.c_double), ("c_av", C.c_double)]


class Orc1Problem(C.Structure):
    _fields_ = [("p", C.c_int32), ("nel", C.c_int32), ("bc_left", C.c_int32),
                ("bc_right", C.c_int32), ("capturing", C.c_int32), ("reserved", C.c_int32),
                ("gamma", C.c_double), ("kappa", C.c_double), ("s0_factor", C.c_double),
                ("eps_floor", C.c_double), ("c_av", C.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_build_operators.argtypes = [C.c_int] + [dp] * 10
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcProblem)] + [dp] * 10
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_metric.restype = dp
        L.orc_metric.argtypes = [C.c_void_p, C.c_int]
        L.orc_residual.argtypes = [C.c_void_p, dp, C.c_double, dp]
        L.orc_set_solid.restype = C.c_int
        L.orc_set_solid.argtypes = [C.c_void_p, C.POINTER(C.c_uint8)]
        L.orc1_set_boundary_state.restype = C.c_int
        L.orc1_set_boundary_state.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_set_boundary.restype = C.c_int
        L.orc_set_boundary.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.orc_set_metrics.restype = C.c_int
        L.orc_set_metrics.argtypes = [C.c_void_p, C.POINTER(dp)]
        L.orc_set_boundary_split.restype = C.c_int
        L.orc_set_boundary_split.argtypes = [C.c_void_p, C.c_int, dp, dp, C.c_int, dp]
        L.orc_sense.argtypes = [C.c_void_p, dp, dp]
        L.orc_element_viscosities.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_apply_br0_eps.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_apply_dd_eps.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_max_wave_speed.restype = C.c_double
        L.orc_max_wave_speed.argtypes = [C.c_void_p, dp]
        L.orc_compute_dt.restype = C.c_double
        L.orc_compute_dt.argtypes = [C.c_void_p, dp, C.c_double]
        L.orc_rk3_step.restype = C.c_int
        L.orc_rk3_step.argtypes = [C.c_void_p, dp, dp, C.c_double]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_double, C.c_double,
                              C.c_int64, C.POINTER(C.c_int64)]
        L.orc_flux_n.argtypes = [C.c_int, C.c_double, dp, dp, C.c_double, C.c_double, dp]
        L.orc1_create.restype = C.c_void_p
        L.orc1_create.argtypes = [C.POINTER(Orc1Problem)] + [dp] * 9
        L.orc1_destroy.argtypes = [C.c_void_p]
        L.orc1_residual.argtypes = [C.c_void_p, dp, C.c_double, dp]
        L.orc1_sense.argtypes = [C.c_void_p, dp, dp]
        L.orc1_element_viscosities.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc1_compute_dt.restype = C.c_double
        L.orc1_compute_dt.argtypes = [C.c_void_p, dp, C.c_double]
        L.orc1_rk3_step.restype = C.c_int
        L.orc1_rk3_step.argtypes = [C.c_void_p, dp, dp, C.c_double]
        L.orc1_run.restype = C.c_int
        L.orc1_run.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_double, C.c_double,
                               C.c_int64, C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def build_operators(p: int) -> dict:
    """The oracle's own operators (Newton Gauss-Legendre); D and V indexed [i, j]."""
    n = p + 1
    arrs = {k: np.zeros(n) for k in ("xi", "w", "lL", "lR", "gLxi", "gRxi", "gL", "gR")}
    D = np.zeros(n * n)
    V = np.zeros(n * n)
    lib().orc_build_operators(p, _p(arrs["xi"]), _p(arrs["w"]), _p(D), _p(arrs["lL"]),
                              _p(arrs["lR"]), _p(arrs["gLxi"]), _p(arrs["gRxi"]),
                              _p(arrs["gL"]), _p(arrs["gR"]), _p(V))
    arrs["D"] = D.reshape(n, n, order="F")
    arrs["V"] = V.reshape(n, n, order="F")
    return arrs


def flux_n(kind: str, gamma: float, uL, uR, nx: float, ny: float) -> np.ndarray:
    uL = np.ascontiguousarray(uL, dtype=np.float64)
    uR = np.ascontiguousarray(uR, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_flux_n(FLUX[kind], gamma, _p(uL), _p(uR), nx, ny, _p(out))
    return out


class Oracle:
    """One oracle problem: operators (xi, w, D, lL, lR, gLxi, gRxi, V as numpy arrays with
    matrices indexed [i, j]) and the mesh vertices are inputs."""

    BC_KIND = {"transmissive": 1, "reflecting": 2, "dirichlet": 3}
    SIDE = {"left": 0, "right": 1, "bottom": 2, "top": 3}

    def __init__(self, p, nx, ny, xv, yv, ops, bc_x="periodic", bc_y="periodic",
                 flux="rusanov", capturing="null", gamma=1.4, kappa=4.0, s0_factor=-4.0,
                 eps_floor=1e-16, c_av=0.1, sides=None, solid=None, metrics=None):
        """sides: optional {"left"|"right"|"bottom"|"top": "transmissive" | "reflecting" |
        ("dirichlet", U[4]) | ("split", (a, b, c, d), U_lo, "reflecting" | "dirichlet", U_hi)}
        for the non-periodic axes (split: U_lo where s < a + (b + c t) / d along the side)."""
        pb = OrcProblem(p, nx, ny, BC_PERIODIC if bc_x == "periodic" else BC_TRANSMISSIVE,
                        BC_PERIODIC if bc_y == "periodic" else BC_TRANSMISSIVE, FLUX[flux],
                        CAPT[capturing], gamma, kappa, s0_factor, eps_floor, c_av)
        self.p, self.nx, self.ny, self.Np, self.Nel = p, nx, ny, p + 1, nx * ny
        self._keep = {k: np.asfortranarray(np.asarray(ops[k], dtype=np.float64))
                      for k in ("xi", "w", "D", "lL", "lR", "gLxi", "gRxi", "V")}
        self._xv = np.ascontiguousarray(xv, dtype=np.float64)
        self._yv = np.ascontiguousarray(yv, dtype=np.float64)
        k = self._keep
        self.h = lib().orc_create(C.byref(pb), _p(k["xi"]), _p(k["w"]), _p(k["D"]), _p(k["lL"]),
                                  _p(k["lR"]), _p(k["gLxi"]), _p(k["gRxi"]), _p(k["V"]),
                                  _p(self._xv), _p(self._yv))
        if not self.h:
            raise ValueError("orc_create failed")
        self._bc_states = []
        for side, spec in (sides or {}).items():
            if not isinstance(spec, str) and spec[0] == "split":
                _, coef, U_lo, kind_hi, U_hi = spec
                arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (coef, U_lo)]
                hi = None if U_hi is None else np.ascontiguousarray(U_hi, dtype=np.float64)
                self._bc_states += arrs + [hi]
                if lib().orc_set_boundary_split(self.h, self.SIDE[side], _p(arrs[0]), _p(arrs[1]),
                                                self.BC_KIND[kind_hi],
                                                None if hi is None else _p(hi)) != 0:
                    raise ValueError(f"orc_set_boundary_split({side}) failed")
                continue
            kind, U = (spec, None) if isinstance(spec, str) else spec
            Ua = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
            self._bc_states.append(Ua)
            if lib().orc_set_boundary(self.h, self.SIDE[side], self.BC_KIND[kind],
                                      None if Ua is None else _p(Ua)) != 0:
                raise ValueError(f"orc_set_boundary({side}, {kind}) failed")
        self._metrics = None
        if metrics is not None:   # MeshMetrics2D-like: the 20 arrays in the reference's field order
            names = ("x", "y", "x_xi", "x_eta", "y_xi", "y_eta", "J", "nx_W", "ny_W", "sJ_W",
                     "nx_E", "ny_E", "sJ_E", "nx_S", "ny_S", "sJ_S", "nx_N", "ny_N", "sJ_N", "h_char")
            self._metrics = [np.asfortranarray(getattr(metrics, n), dtype=np.float64) for n in names]
            arr = (C.POINTER(C.c_double) * 20)(*[_p(a) for a in self._metrics])
            if lib().orc_set_metrics(self.h, arr) != 0:
                raise ValueError("orc_set_metrics failed (split sides take affine coordinates)")
        self._solid = None
        if solid is not None:
            self._solid = np.ascontiguousarray(np.asarray(solid, dtype=bool).astype(np.uint8))
            if self._solid.size != self.Nel or lib().orc_set_solid(
                    self.h, self._solid.ctypes.data_as(C.POINTER(C.c_uint8))) != 0:
                raise ValueError("orc_set_solid failed (mask length, or a periodic axis)")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        return (self.Np, self.Np, self.Nel, 4)

    def _state(self, u):
        u = np.asarray(u, dtype=np.float64)
        if u.shape != self.shape:
            raise ValueError(f"state shape {u.shape} != {self.shape}")
        return np.asfortranarray(u)

    def metric(self, which: int) -> np.ndarray:
        n = self.Np * self.Np * self.Nel if which < 7 else (self.Np * self.Nel if which < 19 else self.Nel)
        ptr = lib().orc_metric(self.h, which)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def residual(self, u, t=0.0):
        u = self._state(u)
        du = np.zeros(self.shape, order="F")
        lib().orc_residual(self.h, _p(u), t, _p(du))
        return du

    def sense(self, u):
        u = self._state(u)
        s = np.zeros(self.Nel)
        lib().orc_sense(self.h, _p(u), _p(s))
        return s

    def element_viscosities(self, u, sigma):
        u = self._state(u)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        eps = np.zeros(self.Nel)
        lib().orc_element_viscosities(self.h, _p(u), _p(sigma), _p(eps))
        return eps

    def apply_br0(self, u, eps, du=None):
        u = self._state(u)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        du = np.zeros(self.shape, order="F") if du is None else np.asfortranarray(du).copy(order="F")
        lib().orc_apply_br0_eps(self.h, _p(u), _p(eps), _p(du))
        return du

    def apply_dd(self, u, eps, du=None):
        u = self._state(u)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        du = np.zeros(self.shape, order="F") if du is None else np.asfortranarray(du).copy(order="F")
        lib().orc_apply_dd_eps(self.h, _p(u), _p(eps), _p(du))
        return du

    def max_wave_speed(self, u):
        return lib().orc_max_wave_speed(self.h, _p(self._state(u)))

    def compute_dt(self, u, cfl):
        return lib().orc_compute_dt(self.h, _p(self._state(u)), cfl)

    def rk3_step(self, u, t, dt):
        """Returns (status, u_new, t_new); ``u`` is not modified."""
        u = self._state(u).copy(order="F")
        tt = C.c_double(t)
        st = lib().orc_rk3_step(self.h, _p(u), C.byref(tt), dt)
        return STATUS[st], u, tt.value

    def run(self, u, t, t_final, cfl=0.2, dt=None, max_steps=10**7):
        """ssp_rk3! loop; returns (status, n_steps, t, u_new)."""
        u = self._state(u).copy(order="F")
        tt = C.c_double(t)
        n = C.c_int64(0)
        st = lib().orc_run(self.h, _p(u), C.byref(tt), t_final, cfl, -1.0 if dt is None else dt,
                           max_steps, C.byref(n))
        return STATUS[st], n.value, tt.value, u


class Oracle1D:
    """The 1D Euler restatement (Sod tube configuration): u[j, e, c] with shape (Np, Nel, 3)."""

    def __init__(self, p, nel, xv, ops, bc="transmissive", capturing="null", gamma=1.4, kappa=4.0,
                 s0_factor=-4.0, eps_floor=1e-16, c_av=0.1, dirichlet=None):
        """dirichlet: optional (U_left, U_right) constant ghost states (None for a transmissive
        end) on a non-periodic mesh."""
        b = BC_PERIODIC if bc == "periodic" else BC_TRANSMISSIVE
        ends = [b, b]
        for k, U in enumerate(dirichlet or (None, None)):
            if U is not None:
                ends[k] = BC_DIRICHLET
        pb = Orc1Problem(p, nel, ends[0], ends[1], CAPT[capturing], 0, gamma, kappa, s0_factor, eps_floor, c_av)
        self.p, self.nel, self.Np = p, nel, p + 1
        self._keep = {k: np.asfortranarray(np.asarray(ops[k], dtype=np.float64))
                      for k in ("xi", "w", "D", "lL", "lR", "gLxi", "gRxi", "V")}
        self._xv = np.ascontiguousarray(xv, dtype=np.float64)
        k = self._keep
        self.h = lib().orc1_create(C.byref(pb), _p(k["xi"]), _p(k["w"]), _p(k["D"]), _p(k["lL"]),
                                   _p(k["lR"]), _p(k["gLxi"]), _p(k["gRxi"]), _p(k["V"]), _p(self._xv))
        if not self.h:
            raise ValueError("orc1_create failed")
        self._bc_states = []
        for k, U in enumerate(dirichlet or (None, None)):
            if U is not None:
                Ua = np.ascontiguousarray(U, dtype=np.float64)
                self._bc_states.append(Ua)
                if lib().orc1_set_boundary_state(self.h, k, _p(Ua)) != 0:
                    raise ValueError("orc1_set_boundary_state failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc1_destroy(self.h)
            self.h = None

    def _state(self, u):
        u = np.asfortranarray(u, dtype=np.float64)
        assert u.shape == (self.Np, self.nel, 3), u.shape
        return u

    def residual(self, u, t=0.0):
        u = self._state(u)
        du = np.zeros_like(u, order="F")
        lib().orc1_residual(self.h, _p(u), t, _p(du))
        return du

    def sense(self, u):
        u = self._state(u)
        s = np.zeros(self.nel)
        lib().orc1_sense(self.h, _p(u), _p(s))
        return s

    def element_viscosities(self, u, sigma):
        u = self._state(u)
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        e = np.zeros(self.nel)
        lib().orc1_element_viscosities(self.h, _p(u), _p(sg), _p(e))
        return e

    def compute_dt(self, u, cfl):
        return lib().orc1_compute_dt(self.h, _p(self._state(u)), cfl)

    def rk3_step(self, u, t, dt):
        u = self._state(u).copy(order="F")
        tt = C.c_double(t)
        st = lib().orc1_rk3_step(self.h, _p(u), C.byref(tt), dt)
        return STATUS[st], u, tt.value

    def run(self, u, t, t_final, cfl=0.2, dt=None, max_steps=10**7):
        """ssp_rk3! 1D; returns (status, n_steps, t, u_new)."""
        u = self._state(u).copy(order="F")
        tt = C.c_double(t)
        n = C.c_int64(0)
        st = lib().orc1_run(self.h, _p(u), C.byref(tt), t_final, cfl, -1.0 if dt is None else dt,
                            max_steps, C.byref(n))
        return STATUS[st], n.value, tt.value, u
