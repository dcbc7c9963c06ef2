"""The hot path: residual, sensor, CFL step and SSP-RK3 stepping on the GPU.

Mirrors FRForge's Julia API for SolutionState2D (names drop Julia's `!`; the first argument
is still mutated in place and returned):

    residual(du, state, eq, method)            residual!        Residual2D.jl:131-416
    sense(sigma, method, u_work, state, eq)    sense!           PerssonAV2D.jl:20-59
    compute_dt(state, eq, cfl=0.2)             compute_dt       SSP_RK3.jl:126-133
    ssp_rk3_step(state, eq, method, dt)        ssp_rk3_step!    SSP_RK3.jl:135-164
    ssp_rk3(state, eq, method, t_final, ...)   ssp_rk3!         SSP_RK3.jl:166-211
    integrate(state, eq, method, t_final, ...) integrate!       Integrate.jl:9-43

Each call uploads ``state.u`` / ``state.t``, runs on the device through libfrforge_hip.so and
downloads what the reference would have mutated.  Every call goes to the GPU.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from .capturing import AbstractCapturingMethod, NullCapturing
from .device import DeviceBatch, DeviceContext
from .equations import Euler2D
from .mesh import DirichletBC, GhostStateBC, PeriodicBC, ReflectingBC, TransmissiveBC
from .state import SolutionState2D


def _bc_name(bc) -> str:
    if isinstance(bc, PeriodicBC):
        return "periodic"
    if isinstance(bc, TransmissiveBC):
        return "transmissive"
    if isinstance(bc, ReflectingBC):
        return "reflecting"
    if isinstance(bc, DirichletBC):
        if bc.state is None:
            raise NotImplementedError("the device path takes constant Dirichlet states "
                                      "(DirichletBC(state=U)), not a u_func")
        return "dirichlet"
    if isinstance(bc, GhostStateBC):
        if bc.coef is None:
            raise NotImplementedError("the device path takes split ghost states "
                                      "(GhostStateBC.split), not a ghost callable")
        return "ghost_split"
    raise NotImplementedError(f"{type(bc).__name__} is not supported by the device path")


def _sides(mesh):
    return {"left": mesh.left_bc, "right": mesh.right_bc, "bottom": mesh.bottom_bc, "top": mesh.top_bc}


def _device(state: SolutionState2D, eq: Euler2D, method: AbstractCapturingMethod) -> DeviceContext:
    """The state's device context, created (or re-created) for this problem on first use, like
    the reference's lazily (re)allocated residual workspace (ResidualWorkspace.jl:72-80)."""
    if not isinstance(eq, Euler2D):
        raise NotImplementedError("the device path implements Euler2D only")
    if state.u.shape[3] != 4:
        raise NotImplementedError("the device path implements Neq = 4 only")
    kind, prm = method.device_capturing()
    mesh = state.mesh
    key = (state.p, mesh.nx, mesh.ny, tuple(mesh.x_vertices), tuple(mesh.y_vertices),
           _bc_name(mesh.left_bc), _bc_name(mesh.right_bc), _bc_name(mesh.bottom_bc),
           _bc_name(mesh.top_bc), state.scheme.flux, state.precision, kind,
           tuple(sorted(prm.items())), float(eq.gamma), id(state.ops),
           getattr(method, "weights_path", None),
           tuple((k, b.state) for k, b in _sides(mesh).items() if isinstance(b, DirichletBC)),
           tuple((k, b.key()) for k, b in _sides(mesh).items() if isinstance(b, GhostStateBC)),
           None if mesh.solid is None else mesh.solid.tobytes(),
           id(state.metrics) if state.metrics.is_curved else None)
    dev = state.device
    if dev is not None and getattr(dev, "key", None) == key:
        return dev
    if dev is not None:
        dev.close()
    dev = DeviceContext(p=state.p, nx=mesh.nx, ny=mesh.ny, bc_left=_bc_name(mesh.left_bc),
                        bc_right=_bc_name(mesh.right_bc), bc_bottom=_bc_name(mesh.bottom_bc),
                        bc_top=_bc_name(mesh.top_bc), flux=state.scheme.flux,
                        precision=state.precision, capturing=kind, gamma=float(eq.gamma), **prm)
    ops = state.ops
    dev.set_operators(ops.xi, ops.w, ops.D, ops.ell_L, ops.ell_R, ops.gL_xi, ops.gR_xi, ops.V_legendre)
    dev.set_geometry(mesh.x_vertices, mesh.y_vertices)
    for side, bc in _sides(mesh).items():
        if isinstance(bc, DirichletBC):
            dev.set_boundary_state(side, bc.state)
        elif isinstance(bc, GhostStateBC):
            dev.set_boundary_split(side, bc.coef, bc.U_lo, bc.hi, bc.U_hi)
    if mesh.solid is not None:
        dev.set_solid_mask(mesh.solid)
    if state.metrics.is_curved:
        jx = np.arange(mesh.n_elements) % mesh.nx
        jy = np.arange(mesh.n_elements) // mesh.nx
        if not np.array_equal(state.metrics.h_char, np.minimum(mesh.dx[jx], mesh.dy[jy])):
            raise NotImplementedError("the device keeps the affine element size min(dx, dy); curved "
                                      "metrics with another h_char (isoparametric D-metrics) are not "
                                      "supported")
        dev.set_curved_metrics(state.metrics)
    if getattr(method, "weights_path", None) is not None:   # learned sensor (frforge_amd.learned)
        dev.set_learned_sensor(method.weights().pack())
    dev.key = key
    state.device = dev
    return dev


def _push(state: SolutionState2D, dev: DeviceContext) -> None:
    if not state.u.flags.f_contiguous:
        state.u = np.asfortranarray(state.u)
    dev.upload(state.u, float(state.t))


def _pull(state: SolutionState2D, dev: DeviceContext) -> None:
    state.t = dev.download(state.u)


def residual(du: np.ndarray, state: SolutionState2D, eq: Euler2D,
             method: Optional[AbstractCapturingMethod] = None) -> np.ndarray:
    """Overwrite ``du`` with the FR residual of ``state`` (Residual2D.jl:131-416)."""
    method = method if method is not None else NullCapturing()
    if du.shape != state.u.shape:
        raise ValueError(f"du shape {du.shape} != state shape {state.u.shape}")
    dev = _device(state, eq, method)
    _push(state, dev)
    out = du if (du.flags.f_contiguous and du.dtype == np.float64) else np.empty(state.u.shape, order="F")
    dev.residual(out)
    if out is not du:
        du[...] = out
    return du


def sense(sigma: np.ndarray, method: AbstractCapturingMethod, u_work: np.ndarray,
          state: SolutionState2D, eq: Euler2D) -> np.ndarray:
    """Persson sensor sigma per element for ``u_work`` (PerssonAV2D.jl:20-59); identity hooks
    give zeros (Interfaces.jl:79-82)."""
    if sigma.shape != (state.mesh.n_elements,):
        raise ValueError(f"σ length {sigma.shape} != Nel {state.mesh.n_elements}")
    kind, _ = method.device_capturing()
    if kind == "null":
        sigma[:] = 0.0
        return sigma
    dev = _device(state, eq, method)
    dev.upload(np.asfortranarray(u_work, dtype=np.float64), float(state.t))
    s, _ = dev.sense()
    sigma[:] = s
    return sigma


def element_viscosities(method: AbstractCapturingMethod, state: SolutionState2D, eq: Euler2D):
    """(sigma, eps) per element for the current state (PerssonAV2D.jl:89-109)."""
    dev = _device(state, eq, method)
    _push(state, dev)
    return dev.sense()


def compute_dt(state: SolutionState2D, eq: Euler2D, cfl: float = 0.2) -> float:
    """cfl * min(h_char) / ((2p+1) max|u|+c) (SSP_RK3.jl:126-133)."""
    dev = state.device if state.device is not None else _device(state, eq, NullCapturing())
    _push(state, dev)
    return dev.compute_dt(float(cfl))


def ssp_rk3_step(state: SolutionState2D, eq: Euler2D, method: AbstractCapturingMethod,
                 dt: float) -> str:
    """One Shu-Osher step; returns "ok" or "diverged" and leaves state.u / state.t where the
    reference leaves them (SSP_RK3.jl:135-164)."""
    dev = _device(state, eq, method)
    _push(state, dev)
    status = dev.rk3_step(float(dt))
    _pull(state, dev)
    return status


def ssp_rk3(state: SolutionState2D, eq: Euler2D, method: AbstractCapturingMethod, t_final: float,
            cfl: float = 0.2, dt: Optional[float] = None, max_steps: int = 10**7,
            progress_every: int = 0, progress_label: str = "") -> dict:
    """Integrate to ``t_final`` (SSP_RK3.jl:166-211); returns {"status", "n_steps", "t"}."""
    dev = _device(state, eq, method)
    _push(state, dev)
    r = dev.run(float(t_final), float(cfl), dt, int(max_steps))
    _pull(state, dev)
    return {"status": r["status"], "n_steps": r["n_steps"], "t": r["t"]}


def integrate(state: SolutionState2D, eq: Euler2D, method: Optional[AbstractCapturingMethod],
              t_final: float, cfl: float = 0.2, dt: Optional[float] = None,
              max_steps: int = 10**7, time: Optional[str] = None, progress_every: int = 0,
              progress_label: str = "") -> dict:
    """Scheme-aware dispatcher (Integrate.jl:9-43): SSP-RK3 on the device."""
    method = method if method is not None else NullCapturing()
    tsym = time if time is not None else state.scheme.time
    if tsym == "ssp_rk3":
        return ssp_rk3(state, eq, method, t_final, cfl=cfl, dt=dt, max_steps=max_steps)
    if tsym == "ssp_rk2":
        raise NotImplementedError("SSP-RK2 is not implemented on the device path")
    raise ValueError(f"Unknown time scheme {tsym} (use ssp_rk3 or ssp_rk2)")


def ssp_rk3_batch(states, eq: Euler2D, methods, t_final, cfl=0.2, dt=None, max_steps=10**7,
                  tune: bool = True):
    """Integrate independent states together, as a sweep of ``ssp_rk3`` calls would, in one
    device loop (frf_batch_run): instances sharing degree, flux and capturing method advance with
    one stage launch per group, replayed from a captured HIP graph.  FP32-mixed states only.

    ``methods``, ``t_final``, ``cfl``, ``dt`` and ``max_steps`` are one per state or a single value
    for all; ``dt=None`` (or a non-positive entry) selects the CFL step.  ``tune`` runs the tile
    autotuner first.  Returns ([{"status", "n_steps", "t"} per state], device milliseconds of the
    stepping loop)."""
    states = list(states)
    n = len(states)
    if n == 0:
        return [], 0.0
    methods = list(methods) if isinstance(methods, (list, tuple)) else [methods] * n
    if len(methods) != n:
        raise ValueError("one capturing method per state")
    if dt is None:
        fixed = np.full(n, -1.0)
    else:
        fixed = np.array([(-1.0 if d is None else float(d)) for d in np.broadcast_to(np.asarray(dt, dtype=object), (n,))])
    devs = []
    for st, m in zip(states, methods):
        if st.precision != "fp32_mixed":
            raise NotImplementedError("the batched form runs FP32-mixed states only")
        dev = _device(st, eq, m)
        _push(st, dev)
        devs.append(dev)
    batch = DeviceBatch(devs)
    try:
        if tune:
            batch.tune()
        res, ms = batch.run(t_final, cfl, fixed, max_steps)
    finally:
        batch.close()
    for st, dev in zip(states, devs):
        _pull(st, dev)
    return [{"status": r["status"], "n_steps": r["n_steps"], "t": r["t"]} for r in res], ms
