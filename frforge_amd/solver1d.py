"""The reference's 1D time-stepping API on the device (SSP_RK3.jl:8-118, Residual.jl:161-214):
residual, compute_dt, one SSP-RK3 step and the integration loop for Euler1D states."""
from __future__ import annotations

import numpy as np

from .capturing import AbstractCapturingMethod, NullCapturing
from .device import DeviceContext1D
from .oned import Euler1D, SolutionState1D


def _device1d(state: SolutionState1D, eq: Euler1D, method: AbstractCapturingMethod) -> DeviceContext1D:
    if not isinstance(eq, Euler1D):
        raise NotImplementedError("the 1D device path implements Euler1D only")
    if state.scheme.flux != "rusanov":
        raise NotImplementedError("the 1D device path implements the Rusanov flux")
    kind, prm = method.device_capturing()
    mesh = state.mesh
    ends = mesh.bc_names()
    states = tuple(b.state if ends[k] == "dirichlet" else None
                   for k, b in enumerate((mesh.left_bc, mesh.right_bc)))
    key = (state.p, mesh.n_elements, tuple(mesh.x_vertices), ends, states, kind,
           tuple(sorted(prm.items())), float(eq.gamma), id(state.ops))
    dev = state.device
    if dev is not None and getattr(dev, "key", None) == key:
        return dev
    if dev is not None:
        dev.close()
    dev = DeviceContext1D(p=state.p, nel=mesh.n_elements, bc=ends[0], bc_right=ends[1], capturing=kind,
                          gamma=float(eq.gamma), **prm)
    ops = state.ops
    dev.set_operators(ops.xi, ops.w, ops.D, ops.ell_L, ops.ell_R, ops.gL_xi, ops.gR_xi, ops.V_legendre)
    dev.set_geometry(mesh.x_vertices)
    for side, U in zip(("left", "right"), states):
        if U is not None:
            dev.set_boundary_state(side, U)
    dev.key = key
    state.device = dev
    return dev


def _push(state, dev):
    if not state.u.flags.f_contiguous:
        state.u = np.asfortranarray(state.u)
    dev.upload(state.u, state.t)


def residual_1d(du: np.ndarray, state: SolutionState1D, eq: Euler1D, method=None) -> np.ndarray:
    method = method if method is not None else NullCapturing()
    dev = _device1d(state, eq, method)
    _push(state, dev)
    out = du if (du.flags.f_contiguous and du.dtype == np.float64) else np.empty(state.u.shape, order="F")
    dev.residual(out)
    if out is not du:
        du[...] = out
    return du


def compute_dt_1d(state: SolutionState1D, eq: Euler1D, cfl: float = 0.2) -> float:
    dev = _device1d(state, eq, NullCapturing()) if state.device is None else state.device
    _push(state, dev)
    return dev.compute_dt(cfl)


def ssp_rk3_step_1d(state: SolutionState1D, eq: Euler1D, method, dt: float) -> str:
    dev = _device1d(state, eq, method)
    _push(state, dev)
    st = dev.rk3_step(dt)
    state.t = dev.download(state.u)
    return st


def ssp_rk3_1d(state: SolutionState1D, eq: Euler1D, method, t_final: float, cfl: float = 0.2,
               dt=None, max_steps: int = 10**7) -> dict:
    """ssp_rk3! for Euler1D; returns {"status", "n_steps", "t"}."""
    dev = _device1d(state, eq, method)
    _push(state, dev)
    r = dev.run(t_final, cfl, dt, max_steps)
    state.t = dev.download(state.u)
    return {"status": r["status"], "n_steps": r["n_steps"], "t": r["t"]}
