"""Capturing-method hook types and the method registry (FRForge `src/capturing/Interfaces.jl`,
`src/capturing/PerssonAV.jl`, `src/methods/Registry.jl`, `src/methods/ScaledPersson.jl`).

The device path implements the hook set of the shipped methods: NullCapturing (identity hooks)
and Persson modal sensor + element artificial viscosity in its two `av_form`s.  A method says
what it runs on the device through ``device_capturing()``, which returns
``(kind, params)`` with kind one of "null", "persson_br0", "persson_dd".  A method whose hooks
cannot be expressed that way raises NotImplementedError from ``device_capturing``: there is no
CPU fallback in this package.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Dict


class AbstractCapturingMethod:
    """Base of all capturing methods (Interfaces.jl:14)."""

    def device_capturing(self):
        raise NotImplementedError(
            f"{type(self).__name__} defines hooks the device path does not implement")


class NullCapturing(AbstractCapturingMethod):
    """Identity hooks: no sensor, no artificial viscosity (Interfaces.jl:21-26)."""

    def device_capturing(self):
        return "null", {}

    def __repr__(self):
        return "NullCapturing()"


@dataclass(frozen=True)
class PerssonSensor:
    """Persson-Peraire modal sensor parameters (PerssonAV.jl:9-22)."""

    kappa: float = 4.0
    s0_factor: float = -4.0
    eps_floor: float = 1e-16


@dataclass(frozen=True)
class ElementArtificialViscosity:
    """eps_e = c_av sigma_e (h_e / p) lambda_max,e; av_form "conservative_br0" (default) or
    "element_local_DD" (PerssonAV.jl:24-46)."""

    c_av: float = 0.1
    av_form: str = "conservative_br0"

    def __post_init__(self):
        if self.av_form not in ("conservative_br0", "element_local_DD"):
            raise ValueError(f"Unsupported av_form={self.av_form} for 2D")


@dataclass(frozen=True)
class PerssonAVMethod(AbstractCapturingMethod):
    """Persson sensor + element AV (PerssonAV.jl:48-69)."""

    sensor: PerssonSensor = PerssonSensor()
    dissip: ElementArtificialViscosity = ElementArtificialViscosity()

    @classmethod
    def create(cls, kappa=4.0, s0_factor=-4.0, c_av=0.1, av_form="conservative_br0"):
        return cls(PerssonSensor(kappa, s0_factor), ElementArtificialViscosity(c_av, av_form))

    def device_capturing(self):
        kind = "persson_br0" if self.dissip.av_form == "conservative_br0" else "persson_dd"
        return kind, {"kappa": self.sensor.kappa, "s0_factor": self.sensor.s0_factor,
                      "eps_floor": self.sensor.eps_floor, "c_av": self.dissip.c_av}

    def method_params(self):
        return {"κ": self.sensor.kappa, "s0_factor": self.sensor.s0_factor,
                "c_av": self.dissip.c_av, "av_form": self.dissip.av_form, "method": "persson_av"}


@dataclass(frozen=True)
class ScaledPerssonMethod(AbstractCapturingMethod):
    """Persson AV with c_av = 0.5 by default, forwarding every hook to the inner method
    (ScaledPersson.jl:15-62)."""

    inner: PerssonAVMethod = PerssonAVMethod.create(c_av=0.5)

    @classmethod
    def create(cls, kappa=4.0, s0_factor=-4.0, c_av=0.5, av_form="conservative_br0"):
        return cls(PerssonAVMethod.create(kappa, s0_factor, c_av, av_form))

    def device_capturing(self):
        return self.inner.device_capturing()

    def method_params(self):
        p = self.inner.method_params()
        p["method"] = "scaled_persson"
        p["note"] = "example inventable method; elevated c_av over baseline persson_av"
        return p


def method_params(m: AbstractCapturingMethod) -> dict:
    """(PerssonAV.jl:79-92)"""
    if isinstance(m, NullCapturing):
        return {"method": "null"}
    f = getattr(m, "method_params", None)
    return f() if f else {}


METHOD_REGISTRY: Dict[str, Callable[..., AbstractCapturingMethod]] = {
    "null": lambda **kw: NullCapturing(),
    "persson_av": lambda **kw: PerssonAVMethod.create(**kw),
    "scaled_persson": lambda **kw: ScaledPerssonMethod.create(**kw),
}


def register_method(name: str, factory: Callable[..., AbstractCapturingMethod]) -> None:
    """register_method! (Interfaces.jl:123-126)."""
    METHOD_REGISTRY[str(name)] = factory


def get_capturing_method(name: str, **kwargs) -> AbstractCapturingMethod:
    """get_capturing_method (Interfaces.jl:128-135); unknown names raise KeyError listing the
    known ones, as the reference's error does."""
    key = str(name)
    if key not in METHOD_REGISTRY:
        known = ", ".join(sorted(METHOD_REGISTRY))
        raise KeyError(f'Unknown capturing method "{key}". Known: {known}')
    return METHOD_REGISTRY[key](**kwargs)


def list_methods():
    """(Registry.jl:14-17)"""
    return sorted(METHOD_REGISTRY)
