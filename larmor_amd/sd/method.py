"""Spherical deconvolution methods and their defaults (src/larmor/sd/method.py)."""

from enum import Enum

from larmor_amd.sd.rf import IsotropicRF, ResponseFunction
from larmor_amd.sd.solvers import DRL, FRACTION_TOL, L2NNLS


class SDMethod(str, Enum):
    """GRL (dRL solver, inner shells weighted 0.2) and mFOD (L2NNLS, equal shell weights)."""

    GRL = "GRL"
    MFOD = "mFOD"

    @classmethod
    def _missing_(cls, value):
        if isinstance(value, str):
            for member in cls:
                if member.value.lower() == value.lower():
                    return member
        valid = ", ".join(m.value for m in cls)
        raise ValueError(f"Unknown method {value!r}; valid methods are {valid}.")

    def __str__(self) -> str:
        return self.value

    def default_solver(self):
        return DRL() if self is SDMethod.GRL else L2NNLS()

    @property
    def inner_shell_weight(self) -> float:
        return 0.2 if self is SDMethod.GRL else 1.0

    @property
    def _fraction_tolerance(self) -> float:
        return FRACTION_TOL if self is SDMethod.GRL else 1e-7

    def _check_rfs(self, rfs: list[ResponseFunction]) -> None:
        n_aniso = sum(1 for rf in rfs if not isinstance(rf, IsotropicRF))
        if self is SDMethod.GRL and n_aniso != 1:
            raise ValueError(f"{self} takes exactly one anisotropic response function, got {n_aniso}.")
        if self is SDMethod.MFOD and n_aniso < 1:
            raise ValueError(f"{self} takes at least one anisotropic response function.")
