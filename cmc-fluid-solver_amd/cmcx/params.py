"""Flow parameters: the reference's ``FluidParams`` (src/Common/Geometry.h:538-562).

Both constructors compute in double; the solver rounds the four fields to its
field type (FTYPE float in FP32 runs), as the reference stores them as FTYPE.
"""
from __future__ import annotations

import dataclasses


@dataclasses.dataclass
class FluidParams:
    v_T: float
    v_vis: float
    t_vis: float
    t_phi: float

    @classmethod
    def normalized(cls, Re: float, Pr: float, lam: float) -> "FluidParams":
        """FluidParams(Re, Pr, lambda) (Geometry.h:545-552)."""
        return cls(1.0, 1.0 / Re, 1.0 / (Re * Pr), (lam - 1) / (lam * Re))

    @classmethod
    def dimensional(cls, vis: float, rho: float, R: float, k: float, cv: float) -> "FluidParams":
        """FluidParams(vis, rho, R, k, cv) (Geometry.h:554-561)."""
        return cls(R, vis / rho, k / (rho * cv), vis / (rho * cv))
