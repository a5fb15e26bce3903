"""Variable indices, floors and the simulation ``Config`` (config.py of the
reference, config.py:10-89).

The names, order and defaults are the reference's, so existing callers keep
working; one field is added at the end: ``precision`` ("fp64", the
reference's arithmetic, or "fp32").
"""
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List


class ConsVar(IntEnum):
    """Conserved variable order (config.py:13-16)."""

    RHO = 0; MX = 1; MY = 2; MZ = 3
    BX = 4; BY = 5; BZ = 6; EN = 7
    PSI = 8; RHOC = 9


class PrimVar(IntEnum):
    """Primitive variable order (config.py:19-22)."""

    RHO = 0; VX = 1; VY = 2; VZ = 3
    BX = 4; BY = 5; BZ = 6; PR = 7
    PSI = 8; CLR = 9


NVAR = 10

RHO, MX, MY, MZ, BX, BY, BZ, EN, PSI, RHOC = tuple(ConsVar)
iRHO, iVX, iVY, iVZ, iBX, iBY, iBZ, iPR, iPSI, iCLR = tuple(PrimVar)

FLOOR_RHO = 1e-12  # config.py:40
FLOOR_PR = 1e-12   # config.py:41


@dataclass
class Config:
    """Simulation configuration (config.py:47-89)."""

    nx: int = 400
    ny: int = 200
    x_min: float = 0.0
    x_max: float = 6.0
    y_min: float = 0.0
    y_max: float = 2.0
    t_end: float = 0.25
    cfl: float = 0.30
    max_steps: int = 200000
    gamma: float = 5.0 / 3.0
    mach: float = 10.0
    interface_x: float = 1.5
    perturbation_amp: float = 0.15
    perturbation_mode: int = 4
    density_ratio: float = 3.0
    interface_width: float = 2.0
    B_transverse: float = 0.0
    glm_ch: float = 0.0
    glm_alpha: float = 1.5
    powell_source: bool = True
    use_char_bc: bool = True
    bc_x_type: str = "auto"
    bc_y_type: str = "periodic"
    enable_smoothing: bool = False
    diag_interval: int = 20
    snapshot_times: List[float] = field(
        default_factory=lambda: [0.0, 0.05, 0.10, 0.15, 0.20, 0.25])
    # added: arithmetic of the solver context ("fp64" or "fp32")
    precision: str = "fp64"

    @property
    def dx(self) -> float:
        return (self.x_max - self.x_min) / self.nx

    @property
    def dy(self) -> float:
        return (self.y_max - self.y_min) / self.ny

    def get_bc_x(self) -> str:
        """config.py:86-89"""
        if self.bc_x_type != "auto":
            return self.bc_x_type
        return "characteristic" if self.use_char_bc else "extrapolation"
