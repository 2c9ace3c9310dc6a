"""cmcx: MI355X-native ADI time step (drop-in for cmc-fluid-solver's AdiSolver3D)."""
from .params import FluidParams  # noqa: F401
from . import domains  # noqa: F401
