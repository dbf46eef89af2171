"""MI355X (gfx950) 3D Yee FDTD solver: HIP kernels behind the reference's simFDTD operator API."""
from .simfdtd import Batch, Server, Simulation, SimFDTDError, simFDTD, lib  # noqa: F401
