"""tetfit_amd - MI355X-native time-stepping hot path of TetFIT (FIT/FDTD,
explicit DGA and CG-solved DGA leapfrog), driven through C ABIs:
libtetfit_host.so (input language, meshers, operator producers) and
libtetfit_hip.so (gfx950 HIP kernels)."""
from .host import Model, Build, TetfitError, write_kuhn_mesh  # noqa: F401
