"""MI355X-native dual simplex: gfx950 HIP kernels behind a C-ABI (libpds.so)."""
from ._native import lib  # noqa: F401
from .lp import LP, gen_bernoulli, gen_kcol, gen_transport, packing_lp, equality_lp  # noqa: F401
from .solver import DualSimplex, LPsolution, default_params, solve_lp  # noqa: F401
