"""pyshocks_amd: the WENO-JS5 / Rusanov / Godunov / SSP-RK3 hot path of
pyshocks for AMD MI355X (gfx950).

The public names follow ``pyshocks/__init__.py:85-127`` for the hot-path
subset; arrays are ``torch.Tensor`` on the HIP device and every numerical
routine runs in ``libpyshocks_amd.so`` (``include/pyshocks_amd.h``).
"""

from __future__ import annotations

from pyshocks_amd.grid import (
    Grid,
    Quadrature,
    UniformGrid,
    cell_average,
    make_leggauss_quadrature,
    make_uniform_cell_grid,
    norm,
    rnorm,
)
from pyshocks_amd.schemes import (
    Boundary,
    BoundaryType,
    ConservationLawScheme,
    FiniteDifferenceSchemeBase,
    FiniteVolumeSchemeBase,
    SchemeBase,
    SchemeT,
    apply_boundary,
    apply_operator,
    bind,
    evaluate_boundary,
    flux,
    numerical_flux,
    predict_timestep,
)
from pyshocks_amd.tools import EOCRecorder, estimate_order_of_convergence

__all__ = (
    "Boundary",
    "BoundaryType",
    "ConservationLawScheme",
    "EOCRecorder",
    "FiniteDifferenceSchemeBase",
    "FiniteVolumeSchemeBase",
    "Grid",
    "Quadrature",
    "SchemeBase",
    "SchemeT",
    "UniformGrid",
    "apply_boundary",
    "apply_operator",
    "bind",
    "cell_average",
    "estimate_order_of_convergence",
    "evaluate_boundary",
    "flux",
    "make_leggauss_quadrature",
    "make_uniform_cell_grid",
    "norm",
    "numerical_flux",
    "predict_timestep",
    "rnorm",
)
