"""geothermalwells.jl_amd -- MI355X (gfx950) implementation of GeothermalWells.jl's
per-timestep hot path (z-diffusion stencil under ROCK2, ADI x/y Thomas solves,
semi-Lagrangian pipe advection, outlet reductions) behind the reference's API.

The compute path is libgeowells_hip.so (HIP kernels + C ABI, include/geowells.h);
this package is the host-side mirror of the reference's Julia interface.
"""
from . import _abi
from .adjoint import outlet_gradient
from .cache import (Cache, CallbackSet, DiscreteCallback, SavedValues, create_batch_cache, create_cache,
                    get_simulation_callback, load_checkpoint, save_checkpoint)
from .grids import compute_domain, create_adaptive_grid_1d, create_uniform_gridz_with_borehole_depths
from .physics import (AbstractBorehole, AbstractMaterialProperties, Borehole, ConstantInlet, CustomInlet,
                      HeatExchangerInlet, HomogenousMaterialProperties, PerWellInlet, StratifiedMaterialProperties,
                      initial_condition_thermal_gradient)
from .util import (data_brown_single_well_b, data_brown_single_well_c, data_hu, data_li, extract_x_profile,
                   get_temperatures_along_z_single_well)
from .solver import (ADI_and_ADV_callback_, ODEProblem, ROCK2, eigen_estimator, rhs_diffusion_z_,
                     rock2_coefficients, solve, step_plan)

LIB_PATH = _abi.LIB_PATH


def load_library():
    """Load libgeowells_hip.so (raises if it is not built)."""
    return _abi.lib()
