"""becgpp_amd -- MI355X-native ground-state solve path of becGPP.

Drop-in for the reference's ground-state API: ``default_cfg``/``run(cfg)``
and the call signatures of its ``grid``, ``operators``, ``interactions`` and
``solvers`` modules, computed by hand-written HIP kernels for gfx950 in
``libbecgpp_hip.so`` (C-ABI: include/becgpp_hip.h).  There is no CPU path.
"""
from .constants import CODE_VERSION, DEV
from .config import CFG_DEFAULTS, default_cfg, canonical_key
from .grid import make_grid, geometry, resolve_kernel, auto_grid, resample_state, dV, norm_of, Lz_op
from .interactions import long_range_phi, kernel_convolution, q2d_kernel, clear_kernel_cache
from .operators import energy_components, energy_t, apply_H, compute_residual, observables, sensitivities
from .thomasfermi import extract_tf, mass_quantiles, mass_radius, center_of_mass
from .vortex import vortex_diagnostic, lll_weight
from .diagnostics import diagnostics, diagnostics_batch, droplet_radius, radial_average
from .seeds import make_seed
from .solvers import (ground_state, ground_state_batch, ground_state_multiseed, ground_state_uneven, solve_native,
                      solve_native_batch)

from .modes import (run, run_id, DISPATCH, zip_results, mode_validate, mode_smoke, mode_single, mode_tf_only,
                    mode_sweep, mode_convergence)
from .units import physical_scales, sim_params_from_physical, outputs_to_physical
from . import paths

__version__ = "0.1.0"
