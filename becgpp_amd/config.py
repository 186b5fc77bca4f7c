"""Run configuration: every tunable key with its default, user aliases, and
unknown-key rejection (becgpp/config.py:15-119).

The key set and defaults are the reference's (a configuration written for the
reference runs unchanged here); ``default_cfg`` applies the same alias and
validation rules.
"""

CFG_DEFAULTS = dict(
    mode="smoke",
    # geometry
    dimension="2D", s=2, trap_coeff=0.5, Omega=0.0,
    # interactions
    beta2=100.0, beta3=0.0, G_C=0.0, kernel="auto", l_z=0.0,
    # grid
    L=10.0, Ngrid=128, pad=2, mem_warn_gb=12.0,
    # solver
    maxit=60000, res_tol=1e-4, energy_tol=1e-9, conv_window=2000, step=1.0, step_max=3.0,
    linesearch_max=8, linesearch_stall=10, precond_shift_min=0.5, cg_beta="pr_precond",
    cg_restart=30, record_trace=False, check=200,
    # seeding
    seed="tf", seed_winding=0, seed_noise=1e-2, seed_phase_noise=5e-2, rng=0, nseeds=1,
    # diagnostics
    want_vortices=True, want_lll=True, tf_cache=True,
    # validate mode
    validate_N=256,
    # sweep mode
    sweep_param="G_C", sweep_values=[1.0, 2.0, 4.0, 8.0, 16.0], sweep_autobox=False,
    sweep_continuation=True,
    # convergence mode
    conv_Ngrids=[96, 128, 192], conv_box_factors=[1.0, 1.2], conv_pads=[2, 3],
    # output
    tag="becgpp", show_inline=True, save_figs=True, save_ckpt=True, zip_output=False,
)

# Keys of this build with no counterpart in the reference: accepted by
# default_cfg when given, absent otherwise (so a reference configuration, and
# its run id, are unchanged).
#   warm_fp32:    run the loop on complex64 fields to warm_res_tol first, then
#                 polish in FP64 (instances with long range stay FP64)
#   host_loop:    keep the loop's scalar logic on the host (one synchronisation
#                 per iteration) instead of the device controllers
#   autotune:     time the candidate FFT stage orders and lines per workgroup
#                 on the first solve of a plan and keep the fastest
#   line_search:  "poly" prices every Armijo trial of an iteration from one
#                 pass of line coefficients (one loop step per iteration);
#                 "trial" evaluates one trial per loop step
CFG_EXTENSIONS = dict(warm_fp32=False, warm_res_tol=1e-4, host_loop=False, autotune=False, line_search="poly")

_ALIASES = {"res": "res_tol", "etol": "energy_tol", "N": "Ngrid", "dim": "dimension",
            "precond_shift": "precond_shift_min"}


def canonical_key(key):
    """User alias -> canonical key (config.py:95-97)."""
    return _ALIASES.get(key, key)


def default_cfg(**overrides):
    """Fresh copy of the defaults with ``overrides`` applied (config.py:100-119).

    Aliases map to their canonical key; passing both an alias and its key, or
    any key that is not a configuration key, raises ``ValueError``.
    """
    given = dict(overrides)
    for alias, key in _ALIASES.items():
        if alias in given:
            if key in given:
                raise ValueError(f"pass only one of {alias!r} or {key!r}")
            given[key] = given.pop(alias)
    unknown = sorted(set(given) - set(CFG_DEFAULTS) - set(CFG_EXTENSIONS))
    if unknown:
        raise ValueError(f"unknown config key(s): {unknown}; "
                         f"see becgpp_amd.config.CFG_DEFAULTS for valid keys")
    cfg = dict(CFG_DEFAULTS)
    cfg.update(given)
    return cfg
