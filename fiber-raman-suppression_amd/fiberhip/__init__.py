"""fiberhip -- host side of libfiberhip, the MI355X GMMNLSE propagation engine.

Mirrors the FiberLab surface that sits on the reference's propagation seam
(``fiber_problem`` / ``propagate`` / ``verify`` and the fixed-step
``_scalar_rk4_output``) over the C ABI in include/fiberhip.h.  All numerics
run in libfiberhip's HIP kernels; this package only builds operator arrays on
the host the way the reference's helpers do and moves fields across the ABI.
"""
from ._lib import FiberHIPError, FiberHIPUnavailable, exported_names, lib  # noqa: F401
from .engine import Context, DeviceArray, Plan, Tape, default_context  # noqa: F401
from .serving import Server  # noqa: F401
from .problem import (  # noqa: F401
    Fiber,
    FiberFieldProblem,
    Grid,
    Pulse,
    fftfreq,
    fiber_field_problem,
    fiber_problem,
    frequency_band_mask,
    get_disp_fiber_params_user_defined,
    get_disp_sim_params,
    get_initial_state,
    raman_fields,
    resolve_grid,
    resolve_sampling_grid,
    self_steepening,
    taylor_phase_basis,
    with_launch,
    with_raman_fraction,
)
from .adjoint import (  # noqa: F401
    AdjointModel,
    FullGridPhase,
    ObjectiveMap,
    PhaseBasis,
    band_gradient,
    check_adjoint_gradient,
    fiber_model,
    mmf_cost_fundamental,
    mmf_cost_sum,
    mmf_cost_worst_mode,
    mmf_objective,
    raman_band_objective,
    run_adjoint_step,
    spectral_band_objective,
)
from .propagation import (  # noqa: F401
    PropagationResult,
    centroid_thz,
    metrics,
    propagate,
    scalar_rk4_output,
    verify,
)
