"""fluxoptics (MI355X): the FluxOptics.jl propagation path on libfluxhip.

Host-side mirror of the reference's interface for the propagators (ASProp, RSProp,
ShiftProp), modulators, BPM volumes (AS_BPM, Shift_BPM), metrics and optimisers;
the arithmetic is the HIP library's (``libfluxhip.so``, C ABI in include/fluxhip.h).
"""

from .fields import (Backward, Direction, Forward, ScalarField, coupling_efficiency, intensity,
                     normalize_power_, phase, power, rs_valid_distance)
from .components import (AS_BPM, BPM, ASProp, ASPropZ, CollinsProp, FourierLens, ParaxialProp, FourierMask, FourierOperator, FourierPhase, FourierWrapper,
                         Mask, Phase, RSProp, ShiftProp, Shift_BPM, TeaDOE, TeaReflector, TiltAnchor,
                         phase_propagate_,
                         propagate_phase_backward_)
from .czt import CZTPlan, plan_czt, plan_czt_
from .group import PropagationGroup
from .api import (FFTPlans, backpropagate, backpropagate_metric, compute_ft_, compute_ift_, compute_metric, dot, get_data, get_lambdas, get_lambdas_collection, get_source, get_tilts, get_tilts_collection, is_on_axis, isbuffered, istrainable, make_fft_plans, offset_tilts_, set_field_data, set_field_ds_, set_field_tilts, spatial_vectors, trainable)
from .padcrop import CropOperator, FieldProbe, PadCropOperator, PadOperator, crop, crop_, pad, pad_
from .metrics import DotProduct, PowerCoupling, SquaredFieldDifference, SquaredIntensityDifference
from .system import OpticalSequence, OpticalSystem, ScalarSource, merge
from .optim import (Adam, ClampProx, CompositeProx, Descent, Fista, IstaProx, Momentum, Nesterov, NoDescent,
                    PositiveProx, ProxRule, compose, make_rules, setup, update_)

__all__ = [
    "ASPropZ",
    "FFTPlans", "backpropagate", "backpropagate_metric", "compute_ft_", "compute_ift_", "compute_metric", "dot", "get_data", "get_lambdas", "get_lambdas_collection", "get_source", "get_tilts", "get_tilts_collection", "is_on_axis", "isbuffered", "istrainable", "make_fft_plans", "offset_tilts_", "set_field_data", "set_field_ds_", "set_field_tilts", "spatial_vectors", "trainable",
    "intensity", "normalize_power_", "phase", "power",
    "SquaredFieldDifference",
    "CropOperator", "FieldProbe", "PadCropOperator", "PadOperator", "crop", "crop_", "pad", "pad_",
    "CZTPlan", "plan_czt", "plan_czt_", "AS_BPM", "BPM", "Shift_BPM", "TiltAnchor", "ASProp", "CollinsProp", "FourierLens", "ParaxialProp", "Backward", "Direction", "FourierOperator", "FourierWrapper", "Forward", "Phase",
    "RSProp", "ScalarField", "ShiftProp", "Mask", "TeaDOE", "TeaReflector", "FourierPhase", "FourierMask",
    "Adam", "ClampProx", "CompositeProx", "Descent", "Fista", "IstaProx", "Momentum", "Nesterov", "NoDescent",
    "PositiveProx", "ProxRule", "compose", "make_rules", "setup", "update_",
    "OpticalSequence", "OpticalSystem", "PropagationGroup", "ScalarSource", "SquaredIntensityDifference", "DotProduct", "PowerCoupling", "merge",
    "coupling_efficiency", "phase_propagate_", "propagate_phase_backward_", "propagate",
    "propagate_", "backpropagate_", "rs_valid_distance",
]


def propagate_(u, component, direction=Forward):
    """``propagate!(u, p, direction)``: in place."""
    return component.propagate_(u, direction)


def propagate(u, component, direction=Forward):
    """``propagate(u, p, direction)``: on a copy."""
    return component.propagate(u, direction)


def backpropagate_(dv, component, direction=Forward):
    """``backpropagate!(dv, p, direction)``."""
    return component.backpropagate_(dv, direction)
