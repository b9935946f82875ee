"""nbodygradient_amd -- NbodyGradient.jl's AHL21 step with Jacobian
propagation and transit timing, running on MI355X (gfx950) through
libnbg_hip.so.  The public names mirror the reference's Julia API."""
from ._lib import Context, NbgError, default_context, lib
from .core import (CartesianOutput, Derivatives, Integrator, State, TransitParameters, TransitSnapshot, TransitTiming,
                   ahl21, dState)
from .ics import CartesianIC, Elements, ElementsIC, hierarchy, init_nbody, kepler_init

__all__ = ["CartesianOutput", "Context", "NbgError", "default_context", "lib", "Derivatives", "Integrator", "State",
           "TransitParameters", "TransitSnapshot", "TransitTiming", "ahl21", "dState", "CartesianIC", "Elements", "ElementsIC",
           "hierarchy", "init_nbody", "kepler_init"]
