"""adfwi_amd -- MI355X-native propagator hot path of ADFWI (acoustic / elastic / VTI).

Public API mirrors ADFWI's propagator layer:

    from adfwi_amd.model import AcousticModel
    from adfwi_amd.survey import Source, Receiver, Survey
    from adfwi_amd.propagator import AcousticPropagator

The time stepping and its discrete adjoint run in libadfwi_amd.so (hand-written
HIP for gfx950, C ABI in include/adfwi_amd.h); there is no CPU fallback.
"""
__version__ = "0.1.0"
