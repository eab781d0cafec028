"""Record and gradient post-processing of the FWI loop on the GPU (ADFWI/fwi/misfit/L2.py,
the trace normalisation of ADFWI/fwi/acoustic_fwi.py, ADFWI/propagator/gradient_process.py)."""
from .gradient_process import GradProcessor, grad_taper, smooth2d
from .misfit import Misfit_waveform_L2, Misfit_waveform_L2_squared, NormalizedWaveformL2, normalize_traces

__all__ = ["GradProcessor", "grad_taper", "smooth2d", "Misfit_waveform_L2", "Misfit_waveform_L2_squared",
           "NormalizedWaveformL2",
           "normalize_traces"]
