"""gbgpu_amd: FastGB waveform + likelihood backend for AMD MI355X (gfx950).

Same user API as GBGPU (``GBGPU``, ``GBGPUThirdBody``, ``run_wave``,
``get_ll``, ``fill_global_template``, ``generate_global_template``,
``inject_signal``, ``information_matrix``, ``swap_likelihood_difference``),
with the waveform, likelihood and template-summation work in hand-written
HIP kernels behind the C ABI in ``include/gbgpu_hip.h``.
"""

from .cutils import KNOWN_BACKENDS, get_backend
from .gbgpu import GBGPU, GBGPUBase
from .orbits import EqualArmlengthOrbits, InterpolatedOrbits, Orbits
from .thirdbody import GBGPUThirdBody

__all__ = [
    "GBGPU",
    "GBGPUBase",
    "GBGPUThirdBody",
    "EqualArmlengthOrbits",
    "InterpolatedOrbits",
    "Orbits",
    "KNOWN_BACKENDS",
    "get_backend",
]
