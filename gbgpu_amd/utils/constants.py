"""Physical constants used by the FastGB path.

The reference star-imports these names from ``lisatools.utils.constants``
(src/gbgpu/gbgpu.py:9, src/gbgpu/utils/utility.py:5, src/gbgpu/thirdbody.py:8).
That package is not vendored, so the values are fixed here.  ``G_SI``,
``MSUN_SI`` and ``PC_SI`` are the values that reproduce, to every printed
digit, the tutorial notebook's ``get_amplitude`` (6.37562267e-23) and
``get_fdot`` (1.73157515e-19) outputs (examples/GBGPU_tutorial.ipynb cells 30
and 32); tests/test_constants.py holds that check.
"""

import numpy as np

C_SI = 299792458.0
G_SI = 6.67408e-11
MSUN_SI = 1.98848e30
PC_SI = 3.085677581491367e16
YRSID_SI = 31558149.763545603
AU_SI = 149597870700.0
PI = np.pi
Mjup = 1.898e27

# LISA constellation constants (src/gbgpu/cutils/LISA.h:12-27)
LARM_SI = 2.5e9  # mean arm length (m)
LISA_EC = 0.0048241852  # orbital eccentricity
LISA_KAPPA = 0.0  # initial azimuthal position of the guiding centre
LISA_LAMBDA = 0.0  # initial orientation of the constellation
LISA_FM = 3.168753575e-8  # modulation frequency (Hz), 1 / sidereal year
FSTAR = C_SI / (2.0 * np.pi * LARM_SI)  # transfer frequency, 0.01908538063694777 Hz

__all__ = [
    "C_SI",
    "G_SI",
    "MSUN_SI",
    "PC_SI",
    "YRSID_SI",
    "AU_SI",
    "PI",
    "Mjup",
    "LARM_SI",
    "LISA_EC",
    "LISA_KAPPA",
    "LISA_LAMBDA",
    "LISA_FM",
    "FSTAR",
]
