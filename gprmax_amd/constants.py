"""Physical constants and the precision contract of the solver.

The constants come from ``scipy.constants`` exactly as the reference takes them
(gprMax/constants.py:19-25), so coefficient tables built here and by the
reference agree on the same SciPy.  Field arrays are float32 and the
dispersive coefficient table is complex64 (gprMax/constants.py:34-41).
"""

import numpy as np
from scipy.constants import c
from scipy.constants import epsilon_0 as e0
from scipy.constants import mu_0 as m0

# Impedance of free space, computed as in gprMax/constants.py:25
z0 = np.sqrt(m0 / e0)

floattype = np.float32
complextype = np.complex64

# Order of the field components in the ID array and in the receiver buffer
# (gprMax/grid.py:169, gprMax/receivers.py:30).
COMPONENTS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')
IDlookup = {name: n for n, name in enumerate(COMPONENTS)}

__all__ = ['c', 'e0', 'm0', 'z0', 'floattype', 'complextype', 'COMPONENTS', 'IDlookup']
