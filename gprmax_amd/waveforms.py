"""Source waveforms, evaluated on the host in float64 (gprMax/waveforms.py:47-122).

The solver never evaluates a transcendental on the GPU: every waveform is
sampled here once per iteration and handed to the device as float32 tables.
"""

import numpy as np


class Waveform:
    """A named waveform shape with amplitude and centre frequency.

    The shapes and their parameters follow gprMax/waveforms.py:26-122:
    ``chi`` is the delay of the pulse centre and ``zeta`` its width parameter.
    """

    types = ('gaussian', 'gaussiandot', 'gaussiandotnorm', 'gaussiandotdot',
             'gaussiandotdotnorm', 'gaussianprime', 'gaussiandoubleprime',
             'ricker', 'sine', 'contsine', 'impulse', 'user')

    def __init__(self, ID=None, type=None, amp=1, freq=None, userfunc=None):
        self.ID = ID
        self.type = type
        self.amp = amp
        self.freq = freq
        self.userfunc = userfunc
        self.chi = 0
        self.zeta = 0

    def calculate_coefficients(self):
        """Set chi/zeta (gprMax/waveforms.py:47-55)."""
        gaussian_family = ('gaussian', 'gaussiandot', 'gaussiandotnorm',
                           'gaussianprime', 'gaussiandoubleprime')
        if self.type in gaussian_family:
            self.chi = 1 / self.freq
            self.zeta = 2 * np.pi**2 * self.freq**2
        elif self.type in ('gaussiandotdot', 'gaussiandotdotnorm', 'ricker'):
            self.chi = np.sqrt(2) / self.freq
            self.zeta = np.pi**2 * self.freq**2

    def calculate_value(self, time, dt):
        """Waveform value at ``time`` seconds (gprMax/waveforms.py:57-122)."""
        self.calculate_coefficients()
        kind = self.type
        zeta = self.zeta
        if kind in ('gaussian', 'gaussiandot', 'gaussianprime', 'gaussiandotnorm',
                    'gaussiandotdot', 'gaussiandoubleprime', 'gaussiandotdotnorm', 'ricker'):
            delay = time - self.chi
            g = np.exp(-zeta * delay**2)
            if kind == 'gaussian':
                value = g
            elif kind in ('gaussiandot', 'gaussianprime'):
                value = -2 * zeta * delay * g
            elif kind == 'gaussiandotnorm':
                value = -2 * zeta * delay * g * np.sqrt(np.exp(1) / (2 * zeta))
            elif kind in ('gaussiandotdot', 'gaussiandoubleprime'):
                value = 2 * zeta * (2 * zeta * delay**2 - 1) * g
            elif kind == 'gaussiandotdotnorm':
                value = 2 * zeta * (2 * zeta * delay**2 - 1) * g * (1 / (2 * zeta))
            else:  # ricker: the negated, normalised second derivative
                value = -(2 * zeta * (2 * zeta * delay**2 - 1) * g) * (1 / (2 * zeta))
        elif kind == 'sine':
            value = np.sin(2 * np.pi * self.freq * time)
            if time * self.freq > 1:
                value = 0
        elif kind == 'contsine':
            ramp = min(0.25 * time * self.freq, 1)
            value = ramp * np.sin(2 * np.pi * self.freq * time)
        elif kind == 'impulse':
            value = 1 if (time == 0 or time < dt) else 0
        elif kind == 'user':
            value = self.userfunc(time)
        else:
            raise ValueError('unknown waveform type {!r}'.format(kind))
        return value * self.amp
