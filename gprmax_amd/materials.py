"""Materials and their update coefficients (gprMax/materials.py:26-241).

Each material row becomes ``updatecoeffsE[m] = (CA, CBx, CBy, CBz, srce)`` and
``updatecoeffsH[m] = (DA, DBx, DBy, DBz, srcm)``; dispersive materials add
``updatecoeffsdispersive[m, 3p:3p+3] = (e0*eqt2_p, eqt_p, zt_p)``.

The arithmetic keeps the reference's dtype flow: Python floats (float64) for the
standard coefficients, complex64 NumPy scalars for the pole terms.  Under NumPy
2 promotion the dispersive correction makes EA/EB float32 for every material of
a model that has poles; that is reproduced by computing it the same way.
"""

import numpy as np

from .constants import complextype, e0, floattype, m0


class Material:
    """One material: constitutive parameters plus derived update coefficients."""

    def __init__(self, numID, ID):
        self.numID = numID
        self.ID = ID
        self.type = ''
        self.averagable = True
        self.er = 1.0
        self.se = 0.0
        self.mr = 1.0
        self.sm = 0.0
        self.poles = 0
        self.deltaer = []
        self.tau = []
        self.alpha = []
        # Number of poles of the model this material belongs to (the reference
        # keeps this as the class attribute Material.maxpoles).
        self.maxpoles = 0

    def calculate_update_coeffsH(self, G):
        """DA, DB*, srcm (gprMax/materials.py:69-82)."""
        HA = (m0 * self.mr / G.dt) + 0.5 * self.sm
        HB = (m0 * self.mr / G.dt) - 0.5 * self.sm
        self.DA = HB / HA
        self.DBx = (1 / G.dx) * 1 / HA
        self.DBy = (1 / G.dy) * 1 / HA
        self.DBz = (1 / G.dz) * 1 / HA
        self.srcm = 1 / HA

    def _pole_terms(self, G):
        """Complex64 pole constants w, q and their exponentials
        (gprMax/materials.py:92-123)."""
        P = self.maxpoles
        ctype = getattr(G, 'complextype', complextype)
        self.w = np.zeros(P, dtype=ctype)
        self.q = np.zeros(P, dtype=ctype)
        self.zt = np.zeros(P, dtype=ctype)
        self.zt2 = np.zeros(P, dtype=ctype)
        self.eqt = np.zeros(P, dtype=ctype)
        self.eqt2 = np.zeros(P, dtype=ctype)
        for p in range(self.poles):
            if 'debye' in self.type:
                self.w[p] = self.deltaer[p] / self.tau[p]
                self.q[p] = -1 / self.tau[p]
            elif 'lorentz' in self.type:
                # tau holds the pole frequency and alpha the damping
                wp2 = (2 * np.pi * self.tau[p])**2
                root = np.sqrt(wp2 - self.alpha[p]**2)
                self.w[p] = -1j * ((wp2 * self.deltaer[p]) / root)
                self.q[p] = -self.alpha[p] + (1j * root)
            elif 'drude' in self.type:
                # tau holds the plasma frequency, alpha the inverse relaxation time
                wp2 = (2 * np.pi * self.tau[p])**2
                self.se += wp2 / self.alpha[p]
                self.w[p] = -(wp2 / self.alpha[p])
                self.q[p] = -self.alpha[p]
            self.eqt[p] = np.exp(self.q[p] * G.dt)
            self.eqt2[p] = np.exp(self.q[p] * (G.dt / 2))
            self.zt[p] = (self.w[p] / self.q[p]) * (1 - self.eqt[p]) / G.dt
            self.zt2[p] = (self.w[p] / self.q[p]) * (1 - self.eqt2[p])

    def calculate_update_coeffsE(self, G):
        """CA, CB*, srce (gprMax/materials.py:84-143)."""
        if self.maxpoles > 0:
            self._pole_terms(G)
            correction = (e0 / G.dt) * np.sum(self.zt2.real)
            EA = (e0 * self.er / G.dt) + 0.5 * self.se - correction
            EB = (e0 * self.er / G.dt) - 0.5 * self.se - correction
        else:
            EA = (e0 * self.er / G.dt) + 0.5 * self.se
            EB = (e0 * self.er / G.dt) - 0.5 * self.se

        if self.ID == 'pec' or self.se == float('inf'):
            self.CA = self.CBx = self.CBy = self.CBz = self.srce = 0
        else:
            self.CA = EB / EA
            self.CBx = (1 / G.dx) * 1 / EA
            self.CBy = (1 / G.dy) * 1 / EA
            self.CBz = (1 / G.dz) * 1 / EA
            self.srce = 1 / EA

    def calculate_er(self, freq):
        """Complex relative permittivity at ``freq`` (gprMax/materials.py:145-174)."""
        er = self.er
        if self.poles > 0:
            w = 2 * np.pi * freq
            er += self.se / (1j * w * e0)
            if 'debye' in self.type:
                for p in range(self.poles):
                    er += self.deltaer[p] / (1 + 1j * w * self.tau[p])
            elif 'lorentz' in self.type:
                for p in range(self.poles):
                    er += (self.deltaer[p] * self.tau[p]**2) / (self.tau[p]**2 + 2j * w * self.alpha[p] - w**2)
            elif 'drude' in self.type:
                total = 0
                for p in range(self.poles):
                    total += self.tau[p]**2 / (w**2 - 1j * w * self.alpha[p])
                    er -= total
        return er


def default_materials():
    """The two materials every model starts with: pec (0) and free_space (1)."""
    pec = Material(0, 'pec')
    pec.se = float('inf')
    pec.type = 'builtin'
    pec.averagable = False
    free = Material(1, 'free_space')
    free.type = 'builtin'
    return [pec, free]


def add_dispersion(material, kind, deltaer, tau, alpha=None):
    """Attach poles to a material, as #add_dispersion_{debye,lorentz,drude} do
    (gprMax/input_cmds_multiuse.py:587-618)."""
    material.type = kind
    material.poles = len(deltaer) if deltaer is not None else len(tau)
    material.deltaer = list(deltaer) if deltaer is not None else []
    material.tau = list(tau)
    material.alpha = list(alpha) if alpha is not None else []
    material.averagable = False


def process_materials(G):
    """Fill G.updatecoeffsE/H (and G.updatecoeffsdispersive) from G.materials
    (gprMax/materials.py:177-208)."""
    nmat = len(G.materials)
    maxpoles = max([m.poles for m in G.materials] + [0])
    G.maxpoles = maxpoles
    ftype = getattr(G, 'floattype', floattype)
    ctype = getattr(G, 'complextype', complextype)
    G.updatecoeffsE = np.zeros((nmat, 5), dtype=ftype)
    G.updatecoeffsH = np.zeros((nmat, 5), dtype=ftype)
    if maxpoles > 0:
        G.updatecoeffsdispersive = np.zeros((nmat, 3 * maxpoles), dtype=ctype)
    else:
        G.updatecoeffsdispersive = None
    for m in G.materials:
        m.maxpoles = maxpoles
        m.calculate_update_coeffsE(G)
        m.calculate_update_coeffsH(G)
        G.updatecoeffsE[m.numID, :] = m.CA, m.CBx, m.CBy, m.CBz, m.srce
        G.updatecoeffsH[m.numID, :] = m.DA, m.DBx, m.DBy, m.DBz, m.srcm
        if maxpoles > 0:
            for p in range(maxpoles):
                G.updatecoeffsdispersive[m.numID, 3 * p:3 * p + 3] = e0 * m.eqt2[p], m.eqt[p], m.zt[p]
