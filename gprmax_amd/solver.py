"""Drop-in solve function for the reference's GPU path.

``solve_hip(currentmodelrun, modelend, G)`` has the signature and side effects
of ``solve_gpu`` (gprMax/model_build_run.py:477-716): it runs all
``G.iterations`` on the device, fills ``rx.outputs`` of every receiver,
``snap.electric/magnetic`` of every snapshot and ``Vtotal``/``Itotal`` of
every transmission line, and returns
``(tsolve_seconds, memsolve_bytes)``.  ``GPU``/``detect_check_gpus`` stand in
for gprMax/utilities.py:341-413 so the ``-gpu`` selection works unchanged.
"""

import os

from .hip import HipSolver, device_info
from .receivers import scatter_rx_buffer
from .snapshots import store_components
from .utilities import GeneralError


class GPU:
    """GPU information (gprMax/utilities.py:341-366)."""

    def __init__(self, deviceID):
        self.deviceID = deviceID
        self.name = None
        self.pcibusID = None
        self.constmem = None
        self.totalmem = None

    def get_gpu_info(self, drv=None):
        info = device_info(self.deviceID)
        self.name = info['name']
        self.totalmem = info['totalmem']
        # Coefficient tables live in LDS, not constant memory; report the
        # table limit the library enforces instead.
        self.constmem = 64 * 1024


def detect_check_gpus(deviceIDs):
    """Return (gpus, text) for the requested device IDs
    (gprMax/utilities.py:369-413)."""
    visible = os.environ.get('HIP_VISIBLE_DEVICES') or os.environ.get('CUDA_VISIBLE_DEVICES')
    deviceIDs = deviceIDs or [0]
    gpus, text = [], []
    for ID in deviceIDs:
        gpu = GPU(ID)
        try:
            gpu.get_gpu_info()
        except GeneralError as exc:
            raise GeneralError('GPU with device ID {} does not exist ({})'.format(ID, exc))
        gpus.append(gpu)
        text.append('{} - {}, {} bytes{}'.format(ID, gpu.name, gpu.totalmem,
                                                 ' (visible: {})'.format(visible) if visible else ''))
    return gpus, text


def solve_hip(currentmodelrun, modelend, G):
    """Solve on the GPU named by ``G.gpu`` (device 0 if unset)."""
    device = G.gpu.deviceID if G.gpu is not None else 0
    with HipSolver(G, device=device) as s:
        ms, _, _ = s.run_timed(0, G.iterations)
        memsolve = s.memory_bytes()
        if G.rxs:
            scatter_rx_buffer(s.download_rx(), G)
        for n, snap in enumerate(G.snapshots):
            store_components(snap, *s.download_snapshot(n))
        for n, tl in enumerate(getattr(G, 'transmissionlines', [])):
            vt, it_, v, i = s.download_transmission_line(n)
            tl.Vtotal[:] = vt
            tl.Itotal[:] = it_
            tl.voltage[:tl.nl] = v
            tl.current[:tl.nl] = i
    return ms * 1e-3, memsolve
