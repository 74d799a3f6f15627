"""Models with transmission lines for the CPU and GPU tests."""

from gprmax_amd.grid import ModelBuilder
from gprmax_amd.pml import CFS


def tl_model(iterations=120, order=1, formulation='HORIPML', extra_sources=True):
    """A 24x22x20 grid with 4-cell PML, a soil layer, a line of each
    polarisation (one with a start/stop window) and one on a grid face,
    point sources away from the lines' edges and receivers, one on a line's
    edge."""
    cfs = None
    if order == 2:
        a, b = CFS(), CFS()
        b.alpha.max = 0.05
        b.kappa.scalingprofile, b.kappa.min, b.kappa.max = 'quadratic', 1, 3
        cfs = [a, b]
    bld = ModelBuilder(24, 22, 20, 1e-3, 1e-3, 1e-3, iterations, pml_cells=4,
                       pml_formulation=formulation, cfs=cfs)
    soil = bld.add_material(5.0, 0.01, 1.0, 0.0, 'soil')
    bld.box(0, 24, 0, 22, 0, 8, soil)
    bld.add_waveform('gaussian', 1, 3e9, 'g')
    bld.add_waveform('ricker', 1, 2e9, 'r')
    bld.add_transmission_line('z', (12, 11, 10), 73, 'g')
    bld.add_transmission_line('x', (8, 9, 12), 50, 'r', start=1e-11, stop=1.5e-10)
    bld.add_transmission_line('y', (15, 6, 14), 120, 'g')
    # on the y = 0 face, inside the PML: its loop current is 0 there
    bld.add_transmission_line('x', (5, 0, 7), 60, 'g')
    if extra_sources:
        bld.add_hertzian_dipole('z', (6, 15, 9), 'r')
        bld.add_magnetic_dipole('x', (17, 14, 6), 'g')
        bld.add_voltage_source('y', (10, 16, 11), 40, 'r')
    bld.add_rx((12, 11, 10))
    bld.add_rx((16, 12, 9))
    return bld.finalise()
