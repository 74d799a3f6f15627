"""Edge material indices built on the GPU (gprmax_amd.edges) against the host
build (grid.build_electric_components / build_magnetic_components, itself
pinned bit for bit to the reference's compiled yee_cell_build_ext in
test_host_model.py)."""

import numpy as np
import pytest

from gprmax_amd.grid import build_electric_components, build_magnetic_components
from models import edge_model

pytestmark = pytest.mark.gpu


def _model(seed, shape=(30, 26, 21), nmat=6, frac=0.05):
    return edge_model(seed, shape, nmat, frac)


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_device_edge_build_matches_host(seed):
    from gprmax_amd.edges import build_components_on_device
    G1, G2 = _model(seed), _model(seed)
    build_electric_components(G1)
    build_magnetic_components(G1)
    build_components_on_device(G2)
    assert np.array_equal(G1.ID, G2.ID)
    assert [m.ID for m in G1.materials] == [m.ID for m in G2.materials]
    assert len(G1.materials) > 50
    for a, b in zip(G1.materials, G2.materials):
        assert (a.er, a.se, a.mr, a.sm) == (b.er, b.se, b.mr, b.sm)


def test_device_edge_build_on_a_2d_grid_and_small_table():
    """A TMz plane, and a hash table smaller than the number of sets fails
    with a clear error."""
    from gprmax_amd.edges import build_components_on_device
    from gprmax_amd.utilities import GeneralError
    G1, G2 = _model(5, shape=(40, 33, 1)), _model(5, shape=(40, 33, 1))
    build_electric_components(G1)
    build_magnetic_components(G1)
    build_components_on_device(G2)
    assert np.array_equal(G1.ID, G2.ID)
    with pytest.raises(GeneralError, match='table'):
        build_components_on_device(_model(6), capacity=4)
