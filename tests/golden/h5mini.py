"""HDF5 reader used by make_fixtures.py: the package's reader
(gprmax_amd/hdf5.py), re-exported so the fixture script keeps its import."""

import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))

from gprmax_amd.hdf5 import H5Error, read_file  # noqa: E402,F401
