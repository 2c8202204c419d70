"""Writes tests/golden/orifice_geometry.mles: the geometry of the reference's
own test mesh src/core/common/tests/assets/orifice.mles (an orifice plate in a
pipe: 1056 cells, 5654 faces, 693 boundary faces in six zones) as the solver
reads it (cells, faces, boundary zones and faces with their
area_over_delta_bf and Gij_bf), written back with this repository's .mles
writer (no node records). GPU tests load it where the reference tree is absent;
reading it back gives bit-identical mesh and operator arrays (checked below).

Usage: python tests/golden/make_orifice_fixture.py [path/to/orifice.mles]
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from nextgen_amd.ideal_gas import Mesh  # noqa: E402

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/core/common/tests/assets/orifice.mles"
OUT = os.path.join(HERE, "orifice_geometry.mles")

m = Mesh.read_mles(SRC)
m.write_mles(OUT)
r = Mesh.read_mles(OUT)
for name in ("x_cv", "vol", "cvofa", "n_fa", "x_fa", "lo", "hi", "n_use", "c_use", "grad_coeff", "cvobf", "zone_bf",
             "n_bf", "aod_bf", "area_bf"):
    assert np.array_equal(m.get(name), r.get(name)), name
assert m.zones == r.zones
print(f"wrote {OUT}: {r.ncv} cells, {r.nfa} faces, {r.nbf} boundary faces in zones {r.zones}")
