"""Writes tests/golden/ref_mesh_m{1,2,3}.npz: the box meshes produced by the
reference's own mesh generator (src/examples/mesh_examples.cpp:120-571, built
from its sources by `make -C oracle ref` into oracle/_ref/ref_mesh_dump), so
the oracle's mesh is pinned by the reference even where the reference tree is
absent; and tests/golden/ref_vtk_m2.npz: input fields and the file the
reference's VisitWriter::write_flow (include/visit/VisitWriter.hpp:127-234,
oracle/_ref/ref_vtk_dump) writes for them on the m = 2 box.  Run from the
repository root:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402


def main():
    if not os.path.exists(po.REF_MESH_DUMP):
        raise SystemExit("oracle/_ref/ref_mesh_dump missing: run `make -C oracle ref` first")
    with tempfile.TemporaryDirectory() as d:
        for m in (1, 2, 3):
            mesh = po.ref_mesh(m, d)
            arrays = {k: np.ascontiguousarray(v) for k, v in mesh.items() if k != "m"}
            out = os.path.join(HERE, "ref_mesh_m%d.npz" % m)
            np.savez_compressed(out, **arrays)
            print("wrote", out, {k: v.shape for k, v in arrays.items()})
        m = 2
        n = 2 * m ** 3
        fields = po.vtk_test_fields(n, 2024)
        vtk = po.ref_vtk(m, fields, d)
        out = os.path.join(HERE, "ref_vtk_m%d.npz" % m)
        np.savez_compressed(out, fields=fields, vtk=np.frombuffer(vtk, dtype=np.uint8), timestep=np.int64(7))
        print("wrote", out, len(vtk), "bytes of VTK")


if __name__ == "__main__":
    main()
