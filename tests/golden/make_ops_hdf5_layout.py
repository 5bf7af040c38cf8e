"""Writes ops_hdf5_layout.json: the metadata layout (attribute names, types,
sizes, string padding, shapes) of an OPS-layout block file held by the
reference, apps/cylinder/incompressible_cylinder/data.h5 -- written by the
reference's output_hdf5 / set_hdf5_metadata (opensbli/utilities/
helperfunctions.py:99-147) in the layout ops_decl_dat_hdf5 reads.

Run with a Python that has h5py (the image's /opt/conda/bin/python3.9):
    /opt/conda/bin/python3.9 tests/golden/make_ops_hdf5_layout.py
Only data (the layout description) is written; no reference source is copied."""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from h5layout import layout, values  # noqa: E402

SRC = "/root/reference/apps/cylinder/incompressible_cylinder/data.h5"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ops_hdf5_layout.json")

if __name__ == "__main__":
    import h5py
    with h5py.File(SRC, "r") as h:
        shape = list(h["opensbliblock00/x0_B0"].shape)
    doc = {"layout": layout(SRC), "attr_values": values(SRC)["attrs"], "x0_B0_shape": shape}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print("wrote", OUT)
