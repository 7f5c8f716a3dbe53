"""Extract the plastic-cone iteration log recorded in the reference notebook.

Reads ``examples/plastic_simulation_cone.ipynb`` (cell 15 outputs) from the
reference checkout given on the command line (default ``/root/reference``)
and writes ``plastic_cone_notebook.json``: one row per logged iteration with
(load index, status, it, area, frac. area, total force, offset, plast. area,
frac. plast. area).  Only the recorded numbers are stored.

Run once in a container that has the reference tree:
    python tests/golden/make_notebook_fixture.py [/root/reference]
"""

import json
import os
import sys


def main(ref_root="/root/reference"):
    nb_path = os.path.join(ref_root, "examples", "plastic_simulation_cone.ipynb")
    with open(nb_path) as fh:
        nb = json.load(fh)
    cell = nb["cells"][15]
    text = "".join("".join(o.get("text", "")) for o in cell.get("outputs", []))
    rows = []
    load = 0
    for line in text.splitlines():
        parts = line.split()
        if not parts or line.startswith("#"):
            continue
        if len(parts) != 8:
            continue
        status, it = parts[0], int(parts[1])
        rows.append(dict(load=load, status=status, it=it, area=int(parts[2]),
                         frac_area=float(parts[3]), total_force=float(parts[4]),
                         offset=float(parts[5]), plastic_area=int(parts[6]),
                         frac_plastic_area=float(parts[7])))
        if status == "CONVERGED":
            load += 1
    out = dict(
        source="examples/plastic_simulation_cone.ipynb cell 15 (recorded outputs)",
        setup=dict(nx=128, ny=128, sx=0.005, sy=0.005, cone_slope=0.05,
                   young=230000.0, hardness=6000.0, substrate="free",
                   external_forces=[0.001, 0.005, 0.01, 0.02, 0.01, 0.005, 0.001],
                   initial_penetration=0.00002, pentol=1e-10, warm_start=True),
        rows=rows)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "plastic_cone_notebook.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"wrote {len(rows)} rows to {dst}")


if __name__ == "__main__":
    main(*sys.argv[1:])
