"""Write tests/golden/reference_outputs.json from numbers the reference's
notebooks printed (recorded cell outputs, read as text; nothing is run).

Sources (paths relative to the reference repository):
  ARTn/Parallel_multigrid_2D.ipynb  cell 50 output: w_2 (2x2), cell 52: w_4 (4x4)
      after one sawtooth cycle on 128^2 (stencil cell 5, initial field cells 15-16)
  ARTn/Parallel_multigrid_1D.ipynb  cell 19 output: "w_32 CNN" (A applied to the
      prolonged w_16) and "w_32 JACOBI" (after the level-32 sweep); cell 21: w_4
  ARTn/Multigrid_3d_7point_source.ipynb cells 4, 5, 6, 10 outputs: kernel sum,
      the w2[0,1] slice, the centre weight and the full 3x3x3 kernel printed by
      CNN3D_A_1.get_weights() (row-major (3,3,3,1,1) layout)

Usage (in a container that holds the reference):
  python tests/golden/extract_reference_outputs.py /root/reference
"""
from __future__ import annotations

import json
import os
import re
import sys

ARTN = os.path.join("Anisotropic Resistivity Tomography", "notebooks")
NUM = r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?"


def cell_outputs(nb_path, idx):
    nb = json.load(open(nb_path))
    c = nb["cells"][idx]
    txt = []
    for o in c.get("outputs", []):
        if "text" in o:
            txt.append("".join(o["text"]))
        elif "data" in o and "text/plain" in o["data"]:
            txt.append("".join(o["data"]["text/plain"]))
    return "\n".join(txt)


def numbers(s):
    return [float(v) for v in re.findall(NUM, s)]


def tensor_block(text, start_marker):
    """Numbers of the first tf.Tensor(...) printout following start_marker."""
    i = text.index(start_marker)
    j = text.index("tf.Tensor(", i)
    k = text.index("shape=", j)
    return numbers(text[j + len("tf.Tensor("):k])


def main(ref):
    out = {}
    nb2 = os.path.join(ref, ARTN, "Parallel_multigrid_2D.ipynb")
    t50 = cell_outputs(nb2, 50)
    t52 = cell_outputs(nb2, 52)
    out["mg2d_w2"] = tensor_block(t50, "tf.Tensor(")
    out["mg2d_w4"] = tensor_block(t52, "tf.Tensor(")
    nb1 = os.path.join(ref, ARTN, "Parallel_multigrid_1D.ipynb")
    t19 = cell_outputs(nb1, 19)
    out["mg1d_w32_cnn"] = tensor_block(t19, "w_32 CNN:")
    out["mg1d_w32_jacobi"] = tensor_block(t19, "w_32 JACOBI:")
    out["mg1d_w4"] = tensor_block(cell_outputs(nb1, 21), "tf.Tensor(")
    nb3 = os.path.join(ref, ARTN, "Multigrid_3d_7point_source.ipynb")
    out["src3d_sum"] = numbers(cell_outputs(nb3, 4))
    out["src3d_slice"] = numbers(cell_outputs(nb3, 5))
    out["src3d_centre"] = numbers(cell_outputs(nb3, 6))
    gw = cell_outputs(nb3, 10)
    gw = gw[gw.index("[array("):]
    out["src3d_getweights"] = numbers(gw[:gw.index("dtype")].replace("array(", ""))
    out["_sources"] = {
        "mg2d_w2": "ARTn/Parallel_multigrid_2D.ipynb#c50 out",
        "mg2d_w4": "ARTn/Parallel_multigrid_2D.ipynb#c52 out",
        "mg1d_w32_cnn": "ARTn/Parallel_multigrid_1D.ipynb#c19 out",
        "mg1d_w32_jacobi": "ARTn/Parallel_multigrid_1D.ipynb#c19 out",
        "mg1d_w4": "ARTn/Parallel_multigrid_1D.ipynb#c21 out",
        "src3d_sum": "ARTn/Multigrid_3d_7point_source.ipynb cell 4 out (raw line 117)",
        "src3d_slice": "ARTn/Multigrid_3d_7point_source.ipynb cell 5 out (raw lines 137-139)",
        "src3d_centre": "ARTn/Multigrid_3d_7point_source.ipynb cell 6 out (raw line 159)",
        "src3d_getweights": "ARTn/Multigrid_3d_7point_source.ipynb cell 10 out",
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_outputs.json")
    json.dump(out, open(dst, "w"), indent=1)
    print("wrote", dst)
    for k, v in out.items():
        if not k.startswith("_"):
            print(k, len(v), v[:6])


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
