"""Write tests/golden/extent_placeholder_expected.json: the expected
detector/image extents of the reference's tests/test_extent_placeholder.py
(fixed image extent 2.0 resp. fixed detector extent 2.0; Nx=400, Ns=100,
Na=180; side ratios, detector centres and image centres as listed in the
JSON).  Only the numbers are taken (parsed with ``ast``); run from the
repository root:  python tests/golden/make_extent_tables.py
"""

import ast
import json
import os

SRC = "/root/reference/tests/test_extent_placeholder.py"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "extent_placeholder_expected.json")


def _table(tree, name):
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign) and any(getattr(t, "id", None) == name for t in node.targets):
            out = {}
            for key, value in zip(node.value.keys, node.value.values):
                placeholder = key.elts[0].attr
                ratio = ast.literal_eval(key.elts[1])
                out[f"{placeholder},{ratio}"] = ast.literal_eval(value)
            return out
    raise KeyError(name)


def _constant(tree, name):
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign) and any(getattr(t, "id", None) == name for t in node.targets):
            return ast.literal_eval(node.value)
    raise KeyError(name)


if __name__ == "__main__":
    tree = ast.parse(open(SRC).read())
    data = {
        "Nx": _constant(tree, "Nx"), "Ns": _constant(tree, "Ns"), "Na": _constant(tree, "Na"),
        "image_centers": _constant(tree, "IMAGE_CENTERS"),
        "detector_centers": _constant(tree, "DETECTOR_CENTERS"),
        "side_ratios": _constant(tree, "SIDE_RATIOS"),
        "detector_placeholder": _table(tree, "DETECTOR_PLACEHOLDER_EXPECTED"),
        "image_placeholder": _table(tree, "IMAGE_PLACEHOLDER_EXPECTED"),
    }
    with open(DST, "w") as f:
        json.dump(data, f, indent=1)
    n = sum(len(v) for k in ("detector_placeholder", "image_placeholder") for v in data[k].values())
    print(DST, n, "cases")
