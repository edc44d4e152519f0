import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "radik.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def known_answers():
    import json
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


def load_cases():
    z = np.load(os.path.join(GOLDEN, "cases.npz"))
    ids = sorted({k.split("_")[0] for k in z.files if k.startswith("c")})
    out = []
    for c in ids:
        idx_key = str(z[c + "_idx"])
        out.append({
            "name": str(z[c + "_name"]),
            "x": z["in_" + str(z[c + "_x"])],
            "lens": z[c + "_lens"],
            "k": int(z[c + "_k"]),
            "flags": int(z[c + "_flags"]),
            "idx": z["in_" + idx_key] if idx_key else None,
            "vals": z[c + "_vals"],
            "inds": z[c + "_inds"],
        })
    return out


@pytest.fixture(scope="session")
def golden_cases():
    return load_cases()
