import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden_cases():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def golden_path(name, which):
    return os.path.join(GOLDEN, "%s.%s.dat" % (name, which))


def bc_of(case):
    return [tuple(b) for b in case["bc"]] if case["bc"] else None


def params_of(case):
    """Model parameters that the reference run used (stock = hard-coded in the example)."""
    m, v, p = case["model"], case["variant"], case["params"]
    if m == "ac":
        return p if v == "config" else [1.0, 1.0, 1.0, 1.0, 0.01]  # allen-cahn.cpp:58-62
    if m == "ch":
        return p if v == "config" else [1.0, 1.0, 1.0, 1.0, 0.01]  # cahn-hilliard.cpp:59-63
    if m == "gg":
        return p if v == "config" else [0.01]  # graingrowth.cpp:84
    if m == "zn":
        return [0.01]  # zener.cpp:115
    if m == "ow":
        return [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]  # ostwald.cpp:68-82
    if m == "ga":
        return [0.01, 8.0, 1.0, 1.0]  # anisotropic graingrowth.cpp:105-106, anisotropy.hpp energy/mobility 1
    raise ValueError(m)


def model_of(case):
    """Product model id for a golden case."""
    if case["model"] == "ch":
        return "ch_dw" if case["variant"] == "config" else "ch"
    return {"zn": "zener", "ow": "ostwald", "ga": "gg_aniso"}.get(case["model"], case["model"])


@pytest.fixture(scope="session")
def cases():
    return golden_cases()
