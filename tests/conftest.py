import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmdff_gpu.so)")


def rel(a, b, floor=0.0):
    """relative error in the 2-norm, with a floor on the denominator"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(float(np.linalg.norm(b)), floor, 1e-300)
    return float(np.linalg.norm(a - b)) / den


def load_case(name):
    import mdff20_amd as m
    d = os.path.join(GOLDEN, name)
    cfg = m.read_posff(os.path.join(d, "POSFF"), 2)
    p = m.read_control(os.path.join(d, "control"), cfg.ntype)
    ref = dict(np.load(os.path.join(d, "ref.npz"), allow_pickle=False))
    return p, cfg, ref


def ref_state(ref, j):
    r = np.stack([ref["rx"][j], ref["ry"][j], ref["rz"][j]])
    v = np.stack([ref["vx"][j], ref["vy"][j], ref["vz"][j]])
    f = np.stack([ref["fx"][j], ref["fy"][j], ref["fz"][j]])
    return r, v, f


SCALAR_KEYS = ["e_kin", "e_nvt", "u_lj", "u_bhmftd", "u_coul", "u_pol", "pvir_lj", "pvir_bhmftd", "pvir_coul"]


def scalars_of(t):
    return np.array([t[k] for k in SCALAR_KEYS])


@pytest.fixture(scope="session")
def gpu_available():
    import mdff20_amd as m
    return m.device_count() > 0
