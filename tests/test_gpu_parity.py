"""HIP engine (FP64 form) against the golden vectors, through the C ABI.
Cases without polarisable ions were written by the reference's own routines
(tests/golden, source "reference"); the polarisable ones by the oracle.  Tolerance: 1e-10 relative (2-norm) on energies,
forces, dipoles and stress; trajectories within the horizon the golden
vectors cover."""
import numpy as np
import pytest

import mdff20_amd as m
from conftest import SCALAR_KEYS, load_case, ref_state, rel, scalars_of

pytestmark = pytest.mark.gpu

TOL = 1e-10

STATIC = ["twocharges", "onedipole", "twodipoles", "twocharges_twopol", "sio2_102_static", "lj_fcc_static",
          "pim192_static", "pim192_fixed", "sio2_102_ko_static", "melt216_quad_static"]
MD = ["melt216_nve", "melt216_raw", "melt216_dip_nve", "melt216_bhmftd_nve", "melt216_nvt", "melt216_lf",
      "lj_fcc_nve", "sio2_102_nvt", "bsio_200_nvt", "pim192_nvt", "sio2_102_ko_nvt", "bsio_200_ko_nvt"]


def _engine(p, cfg, precision=m.FP64):
    return m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=precision)


class _Evals:
    """the golden n_scf counts the induced-dipole real-space Ewald calls of
    an SCF: one per iteration, plus every inner sweep in scfkO (the engine
    counts those in its scf_inner counter)"""

    def __init__(self, case, e):
        self.ko, self.e = "_ko_" in case, e
        self.c0 = e.counters()["scf_inner"]

    def __call__(self, t):
        c1 = self.e.counters()["scf_inner"]
        n = t["n_scf"] + (c1 - self.c0 if self.ko else 0)
        self.c0 = c1
        return n


@pytest.mark.parametrize("case", STATIC)
def test_static_vs_golden(case):
    print(f"[gpu] static {case}", flush=True)
    p, cfg, ref = load_case(case)
    with _engine(p, cfg) as e:
        ev = _Evals(case, e)
        t = e.engforce(0)
        sc = ref["scalars"][0]
        mine = scalars_of(t)
        keys = list(range(len(SCALAR_KEYS)))
        if case == "twocharges_twopol":
            keys.remove(5)  # u_pol of a singular polarisability tensor
        for k in keys:
            assert abs(mine[k] - sc[k]) <= TOL * max(abs(sc[k]), 1e-12), (SCALAR_KEYS[k], mine[k], sc[k])
        _, _, fr = ref_state(ref, 0)
        floor = 1e-4 * np.sqrt(cfg.nion)
        assert rel(e.forces(), fr, floor) < TOL
        if "mu" in ref:
            assert rel(e.dipoles().ravel(), ref["mu"][0], 1e-12) < TOL
        assert rel(np.array(t["tau_coul"]), ref["tau_coul"][0], 1e-12) < TOL
        assert rel(np.array(t["tau_bhmftd"]), ref["tau_bhmftd"][0], 1e-12) < TOL
        assert ev(t) == int(ref["n_scf"][0])
        a, k, nk = e.ewald()
        assert abs(a - float(ref["alphaES"])) <= 1e-15 * max(a, 1.0) and k == list(ref["kES"]) and nk == int(ref["nk"])


@pytest.mark.parametrize("case", MD)
def test_md_vs_golden(case):
    print(f"[gpu] md {case}", flush=True)
    p, cfg, ref = load_case(case)
    with _engine(p, cfg) as e:
        e.init_velocities()
        v0 = e.velocities()
        assert rel(v0, np.stack([ref["v0x"], ref["v0y"], ref["v0z"]])) == 0.0
        worst = 0.0
        ev = _Evals(case, e)
        for k, s in enumerate(ref["step"]):
            t = e.md_init() if s == 0 else e.md_step(int(s))
            sc = ref["scalars"][k]
            mine = scalars_of(t)
            err = np.max(np.abs(mine - sc) / np.maximum(np.abs(sc), 1e-9))
            worst = max(worst, err)
            ns = ev(t)
            assert ns == int(ref["n_scf"][k]), (s, ns, ref["n_scf"][k])
            assert t["verlet_updates"] == int(ref["verlet_updates"][k]), s
            if s in list(ref["state_step"]):
                j = list(ref["state_step"]).index(s)
                rr, vr, fr = ref_state(ref, j)
                assert rel(e.positions(), rr) < TOL, s
                assert rel(e.velocities(), vr) < TOL, s
                assert rel(e.forces(), fr, 1e-4 * np.sqrt(cfg.nion)) < TOL, s
        assert worst < TOL, worst
