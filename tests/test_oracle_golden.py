"""The oracle (CPU restatement, oracle/mdff_oracle.c) against the golden
vectors of tests/golden/make_fixtures.py.  Cases whose ref.npz records source
"reference" were written by the reference's own routines (no polarisable
ions: twocharges, onedipole, twodipoles, the melt216 family, lj_fcc) and pin
the oracle; the polarisable cases were written by the oracle itself and are
regression vectors (the polarisable path is pinned by test_oracle_pins.py)."""
import numpy as np
import pytest

from conftest import SCALAR_KEYS, load_case, ref_state, rel, scalars_of

Oracle = pytest.importorskip("oracle.oracle").Oracle

TOL = 1e-10
STATIC = ["twocharges", "onedipole", "twodipoles", "twocharges_twopol", "sio2_102_static", "lj_fcc_static",
          "pim192_static", "pim192_fixed", "sio2_102_ko_static", "melt216_quad_static"]
MD = ["melt216_nve", "melt216_raw", "melt216_dip_nve", "melt216_bhmftd_nve", "melt216_nvt", "melt216_lf",
      "lj_fcc_nve", "sio2_102_nvt", "bsio_200_nvt", "pim192_nvt", "sio2_102_ko_nvt", "bsio_200_ko_nvt"]


def n_evals(case, t, o):
    """the golden n_scf counts the induced-dipole real-space Ewald calls of
    an SCF: one per iteration, plus every inner sweep in scfkO"""
    return t["n_scf"] + (o.ko_inner() if "_ko_" in case else 0)


REFERENCE_CASES = {"twocharges", "onedipole", "twodipoles", "lj_fcc_static", "melt216_nve", "melt216_raw",
                   "melt216_dip_nve", "melt216_bhmftd_nve", "melt216_nvt", "melt216_lf", "lj_fcc_nve",
                   "melt216_quad_static"}


@pytest.mark.parametrize("case", STATIC + MD)
def test_fixture_source(case):
    """which fixtures the reference wrote"""
    _, _, ref = load_case(case)
    assert str(ref["source"]) == ("reference" if case in REFERENCE_CASES else "oracle")


@pytest.mark.parametrize("case", STATIC)
def test_static(case):
    p, cfg, ref = load_case(case)
    o = Oracle(p, cfg.typia, cfg.A, cfg.r, threads=4)
    t = o.engforce(0)
    mine, sc = scalars_of(t), ref["scalars"][0]
    for k in range(len(SCALAR_KEYS)):
        if case == "twocharges_twopol" and SCALAR_KEYS[k] == "u_pol":
            continue  # singular polarisability: the reference's LAPACK inverse is not reproduced
        assert abs(mine[k] - sc[k]) <= TOL * max(abs(sc[k]), 1e-12), (SCALAR_KEYS[k], mine[k], sc[k])
    _, _, fr = ref_state(ref, 0)
    assert rel(o.forces(), fr, 1e-4 * np.sqrt(cfg.nion)) < TOL
    if "mu" in ref:
        assert rel(o.dipoles().ravel(), ref["mu"][0], 1e-12) < TOL
    assert rel(np.array(t["tau_coul"]), ref["tau_coul"][0], 1e-12) < TOL
    assert rel(np.array(t["tau_bhmftd"]), ref["tau_bhmftd"][0], 1e-12) < TOL
    assert rel(np.array(t["tau_lj"]), ref["tau_lj"][0], 1e-12) < TOL
    assert n_evals(case, t, o) == int(ref["n_scf"][0])
    a, k, nk = o.ewald()
    assert a == pytest.approx(float(ref["alphaES"]), rel=1e-15) and k == list(ref["kES"]) and nk == int(ref["nk"])


@pytest.mark.parametrize("case", MD)
def test_trajectory(case):
    p, cfg, ref = load_case(case)
    o = Oracle(p, cfg.typia, cfg.A, cfg.r, threads=4)
    o.init_velocities()
    assert rel(o.velocities(), np.stack([ref["v0x"], ref["v0y"], ref["v0z"]])) == 0.0
    for k, s in enumerate(ref["step"]):
        t = o.md_init() if s == 0 else o.md_step(int(s))
        mine, sc = scalars_of(t), ref["scalars"][k]
        err = np.max(np.abs(mine - sc) / np.maximum(np.abs(sc), 1e-9))
        assert err < TOL, (s, err)
        assert n_evals(case, t, o) == int(ref["n_scf"][k])
        assert t["verlet_updates"] == int(ref["verlet_updates"][k])
        if s in list(ref["state_step"]):
            j = list(ref["state_step"]).index(s)
            rr, vr, fr = ref_state(ref, j)
            r, v, f = o.state()
            assert rel(r, rr) < TOL and rel(v, vr) < TOL and rel(f, fr, 1e-4 * np.sqrt(cfg.nion)) < TOL
