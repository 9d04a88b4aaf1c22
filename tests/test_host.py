"""Host front end and the C-ABI library (no GPU work)."""
import os
import re
import tempfile

import numpy as np
import pytest

import mdff20_amd as m
from conftest import GOLDEN, ROOT, load_case

HEADERS = ["mdff_gpu.h", "mdff_host.h", "mdff_batch.h", "mdff_rdf.h"]


def declared_functions():
    names = []
    for h in HEADERS:
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for mt in re.finditer(r"^\s*[A-Za-z_][\w \*]*?\b(mdff_\w+)\s*\(", txt, flags=re.M):
            names.append(mt.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = m.lib()
    names = declared_functions()
    assert len(names) > 30
    for n in names:
        assert hasattr(L, n), n


def test_struct_layouts_match():
    L = m.lib()
    import ctypes
    assert L.mdff_sizeof_params() == ctypes.sizeof(m.Params)
    assert L.mdff_sizeof_thermo() == ctypes.sizeof(m.Thermo)


def test_parse_sio2_control():
    p = m.read_control(os.path.join(GOLDEN, "sio2_102_static", "control"), 2)
    # values of control_SiO.s; the Si-Si entries are missing so stay 0 (bhmftd.c:175)
    assert p.lbhmftd == 1 and p.lcoulombic == 1 and p.lautoES == 1
    assert p.Abhmftd[0][0] == 7902.1862 and p.Abhmftd[0][1] == 1311.5888 and p.Abhmftd[1][0] == 1311.5888
    assert p.Abhmftd[1][1] == 0.0 and p.Bbhmftd[1][1] == 0.0
    assert p.massit[0] == 15.99994 and p.qit[1] == 4.0
    assert p.lpolar[0] == 1 and p.lpolar[1] == 0
    # damping table symmetrised upper -> lower (field.c:209-219)
    assert p.pol_damp_b[0][1][0] == p.pol_damp_b[0][0][1] == 3.66480
    assert p.pol_damp_k[0][0][0] == 4 and p.lpoldamping[0][1][0] == 1
    assert p.integrator == m.NVT_NHCN and p.nhc_n == 3 and p.nhc_yosh_order == 3 and p.nhc_mults == 2
    assert p.algo_pim == m.PIM_SCF and p.extrapolate_order == 4
    assert p.min_scf_pim_iter == 4 and p.max_scf_pim_iter == 50 and p.conv_tol_ind == 1e-5
    assert p.cutshortrange == 5.0 and p.skindiff == 1.0


def test_parse_defaults_and_comment_keywords():
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "c")
        # "nprint" inside a comment is followed by a keyword: the read fails
        # and consumes nothing (fscanf semantics)
        open(f, "w").write("nprint 7 // info each nprint\nfprint 3\nlnmlj true\nsigmalj 2.0\nepsw 1e-6 0.5\n")
        p = m.read_control(f, 1)
    assert p.nprint == 7 and p.fprint == 3 and p.lnmlj == 1
    assert p.sigmalj[0][0] == 2.0 and p.epslj[0][0] == 0.010323576 and p.plj[0][0] == 6.0
    assert p.epsw == 0.5  # read_field tests "epsw" twice (field.c:129-134)
    assert p.lverletL == 1 and p.skindiff == 0.15 and p.max_scf_pim_iter == 100 and p.Fposff == 2


def test_parse_rejects_bad_values():
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "c")
        open(f, "w").write("lverletL maybe\n")
        with pytest.raises(m.MdffError):
            m.read_control(f, 1)
        open(f, "w").write("integrator nve-xx\n")
        with pytest.raises(m.MdffError):
            m.read_control(f, 1)


def test_posff_read_wraps_and_roundtrips():
    cfg = m.read_posff(os.path.join(GOLDEN, "sio2_102_static", "POSFF"))
    assert cfg.nion == 102 and cfg.atypit == ["O", "Si"] and cfg.nionit == [68, 34]
    L = cfg.A[0, 0]
    s = cfg.r / L
    assert np.all(s >= -0.5) and np.all(s <= 0.5)
    assert list(cfg.typia[:68]) == [0] * 68 and list(cfg.typia[68:]) == [1] * 34
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "CONTFF")
        m.write_config(f, cfg)
        back = m.read_posff(f)
    assert np.max(np.abs(back.r - cfg.r)) < 1e-11  # %19.12f columns


def test_direct_coordinates():
    cfg = m.read_posff(os.path.join(GOLDEN, "lj_fcc_static", "POSFF"))
    assert cfg.nion == 256
    assert np.allclose(cfg.r[:, 1], [0.125 * 20.993868, 0.125 * 20.993868, 0.0])


@pytest.mark.parametrize("case", ["melt216_nve", "pim192_nvt", "sio2_102_nvt"])
def test_initial_velocities_bitwise(case):
    p, cfg, ref = load_case(case)
    mass = np.array([p.massit[t] for t in cfg.typia])
    v = m.init_velocities_host(p, mass)
    assert np.array_equal(v, np.stack([ref["v0x"], ref["v0y"], ref["v0z"]]))


def test_tt_damping_closed_forms():
    # order 0: f = 1 - c exp(-b r), f' = c b exp(-b r)
    f, fd = m.tt_damping(2.0, 1.5, 1.3, 0)
    assert f == pytest.approx(1 - 1.5 * np.exp(-2.6), rel=1e-15)
    assert fd == pytest.approx(1.5 * 2.0 * np.exp(-2.6), rel=1e-15)
    # f -> 1 far away, f' = df/dr by central differences
    for k in range(0, 9):
        f, _ = m.tt_damping(2.0, 1.0, 40.0, k)
        assert abs(1 - f) < 1e-20
        h = 1e-5
        fp, _ = m.tt_damping(3.1, 1.2, 1.7 + h, k)
        fm, _ = m.tt_damping(3.1, 1.2, 1.7 - h, k)
        _, fd = m.tt_damping(3.1, 1.2, 1.7, k)
        assert (fp - fm) / (2 * h) == pytest.approx(fd, rel=1e-7)
    # c = 1, order n: 1 - f = exp(-x) sum_{k<=n} x^k/k!  (incomplete gamma)
    from math import exp, factorial
    x = 4.2
    f, _ = m.tt_damping(x, 1.0, 1.0, 6)
    assert 1 - f == pytest.approx(exp(-x) * sum(x ** k / factorial(k) for k in range(7)), rel=1e-14)
