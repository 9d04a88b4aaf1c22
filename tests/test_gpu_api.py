"""C-ABI entry points beyond engforce/MD: multipole_ES with its flag
combinations, get_dipoles, the short-range pair part, and the mixed form's
accuracy against the FP64 form."""
import numpy as np
import pytest

import mdff20_amd as m
from conftest import load_case, rel, scalars_of

pytestmark = pytest.mark.gpu
Oracle = pytest.importorskip("oracle.oracle").Oracle
ES = m.ES

COMBOS = [
    ES["LQ"] | ES["EF"] | ES["DIR"] | ES["REC"] | ES["LDAMP"],
    ES["LD"] | ES["EF"] | ES["DIR"] | ES["REC"],
    ES["LQ"] | ES["LD"] | ES["LDAMP"] | ES["FORCES"] | ES["STRESS"] | ES["DIR"] | ES["REC"],
    ES["LQ"] | ES["LD"] | ES["EF"] | ES["EFG"] | ES["FORCES"] | ES["STRESS"] | ES["DIR"] | ES["REC"],
    ES["LQ"] | ES["LD"] | ES["EF"] | ES["EFG"] | ES["DIR"],
    ES["LQ"] | ES["LD"] | ES["EF"] | ES["FORCES"] | ES["REC"],
]


@pytest.mark.parametrize("flags", COMBOS)
def test_multipole_es_matches_oracle(flags):
    print(f"[gpu] multipole_es flags {flags}", flush=True)
    p, cfg, _ = load_case("pim192_static")
    rng = np.random.default_rng(flags)
    q = np.array([p.qit[t] for t in cfg.typia])
    mu = rng.normal(0, 0.05, (cfg.nion, 3))
    mu[cfg.typia == 1] = 0.0  # cations carry no dipole
    o = Oracle(p, cfg.typia, cfg.A, cfg.r)
    with m.Engine(p, cfg.typia, cfg.A, cfg.r) as e:
        a = e.multipole_es(q, mu, flags)
        b = o.multipole_es(q, mu, flags)
    assert a["u"] == pytest.approx(b["u"], rel=1e-10)
    assert rel(a["ef"], b["ef"], 1e-12) < 1e-10
    assert rel(a["efg"], b["efg"], 1e-12) < 1e-10
    assert rel(a["f"], b["f"], 1e-12) < 1e-10
    if flags & ES["STRESS"]:
        assert rel(a["tau"], b["tau"], 1e-12) < 1e-10
        assert a["pvir"] == pytest.approx(b["pvir"], rel=1e-10)


def test_pair_short_and_dipoles_match_oracle():
    p, cfg, _ = load_case("pim192_static")
    o = Oracle(p, cfg.typia, cfg.A, cfg.r)
    to = o.engforce(0)
    with m.Engine(p, cfg.typia, cfg.A, cfg.r) as e:
        u, pv, tau = e.pair_short()
        assert u == pytest.approx(to["u_bhmftd"], rel=1e-12)
        assert pv == pytest.approx(to["pvir_bhmftd"], rel=1e-12)
        mu, upol = e.dipoles_scf(0)
        assert rel(mu, o.dipoles()) < 1e-10
        assert upol == pytest.approx(to["u_pol"], rel=1e-10)


@pytest.mark.parametrize("lanes", [8, 16, 32, 64])
def test_lane_widths_agree(lanes):
    p, cfg, _ = load_case("pim192_static")
    with m.Engine(p, cfg.typia, cfg.A, cfg.r) as e0, m.Engine(p, cfg.typia, cfg.A, cfg.r) as e1:
        e1.set_tuning(lanes=lanes)
        t0, t1 = e0.engforce(0), e1.engforce(0)
        assert rel(e0.forces(), e1.forces()) < 1e-12
        assert t0["u_tot"] == pytest.approx(t1["u_tot"], rel=1e-13)


@pytest.mark.parametrize("katoms", [0, -1])
@pytest.mark.parametrize("lanes", [8, 16, 32, 64])
def test_lane_widths_agree_mixed(lanes, katoms):
    """the mixed form's pair kernels and its CG steps' pair-tensor field at
    every lane width: only the FP32 summation order differs.  katoms 0: the
    tensor field in the back-projection's extra rows; -1: in the finish
    kernel after the vector-ALU back-projection (lanes per ion apply)"""
    p, cfg, _ = load_case("pim192_static")
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as e0, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as e1:
        e0.set_tuning(katoms=katoms)
        e1.set_tuning(lanes=lanes, katoms=katoms)
        t0, t1 = e0.engforce(0), e1.engforce(0)
        assert rel(e0.forces(), e1.forces()) < 1e-6
        assert rel(e0.dipoles(), e1.dipoles()) < 1e-6
        assert t0["u_tot"] == pytest.approx(t1["u_tot"], rel=1e-8)


@pytest.mark.parametrize("case", ["pim192_static", "bsio_200_nvt"])
def test_fused_kspace_kernels_match_separate(case):
    """the mixed form's fused kernels (the forward's partial sums inside the
    S(k) reduction, the dipole channels folded into rho = i k.S_mu by the
    forward sum itself, the CG direction's pair-tensor field in the back-
    projection's launch) against separate kernels (katoms 4).  The fused
    forward rounds rho to FP32 once per K range where the separate one
    rounds each channel's partial sum: the two agree to the FP32 rounding of
    those partials (measured: dipoles 9e-9, forces up to 2e-7 relative, the
    forces of bsio_200 amplifying the dipoles' difference), far inside the
    mixed form's 1e-5"""
    p, cfg, _ = load_case(case)
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as ef, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as es:
        es.set_tuning(katoms=4)
        tf, ts = ef.engforce(0), es.engforce(0)
        print(f"[gpu] fused vs separate {case}: forces {rel(ef.forces(), es.forces()):.2e}, "
              f"dipoles {rel(ef.dipoles(), es.dipoles()):.2e}", flush=True)
        assert rel(ef.forces(), es.forces()) < 1e-6
        assert rel(ef.dipoles(), es.dipoles()) < 1e-7
        assert tf["u_tot"] == pytest.approx(ts["u_tot"], rel=1e-9)


@pytest.mark.parametrize("katoms", [1, 2, -1])
@pytest.mark.parametrize("case", ["pim192_static", "sio2_102_static", "bsio_200_nvt"])
def test_mixed_form_accuracy(case, katoms):
    """mixed form (FP32 real space, low-precision S(k); anchored-CG steps
    split-FP16 (katoms 1) or FP32 (2) on the matrix cores, or FP32 on the
    vector ALU (-1)) against the FP64 form: force RMS <= 1e-5 relative,
    dipoles <= 1e-5, energy <= 1e-6"""
    print(f"[gpu] mixed {case} katoms={katoms}", flush=True)
    p, cfg, _ = load_case(case)
    p.conv_tol_ind = 1e-9
    p.max_scf_pim_iter = 400
    p.min_scf_pim_iter = 1
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.FP64) as ef, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as em:
        em.set_tuning(katoms=katoms)
        tf, tm = ef.engforce(0), em.engforce(0)
        assert tf["status"] == 0 and tm["status"] == 0
        Ff, Fm = ef.forces(), em.forces()
        frms = np.sqrt(np.mean((Fm - Ff) ** 2)) / np.sqrt(np.mean(Ff ** 2))
        assert frms < 1e-5, frms
        assert rel(em.dipoles(), ef.dipoles()) < 1e-5
        assert tm["u_tot"] == pytest.approx(tf["u_tot"], rel=1e-6)
        assert tm["rmsd"] <= 1e-9


@pytest.mark.parametrize("case", ["pim192_static", "sio2_102_static", "bsio_200_nvt", "melt216_nve"])
@pytest.mark.parametrize("modes", [(m.KS_F16X3, m.KS_F16X3), (m.KS_F16X3, m.KS_F32)])
def test_mixed_form_split_fp16(case, modes):
    """split-FP16 matrix-core sums (power-of-two scaled operands, three
    products) for the static field and force call, with split-FP16 or FP32
    SCF steps: the same bounds against the FP64 form as the FP32 form, and
    within a few 1e-6 of the FP32 matrix-core form"""
    print(f"[gpu] split fp16 {case} {modes}", flush=True)
    p, cfg, _ = load_case(case)
    p.conv_tol_ind = 1e-9
    p.max_scf_pim_iter = 400
    p.min_scf_pim_iter = 1
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.FP64) as ef, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as e32, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as e16:
        e32.set_kspace_modes(m.KS_F32, m.KS_F32)
        e16.set_kspace_modes(*modes)
        tf, t32, t16 = ef.engforce(0), e32.engforce(0), e16.engforce(0)
        assert tf["status"] == 0 and t32["status"] == 0 and t16["status"] == 0
        Ff, F16 = ef.forces(), e16.forces()
        frms = np.sqrt(np.mean((F16 - Ff) ** 2)) / np.sqrt(np.mean(Ff ** 2))
        assert frms < 1e-5, frms
        assert rel(e16.dipoles(), ef.dipoles()) < 1e-5
        assert t16["u_tot"] == pytest.approx(tf["u_tot"], rel=1e-6)
        assert rel(F16, e32.forces()) < 5e-6
        assert t16["rmsd"] <= 1e-9


@pytest.mark.parametrize("case", ["pim192_static", "bsio_200_nvt", "melt216_nve"])
def test_mixed_scf_solvers_reach_the_fp64_fixed_point(case):
    """the mixed form's two SCF solvers: the reference's Jacobi schedule on
    FP64 fields (0) and anchored CG (2: one split-FP16 anchor evaluation by
    default, the field carried along the CG steps): both reach the FP64
    form's dipoles and forces to the SCF tolerance; solver 0 evaluates
    every field in FP64, the anchored solve none; other solver numbers are
    rejected"""
    print(f"[gpu] mixed solvers {case}", flush=True)
    p, cfg, _ = load_case(case)
    p.conv_tol_ind = 1e-10
    p.max_scf_pim_iter = 400
    p.min_scf_pim_iter = 1
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.FP64) as ef, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as ej, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as ea:
        ej.set_scf_solver(0, 6)
        ea.set_scf_solver(2, 6)
        for bad in (1, 3):
            with pytest.raises(m.MdffError):
                ea.set_scf_solver(bad, 6)
        tf, tj, ta = ef.engforce(0), ej.engforce(0), ea.engforce(0)
        assert all(t["status"] == 0 for t in (tf, tj, ta))
        assert all(t["rmsd"] <= 1e-10 for t in (tj, ta))
        muf = ef.dipoles()
        for e in (ej, ea):
            assert rel(e.dipoles(), muf) < 1e-5
            assert rel(e.forces(), ef.forces()) < 1e-5
        assert rel(ea.dipoles(), ej.dipoles()) < 1e-5
        cj, ca = ej.counters(), ea.counters()
        if cj["npol"] > 0:
            assert cj["scf_fp64"] > 0 and cj["scf_inner"] == 0
        # the anchored solve makes no FP64 field evaluation
        assert ca["scf_fp64"] == 0


@pytest.mark.parametrize("case", ["pim192_static", "bsio_200_nvt", "melt216_nve"])
@pytest.mark.parametrize("anchor,inner", [(0, 3), (3, 2), (3, 3)])
def test_anchored_cg_anchor_and_step_modes(case, anchor, inner):
    """anchored CG with the anchor evaluation in FP64 (0) or split FP16 over
    the FP32 tables (3), and FP32 (2) or split-FP16 (3) steps (the
    latter scaled by the channel maxima that the CG kernels write beside p):
    dipoles and forces reach the FP64 form's to the SCF tolerance, and a
    split-FP16 anchor needs no FP64 field evaluation"""
    print(f"[gpu] anchor {case} {anchor} {inner}", flush=True)
    p, cfg, _ = load_case(case)
    p.conv_tol_ind = 1e-9
    p.max_scf_pim_iter = 400
    p.min_scf_pim_iter = 1
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.FP64) as ef, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as ea:
        ea.set_scf_solver(2, 6)
        ea.set_kspace_modes(0, inner)
        ea.set_scf_anchor(anchor)
        tf, ta = ef.engforce(0), ea.engforce(0)
        assert tf["status"] == 0 and ta["status"] == 0
        assert ta["rmsd"] <= 1e-9
        Ff, Fa = ef.forces(), ea.forces()
        frms = np.sqrt(np.mean((Fa - Ff) ** 2)) / np.sqrt(np.mean(Ff ** 2))
        assert frms < 1e-5, frms
        assert rel(ea.dipoles(), ef.dipoles()) < 1e-5
        assert ta["u_tot"] == pytest.approx(tf["u_tot"], rel=1e-6)
        cnt = ea.counters()
        assert cnt["scf_fp64"] == (1 if anchor == 0 and cnt["npol"] > 0 else 0)


@pytest.mark.parametrize("case", ["pim192_static", "bsio_200_nvt"])
def test_mixed_form_on_a_flat_k_box(case):
    """a k box with Kz = 0 (one ez row per ion): a forward-sum block's table
    slice no longer fits its LDS staging area, so the split-FP16 forward sum
    takes the stored point-major operand instead; the mixed form still
    agrees with the FP64 form of the same k box"""
    print(f"[gpu] flat k box {case}", flush=True)
    p, cfg, _ = load_case(case)
    p.lautoES = 0
    p.kES[0] = max(int(p.kES[0]), 3)
    p.kES[1] = max(int(p.kES[1]), 3)
    p.kES[2] = 0
    p.conv_tol_ind = 1e-9
    p.max_scf_pim_iter = 400
    p.min_scf_pim_iter = 1
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.FP64) as ef, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as em:
        tf, tm = ef.engforce(0), em.engforce(0)
        assert tf["status"] == 0 and tm["status"] == 0
        Ff, Fm = ef.forces(), em.forces()
        frms = np.sqrt(np.mean((Fm - Ff) ** 2)) / np.sqrt(np.mean(Ff ** 2))
        assert frms < 1e-5, frms
        if tf["u_pol"]:
            assert rel(em.dipoles(), ef.dipoles()) < 1e-5
        assert tm["u_tot"] == pytest.approx(tf["u_tot"], rel=1e-6)


@pytest.mark.parametrize("case", ["pim192_static", "bsio_200_nvt", "melt216_nve"])
def test_mfma_kspace_matches_vector_kernels_fp32(case):
    """FP32 S(k) forward and back-projection on the matrix cores against the
    FP32 vector-ALU kernels in the mixed form: both are FP32 roundings of the
    FP64 sums (the matrix cores accumulate a K range of up to a few hundred
    ions in FP32 before the FP64 reduction), so they agree to a few 1e-6"""
    print(f"[gpu] mfma vs valu fp32 {case}", flush=True)
    p, cfg, _ = load_case(case)
    p.conv_tol_ind = 1e-10
    p.max_scf_pim_iter = 400
    p.min_scf_pim_iter = 1
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as e0, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as e1:
        e0.set_tuning(katoms=-1)
        e1.set_tuning(katoms=2)
        t0, t1 = e0.engforce(0), e1.engforce(0)
        assert rel(e1.forces(), e0.forces()) < 5e-6
        assert t1["u_coul"] == pytest.approx(t0["u_coul"], rel=1e-7)
        if t0["u_pol"]:
            assert rel(e1.dipoles(), e0.dipoles()) < 5e-6


@pytest.mark.parametrize("case", ["pim192_static", "bsio_200_nvt", "melt216_nve"])
def test_mfma_kspace_matches_vector_kernels(case):
    """S(k) forward and back-projection on the matrix cores (FP64) against
    the vector-ALU kernels"""
    print(f"[gpu] mfma vs valu {case}", flush=True)
    p, cfg, _ = load_case(case)
    with m.Engine(p, cfg.typia, cfg.A, cfg.r) as e0, m.Engine(p, cfg.typia, cfg.A, cfg.r) as e1:
        e0.set_tuning(katoms=-1)
        e1.set_tuning(katoms=1)
        t0, t1 = e0.engforce(0), e1.engforce(0)
        assert rel(e1.forces(), e0.forces()) < 1e-12
        assert t1["u_coul"] == pytest.approx(t0["u_coul"], rel=1e-13)
        assert t1["n_scf"] == t0["n_scf"]
        assert rel(np.array(t1["tau_coul"]), np.array(t0["tau_coul"])) < 1e-12
        if t0["u_pol"]:
            assert rel(e1.dipoles(), e0.dipoles()) < 1e-12


def test_quadrupoles_take_no_part():
    """static quadrupoles (quadit, theta) enter no term of the reference
    (ewald.c:646-688 builds T5 and nothing reads it; ewald.c:379-491 never
    uses theta): results with and without them are identical"""
    p, cfg, _ = load_case("pim192_static")
    q = m.Params()
    import ctypes as C
    C.pointer(q)[0] = p
    for a in range(3):
        for b in range(3):
            q.quadit[0][a][b] = 0.3 if a == b else 0.1
    q.quadit[0][2][2] = -0.6
    with m.Engine(p, cfg.typia, cfg.A, cfg.r) as e0, m.Engine(q, cfg.typia, cfg.A, cfg.r) as e1:
        t0, t1 = e0.engforce(0), e1.engforce(0)
        assert np.array_equal(e0.forces(), e1.forces())
        assert t0["u_tot"] == t1["u_tot"]
        qch = np.array([p.qit[t] for t in cfg.typia])
        mu = e0.dipoles()
        theta = np.tile(np.array([[0.3, 0.1, 0.1], [0.1, 0.3, 0.1], [0.1, 0.1, -0.6]]), (cfg.nion, 1, 1))
        flags = ES["LQ"] | ES["LD"] | ES["EF"] | ES["EFG"] | ES["FORCES"] | ES["STRESS"] | ES["DIR"] | ES["REC"]
        a = e0.multipole_es(qch, mu, flags)
        b = e0.multipole_es(qch, mu, flags | ES["LT"], theta=theta)
        for k in ("ef", "efg", "f", "tau"):
            assert np.array_equal(a[k], b[k]), k
        assert a["u"] == b["u"]


def test_fp64_anchored_cg_reaches_the_oracle_fixed_point():
    """the FP64 form with the anchored CG solver (mdff_set_scf_solver 2):
    the same fixed point as the reference's Jacobi SCF in the oracle, in
    fewer iterations"""
    import ctypes as C
    p, cfg, _ = load_case("pim192_static")
    q = m.Params()
    C.pointer(q)[0] = p
    q.conv_tol_ind = 1e-11
    q.max_scf_pim_iter = 1000
    o = Oracle(q, cfg.typia, cfg.A, cfg.r)
    to = o.engforce(0)
    with m.Engine(q, cfg.typia, cfg.A, cfg.r) as e:
        e.set_scf_solver(2, 0)
        t = e.engforce(0)
        print(f"[gpu] FP64 anchored CG: {t['n_scf']} iterations, Jacobi {to['n_scf']}; "
              f"dipoles {rel(e.dipoles(), o.dipoles()):.2e}, forces {rel(e.forces(), o.forces()):.2e}", flush=True)
        assert t["n_scf"] < to["n_scf"]
        assert t["rmsd"] <= 1e-11
        assert rel(e.dipoles(), o.dipoles()) < 1e-9
        assert rel(e.forces(), o.forces()) < 1e-9
        assert t["u_tot"] == pytest.approx(to["u_tot"], rel=1e-12)


def test_cell_grid_list_equals_all_pairs():
    """the cell-grid neighbour build against the all-pairs scan on the
    1536-ion PIM melt (5 x 5 x 5 cells, asked for): both leave each row in ascending
    neighbour index, so 20 FP64 NVT steps give bitwise the same energies,
    forces, rebuilds and SCF counts"""
    import bench
    L, pos, typ = bench.system(512, 1536)
    r = bench.equilibrate(L, pos, typ, 100, kmax=10)
    p = bench.params_of(bench.control(True, alpha=0.3, kmax=10, rc=9.0))
    A = np.diag([L, L, L])
    with m.Engine(p, typ, A, r) as ec, m.Engine(p, typ, A, r) as ea:
        assert not ec.list_info()["cells"]   # 125 cells, all in the stencil: all pairs by default
        ec.set_list_method(2)
        ea.set_list_method(1)
        ec.check_verlet()
        ic, ia = ec.list_info(), ea.list_info()
        print(f"[gpu] cell grid {ic}, all pairs {ia}", flush=True)
        assert ic["cells"] and ic["nc"] == [5, 5, 5] and ic["reach"] == [2, 2, 2] and not ia["cells"]
        for e in (ec, ea):
            e.init_velocities()
        tc, ta = ec.md_init(), ea.md_init()
        for s in range(1, 21):
            tc, ta = ec.md_step(s), ea.md_step(s)
            assert tc["verlet_updates"] == ta["verlet_updates"] and tc["n_scf"] == ta["n_scf"], s
            assert tc["u_tot"] == ta["u_tot"], s
        assert tc["verlet_updates"] >= 2
        assert np.array_equal(ec.forces(), ea.forces())
        assert np.array_equal(ec.positions(), ea.positions())
        tcell, tall = ec.time_kernel(7, iters=5), ea.time_kernel(7, iters=5)
        print(f"[gpu] list rebuild at N=1536: cells {tcell * 1e3:.1f} us, all pairs {tall * 1e3:.1f} us", flush=True)


def test_nonsymmetric_polarisability():
    """a polarisability tensor that is not symmetric, read transposed by the
    SCF (mu = pol^T E, pim.c:637).  The rmsd of get_rmsd_mu (pim.c:431-449)
    divides by the diagonal entries only, so it does not vanish at the fixed
    point: the loop runs to max_scf_pim_iter and reports it unconverged, as
    the reference does.  The FP64 form equals the oracle (1e-10, the same
    count); the mixed form cannot run the anchored CG (its system would not
    be symmetric) and takes the reference's Jacobi schedule on FP64 fields,
    within 1e-5 of the FP64 form"""
    import ctypes as C
    p, cfg, _ = load_case("pim192_static")
    q = m.Params()
    C.pointer(q)[0] = p
    q.conv_tol_ind = 1e-11
    q.max_scf_pim_iter = 80
    pt = [t for t in range(q.ntype) if q.polit[t][0][0] != 0.0][0]
    q.polit[pt][0][1] += 0.1 * q.polit[pt][0][0]
    q.polit[pt][2][0] -= 0.05 * q.polit[pt][0][0]
    o = Oracle(q, cfg.typia, cfg.A, cfg.r)
    to = o.engforce(0)
    with m.Engine(q, cfg.typia, cfg.A, cfg.r) as ef, \
            m.Engine(q, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as em:
        tf, tm = ef.engforce(0), em.engforce(0)
        print(f"[gpu] non-symmetric pol: FP64 {tf['n_scf']} / oracle {to['n_scf']} iterations, "
              f"mixed {tm['n_scf']}; FP64 dipoles {rel(ef.dipoles(), o.dipoles()):.2e}, "
              f"mixed forces {rel(em.forces(), ef.forces()):.2e}", flush=True)
        assert tf["status"] == m.ERR_SCF and tm["status"] == m.ERR_SCF and to["status"] == m.ERR_SCF
        assert tf["n_scf"] == to["n_scf"] == tm["n_scf"]
        assert rel(ef.dipoles(), o.dipoles()) < 1e-10
        assert rel(ef.forces(), o.forces()) < 1e-10
        assert tf["u_tot"] == pytest.approx(to["u_tot"], rel=1e-10)
        assert rel(em.dipoles(), ef.dipoles()) < 1e-5
        assert rel(em.forces(), ef.forces()) < 1e-5
        assert tm["u_tot"] == pytest.approx(tf["u_tot"], rel=1e-6)
        cm = em.counters()
        assert cm["scf_fp64"] > 0 and cm["scf_inner"] == 0


def test_kbox_beyond_the_matrix_core_tiles():
    """kES = (18, 8, 8): the force call's forward sum (4 channels, 2 x 4 x 19
    = 152 rows) is wider than the matrix-core tiles take (144) and every
    back-projection has more nx planes (19) than its staged ex table holds
    (17), so those run on the vector ALU while the SCF's forward sums (3
    channels, 114 rows) stay on the matrix cores: FP64 equals the oracle
    (1e-10, the same SCF count) and the mixed form stays within 1e-5 of
    FP64"""
    import ctypes as C
    p, cfg, _ = load_case("pim192_static")
    q = m.Params()
    C.pointer(q)[0] = p
    q.lautoES = 0
    q.kES[0], q.kES[1], q.kES[2] = 18, 8, 8
    o = Oracle(q, cfg.typia, cfg.A, cfg.r)
    to = o.engforce(0)
    with m.Engine(q, cfg.typia, cfg.A, cfg.r) as ef, \
            m.Engine(q, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as em:
        tf, tm = ef.engforce(0), em.engforce(0)
        print(f"[gpu] k box 18x8x8: FP64 vs oracle forces {rel(ef.forces(), o.forces()):.2e}, "
              f"mixed vs FP64 {rel(em.forces(), ef.forces()):.2e}", flush=True)
        assert tf["status"] == 0 and tm["status"] == 0
        assert tf["n_scf"] == to["n_scf"]
        assert rel(ef.forces(), o.forces()) < 1e-10
        assert rel(ef.dipoles(), o.dipoles()) < 1e-10
        assert tf["u_tot"] == pytest.approx(to["u_tot"], rel=1e-10)
        assert rel(np.array(tf["tau_coul"]), np.array(to["tau_coul"])) < 1e-10
        assert rel(em.forces(), ef.forces()) < 1e-5
        assert rel(em.dipoles(), ef.dipoles()) < 1e-5
        assert tm["u_tot"] == pytest.approx(tf["u_tot"], rel=1e-6)


def _sheared(cfg, shear=((1, 0, 0.15), (2, 0, 0.10), (2, 1, 0.05))):
    """the fixture's ions at the same fractional coordinates in a triclinic
    cell: lattice vector i gains shear x L along axis j"""
    A = np.array(cfg.A, dtype=np.float64)
    L = A[0][0]
    s = np.linalg.solve(A.T, np.asarray(cfg.r))
    A2 = A.copy()
    for i, j, f in shear:
        A2[i][j] += f * L
    return A2, A2.T @ s


def test_triclinic_cell():
    """a triclinic cell (reciprocal basis with off-diagonal terms in every
    k-vector, kardir/dirkar cell.c:125-161): FP64 equals the oracle (1e-10,
    the same SCF count) and the mixed form stays within 1e-5 of FP64"""
    p, cfg, _ = load_case("pim192_static")
    A2, r2 = _sheared(cfg)
    o = Oracle(p, cfg.typia, A2, r2)
    to = o.engforce(0)
    with m.Engine(p, cfg.typia, A2, r2) as ef, m.Engine(p, cfg.typia, A2, r2, precision=m.MIXED) as em:
        tf, tm = ef.engforce(0), em.engforce(0)
        print(f"[gpu] triclinic: FP64 vs oracle forces {rel(ef.forces(), o.forces()):.2e}, "
              f"mixed vs FP64 {rel(em.forces(), ef.forces()):.2e}", flush=True)
        assert tf["status"] == 0 and tm["status"] == 0 and to["status"] == 0
        assert tf["n_scf"] == to["n_scf"]
        assert rel(ef.forces(), o.forces()) < 1e-10
        assert rel(ef.dipoles(), o.dipoles()) < 1e-10
        assert tf["u_tot"] == pytest.approx(to["u_tot"], rel=1e-10)
        assert rel(np.array(tf["tau_coul"]), np.array(to["tau_coul"])) < 1e-10
        assert rel(em.forces(), ef.forces()) < 1e-5
        assert rel(em.dipoles(), ef.dipoles()) < 1e-5
        assert tm["u_tot"] == pytest.approx(tf["u_tot"], rel=1e-6)


def test_neighbour_list_grows_on_overflow():
    """ions packed into a fifth of a long box: the list capacity estimated
    from the mean density (1.3 x the mean count + 64, rounded up to 64) is
    too small, so the first build overflows and the rows grow (build_list);
    three NVT steps then equal the oracle's (1e-10, the same SCF counts and
    rebuilds)"""
    p, cfg, _ = load_case("bsio_200_nvt")
    A = np.asarray(cfg.A, dtype=np.float64)
    r = np.asarray(cfg.r)
    typ = np.asarray(cfg.typia)
    rs, ts = [], []
    for ix in range(2):
        for iy in range(2):
            for iz in range(2):
                rs.append(r + (ix * A[0] + iy * A[1] + iz * A[2])[:, None])
                ts.append(typ)
    rr, tt = np.concatenate(rs, axis=1), np.concatenate(ts)
    order = np.argsort(tt, kind="stable")
    rr, tt = rr[:, order], tt[order].astype(np.int32)
    A2 = 2.0 * A
    A2[0] *= 5.0
    o = Oracle(p, tt, A2, rr)
    rc = max(p.cutshortrange, p.cutlongrange) + p.skindiff
    est = 4.0 / 3.0 * np.pi * rc ** 3 * len(tt) / abs(np.linalg.det(A2))
    cap0 = (int(est * 1.3) + 64 + 63) // 64 * 64
    with m.Engine(p, tt, A2, rr) as e:
        e.init_velocities()
        o.init_velocities()
        assert rel(e.velocities(), o.velocities()) == 0.0
        for s in range(4):
            t = e.md_init() if s == 0 else e.md_step(s)
            to = o.md_init() if s == 0 else o.md_step(s)
            if s == 0:
                cap1 = e.list_info()["cap"]
                print(f"[gpu] list capacity {cap0} -> {cap1}", flush=True)
                assert cap1 > cap0
            assert t["n_scf"] == to["n_scf"], s
            assert t["verlet_updates"] == to["verlet_updates"], s
            assert rel(scalars_of(t), scalars_of(to), 1e-12) < 1e-10, s
            assert rel(e.forces(), o.forces()) < 1e-10, s
    o.close()


@pytest.mark.parametrize("case", ["lj_fcc_static", "sio2_102_ko_static", "melt216_bhmftd_nve", "melt216_dip_nve",
                                  "melt216_quad_static"])
def test_mixed_form_other_models(case):
    """the mixed form on the models the accuracy tests above leave out:
    NMLJ pairs (lj_fcc: FP64 in both forms, so equal), scfkO (FP64 SCF,
    low-precision static field),
    undamped BHMFTD, static dipoles without polarisation, static
    quadrupoles: forces within 1e-5 (RMS, relative) of the FP64 form,
    energy within 1e-6"""
    print(f"[gpu] mixed other models {case}", flush=True)
    p, cfg, _ = load_case(case)
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.FP64) as ef, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as em:
        tf, tm = ef.engforce(0), em.engforce(0)
        assert tf["status"] == 0 and tm["status"] == 0
        Ff, Fm = ef.forces(), em.forces()
        frms = np.sqrt(np.mean((Fm - Ff) ** 2)) / np.sqrt(np.mean(Ff ** 2))
        print(f"  forces {frms:.2e}, energy {abs(tm['u_tot'] - tf['u_tot']) / abs(tf['u_tot']):.2e}, "
              f"n_scf {tf['n_scf']} / {tm['n_scf']}", flush=True)
        assert frms < 1e-5, frms
        assert tm["u_tot"] == pytest.approx(tf["u_tot"], rel=1e-6)
        if tf["u_pol"]:
            assert rel(em.dipoles(), ef.dipoles()) < 1e-5


@pytest.mark.parametrize("chain", [2, 5, 10])
def test_nhc_chain_lengths(chain):
    """Nose-Hoover chains of other lengths than the fixtures' 3 (the chain
    kernel is instantiated for up to 4, 8 and 16 thermostats): 20 NVT steps
    equal the oracle's chain_nhcn (integration.c:189-294), 1e-10 on every
    thermo scalar, positions and velocities; chains shorter than 2 (the
    reference's sweep reads thermostat nhc_n - 2, integration.c:251) or
    longer than 16 are rejected at creation"""
    import ctypes as C
    p, cfg, _ = load_case("melt216_nvt")
    q = m.Params()
    C.pointer(q)[0] = p
    for bad in (1, 17):
        q.nhc_n = bad
        with pytest.raises(m.MdffError):
            m.Engine(q, cfg.typia, cfg.A, cfg.r)
    q.nhc_n = chain
    o = Oracle(q, cfg.typia, cfg.A, cfg.r)
    with m.Engine(q, cfg.typia, cfg.A, cfg.r) as e:
        e.init_velocities()
        o.init_velocities()
        for s in range(21):
            t = e.md_init() if s == 0 else e.md_step(s)
            to = o.md_init() if s == 0 else o.md_step(s)
            assert rel(scalars_of(t), scalars_of(to), 1e-12) < 1e-10, (chain, s)
        print(f"[gpu] NHC-{chain}: e_nvt {t['e_nvt']:.10e} / {to['e_nvt']:.10e}", flush=True)
        assert rel(e.positions(), o.positions()) < 1e-10
        assert rel(e.velocities(), o.velocities()) < 1e-10
    o.close()


@pytest.mark.parametrize("order,mults", [(1, 1), (3, 2), (5, 1), (7, 3), (9, 1)])
def test_nhc_yoshida_orders(order, mults):
    """the Yoshida-Suzuki weights of chain_nhcn (integration.c:205-236:
    orders 1, 3, 5, 7, and 9, whose weights the reference leaves at zero)
    with nhc_mults sub-steps: 20 NVT steps equal the oracle, 1e-10 on every
    thermo scalar, positions and velocities"""
    import ctypes as C
    p, cfg, _ = load_case("melt216_nvt")
    q = m.Params()
    C.pointer(q)[0] = p
    q.nhc_yosh_order = order
    q.nhc_mults = mults
    o = Oracle(q, cfg.typia, cfg.A, cfg.r)
    with m.Engine(q, cfg.typia, cfg.A, cfg.r) as e:
        e.init_velocities()
        o.init_velocities()
        for s in range(21):
            t = e.md_init() if s == 0 else e.md_step(s)
            to = o.md_init() if s == 0 else o.md_step(s)
            assert rel(scalars_of(t), scalars_of(to), 1e-12) < 1e-10, (order, mults, s)
        print(f"[gpu] Yoshida {order} x {mults}: e_nvt {t['e_nvt']:.10e} / {to['e_nvt']:.10e}", flush=True)
        assert rel(e.positions(), o.positions()) < 1e-10
        assert rel(e.velocities(), o.velocities()) < 1e-10
    o.close()


@pytest.mark.parametrize("order", [0, 1, 3, 4, 5])
def test_aspc_orders_md(order):
    """the ASPC predictor-corrector (pim.c:314-424) at the orders the
    fixtures leave out (they use 2), with its ramp over the first steps and
    the history shift: 12 PIM NVT steps equal the oracle, the same SCF
    count every step, 1e-10 on every thermo scalar and the dipoles"""
    import ctypes as C
    p, cfg, _ = load_case("pim192_nvt")
    q = m.Params()
    C.pointer(q)[0] = p
    q.extrapolate_order = order
    o = Oracle(q, cfg.typia, cfg.A, cfg.r)
    with m.Engine(q, cfg.typia, cfg.A, cfg.r) as e:
        e.init_velocities()
        o.init_velocities()
        ns = []
        for s in range(13):
            t = e.md_init() if s == 0 else e.md_step(s)
            to = o.md_init() if s == 0 else o.md_step(s)
            ns.append(t["n_scf"])
            assert t["n_scf"] == to["n_scf"], (order, s, t["n_scf"], to["n_scf"])
            assert rel(scalars_of(t), scalars_of(to), 1e-12) < 1e-10, (order, s)
        print(f"[gpu] ASPC order {order}: SCF iterations {ns}", flush=True)
        assert rel(e.dipoles(), o.dipoles()) < 1e-10
        assert rel(e.positions(), o.positions()) < 1e-10
    o.close()


@pytest.mark.parametrize("precision", [m.FP64, m.MIXED])
def test_ion_order_invariance(precision):
    """the ions in a random order (types interleaved instead of in blocks):
    energies, the SCF count and the per-ion forces and dipoles (permuted
    back) equal those of the fixture's order, to 1e-10 in FP64 and 1e-5 in
    the mixed form (whose sums change their order)"""
    p, cfg, _ = load_case("bsio_200_nvt")
    typ = np.asarray(cfg.typia, dtype=np.int32)
    r = np.asarray(cfg.r)
    perm = np.random.default_rng(11).permutation(len(typ))
    tol = 1e-10 if precision == m.FP64 else 1e-5
    with m.Engine(p, typ, cfg.A, r, precision=precision) as e0, \
            m.Engine(p, typ[perm].copy(), cfg.A, r[:, perm].copy(), precision=precision) as e1:
        t0, t1 = e0.engforce(0), e1.engforce(0)
        print(f"[gpu] ion order ({'FP64' if precision == m.FP64 else 'mixed'}): forces "
              f"{rel(e1.forces(), e0.forces()[:, perm]):.2e}, SCF {t0['n_scf']} / {t1['n_scf']}", flush=True)
        assert t0["status"] == 0 and t1["status"] == 0
        if precision == m.FP64:
            assert t1["n_scf"] == t0["n_scf"]
        assert t1["u_tot"] == pytest.approx(t0["u_tot"], rel=tol)
        assert rel(e1.forces(), e0.forces()[:, perm]) < tol
        assert rel(e1.dipoles(), e0.dipoles()[perm]) < tol


def test_cell_grid_list_triclinic():
    """the cell-grid neighbour build in a triclinic cell (cells along the
    fractional axes, stencil reach from the perpendicular widths) against
    the all-pairs scan: the 1536-ion melt sheared, 10 FP64 NVT steps give
    bitwise the same energies, forces, rebuilds and SCF counts"""
    import bench
    L, pos, typ = bench.system(512, 1536)
    r = bench.equilibrate(L, pos, typ, 60, kmax=10)
    p = bench.params_of(bench.control(True, alpha=0.3, kmax=10, rc=9.0))
    A = np.diag([L, L, L])
    s = np.linalg.solve(A.T, r)
    A2 = A.copy()
    A2[1][0] += 0.12 * L
    A2[2][0] += 0.08 * L
    A2[2][1] -= 0.06 * L
    r2 = A2.T @ s
    with m.Engine(p, typ, A2, r2) as ec, m.Engine(p, typ, A2, r2) as ea:
        ec.set_list_method(2)
        ea.set_list_method(1)
        ec.check_verlet()
        ic = ec.list_info()
        print(f"[gpu] triclinic cell grid {ic}", flush=True)
        assert ic["cells"]
        for e in (ec, ea):
            e.init_velocities()
        tc, ta = ec.md_init(), ea.md_init()
        for st in range(1, 11):
            tc, ta = ec.md_step(st), ea.md_step(st)
            assert tc["verlet_updates"] == ta["verlet_updates"] and tc["n_scf"] == ta["n_scf"], st
            assert tc["u_tot"] == ta["u_tot"], st
        assert np.array_equal(ec.forces(), ea.forces())
        assert np.array_equal(ec.positions(), ea.positions())


@pytest.mark.parametrize("cw", [1, 4])
@pytest.mark.parametrize("case", ["pim192_static", "sio2_102_static"])
def test_kspace_shape_column_tiles(case, cw):
    """the forward sum's other column-tile widths (1 or 4 tiles of 16 box
    points per wavefront instead of 2; the table slice a block stages grows
    with them) and fixed block counts: the mixed form within 1e-6 of its
    default shape and 1e-5 of the FP64 form"""
    print(f"[gpu] k-space shape {case} cw={cw}", flush=True)
    p, cfg, _ = load_case(case)
    p.conv_tol_ind = 1e-9   # the fixtures' looser tolerances separate the forms' SCF end points more
    p.max_scf_pim_iter = 400
    p.min_scf_pim_iter = 1
    with m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.FP64) as ef, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as e2, \
            m.Engine(p, cfg.typia, cfg.A, cfg.r, precision=m.MIXED) as ew:
        ew.set_kspace_shape(cw64=cw, cw32=cw, fwd_blocks=300, back_blocks=200)
        tf, t2, tw = ef.engforce(0), e2.engforce(0), ew.engforce(0)
        assert tf["status"] == 0 and t2["status"] == 0 and tw["status"] == 0
        assert rel(ew.forces(), e2.forces()) < 1e-6
        assert rel(ew.dipoles(), e2.dipoles()) < 1e-6
        assert rel(ew.forces(), ef.forces()) < 1e-5
        assert tw["u_tot"] == pytest.approx(tf["u_tot"], rel=1e-6)
