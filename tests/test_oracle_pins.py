"""Pins of the oracle (oracle/mdff_oracle.c) that do not go through the
reference build: closed forms from the mathematics, a textbook special
function, the published ASPC coefficients, and the reference's own ASPC
predictor run alone.  Each test names the oracle routine it pins.

  engforce_bhmftd           per-pair BHMFT / BHMFTD closed form (the Tang-
                            Toennies f_n as the regularised incomplete gamma
                            function), forces, stress in GPa
  ewald_dir (damping)       the Tang-Toennies damped charge-dipole field,
                            energy and forces of ewald.c:504-541, 700-812
  moment_pola_scf,          two polarisable ions in the field of two charges:
  induced_moment,           the Jacobi iterates of momentpolaSCF form a
  get_rmsd, get_upol        geometric series, which fixes the iteration
                            count, the rmsd and u_pol of pim.c:177-306, 431-449
  aspc_coeffs, aspc_order,  Kolafa's ASPC coefficients (J. Comput. Chem. 25,
  aspc_predict              335 (2004)) and the reference's own predictor
                            (oracle/_ref/ref_aspc)
  moment_pola_scf_ko        its fixed point equals momentpolaSCF's
  ewald_dir/_rec stress,    the charge-only stress equals minus the strain
  engforce_bhmftd stress    derivative of the energy
"""
import ctypes as C
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest
from scipy.special import gammainc, gammaincc

import mdff20_amd as m
from conftest import SCALAR_KEYS, load_case, ref_state, rel, scalars_of

orc = pytest.importorskip("oracle.oracle")
Oracle = orc.Oracle
FL = orc.FLAGS

PRESS_GPA = 0.00624150964712042   # eV/A^3 per GPa as the reference states it (constants.c)


def params_from(text, ntype):
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "control")
        open(f, "w").write(text)
        return m.read_control(f, ntype)


def min_image(d, L):
    return d - L * np.rint(d / L)


# ------------------------------------------------------------ BHMFT(D) pairs
def _bhm_u_du(d, A, B, Cc, D, BD, damped):
    """U(d) = A e^{-Bd} - f6 C/d^6 - f8 D/d^8 and dU/dd; f_n(x) = P(n+1, x),
    the regularised lower incomplete gamma function (1 - e^-x sum x^k/k!),
    f_n'(x) = x^n e^-x / n!"""
    if damped:
        x = BD * d
        f6, f8 = gammainc(7, x), gammainc(9, x)
        g6 = x ** 6 * math.exp(-x) / math.factorial(6) * BD
        g8 = x ** 8 * math.exp(-x) / math.factorial(8) * BD
    else:
        f6 = f8 = 1.0
        g6 = g8 = 0.0
    u = A * math.exp(-B * d) - f6 * Cc / d ** 6 - f8 * D / d ** 8
    du = -A * B * math.exp(-B * d) - Cc * (g6 / d ** 6 - 6 * f6 / d ** 7) - D * (g8 / d ** 8 - 8 * f8 / d ** 9)
    return u, du


@pytest.mark.parametrize("damped", [False, True])
def test_bhmft_pairs_closed_form(damped):
    """engforce_bhmftd (bhmftd.c:117-271): three ions, Coulomb off"""
    A = {(0, 0): 1200.0, (0, 1): 9000.0, (1, 1): 3000.0}
    B = {(0, 0): 3.1, (0, 1): 3.7, (1, 1): 2.9}
    Cc = {(0, 0): 40.0, (0, 1): 25.0, (1, 1): 60.0}
    D = {(0, 0): 90.0, (0, 1): 35.0, (1, 1): 120.0}
    BD = {(0, 0): 2.2, (0, 1): 3.3, (1, 1): 1.9}
    tri = lambda t: " ".join(repr(t[k]) for k in ((0, 0), (0, 1), (1, 1)))
    ctl = (f"lstatic true\nlverletL false\ncutshortrange 9.0\nlcoulombic false\n"
           f"{'lbhmftd' if damped else 'lbhmft'} true\nmassit 16.0\nmassit 28.0\n"
           f"Abhmftd {tri(A)}\nBbhmftd {tri(B)}\nCbhmftd {tri(Cc)}\nDbhmftd {tri(D)}\nBDbhmftd {tri(BD)}\n")
    p = params_from(ctl, 2)
    L = 30.0
    pos = np.array([[0.0, 0.0, 0.0], [-1.1, 1.9, 2.4], [2.1, 0.3, -0.5]])
    typ = np.array([0, 0, 1], np.int32)
    o = Oracle(p, typ, np.diag([L] * 3), pos.T.copy(), threads=1)
    t = o.engforce(0)
    U, F, W = 0.0, np.zeros((3, 3)), np.zeros((3, 3))
    for i in range(3):
        for j in range(i + 1, 3):
            key = tuple(sorted((typ[i], typ[j])))
            r = min_image(pos[i] - pos[j], L)
            d = np.linalg.norm(r)
            u, du = _bhm_u_du(d, A[key], B[key], Cc[key], D[key], BD[key], damped)
            U += u
            f = -du * r / d         # on i from j
            F[i] += f
            F[j] -= f
            W += np.outer(r, f)
    assert t["u_bhmftd"] == pytest.approx(U, rel=1e-13, abs=1e-13)
    assert rel(o.forces().T, F) < 1e-12
    tau = W / (L ** 3 * PRESS_GPA)
    assert rel(np.array(t["tau_bhmftd"]), tau) < 1e-12
    assert t["pvir_bhmftd"] == pytest.approx(np.trace(tau) / 3, rel=1e-12)
    o.close()


# ------------------------------------------------ damped charge-dipole field
DAMP_CTL = """lstatic true
lreduced true
lverletL false
cutlongrange 7.0
lcoulombic true
alphaES 0.6
kES 9 9 9
massit 16.0
massit 28.0
qit -1.0
qit 2.0
polit 1.1 0.0 0.0 0.0 1.1 0.0 0.0 0.0 1.1
polit 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0
lpoldamping 0 0 0 true
pol_damp_b 0 0 0 2.0
pol_damp_c 0 0 0 1.5
pol_damp_k 0 0 0 4
lpoldamping 0 0 1 true
pol_damp_b 0 0 1 1.7
pol_damp_c 0 0 1 1.2
pol_damp_k 0 0 1 3
"""


def _tt(b, c, k, d):
    """Tang-Toennies f = 1 - c e^-bd sum_{j<=k} (bd)^j / j! = 1 - c Q(k+1, bd)"""
    return 1.0 - c * gammaincc(k + 1, b * d)


def _damp_delta(pos, typ, q, mu, p, L, rc):
    """change of field at every ion and of the energy that the damping makes:
    dE_a = -(1 - f_ab(d)) q_b (r_a - r_b)/d^3 with the parameters
    [t_a][t_a][t_b] (ewald.c:504-513); dU = -sum_a mu_a . dE_a"""
    n = len(pos)
    dE = np.zeros((n, 3))
    for a in range(n):
        for b in range(n):
            if a == b:
                continue
            ta, tb = typ[a], typ[b]
            if not (p.lpoldamping[ta][ta][tb] or p.lpoldamping[tb][ta][tb]):
                continue
            r = min_image(pos[a] - pos[b], L)
            d = np.linalg.norm(r)
            if d * d >= rc * rc:
                continue
            f = _tt(p.pol_damp_b[ta][ta][tb], p.pol_damp_c[ta][ta][tb], p.pol_damp_k[ta][ta][tb], d)
            dE[a] += -(1.0 - f) * q[b] * r / d ** 3
    return dE, -float(np.sum(mu * dE))


def test_damped_charge_dipole_field_closed_form():
    """ewald_dir with ldamp: the damped part of the field at the polarisable
    ions, of the charge-dipole energy, and forces = -grad of that energy"""
    p = params_from(DAMP_CTL, 2)
    L, rc = 15.0, 7.0
    rng = np.random.default_rng(5)
    typ = np.array([0, 0, 0, 1, 1], np.int32)
    pos = np.array([[0.0, 0.0, 0.0], [2.2, 0.4, -0.3], [-0.5, 1.8, 1.1], [1.0, -1.3, 0.6], [-1.9, -0.2, -1.4]])
    q = np.array([p.qit[t] for t in typ])
    mu = rng.normal(0.0, 0.3, (5, 3))
    mu[typ == 1] = 0.0
    o = Oracle(p, typ, np.diag([L] * 3), pos.T.copy(), threads=1)
    base = FL["LQ"] | FL["LD"] | FL["EF"] | FL["FORCES"] | FL["DIR"]
    a = o.multipole_es(q, mu, base | FL["LDAMP"])
    b = o.multipole_es(q, mu, base)
    dE, dU = _damp_delta(pos, typ, q, mu, p, L, rc)
    assert np.abs(dE).max() > 1e-3   # the damping is active
    assert np.allclose(a["ef"] - b["ef"], dE, rtol=1e-12, atol=1e-13)
    assert (a["u"] - b["u"]) == pytest.approx(dU, rel=1e-11, abs=1e-13)
    # forces: minus the gradient of the damped energy (central differences)
    h = 1e-6
    dF = (a["f"] - b["f"]).T
    for i in range(5):
        for k in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, k] += h
            pm[i, k] -= h
            up = _damp_delta(pp, typ, q, mu, p, L, rc)[1]
            um = _damp_delta(pm, typ, q, mu, p, L, rc)[1]
            assert dF[i, k] == pytest.approx(-(up - um) / (2 * h), rel=1e-6, abs=1e-8), (i, k)
    o.close()


# ------------------------------------------------ PIM SCF: a geometric series
def _two_ion_case(alpha, d, s, Q, tol, minit=1, maxit=200):
    """polarisable ions at z = -d/2, +d/2 (type 0, no charge), +Q at z = -s
    and -Q at z = +s; every field is along z and equal at the two ions"""
    L = 90.0
    ctl = f"""lstatic true
lreduced true
lverletL false
cutlongrange 44.99
lcoulombic true
alphaES 0.2
kES 36 36 36
massit 1.0
massit 1.0
massit 1.0
qit 0.0
qit {Q!r}
qit {-Q!r}
polit {alpha!r} 0.0 0.0 0.0 {alpha!r} 0.0 0.0 0.0 {alpha!r}
polit 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0
polit 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0
algo_pim scf
algo_extrapolate_dipole aspc
extrapolate_order 2
conv_tol_ind {tol!r}
min_scf_pim_iter {minit}
max_scf_pim_iter {maxit}
"""
    p = params_from(ctl, 3)
    typ = np.array([0, 0, 1, 2], np.int32)
    pos = np.array([[0, 0, -d / 2], [0, 0, d / 2], [0, 0, -s], [0, 0, s]], dtype=float)
    return p, typ, np.diag([L] * 3), pos, L


def test_scf_jacobi_iterates_form_a_geometric_series():
    """momentpolaSCF from zero history (istep 0): iteration 0 leaves mu = 0
    (ASPC order 0, corrector weight 0), iteration k >= 1 sets mu = alpha E.
    With x the common z dipole, E0 the static field (bare field of the
    charges plus the tin-foil term 4 pi M/3V of the charges' moment) and
    g = 2/d^3 + 8 pi/3V, x_k = alpha E0 (1 - (alpha g)^k)/(1 - alpha g) and
    rmsd_k = |E0| (alpha g)^k, so the loop stops at the first k with
    rmsd_k <= tol; u_pol = x^2/alpha (two ions)"""
    alpha, d, s, Q = 2.4, 2.0, 1.7, 1.0
    p, typ, A, pos, L = _two_ion_case(alpha, d, s, Q, 1e-9)
    V = L ** 3
    za = -d / 2
    e_bare = Q * (za + s) / abs(za + s) ** 3 - Q * (za - s) / abs(za - s) ** 3
    e0 = e_bare + 4 * np.pi / (3 * V) * (Q * (-s) - Q * s)
    g = 2.0 / d ** 3 + 8 * np.pi / (3 * V)
    rho = alpha * g
    assert 0.5 < rho < 0.9
    rm = [abs(e0) * rho ** k for k in range(200)]
    n_expect = next(k for k in range(200) if rm[k] <= 1e-9) + 1
    assert min(abs(math.log(rm[k] / 1e-9)) for k in range(200)) > 1e-3   # no borderline iteration
    o = Oracle(p, typ, A, pos.T.copy(), threads=1)
    t = o.engforce(0)
    x = alpha * e0 * (1 - rho ** (n_expect - 1)) / (1 - rho)
    mu = o.dipoles()
    assert t["n_scf"] == n_expect
    # image terms beyond the tin-foil one are O(size^3 / L^5) ~ 1e-7 of the
    # fields; rmsd_k carries k times their relative effect on rho
    assert t["rmsd"] == pytest.approx(rm[n_expect - 1], rel=1e-4)
    assert mu[0, 2] == pytest.approx(x, rel=1e-6) and mu[1, 2] == pytest.approx(x, rel=1e-6)
    assert np.abs(mu[:, :2]).max() < 1e-12 and np.abs(mu[2:]).max() == 0.0
    assert t["u_pol"] == pytest.approx(x * x / alpha, rel=2e-6)
    o.close()
    # min_scf_pim_iter forces more iterations past convergence; max stops early
    for minit, maxit, n in ((n_expect + 3, 200, n_expect + 3), (1, n_expect - 2, n_expect - 2)):
        p, typ, A, pos, L = _two_ion_case(alpha, d, s, Q, 1e-9, minit, maxit)
        o = Oracle(p, typ, A, pos.T.copy(), threads=1)
        t = o.engforce(0)
        assert t["n_scf"] == n
        assert t["rmsd"] == pytest.approx(rm[n - 1], rel=1e-4)
        o.close()


def test_scfko_fixed_point_equals_scf():
    """moment_pola_scf_ko (pim.c:457-748) converges to the fixed point of
    momentpolaSCF (its iteration schedule is not pinned here), on the
    reference's own scfkO input (src/control_SiO.s)"""
    p, cfg, _ = load_case("sio2_102_ko_static")
    res = []
    for algo in (0, 1):
        q = m.Params()
        C.pointer(q)[0] = p
        q.algo_pim = algo
        q.conv_tol_ind = 1e-11
        q.max_scf_pim_iter = 2000
        o = Oracle(q, cfg.typia, cfg.A, cfg.r, threads=8)
        t = o.engforce(0)
        res.append((o.dipoles(), o.forces(), t["u_tot"]))
        o.close()
    assert rel(res[1][0], res[0][0]) < 1e-9
    assert rel(res[1][1], res[0][1]) < 1e-9
    assert res[1][2] == pytest.approx(res[0][2], rel=1e-11)


# ------------------------------------------------------------------ ASPC
def _kolafa(ext_ord):
    """Kolafa's always-stable predictor-corrector of order k = ext_ord - 1:
    B_j = (-1)^(j+1) j C(2k+4, k+2-j) / C(2k+2, k+1), j = 1..k+2, and the
    corrector weight (k+2)/(2k+3); order 0 is the previous step (B = 1,
    no correction)"""
    if ext_ord == 0:
        return [1.0], 0.0
    k = ext_ord - 1
    B = [(-1) ** (j + 1) * j * math.comb(2 * k + 4, k + 2 - j) / math.comb(2 * k + 2, k + 1)
         for j in range(1, k + 3)]
    return B, (k + 2) / (2 * k + 3)


def _oracle_aspc(istep, order, hist):
    L = orc.load()
    L.orc_aspc_predict1.restype = C.c_int
    L.orc_aspc_predict1.argtypes = [C.c_int, C.c_int] + [C.POINTER(C.c_double)] * 3
    h = np.ascontiguousarray(hist, dtype=np.float64).copy()
    mu = np.zeros(3)
    W = C.c_double()
    eo = L.orc_aspc_predict1(istep, order, h.ctypes.data_as(C.POINTER(C.c_double)),
                             mu.ctypes.data_as(C.POINTER(C.c_double)), C.byref(W))
    return eo, mu, h, W.value


@pytest.mark.parametrize("order", range(6))
def test_aspc_coefficients_are_kolafas(order):
    """the full-order predictor reproduces Kolafa's B_j (unit histories)"""
    B, W = _kolafa(order)
    for j in range(order + 1):
        hist = np.zeros((6, 3))
        hist[j] = (1.0, -2.0, 0.5)
        eo, mu, _, w = _oracle_aspc(50, order, hist)
        assert eo == order
        assert np.allclose(mu, B[j] * hist[j], rtol=1e-15, atol=0)
        assert w == pytest.approx(W, rel=1e-15)


REF_ASPC = os.path.join(os.path.dirname(orc.__file__), "_ref", "ref_aspc")


@pytest.mark.parametrize("order", range(6))
def test_aspc_predictor_equals_the_references(order):
    """extrapolate_dipole_aspc of the reference itself (its pim.c compiled
    here, one ion): the predicted dipole and the shifted history at every
    step of the order ramp (pim.c:324-334, 393-398)"""
    if not os.path.exists(REF_ASPC):
        pytest.skip("oracle/_ref/ref_aspc not built (needs the reference tree at build time)")
    rng = np.random.default_rng(order)
    for istep in range(0, order + 4):
        hist = rng.normal(size=(6, 3))
        out = subprocess.run([REF_ASPC, str(order), str(istep)] + [repr(float(x)) for x in hist.ravel()],
                             capture_output=True, text=True, check=True).stdout.split()
        ref = np.array([float(x) for x in out])
        _, mu, h, _ = _oracle_aspc(istep, order, hist)
        assert np.array_equal(mu, ref[:3]), (istep, mu, ref[:3])
        assert np.array_equal(h.ravel(), ref[3:]), istep


# --------------------------------------------- stress = -d U / d strain
def test_charge_stress_is_minus_strain_derivative():
    """ewald_dir / ewald_rec / engforce_bhmftd stress of a rigid-ion melt
    (charges only: the stress is exact there, ewald.c:824-830, 999-1006):
    V tau_ab (eV) = -dU/d eps_ab under affine strain, to 1e-6"""
    p, cfg, _ = load_case("melt216_nve")
    L = cfg.A[0, 0]
    r = cfg.r.copy()
    # the cutoff in the widest gap between pair distances near 7 A: strains of
    # 1e-6 move distances by less than 1e-5 A, so no pair crosses it (a
    # crossing would break the derivative)
    dd = cfg.r[:, :, None] - cfg.r[:, None, :]
    dd = np.sqrt(np.sum(min_image(dd, L) ** 2, axis=0))
    ds = np.sort(dd[(dd > 6.8) & (dd < 7.2)])
    k = int(np.argmax(np.diff(ds)))
    rc = 0.5 * (ds[k] + ds[k + 1])
    assert ds[k + 1] - ds[k] > 2e-4
    p.alphaES = 0.7
    p.kES[0] = p.kES[1] = p.kES[2] = 16
    p.lverletL = 0
    p.cutshortrange = p.cutlongrange = rc
    o = Oracle(p, cfg.typia, cfg.A, r, threads=8)
    t = o.engforce(0)
    tau = (np.array(t["tau_coul"]) + np.array(t["tau_bhmftd"])) * L ** 3 * PRESS_GPA
    o.close()
    h = 1e-6

    def energy(eps):
        Fm = np.eye(3) + eps
        oo = Oracle(p, cfg.typia, cfg.A @ Fm.T, Fm @ r, threads=8)
        u = oo.engforce(0)["u_tot"]
        oo.close()
        return u

    for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)):
        E = np.zeros((3, 3))
        E[a, b] = E[b, a] = h
        dU = (energy(E) - energy(-E)) / (2 * h)
        want = -dU if a == b else -dU / 2
        assert tau[a, b] == pytest.approx(want, rel=1e-6, abs=1e-6 * np.abs(tau).max()), (a, b, tau[a, b], want)


def test_reference_ignores_static_quadrupoles():
    """the reference's own static run with quadit on both types (lqua set,
    field.c:201; the rank-5 tensor built, ewald.c:646-688) equals the oracle
    with the quadrupoles cleared: no energy, force, stress or Ewald setting of
    the reference reads them"""
    import ctypes as C
    p, cfg, ref = load_case("melt216_quad_static")
    assert str(ref["source"]) == "reference"
    assert p.nquadit == 2 and p.quadit[0][2][2] == -0.5
    q = m.Params()
    C.pointer(q)[0] = p
    for t in range(2):
        for i in range(3):
            for j in range(3):
                q.quadit[t][i][j] = 0.0
    o = Oracle(q, cfg.typia, cfg.A, cfg.r, threads=1)
    t = o.engforce(0)
    sc = ref["scalars"][0]
    mine = scalars_of(t)
    for k in range(len(SCALAR_KEYS)):
        assert abs(mine[k] - sc[k]) <= 1e-12 * max(abs(sc[k]), 1e-12), (SCALAR_KEYS[k], mine[k], sc[k])
    _, _, fr = ref_state(ref, 0)
    assert rel(o.forces(), fr, 1e-4) < 1e-12
    assert rel(np.array(t["tau_coul"]), ref["tau_coul"][0], 1e-12) < 1e-12
