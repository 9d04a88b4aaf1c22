"""Batched engine and force-matching adjoint (include/mdff_batch.h) through
the C ABI.

Parity: every configuration of a batch of uneven size against the oracle run
on that configuration alone (FP64, 1e-10 relative on energies, forces,
dipoles and stress, equal SCF iteration counts).  Mixed form: force RMS
<= 1e-5 relative to the FP64 batch.  Adjoint: the gradient of
L = sum |F - F_ref|^2 against central finite differences of L (step 1e-6
relative) to 1e-6 relative, the batch forces against -dU/dr by finite
differences, and 50 Adam steps through the torch.autograd.Function."""
import ctypes as C

import numpy as np
import pytest

import mdff20_amd as m
from conftest import load_case, rel

pytestmark = pytest.mark.gpu


def _supercell(cfg, reps, noise, seed):
    """reps (nx, ny, nz) copies of a fixture cell, positions jittered"""
    A = np.asarray(cfg.A)
    r = np.asarray(cfg.r)
    typ = np.asarray(cfg.typia)
    rs, ts = [], []
    for ix in range(reps[0]):
        for iy in range(reps[1]):
            for iz in range(reps[2]):
                sh = ix * A[0] + iy * A[1] + iz * A[2]
                rs.append(r + sh[:, None])
                ts.append(typ)
    rr = np.concatenate(rs, axis=1)
    tt = np.concatenate(ts)
    # keep the type-ordered layout POSFF gives (types by blocks)
    order = np.argsort(tt, kind="stable")
    rr, tt = rr[:, order], tt[order]
    rng = np.random.default_rng(seed)
    rr = rr + noise * rng.standard_normal(rr.shape)
    AA = A * np.array(reps)[:, None]
    return tt.astype(np.int32), AA, rr


def _configs():
    """uneven batch built from the pim192 fixture: itself, jittered copies,
    a 2x1x1 and a 2x2x1 supercell and a stretched cell"""
    p, cfg, _ = load_case("pim192_static")
    base = dict(alpha=0.35, kES=(6, 6, 6), rcut_long=7.0, rcut_short=7.0)
    out = []
    t, A, r = _supercell(cfg, (1, 1, 1), 0.0, 0)
    out.append(((t, A, r), base))
    t, A, r = _supercell(cfg, (1, 1, 1), 0.03, 1)
    out.append(((t, A, r), dict(base, alpha=0.4, kES=(7, 7, 7))))
    t, A, r = _supercell(cfg, (2, 1, 1), 0.03, 2)
    out.append(((t, A, r), dict(base, kES=(12, 6, 6))))
    t, A, r = _supercell(cfg, (1, 1, 1), 0.02, 3)
    s = 1.03
    out.append(((t, A * s, r * s), dict(base, alpha=0.38, kES=(6, 6, 6), rcut_long=7.2, rcut_short=6.5)))
    t, A, r = _supercell(cfg, (2, 2, 1), 0.03, 4)
    out.append(((t, A, r), dict(base, kES=(12, 12, 6))))
    return p, out


def _oracle_params(p, e):
    q = m.Params()
    C.pointer(q)[0] = p
    q.alphaES = e["alpha"]
    for k in range(3):
        q.kES[k] = int(e["kES"][k])
    q.cutlongrange = e["rcut_long"]
    q.cutshortrange = e["rcut_short"]
    q.lautoES = 0
    return q


KEYS = ["u_bhmftd", "u_coul", "u_pol", "u_tot"]


def test_batch_fp64_matches_oracle_per_configuration():
    from oracle.oracle import Oracle
    p, confs = _configs()
    print(f"[gpu] batch fp64 parity: sizes {[len(c[0][0]) for c in confs]}", flush=True)
    with m.batch.Batch(p, [c[0] for c in confs], [c[1] for c in confs], precision=m.FP64) as b:
        th = b.engforce()
        F = b.split(b.forces())
        mu = b.dipoles()
        for c, ((t, A, r), e) in enumerate(confs):
            q = _oracle_params(p, e)
            o = Oracle(q, t, A, r)
            to = o.engforce(0)
            for k in KEYS:
                assert th[c][k] == pytest.approx(to[k], rel=1e-10, abs=1e-12), (c, k, th[c][k], to[k])
            assert th[c]["n_scf"] == to["n_scf"], (c, th[c]["n_scf"], to["n_scf"])
            assert rel(F[c], o.forces()) < 1e-10, c
            mc = mu[b.off[c]:b.off[c + 1]]
            assert rel(mc, o.dipoles()) < 1e-10, c
            assert rel(np.array(th[c]["tau_coul"]), np.array(to["tau_coul"])) < 1e-10, c
            assert rel(np.array(th[c]["tau_bhmftd"]), np.array(to["tau_bhmftd"])) < 1e-10, c
            o.close()


def test_batch_autotune_keeps_results():
    p, confs = _configs()
    print("[gpu] batch autotune", flush=True)
    with m.batch.Batch(p, [c[0] for c in confs], [c[1] for c in confs]) as b:
        b.engforce()
        F0 = b.forces()
        tune = b.autotune()
        assert tune["lanes"] in (16, 32, 64) and tune["kpb"] in (32, 64, 128) and tune["ipb"] in (16, 32, 64, 128, 256)
        b.reset_dipoles()
        b.engforce()
        assert rel(b.forces(), F0) < 1e-12
        for lanes, kpb, ipb, bpc in [(32, 32, 128, 2), (64, 128, 256, 8), (16, 0, 32, 0), (0, 0, 16, 0)]:
            b.set_tuning(lanes, kpb, ipb, bpc)
            b.reset_dipoles()
            b.engforce()
            assert rel(b.forces(), F0) < 1e-12


def test_batch_mixed_accuracy():
    p, confs = _configs()
    print("[gpu] batch mixed", flush=True)
    q = m.Params()
    C.pointer(q)[0] = p
    q.conv_tol_ind = 1e-10
    with m.batch.Batch(q, [c[0] for c in confs], [c[1] for c in confs], precision=m.FP64) as bf, \
            m.batch.Batch(q, [c[0] for c in confs], [c[1] for c in confs], precision=m.MIXED) as bm:
        tf, tm = bf.engforce(), bm.engforce()
        Ff, Fm = bf.split(bf.forces()), bm.split(bm.forces())
        for c in range(len(confs)):
            frms = np.sqrt(np.mean((Fm[c] - Ff[c]) ** 2)) / np.sqrt(np.mean(Ff[c] ** 2))
            assert frms < 1e-5, (c, frms)
            assert tm[c]["u_tot"] == pytest.approx(tf[c]["u_tot"], rel=1e-6)


def _tight(p):
    q = m.Params()
    C.pointer(q)[0] = p
    q.conv_tol_ind = 1e-14
    q.max_scf_pim_iter = 2000
    q.min_scf_pim_iter = 1
    return q


def _perturbed(spec, seed):
    rng = np.random.default_rng(seed)
    v = spec.values()
    return spec.params(v + np.log(rng.uniform(0.9, 1.1, v.size)))


def test_adjoint_matches_finite_differences():
    p, confs = _configs()
    confs = confs[:3]
    p = _tight(p)
    print("[gpu] adjoint vs finite differences", flush=True)
    with m.batch.Batch(p, [c[0] for c in confs], [c[1] for c in confs]) as b:
        spec = m.batch.ParamSpec.default(p, p.ntype, with_D=False)
        kinds = {e[0] for e in spec.entries}
        assert {"A", "B", "C", "pol", "damp_b", "damp_c"} <= kinds, spec.entries
        # targets: FP64 forces under a perturbed parameter set
        b.set_params(_perturbed(spec, 7))
        b.engforce()
        fref = b.forces()
        b.set_params(p)
        g = b.adjoint(fref)
        assert g["adjoint_residual"] < 1e-12
        spec_lin = m.batch.ParamSpec(p, spec.entries, log=False)
        th0 = spec_lin.values()
        grad = spec_lin.gather(g)
        for k, e in enumerate(spec.entries):
            h = 1e-6 * abs(th0[k])
            L = []
            for sgn in (1, -1):
                th = th0.copy()
                th[k] += sgn * h
                b.set_params(spec_lin.params(th))
                b.engforce()
                d = b.forces() - fref
                L.append(float(np.sum(d * d)))
            fd = (L[0] - L[1]) / (2 * h)
            print(f"  {e}: adjoint {grad[k]:.10e} fd {fd:.10e}", flush=True)
            assert abs(grad[k] - fd) <= 1e-6 * max(abs(fd), 1e-8 * abs(g['loss']) / max(abs(th0[k]), 1e-30)), (e, grad[k], fd)
        b.set_params(p)


def test_batch_forces_are_minus_energy_gradient():
    p, confs = _configs()
    p = _tight(p)
    (t, A, r), e = confs[1]
    print("[gpu] batch forces vs -dU/dr", flush=True)
    with m.batch.Batch(p, [(t, A, r)], [e]) as b:
        b.engforce()
        F = b.forces()
        h = 1e-5
        for i in (0, 17, 150):
            for a in range(3):
                U = []
                for sgn in (1, -1):
                    rr = r.copy()
                    rr[a, i] += sgn * h
                    b.set_positions(rr)
                    U.append(b.engforce()[0]["u_tot"])
                fd = -(U[0] - U[1]) / (2 * h)
                assert abs(fd - F[a, i]) <= 2e-6 * max(1.0, abs(F[a, i])), (i, a, fd, F[a, i])


def test_adam_through_autograd_function():
    import torch  # noqa: F401
    from mdff20_amd.adjoint import ForceMatchLoss, fit
    p, confs = _configs()
    confs = confs[:2]
    q = m.Params()
    C.pointer(q)[0] = p
    q.conv_tol_ind = 1e-10
    print("[gpu] 50 Adam steps", flush=True)
    with m.batch.Batch(q, [c[0] for c in confs], [c[1] for c in confs]) as b:
        spec = m.batch.ParamSpec.default(q, q.ntype)
        b.set_params(_perturbed(spec, 11))
        b.engforce()
        fref = b.forces()
        b.set_params(q)
        theta = torch.tensor(spec.values(), dtype=torch.float64, requires_grad=True)
        loss = ForceMatchLoss.apply(theta, b, spec, fref)
        loss.backward()
        g = theta.grad.numpy().copy()
        assert np.all(np.isfinite(g)) and np.linalg.norm(g) > 0
        th, hist = fit(b, spec, fref, steps=50, lr=0.01)
        print(f"  loss {hist[0]:.6e} -> {hist[-1]:.6e}", flush=True)
        assert len(hist) == 50 and hist[-1] < 0.5 * hist[0]


def test_adjoint_pol_gradient_anisotropic():
    """a diagonal, anisotropic polarisability: the 'pol' entry scales the
    whole tensor, and its gradient from the adjoint (sum of dpol times the
    tensor over its [0][0] entry) equals the finite difference"""
    p, confs = _configs()
    confs = confs[:2]
    p = _tight(p)
    for a, v in enumerate((1.2, 1.0, 0.8)):
        p.polit[0][a][a] = v
    print("[gpu] anisotropic polarisability gradient", flush=True)
    with m.batch.Batch(p, [c[0] for c in confs], [c[1] for c in confs]) as b:
        spec = m.batch.ParamSpec(p, [("pol", 0)], log=False)
        b.set_params(spec.params(spec.values() * 1.05))
        b.engforce()
        fref = b.forces()
        b.set_params(p)
        g = b.adjoint(fref)
        grad = spec.gather(g)[0]
        x0 = spec.values()[0]
        h = 1e-6 * x0
        L = []
        for sgn in (1, -1):
            b.set_params(spec.params(np.array([x0 + sgn * h])))
            b.engforce()
            d = b.forces() - fref
            L.append(float(np.sum(d * d)))
        fd = (L[0] - L[1]) / (2 * h)
        print(f"  adjoint {grad:.10e} fd {fd:.10e}", flush=True)
        assert abs(grad - fd) <= 1e-6 * abs(fd)


def test_nonsymmetric_polarisability_batch():
    """a polarisability tensor that is not symmetric (the SCF reads it
    transposed, pim.c:637): the conjugate-gradient solves need a symmetric
    system, so the mixed SCF and the adjoint solve run their Jacobi
    sequences (the SCF to max_scf_pim_iter: the rmsd of pim.c:431-449
    divides by the diagonal entries only and does not vanish); the mixed
    forces stay within 1e-5 of the FP64 batch and the adjoint's 'pol'
    gradient equals the finite difference"""
    p, confs = _configs()
    confs = confs[:2]
    p = _tight(p)
    for a, v in enumerate((1.2, 1.0, 0.8)):
        p.polit[0][a][a] = v
    p.polit[0][0][1] = 0.15
    p.polit[0][2][0] = -0.1
    print("[gpu] non-symmetric polarisability", flush=True)
    q = m.Params()
    C.pointer(q)[0] = p
    q.conv_tol_ind = 1e-10
    with m.batch.Batch(q, [c[0] for c in confs], [c[1] for c in confs], precision=m.FP64) as bf, \
            m.batch.Batch(q, [c[0] for c in confs], [c[1] for c in confs], precision=m.MIXED) as bm:
        tf, tm = bf.engforce(), bm.engforce()
        Ff, Fm = bf.split(bf.forces()), bm.split(bm.forces())
        for c in range(len(confs)):
            frms = np.sqrt(np.mean((Fm[c] - Ff[c]) ** 2)) / np.sqrt(np.mean(Ff[c] ** 2))
            assert frms < 1e-5, (c, frms)
            assert tm[c]["u_tot"] == pytest.approx(tf[c]["u_tot"], rel=1e-6)
    with m.batch.Batch(p, [c[0] for c in confs], [c[1] for c in confs]) as b:
        spec = m.batch.ParamSpec(p, [("pol", 0)], log=False)
        b.set_params(spec.params(spec.values() * 1.05))
        b.engforce()
        fref = b.forces()
        b.set_params(p)
        g = b.adjoint(fref)
        grad = spec.gather(g)[0]
        x0 = spec.values()[0]
        h = 1e-6 * x0
        L = []
        for sgn in (1, -1):
            b.set_params(spec.params(np.array([x0 + sgn * h])))
            b.engforce()
            d = b.forces() - fref
            L.append(float(np.sum(d * d)))
        fd = (L[0] - L[1]) / (2 * h)
        print(f"  adjoint {grad:.10e} fd {fd:.10e}, {g['adjoint_iterations']} sweeps", flush=True)
        assert abs(grad - fd) <= 1e-6 * abs(fd)


def test_batch_vector_alu_kspace():
    """a configuration whose k box is wider than the matrix-core tiles take
    (kES 16 in x: 17 nx planes) sends the whole batch's k-space sums to the
    vector-ALU kernels: FP64 against the oracle per configuration (1e-10,
    the same SCF counts), the mixed batch within 1e-5 of it, and the
    adjoint's 'pol' gradient against the finite difference"""
    from oracle.oracle import Oracle
    p, confs = _configs()
    confs = [confs[0], (confs[2][0], dict(confs[2][1], kES=(16, 6, 6)))]
    print("[gpu] batch vector-ALU k space", flush=True)
    with m.batch.Batch(p, [c[0] for c in confs], [c[1] for c in confs], precision=m.FP64) as b, \
            m.batch.Batch(p, [c[0] for c in confs], [c[1] for c in confs], precision=m.MIXED) as bm:
        th, tm = b.engforce(), bm.engforce()
        F, Fm = b.split(b.forces()), bm.split(bm.forces())
        for c, ((t, A, r), e) in enumerate(confs):
            o = Oracle(_oracle_params(p, e), t, A, r)
            to = o.engforce(0)
            for k in KEYS:
                assert th[c][k] == pytest.approx(to[k], rel=1e-10, abs=1e-12), (c, k, th[c][k], to[k])
            assert th[c]["n_scf"] == to["n_scf"], c
            assert rel(F[c], o.forces()) < 1e-10, c
            o.close()
            frms = np.sqrt(np.mean((Fm[c] - F[c]) ** 2)) / np.sqrt(np.mean(F[c] ** 2))
            assert frms < 1e-5, (c, frms)
    q = _tight(p)
    with m.batch.Batch(q, [c[0] for c in confs], [c[1] for c in confs]) as b:
        spec = m.batch.ParamSpec(q, [("pol", 0)], log=False)
        b.set_params(spec.params(spec.values() * 1.05))
        b.engforce()
        fref = b.forces()
        b.set_params(q)
        g = b.adjoint(fref)
        grad = spec.gather(g)[0]
        x0 = spec.values()[0]
        h = 1e-6 * x0
        L = []
        for sgn in (1, -1):
            b.set_params(spec.params(np.array([x0 + sgn * h])))
            b.engforce()
            d = b.forces() - fref
            L.append(float(np.sum(d * d)))
        fd = (L[0] - L[1]) / (2 * h)
        print(f"  adjoint {grad:.10e} fd {fd:.10e}", flush=True)
        assert abs(grad - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("kx", [6, 16])
def test_batch_triclinic_configuration(kx):
    """a triclinic configuration beside a cubic one (each with its own
    reciprocal basis in the back-projection's epilogue; kES 16 in x sends
    the batch to the vector-ALU k-space sums): FP64 against the oracle per
    configuration (1e-10, the same SCF counts) and the mixed batch within
    1e-5"""
    from oracle.oracle import Oracle
    p, confs = _configs()
    (t, A, r), e = confs[0]
    e = dict(e, kES=(kx, 6, 6))
    A = np.asarray(A, dtype=np.float64)
    s = np.linalg.solve(A.T, np.asarray(r))
    A2 = A.copy()
    A2[1][0] += 0.15 * A[0][0]
    A2[2][0] += 0.10 * A[0][0]
    A2[2][1] += 0.05 * A[0][0]
    confs = [confs[1], ((t, A2, A2.T @ s), e)]
    print("[gpu] batch triclinic", flush=True)
    with m.batch.Batch(p, [c[0] for c in confs], [c[1] for c in confs], precision=m.FP64) as b, \
            m.batch.Batch(p, [c[0] for c in confs], [c[1] for c in confs], precision=m.MIXED) as bm:
        th, tm = b.engforce(), bm.engforce()
        F, Fm = b.split(b.forces()), bm.split(bm.forces())
        for c, ((tt, AA, rr), ee) in enumerate(confs):
            o = Oracle(_oracle_params(p, ee), tt, AA, rr)
            to = o.engforce(0)
            for k in KEYS:
                assert th[c][k] == pytest.approx(to[k], rel=1e-10, abs=1e-12), (c, k, th[c][k], to[k])
            assert th[c]["n_scf"] == to["n_scf"], c
            assert rel(F[c], o.forces()) < 1e-10, c
            assert rel(np.array(th[c]["tau_coul"]), np.array(to["tau_coul"])) < 1e-10, c
            o.close()
            frms = np.sqrt(np.mean((Fm[c] - F[c]) ** 2)) / np.sqrt(np.mean(F[c] ** 2))
            assert frms < 1e-5, (c, frms)


def test_set_params_switches_the_scf_solver():
    """set_params from symmetric to non-symmetric polarisabilities and back:
    the solver follows the tensors (conjugate gradients only while every
    polarisability is symmetric positive definite), so a batch whose
    parameters were changed gives the forces of a batch created with them
    (mixed form, both from zero dipoles: 1e-12 relative)"""
    p, confs = _configs()
    confs = confs[:2]
    q = m.Params()
    C.pointer(q)[0] = p
    q.conv_tol_ind = 1e-10
    q.max_scf_pim_iter = 400
    ns = m.Params()
    C.pointer(ns)[0] = q
    ns.polit[0][0][1] = 0.1 * ns.polit[0][0][0]
    print("[gpu] set_params symmetric <-> non-symmetric", flush=True)
    args = ([c[0] for c in confs], [c[1] for c in confs])
    with m.batch.Batch(q, *args, precision=m.MIXED) as b, m.batch.Batch(ns, *args, precision=m.MIXED) as bn:
        b.engforce()
        b.set_params(ns)
        b.reset_dipoles()
        b.engforce()
        bn.engforce()
        assert rel(b.forces(), bn.forces()) < 1e-12
        b.set_params(q)
        b.reset_dipoles()
        b.engforce()
        with m.batch.Batch(q, *args, precision=m.MIXED) as bq:
            bq.engforce()
            assert rel(b.forces(), bq.forces()) < 1e-12


def test_set_params_switches_damping():
    """set_params that turns the Tang-Toennies damping of the charge-dipole
    terms off and on again: the batch gives the forces of a batch created
    with those parameters (FP64, both from zero dipoles: 1e-12)"""
    p, confs = _configs()
    confs = confs[:2]
    nd = m.Params()
    C.pointer(nd)[0] = p
    for t in range(nd.ntype):
        for i in range(nd.ntype):
            for k in range(nd.ntype):
                nd.lpoldamping[t][i][k] = 0
    print("[gpu] set_params damping off / on", flush=True)
    args = ([c[0] for c in confs], [c[1] for c in confs])
    with m.batch.Batch(p, *args) as b, m.batch.Batch(nd, *args) as bn, m.batch.Batch(p, *args) as bp:
        b.engforce()
        b.set_params(nd)
        b.reset_dipoles()
        b.engforce()
        bn.engforce()
        bp.engforce()
        assert rel(b.forces(), bn.forces()) < 1e-12
        assert rel(bn.forces(), bp.forces()) > 1e-8   # the damping matters here
        b.set_params(p)
        b.reset_dipoles()
        b.engforce()
        assert rel(b.forces(), bp.forces()) < 1e-12
        # what the batch does not run is refused, as at creation: scfkO
        ko = m.Params()
        C.pointer(ko)[0] = p
        ko.algo_pim = m.PIM_SCFKO
        with pytest.raises(m.MdffError):
            b.set_params(ko)


def test_adjoint_triclinic():
    """the adjoint of a batch holding a triclinic configuration (its real-
    space, reciprocal-space and damping terms through the configuration's
    own cell): the gradients of A, pol and the damping b against central
    finite differences, as test_adjoint_matches_finite_differences"""
    p, confs = _configs()
    (t, A, r), e = confs[0]
    A = np.asarray(A, dtype=np.float64)
    s = np.linalg.solve(A.T, np.asarray(r))
    A2 = A.copy()
    A2[1][0] += 0.15 * A[0][0]
    A2[2][1] += 0.05 * A[0][0]
    confs = [confs[1], ((t, A2, A2.T @ s), e)]
    q = _tight(p)
    print("[gpu] adjoint, triclinic configuration", flush=True)
    entries = [("A", 0, 1), ("pol", 0), ("damp_b", 0, 1)]
    with m.batch.Batch(q, [c[0] for c in confs], [c[1] for c in confs]) as b:
        b.set_params(_perturbed(m.batch.ParamSpec(q, entries), 5))
        b.engforce()
        fref = b.forces()
        b.set_params(q)
        g = b.adjoint(fref)
        spec = m.batch.ParamSpec(q, entries, log=False)
        th0 = spec.values()
        grad = spec.gather(g)
        for k, en in enumerate(entries):
            h = 1e-6 * abs(th0[k])
            L = []
            for sgn in (1, -1):
                th = th0.copy()
                th[k] += sgn * h
                b.set_params(spec.params(th))
                b.engforce()
                d = b.forces() - fref
                L.append(float(np.sum(d * d)))
            fd = (L[0] - L[1]) / (2 * h)
            print(f"  {en}: adjoint {grad[k]:.10e} fd {fd:.10e}", flush=True)
            assert abs(grad[k] - fd) <= 1e-6 * max(abs(fd), 1e-8 * abs(g["loss"]) / max(abs(th0[k]), 1e-30)), en
