"""Write the golden vectors of tests/golden/ from the reference build.

TEST INFRASTRUCTURE, run in the development container only (it needs the
reference tree for oracle/_ref and the reference's data files):

    make -C oracle all ref && python tests/golden/make_fixtures.py

Every case directory holds the inputs (`control`, `POSFF`) and `ref.npz`
at %.17g.  Cases without polarisable ions are written by
oracle/_ref/mdff20_ref_driver: the reference's own routines, its
init_config included (see oracle/ref_driver.c), compiled from its sources;
their ref.npz records source "reference" and they pin the oracle.  A
polarisable input needs the reference's LAPACK 3x3 inverse
(inverse_matrix.c), which is absent, so the reference build cannot run it:
those cases are written by the oracle itself (source "oracle") and are
regression vectors only; the oracle's polarisable path is pinned by the
closed forms of tests/test_oracle_pins.py.  Inputs are either data files of the
reference (src/*.s, src/POSFF.*, examples/ex0, ex1) or synthetic melts made
here from a fixed seed (SiO2-like: cations q=+2.4, anions q=-1.2, 2.2 g/cm^3,
uniform positions with 1.6 A minimum-image rejection).  `ccd_lj.npz` holds
the Cambridge Cluster Database energies and structures that the reference's
example 2 compares against (examples/ex2/config/REFERENCE and <N> files).
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)

from oracle import refrun  # noqa: E402

M_O, M_SI = 15.99994, 28.084
RHO_UNIT = 1.660538782


def sio2_box(n_units, rho=2.2):
    mtot = n_units * (M_SI + 2 * M_O)
    return (mtot / (rho / RHO_UNIT)) ** (1.0 / 3.0)


def random_positions(n, L, rng, dmin=1.6, max_tries=200000):
    """uniform positions, minimum-image rejection at dmin"""
    pos = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("packing failed")
        p = rng.uniform(-0.5 * L, 0.5 * L, 3)
        if k:
            d = pos[:k] - p
            d -= L * np.rint(d / L)
            if np.min(np.einsum("ij,ij->i", d, d)) < dmin * dmin:
                continue
        pos[k] = p
        k += 1
    return pos


def write_posff(path, name, L, types, counts, pos, vel=None):
    with open(path, "w") as fh:
        fh.write(f"{len(pos)}\n{name}\n")
        fh.write(f"{L:.12f} 0.0 0.0\n0.0 {L:.12f} 0.0\n0.0 0.0 {L:.12f}\n")
        fh.write(f"{len(types)}\n{' '.join(types)}\n{' '.join(str(c) for c in counts)}\nCartesian\n")
        ti = np.repeat(np.arange(len(types)), counts)
        for i, p in enumerate(pos):
            v = vel[i] if vel is not None else (0.0, 0.0, 0.0)
            fh.write(f"{types[ti[i]]} {p[0]:.15f} {p[1]:.15f} {p[2]:.15f} "
                     f"{v[0]:.15f} {v[1]:.15f} {v[2]:.15f} 0.0 0.0 0.0\n")


def pair_params(seed):
    """BHMFT per-type-pair parameters A in [1e3,2e4] eV, B in [2.5,5] 1/A,
    C in [0,150] eV A^6, D = 0, drawn from `seed`; a draw is kept only if
    every pair (with its Coulomb term) is repulsive on [0.6, 1.5] A, so no
    pair can collapse in a 3000 K melt.  Returns (seed used, A, B, C) with
    upper triangles in the order O-O, O-Si, Si-Si."""
    q = {0: -1.2, 1: 2.4}
    pairs = [(0, 0), (0, 1), (1, 1)]
    r = np.linspace(0.6, 1.5, 400)
    while True:
        rng = np.random.default_rng(seed)
        A = rng.uniform(1e3, 2e4, 3)
        B = rng.uniform(2.5, 5.0, 3)
        Cc = rng.uniform(0.0, 150.0, 3)
        ok = True
        for k, (i, j) in enumerate(pairs):
            dV = -A[k] * B[k] * np.exp(-B[k] * r) + 6 * Cc[k] / r ** 7 - 14.3996441494161 * q[i] * q[j] / r ** 2
            if np.any(dV >= 0):
                ok = False
        if ok:
            return seed, A, B, Cc
        seed += 1


def control_text(**kw):
    lines = []
    for k, v in kw.items():
        if isinstance(v, bool):
            v = "true" if v else "false"
        if isinstance(v, (list, tuple)):
            v = " ".join(str(x) for x in v)
        lines.append(f"{k} {v}")
    return "\n".join(lines) + "\n"


def sio2_field(seed_used, A, B, Cc, pim, alpha, kmax, rc, extra=""):
    t = [
        "# SiO2-like field: type 0 = O (anion), type 1 = Si (cation)",
        f"# BHMFT parameters drawn with seed {seed_used} (see tests/golden/make_fixtures.py)",
        "massit 15.99994", "massit 28.084", "qit -1.2", "qit 2.4",
        "lbhmft true",
        "Abhmftd " + " ".join(repr(float(x)) for x in A),
        "Bbhmftd " + " ".join(repr(float(x)) for x in B),
        "Cbhmftd " + " ".join(repr(float(x)) for x in Cc),
        "Dbhmftd 0.0 0.0 0.0",
        "lcoulombic true", "lautoES false", f"alphaES {alpha}", f"kES {kmax} {kmax} {kmax}",
        f"cutshortrange {float(rc)!r}", f"cutlongrange {float(rc)!r}",
    ]
    if pim:
        t += ["polit 1.2 0.0 0.0 0.0 1.2 0.0 0.0 0.0 1.2",
              "polit 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0",
              "lpoldamping 0 0 0 true", "pol_damp_b 0 0 0 2.0", "pol_damp_c 0 0 0 1.5", "pol_damp_k 0 0 0 4",
              "lpoldamping 0 0 1 true", "pol_damp_b 0 0 1 2.0", "pol_damp_c 0 0 1 1.5", "pol_damp_k 0 0 1 4",
              "algo_pim scf"]
    return "\n".join(t) + "\n" + extra


def oracle_run(workdir, mode, *args):
    """the driver's dump written by the oracle (threads 1): same records as
    refrun.parse returns"""
    sys.path.insert(0, ROOT)
    import mdff20_amd as m
    from oracle.oracle import Oracle
    cfg = m.read_posff(os.path.join(workdir, "POSFF"), 2)
    p = m.read_control(os.path.join(workdir, "control"), cfg.ntype)
    o = Oracle(p, cfg.typia, cfg.A, cfg.r, threads=1)
    a, k, nk = o.ewald()
    res = {"alphaES": a, "kES": k, "nk": nk, "steps": []}
    ko = p.algo_pim == 1

    def record(step, t, dump):
        rec = {"step": step, "n_scf": t["n_scf"] + (o.ko_inner() if ko else 0),
               "verlet_updates": t["verlet_updates"],
               "scalars": np.array([t[k] for k in ("e_kin", "e_nvt", "u_lj", "u_bhmftd", "u_coul", "u_pol",
                                                   "pvir_lj", "pvir_bhmftd", "pvir_coul")])}
        for key in ("tau_lj", "tau_bhmftd", "tau_coul"):
            rec[key] = np.array(t[key]).reshape(9)
        if dump:
            r, v, f = o.state()
            for i, c in enumerate("xyz"):
                rec["r" + c], rec["v" + c], rec["f" + c] = r[i], v[i], f[i]
            if p.lcoulombic:
                rec["mu"] = o.dipoles().ravel()
        res["steps"].append(rec)

    if mode == "static":
        record(0, o.engforce(0), True)
    else:
        nsteps, period = int(args[0]), int(args[1])
        o.init_velocities()
        v0 = o.velocities()
        res["v0x"], res["v0y"], res["v0z"] = v0[0], v0[1], v0[2]
        record(0, o.md_init(), period > 0)
        for s in range(1, nsteps + 1):
            record(s, o.md_step(s), (period > 0 and s % period == 0) or s == nsteps)
    o.close()
    return res


def run_case(name, control, posff_src=None, posff_text=None, mode="static", args=(), posff_writer=None):
    d = os.path.join(HERE, name)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "control"), "w") as fh:
        fh.write(control)
    if posff_src:
        shutil.copy(posff_src, os.path.join(d, "POSFF"))
    elif posff_writer:
        posff_writer(os.path.join(d, "POSFF"))
    with tempfile.TemporaryDirectory() as tmp:
        shutil.copy(os.path.join(d, "control"), os.path.join(tmp, "control"))
        shutil.copy(os.path.join(d, "POSFF"), os.path.join(tmp, "POSFF"))
        try:
            res, source = refrun.run(tmp, mode, *args, threads=1), "reference"
        except subprocess.CalledProcessError as ex:
            if ex.returncode != 5:   # 5: polarisable input, unbuildable reference path
                raise
            res, source = oracle_run(tmp, mode, *args), "oracle"
    out = {"mode": np.array(mode), "source": np.array(source), "alphaES": np.array(res["alphaES"]),
           "kES": np.array(res["kES"]), "nk": np.array(res["nk"])}
    for key in ("v0x", "v0y", "v0z"):
        if key in res:
            out[key] = res[key]
    steps = res["steps"]
    out["step"] = np.array([s["step"] for s in steps])
    out["n_scf"] = np.array([s["n_scf"] for s in steps])
    out["verlet_updates"] = np.array([s["verlet_updates"] for s in steps])
    out["scalars"] = np.stack([s["scalars"] for s in steps])
    for k in ("tau_lj", "tau_bhmftd", "tau_coul"):
        out[k] = np.stack([s[k].reshape(3, 3) for s in steps])
    dumped = [s for s in steps if "rx" in s]
    out["state_step"] = np.array([s["step"] for s in dumped])
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz", "mu"):
        if dumped and k in dumped[0]:
            out[k] = np.stack([s[k] for s in dumped])
    for k, v in out.items():
        if np.issubdtype(np.asarray(v).dtype, np.floating) and not np.all(np.isfinite(v)):
            print(f"  warning: non-finite values in {name}:{k}")
    np.savez_compressed(os.path.join(d, "ref.npz"), **out)
    print(f"{name} ({source}): {len(steps)} records, n_scf {out['n_scf'][:5]}")
    return out


def parse_trajff(path, n):
    """TRAJFF records (write_config config.c:291-345) -> [(A (3, 3), r (3, n))],
    the values as read_next_traj (config.c:448-502) reads them"""
    lines = open(path).read().split("\n")
    out, i = [], 0
    while i < len(lines) and lines[i].strip():
        A = np.array([[float(x) for x in lines[i + 2 + k].split()] for k in range(3)])
        rows = lines[i + 9:i + 9 + n]
        r = np.array([[float(x) for x in ln.split()[1:4]] for ln in rows]).T.copy()
        out.append((A, r))
        i += 9 + n
    return out


def rdf_case(name, control, posff_text, frames, resg, seed_note):
    """the reference's RDF mode (mdff20.c:108-110) over TRAJFF frames written
    here: gr counts as get_gr leaves them (with its extra all-pair count in
    gr[.][0][0]) and the positions after the last frame"""
    import mdff20_amd as m
    d = os.path.join(HERE, name)
    os.makedirs(d, exist_ok=True)
    ctl = control + f"lrdf true\nresg {resg}\nnconf {len(frames)}\n"
    open(os.path.join(d, "control"), "w").write(ctl)
    open(os.path.join(d, "POSFF"), "w").write(posff_text)
    P = m.read_posff(os.path.join(d, "POSFF"))
    with tempfile.TemporaryDirectory() as tmp:
        for f in ("control", "POSFF"):
            shutil.copy(os.path.join(d, f), os.path.join(tmp, f))
        for A, r in frames:
            P.A = np.asarray(A, dtype=float)
            P.r = r
            P.v = np.zeros_like(r)
            P.f = np.zeros_like(r)
            m.write_config(os.path.join(tmp, "TRAJFF"), P, mode="a")
        read = parse_trajff(os.path.join(tmp, "TRAJFF"), P.nion)
        res = refrun.run(tmp, "rdf", threads=1)
    nb, nt = int(res["rdf_nbins"]), int(res["rdf_ntype"])
    out = dict(source=np.array("reference"), typia=np.asarray(P.typia, dtype=np.int32),
               atypit=np.array(P.atypit), nionit=np.array(P.nionit, dtype=np.int32),
               A=np.stack([a for a, _ in read]), frames=np.stack([r for _, r in read]),
               resg=np.array(float(res["resg"])), cutgr=np.array(float(res["cutgr"])), nbins=np.array(nb),
               omega=np.array(float(res["omega"])), nconf=np.array(len(frames)),
               gr=res["gr"].reshape(nb, nt, nt).astype(np.int64),
               r_after=np.stack([res["rx"], res["ry"], res["rz"]]), note=np.array(seed_note))
    np.savez_compressed(os.path.join(d, "ref.npz"), **out)
    print(f"{name} (reference rdf): {len(frames)} frames, {nb} bins, {int(out['gr'][:, 0, 0].sum())} counts in gr[.][0][0]")


def melt(n_units, seed, pos_seed):
    L = sio2_box(n_units)
    rng = np.random.default_rng(pos_seed)
    # O first, then Si
    pos = random_positions(3 * n_units, L, rng)
    return L, pos


def equilibrate(L, pos, field, steps=300, temp=3000.0):
    """rigid-ion pre-equilibration of a random packing (input generation
    only) with the oracle restatement: NVE velocity Verlet with a Berendsen
    rescale every step, dt 1 fs"""
    sys.path.insert(0, ROOT)
    import mdff20_amd as m
    from oracle.oracle import Oracle
    with tempfile.TemporaryDirectory() as tmp:
        ctl = os.path.join(tmp, "control")
        with open(ctl, "w") as fh:
            fh.write(field + control_text(integrator="nve-vv", dt=0.001, temp=temp, tauTberendsen=0.001,
                                          nequil=steps + 1, nequilT=1, lverletL=True, skindiff=0.5))
        p = m.read_control(ctl, 2)
    n = len(pos)
    nu = n // 3
    typ = np.array([0] * (2 * nu) + [1] * nu, dtype=np.int32)
    A = np.diag([L, L, L])
    o = Oracle(p, typ, A, pos.T.copy(), threads=8)
    o.init_velocities()
    o.md_init()
    for s in range(1, steps + 1):
        o.md_step(s)
    r = o.positions().T
    r -= L * np.rint(r / L)
    o.close()
    return r


def main():
    only = set(sys.argv[1:])
    want = lambda n: not only or n in only
    src = os.path.join(REF, "src")
    # ---- tiny analytic inputs of the reference (src/)
    if want("twocharges"):
        run_case("twocharges", open(os.path.join(src, "control_twocharges.s")).read(),
                 os.path.join(src, "POSFF.twocharges"))
    if want("onedipole"):
        run_case("onedipole", open(os.path.join(src, "control_onedipoles.s")).read(),
                 os.path.join(src, "POSFF.onedipole"))
    if want("twocharges_twopol"):
        run_case("twocharges_twopol", open(os.path.join(src, "control_twocharges_twopol.s")).read(),
                 os.path.join(src, "POSFF.twocharges_twopol"))
    # ---- SiO2 102 ions and BSiO 200 ions melts of src/ with the scf algorithm
    if want("sio2_102_static") or want("sio2_102_nvt"):
        base = open(os.path.join(src, "control_SiO.s")).read().replace("algo_pim scfkO", "algo_pim scf")
        if want("sio2_102_static"):
            run_case("sio2_102_static", base.replace("lstatic false", "lstatic true"),
                     os.path.join(src, "POSFF.SiO2"))
        if want("sio2_102_nvt"):
            run_case("sio2_102_nvt", base, os.path.join(src, "POSFF.SiO2"), mode="md", args=(12, 6))
    if want("bsio_200_nvt"):
        base = open(os.path.join(src, "control_BSiO.s")).read().replace("algo_pim scfkO", "algo_pim scf")
        run_case("bsio_200_nvt", base, os.path.join(src, "POSFF.BSiO"), mode="md", args=(6, 3))
    # ---- the same two inputs with their own SCF algorithm, scfkO (pim.c:457-748)
    if want("sio2_102_ko_static") or want("sio2_102_ko_nvt"):
        base = open(os.path.join(src, "control_SiO.s")).read()
        if want("sio2_102_ko_static"):
            run_case("sio2_102_ko_static", base.replace("lstatic false", "lstatic true"),
                     os.path.join(src, "POSFF.SiO2"))
        if want("sio2_102_ko_nvt"):
            run_case("sio2_102_ko_nvt", base, os.path.join(src, "POSFF.SiO2"), mode="md", args=(12, 6))
    if want("bsio_200_ko_nvt"):
        base = open(os.path.join(src, "control_BSiO.s")).read()
        run_case("bsio_200_ko_nvt", base, os.path.join(src, "POSFF.BSiO"), mode="md", args=(6, 3))
    # ---- Lennard-Jones FCC of examples/ex0 (static) and ex1 (NVE with rescale)
    if want("lj_fcc_static"):
        run_case("lj_fcc_static", open(os.path.join(REF, "examples/ex0/config/control.s")).read(),
                 os.path.join(REF, "examples/ex0/config/POSFF"))
    if want("lj_fcc_nve"):
        run_case("lj_fcc_nve", open(os.path.join(REF, "examples/ex1/config/control.s")).read(),
                 os.path.join(REF, "examples/ex1/config/POSFF"), mode="md", args=(40, 20))
    # ---- synthetic SiO2-like melts
    seed_used, A, B, Cc = pair_params(20260)
    if want("melt216_nve") or want("melt216_raw"):
        # rigid ions, N=216, Ewald alpha 0.3, rc = L/2, kmax 8, NVE VV 1 fs, 3000 K
        nu = 72
        L = sio2_box(nu)
        _, pos = melt(nu, seed_used, 11)
        field = sio2_field(seed_used, A, B, Cc, False, 0.3, 8, 0.5 * L)
        ctl = field + control_text(lverletL=True, skindiff=0.5, integrator="nve-vv", dt=0.001, temp=3000.0,
                                   tauTberendsen=0.001, npas=200, nprint=1000, fprint=1000, cprint=1000,
                                   tprint=1000)
        if want("melt216_raw"):
            # the raw 1.6 A packing is far from equilibrium and the trajectory
            # is strongly chaotic: 40 steps only
            run_case("melt216_raw", ctl, mode="md", args=(40, 20),
                     posff_writer=lambda p: write_posff(p, "melt216raw", L, ["O", "Si"], [2 * nu, nu], pos))
        if want("melt216_nve"):
            eq = equilibrate(L, pos, field, steps=300)
            run_case("melt216_nve", ctl, mode="md", args=(200, 100),
                     posff_writer=lambda p: write_posff(p, "melt216", L, ["O", "Si"], [2 * nu, nu], eq))
    if want("pim192_static") or want("pim192_nvt") or want("pim192_fixed"):
        # polarisable ions, N=192, alpha_O 1.2 A^3, TT b=2 c=1.5 k=4, rc 7 A,
        # kmax 6, pre-equilibrated rigid-ion start
        nu = 64
        L = sio2_box(nu)
        _, pos = melt(nu, seed_used, 23)
        rig = sio2_field(seed_used, A, B, Cc, False, 0.35, 6, 7.0)
        pos = equilibrate(L, pos, rig, steps=300)
        field = sio2_field(seed_used, A, B, Cc, True, 0.35, 6, 7.0)
        wr = lambda p: write_posff(p, "pim192", L, ["O", "Si"], [2 * nu, nu], pos)
        if want("pim192_static"):
            ctl = field + control_text(lverletL=True, skindiff=1.0, lstatic=True, conv_tol_ind=1e-8,
                                       min_scf_pim_iter=3, max_scf_pim_iter=400)
            run_case("pim192_static", ctl, posff_writer=wr)
        if want("pim192_fixed"):
            # fixed SCF schedule: exactly 12 iterations, independent of round-off
            ctl = field + control_text(lverletL=True, skindiff=1.0, lstatic=True, conv_tol_ind=1e-30,
                                       min_scf_pim_iter=12, max_scf_pim_iter=12)
            run_case("pim192_fixed", ctl, posff_writer=wr)
        if want("pim192_nvt"):
            ctl = field + control_text(lverletL=True, skindiff=1.0, integrator="nvt-nhcn", nhc_n=3,
                                       nhc_yosh_order=3, nhc_mults=2, timesca_thermo=0.1, temp=2500.0, dt=0.001,
                                       conv_tol_ind=1e-8, min_scf_pim_iter=3, max_scf_pim_iter=400,
                                       algo_extrapolate_dipole="aspc", extrapolate_order=2, nprint=1000)
            run_case("pim192_nvt", ctl, mode="md", args=(10, 5), posff_writer=wr)
    # ---- variants of the relaxed 216-ion melt the reference runs unmodified
    # (no polarisable ions): static dipoles, damped BHMFTD dispersion, the
    # Nose-Hoover chain and leap-frog
    m216 = os.path.join(HERE, "melt216_nve")
    if any(want(c) for c in ("melt216_dip_nve", "melt216_bhmftd_nve", "melt216_nvt", "melt216_lf",
                             "melt216_quad_static")):
        base = open(os.path.join(m216, "control")).read()
        tmp = tempfile.mkdtemp()
        shutil.copy(os.path.join(m216, "POSFF"), os.path.join(tmp, "POSFF"))
        pos216 = os.path.join(tmp, "POSFF")

        def vary(text, **kw):
            keep = [ln for ln in text.splitlines() if ln.split()[:1] not in ([k] for k in kw)]
            return "\n".join(keep) + "\n" + control_text(**kw)
        if want("melt216_dip_nve"):
            # a static dipole on every anion: charge-dipole and dipole-dipole
            # Ewald terms, forces and stress (ewald.c:737-830, 999-1006)
            run_case("melt216_dip_nve", vary(base, npas=50) + "dipit 0.05 -0.03 0.08\ndipit 0.0 0.0 0.0\n",
                     pos216, mode="md", args=(50, 25))
        if want("melt216_quad_static"):
            # static quadrupoles on both types: the reference sets lqua
            # (field.c:201) and builds its rank-5 tensor (ewald.c:646-688),
            # but no energy, field or force term reads the quadrupoles
            run_case("melt216_quad_static", vary(base, npas=0) + "lstatic true\n"
                     "quadit 0.3 0.0 0.1 0.0 0.2 0.0 0.1 0.0 -0.5\nquadit -0.2 0.05 0.0 0.05 0.4 0.0 0.0 0.0 -0.2\n",
                     pos216)
        if want("melt216_bhmftd_nve"):
            # BHMFTD with Tang-Toennies damped C6/C8 terms (bhmftd.c:117-271)
            run_case("melt216_bhmftd_nve",
                     vary(base.replace("lbhmft true", "lbhmftd true"), Dbhmftd="40.0 55.0 70.0",
                          BDbhmftd="3.1 2.7 3.4", npas=50), pos216, mode="md", args=(50, 25))
        if want("melt216_nvt"):
            # Nose-Hoover chain of length 3, Yoshida order 3, 2 multiple steps
            # (integration.c:140-294, with the second half-sweep of :269-277)
            run_case("melt216_nvt", vary(base, integrator="nvt-nhcn", nhc_n=3, nhc_yosh_order=3, nhc_mults=2,
                                         timesca_thermo=0.02, temp=2500.0, npas=100), pos216,
                     mode="md", args=(100, 50))
        if want("melt216_lf"):
            # leap-frog (integration.c:93-137) from r(t-dt) = 0 as the
            # reference allocates it (config.c:61-70), no equilibration
            run_case("melt216_lf", vary(base, integrator="nve-lf", nequil=0, npas=4), pos216,
                     mode="md", args=(4, 2))
    if want("twodipoles"):
        # src/control_twodipoles.s with two unit dipoles along x
        run_case("twodipoles", open(os.path.join(src, "control_twodipoles.s")).read(),
                 posff_writer=lambda p: open(p, "w").write(
                     "2\nTWODIP\n10.0 0.0 0.0\n0.0 10.0 0.0\n0.0 0.0 10.0\n2\nA B\n1 1\nCartesian\n"
                     "A 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0\n"
                     "B 1.3 0.7 -0.4 0.0 0.0 0.0 0.0 0.0 0.0\n"))
    # ---- radial distribution (radial_distribution.c): TRAJFF frames made here
    if want("rdf_melt216"):
        # the melt216 configuration and three frames displaced by up to 0.3 A
        g = os.path.join(HERE, "melt216_nve")
        ctl = open(os.path.join(g, "control")).read()
        txt = open(os.path.join(g, "POSFF")).read()
        import mdff20_amd as m
        P = m.read_posff(os.path.join(g, "POSFF"))
        rng = np.random.default_rng(20216)
        r, frames = P.r.copy(), []
        for _ in range(3):
            r = r + rng.uniform(-0.3, 0.3, r.shape)
            frames.append((P.A.copy(), r.copy()))
        rdf_case("rdf_melt216", ctl, txt, frames, 0.1, "melt216 POSFF, uniform +-0.3 A steps, seed 20216")
    if want("rdf_tric"):
        # a triclinic cell, 150 O + 90 Si uniform in fractional coordinates,
        # two frames, bins of 0.15 A
        g = os.path.join(HERE, "melt216_nve")
        ctl = open(os.path.join(g, "control")).read()
        rng = np.random.default_rng(7150)
        A = np.array([[13.0, 0.0, 0.0], [2.6, 12.4, 0.0], [-1.9, 3.1, 11.7]])
        lines = ["240", "tric240"] + [" ".join(f"{x:.12f}" for x in row) for row in A]
        lines += ["2", "O Si", "150 90", "Direct"]
        s0 = rng.uniform(-0.5, 0.5, (240, 3))
        lines += [f"{'O' if i < 150 else 'Si'} {s0[i, 0]:.12f} {s0[i, 1]:.12f} {s0[i, 2]:.12f} "
                  "0.0 0.0 0.0 0.0 0.0 0.0" for i in range(240)]
        txt = "\n".join(lines) + "\n"
        frames = []
        for _ in range(2):
            s1 = rng.uniform(-0.5, 0.5, (240, 3))
            frames.append((A, (s1 @ A).T.copy()))
        rdf_case("rdf_tric", ctl, txt, frames, 0.15, "triclinic 240 ions, uniform fractional, seed 7150")
    # ---- CCD LJ clusters (examples/ex2): reference energies and structures
    if want("ccd"):
        cdir = os.path.join(REF, "examples/ex2/config")
        names, energies, coords = [], [], {}
        with open(os.path.join(cdir, "REFERENCE")) as fh:
            for line in fh:
                if line.startswith("#") or not line.strip():
                    continue
                t = line.split()
                names.append(t[0])
                energies.append(float(t[1]))
        for nm in names:
            f = os.path.join(cdir, nm)
            if os.path.exists(f):
                coords[nm] = np.loadtxt(f).reshape(-1, 3)
        np.savez_compressed(os.path.join(HERE, "ccd_lj.npz"), names=np.array(names), energies=np.array(energies),
                            **{f"xyz_{k}": v for k, v in coords.items()})
        print(f"ccd: {len(names)} energies, {len(coords)} structures")


if __name__ == "__main__":
    main()
