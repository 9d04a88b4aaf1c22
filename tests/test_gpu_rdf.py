"""The device radial distribution (mdff_rdf.h) and the device trajectory
against the oracle's get_gr restatement (radial_distribution.c:180-231,
pinned to the reference's own counts in test_oracle_rdf.py): ordered pair
counts equal integer for integer, and the positions get_gr leaves equal bit
for bit."""
import os

import numpy as np
import pytest

import mdff20_amd as m
from conftest import GOLDEN, load_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.mark.parametrize("case", ["rdf_melt216", "rdf_tric"])
def test_rdf_frames_vs_oracle_and_reference(case):
    z = dict(np.load(os.path.join(GOLDEN, case, "ref.npz"), allow_pickle=False))
    nb, nt, resg = int(z["nbins"]), len(z["nionit"]), float(z["resg"])
    cut, nbo = m.rdf_defaults(z["A"][0])
    assert cut == float(z["cutgr"]) and nbo == nb
    say(f"rdf {case}: {len(z['frames'])} frames, {nb} bins")
    oc = np.zeros((nb, nt, nt), dtype=np.int64)
    with m.Rdf(z["typia"], nt, nb, resg, cut) as h:
        for A, r in zip(z["A"], z["frames"]):
            after = h.add_frame(A, r, write_back=True)
            oc, oafter, _ = O.rdf_frame(z["typia"], nt, A, r, cut, resg, nb, oc)
            assert np.array_equal(after, oafter)
        c, (frames, out) = h.counts()
    assert frames == len(z["frames"]) and out == 0
    assert np.array_equal(c, oc)
    q = c.copy()
    q[:, 0, 0] += c.sum(axis=(1, 2))
    assert np.array_equal(q, z["gr"])
    assert np.array_equal(after, z["r_after"])


def test_rdf_normalisation_vs_oracle():
    z = dict(np.load(os.path.join(GOLDEN, "rdf_melt216", "ref.npz"), allow_pickle=False))
    nb, nt, resg = int(z["nbins"]), len(z["nionit"]), float(z["resg"])
    with m.Rdf(z["typia"], nt, nb, resg, float(z["cutgr"])) as h:
        for A, r in zip(z["A"], z["frames"]):
            h.add_frame(A, r)
        c, (frames, _) = h.counts()
        r, g = h.gr(c, frames, float(z["omega"]))
    ro, go = O.rdf_normalise(c, frames, resg, float(z["omega"]), z["nionit"])
    assert np.array_equal(r, ro)
    assert np.allclose(g, go, rtol=1e-15, atol=0.0)


def test_rdf_12288_ions_and_three_types_vs_oracle():
    """one frame at the 12288-ion melt's size (1.5e8 ordered pairs) and a
    three-type triclinic system, integer for integer"""
    import bench
    L, pos, typ = bench.system(4096, 2026)
    pos = pos.T.copy()
    A = np.diag([L, L, L])
    cut, nb = m.rdf_defaults(A)
    say(f"rdf N={len(typ)}: {nb} bins, cutgr {cut:.3f}")
    with m.Rdf(typ, 2, nb, 0.1, cut) as h:
        h.add_frame(A, pos)
        c, (_, out) = h.counts()
    oc, _, oout = O.rdf_frame(typ, 2, A, pos, cut, 0.1, nb)
    assert out == 0 and oout == 0
    assert np.array_equal(c, oc)
    n = len(typ)
    rng = np.random.default_rng(5)
    At = np.array([[15.0, 0.0, 0.0], [3.0, 14.0, 0.0], [-2.0, 2.5, 13.5]])
    t3 = rng.integers(0, 3, 2000).astype(np.int32)
    r3 = (rng.uniform(-0.5, 0.5, (2000, 3)) @ At).T.copy()
    cut3, nb3 = m.rdf_defaults(At)
    with m.Rdf(t3, 3, 2 * nb3, 0.05, cut3) as h:
        h.add_frame(At, r3)
        c3, _ = h.counts()
    o3, _, _ = O.rdf_frame(t3, 3, At, r3, cut3, 0.05, 2 * nb3)
    assert np.array_equal(c3, o3)
    assert n == 12288


def test_device_trajectory_and_rdf_of_it():
    """md_run keeps the TRAJFF frames (iopsnode cadence) in device memory:
    each equals the state an engine stepped alongside reaches at that step,
    and the RDF of all frames in one launch equals the frames added one by
    one (and the engine's current positions sampled in place)"""
    p, cfg, _ = load_case("melt216_nve")
    npas, tprint, tstart = 30, 7, 5
    steps = [s for s in range(1, npas + 1) if (s % tprint == 0 or s == npas) and s >= tstart]
    with m.Engine(p, cfg.typia, cfg.A, cfg.r) as e, m.Engine(p, cfg.typia, cfg.A, cfg.r) as w:
        e.set_trajectory(tprint, tstart, npas, capacity=len(steps))
        for x in (e, w):
            x.init_velocities()
            x.md_init()
        e.md_run(1, npas)
        info = e.traj_info()
        assert info == dict(frames=len(steps), capacity=len(steps), dropped=0)
        cut, nb = m.rdf_defaults(cfg.A)
        k = 0
        with m.Rdf(cfg.typia, cfg.ntype, nb, 0.1, cut) as h1, m.Rdf(cfg.typia, cfg.ntype, nb, 0.1, cut) as h2:
            for s in range(1, npas + 1):
                w.md_step(s)
                if s in steps:
                    st, r, v, f = e.traj_frame(k)
                    assert st == s
                    assert np.array_equal(r, w.positions())
                    assert np.array_equal(v, w.velocities())
                    assert np.array_equal(f, w.forces())
                    h1.add_frame(cfg.A, r)
                    k += 1
            h2.sample_trajectory(e)
            c1, (f1, _) = h1.counts()
            c2, (f2, _) = h2.counts()
            assert f1 == f2 == len(steps) and np.array_equal(c1, c2)
            h2.reset()
            h2.sample(e)
            h1.reset()
            h1.add_frame(cfg.A, e.positions())
            assert np.array_equal(h1.counts()[0], h2.counts()[0])
        # a second run with room for one frame of five: the rest are counted
        e.set_trajectory(2, 0, 0, capacity=1)
        e.md_run(npas + 1, 10)
        assert e.traj_info() == dict(frames=1, capacity=1, dropped=4)
        assert e.traj_frame(0)[0] == npas + 2
        e.traj_clear()
        assert e.traj_info()["frames"] == 0
        e.set_trajectory(0, 0, 0, 0)
