"""The oracle's radial distribution (get_gr / write_rdf of
radial_distribution.c) against the reference build's own counts, a brute
force over cartesian distances, and the ideal-gas limit of g(r)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O

RDF_CASES = ["rdf_melt216", "rdf_tric"]


def load(case):
    return dict(np.load(os.path.join(GOLDEN, case, "ref.npz"), allow_pickle=False))


@pytest.mark.parametrize("case", RDF_CASES)
def test_oracle_rdf_matches_reference(case):
    """counts of every frame, the reference's extra all-pair count in
    gr[.][0][0] added, equal the reference's gr; the positions get_gr leaves
    (kardir then dirkar) are bitwise the reference's"""
    z = load(case)
    assert str(z["source"]) == "reference"
    nb, nt = int(z["nbins"]), len(z["nionit"])
    cut, nbo = O.rdf_defaults(z["A"][0])
    assert cut == float(z["cutgr"]) and nbo == nb
    counts = np.zeros((nb, nt, nt), dtype=np.int64)
    for A, r in zip(z["A"], z["frames"]):
        counts, after, out = O.rdf_frame(z["typia"], nt, A, r, cut, float(z["resg"]), nb, counts)
        assert out == 0
    q = counts.copy()
    q[:, 0, 0] += counts.sum(axis=(1, 2))
    assert np.array_equal(q, z["gr"])
    assert np.array_equal(after, z["r_after"])
    # ordered pairs: the type-swapped counts are equal
    assert np.array_equal(counts, counts.transpose(0, 2, 1))


def test_oracle_rdf_brute_force_orthorhombic():
    """an independent count: cartesian minimum image d - L round(d / L) in
    an orthorhombic cell, every ordered pair, numpy's histogram of r / resg"""
    rng = np.random.default_rng(3)
    Ls = np.array([11.0, 12.5, 9.75])
    n, nt = 300, 3
    typ = rng.integers(0, nt, n).astype(np.int32)
    r = (rng.uniform(-0.5, 0.5, (n, 3)) * Ls).T.copy()
    A = np.diag(Ls)
    cut, nb = O.rdf_defaults(A)
    assert cut == 0.5 * 9.75 and nb == int(cut / 0.1) + 1
    resg = 0.1
    counts, _, out = O.rdf_frame(typ, nt, A, r, cut, resg, nb)
    assert out == 0
    d = r.T[:, None, :] - r.T[None, :, :]
    d -= Ls * np.round(d / Ls)
    dist = np.sqrt((d * d).sum(-1))
    ref = np.zeros((nb, nt, nt), dtype=np.int64)
    i, j = np.nonzero((dist * dist < cut * cut) & ~np.eye(n, dtype=bool))
    b = (dist[i, j] / resg).astype(int)
    np.add.at(ref, (b, typ[i], typ[j]), 1)
    assert np.array_equal(counts, ref)
    # the pairs within cutgr, each ordered pair once
    assert counts.sum() == len(i)


def test_oracle_rdf_ideal_gas():
    """uniform random positions: write_rdf's normalisation gives g(r) = 1 for
    every type pair within the counting noise (the shell volume, the frame
    count, the per-type densities and the cell volume all enter)"""
    rng = np.random.default_rng(11)
    L = 14.0
    A = np.diag([L, L, L])
    n, nconf = 400, 12
    typ = np.repeat([0, 1], [250, 150]).astype(np.int32)
    cut, nb = O.rdf_defaults(A)
    resg = 0.35
    counts = np.zeros((nb, 2, 2), dtype=np.int64)
    for _ in range(nconf):
        r = rng.uniform(-0.5 * L, 0.5 * L, (3, n))
        counts, _, _ = O.rdf_frame(typ, 2, A, r, cut, resg, nb, counts)
    rr, g = O.rdf_normalise(counts, nconf, resg, L ** 3, [250, 150])
    assert np.allclose(rr, (np.arange(nb) + 0.5) * resg)
    used = (rr > 2.0) & (rr + 0.5 * resg < cut)
    for a in range(2):
        for b in range(2):
            m = g[used, a, b]
            assert abs(m.mean() - 1.0) < 0.02, (a, b, m.mean())
    # a factor of two in any term would be far outside the noise
    assert np.all(np.abs(g[used].mean(axis=0) - 1.0) < 0.02)
