"""mdff_check_verlet_noalloc against mdff_check_verlet through the C ABI:
twin contexts stepped by prop_agate alone (mdff_propagate), one checking its
list with each entry point; the rebuild counts and (bitwise) the forces and
positions agree.  Two calls in a row: the 12288-ion melt (cell grid), then
the smaller 1536-ion melt (cell grid asked for), on the same GPU, so the
second call reuses the scratch the first one left."""
import ctypes as C

import numpy as np
import pytest

import bench
import mdff20_amd as m

pytestmark = pytest.mark.gpu


def _system(units, seed, kmax, method):
    L, pos, typ = bench.system(units, seed)
    p = bench.params_of(bench.control(False, kmax=kmax))
    return p, typ, np.diag([L, L, L]), pos.T.copy(), method


@pytest.mark.parametrize("units,seed,kmax,method", [(4096, 7, 16, 0), (512, 8, 10, 2)])
def test_check_verlet_noalloc_matches_check_verlet(units, seed, kmax, method):
    p, typ, A, r, method = _system(units, seed, kmax, method)
    L = m.lib()
    print(f"[gpu] no-alloc Verlet check: N={len(typ)}, list method {method}", flush=True)
    with m.Engine(p, typ, A, r) as er, m.Engine(p, typ, A, r) as en:
        for e in (er, en):
            e.set_list_method(method)
            e.init_velocities()
            e.md_init()
        for s in range(1, 41):
            tr, tn = er.propagate(s), en.propagate(s)
            assert L.mdff_check_verlet(er._h) == 0
            rc = L.mdff_check_verlet_noalloc(en._h)
            assert rc == 0, (s, rc)
            assert np.array_equal(er.forces(), en.forces()), s
            assert tr["u_tot"] == tn["u_tot"], s
        ur, un = er.thermo()["verlet_updates"], en.thermo()["verlet_updates"]
        assert ur == un and un >= 2, (ur, un)
        assert np.array_equal(er.positions(), en.positions())
        print(f"[gpu] N={len(typ)}: grid {er.list_info()}, {un} builds, forces bitwise equal", flush=True)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    """an ion count above the scratch's maximum: a clear status, nothing run"""
    L, pos, typ = bench.system(45000, 9)   # 135000 ions > 131072
    p = bench.params_of(bench.control(False, kmax=4, rc=5.0))
    with m.Engine(p, typ, np.diag([L, L, L]), pos.T.copy()) as e:
        assert m.lib().mdff_check_verlet_noalloc(e._h) == 10
