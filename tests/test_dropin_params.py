"""The same-name drop-ins build their engine parameters from the reference's
parsed globals (mdff_dropin_params, include/mdff_dropin.h).  Run through the
reference's own driver linked with the drop-ins (oracle/_ref, no GPU call),
they must equal what this package's control-file parser gives for every
golden case, apart from the MD keys the force context does not use and the
Ewald settings set_autoES already applied (compared with the reference's
own alphaES/kES)."""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

import mdff20_amd as m
from conftest import GOLDEN, load_case
from oracle import refrun

IGNORED = {"nmassit", "nqit", "ndipit", "nquadit", "npolit", "integrator", "dt", "temp", "tauTberendsen",
           "timesca_thermo", "timesca_baro", "npas", "nprint", "fprint", "cprint", "tprint", "tstart", "nequil",
           "nequilT", "nhc_n", "nhc_mults", "nhc_yosh_order", "lautoES", "alphaES", "kES"}
LJ = {"sigmalj", "epslj", "plj", "qlj"}
PIM = {"algo_pim", "algo_extrapolate_dipole", "extrapolate_order", "min_scf_pim_iter", "max_scf_pim_iter",
       "conv_tol_ind", "omegakO"}
CASES = sorted(c for c in os.listdir(GOLDEN)
               if os.path.exists(os.path.join(GOLDEN, c, "control")) and not c.startswith("rdf_"))


@pytest.mark.parametrize("case", CASES)
def test_dropin_params_equal_parser(case, tmp_path):
    if not os.path.exists(refrun.DRIVER_DROPIN):
        pytest.skip("oracle/_ref/mdff20_dropin_driver not built (needs the reference tree at build time)")
    d = str(tmp_path)
    for f in ("control", "POSFF"):
        shutil.copy(os.path.join(GOLDEN, case, f), d)
    out = refrun.run(d, "params", C.sizeof(m.Params), exe=refrun.DRIVER_DROPIN, timeout=60)
    q = m.Params.from_buffer_copy(bytes.fromhex(out["params"]))
    p, cfg, ref = load_case(case)
    for name, _ in m.Params._fields_:
        if name in IGNORED or (name in LJ and not p.lnmlj) or (name in PIM and not any(p.lpolar[:p.ntype])):
            continue
        a, b = getattr(p, name), getattr(q, name)
        if hasattr(a, "_length_"):
            # per-type tables: the entries of the types in use
            A, B = np.ctypeslib.as_array(a), np.ctypeslib.as_array(b)
            nt = p.ntype
            sl = tuple(slice(0, nt) if s_ == m._types.NT else slice(None) for s_ in A.shape)
            assert np.array_equal(A[sl], B[sl]), name
        else:
            assert a == b, (name, a, b)
    assert q.lautoES == 0
    assert q.alphaES == pytest.approx(float(ref["alphaES"]), rel=1e-15)
    assert list(q.kES) == [int(k) for k in ref["kES"]]
    assert out["alphaES"] == pytest.approx(q.alphaES, rel=1e-15)
