"""Batched engine and force-matching adjoint (C ABI of include/mdff_batch.h).

Batch evaluates many independent periodic configurations of uneven size with
one force-field parameter set in one set of launches; Batch.adjoint returns
the loss sum |F - F_ref|^2 and its gradient with respect to the BHMFTD A, B,
C, D, the polarisabilities and the Tang-Toennies charge-dipole damping b, c
(one implicit adjoint solve with the dipole tensor).  ParamSpec maps a flat
parameter vector onto Params for the torch.autograd.Function in
mdff20_amd.adjoint.
"""
import ctypes as C

import numpy as np

from . import FP64, MIXED, MdffError, Params, Thermo, _check, _dp, lib
from ._types import NT

__all__ = ["Batch", "BatchEwald", "Grad", "ParamSpec", "ewald_for_accuracy"]

_d = C.c_double
_T2d = (_d * NT) * NT


class BatchEwald(C.Structure):
    _fields_ = [("alpha", _d), ("kES", C.c_int * 3), ("rcut_long", _d), ("rcut_short", _d)]


class Grad(C.Structure):
    _fields_ = [("loss", _d), ("dA", _T2d), ("dB", _T2d), ("dC", _T2d), ("dD", _T2d),
                ("dpol", ((_d * 3) * 3) * NT), ("ddamp_b", _T2d), ("ddamp_c", _T2d),
                ("adjoint_iterations", C.c_int), ("adjoint_residual", _d)]

    def arrays(self):
        return dict(loss=self.loss, dA=np.ctypeslib.as_array(self.dA).copy(),
                    dB=np.ctypeslib.as_array(self.dB).copy(), dC=np.ctypeslib.as_array(self.dC).copy(),
                    dD=np.ctypeslib.as_array(self.dD).copy(), dpol=np.ctypeslib.as_array(self.dpol).copy(),
                    ddamp_b=np.ctypeslib.as_array(self.ddamp_b).copy(),
                    ddamp_c=np.ctypeslib.as_array(self.ddamp_c).copy(),
                    adjoint_iterations=self.adjoint_iterations, adjoint_residual=self.adjoint_residual)


def _batch_lib():
    L = lib()
    if getattr(L, "_batch_ready", False):
        return L
    P = C.POINTER
    dp, ip, vp = P(C.c_double), P(C.c_int), C.c_void_p
    sig = {
        "mdff_batch_create": (C.c_int, [P(vp), P(Params), C.c_int, ip, ip, dp, dp, dp, dp, P(BatchEwald), C.c_int,
                                        C.c_int]),
        "mdff_batch_destroy": (None, [vp]),
        "mdff_batch_set_positions": (C.c_int, [vp, dp, dp, dp]),
        "mdff_batch_set_params": (C.c_int, [vp, P(Params)]),
        "mdff_batch_reset_dipoles": (C.c_int, [vp]),
        "mdff_batch_engforce": (C.c_int, [vp, P(Thermo)]),
        "mdff_batch_get_forces": (C.c_int, [vp, dp, dp, dp]),
        "mdff_batch_get_dipoles": (C.c_int, [vp, dp]),
        "mdff_batch_adjoint": (C.c_int, [vp, dp, dp, dp, P(Grad)]),
        "mdff_batch_autotune": (C.c_int, [vp, ip]),
        "mdff_batch_set_tuning": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "mdff_batch_sizeof_grad": (C.c_size_t, []),
        "mdff_batch_get_ewald": (C.c_int, [vp, C.c_int, dp, ip, ip]),
        "mdff_batch_get_timing": (C.c_int, [vp, dp, P(C.c_longlong)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.mdff_batch_sizeof_grad() != C.sizeof(Grad):
        raise RuntimeError("struct layout mismatch between batch.py and mdff_batch.h")
    L._batch_ready = True
    return L


def ewald_for_accuracy(L, rc_max=9.0, eps=1e-6, rs_max=None):
    """Ewald settings of a cubic cell of side L for relative accuracy eps:
    rc = min(rc_max, L/2), alpha = sqrt(-ln eps)/rc, K = ceil(alpha L sqrt(-ln eps)/pi)."""
    rc = min(rc_max, 0.5 * L * (1.0 - 1e-9))
    s = np.sqrt(-np.log(eps))
    alpha = s / rc
    k = int(np.ceil(alpha * L * s / np.pi))
    rs = rc if rs_max is None else min(rs_max, rc)
    return dict(alpha=alpha, kES=(k, k, k), rcut_long=rc, rcut_short=rs)


class Batch:
    """Many configurations, one parameter set, one GPU (mdff_batch.h).
    configs: list of (typia[n], A[3,3], r[3,n]); ewald: list of dicts with
    alpha, kES, rcut_long, rcut_short (or None for the params' values)."""

    def __init__(self, params, configs, ewald=None, precision=FP64, device=0):
        L = _batch_lib()
        self._L = L
        self.params = params
        self.nconf = len(configs)
        self.nion = np.array([len(c[0]) for c in configs], dtype=np.int32)
        self.off = np.concatenate([[0], np.cumsum(self.nion)]).astype(np.int64)
        self.n = int(self.off[-1])
        typ = np.ascontiguousarray(np.concatenate([np.asarray(c[0], dtype=np.int32) for c in configs]))
        A = np.ascontiguousarray(np.concatenate([np.asarray(c[1], dtype=np.float64).reshape(9) for c in configs]))
        r = np.concatenate([np.asarray(c[2], dtype=np.float64).reshape(3, -1) for c in configs], axis=1)
        rx, ry, rz = (np.ascontiguousarray(r[k]) for k in range(3))
        ew = None
        if ewald is not None:
            ew = (BatchEwald * self.nconf)()
            for i, e in enumerate(ewald):
                ew[i].alpha = e["alpha"]
                for k in range(3):
                    ew[i].kES[k] = int(e["kES"][k])
                ew[i].rcut_long = e["rcut_long"]
                ew[i].rcut_short = e["rcut_short"]
        h = C.c_void_p()
        rc = L.mdff_batch_create(C.byref(h), C.byref(params), self.nconf,
                                 self.nion.ctypes.data_as(C.POINTER(C.c_int)),
                                 typ.ctypes.data_as(C.POINTER(C.c_int)), _dp(A), _dp(rx), _dp(ry), _dp(rz),
                                 ew, precision, device)
        _check(rc, "batch_create", False)
        self._h = h
        self.typ = typ

    def close(self):
        if getattr(self, "_h", None):
            self._L.mdff_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def split(self, a):
        """split a per-ion array (last axis = ions) into configurations"""
        return [a[..., self.off[c]:self.off[c + 1]] for c in range(self.nconf)]

    def set_positions(self, r):
        r = np.asarray(r, dtype=np.float64).reshape(3, self.n)
        a = [np.ascontiguousarray(r[k]) for k in range(3)]
        _check(self._L.mdff_batch_set_positions(self._h, _dp(a[0]), _dp(a[1]), _dp(a[2])), "batch_set_positions",
               False)

    def set_params(self, params):
        _check(self._L.mdff_batch_set_params(self._h, C.byref(params)), "batch_set_params", False)
        self.params = params

    def reset_dipoles(self):
        _check(self._L.mdff_batch_reset_dipoles(self._h), "batch_reset_dipoles", False)

    def engforce(self):
        th = (Thermo * self.nconf)()
        rc = _check(self._L.mdff_batch_engforce(self._h, th), "batch_engforce")
        out = [t.as_dict() for t in th]
        for d in out:
            d["status"] = rc
        return out

    def forces(self):
        a = [np.zeros(self.n) for _ in range(3)]
        _check(self._L.mdff_batch_get_forces(self._h, _dp(a[0]), _dp(a[1]), _dp(a[2])), "batch_get_forces", False)
        return np.stack(a)

    def dipoles(self):
        mu = np.zeros((self.n, 3))
        _check(self._L.mdff_batch_get_dipoles(self._h, _dp(mu)), "batch_get_dipoles", False)
        return mu

    def adjoint(self, fref):
        """loss and gradient (dict of arrays) against target forces fref[3, n]"""
        fref = np.asarray(fref, dtype=np.float64).reshape(3, self.n)
        a = [np.ascontiguousarray(fref[k]) for k in range(3)]
        g = Grad()
        _check(self._L.mdff_batch_adjoint(self._h, _dp(a[0]), _dp(a[1]), _dp(a[2]), C.byref(g)), "batch_adjoint")
        return g.arrays()

    def autotune(self):
        out = (C.c_int * 4)()
        _check(self._L.mdff_batch_autotune(self._h, out), "batch_autotune", False)
        return dict(lanes=out[0], kpb=out[1], ipb=out[2], bpc=out[3])

    def set_tuning(self, lanes=0, kpb=0, ipb=0, bpc=0):
        _check(self._L.mdff_batch_set_tuning(self._h, lanes, kpb, ipb, bpc), "batch_set_tuning", False)

    def ewald(self, c):
        a, k, nk = C.c_double(), (C.c_int * 3)(), C.c_int()
        _check(self._L.mdff_batch_get_ewald(self._h, c, C.byref(a), k, C.byref(nk)), "batch_get_ewald", False)
        return a.value, list(k), nk.value

    def timing(self):
        ms, sw = C.c_double(), C.c_longlong()
        _check(self._L.mdff_batch_get_timing(self._h, C.byref(ms), C.byref(sw)), "batch_get_timing", False)
        return ms.value, sw.value


class ParamSpec:
    """Flat parameter vector <-> Params for force matching.  Entries:
    ('A'|'B'|'C'|'D', ti, tj) BHMFTD pair values (symmetric), ('pol', t)
    isotropic polarisability (polit = a I), ('damp_b'|'damp_c', t, s)
    charge-dipole damping of the field at type t from type s
    (pol_damp_*[t][t][s]).  log=True works on log(value)."""

    def __init__(self, params, entries, log=True):
        self.base = params
        self.entries = list(entries)
        self.log = log

    @staticmethod
    def default(params, ntype, with_D=False):
        """every nonzero BHMFTD A, B, C (D if with_D), every polarisable type
        and every damped (t, s) pair of params"""
        ent = []
        for key, fld in (("A", "Abhmftd"), ("B", "Bbhmftd"), ("C", "Cbhmftd")) + ((("D", "Dbhmftd"),) if with_D else ()):
            for i in range(ntype):
                for j in range(i, ntype):
                    if getattr(params, fld)[i][j] != 0.0:
                        ent.append((key, i, j))
        for t in range(ntype):
            if params.lpolar[t] and params.polit[t][0][0] != 0.0:
                ent.append(("pol", t))
        for t in range(ntype):
            for s in range(ntype):
                if params.lpoldamping[t][t][s] or params.lpoldamping[s][t][s]:
                    if params.pol_damp_b[t][t][s] != 0.0:
                        ent.append(("damp_b", t, s))
                    if params.pol_damp_c[t][t][s] != 0.0:
                        ent.append(("damp_c", t, s))
        return ParamSpec(params, ent)

    def values(self, params=None):
        p = self.base if params is None else params
        v = []
        for e in self.entries:
            k = e[0]
            if k in "ABCD":
                v.append(getattr(p, k + "bhmftd")[e[1]][e[2]])
            elif k == "pol":
                v.append(p.polit[e[1]][0][0])
            elif k == "damp_b":
                v.append(p.pol_damp_b[e[1]][e[1]][e[2]])
            else:
                v.append(p.pol_damp_c[e[1]][e[1]][e[2]])
        v = np.array(v, dtype=np.float64)
        return np.log(v) if self.log else v

    def params(self, theta):
        """Params with the entries set from theta"""
        v = np.exp(np.asarray(theta, dtype=np.float64)) if self.log else np.asarray(theta, dtype=np.float64)
        p = Params()
        C.pointer(p)[0] = self.base
        for e, x in zip(self.entries, v):
            k = e[0]
            x = float(x)
            if k in "ABCD":
                f = getattr(p, k + "bhmftd")
                f[e[1]][e[2]] = x
                f[e[2]][e[1]] = x
            elif k == "pol":
                old = p.polit[e[1]][0][0]
                for a in range(3):
                    for b in range(3):
                        p.polit[e[1]][a][b] = p.polit[e[1]][a][b] * (x / old) if old != 0.0 else (x if a == b else 0.0)
            elif k == "damp_b":
                p.pol_damp_b[e[1]][e[1]][e[2]] = x
            else:
                p.pol_damp_c[e[1]][e[1]][e[2]] = x
        return p

    def gather(self, grad, theta=None):
        """d loss / d theta from the gradient dict of Batch.adjoint"""
        g = []
        for e in self.entries:
            k = e[0]
            if k in "ABCD":
                g.append(grad["d" + k][e[1]][e[2]])
            elif k == "pol":
                # params() scales the whole tensor polit_base by x / polit_base[0][0]
                base = np.array([[self.base.polit[e[1]][a][b] for b in range(3)] for a in range(3)])
                g.append(float(np.sum(grad["dpol"][e[1]] * base)) / base[0, 0])
            elif k == "damp_b":
                g.append(grad["ddamp_b"][e[1]][e[2]])
            else:
                g.append(grad["ddamp_c"][e[1]][e[2]])
        g = np.array(g, dtype=np.float64)
        if self.log:
            v = np.exp(np.asarray(theta, dtype=np.float64)) if theta is not None else np.exp(self.values())
            g = g * v
        return g
