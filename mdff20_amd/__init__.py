"""mdff20_amd -- MI355X engine for mdff20's force and integration path.

Python side of the C ABI in include/mdff_gpu.h and include/mdff_host.h.
The engine lives in libmdff_gpu.so next to this file (built by
__graft_entry__.build()); importing a compute entry point without it raises
instead of falling back to any CPU code.
"""
import ctypes as C
import os

import numpy as np

from ._types import (ES, ERR_SCF, FP64, KS_F16X3, KS_F32, MIXED, NVE_LF, NVE_VV, NVT_NHCN, PIM_PCG, PIM_SCF,
                     PIM_SCFKO, STATUS, Config, Params, Thermo)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmdff_gpu.so")
_lib = None

__all__ = ["lib", "MdffError", "read_control", "read_posff", "Engine", "Params", "Thermo",
           "FP64", "MIXED", "KS_F32", "KS_F16X3", "ES", "NVE_LF", "NVE_VV", "NVT_NHCN", "PIM_SCF", "PIM_SCFKO", "PIM_PCG",
           "device_count", "tt_damping", "init_velocities_host", "Rdf", "rdf_defaults"]


class MdffError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"mdff error {code} ({STATUS.get(code, '?')}) in {where}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def lib():
    """Load libmdff_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    dp, ip, vp = P(C.c_double), P(C.c_int), C.c_void_p
    sig = {
        "mdff_version": (C.c_char_p, []),
        "mdff_device_count": (C.c_int, []),
        "mdff_sizeof_params": (C.c_size_t, []),
        "mdff_sizeof_thermo": (C.c_size_t, []),
        "mdff_sizeof_config": (C.c_size_t, []),
        "mdff_params_default": (None, [P(Params)]),
        "mdff_read_control": (C.c_int, [C.c_char_p, C.c_int, P(Params)]),
        "mdff_read_posff": (C.c_int, [C.c_char_p, C.c_int, P(Config)]),
        "mdff_write_config": (C.c_int, [C.c_char_p, C.c_char_p, P(Config)]),
        "mdff_config_alloc": (C.c_int, [P(Config), C.c_int]),
        "mdff_config_free": (None, [P(Config)]),
        "mdff_host_init_velocities": (C.c_int, [P(Params), C.c_int, dp, dp, dp, dp, C.c_int]),
        "mdff_tt_damping": (None, [C.c_double, C.c_double, C.c_double, dp, dp, C.c_int]),
        "mdff_create": (C.c_int, [P(vp), P(Params), C.c_int, ip, dp, dp, dp, dp, C.c_int, C.c_int]),
        "mdff_destroy": (None, [vp]),
        "mdff_set_positions": (C.c_int, [vp, dp, dp, dp]),
        "mdff_set_velocities": (C.c_int, [vp, dp, dp, dp]),
        "mdff_get_positions": (C.c_int, [vp, dp, dp, dp]),
        "mdff_get_velocities": (C.c_int, [vp, dp, dp, dp]),
        "mdff_get_forces": (C.c_int, [vp, dp, dp, dp]),
        "mdff_get_dipoles": (C.c_int, [vp, dp]),
        "mdff_get_field": (C.c_int, [vp, dp]),
        "mdff_get_thermo": (C.c_int, [vp, P(Thermo)]),
        "mdff_get_ewald": (C.c_int, [vp, dp, ip, ip]),
        "mdff_init_velocities": (C.c_int, [vp]),
        "mdff_check_verlet": (C.c_int, [vp]),
        "mdff_check_verlet_noalloc": (C.c_int, [vp]),
        "mdff_engforce": (C.c_int, [vp, C.c_int, P(Thermo)]),
        "mdff_pair_short": (C.c_int, [vp, dp, dp, dp]),
        "mdff_multipole_es": (C.c_int, [vp, dp, dp, dp, C.c_int, dp, dp, dp, dp, dp, dp, dp, dp]),
        "mdff_get_dipoles_scf": (C.c_int, [vp, C.c_int, dp, dp]),
        "mdff_md_init": (C.c_int, [vp, P(Thermo)]),
        "mdff_md_step": (C.c_int, [vp, C.c_int, P(Thermo)]),
        "mdff_propagate": (C.c_int, [vp, C.c_int, P(Thermo)]),
        "mdff_md_run": (C.c_int, [vp, C.c_int, C.c_int, P(Thermo)]),
        "mdff_set_tuning": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "mdff_set_kspace_shape": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "mdff_set_kspace_modes": (C.c_int, [vp, C.c_int, C.c_int]),
        "mdff_set_scf_anchor": (C.c_int, [vp, C.c_int]),
        "mdff_get_timing": (C.c_int, [vp, dp]),
        "mdff_time_kernel": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp]),
        "mdff_get_counters": (C.c_int, [vp, P(C.c_longlong)]),
        "mdff_set_scf_solver": (C.c_int, [vp, C.c_int, C.c_int]),
        "mdff_set_scf_check": (C.c_int, [vp, C.c_int]),
        "mdff_set_scf_refine": (C.c_int, [vp, C.c_int]),
        "mdff_set_trajectory": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "mdff_traj_info": (C.c_int, [vp, ip]),
        "mdff_traj_frame": (C.c_int, [vp, C.c_int, ip, P(dp), P(dp), P(dp)]),
        "mdff_traj_clear": (C.c_int, [vp]),
        "mdff_rdf_create": (vp, [C.c_int, C.c_int, ip, C.c_int, C.c_double, C.c_double, ip]),
        "mdff_rdf_destroy": (None, [vp]),
        "mdff_rdf_defaults": (C.c_int, [dp, dp, ip]),
        "mdff_rdf_reset": (C.c_int, [vp]),
        "mdff_rdf_add_frame": (C.c_int, [vp, dp, dp, dp, dp, C.c_int]),
        "mdff_rdf_sample": (C.c_int, [vp, vp]),
        "mdff_rdf_time": (C.c_int, [vp, vp, C.c_int, dp]),
        "mdff_rdf_sample_trajectory": (C.c_int, [vp, vp]),
        "mdff_rdf_counts": (C.c_int, [vp, P(C.c_longlong), P(C.c_longlong)]),
        "mdff_rdf_gr": (C.c_int, [vp, P(C.c_longlong), C.c_longlong, C.c_double, dp, dp]),
        "mdff_set_list_method": (C.c_int, [vp, C.c_int]),
        "mdff_list_info": (C.c_int, [vp, ip]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.mdff_sizeof_params() != C.sizeof(Params) or L.mdff_sizeof_thermo() != C.sizeof(Thermo) \
            or L.mdff_sizeof_config() != C.sizeof(Config):
        raise RuntimeError("struct layout mismatch between _types.py and mdff_types.h")
    _lib = L
    return L


def device_count():
    return lib().mdff_device_count()


def _check(rc, where, allow_scf=True):
    if rc != 0 and not (allow_scf and rc == ERR_SCF):
        raise MdffError(rc, where)
    return rc


def read_control(path, ntype):
    """Parse an mdff20 control file (keywords of global.c, md.c, field.c,
    bhmftd.c, nmlj.c, pim.c) into Params."""
    p = Params()
    _check(lib().mdff_read_control(os.fsencode(path), ntype, C.byref(p)), "read_control", False)
    return p


def default_params():
    p = Params()
    lib().mdff_params_default(C.byref(p))
    return p


class PosffData:
    """Content of a POSFF file with numpy arrays (positions wrapped to the
    primary cell as read_config does)."""

    def __init__(self, cfg):
        n = cfg.nion
        self.nion = n
        self.ntype = cfg.ntype
        self.name = cfg.name.decode()
        self.A = np.array([[cfg.A[i][j] for j in range(3)] for i in range(3)])
        self.atypit = [cfg.atypit[t].value.decode() for t in range(cfg.ntype)]
        self.nionit = [cfg.nionit[t] for t in range(cfg.ntype)]
        self.typia = np.ctypeslib.as_array(cfg.typia, (n,)).copy()
        g = lambda p_: np.ctypeslib.as_array(p_, (n,)).copy()
        self.r = np.stack([g(cfg.rx), g(cfg.ry), g(cfg.rz)])
        self.v = np.stack([g(cfg.vx), g(cfg.vy), g(cfg.vz)])
        self.f = np.stack([g(cfg.fx), g(cfg.fy), g(cfg.fz)])


def read_posff(path, Fposff=2):
    cfg = Config()
    L = lib()
    _check(L.mdff_read_posff(os.fsencode(path), Fposff, C.byref(cfg)), "read_posff", False)
    try:
        return PosffData(cfg)
    finally:
        L.mdff_config_free(C.byref(cfg))


def write_config(path, data, mode="w"):
    """CONTFF/TRAJFF writer (config.c:291-345)."""
    L = lib()
    cfg = Config()
    _check(L.mdff_config_alloc(C.byref(cfg), data.nion), "config_alloc", False)
    try:
        cfg.ntype = data.ntype
        cfg.name = data.name.encode()[:80]
        for i in range(3):
            for j in range(3):
                cfg.A[i][j] = float(data.A[i, j])
        for t in range(data.ntype):
            cfg.atypit[t].value = data.atypit[t].encode()[:80]
            cfg.nionit[t] = data.nionit[t]
        for k, name in enumerate(("rx", "ry", "rz")):
            C.memmove(getattr(cfg, name), np.ascontiguousarray(data.r[k]).ctypes.data, 8 * data.nion)
        for k, name in enumerate(("vx", "vy", "vz")):
            C.memmove(getattr(cfg, name), np.ascontiguousarray(data.v[k]).ctypes.data, 8 * data.nion)
        for k, name in enumerate(("fx", "fy", "fz")):
            C.memmove(getattr(cfg, name), np.ascontiguousarray(data.f[k]).ctypes.data, 8 * data.nion)
        C.memmove(cfg.typia, np.ascontiguousarray(data.typia, dtype=np.int32).ctypes.data, 4 * data.nion)
        _check(L.mdff_write_config(os.fsencode(path), mode.encode(), C.byref(cfg)), "write_config", False)
    finally:
        L.mdff_config_free(C.byref(cfg))


def tt_damping(b, c, r, order):
    f, fd = C.c_double(), C.c_double()
    lib().mdff_tt_damping(b, c, r, C.byref(f), C.byref(fd), order)
    return f.value, fd.value


def init_velocities_host(params, mass):
    """Initial velocities of an MD start (kinetic.c:15-28) from rand() after
    srand(1); returns (3, n)."""
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    n = mass.size
    v = np.zeros((3, n))
    vx, vy, vz = v[0].copy(), v[1].copy(), v[2].copy()
    _check(lib().mdff_host_init_velocities(C.byref(params), n, _dp(mass), _dp(vx), _dp(vy), _dp(vz), 1),
           "init_velocities", False)
    return np.stack([vx, vy, vz])


class Engine:
    """One periodic system on one GPU (handle API of mdff_gpu.h)."""

    def __init__(self, params, typia, A, r, precision=FP64, device=0):
        L = lib()
        self._L = L
        self.params = params
        typia = np.ascontiguousarray(typia, dtype=np.int32)
        self.n = typia.size
        A = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(9))
        r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(3, self.n))
        h = C.c_void_p()
        rx, ry, rz = (np.ascontiguousarray(r[k]) for k in range(3))
        rc = L.mdff_create(C.byref(h), C.byref(params), self.n, typia.ctypes.data_as(C.POINTER(C.c_int)),
                           _dp(A), _dp(rx), _dp(ry), _dp(rz), precision, device)
        _check(rc, "create", False)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.mdff_destroy(self._h)
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

    # ---- state
    def _get3(self, fn):
        a = [np.zeros(self.n) for _ in range(3)]
        _check(fn(self._h, _dp(a[0]), _dp(a[1]), _dp(a[2])), fn.__name__, False)
        return np.stack(a)

    def _set3(self, fn, v):
        v = np.asarray(v, dtype=np.float64).reshape(3, self.n)
        a = [np.ascontiguousarray(v[k]) for k in range(3)]
        _check(fn(self._h, _dp(a[0]), _dp(a[1]), _dp(a[2])), fn.__name__, False)

    def set_positions(self, r):
        self._set3(self._L.mdff_set_positions, r)

    def set_velocities(self, v):
        self._set3(self._L.mdff_set_velocities, v)

    def positions(self):
        return self._get3(self._L.mdff_get_positions)

    def velocities(self):
        return self._get3(self._L.mdff_get_velocities)

    def forces(self):
        return self._get3(self._L.mdff_get_forces)

    def dipoles(self):
        mu = np.zeros((self.n, 3))
        _check(self._L.mdff_get_dipoles(self._h, _dp(mu)), "get_dipoles", False)
        return mu

    def field(self):
        ef = np.zeros((self.n, 3))
        _check(self._L.mdff_get_field(self._h, _dp(ef)), "get_field", False)
        return ef

    def thermo(self):
        t = Thermo()
        _check(self._L.mdff_get_thermo(self._h, C.byref(t)), "get_thermo", False)
        return t.as_dict()

    def ewald(self):
        a, k, nk = C.c_double(), (C.c_int * 3)(), C.c_int()
        _check(self._L.mdff_get_ewald(self._h, C.byref(a), k, C.byref(nk)), "get_ewald", False)
        return a.value, list(k), nk.value

    def timing(self):
        t = np.zeros(6)
        _check(self._L.mdff_get_timing(self._h, _dp(t)), "get_timing", False)
        return t

    def time_kernel(self, which, lo=0, iters=10):
        ms = C.c_double()
        _check(self._L.mdff_time_kernel(self._h, which, lo, iters, C.byref(ms)), "time_kernel", False)
        return ms.value

    def counters(self):
        c = (C.c_longlong * 4)()
        _check(self._L.mdff_get_counters(self._h, c), "get_counters", False)
        return dict(force_calls=c[0], scf_fp64=c[1], scf_inner=c[2], npol=c[3])

    def set_scf_solver(self, solver, inner):
        """0 the reference's Jacobi schedule with FP64 fields (the FP64 form's default), 2 anchored
        CG (the mixed form's default; symmetric polarisabilities); inner = CG steps launched before
        the first convergence poll"""
        _check(self._L.mdff_set_scf_solver(self._h, solver, inner), "set_scf_solver", False)

    def set_scf_check(self, every):
        """mixed form: the FP64 residual of the SCF's dipoles every `every`-th force call (thermo rmsd_fp64)"""
        _check(self._L.mdff_set_scf_check(self._h, every), "set_scf_check", False)

    def set_scf_refine(self, on=True):
        """mixed form, anchored CG: an FP64 refinement pass after the SCF (FP64 static field and anchor,
        low-precision correction steps until the FP64-anchored residual meets the tolerance)"""
        _check(self._L.mdff_set_scf_refine(self._h, 1 if on else 0), "set_scf_refine", False)

    def set_trajectory(self, tprint, tstart=0, npas=0, capacity=16):
        """device trajectory: r, v, f copied in device memory after every TRAJFF step of run_md
        (istep % tprint == 0 or istep == npas, istep >= tstart), up to `capacity` frames"""
        _check(self._L.mdff_set_trajectory(self._h, tprint, tstart, npas, capacity), "set_trajectory", False)

    def traj_info(self):
        o = (C.c_int * 3)()
        _check(self._L.mdff_traj_info(self._h, o), "traj_info", False)
        return dict(frames=o[0], capacity=o[1], dropped=o[2])

    def traj_frame(self, k):
        """(istep, r, v, f) of frame k, each (3, n)"""
        n = self.n
        out = np.zeros((3, 3, n))
        ptrs = [(C.POINTER(C.c_double) * 3)(*[_dp(out[q, a]) for a in range(3)]) for q in range(3)]
        st = C.c_int()
        _check(self._L.mdff_traj_frame(self._h, k, C.byref(st), *ptrs), "traj_frame", False)
        return st.value, out[0], out[1], out[2]

    def traj_clear(self):
        _check(self._L.mdff_traj_clear(self._h), "traj_clear", False)

    def set_list_method(self, method):
        """0: cell-grid neighbour build where it pays (default), 1: all pairs, 2: cell grid wherever it fits"""
        _check(self._L.mdff_set_list_method(self._h, method), "set_list_method", False)

    def list_info(self):
        o = (C.c_int * 8)()
        _check(self._L.mdff_list_info(self._h, o), "list_info", False)
        return dict(cells=bool(o[0]), nc=list(o[1:4]), reach=list(o[4:7]), cap=o[7])

    def set_tuning(self, lanes=0, katoms=0, ksplit=0):
        _check(self._L.mdff_set_tuning(self._h, lanes, katoms, ksplit), "set_tuning", False)

    def set_kspace_modes(self, eval_mode=0, inner_mode=0):
        """mixed form's matrix-core arithmetic: KS_F32 or KS_F16X3 for the static field and
        force call (eval_mode) and the low-precision SCF steps (inner_mode); 0 keeps"""
        _check(self._L.mdff_set_kspace_modes(self._h, eval_mode, inner_mode), "set_kspace_modes", False)

    def set_scf_anchor(self, anchor_mode=0):
        """anchored CG (SCF solver 2): the anchor field evaluation in FP64 (0) or in a matrix-core mode
        (KS_F32, KS_F16X3) over the FP32 tables"""
        _check(self._L.mdff_set_scf_anchor(self._h, anchor_mode), "set_scf_anchor", False)

    def set_kspace_shape(self, cw64=0, cw32=0, fwd_blocks=0, back_blocks=0):
        _check(self._L.mdff_set_kspace_shape(self._h, cw64, cw32, fwd_blocks, back_blocks), "set_kspace_shape",
               False)

    # ---- compute
    def init_velocities(self):
        _check(self._L.mdff_init_velocities(self._h), "init_velocities", False)

    def check_verlet(self):
        _check(self._L.mdff_check_verlet(self._h), "check_verlet", False)

    def check_verlet_noalloc(self):
        """the Verlet check and rebuild with no device allocation (fixed scratch per GPU)"""
        _check(self._L.mdff_check_verlet_noalloc(self._h), "check_verlet_noalloc", False)

    def engforce(self, istep=0):
        t = Thermo()
        rc = _check(self._L.mdff_engforce(self._h, istep, C.byref(t)), "engforce")
        d = t.as_dict()
        d["status"] = rc
        return d

    def pair_short(self):
        u, pv = C.c_double(), C.c_double()
        tau = np.zeros(9)
        _check(self._L.mdff_pair_short(self._h, C.byref(u), C.byref(pv), _dp(tau)), "pair_short", False)
        return u.value, pv.value, tau.reshape(3, 3)

    def dipoles_scf(self, istep=0):
        mu = np.zeros((self.n, 3))
        up = C.c_double()
        _check(self._L.mdff_get_dipoles_scf(self._h, istep, _dp(mu), C.byref(up)), "get_dipoles")
        return mu, up.value

    def multipole_es(self, q, mu, flags, forces=None, theta=None):
        """multipole_ES on host arrays; returns dict(u, pvir, tau, ef, efg,
        f) with f the coulombic-unit forces added to `forces` (or zeros).
        theta (quadrupoles, [n][3][3]) takes no part in any term, as in the
        reference."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        mu = np.ascontiguousarray(np.asarray(mu, dtype=np.float64).reshape(self.n, 3))
        theta = np.zeros((self.n, 9)) if theta is None else \
            np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(self.n, 9))
        u, pv = C.c_double(), C.c_double()
        tau = np.zeros(9)
        ef = np.zeros((self.n, 3))
        efg = np.zeros((self.n, 9))
        f = np.zeros((3, self.n)) if forces is None else np.array(forces, dtype=np.float64).reshape(3, self.n)
        fx, fy, fz = (np.ascontiguousarray(f[k]) for k in range(3))
        _check(self._L.mdff_multipole_es(self._h, _dp(q), _dp(mu), _dp(theta), flags, C.byref(u), C.byref(pv),
                                         _dp(tau), _dp(ef), _dp(efg), _dp(fx), _dp(fy), _dp(fz)),
               "multipole_es", False)
        return dict(u=u.value, pvir=pv.value, tau=tau.reshape(3, 3), ef=ef, efg=efg.reshape(self.n, 3, 3),
                    f=np.stack([fx, fy, fz]))

    def md_init(self):
        t = Thermo()
        _check(self._L.mdff_md_init(self._h, C.byref(t)), "md_init")
        return t.as_dict()

    def md_step(self, istep):
        t = Thermo()
        _check(self._L.mdff_md_step(self._h, istep, C.byref(t)), "md_step")
        return t.as_dict()

    def propagate(self, istep):
        """prop_agate: one integrator step with its force call (no list check, no rescale)"""
        t = Thermo()
        _check(self._L.mdff_propagate(self._h, istep, C.byref(t)), "propagate")
        return t.as_dict()

    def md_run(self, first_step, nsteps):
        t = Thermo()
        _check(self._L.mdff_md_run(self._h, first_step, nsteps, C.byref(t)), "md_run")
        return t.as_dict()


from . import batch  # noqa: E402  (batched engine and adjoint, include/mdff_batch.h)


def rdf_defaults(A):
    """cutgr and nbins of default_rdf (radial_distribution.c:37-42)"""
    A = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(9))
    cut, nb = C.c_double(), C.c_int()
    _check(lib().mdff_rdf_defaults(_dp(A), C.byref(cut), C.byref(nb)), "rdf_defaults", False)
    return cut.value, nb.value


class Rdf:
    """radial distribution histogram on the device (mdff_rdf.h; get_gr
    radial_distribution.c:180-231)"""

    def __init__(self, typia, ntype, nbins, resg, cutgr):
        L = lib()
        self._L = L
        typia = np.ascontiguousarray(typia, dtype=np.int32)
        self.n, self.ntype, self.nbins, self.resg, self.cutgr = typia.size, ntype, nbins, resg, cutgr
        st = C.c_int()
        h = L.mdff_rdf_create(self.n, ntype, typia.ctypes.data_as(C.POINTER(C.c_int)), nbins, resg, cutgr,
                              C.byref(st))
        if not h:
            raise MdffError(st.value, "rdf_create")
        self._h = C.c_void_p(h)

    def close(self):
        if getattr(self, "_h", None):
            self._L.mdff_rdf_destroy(self._h)
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

    def add_frame(self, A, r, write_back=False):
        """one host frame; returns the positions get_gr leaves (kardir + dirkar) when write_back"""
        A = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(9))
        r = np.asarray(r, dtype=np.float64)
        if r.shape != (3, self.n):
            raise ValueError(f"positions must be (3, {self.n}), got {r.shape}")
        rx, ry, rz = (np.ascontiguousarray(r[k]).copy() for k in range(3))
        _check(self._L.mdff_rdf_add_frame(self._h, _dp(A), _dp(rx), _dp(ry), _dp(rz), 1 if write_back else 0),
               "rdf_add_frame", False)
        return np.stack([rx, ry, rz]) if write_back else None

    def sample(self, engine):
        _check(self._L.mdff_rdf_sample(self._h, engine._h), "rdf_sample", False)

    def sample_trajectory(self, engine):
        _check(self._L.mdff_rdf_sample_trajectory(self._h, engine._h), "rdf_sample_trajectory", False)

    def reset(self):
        _check(self._L.mdff_rdf_reset(self._h), "rdf_reset", False)

    def time(self, engine, iters=10):
        """ms of device time per frame of the engine's positions (HIP events); resets the counts"""
        ms = C.c_double()
        _check(self._L.mdff_rdf_time(self._h, engine._h, iters, C.byref(ms)), "rdf_time", False)
        return ms.value

    def counts(self):
        """ordered pair counts [nbins][ntype][ntype] and (frames, pairs out of range)"""
        c = np.zeros((self.nbins, self.ntype, self.ntype), dtype=np.int64)
        info = (C.c_longlong * 2)()
        _check(self._L.mdff_rdf_counts(self._h, c.ctypes.data_as(C.POINTER(C.c_longlong)), info), "rdf_counts",
               False)
        return c, (info[0], info[1])

    def gr(self, counts, nconf, omega):
        """write_rdf's normalisation: r[nbins], g[nbins][ntype][ntype]"""
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        r = np.zeros(self.nbins)
        g = np.zeros((self.nbins, self.ntype, self.ntype))
        _check(self._L.mdff_rdf_gr(self._h, counts.ctypes.data_as(C.POINTER(C.c_longlong)), nconf, omega, _dp(r),
                                   _dp(g)), "rdf_gr", False)
        return r, g
