#!/usr/bin/env python3
"""Benchmark: MD steps/s of the 12,288-ion polarisable-ion (PIM) SiO2-like
melt on one MI355X, mixed form vs FP64 form.

Workload (synthetic, seeded): N = 12288 ions (4096 cations q=+2.4, 8192
anions q=-1.2, anion polarisability 1.2 A^3, Tang-Toennies b=2 c=1.5 k=4),
2.2 g/cm^3 (L = 57.06 A), BHMFT pair potential (the seeded parameter set of
tests/golden), multipole Ewald alpha = 0.3 1/A, rc = 9 A, kmax = 16 (17968
k-vectors), Verlet skin 1 A, SCF to 1e-8 on the field, NVT Nose-Hoover chain
(length 3) at 2500 K, dt = 1 fs.  The random 1.6 A packing is relaxed with
300 rigid-ion steps on the GPU before anything is timed.

Prints one JSON line.  `value` is the mixed form's steps/s; the FP64 form's
steps/s, the speed ratio and the force/dipole/energy/stress errors of the
mixed form against the FP64 form are in the same line.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

M_O, M_SI = 15.99994, 28.084
RHO_UNIT = 1.660538782
PEAK = {"f64": 78.6e12, "f32": 157.3e12, "f16": 2.5e15, "hbm": 8.0e12}
# BHMFT parameters of tests/golden (make_fixtures.pair_params, seed 20349):
# upper triangles O-O, O-Si, Si-Si
PAIR_A = (13843.775422359393, 9907.395718764405, 13891.629698472487)
PAIR_B = (2.6317230798523927, 2.977352637018226, 2.7189376696105843)
PAIR_C = (13.401132428691126, 18.44105648347189, 31.384716111068446)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def packing(n, L, seed, dmin=1.6):
    """uniform positions with minimum-image rejection, on a cell grid"""
    rng = np.random.default_rng(seed)
    nc = max(1, int(L // dmin))
    cs = L / nc
    grid = {}
    pos = np.zeros((n, 3))
    k = 0
    while k < n:
        p = rng.uniform(0.0, L, 3)
        c = (p // cs).astype(int) % nc
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = ((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc)
                    for j in grid.get(key, ()):
                        d = pos[j] - p
                        d -= L * np.rint(d / L)
                        if d @ d < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        pos[k] = p
        grid.setdefault(tuple(c), []).append(k)
        k += 1
    pos -= 0.5 * L
    return pos


def control(pim, alpha=0.3, kmax=16, rc=9.0, extra=""):
    t = ["massit 15.99994", "massit 28.084", "qit -1.2", "qit 2.4", "lbhmft true",
         "Abhmftd " + " ".join(repr(x) for x in PAIR_A),
         "Bbhmftd " + " ".join(repr(x) for x in PAIR_B),
         "Cbhmftd " + " ".join(repr(x) for x in PAIR_C),
         "Dbhmftd 0.0 0.0 0.0", "lcoulombic true", f"alphaES {alpha}", f"kES {kmax} {kmax} {kmax}",
         f"cutshortrange {rc}", f"cutlongrange {rc}", "lverletL true", "skindiff 1.0"]
    if pim:
        t += ["polit 1.2 0.0 0.0 0.0 1.2 0.0 0.0 0.0 1.2", "polit 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0",
              "lpoldamping 0 0 0 true", "pol_damp_b 0 0 0 2.0", "pol_damp_c 0 0 0 1.5", "pol_damp_k 0 0 0 4",
              "lpoldamping 0 0 1 true", "pol_damp_b 0 0 1 2.0", "pol_damp_c 0 0 1 1.5", "pol_damp_k 0 0 1 4",
              "algo_pim scf", "conv_tol_ind 1e-8", "min_scf_pim_iter 1", "max_scf_pim_iter 400",
              "algo_extrapolate_dipole aspc", "extrapolate_order 2",
              "integrator nvt-nhcn", "nhc_n 3", "nhc_yosh_order 3", "nhc_mults 2", "timesca_thermo 0.1",
              "temp 2500.0", "dt 0.001"]
    else:
        t += ["integrator nve-vv", "temp 2500.0", "dt 0.001", "tauTberendsen 0.001", "nequil 100000",
              "nequilT 1"]
    return "\n".join(t) + "\n" + extra


def params_of(text):
    import mdff20_amd as m
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "control")
        open(f, "w").write(text)
        return m.read_control(f, 2)


def system(nunits, seed):
    L = (nunits * (M_SI + 2 * M_O) / (2.2 / RHO_UNIT)) ** (1.0 / 3.0)
    pos = packing(3 * nunits, L, seed)
    typ = np.array([0] * (2 * nunits) + [1] * nunits, dtype=np.int32)
    return L, pos, typ


def equilibrate(L, pos, typ, steps, kmax=16):
    import mdff20_amd as m
    p = params_of(control(False, kmax=kmax))
    A = np.diag([L, L, L])
    e = m.Engine(p, typ, A, pos.T.copy(), precision=m.FP64)
    e.init_velocities()
    e.md_init()
    t0 = time.time()
    th = e.md_run(1, steps)
    r = e.positions()
    e.close()
    log(f"rigid-ion relaxation: {steps} steps in {time.time() - t0:.1f} s, T = {th['temp_r']:.0f} K")
    return r


def batch_systems(nconf, seed, nmin=96, nmax=1536, rho=(2.0, 2.6), eps=1e-6):
    """uneven batch of SiO2-like cells: N log-uniform in [nmin, nmax] (a
    multiple of 3), density uniform in rho (g/cm^3), seeded 1.6 A packings;
    per-cell Ewald alpha / kmax / cutoffs for relative accuracy eps
    (rc = min(9 A, L/2))"""
    import mdff20_amd as m
    rng = np.random.default_rng(seed)
    confs, ews = [], []
    for c in range(nconf):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        nu = max(nmin // 3, min(nmax // 3, int(round(n / 3))))
        dens = rng.uniform(*rho)
        L = (nu * (M_SI + 2 * M_O) / (dens / RHO_UNIT)) ** (1.0 / 3.0)
        pos = packing(3 * nu, L, int(rng.integers(1 << 30)))
        typ = np.array([0] * (2 * nu) + [1] * nu, dtype=np.int32)
        confs.append((typ, np.diag([L, L, L]), pos.T.copy()))
        ews.append(m.batch.ewald_for_accuracy(L, eps=eps))
    return confs, ews


def batch_relax(confs, ews, iters=30, gamma=0.01, cap=0.05):
    """steepest descent of the rigid-ion model on a whole batch (moves
    capped at `cap` A per ion and iteration): separates the close pairs of
    a raw packing before the polarisable model sees it"""
    import mdff20_amd as m
    p = params_of(control(False))
    t0 = time.time()
    with m.batch.Batch(p, confs, ews, precision=m.FP64) as b:
        r = np.concatenate([c[2] for c in confs], axis=1)
        for _ in range(iters):
            b.set_positions(r)
            b.engforce()
            d = gamma * b.forces()
            s = np.sqrt(np.sum(d * d, axis=0))
            d *= np.minimum(1.0, cap / np.maximum(s, 1e-300))
            r = r + d
        parts = b.split(r)
    log(f"batch relaxation: {len(confs)} cells, {iters} steepest-descent steps in {time.time() - t0:.1f} s")
    return [(t, A, rr.copy()) for (t, A, _), rr in zip(confs, parts)]


def batch_bench(seed=4096, which=("fp64", "mixed", "cfg5")):
    """the batched engine on its two workloads: 512 uneven cells (N log-uniform
    in [96, 1536], per-cell Ewald for 1e-6) force calls in the FP64 and mixed
    forms, and the force-matching adjoint on 64 cells of 648 ions from NVT
    dynamics (Adam steps through the autograd wrapper)"""
    import mdff20_amd as m
    out = {}
    confs, ews = batch_systems(512, seed)
    confs = batch_relax(confs, ews)
    p = params_of(control(True))
    nion = sum(len(c[0]) for c in confs)
    for prec, name in ((m.FP64, "fp64"), (m.MIXED, "mixed")):
        if name not in which:
            continue
        with m.batch.Batch(p, confs, ews, precision=prec) as b:
            b.engforce()
            # each call a fresh static evaluation, SCF from zero dipoles as
            # the reference's engforce on a new configuration; then calls
            # that start from the previous call's dipoles
            t0 = time.perf_counter()
            for _ in range(3):
                b.reset_dipoles()
                th = b.engforce()
            dt = (time.perf_counter() - t0) / 3
            nscf = [t["n_scf"] for t in th]
            t0 = time.perf_counter()
            for _ in range(3):
                b.engforce()
            dw = (time.perf_counter() - t0) / 3
        out[f"cfg4_{name}"] = dict(ms_per_force_call=1e3 * dt, configs_per_s=512 / dt, ions_per_s=nion / dt,
                                   scf_sweeps_mean=float(np.mean(nscf)), scf_sweeps_max=int(np.max(nscf)),
                                   warm_ms_per_force_call=1e3 * dw)
    out["cfg4_cells"] = dict(n=512, ions=nion)
    if "cfg5" not in which:
        return out
    # cfg5: 4 seeded starts, 16 cells each from NVT dynamics (kmax 8, 21.4 A cell)
    q = params_of(control(True, alpha=0.3, kmax=8, rc=9.0))
    q.conv_tol_ind = 1e-10
    cells = []
    for s5 in range(4):
        L, pos, typ = system(216, 648 + s5)
        r = equilibrate(L, pos, typ, 200, kmax=8)
        A = np.diag([L, L, L])
        with m.Engine(q, typ, A, r, precision=m.FP64) as e:
            e.init_velocities()
            e.md_init()
            for k in range(16):
                e.md_run(10 * k + 1, 10)
                cells.append((typ, A, e.positions()))
    ew = dict(alpha=q.alphaES, kES=tuple(q.kES), rcut_long=q.cutlongrange, rcut_short=q.cutshortrange)
    try:
        from mdff20_amd.adjoint import fit
    except ImportError:
        return out
    with m.batch.Batch(q, cells, [ew] * len(cells)) as b:
        spec = m.batch.ParamSpec.default(q, q.ntype)
        rng = np.random.default_rng(11)
        v = spec.values()
        b.set_params(spec.params(v + np.log(rng.uniform(0.9, 1.1, v.size))))
        b.engforce()
        fref = b.forces()
        b.set_params(q)
        fit(b, spec, fref, steps=2, lr=0.01)
        t0 = time.perf_counter()
        _, hist = fit(b, spec, fref, steps=20, lr=0.01)
        dt = time.perf_counter() - t0
    out["cfg5_adjoint"] = dict(cells=64, ions_per_cell=648, adam_steps=20, steps_per_s=20 / dt,
                               loss_first=float(hist[0]), loss_last=float(hist[-1]))
    log(f"batch: {out}")
    return out


def timed_run(p, typ, A, r, precision, steps, warmup, fp64_cg=False, list_method=0, check=True, refine=False):
    import mdff20_amd as m
    e = m.Engine(p, typ, A, r, precision=precision)
    if list_method:
        e.set_list_method(list_method)
    if precision == m.MIXED:
        e.set_kspace_modes(VARIANT[MIXED_EVAL][1], VARIANT[MIXED_INNER][1])
        e.set_scf_anchor(VARIANT[MIXED_ANCHOR][1])
        if refine:
            e.set_scf_refine(True)
    elif fp64_cg:
        e.set_scf_solver(2, 0)
    e.init_velocities()
    e.md_init()
    if warmup:
        e.md_run(1, warmup)
    c0 = e.counters()
    t0 = time.perf_counter()
    th = e.md_run(warmup + 1, steps)
    t1 = time.perf_counter()
    c1 = e.counters()
    stats = dict(scf_fp64_per_step=(c1["scf_fp64"] - c0["scf_fp64"]) / steps,
                 scf_inner_per_step=(c1["scf_inner"] - c0["scf_inner"]) / steps,
                 rmsd=th["rmsd"], temp=th["temp_r"], h_tot=th["h_tot"])
    if precision == m.MIXED and check:
        # after the timed region: the FP64 residual of the SCF's final
        # dipoles (the SCF stops on its carried low-precision residual)
        e.set_scf_check(1)
        chk = []
        t2 = time.perf_counter()
        for k in range(5):
            chk.append(e.md_step(warmup + steps + 1 + k)["rmsd_fp64"])
        stats["scf_rmsd_fp64"] = dict(steps=5, mean=float(np.mean(chk)), max=float(np.max(chk)),
                                      tol=p.conv_tol_ind,
                                      ms_per_step_with_check=1e3 * (time.perf_counter() - t2) / 5)
        e.set_scf_check(0)
    return e, (t1 - t0), stats


def cpu_baseline(p_text, L, pos, typ, n_scf, threads):
    """the reference's own routines (oracle/_ref) on the host cores: one
    static field, n_ind induced-dipole Ewald calls and the final force call
    at this N; steps/s from the SCF iteration count of the FP64 GPU run"""
    from oracle import refrun
    if not refrun.available():
        return None
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "control"), "w").write(p_text + "lstatic true\n")
        nu = len(typ) // 3
        with open(os.path.join(d, "POSFF"), "w") as fh:
            fh.write(f"{len(typ)}\nbench\n{L!r} 0.0 0.0\n0.0 {L!r} 0.0\n0.0 0.0 {L!r}\n2\nO Si\n{2 * nu} {nu}\n"
                     "Cartesian\n")
            for i in range(len(typ)):
                fh.write(f"{'O' if typ[i] == 0 else 'Si'} {pos[0, i]!r} {pos[1, i]!r} {pos[2, i]!r}\n")
        n_ind = 2
        t0 = time.time()
        res = refrun.run(d, "time", n_ind, threads=threads, o3=refrun.available(o3=True), timeout=900)
        o3 = refrun.available(o3=True)
        wall = time.time() - t0
    t_step = res["t_short"] + res["t_static"] + n_scf * res["t_induced"] + res["t_force"]
    return dict(value=1.0 / t_step, unit="steps/s", cores=int(res["threads"]), kind="reference",
                sample=(f"reference build ({'-O3 -Ofast' if o3 else 'shipped -O0'} flags, OpenMP) at N={len(typ)}: "
                        f"short range, static field, "
                        f"{n_ind} induced-dipole Ewald calls and the force call timed once ({wall:.1f} s wall); "
                        f"step = t_short + t_static + {n_scf:.1f} x t_induced + t_force "
                        f"(SCF iterations per step of the FP64 GPU run)"),
                t_short=res["t_short"], t_static=res["t_static"], t_induced=res["t_induced"],
                t_force=res["t_force"])


# the S(k) GEMM kernels behind mdff_time_kernel's (which, lo): kernel-name
# prefix (the template's first argument is the arithmetic mode: 0 FP64,
# 2 FP32, 3 split-FP16)
TIMED = {"sk_forward_scf": (0, "k_gemm_fwd"), "sk_back_scf": (1, "k_gemm_back"),
         "sk_forward_force": (3, "k_gemm_fwd"), "sk_back_force": (4, "k_gemm_back")}
VARIANT = {"f64": (0, 0), "f32": (1, 2), "f16x3": (3, 3)}   # name -> (lo, template mode)
# arithmetic of the mixed form's matrix-core sums: static field and force
# call (eval), SCF steps (inner) and the anchor evaluation of the SCF
# (booked with the inner evaluations when it is not FP64)
MIXED_EVAL, MIXED_INNER, MIXED_ANCHOR = "f16x3", "f16x3", "f16x3"
CALIB_BYTES = 512 << 20


# the real-space pair kernels behind mdff_time_kernel's which 5 and 6
PAIR = {"short": (5, "k_short"), "dir_force": (6, "k_ewald_dir")}


def kernel_prefix(kname):
    name, var = kname.split("/")
    if name in PAIR:
        which, pre = PAIR[name]
        lo = VARIANT[var][0]
        return which, lo, f"{pre}<{'float' if lo else 'double'},"
    which, pre = TIMED[name]
    lo, mode = VARIANT[var]
    return which, lo, f"{pre}<{mode},"


def traffic_file(kname):
    name, var = kname.split("/")
    return os.path.join(ROOT, "profiles", f"{name}_{var}_pim12288_traffic.json")


def pmc_run(args, kname):
    """one plain device copy of CALIB_BYTES (the counters' scale) and then
    `--pmc-iters` launches of one S(k) GEMM kernel at the benchmark state;
    run under rocprofv3 --pmc FETCH_SIZE and again with WRITE_SIZE"""
    import ctypes
    import mdff20_amd as m
    which, lo, _ = kernel_prefix(kname)
    L, pos, typ = system(args.units, args.seed)
    if args.state and os.path.exists(args.state):
        r = np.load(args.state, allow_pickle=False)["r"]
    else:
        r = equilibrate(L, pos, typ, args.relax)
    A = np.diag([L, L, L])
    p = params_of(control(True))
    e = m.Engine(p, typ, A, r, precision=m.MIXED)
    e.set_kspace_modes(VARIANT[MIXED_EVAL][1], VARIANT[MIXED_INNER][1])
    e.set_scf_anchor(VARIANT[MIXED_ANCHOR][1])
    e.engforce(0)
    # the runtime's own device-to-device copy kernel over CALIB_BYTES
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(CALIB_BYTES)) != 0:
            raise RuntimeError("hipMalloc failed")
    if (hip.hipMemset(src, 0, ctypes.c_size_t(CALIB_BYTES)) != 0 or hip.hipDeviceSynchronize() != 0 or
            hip.hipMemcpy(dst, src, ctypes.c_size_t(CALIB_BYTES), 3) != 0 or hip.hipDeviceSynchronize() != 0):
        raise RuntimeError("calibration copy failed")
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"calibration copy: {CALIB_BYTES} bytes read and written")
    ms = e.time_kernel(which, lo=lo, iters=args.pmc_iters)
    log(f"{kname}: {args.pmc_iters} launches, {ms:.4f} ms each")
    e.close()


def read_counters(d):
    import csv
    import glob
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        rows += list(csv.DictReader(open(f)))
    return rows


def pmc_summarize(args, kname):
    """traffic of one kernel launch from the two counter runs: per-dispatch
    FETCH_SIZE and WRITE_SIZE (KiB) of the kernel's timed launches, each
    divided by what the same counter reads for the calibration copy per byte
    it moves"""
    _, _, prefix = kernel_prefix(kname)
    out = {"kernel": kname, "kernel_prefix": prefix}
    for counter, d in (("FETCH_SIZE", args.pmc_summarize[0]), ("WRITE_SIZE", args.pmc_summarize[1])):
        rows = [x for x in read_counters(d) if x.get("Counter_Name") == counter]
        by = {}
        for x in rows:
            by.setdefault((x["Dispatch_Id"], x["Kernel_Name"]), 0.0)
            by[(x["Dispatch_Id"], x["Kernel_Name"])] += float(x["Counter_Value"])
        disp = sorted(((int(k[0]), k[1], v) for k, v in by.items()), key=lambda t: t[0])
        copies = [v for _, nm, v in disp if "opy" in nm and v * 1024 > 0.5 * CALIB_BYTES]
        if not copies:
            raise RuntimeError(f"{counter}: no calibration copy dispatch found in {d}")
        factor = max(copies) * 1024.0 / CALIB_BYTES
        ks = [v for _, nm, v in disp if prefix in nm][-args.pmc_iters:]
        if not ks:
            raise RuntimeError(f"{counter}: no dispatch of {prefix} in {d}")
        kib = sum(ks) / len(ks)
        out[counter] = dict(kib_per_dispatch=kib, dispatches=len(ks), copy_factor=factor,
                            bytes=kib * 1024.0 / factor)
    out["traffic_bytes"] = out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]
    with open(traffic_file(kname), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


def pmc_table(args, kname, dirs):
    """per-dispatch averages of every counter found in the rocprofv3 --pmc
    output directories `dirs` for the timed launches of one kernel"""
    _, _, prefix = kernel_prefix(kname)
    out = {}
    for d in dirs:
        rows = read_counters(d)
        by = {}
        for x in rows:
            if prefix not in x["Kernel_Name"]:
                continue
            key = (x["Counter_Name"], int(x["Dispatch_Id"]))
            by[key] = by.get(key, 0.0) + float(x["Counter_Value"])
        for cname in sorted({k[0] for k in by}):
            vals = [v for (c, _), v in sorted(by.items(), key=lambda t: t[0][1]) if c == cname][-args.pmc_iters:]
            out[cname] = sum(vals) / len(vals)
    return out


def trace_summarize(d, top=40, kname=None, last=5):
    """markdown table of a rocprofv3 --kernel-trace --stats run; with kname
    (a --pmc-kernel run) also the average duration of that kernel's last
    `last` dispatches, the launches its time_kernel call timed"""
    import csv
    import glob
    f = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)[0]
    rows = list(csv.DictReader(open(f)))
    print("| kernel | calls | total ms | % | avg us | min us | max us |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for x in rows[:top]:
        print(f"| `{x['Name'][:100]}` | {x['Calls']} | {float(x['TotalDurationNs']) / 1e6:.2f} | "
              f"{float(x['Percentage']):.1f} | {float(x['AverageNs']) / 1e3:.2f} | {float(x['MinNs']) / 1e3:.2f} | "
              f"{float(x['MaxNs']) / 1e3:.2f} |")
    if kname:
        _, _, prefix = kernel_prefix(kname)
        t = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)[0]
        disp = sorted((int(x["Start_Timestamp"]), int(x["End_Timestamp"])) for x in csv.DictReader(open(t))
                      if prefix.replace(" ", "") in x["Kernel_Name"].replace(" ", ""))[-last:]
        us = [(b - a) / 1e3 for a, b in disp]
        print(f"\n{kname} ({prefix}...): last {len(us)} dispatches, average {sum(us) / len(us):.2f} us "
              f"(min {min(us):.2f}, max {max(us):.2f})")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--units", type=int, default=4096, help="SiO2 units (N = 3 x units)")
    ap.add_argument("--relax", type=int, default=300)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--forms", default="fp64,fp64_cg,mixed,mixed_refine",
                    help="fp64 (the reference's Jacobi SCF), fp64_cg (FP64 with the anchored CG SCF), mixed, "
                         "mixed_refine (mixed with the FP64 refinement pass of the SCF)")
    ap.add_argument("--pmc-kernel", default=None, help="e.g. sk_forward_scf/f16x3 (counter runs)")
    ap.add_argument("--pmc-iters", type=int, default=5)
    ap.add_argument("--pmc-summarize", nargs=2, default=None, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--pmc-table", nargs="+", default=None, metavar="DIR",
                    help="per-dispatch averages of all counters of --pmc-kernel in these rocprofv3 output dirs")
    ap.add_argument("--trace-summarize", default=None, metavar="DIR")
    ap.add_argument("--list-method", type=int, default=0, help="0 cell-grid neighbour build (default), 1 all pairs")
    ap.add_argument("--no-kernels", action="store_true",
                    help="skip the per-kernel timing of every arithmetic variant, the accuracy call and the "
                         "FP64 residual check (a kernel trace then holds the timed MD steps of the chosen forms only)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-engine workloads (512 cells, adjoint)")
    ap.add_argument("--batch-only", default=None, metavar="LIST",
                    help="run only these batched-engine workloads (comma-separated: fp64, mixed, cfg5) and print "
                         "their figures (kernel traces of one form)")
    ap.add_argument("--state", default=None, metavar="NPZ",
                    help="relaxed positions: read from NPZ when it exists, else relaxed and written there")
    args = ap.parse_args()
    if args.trace_summarize:
        trace_summarize(args.trace_summarize, kname=args.pmc_kernel, last=args.pmc_iters)
        return
    if args.pmc_summarize:
        pmc_summarize(args, args.pmc_kernel)
        return
    if args.pmc_table:
        print(json.dumps({"kernel": args.pmc_kernel, "prefix": kernel_prefix(args.pmc_kernel)[2],
                          "per_dispatch": pmc_table(args, args.pmc_kernel, args.pmc_table)}, indent=1))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (an engine runs one system "
              f"or one batch on one GPU)", file=sys.stderr)
        sys.exit(2)
    import mdff20_amd as m
    if m.device_count() < 1:
        raise RuntimeError("no HIP device")
    if args.pmc_kernel:
        pmc_run(args, args.pmc_kernel)
        return
    if args.batch_only:
        print(json.dumps(batch_bench(which=tuple(args.batch_only.split(",")))))
        return
    L, pos, typ = system(args.units, args.seed)
    n = len(typ)
    log(f"N = {n}, L = {L:.3f} A")
    if args.state and os.path.exists(args.state):
        r = np.load(args.state, allow_pickle=False)["r"]
        log(f"relaxed positions read from {args.state}")
    else:
        r = equilibrate(L, pos, typ, args.relax)
        if args.state:
            np.savez(args.state, r=r)
    A = np.diag([L, L, L])
    p_text = control(True)
    p = params_of(p_text)
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    res = {}
    eng = {}
    # the engine of the reported form stays open for the kernel timings below;
    # the others are closed after their timed run (each form is timed alone)
    head = "mixed" if "mixed" in forms else forms[0]
    for f in forms:
        prec = m.MIXED if f in ("mixed", "mixed_refine") else m.FP64
        e, dt, stats = timed_run(p, typ, A, r, prec, args.steps, args.warmup, fp64_cg=(f == "fp64_cg"),
                                 list_method=args.list_method, check=not args.no_kernels,
                                 refine=(f == "mixed_refine"))
        res[f] = dict(seconds=dt, steps_per_s=args.steps / dt, **stats)
        if f == head:
            eng[f] = e
        else:
            e.close()
        log(f"{f}: {args.steps} steps in {dt:.3f} s = {args.steps / dt:.2f} steps/s, "
            f"SCF FP64 evaluations/step {stats['scf_fp64_per_step']:.1f}, inner {stats['scf_inner_per_step']:.1f}")
    # accuracy of the mixed form against the FP64 form on one configuration
    acc = None
    acc_refine = None
    if "fp64" in forms and "mixed" in forms and not args.no_kernels:
        ef = m.Engine(p, typ, A, r, precision=m.FP64)
        tf = ef.engforce(0)
        Ff, muf = ef.forces(), ef.dipoles()

        def accuracy(refine):
            em = m.Engine(p, typ, A, r, precision=m.MIXED)
            if refine:
                em.set_scf_refine(True)
            tm = em.engforce(0)
            Fm, mum = em.forces(), em.dipoles()
            em.close()
            return dict(
                force_rms_rel=float(np.sqrt(np.mean((Fm - Ff) ** 2)) / np.sqrt(np.mean(Ff ** 2))),
                dipole_rms_rel=float(np.linalg.norm(mum - muf) / np.linalg.norm(muf)),
                energy_rel=float(abs(tm["u_tot"] - tf["u_tot"]) / abs(tf["u_tot"])),
                stress_rel=float(np.linalg.norm(np.array(tm["tau_coul"]) - np.array(tf["tau_coul"]))
                                 / np.linalg.norm(np.array(tf["tau_coul"]))))
        acc = accuracy(False)
        log(f"mixed vs fp64: {acc}")
        if "mixed_refine" in forms:
            acc_refine = accuracy(True)
            log(f"mixed with FP64 refinement vs fp64: {acc_refine}")
        ef.close()
    # kernels of the reciprocal-space path, timed live with HIP events on the
    # context stream, in every arithmetic variant the run uses; the roofline
    # entry is the variant with the largest share of the step time
    e = eng[head]
    a_, k_, nk = e.ewald()
    npol = e.counters()["npol"]
    st = res[head]
    variants = [(0, "f64", PEAK["f64"])] + ([(1, "f32", PEAK["f32"]), (3, "f16x3", PEAK["f16"])]
                                           if head == "mixed" else [])
    kern = {}
    for lo, vname, peak in ([] if args.no_kernels else variants):
        for which, name, ch in [(0, "sk_forward_scf", 3), (1, "sk_back_scf", 3), (2, "dir_induced", 0),
                                (3, "sk_forward_force", 4), (4, "sk_back_force", 9)]:
            if which == 2 and lo >= 2:
                continue
            try:
                ms = e.time_kernel(which, lo=lo, iters=10)
            except Exception as ex:  # noqa
                log(f"kernel {name}/{vname}: {ex}")
                continue
            nsel = npol if which in (0, 1, 2) else n
            flops = 8.0 * nsel * nk * ch if ch else None
            # launches per MD step of this kernel in this variant (SCF:
            # FP64 evaluations, and in the mixed form FP32 real space and
            # MIXED_INNER k-space for the low-precision SCF steps; force
            # form: the final force call, MIXED_EVAL k-space in the mixed form)
            if which in (0, 1, 2):
                if head == "mixed":
                    inner_lo = 1 if which == 2 else VARIANT[MIXED_INNER][0]
                    per = (st["scf_fp64_per_step"] if lo == 0 else
                           st["scf_inner_per_step"] if lo == inner_lo else 0.0)
                else:
                    per = st["scf_fp64_per_step"]
            else:
                per = (1.0 if lo == VARIANT[MIXED_EVAL][0] else 0.0) if head == "mixed" else 1.0
            kern[f"{name}/{vname}"] = dict(ms=ms, flops=flops, per_step=per, peak=peak,
                                           tflops=(flops / (ms * 1e-3) / 1e12) if flops else None)
    log(f"kernels ({head}): {json.dumps(kern)}")
    # neighbour-list rebuild at this N: the cell-grid build against the
    # all-pairs scan (the list is rebuilt at the current positions)
    rebuild = None
    if not args.no_kernels:
        info = e.list_info()
        t_cells = e.time_kernel(7, iters=5)
        e.set_list_method(1)
        t_all = e.time_kernel(7, iters=5)
        e.set_list_method(0)
        rebuild = dict(cells_ms=t_cells, all_pairs_ms=t_all, grid=info["nc"], reach=info["reach"])
        log(f"neighbour rebuild: {rebuild}")
    # the radial distribution of the run's configuration (get_gr's histogram
    # of every pair within half the box), timed live on the engine's stream;
    # its CPU baseline is the oracle's get_gr restatement on one host core
    rdf = None
    if not args.no_kernels:
        cut, nb = m.rdf_defaults(A)
        with m.Rdf(typ, 2, nb, 0.1, cut) as h:
            ms = h.time(e, iters=10)
        pairs = n * (n - 1) / 2
        rdf = dict(ms_per_frame=ms, frames_per_s=1e3 / ms, pair_tests_per_s=pairs / (ms * 1e-3), nbins=nb,
                   cutgr=cut, pair_tests_per_frame=pairs)
        if not args.no_cpu:
            from oracle import oracle as O
            r_now = e.positions()
            t0 = time.perf_counter()
            O.rdf_frame(typ, 2, A, r_now, cut, 0.1, nb)
            rdf["cpu_baseline"] = dict(value=1.0 / (time.perf_counter() - t0), unit="frames/s", cores=1, kind="port",
                                       sample=f"oracle get_gr restatement, one {n}-ion frame, ordered pairs")
        log(f"rdf: {rdf}")
    batch = None if (args.no_kernels or args.no_batch) else batch_bench()
    roof = None
    mm = [(v["ms"] * v["per_step"], k) for k, v in kern.items() if v["flops"]]
    if mm:
        _, kname = max(mm)
        kv = kern[kname]
        ach = kv["flops"] / (kv["ms"] * 1e-3)
        name, vname = kname.split("/")
        traffic, tsrc = None, None
        tf = traffic_file(kname)
        if os.path.exists(tf):
            traffic = json.load(open(tf))["traffic_bytes"]
            tsrc = os.path.relpath(tf, ROOT)
        roof = dict(bound="mfma", achieved=ach / 1e12, peak=kv["peak"] / 1e12, unit="TFLOP/s",
                    frac=ach / kv["peak"], traffic=traffic, traffic_source=tsrc, kernel=kname,
                    work=(f"8 flops x {npol if 'scf' in name else n} ions x {nk} k-vectors x "
                          f"{3 if 'scf' in name else (4 if 'forward' in name else 9)} channels per launch"))
    cpu = None
    if not args.no_cpu and "fp64" in res:
        try:
            cpu = cpu_baseline(p_text, L, r, typ, res["fp64"]["scf_fp64_per_step"] + 1, args.cpu_threads)
        except Exception as ex:  # noqa
            log(f"cpu baseline failed: {ex}")
    head_res = res[head]
    out = {
        "metric": "MD steps/s, 12k-ion PIM melt: mixed vs FP64 GPU at force err<=1e-5",
        "value": head_res["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / head_res["steps_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16x3/f32/f64" if head == "mixed" else "f64",
        "data": "synthetic (seeded SiO2-like PIM melt, rigid-ion relaxed)",
        "config": {"workload": f"PIM multipole-Ewald NVT MD, N={n}, kmax=16 ({nk} k), alpha=0.3, rc=9 A, "
                               f"skin 1 A, SCF tol 1e-8, NHC-3 2500 K, dt 1 fs",
                   "form": head},
        "ns_per_day": head_res["steps_per_s"] * 1.0 * 0.0864,
        "forms": res,
        "mixed_over_fp64": (res["mixed"]["steps_per_s"] / res["fp64"]["steps_per_s"])
        if ("mixed" in res and "fp64" in res) else None,
        # the speed-up split into the solver (anchored CG against the
        # reference's Jacobi schedule, both FP64) and the arithmetic (mixed
        # against FP64, both anchored CG)
        "fp64cg_over_fp64": (res["fp64_cg"]["steps_per_s"] / res["fp64"]["steps_per_s"])
        if ("fp64_cg" in res and "fp64" in res) else None,
        "mixed_over_fp64cg": (res["mixed"]["steps_per_s"] / res["fp64_cg"]["steps_per_s"])
        if ("mixed" in res and "fp64_cg" in res) else None,
        "mixed_refine_over_fp64": (res["mixed_refine"]["steps_per_s"] / res["fp64"]["steps_per_s"])
        if ("mixed_refine" in res and "fp64" in res) else None,
        "accuracy_mixed_vs_fp64": acc,
        "accuracy_mixed_refine_vs_fp64": acc_refine,
        "kernels": kern,
        "neighbour_rebuild": rebuild,
        "rdf": rdf,
        "batch": batch,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    for e in eng.values():
        e.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
