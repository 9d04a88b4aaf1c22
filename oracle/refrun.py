"""Runner of the reference build in oracle/_ref (TEST INFRASTRUCTURE).

Runs _ref/mdff20_ref_driver on a directory holding `control` and `POSFF`
and parses its %.17g dump.  Used by tests/golden/make_fixtures.py to write
golden vectors, and by bench.py's CPU-baseline leg to time the reference's
own routines.  _ref/ is built by `make -C oracle ref` in the development
container and travels as built files.
"""
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(_HERE, "_ref", "mdff20_ref_driver")
DRIVER_O3 = os.path.join(_HERE, "_ref", "mdff20_ref_driver_O3")
# the same driver with the force path replaced by the same-name drop-ins
DRIVER_DROPIN = os.path.join(_HERE, "_ref", "mdff20_dropin_driver")


def available(o3=False):
    return os.path.exists(DRIVER_O3 if o3 else DRIVER)


def run(workdir, mode, *args, threads=1, o3=False, timeout=3600, exe=None, env_extra=None):
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    if env_extra:
        env.update(env_extra)
    exe = exe or (DRIVER_O3 if o3 else DRIVER)
    subprocess.run([exe, workdir, mode] + [str(a) for a in args], check=True, env=env, timeout=timeout)
    return parse(os.path.join(workdir, "ref_out.txt"))


def parse(path):
    """-> dict with scalars and a list of per-step records."""
    head, steps = {}, []
    cur = head
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            key = t[0]
            if key == "step":
                cur = {"step": int(t[1])}
                steps.append(cur)
                continue
            if key == "params":
                cur[key] = t[2]
            elif key in ("n_scf", "verlet_updates", "nk", "threads", "n_induced"):
                cur[key] = int(t[1])
            elif key == "kES":
                cur[key] = [int(x) for x in t[1:4]]
            elif len(t) == 2:
                cur[key] = float(t[1])
            else:
                cur[key] = np.array([float(x) for x in t[2:]])
    head["steps"] = steps
    return head
