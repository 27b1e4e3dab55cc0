"""Write the golden fixtures of the bench-size parity tests from the oracle.

The oracle (oracle/gatt_oracle.c, single-threaded C) needs minutes for these sizes, so
it runs here once and the GPU tests compare against what it wrote:

* box256_step1.npz: the 256^3 forced turbulence box (cases.turbulence_box(256)) after
  one main-loop step from its seeded initial state;
* bl1024x256_steps1000.npz: the 1024 x 256 boundary layer (cases.boundary_layer()) after
  1000 steps from its initial state with the boundary conditions applied.

Each file holds 4096 sampled 0-based column-major point indices (`idx`), the five fields
there after the run (`vals`, rows U, V, W, R, E), the largest magnitude of each field
over the whole final state (`qmax`), the max changes of the last step (`mc`), the steps
run (`steps`) and the sums of the initial fields (`check`), which the tests compare with
the case they rebuild to make sure the inputs are the same.

    python tests/golden/make_fixtures.py box
    python tests/golden/make_fixtures.py bl
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gatt_dns_amd.prep import cases  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = "UVWRE"


def sample(case, seed=2026, count=4096):
    return np.sort(np.random.default_rng(seed).choice(case.npoints, size=count, replace=False)).astype(np.int64)


def initial_check(case):
    return np.array([np.sum(case.state0[k]) for k in FIELDS])


def write(name, case, st, mc, steps):
    idx = sample(case)
    vals = np.stack([np.asarray(st[k]).ravel(order="F")[idx] for k in FIELDS])
    qmax = np.array([np.max(np.abs(st[k])) for k in FIELDS])
    path = os.path.join(HERE, name)
    np.savez_compressed(path, idx=idx, vals=vals, qmax=qmax, mc=np.asarray(mc), steps=steps,
                        check=initial_check(case))
    print("wrote", path, flush=True)


def box():
    case = cases.turbulence_box(256)
    o = Oracle(case)
    t = time.time()
    st, _, mc, dv = o.step(case.state0, 0.0, case.dt)
    assert not dv
    print("oracle step: %.1f s" % (time.time() - t), flush=True)
    write("box256_step1.npz", case, st, mc, 1)


def bl(steps=1000):
    case = cases.boundary_layer()
    o = Oracle(case)
    q0 = o.apply_bound_cond(case.state0, 0.0)
    o.init_mean(q0)
    t = time.time()
    st, mc, dv = o.run(q0, 0.0, case.dt, steps)
    assert not dv
    print("oracle %d steps: %.1f s" % (steps, time.time() - t), flush=True)
    write("bl1024x256_steps%d.npz" % steps, case, st, mc, steps)


if __name__ == "__main__":
    which = sys.argv[1:] or ["box", "bl"]
    for w in which:
        {"box": box, "bl": bl}[w]()
