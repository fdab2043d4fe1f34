"""Write tests/golden/batch512_device_order.npz: the oracle PCG-DIC on every
system of the 512-system batch (the systems of make_batch_outcomes.py), each
solved alone to rel 1e-8 (tolerance 1e-300, max_iter 30,000) with its dot
products summed in the batched device solver's order for 32-slice tiles
(``oracle.krylov.dot_device(n, tile_slices=32)``).  Per system: iterations
(len(history)), converged flag, final residual, initial residual and a SHA-1
of the solution's bytes, so that a batched device solve with 32-slice tiles
can be checked system by system, bit for bit.

Run in the build container (CPU only, ~10-20 minutes with 7 workers):

    python tests/golden/make_batch_device_order.py
"""

from __future__ import annotations

import hashlib
import os
import sys
import time
from pathlib import Path

for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(v, "1")

import numpy as np  # noqa: E402

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE))

from make_batch_outcomes import MAX_ITER, assemble, system_digest  # noqa: E402
from oracle import krylov as okry  # noqa: E402
from oracle import ldu as oldu  # noqa: E402
from oracle import precond as opre  # noqa: E402

TILE_SLICES = 32


def solve_one(args):
    k, c = args
    d, lo, up = assemble(c)
    rD = opre.dic_build(d, up, c.owner, c.neighbour)
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    x, it, res, info = okry.pcg(mv, lambda r: r / rD, c.b, np.zeros(c.n_cells), tolerance=1e-300,
                                rel_tol=1e-8, max_iter=MAX_ITER,
                                dot=okry.dot_device(c.n_cells, tile_slices=TILE_SLICES))
    xd = hashlib.sha1(np.ascontiguousarray(x, dtype=np.float64).tobytes()).digest()
    return (k, system_digest(d, lo, up, c.b), len(info.residual_history), bool(info.converged),
            float(res), float(info.initial_residual), xd)


def main():
    from concurrent.futures import ProcessPoolExecutor

    from pyopenfoam_amd.cases import poisson_batch
    from tests.helpers import oracle_geometry

    t0 = time.time()
    bc = poisson_batch(512, 8, 48, seed=0, jitter=0.2, sigma=1.0, workers=7,
                       geometry=lambda mesh, progress=None: oracle_geometry(mesh),
                       keep_systems=True)
    cases = bc.systems
    ns = bc.n_systems
    its = np.zeros(ns, np.int32)
    conv = np.zeros(ns, bool)
    fres = np.zeros(ns)
    ires = np.zeros(ns)
    dig = np.zeros((ns, 20), np.uint8)
    xdig = np.zeros((ns, 20), np.uint8)
    order = sorted(range(ns), key=lambda k: -cases[k].n_cells)
    done = 0
    with ProcessPoolExecutor(max_workers=7) as ex:
        for k, dg, it, cv, r, r0, xd in ex.map(solve_one, [(k, cases[k]) for k in order],
                                               chunksize=1):
            its[k], conv[k], fres[k], ires[k] = it, cv, r, r0
            dig[k] = np.frombuffer(dg, np.uint8)
            xdig[k] = np.frombuffer(xd, np.uint8)
            done += 1
            if done % 32 == 0:
                print(f"{done}/{ns} systems, {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(HERE / "batch512_device_order.npz", sizes=bc.sizes,
                        tile_slices=np.int32(TILE_SLICES), iterations=its, converged=conv,
                        final_residual=fres, initial_residual=ires, digest=dig, x_digest=xdig,
                        max_iter=np.int32(MAX_ITER))
    print(f"wrote batch512_device_order.npz in {time.time() - t0:.0f} s; not converged: "
          f"{np.nonzero(~conv)[0].tolist()}", flush=True)


if __name__ == "__main__":
    main()
