"""CPU oracle of the pressure-equation hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
reference results are compared against (or the timed CPU baseline); the
product package ``pyopenfoam_amd`` never imports it and has no CPU fallback
through it.

Each function restates one reference function (paths relative to
``src/pyfoam/`` of pyOpenFOAM) in float64 with the reference's operation order:
numpy where the reference uses vectorised torch ops, plain C (``c/oracle.c``,
built by ``make -C oracle``) where the reference runs per-element Python loops
that are too slow at the benchmark sizes, and pure-Python loops for small
cases.  How each function is pinned (golden vectors written from the
reference, known answers of the reference's own tests, closed forms, brute
force) is recorded in the test that checks it; reductions (dot products,
norms) follow numpy/BLAS summation order, which the reference leaves to
torch, so solver histories are compared to a tolerance and iteration counts
to within one.
"""
