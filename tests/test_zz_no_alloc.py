"""nkd_cggo_noalloc (the FP64 Jacobi PCG with its buffers in the library's
static device array, for callers that cannot allocate) against nkd_cggo,
through the C ABI: two calls in a row, the second on a smaller mesh, each
bit-identical to the ordinary entry point's solve (same kernels, same
partial-sum order) in solution, iteration count, norms and history; a mesh
above the array's limit is refused with an error, not run.
"""
import sys

import numpy as np
import pytest

from problems import Problem, box

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host(t):
    return t.detach().cpu().numpy()


def _solve_pair(gpu, b, seed, h1v, h2v, tin):
    P = Problem(b)
    F = gpu.box_field(b)
    # cggo's input: an assembled, masked rhs (what hmholtz hands it)
    f = P.field.dssum(P.bm1 * P.uniform_field(seed)) * P.mask
    h1, h2 = F.zeros() + h1v, F.zeros() + h2v
    outs = {}
    for noalloc in (False, True):
        say(f"[noalloc] nel={b.nel} lx1={b.lx1} noalloc={noalloc}")
        x = F.zeros()
        out = F.cggo(x, F.to_dev(f), h1, h2, F.mask, F.mult, tin, 900, history=True,
                     noalloc=noalloc)
        outs[noalloc] = (host(x), out)
    F.destroy()
    return outs


def test_noalloc_equals_cggo_twice(gpu):
    # first call: 6x5x4 elements at lx1 = 8 (61,440 local nodes), deformed,
    # h2 = 10; second: a smaller 3x3x2 mesh at lx1 = 6, jittered
    for b, seed, h1v, h2v, tin in ((box((6, 5, 4), 8, deform=0.1), 3, 1.0, 10.0, -1e-8),
                                   (box((3, 3, 2), 6, jitter=0.05, seed=4), 5, 0.5, 0.0,
                                    -1e-10)):
        outs = _solve_pair(gpu, b, seed, h1v, h2v, tin)
        (x0, o0), (x1, o1) = outs[False], outs[True]
        say(f"[noalloc] iterations {o0['niter']} / {o1['niter']}, rbn2 {o0['rbn2']:.6e} / "
            f"{o1['rbn2']:.6e}")
        assert o0["converged"] and o1["converged"]
        assert o1["niter"] == o0["niter"]
        assert o1["rbn0"] == o0["rbn0"] and o1["rbn2"] == o0["rbn2"] and o1["tol"] == o0["tol"]
        assert np.array_equal(o1["history"], o0["history"])
        assert np.array_equal(x1, x0)


def test_noalloc_refuses_large_mesh(gpu):
    # 1,029 elements at lx1 = 8 = 526,848 local nodes > 2^19: refused on the
    # host before any launch
    b = box((21, 7, 7), 8)
    F = gpu.box_field(b)
    f = F.zeros()
    with pytest.raises(gpu.NekAmdError, match="static scratch"):
        F.cggo(F.zeros(), f, F.zeros() + 1.0, F.zeros(), F.mask, F.mult, -1e-8, 10,
               noalloc=True)
    F.destroy()
