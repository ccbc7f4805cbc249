"""Batched hmholtz solves (nkh_batch_create / nkd_hmholtz_batch): solves on
different meshes, orders, coefficients and masks -- one mesh used twice with
two masks -- iterate together with the steady-state iteration of the whole
batch captured in one hipGraph and replayed.  Every solve's iteration count
and solution equal, bit for bit, those of nkd_hmholtz on the same solve
(the same kernels in the same order), and agree with the oracle."""
import sys

import numpy as np
import pytest

from problems import Problem, box, mask_from_fbc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host(t):
    return t.detach().cpu().numpy()


def test_batch_equals_single_solves(gpu):
    bA = box((4, 3, 3), 6, deform=0.1)
    bB = box((3, 3, 2), 8, jitter=0.05, seed=3)
    bC = box(3, 5, periodic_x=True)
    PA, PB, PC = Problem(bA), Problem(bB), Problem(bC)
    FA, FB, FC = gpu.box_field(bA), gpu.box_field(bB), gpu.box_field(bC)
    # a second mask on mesh A: Neumann on its top faces
    fbc = bA.fast_bc()
    top = (np.arange(bA.nel) // (bA.nelx * bA.nely)) == bA.nelz - 1
    fbc[top, 5] = 2
    m2 = FA.to_dev(mask_from_fbc(fbc, bA.lx1))
    FA.dsop(m2, "*")
    say("[batch] meshes A (4,3,3) lx1=6, B (3,3,2) lx1=8, C 3^3 lx1=5 periodic; A twice")
    cases = [
        (FA, PA, FA.mask, 1.0, 0.0, 11, -1e-10),
        (FA, PA, m2, 1.0, 5.0, 12, -1e-9),
        (FB, PB, FB.mask, None, 2.0, 13, -1e-10),
        (FC, PC, FC.mask, 2.0, 1.0, 14, 1e-8),
    ]
    prm = gpu.default_params(istep=20, param22=0.0, restol=0.0)
    solves, ref = [], []
    for F, P, mask, h1v, h2v, seed, tli in cases:
        h1 = F.to_dev(1.0 + 0.5 * (P.uniform_field(seed + 100) + 1.0)) if h1v is None \
            else F.zeros() + h1v
        h2 = F.zeros() + h2v
        b = F.to_dev(P.bm1 * P.uniform_field(seed))
        u = F.zeros()
        out = F.hmholtz("VELX", u, b.clone(), h1, h2, mask, F.mult, tli, 900, params=prm)
        ref.append((host(u), out))
        solves.append(dict(u=F.zeros(), rhs=b.clone(), h1=h1, h2=h2, mask=mask, mult=F.mult,
                           tli=tli))
    B = gpu.Batch([c[0] for c in cases])
    for rep in range(2):  # the batch again (a new graph) gives the same
        for q in solves:
            q["u"].zero_()
        outs, launches = B.hmholtz(
            [dict(q, rhs=q["rhs"].clone()) for q in solves], 900, params=prm)
        say(f"[batch] rep {rep}: graph launches {launches}, iterations "
            f"{[o['niter'] for o in outs]} vs single {[r[1]['niter'] for r in ref]}")
        for k, (q, o) in enumerate(zip(solves, outs)):
            assert o["niter"] == ref[k][1]["niter"], k
            assert o["converged"] == ref[k][1]["converged"]
            assert o["rbn2"] == ref[k][1]["rbn2"]
            assert np.array_equal(host(q["u"]), ref[k][0]), k
        assert launches >= max(r[1]["niter"] for r in ref)
    B.destroy()
    for F in (FA, FB, FC):
        F.destroy()


def test_batch_against_oracle(gpu):
    b = box((4, 4, 2), 7, deform=0.1)
    P = Problem(b)
    F = gpu.box_field(b)
    from oracle import oracle as O

    prm_o = O.Params(20, 0.0, 0.0, 0.0)
    solves, exp = [], []
    for k in range(3):  # three "velocity components": one mesh, one operator
        rhs = P.bm1 * P.uniform_field(30 + k)
        uo, _, ito, _, _ = P.field.hmholtz(rhs, np.ones(P.shape), np.full(P.shape, 3.0), -1e-10,
                                           900, prm_o)
        exp.append((uo, ito))
        solves.append(dict(u=F.zeros(), rhs=F.to_dev(rhs), h1=F.zeros() + 1.0,
                           h2=F.zeros() + 3.0, mask=F.mask, mult=F.mult, tli=-1e-10))
    B = gpu.Batch([F, F, F])
    outs, _ = B.hmholtz(solves, 900, params=gpu.default_params(istep=20, param22=0.0,
                                                                   restol=0.0))
    for q, o, (uo, ito) in zip(solves, outs, exp):
        assert abs(o["niter"] - ito) <= 1
        e = float(np.linalg.norm(host(q["u"]) - uo) / np.linalg.norm(uo))
        assert e < 1e-9, e
    B.destroy()
    F.destroy()


def test_batch_errors(gpu):
    F = gpu.box_field(box(2, 4))
    with pytest.raises(gpu._lib.NekAmdError):
        gpu.Batch([])
    B = gpu.Batch([F])
    with pytest.raises(gpu._lib.NekAmdError, match="precision"):
        B.hmholtz([dict(u=F.zeros(), rhs=F.zeros(), h1=F.zeros() + 1, h2=F.zeros(),
                        mask=F.mask, mult=F.mult, tli=1e-8)], 10,
                  params=gpu.default_params(precision=1))
    B.destroy()
    F.destroy()
