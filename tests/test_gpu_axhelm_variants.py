"""The axhelm kernels (element staged in LDS / one t-plane at a time, each
with the default or the wide packing of thread slabs per block; the r/s
contractions on the FP64 matrix cores) and the per-order autotuner that
picks between them, through the C ABI:

  * the plain apply of variants 0-3 is bit-identical and within 1e-13 of
    the oracle, for every coefficient mode, at orders 3..16; the matrix-core
    variants 4 and 5 bit-identical to each other and within 1e-13 of the
    oracle and of variant 0;
  * the Jacobi PCG (the rho-fused apply) and the multigrid flexible CG (the
    p-update-fused apply) give bit-identical iterates with variants 0 and 1,
    and with the wide packings and the matrix cores the same iteration
    counts and iterates up to rounding;
  * the first apply of an order on a mesh of >= 2^20 nodes tunes it, after
    which the choice is reported and the results do not depend on it.
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


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("lx1", [3, 5, 7, 8, 10, 12, 16])
def test_plain_apply_variants_identical(gpu, lx1):
    b = box((3, 2, 2), lx1, deform=0.1)
    P = Problem(b)
    F = gpu.box_field(b)
    rng = np.random.default_rng(lx1)
    u = rng.standard_normal(P.shape)
    h1f = 1.0 + rng.random(P.shape)
    h2f = rng.random(P.shape)
    one, zero = np.ones(P.shape), np.zeros(P.shape)
    saved = gpu.axhelm_tuning(lx1)
    try:
        for h1, h2, tag in ((one, zero, "h1=1,h2=0"), (one * 0.3, one * 10.0, "uniform h2"),
                            (h1f, h2f, "field h1,h2")):
            ref = P.field.axhelm(u, h1, h2)
            outs = []
            for variant, grid in ((0, 0), (1, 0), (1, 2), (2, 0), (2, 3), (3, 0), (4, 0),
                                  (4, 5), (5, 0), (5, 3)):
                say(f"[axhelm variants] lx1={lx1} {tag} variant={variant} grid={grid}")
                gpu.set_axhelm_tuning(lx1, variant, grid)
                au = F.zeros()
                F.axhelm(au, F.to_dev(u), F.to_dev(h1), F.to_dev(h2))
                outs.append(host(au))
            assert rel(outs[0], ref) < 1e-13, tag
            for o in outs[1:6]:
                assert np.array_equal(o, outs[0]), tag
            # the matrix cores sum each contraction in their own order
            for o in outs[6:]:
                assert rel(o, ref) < 1e-13 and rel(o, outs[0]) < 1e-13, tag
                if lx1 < 8:  # no matrix-core kernel: variant 0 runs
                    assert np.array_equal(o, outs[0]), tag
            for o in outs[7:]:
                assert np.array_equal(o, outs[6]), tag
    finally:
        gpu.set_axhelm_tuning(lx1, saved[0], saved[1])
    F.destroy()


def test_jacobi_pcg_variants_identical(gpu):
    b = box((4, 3, 3), 8, deform=0.1)
    P = Problem(b)
    F = gpu.box_field(b)
    h1 = F.to_dev(1.0 + 0.5 * (P.uniform_field(11) + 1.0))
    h2 = F.zeros() + 2.0
    rhs = F.to_dev(P.bm1 * P.uniform_field(12))
    saved = gpu.axhelm_tuning(8)
    outs = []
    try:
        for variant in (0, 1):
            say(f"[pcg variants] variant={variant}")
            gpu.set_axhelm_tuning(8, variant, 0)
            u = F.zeros()
            out = F.hmholtz("VELX", u, rhs.clone(), h1, h2, F.mask, F.mult, -1e-10, 500,
                            history=True)
            assert out["converged"]
            outs.append((host(u), out))
    finally:
        gpu.set_axhelm_tuning(8, saved[0], saved[1])
    assert outs[0][1]["niter"] == outs[1][1]["niter"]
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(np.asarray(outs[0][1]["history"]), np.asarray(outs[1][1]["history"]))
    F.destroy()


@pytest.mark.parametrize("lx1", [7, 9])
def test_flex_cg_variants_identical(gpu, lx1):
    b = box((3, 3, 3), lx1, deform=0.1)
    P = Problem(b)
    F = gpu.box_field(b)
    x, y, z = b.xyz()
    F.h1mg_setup(b.fast_bc(), b.vertex, x, y, z)
    h1, h2 = F.zeros() + 1.0, F.zeros()
    rhs = F.to_dev(P.field.dssum(P.bm1 * P.uniform_field(3)) * P.mask)
    saved = gpu.axhelm_tuning(lx1)
    outs = []
    try:
        for variant in (0, 1, 2, 3, 4, 5):
            say(f"[flex-cg variants] lx1={lx1} variant={variant}")
            gpu.set_axhelm_tuning(lx1, variant, 0)
            res = rhs.clone()
            prm = gpu.default_params(istep=1, param42=2, param21=-1e-10)
            out = F.hmh_flex_cg(res, h1, h2, F.mult, 100, params=prm, history=True)
            assert out["converged"]
            outs.append((host(res), out))
    finally:
        gpu.set_axhelm_tuning(lx1, saved[0], saved[1])
    assert outs[0][1]["niter"] == outs[1][1]["niter"]
    assert np.array_equal(outs[0][0], outs[1][0])
    # the wide packings (lx1 = 7: nine slabs per block) round the partial
    # sums of p.w in another grouping, the matrix cores (lx1 = 9) the
    # contractions
    for o in outs[2:]:
        assert o[1]["niter"] == outs[0][1]["niter"]
        assert rel(o[0], outs[0][0]) < 1e-10
    assert np.array_equal(outs[5][0], outs[4][0])  # the two matrix-core stagings
    F.destroy()


@pytest.mark.parametrize("lx1", [5, 10, 13, 16])
def test_jacobi_pcg_wide_variants(gpu, lx1):
    """variants 2, 3 and 4 in the CG-fused apply: the same iterates as 0 up
    to the rounding of the partial sums (and of the matrix-core
    contractions)"""
    b = box((3, 3, 2), lx1, deform=0.1)
    P = Problem(b)
    F = gpu.box_field(b)
    h1 = F.to_dev(1.0 + 0.5 * (P.uniform_field(21) + 1.0))
    h2 = F.zeros() + 1.0
    rhs = F.to_dev(P.bm1 * P.uniform_field(22))
    saved = gpu.axhelm_tuning(lx1)
    outs = []
    try:
        for variant in (0, 2, 3, 4, 5):
            say(f"[pcg wide variants] lx1={lx1} variant={variant}")
            gpu.set_axhelm_tuning(lx1, variant, 0)
            u = F.zeros()
            out = F.hmholtz("VELX", u, rhs.clone(), h1, h2, F.mask, F.mult, -1e-10, 500,
                            history=True)
            assert out["converged"]
            outs.append((host(u), out))
    finally:
        gpu.set_axhelm_tuning(lx1, saved[0], saved[1])
    for o in outs[1:]:
        assert o[1]["niter"] == outs[0][1]["niter"]
        assert rel(o[0], outs[0][0]) < 1e-10
        # histories: rounding differences grow as the iteration proceeds (CG
        # scalars near convergence differ by 1e-4 here); the first 20
        # iterations agree to 1e-8
        h0, h = np.asarray(outs[0][1]["history"])[:20], np.asarray(o[1]["history"])[:20]
        big = np.abs(h0) > 0
        assert np.all(np.abs(h - h0)[big] <= 1e-8 * np.abs(h0)[big])
    F.destroy()


def test_autotuner_picks_and_keeps_results(gpu):
    # 12 x 12 x 10 elements at lx1 = 9: 1,049,760 nodes, above the 2^20 threshold
    lx1 = 9
    b = box((12, 12, 10), lx1, deform=0.05)
    F = gpu.box_field(b)
    rng = np.random.default_rng(1)
    u = F.to_dev(rng.standard_normal(F.shape))
    h1, h2 = F.zeros() + 1.0, F.zeros() + 0.5
    before = gpu.axhelm_tuning(lx1)
    say(f"[autotune] lx1={lx1} nodes={F.shape} before={before}")
    au = F.zeros()
    F.axhelm(au, u, h1, h2)
    variant, grid, tuned = gpu.axhelm_tuning(lx1)
    v_cg, v_flex = gpu.axhelm_tuning_fused(lx1)
    say(f"[autotune] chosen variant={variant} grid={grid} tuned={tuned}, fused {v_cg} {v_flex}")
    assert tuned and variant in range(6) and grid >= 0 and v_cg in range(6)
    assert v_flex in range(6)
    a_tuned = host(au)
    for v in range(6):
        gpu.set_axhelm_tuning(lx1, v, 0)
        au2 = F.zeros()
        F.axhelm(au2, u, h1, h2)
        if (v >= 4) == (variant >= 4):
            assert np.array_equal(host(au2), a_tuned)
        else:  # matrix cores against vector FMA: equal to rounding
            assert rel(host(au2), a_tuned) < 1e-13
    gpu.set_axhelm_tuning(lx1, variant, grid)
    assert gpu.axhelm_tuning_fused(lx1) == (variant, variant)
    F.destroy()


def test_tuning_arguments_checked(gpu):
    with pytest.raises(gpu.NekAmdError):
        gpu.set_axhelm_tuning(17, 0, 0)
    with pytest.raises(gpu.NekAmdError):
        gpu.set_axhelm_tuning(8, 6, 0)
    with pytest.raises(gpu.NekAmdError):
        gpu.set_axhelm_tuning(8, 0, -1)
    with pytest.raises(gpu.NekAmdError):
        gpu.axhelm_tuning(1)


@pytest.mark.parametrize("variant", [4, 5])
def test_gated_apply_matrix_core_variants(gpu, variant):
    """The gated apply (the mixed Jacobi PCG's FP64 residual refreshes) on the
    matrix cores: the mixed solve's true FP64 residual, recomputed with the
    oracle's operator, meets the tolerance, in about the FP64 solve's
    iteration count."""
    lx1 = 10
    b = box((3, 3, 2), lx1, deform=0.1)
    P = Problem(b)
    F = gpu.box_field(b)
    h1v, h2v, tli = 0.7, 20.0, -1e-9
    rhs = P.bm1 * P.uniform_field(31)
    h1, h2 = F.zeros() + h1v, F.zeros() + h2v
    saved = gpu.axhelm_tuning(lx1)
    outs = {}
    try:
        for tag, v, prec in (("fp64", 0, 0), ("mixed", variant, 2)):
            gpu.set_axhelm_tuning(lx1, v, 0)
            say(f"[gated variants] lx1={lx1} {tag} variant={v}")
            u = F.zeros()
            r = F.to_dev(rhs)
            prm = gpu.default_params(istep=20, param22=0.0, restol=0.0, precision=prec)
            out = F.hmholtz("VELX", u, r, h1, h2, F.mask, F.mult, tli, 900, params=prm)
            outs[tag] = (host(u), out, host(r))
    finally:
        gpu.set_axhelm_tuning(lx1, saved[0], saved[1])
    u64, o64, f = outs["fp64"]
    u32, o32, _ = outs["mixed"]
    assert o64["converged"] and o32["converged"]
    assert o32["niter"] <= int(1.1 * o64["niter"]) + 3
    au = P.field.axhelm(u32, np.full(P.shape, h1v), np.full(P.shape, h2v))
    res = f - P.mask * P.field.dssum(au)
    rt = float(np.sqrt(np.sum(P.field.binv * P.mult * res * res) / P.field.vol))
    assert rt <= o32["tol"] * (1 + 1e-6)
    F.destroy()
