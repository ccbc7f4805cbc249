"""Real Nek meshes through the drop-in: the reference's own .re2 files
(ethier: 32 curved elements with 'm' midside edges; lpm_one: a 5^3 box)
read by the library, connected by corner coincidence, given GLL
coordinates, geometry on the GPU, then solved by hmholtz (Jacobi PCG) and by
the PnPn pressure path (flexible CG + h1mg), against the oracle on the same
arrays; and synthetic curved cases ('C' arcs, 's' sphere faces) whose
vertex ids come from a .ma2 map file.  Every boundary face is Dirichlet.
Tolerances as in the box tests:
iteration counts within +-1, solutions within 1e-10 (Jacobi) and 1e-9
(multigrid) relative L2.
"""
import os
import sys

import numpy as np
import pytest

from nek5000_amd.mesh import Re2Mesh
from oracle import oracle as O
from problems import Problem, mg_for

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "re2")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host(t):
    return t.detach().cpu().numpy()


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name,lx1", [("ethier.re2", 8), ("lpm_one_uniform.re2", 6)])
def test_jacobi_hmholtz_on_case_mesh(gpu, name, lx1):
    m = Re2Mesh(os.path.join(GOLD, name), lx1)
    say(f"[re2 hmholtz] {name} lx1={lx1} nel={m.nel}")
    P = Problem(m)
    F = gpu.box_field(m)
    assert np.array_equal(host(F.mask), P.mask)
    np.testing.assert_allclose(host(F.mult), P.mult, rtol=0, atol=0)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 2.0)
    b = P.bm1 * P.uniform_field(5)
    prm_o = O.Params(20, 0.0, 0.0, 1e-10)
    uo, _, ito, _, _ = P.field.hmholtz(b, h1, h2, 1e-10, 900, prm_o)
    u = F.zeros()
    out = F.hmholtz("VELX", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult, 1e-10,
                    900, params=gpu.default_params(istep=20, param22=0.0, restol=1e-10))
    say(f"[re2 hmholtz] iterations {out['niter']} vs {ito}")
    assert abs(out["niter"] - ito) <= 1
    assert l2rel(host(u), uo) < 1e-10
    F.destroy()


@pytest.mark.parametrize("name,lx1", [("ethier.re2", 8), ("lpm_one_uniform.re2", 6)])
def test_pressure_flex_cg_on_case_mesh(gpu, name, lx1):
    m = Re2Mesh(os.path.join(GOLD, name), lx1)
    say(f"[re2 flex_cg] {name} lx1={lx1}")
    P = Problem(m)
    mg = mg_for(P)
    F = gpu.box_field(m, multigrid=True)
    info = F.h1mg_info()
    assert info["nh"] == mg.nh
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    b = P.bm1 * P.uniform_field(8)
    rhs = P.field.dssum(b) * P.mask
    uo, ito, _, _ = mg.flex_cg(rhs, h1, h2, 200, -1e-10)
    u = F.zeros()
    out = F.hmholtz("PRES", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult, -1e-10,
                    200, params=gpu.default_params(param42=2, param21=-1e-10))
    say(f"[re2 flex_cg] iterations {out['niter']} vs {ito}")
    assert abs(out["niter"] - ito) <= 1
    assert l2rel(host(u), uo) < 1e-9
    F.destroy()


@pytest.mark.parametrize("kind,lx1", [("annulus", 8), ("shell", 7)])
def test_curved_case_with_map_file(gpu, tmp_path, kind, lx1):
    """'C' arcs (annulus sector) and 's' sphere faces (shell patch) written
    as .re2 with a genmap-layout .ma2 beside it: coordinates by the
    library's genxyz, vertex ids from the map, then hmholtz (Jacobi PCG) and
    the pressure path (flexible CG + h1mg) against the oracle."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import case_files as CF

    c = (CF.annulus_sector(3, 4, 2) if kind == "annulus" else CF.sphere_shell(3, 2))
    fn = str(tmp_path / f"{kind}.re2")
    CF.write_case(fn, c)
    ids = Re2Mesh(fn, lx1).vertex.reshape(c["nel"], 8)
    CF.write_ma2(str(tmp_path / f"{kind}.ma2"), np.zeros(c["nel"], np.int32), ids + 100)
    m = Re2Mesh(fn, lx1, map_path="auto")
    assert m.map_path.endswith(".ma2")
    say(f"[curved] {kind} lx1={lx1} nel={m.nel}")
    P = Problem(m)
    F = gpu.box_field(m)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 2.0)
    b = P.bm1 * P.uniform_field(11)
    uo, _, ito, _, _ = P.field.hmholtz(b, h1, h2, 1e-10, 900, O.Params(20, 0.0, 0.0, 1e-10))
    u = F.zeros()
    out = F.hmholtz("VELX", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult, 1e-10,
                    900, params=gpu.default_params(istep=20, param22=0.0, restol=1e-10))
    say(f"[curved] jacobi iterations {out['niter']} vs {ito}")
    assert abs(out["niter"] - ito) <= 1
    assert l2rel(host(u), uo) < 1e-10
    F.destroy()
    mg = mg_for(P)
    F = gpu.box_field(m, multigrid=True)
    h2 = np.zeros(P.shape)
    rhs = P.field.dssum(b) * P.mask
    uo, ito, _, _ = mg.flex_cg(rhs, h1, h2, 200, -1e-10)
    u = F.zeros()
    out = F.hmholtz("PRES", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult, -1e-10,
                    200, params=gpu.default_params(param42=2, param21=-1e-10))
    say(f"[curved] flex_cg iterations {out['niter']} vs {ito}")
    assert abs(out["niter"] - ito) <= 1
    assert l2rel(host(u), uo) < 1e-9
    F.destroy()
