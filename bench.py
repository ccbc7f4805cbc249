#!/usr/bin/env python3
"""Throughput of the Helmholtz solve path on one MI355X.

Workload (the metric's configuration): E = 64^3 = 262,144 hexahedra, N = 7
(lx1 = 8), every GLL node mapped by x_i <- x_i + 0.1 h sin(2 pi x) sin(2 pi y)
sin(2 pi z), h1 = 1, h2 = 10, f ~ U(-1,1) per unique node (seeded), Dirichlet
on all faces, solved in FP64 to 1e-8 relative residual by hmholtz('PRES')
-> flexible CG with the h1mg V-cycle (orders 7 -> 3 -> 1 + coarse solve;
param(42) = 2), or with --solver jacobi by hmholtz -> Jacobi-PCG cggo.  One
step = one full solve with all inputs resident in HBM (set-up excluded).

value = (unique free DOF x CG iterations) / wall time, in GDOF*iter/s.
roofline: the kernel of the iteration that takes longest (from the library's
HIP-event timers on the stream it runs on, over the timed solves): the
fine-level local FDM solves, the flexible-CG axhelm or the assembly, with
its algorithmic bytes per launch; "kernels" lists all three.
cpu_baseline: the oracle (C, OpenMP over all host threads) solving the same
operator with the same solver on a bounded sample (E = 16^3).
targets: the mixed-precision solve's time to solution against FP64's, the
five-solve batch (mixed, batched) against five FP64 solves one after
another, and the adjoint gradient's finite-difference check.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PCG GDOF·iter/s at E=262144,N=7: FP64 vs mixed @1e-8 rel; %HBM roofline"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_rhs(box, glo, ngv, bm1_host=None, seed=2024):
    from nek5000_amd.mesh import hash_uniform

    uid = box.unique_ids(glo, ngv).reshape(-1)
    f = 2.0 * hash_uniform(seed, uid) - 1.0
    return f.reshape(box.nel, box.lx1, box.lx1, box.lx1)


def fused_axhelm_bytes_per_node(lx1, h1_field=False, h2_kind="uniform"):
    """Algorithmic HBM bytes per local node of the CG-fused axhelm kernel:
    reads r, d, p (3 words), g1..g6 (6), bm1 when h2 != 0 (1), h1 and h2
    when they are fields; writes p and w (2)."""
    words = 3 + 6 + 2
    if h2_kind == "uniform":
        words += 1
    elif h2_kind == "field":
        words += 2
    if h1_field:
        words += 1
    return 8.0 * words


def fdm_bytes_per_element(nh):
    """Algorithmic HBM bytes per element of the fine-level local FDM solves
    (k_sch_fdm): reads r (nh^3), the neighbours' first interior layers
    (6 nh^2, gathered), the element's three eigenvector matrices (3 (nh+2)^2)
    and eigenvalues (3 (nh+2) + 1), writes the local solution's interior
    (nh^3) and its ghost faces (6 nh^2), all 8-byte words; plus the two
    int32 ghost maps (gin, gout: 6 nh^2 each)."""
    nl = nh + 2
    words = 2 * nh ** 3 + 12 * nh ** 2 + 3 * nl ** 2 + 3 * nl + 1
    return 8.0 * words + 4.0 * 12 * nh ** 2


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measured_traffic(name):
    """HBM bytes per launch of the roofline kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload (scaled by the
    calibration factors stated there), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as fh:
            for line in fh:
                if "traffic_bytes_per_launch:" in line:
                    return float(line.split(":")[1])
    except OSError:
        return None
    return None


def cpu_baseline(nel1d, lx1, deform, target_seconds, solver="hsmg"):
    """The oracle's solve of the same operator on E=nel1d^3, bounded by maxit
    so that the timed part takes about target_seconds: hmholtz -> cggo
    (Jacobi) or the flexible CG with the h1mg V-cycle (set-up not timed)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import oracle as O
    from problems import Problem, box as mkbox, mg_for

    P = Problem(mkbox(nel1d, lx1, deform=deform))
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 10.0)
    b = P.bm1 * P.uniform_field(2024)
    free = P.free_dof()
    if solver == "hsmg":
        mg = mg_for(P)
        rhs = P.field.dssum(b) * P.mask
        t = time.perf_counter()
        _, n_cal, _, _ = mg.flex_cg(rhs, h1, h2, 2, -1e-30)
        dt = (time.perf_counter() - t) / max(n_cal, 1)
        maxit = int(max(3, min(200, target_seconds / max(dt, 1e-6))))
        t = time.perf_counter()
        _, niter, _, _ = mg.flex_cg(rhs, h1, h2, maxit, -1e-8)
        el = time.perf_counter() - t
        what = f"flexible CG + h1mg V-cycle, {niter} iterations"
    else:
        t = time.perf_counter()
        _, _, n_cal, _, _ = P.field.hmholtz(b, h1, h2, -1e-30, 5)
        dt = (time.perf_counter() - t) / max(n_cal, 1)
        maxit = int(max(10, min(900, target_seconds / max(dt, 1e-6))))
        t = time.perf_counter()
        _, _, niter, _, _ = P.field.hmholtz(b, h1, h2, -1e-8, maxit)
        el = time.perf_counter() - t
        what = f"hmholtz/cggo Jacobi-PCG, {niter} iterations"
    return dict(value=free * niter / el / 1e9, unit="GDOF*iter/s", cores=O.num_threads(),
                kind="port", cpu=cpu_model(),
                sample=(f"oracle (C restatement, OpenMP) on E={nel1d}^3 N={lx1-1} deformed, "
                        f"h1=1 h2=10: {what} ({free:.0f} free DOF) in {el:.1f} s"))


def order_sweep(orders=range(3, 16)):
    """Polynomial-order sweep (jittered boxes, E^(1/3) ~ 368/N so every order
    carries about 5e7 unique DOF; h1 = h2 = 1; f ~ U(-1,1)): per order the
    autotuner's choice (made on the first apply), the CG-fused axhelm
    kernel's time per kernel variant (0/1 element/plane staged, 2/3 the
    same with the wide slab packing, 4/5 the r/s contractions on the FP64
    matrix cores with the element / groups of t-planes staged, at N = 7..15;
    HIP events of the library's timers
    over 20 Jacobi-PCG iterations each), and 100 fixed Jacobi-PCG iterations
    of hmholtz with the FP64 operator (its axhelm kernel time and HBM
    fraction at 12 words per node) and with the FP32 one (mixed precision),
    the latter also with each kernel of the FP32 operator (vector FMA,
    matrix cores).
    Prints one JSON line with the table."""
    import ctypes

    import torch
    from nek5000_amd import _lib, solver
    from nek5000_amd.mesh import BoxMesh

    solver.init(0)
    lib = _lib.load()

    def timed_pcg(F, u, rhs, one, iters, params=None):
        """wall seconds, and the axhelm kernel's mean ms from the timers"""
        lib.nkh_timers_reset()
        lib.nkh_timers_enable(1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = F.hmholtz("TEMP", u, rhs.clone(), one, one, F.mask, F.mult, -1e-30, iters,
                        params=params)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        lib.nkh_timers_enable(0)
        sec, calls = ctypes.c_double(), ctypes.c_int64()
        lib.nkh_timers_read(b"axhm", ctypes.byref(sec), ctypes.byref(calls))
        return wall, sec.value / max(calls.value, 1) * 1e3, out
    rows = []
    for N in orders:
        e1 = {3: 110, 15: 24}.get(N, int(round(368.0 / N)))
        t0 = time.time()
        box = BoxMesh(e1, e1, e1, N + 1, jitter=0.05, seed=N)
        F = solver.box_field(box)
        f = make_rhs(box, F.glo_num, F.ngv, seed=N)
        bm1, _ = F.mass()
        rhs = bm1 * torch.as_tensor(f, device="cuda")
        del f
        one = torch.ones_like(rhs)
        free = float(torch.sum(F.mask * F.mult))
        setup = time.time() - t0
        au = torch.zeros_like(rhs)
        F.axhelm(au, rhs, one, one)  # the first apply of the order tunes it
        torch.cuda.synchronize()
        chosen = solver.axhelm_tuning(N + 1)
        chosen_fused = solver.axhelm_tuning_fused(N + 1)

        u = torch.zeros_like(rhs)
        per_variant = {}
        for v in (0, 1, 2, 3, 4, 5):
            solver.set_axhelm_tuning(N + 1, v, 0)
            timed_pcg(F, u, rhs, one, 5)
            per_variant[f"variant{v}_cg_ms"] = timed_pcg(F, u, rhs, one, 20)[1]
        solver.set_axhelm_tuning(N + 1, chosen_fused[0], chosen[1])
        timed_pcg(F, u, rhs, one, 100)
        pcg_s, ax_ms, out = timed_pcg(F, u, rhs, one, 100)
        # CG-fused axhelm (h2 uniform nonzero): r, d, p, g1..g6, bm1 in;
        # p, w out -- 12 words per local node
        ax_gbs = fused_axhelm_bytes_per_node(N + 1) * box.n / (ax_ms * 1e-3) / 1e9
        # the same 100 iterations with the FP32 operator (mixed Jacobi PCG)
        p32 = solver.default_params(precision=2)
        timed_pcg(F, u, rhs, one, 100, p32)
        pcg32_s, _, out32 = timed_pcg(F, u, rhs, one, 100, p32)
        # and with each kernel of the FP32 operator (vector FMA / matrix cores)
        v32 = solver.axhelm32_variant(N + 1)
        per32 = {}
        for v in ((0, 1) if N + 1 >= 8 else (0,)):
            solver.set_axhelm32_variant(N + 1, v)
            timed_pcg(F, u, rhs, one, 20, p32)
            t32, _, o32 = timed_pcg(F, u, rhs, one, 100, p32)
            per32[f"pcg32_variant{v}_GDOF_iter_s"] = free * o32["niter"] / t32 / 1e9
        solver.set_axhelm32_variant(N + 1, v32)
        row = dict(N=N, E=box.nel, local_nodes=box.n, free_dof=free, setup_s=round(setup, 1),
                   axhelm_variant=chosen[0], axhelm_grid=chosen[1],
                   axhelm_cg_variant=chosen_fused[0], **per_variant,
                   axhelm_cg_ms=ax_ms, axhelm_cg_GBs=ax_gbs,
                   axhelm_cg_hbm_frac=ax_gbs / HBM_PEAK_GBS,
                   pcg_iterations=out["niter"], pcg_ms=pcg_s * 1e3,
                   pcg_GDOF_iter_s=free * out["niter"] / pcg_s / 1e9,
                   pcg32_ms=pcg32_s * 1e3,
                   pcg32_GDOF_iter_s=free * out32["niter"] / pcg32_s / 1e9,
                   axhelm32_variant=v32, **per32)
        log(f"[sweep] {row}")
        rows.append(row)
        F.destroy()
        del F, rhs, one, au, u, bm1
        torch.cuda.empty_cache()
    print(json.dumps({"sweep": "order", "precision": "fp64", "rows": rows}), flush=True)


def adjoint_run(nel1d=32, lx1=8):
    """Adjoint coefficient gradient at E = 32^3, N = 7: -div(h1 grad u) = 1
    (h2 = 0, Dirichlet), h1 = exp(g), g = sum of 16 seeded Fourier modes of
    amplitude 0.5 with wavenumbers 1..4; J = 1/2 (u - u*)^T B (u - u*) with
    u* the solution for a second seeded coefficient field.  Forward and
    adjoint by the flexible CG with the h1mg V-cycle to 1e-12 relative (the
    same solver for both: the operator is self-adjoint), then the gradient
    kernel; central differences (step 1e-6) along 20 seeded random
    directions check it.  Prints one JSON line."""
    import torch
    from nek5000_amd import solver
    from nek5000_amd.mesh import BoxMesh

    solver.init(0)
    box = BoxMesh(nel1d, nel1d, nel1d, lx1, deform=0.1)
    F = solver.box_field(box, multigrid=True)
    x, y, z = box.xyz()

    def fourier(seed, nmodes):
        rng = np.random.default_rng(seed)
        s = np.zeros_like(x)
        for _ in range(nmodes):
            k = rng.integers(1, 5, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            s += 0.5 * np.cos(np.pi * k[0] * x + ph[0]) * np.cos(np.pi * k[1] * y + ph[1]) * \
                np.cos(np.pi * k[2] * z + ph[2])
        return s

    h1, h1star = F.to_dev(np.exp(fourier(5, 16))), F.to_dev(np.exp(fourier(6, 16)))
    h2 = F.zeros()
    bm1, _ = F.mass()
    f = bm1.clone()
    prm = solver.default_params(istep=20, param42=2, param21=-1e-12)

    def hsolve(rhs, h):
        u = F.zeros()
        out = F.hmholtz("PRES", u, rhs.clone(), h, h2, F.mask, F.mult, -1e-12, 2000, params=prm)
        assert out["converged"]
        return u, out["niter"]

    ustar, _ = hsolve(f, h1star)

    def J(h):
        u, _ = hsolve(f, h)
        d = u - ustar
        return 0.5 * float(torch.sum(bm1 * d * d))

    hsolve(f, h1)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u, it_f = hsolve(f, h1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lam, it_a = hsolve(bm1 * (u - ustar), h1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g = F.zeros()
    F.adjoint_coef_grad(g, u, lam)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for _ in range(5):
        F.adjoint_coef_grad(g, u, lam)
    torch.cuda.synchronize()
    kern_ms = (time.perf_counter() - t3) / 5 * 1e3
    checks = []
    t4 = time.perf_counter()
    for seed in range(11, 31):
        # seeded smooth relative perturbations (4 random Fourier modes): a
        # pointwise-random sign would cancel dJ down to the round-off of J
        dh = F.to_dev(fourier(seed, 4)) * h1
        eps = 1e-6
        fd = (J(h1 + eps * dh) - J(h1 - eps * dh)) / (2 * eps)
        ad = float(torch.sum(g * dh))
        checks.append(dict(seed=seed, adjoint=ad, central_difference=fd,
                           rel_diff=abs(fd - ad) / abs(ad)))
    fd_s = time.perf_counter() - t4
    del x, y, z
    line = dict(run="adjoint", workload=f"E={nel1d}^3 N={lx1 - 1} deformed box, h1=exp(Fourier)",
                forward_ms=(t1 - t0) * 1e3, forward_iterations=it_f,
                adjoint_ms=(t2 - t1) * 1e3, adjoint_iterations=it_a,
                gradient_kernel_ms=kern_ms,
                gradient_kernel_GBs=9 * 8 * box.n / (kern_ms * 1e-3) / 1e9,
                fd_directions=len(checks), fd_step=1e-6,
                fd_max_rel_diff=max(c["rel_diff"] for c in checks), fd_seconds=fd_s,
                finite_difference_checks=checks)
    F.destroy()
    return line


def batch_run(steps=2):
    """Five solves of one time step (three velocity components on a 48^3 box,
    h1 = 1e-3; a passive scalar on 24^3, h1 = 1e-3; one on a 40x40x20
    x-periodic box, h1 = 1e-2; every field with the BDF3 mass term of
    dt = 1e-3, h2 = bd(1)/dt = 11/(6e-3), as sethlm (subs1.f:1013-1036) sets
    it for any field of a transient run; N = 7, f ~ U(-1,1) per unique node,
    Jacobi PCG to 1e-8 relative, at most 900 iterations as cggo's maxcg):
    one after another through hmholtz in FP64 (the baseline), and batched
    (nkd_hmholtz_batch: the iteration of all five in one hipGraph) in FP64
    and in mixed precision (FP32 iteration, FP64 refreshes).  Prints one
    JSON line."""
    import torch
    from nek5000_amd import solver
    from nek5000_amd.mesh import BoxMesh

    solver.init(0)
    h2_bdf3 = 11.0 / (6.0 * 1e-3)
    specs = [("vel", BoxMesh(48, 48, 48, 8, deform=0.1), 1e-3, h2_bdf3),
             ("A", BoxMesh(24, 24, 24, 8, deform=0.1), 1e-3, h2_bdf3),
             ("B", BoxMesh(40, 40, 20, 8, periodic_x=True), 1e-2, h2_bdf3)]
    fields, solves, dofs = [], [], []
    for name, box, h1v, h2v in specs:
        F = solver.box_field(box)
        bm1, _ = F.mass()
        free = float(torch.sum(F.mask * F.mult))
        ncomp = 3 if name == "vel" else 1
        for c in range(ncomp):
            f = make_rhs(box, F.glo_num, F.ngv, seed=100 + len(solves))
            rhs = bm1 * torch.as_tensor(f, device="cuda")
            fields.append(F)
            solves.append(dict(u=F.zeros(), rhs=rhs, h1=F.zeros() + h1v, h2=F.zeros() + h2v,
                               mask=F.mask, mult=F.mult, tli=-1e-8))
            dofs.append(free)
    B = solver.Batch(fields)
    prm = {p: solver.default_params(istep=20, param22=0.0, restol=0.0, precision=p)
           for p in (0, 2)}

    def run_batch(precision):
        qs = [dict(q, rhs=q["rhs"].clone()) for q in solves]
        torch.cuda.synchronize()
        t = time.perf_counter()
        outs, launches = B.hmholtz(qs, 900, params=prm[precision])
        torch.cuda.synchronize()
        return time.perf_counter() - t, outs, launches

    def run_seq(precision):
        qs = [dict(q, rhs=q["rhs"].clone()) for q in solves]
        torch.cuda.synchronize()
        t = time.perf_counter()
        outs = [F.hmholtz("VELX", q["u"], q["rhs"], q["h1"], q["h2"], q["mask"], q["mult"],
                          q["tli"], 900, params=prm[precision]) for F, q in zip(fields, qs)]
        torch.cuda.synchronize()
        return time.perf_counter() - t, outs, None

    rows = {}
    sols = {}
    for tag, fn, precision in (("sequential_fp64", run_seq, 0), ("batched_fp64", run_batch, 0),
                               ("batched_mixed", run_batch, 2),
                               ("sequential_mixed", run_seq, 2)):
        fn(precision)  # warm-up (graph capture, autotuning, FP32 factors)
        best = None
        for _ in range(steps):
            r = fn(precision)
            if best is None or r[0] < best[0]:
                best = r
        t, outs, launches = best
        its = [o["niter"] for o in outs]
        work = sum(d * i for d, i in zip(dofs, its))
        rows[tag] = dict(seconds=t, iterations=its, graph_launches=launches,
                         GDOF_iter_s=work / t / 1e9,
                         converged=[bool(o["converged"]) for o in outs],
                         rel_residual=[o["rbn2"] / o["rbn0"] for o in outs])
        sols[tag] = [q["u"].clone() for q in solves]
        log(f"[batch] {tag}: {rows[tag]}")
    base = rows["sequential_fp64"]["seconds"]
    for tag, row in rows.items():
        row["speedup_vs_sequential_fp64"] = base / row["seconds"]
        row["max_rel_diff_vs_fp64"] = max(
            float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
            for a, b in zip(sols[tag], sols["sequential_fp64"]))
    line = dict(run="batch", workload="3 x velocity 48^3 + scalar 24^3 + scalar 40x40x20 "
                                      "periodic, N=7, h2=bd(1)/dt=11/(6e-3) for all, Jacobi PCG "
                                      "to 1e-8 relative (maxcg 900)",
                local_nodes=[int(F.n) for F in fields], free_dof=dofs, variants=rows)
    B.destroy()
    for F in {id(F): F for F in fields}.values():
        F.destroy()
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nel", type=int, default=64, help="elements per direction")
    ap.add_argument("--lx1", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-nel", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--maxit", type=int, default=900,
                    help="iteration cap per solve (900 = the reference's maxcg)")
    ap.add_argument("--crs", type=int, default=0,
                    help="coarse solver of the V-cycle: 0 solved to 1e-13 (the reference's "
                         "exact XXT), 1 one AMG V-cycle (its param(40)=1)")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 1, 2),
                    help="main line's nkh_params.precision (profiling the mixed solvers)")
    ap.add_argument("--fdm-variant", type=int, default=0, choices=(0, 1, 2, 3),
                    help="kernel of the local FDM solves for the main line (nkh_fdm_set_variant:"
                         " 0 vector, 1/2 FP32/bf16 matrix cores for the single-precision "
                         "smoother, 3 FP64 matrix cores)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the extra short run of the other coarse solver")
    ap.add_argument("--batch", action="store_true",
                    help="instead of the bench line: five solves of a time step batched in one "
                         "graph vs one after another")
    ap.add_argument("--adjoint", action="store_true",
                    help="instead of the bench line: adjoint gradient timing and finite-"
                         "difference check at E=32^3, N=7")
    ap.add_argument("--sweep", action="store_true",
                    help="instead of the bench line: polynomial-order sweep N=3..15 of the "
                         "CG-fused axhelm (both kernel variants) and 100 Jacobi-PCG "
                         "iterations with the FP64 and the FP32 operator")
    ap.add_argument("--sweep-orders", default="3-15",
                    help="polynomial orders of --sweep, as first-last")
    ap.add_argument("--solver", choices=("hsmg", "jacobi"), default="hsmg",
                    help="hsmg: hmholtz('PRES') -> hmh_flex_cg with h1mg (param42=2); "
                         "jacobi: hmholtz -> cggo")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (elements "
              "would be partitioned across GPUs with a halo exchange of the "
              "interface nodes).", file=sys.stderr)
        sys.exit(2)
    if args.sweep:
        lo, hi = (int(x) for x in args.sweep_orders.split("-"))
        order_sweep(range(lo, hi + 1))
        return
    if args.adjoint:
        print(json.dumps(adjoint_run()), flush=True)
        return
    if args.batch:
        print(json.dumps(batch_run()), flush=True)
        return

    import torch
    from nek5000_amd import _lib, solver
    from nek5000_amd.mesh import BoxMesh

    solver.init(0)
    lib = _lib.load()
    t0 = time.time()
    box = BoxMesh(args.nel, args.nel, args.nel, args.lx1, deform=0.1)
    log(f"[bench] mesh E={box.nel} lx1={box.lx1} n={box.n}")
    F = solver.box_field(box)
    log(f"[bench] field set up in {time.time() - t0:.1f} s")
    hsmg = args.solver == "hsmg"
    if hsmg:
        t1 = time.time()
        x, y, z = box.xyz()
        F.h1mg_setup(box.fast_bc(), box.vertex, x, y, z)
        del x, y, z
        log(f"[bench] h1mg set up in {time.time() - t1:.1f} s: {F.h1mg_info()}")
    f = make_rhs(box, F.glo_num, F.ngv)
    bm1, _ = F.mass()
    rhs0 = bm1 * torch.as_tensor(f, device="cuda")
    del f
    h1 = torch.ones_like(rhs0)
    h2 = torch.full_like(rhs0, 10.0)
    free_dof = float(torch.sum(F.mask * F.mult))
    log(f"[bench] free unique DOF = {free_dof:.0f}")

    def params_for(crs, precision=0):
        if hsmg:
            return solver.default_params(param42=2, param21=-args.tol, param40=crs,
                                         precision=precision)
        return solver.default_params(precision=precision)

    prm = params_for(args.crs, args.precision)
    solver.set_fdm_variant(args.fdm_variant)

    def solve(r, p=None):
        u = torch.zeros_like(rhs0)
        return F.hmholtz("PRES" if hsmg else "TEMP", u, r, h1, h2, F.mask, F.mult, -args.tol,
                         args.maxit, params=p if p is not None else prm)

    # plain copy of 2 GiB by torch: the achievable-bandwidth reference for
    # the roofline numbers (and the calibration kernel of the PMC byte counts)
    src = torch.empty(2 ** 28, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    tc = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * src.numel() * 8 / (time.perf_counter() - tc) / 1e9
    del src, dst
    log(f"[bench] torch copy: {copy_gbs:.0f} GB/s (read+write)")

    for w in range(args.warmup):
        out = solve(rhs0.clone())
        log(f"[bench] warmup {w}: niter={out['niter']}")
    rhss = [rhs0.clone() for _ in range(args.steps)]
    lib.nkh_timers_enable(1)
    lib.nkh_timers_reset()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    iters = []
    for k in range(args.steps):
        out = solve(rhss[k])
        iters.append(out["niter"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    lib.nkh_timers_enable(0)
    log(f"[bench] timed {args.steps} solves: iters={iters} in {elapsed:.3f} s")

    import ctypes

    sec = ctypes.c_double()
    calls = ctypes.c_int64()
    lib.nkh_timers_read(b"axhm", ctypes.byref(sec), ctypes.byref(calls))
    t_ax = sec.value / max(calls.value, 1)
    lib.nkh_timers_read(b"dsum", ctypes.byref(sec), ctypes.byref(calls))
    t_gs = sec.value / max(calls.value, 1)
    # the kernel the autotuner chose for this order's fused apply in the
    # warm-up solve
    v_cg, v_flex = solver.axhelm_tuning_fused(args.lx1)
    kname = {0: "k_axhelm", 1: "k_axhelm_pl", 2: "k_axhelm_pl", 3: "k_axhelm",
             4: "k_axhelm_mx", 5: "k_axhelm_mxp"}[v_flex if hsmg else v_cg]
    if hsmg:
        # axhelm of ax (gmres.f:284) with the p update and the previous
        # iteration's x += alpha p fused: reads p, z, x, g1..g6, bm1; writes
        # p, x, w
        bpn = 8.0 * 13
        ax_kernel = (f"{kname}<{args.lx1},2,false,1> (flex-CG axhelm, p = beta p + z and "
                     f"x += alpha p fused)")
    else:
        bpn = fused_axhelm_bytes_per_node(args.lx1)
        ax_kernel = f"{kname}<{args.lx1},1,false,1> (CG-fused axhelm)"
    ax_bytes = bpn * box.n
    # (the mixed solver's FP32 operator records no axhelm events: no roofline
    # entry for a --precision 2 main line)
    # per-kernel table from the library's HIP-event timers over the timed
    # solves, each with its algorithmic bytes per launch
    fb = 1.0 - ((args.lx1 - 2) / args.lx1) ** 3  # element-boundary fraction of the nodes
    kernels = [dict(kernel=ax_kernel, ms=t_ax * 1e3, bytes_per_launch=ax_bytes),
               dict(kernel="k_gs_assemble<double> (w = mask dssum(w))", ms=t_gs * 1e3,
                    bytes_per_launch=box.n * fb * (8.0 * 2 + 4.0))]
    if hsmg:
        lib.nkh_timers_read(b"fdm", ctypes.byref(sec), ctypes.byref(calls))
        t_fdm = sec.value / max(calls.value, 1)
        lib.nkh_timers_read(b"mgv", ctypes.byref(sec), ctypes.byref(calls))
        t_mgv = sec.value / max(calls.value, 1)
        nh = args.lx1
        kernels.append(dict(kernel=f"k_sch_fdm_d<{nh + 2}> (fine-level local FDM solves)",
                            ms=t_fdm * 1e3,
                            bytes_per_launch=fdm_bytes_per_element(nh) * box.nel))
    for k in kernels:
        k["GBs"] = k["bytes_per_launch"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else None
        k["hbm_frac"] = k["GBs"] / HBM_PEAK_GBS if k["GBs"] else None
    dom = max(kernels, key=lambda k: k["ms"])
    achieved = dom["GBs"]
    total_iters = sum(iters)
    value = free_dof * total_iters / elapsed / 1e9
    mg_info = F.h1mg_info() if hsmg else None
    variants = []
    if not args.no_variants:
        # beside the main line (not its value): the other coarse solver, and
        # the mixed-precision smoother; one warm-up and two timed solves each
        def measure(label, p):
            solve(rhs0.clone(), p)
            rv = [rhs0.clone() for _ in range(2)]
            torch.cuda.synchronize()
            tv = time.perf_counter()
            outs = [solve(rv[k], p) for k in range(2)]
            torch.cuda.synchronize()
            el_v = time.perf_counter() - tv
            it_v = [o["niter"] for o in outs]
            # the same tolerance reached: the final (for the mixed solver the
            # true FP64) residual norm over the initial one
            v = {"variant": label, "iterations_per_solve": it_v, "seconds_per_solve": el_v / 2,
                 "value": free_dof * sum(it_v) / el_v / 1e9,
                 "final_rel_residual": outs[-1]["rbn2"] / outs[-1]["rbn0"],
                 "converged": all(o["converged"] for o in outs),
                 "time_to_solution_vs_fp64": (elapsed / args.steps) / (el_v / 2)}
            log(f"[bench] {v}")
            variants.append(v)

        if hsmg:
            other = 1 - args.crs
            measure("coarse: " + ("one AMG V-cycle (param40=1)" if other == 1 else
                                  "AMG-PCG to 1e-13"), params_for(other))
            measure("mixed: FP32 fine-level smoother, FP64 Krylov and coarse",
                    params_for(args.crs, 1))
            measure("mixed + one-V-cycle coarse", params_for(1, 1))
            measure("mixed: FP32 inner flexible CG (operator, fine-level V-cycle, vectors) in "
                    "FP64 defect correction, FP64 coarse", params_for(args.crs, 2))
            measure("mixed FP32 inner solve + one-V-cycle coarse", params_for(1, 2))
            # the matrix-core trial of the FP32 local solves (bf16 operands,
            # FP32 accumulation); the default stays the vector kernel
            solver.set_fdm_variant(2)
            measure("mixed FP32 inner solve, FDM contractions in bf16 on the matrix cores",
                    params_for(args.crs, 2))
            solver.set_fdm_variant(1)
            measure("mixed FP32 inner solve, FDM contractions in FP32 on the matrix cores",
                    params_for(args.crs, 2))
            # the FP64 smoother's contractions on the FP64 matrix cores
            solver.set_fdm_variant(3)
            measure("FP64 solve, FDM contractions in FP64 on the matrix cores",
                    params_for(args.crs, 0))
            solver.set_fdm_variant(0)
        else:
            measure("mixed: FP32 Jacobi-PCG iteration (operator, vectors), FP64 dots and "
                    "residual refreshes", params_for(0, 2))

    host_io = None
    if not args.no_variants:
        # the same solve through the host-array entry point (nkh_hmholtz:
        # rhs, h1, h2, mask, mult in and u, rhs out over PCIe, pageable
        # numpy memory): the PCIe-inclusive rate, never the bench value
        arrs = {k: np.ascontiguousarray(t.cpu().numpy()) for k, t in
                (("rhs", rhs0), ("h1", h1), ("h2", h2), ("mask", F.mask), ("mult", F.mult))}
        u_h = np.zeros_like(arrs["rhs"])
        res = _lib.NkhResult()

        def host_solve():
            r_h = arrs["rhs"].copy()
            t = time.perf_counter()
            _lib.check(lib.nkh_hmholtz(F.handle, b"PRES" if hsmg else b"TEMP",
                                       u_h.ctypes.data, r_h.ctypes.data, arrs["h1"].ctypes.data,
                                       arrs["h2"].ctypes.data, arrs["mask"].ctypes.data,
                                       arrs["mult"].ctypes.data, -args.tol, args.maxit, 1,
                                       ctypes.byref(prm), ctypes.byref(res), None), "nkh_hmholtz")
            return time.perf_counter() - t

        host_solve()
        th = host_solve()
        host_io = {"entry": "nkh_hmholtz (host arrays: 5 fields in, 2 out)",
                   "seconds_per_solve": th, "iterations": res.niter,
                   "value": free_dof * res.niter / th / 1e9,
                   "vs_device_arrays": (elapsed / args.steps) / th}
        log(f"[bench] host arrays {host_io}")
        del arrs, u_h

    traffic = None
    if hsmg and args.nel == 64 and args.lx1 == 8 and args.precision == 0:
        traffic = measured_traffic("sch_fdm_E262144_N7_pmc.md" if "sch_fdm" in dom["kernel"]
                                   else "axhelm_flexcg_E262144_N7_pmc.md")

    targets = None
    if not args.no_variants and hsmg and args.precision == 0:
        # the stated targets of this workload, recorded with the bench line
        mixed = [v for v in variants if v["variant"].startswith("mixed: FP32 inner flexible CG")]
        F.destroy()
        del rhs0, rhss, h1, h2, bm1
        torch.cuda.empty_cache()
        log("[bench] targets: batch ...")
        bl = batch_run()
        bv = bl["variants"]
        log("[bench] targets: adjoint ...")
        al = adjoint_run()
        targets = {
            "mixed_time_to_solution_vs_fp64": mixed[0]["time_to_solution_vs_fp64"] if mixed else None,
            "mixed_value": mixed[0]["value"] if mixed else None,
            "mixed_final_rel_residual": mixed[0]["final_rel_residual"] if mixed else None,
            "target_mixed_ratio": 1.6,
            "batch": {"workload": bl["workload"],
                      "sequential_fp64_s": bv["sequential_fp64"]["seconds"],
                      "batched_fp64_s": bv["batched_fp64"]["seconds"],
                      "batched_mixed_s": bv["batched_mixed"]["seconds"],
                      "batched_mixed_vs_sequential_fp64":
                          bv["batched_mixed"]["speedup_vs_sequential_fp64"],
                      "batched_mixed_GDOF_iter_s": bv["batched_mixed"]["GDOF_iter_s"],
                      "sequential_fp64_GDOF_iter_s": bv["sequential_fp64"]["GDOF_iter_s"],
                      "batched_mixed_max_rel_diff_vs_fp64":
                          bv["batched_mixed"]["max_rel_diff_vs_fp64"],
                      "converged": all(all(r["converged"]) for r in bv.values())},
            "adjoint": {k: al[k] for k in ("workload", "fd_directions", "fd_step",
                                           "fd_max_rel_diff", "forward_ms", "adjoint_ms",
                                           "gradient_kernel_ms", "gradient_kernel_GBs")},
        }
        log(f"[bench] targets {targets}")

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(args.cpu_nel, args.lx1, 0.1, args.cpu_seconds, args.solver)
        log(f"[bench] cpu baseline {cpu}")

    prec_name = {0: "fp64", 1: "fp64 (FP32 fine-level smoother)",
                 2: "mixed (FP32 inner iteration, FP64 residuals)"}[args.precision]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "GDOF*iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"hmholtz E={args.nel}^3={box.nel} N={args.lx1 - 1} deformed box "
                         f"(0.1h sin map), h1=1 h2=10, f~U(-1,1), Dirichlet, "
                         + (f"flexible CG + h1mg V-cycle (PRES, param42=2) {prec_name} to "
                            f"{args.tol:g} relative" if hsmg else
                            f"Jacobi-PCG {prec_name} to {args.tol:g} relative")),
            "nel": box.nel, "lx1": args.lx1,
            "preconditioner": "h1mg" if hsmg else "jacobi", "precision": prec_name,
            "fdm_variant": args.fdm_variant,
            "multigrid": mg_info,
            "coarse_solver": (None if not hsmg else
                              ("one AMG V-cycle (param40=1)" if args.crs == 1 else
                               "AMG-preconditioned CG to 1e-13 (exact, as XXT)")),
            "variants": variants,
            "seconds_per_solve": elapsed / args.steps,
            "free_dof": free_dof, "iterations_per_solve": iters,
            "axhelm_ms": t_ax * 1e3, "gs_ms": t_gs * 1e3,
            "vcycle_ms": (t_mgv * 1e3 if hsmg else None),
            "kernels": kernels,
            "ms_per_iteration": elapsed / total_iters * 1e3,
            "copy_GBs": copy_gbs,
            "host_arrays": host_io,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
            "kernel": dom["kernel"], "ms_per_launch": dom["ms"],
            "bytes_per_launch": dom["bytes_per_launch"],
        },
        "cpu_baseline": cpu,
        "targets": targets,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
