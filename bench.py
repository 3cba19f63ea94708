"""Benchmark: residual Mcell-updates/s of the MI355X residual path.

Headline workload (one GPU): 3D Euler, 100^3 hexes split into 6,000,000 tets on a periodic
[0,10]^3 box (vertices jittered +-5% from seed 2, cells in Hilbert order on upload; the
reverse Cuthill-McKee order is timed beside it), isentropic vortex (strength 5) advected at
M=0.5, second-order MUSCL with LSQ gradients and the Venkatakrishnan limiter (K=5), HLLC,
classical RK4, CFL 0.8.  A step is one RK4 step = 4 residual evaluations (time step on the
device included); Mcell-updates/s = cells * 4 * steps / time.  `value` is the FP64 form; the
mixed form (FP32 fluxes and gradients; FP64 state, RK accumulation and primitive recovery), the
mixed form that keeps the stage-increment sum in FP32, their per-field errors after 1000 steps and
the other workloads are reported beside it:

- c1: 2D channel, 256 x 128 jittered quads, first-order HLLC + RK2, 500 steps, FP64, with its
  flux kernel's roofline (the CPU oracle runs it at full size beside the GPU);
- adjoint / c4: the 249,480-tet bump channel (first-order HLLC + RK2): primal rate and roofline,
  adjoint/primal kernel time, and the wall-force gradient (64 wall vertices after 20 steps, the
  inflow state after 400) against central differences and the tangent-linear product;
- implicit: 160^3 = 4.1M-hex Navier-Stokes Taylor-Green, matrix-free backward Euler, with the
  roofline of the JVP's residual evaluation;
- ns_16m: 256^3 = 16.8M-hex Navier-Stokes, RK4, FP64 and mixed, its own roofline and CPU
  baseline, adjoint/primal ratio, a 200-step checkpointed adjoint, and the device mesh build of
  its box beside the host (numpy) build of a 64^3 sample.

The large meshes are built on the device (faces and geometry: mesh.build_ustmesh(..., device=0),
the host build's arrays bit for bit).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 100] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "residual Mcell-updates/s, FP64 vs FP32+refine at stated err; adjoint/primal ratio"
HBM_PEAK_GBS = 8000.0


def byte_model(nc, nf, nb, dim=3, nv=5, mixed=False, rk4=True, gradients=True, f32sum=False):
    """Algorithmic HBM bytes per launch of the two residual kernels (compulsory traffic, each
    array once per pass, int32 indices): gradient+limiter pass and flux+gather+RK-update pass.
    gradients=False: a first-order inviscid residual (no gradient pass); rk4=False: the RK2
    (midpoint) combine instead of RK4's.  RK4 per step and value: U0 read by each stage and written
    by the last (5 x 8 B), the stage-increment sum written by stages 1-3 and read by 2-4 (6 x 8 B;
    6 x 4 B with f32sum, the mixed form that keeps that sum in FP32) and three stage states
    (3 x 4 or 8 B), a quarter of it per launch."""
    ps = pg = px = 4 if mixed else 8
    pa = 4 if mixed and f32sum else 8
    E = 2 * nf - nb
    b_grad = nc * (nv * ps + 8 * dim + dim * (dim + 1) // 2 * pg + dim * nv * pg + 4) + 8 * nf + 4 * E
    b_flux = nc * (nv * ps + 8 + 4) + nf * (dim * px + 8) + 4 * E
    if gradients:
        b_flux += nc * (dim * nv * pg + 8 * dim) + nf * dim * px
    else:
        b_grad = 0
    b_rk = nc * nv * (10 + 1.5 * pa + 0.75 * ps) if rk4 else nc * nv * (ps + 16) / 2
    return dict(grad=b_grad, flux=b_flux + b_rk, residual_only=b_grad + b_flux, total=b_grad + b_flux + b_rk)


def flux_roofline(bytes_per_launch, tm, what):
    """roofline object of the flux kernel from its event-timed launches"""
    ms = tm["flux_ms"] / max(1, tm["flux_n"])
    ach = bytes_per_launch / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "k_flux", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes_per_launch": int(bytes_per_launch),
            "avg_launch_ms": round(ms, 5), "model": what}


def run_gpu(m, s, g, u, steps, warmup, precision, order="hilbert"):
    from typhon_amd.solver import Solver
    sc = dict(s)
    sc["precision"] = precision
    solver = Solver(m, sc, g, cell_order=order)
    solver.set_state(u)
    solver.advance(warmup)
    solver.sync()
    solver.timing(True)  # reset and enable per-kernel events
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731
    sync()
    solver.sync()
    t0 = time.perf_counter()
    solver.advance(steps)
    solver.sync()
    sync()
    t1 = time.perf_counter()
    tm = solver.timing(False)
    state = solver.get_state()
    info = solver.mesh_info()
    solver.close()
    return t1 - t0, tm, state, info


def cpu_baseline(n_cpu, budget_s=20.0):
    """Oracle C restatement (OpenMP over the host's cores: per-face and per-cell loops in
    parallel, face->cell scatters serial) on a bounded sample of the same workload."""
    import cases as K
    from oracle import oracle as OR
    threads = int(OR.lib().or_threads())
    m, s, g, u = K.vortex_case(n_cpu, L=10.0)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    steps = 0
    t0 = time.perf_counter()
    while True:
        dt = o.timestep(s["cfl"])
        o.rk_step(dt)
        steps += 1
        if time.perf_counter() - t0 > budget_s or steps >= 200:
            break
    el = time.perf_counter() - t0
    rate = m.ncell_int * 4 * steps / el / 1e6
    return dict(value=rate, unit="Mcell-updates/s", cores=threads, kind="port",
                us_per_cell_it_x_ncpu=round(threads * el / steps / m.ncell_int * 1e6, 4),
                sample=f"oracle C restatement, {threads} OpenMP threads, same scheme on {n_cpu}^3 hexes x 6 = "
                       f"{m.ncell_int} periodic tets, {steps} RK4 steps ({el:.1f} s)")


def field_errors(a, b, gamma=1.4):
    """relative L2 and Linf errors of state a against state b [5, n] per field: density, total
    energy, momentum (one vector field: norms of the vector difference over norms of the vector)
    and pressure; L2 = ||a - b||_2 / ||b||_2, Linf = max |a - b| / max |b|"""
    def p(u):
        return (gamma - 1.0) * (u[1] - 0.5 * (u[2] ** 2 + u[3] ** 2 + u[4] ** 2) / u[0])

    out = {}
    for name, fa, fb in (("rho", a[0], b[0]), ("rhoE", a[1], b[1]), ("momentum", a[2:5], b[2:5]),
                         ("p", p(a), p(b))):
        d = fa - fb
        if fa.ndim == 2:
            dn, bn = np.sqrt((d * d).sum(0)), np.sqrt((fb * fb).sum(0))
        else:
            dn, bn = np.abs(d), np.abs(fb)
        out[name] = dict(l2=float(np.sqrt((dn * dn).sum() / (bn * bn).sum())), linf=float(dn.max() / bn.max()))
    return out


def run_steps(solver, u, steps):
    """state after ``steps`` device-resident steps from u (time step on the device)"""
    solver.set_state(u)
    solver.advance(steps)
    solver.sync()
    return solver.get_state()


def mixed_accuracy(m, s, g, u, steps=1000, order="hilbert", forms=(1,)):
    """FP64 and the mixed ``forms`` (precision codes) advanced ``steps`` steps from the same state;
    per-field errors of each mixed form against FP64, one dict per form"""
    from typhon_amd import scheme as S
    from typhon_amd.solver import Solver
    st = {}
    for prec in (S.PREC_FP64,) + tuple(forms):
        solver = Solver(m, dict(s, precision=prec), g, cell_order=order)
        st[prec] = run_steps(solver, u, steps)
        solver.close()
    return [dict(steps=steps, reference="the FP64 GPU form (parity-tied to the oracle at 1e-12)",
                 fields=field_errors(st[f], st[S.PREC_FP64])) for f in forms]


def c1_bench(steps=500, reps=3, cpu=True):
    """2D channel (256 x 128 jittered quads, slip walls, Dirichlet inlet, pressure outlet, density
    bump), first-order HLLC + RK2 (reference midpoint), CFL 0.8, global time step: ``steps`` steps
    device-resident (best of ``reps``), and the CPU oracle over the same steps at full size"""
    import cases as K
    from typhon_amd.solver import Solver
    m, s, g, u = K.channel_case()
    solver = Solver(m, s, g)
    run_steps(solver, u, 20)
    best = np.inf
    for _ in range(reps):
        solver.set_state(u)
        solver.sync()
        t0 = time.perf_counter()
        solver.advance(steps)
        solver.sync()
        best = min(best, time.perf_counter() - t0)
    solver.set_state(u)
    solver.sync()
    solver.timing(True)
    solver.advance(steps)
    solver.sync()
    tm = solver.timing(False)
    info = solver.mesh_info()
    gpu_state = run_steps(solver, u, steps)
    solver.close()
    nc = m.ncell_int
    bm = byte_model(nc, info["nface"], info["nbface"], dim=2, nv=4, rk4=False, gradients=False)
    out = dict(workload=f"2D Euler channel, 256 x 128 = {nc} jittered quads, first-order HLLC + RK2, CFL 0.8, "
                        f"{steps} steps", cells=nc, steps=steps, value=round(nc * 2 * steps / best / 1e6, 1),
               unit="Mcell-updates/s", ms_per_step=round(best / steps * 1e3, 4), dtype="f64",
               us_per_cell_it=round(best / steps / nc * 1e6, 6),
               roofline=flux_roofline(bm["flux"], tm, f"first-order flux + RK2 combine, "
                                                      f"{bm['flux'] / nc:.1f} B/cell per launch"))
    if cpu:
        from oracle import oracle as OR
        o = K.make_oracle(m, s, g)
        o.set_cons(u)
        t0 = time.perf_counter()
        for _ in range(steps):
            o.rk_step(o.timestep(s["cfl"]))
        el = time.perf_counter() - t0
        threads = int(OR.lib().or_threads())
        out["cpu_baseline"] = dict(value=round(nc * 2 * steps / el / 1e6, 2), unit="Mcell-updates/s", cores=threads,
                                   kind="port", sample=f"oracle C restatement, {threads} OpenMP threads, the full "
                                                       f"workload ({steps} steps, {el:.1f} s)",
                                   us_per_cell_it_x_ncpu=round(threads * el / steps / nc * 1e6, 4))
        out["gpu_vs_oracle_after_steps"] = field_errors(gpu_state, o.get_cons())
    return out


def c4_bench(rate_steps=500, nvert=64, log=None):
    """the adjoint workload (249,480-tet bump channel, first-order HLLC + RK2): primal rate over
    ``rate_steps`` device-resident steps with its flux kernel's roofline, and the wall-force gradient
    checks (bump_gradient.check): ``nvert`` bump-wall vertices after 20 steps, the inflow rho, u, p
    after 400 steps"""
    import cases as K
    import bump_gradient as BG
    from typhon_amd.solver import Solver
    m, s, g, u = K.bump_case()
    solver = Solver(m, s, g)
    run_steps(solver, u, 10)
    solver.set_state(u)
    solver.sync()
    t0 = time.perf_counter()
    solver.advance(rate_steps)
    solver.sync()
    el = time.perf_counter() - t0
    solver.set_state(u)
    solver.sync()
    solver.timing(True)
    solver.advance(rate_steps)
    solver.sync()
    tm = solver.timing(False)
    info = solver.mesh_info()
    solver.close()
    nc = m.ncell_int
    bm = byte_model(nc, info["nface"], info["nbface"], rk4=False, gradients=False)
    out = dict(primal=dict(value=round(nc * 2 * rate_steps / el / 1e6, 1), unit="Mcell-updates/s",
                           ms_per_step=round(el / rate_steps * 1e3, 4), steps=rate_steps, dtype="f64",
                           roofline=flux_roofline(bm["flux"], tm, f"first-order flux + RK2 combine, "
                                                                  f"{bm['flux'] / nc:.1f} B/cell per launch")))
    for key, kw in (("gradient", dict(nvert=nvert)),
                    ("gradient_inflow", dict(nsteps=400, nvert=0, inflow=True))):
        rep = BG.check(log=log, **kw)
        rows = rep.pop("rows")
        rep["table"] = [[r["param"], r["adjoint"], r["fd"], r["rel_diff"], r["rel_to_norm"]] for r in rows]
        out[key] = rep
    return out


def ns_cpu_baseline(n=48, budget_s=15.0, max_steps=200):
    """the oracle on a bounded sample of the 16.8M-hex NS workload: the same scheme on n^3 hexes"""
    import cases as K
    from oracle import oracle as OR
    threads = int(OR.lib().or_threads())
    m, s, g, u = K.tgv_case(n, walls=True)
    o = K.make_oracle(m, s, g)
    o.set_cons(u)
    dt = o.timestep(s["cfl"])
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.rk_step(dt)
        steps += 1
        if time.perf_counter() - t0 > budget_s or steps >= max_steps:
            break
    el = time.perf_counter() - t0
    return dict(value=round(m.ncell_int * 4 * steps / el / 1e6, 3), unit="Mcell-updates/s", cores=threads,
                kind="port", us_per_cell_it_x_ncpu=round(threads * el / steps / m.ncell_int * 1e6, 4),
                sample=f"oracle C restatement, {threads} OpenMP threads, same scheme on {n}^3 = {m.ncell_int} hexes, "
                       f"{steps} RK4 steps ({el:.1f} s)")


_LIGHT = "--no-cpu --no-implicit --no-c1 --no-rcm --no-f32sum --accuracy-steps 0 --c4-vertices -1"
PROFILE_CMDS = {
    "trace": "rocprofv3 --kernel-trace --stats -f csv -d {d}/trace -o a -- python3 bench.py --steps 20 --warmup 2 "
             + _LIGHT + " --ns-steps 20 --ns-adj-steps 4",
    "fetch": "rocprofv3 --pmc FETCH_SIZE -f csv -d {d}/fetch -o a -- python3 bench.py --steps 2 --warmup 1 "
             + _LIGHT + " --ns-steps 2 --ns-adj-steps 2",
    "write": "rocprofv3 --pmc WRITE_SIZE -f csv -d {d}/write -o a -- python3 bench.py --steps 2 --warmup 1 "
             + _LIGHT + " --ns-steps 2 --ns-adj-steps 2",
    "sq": "rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_WAVE_CYCLES SQ_BUSY_CYCLES "
          "SQ_WAIT_INST_ANY SQ_ACTIVE_INST_VALU -f csv -d {d}/sq -o a -- python3 bench.py --steps 2 --warmup 1 "
          + _LIGHT + " --ns-steps 2 --ns-adj-steps 2",
    "imp": "rocprofv3 --kernel-trace --stats -f csv -d {d}/imp -o imp -- python3 bench.py --steps 1 --warmup 1 "
           "--no-cpu --no-mixed --no-adjoint --no-ns --no-c1 --no-rcm --accuracy-steps 0",
    "cfetch": "rocprofv3 --pmc FETCH_SIZE -f csv -d {d}/cfetch -o copy -- python3 bench.py --copy-calibration 1024",
    "cwrite": "rocprofv3 --pmc WRITE_SIZE -f csv -d {d}/cwrite -o copy -- python3 bench.py --copy-calibration 1024",
}


def _csv_rows(path):
    import csv
    with open(path) as f:
        return list(csv.DictReader(f))


def _pmc(path, counter):
    """kernel name -> per-dispatch values of one counter (in the counter's unit)"""
    out = {}
    for r in _csv_rows(path):
        if r["Counter_Name"] == counter:
            out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return out


def _mean(v):
    return sum(v) / len(v) if v else float("nan")


def adjoint_byte_model(nc, nf, nb, dim=3, nv=5):
    """Algorithmic HBM bytes of one adjoint evaluation (lambda -> (dR/dU)^T lambda) of the first-order
    viscous residual, priced with the same rule as the primal's: what the primal residual reads
    (gradient pass and flux pass, residual only) plus lambda read and the result written, FP64"""
    bm = byte_model(nc, nf, nb, dim, nv, mixed=False)
    return bm["residual_only"] + 2 * nc * nv * 8


def krylov_bytes_section(trace, ncell=160 ** 3, nv=5, pk=4):
    """Bytes of the FP32 Krylov vector kernels of the implicit profile, launch by launch.  An
    Arnoldi step is: the preconditioner, the JVP (its FD step size takes one multi-dot of a single
    vector), the Gram-Schmidt multi-dot over the step's j + 1 basis vectors, two CGS sweeps; a cycle
    ends with one linear combination.  So the basis size of a step is the number of first CGS sweeps
    since the cycle's last linear combination, and the multi-dot right before each first sweep is the
    Gram-Schmidt one."""
    if not os.path.exists(trace):
        return []
    rows = sorted(_csv_rows(trace), key=lambda r: int(r["Start_Timestamp"]))
    vec = ncell * nv * pk
    impl = alg = 0.0
    t_ns = 0
    steps = 0
    j1 = 0  # basis vectors of the current Arnoldi step
    last_dot = None
    for r in rows:
        name = r["Kernel_Name"]
        dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        if "k_multidot<float>" in name:
            if last_dot is not None:  # the previous one was a single-vector dot
                impl += vec
            last_dot = dur
            t_ns += dur
        elif "k_cgs_pass<float, true" in name:
            j1 += 1
            steps += 1
            impl += (j1 + 1) * vec + (j1 + 2) * vec  # its multi-dot (w, V_0..V_j) and this sweep
            alg += 2 * j1 * vec  # the model's orthogonalisation: 2 j vectors per step
            last_dot = None
            t_ns += dur
        elif "k_cgs_pass<float, false" in name:
            impl += (j1 + 2) * vec
            t_ns += dur
        elif "k_lincomb<float>" in name:
            impl += (j1 + 1) * vec
            alg += (j1 + 1) * vec
            t_ns += dur
            j1 = 0
    if last_dot is not None:
        impl += vec
    if not steps or not t_ns:
        return []
    t = t_ns * 1e-9
    return ["", "## FP32 Krylov vector kernels (multi-dots, the two CGS2 sweeps, the cycle's linear combination)", "",
            f"{steps} Arnoldi steps in this run, {t * 1e3:.1f} ms in these kernels ({t / steps * 1e3:.3f} ms per step); "
            f"a vector is {vec / 1e6:.1f} MB (5 x 4,096,000 FP32).",
            f"- bytes of the implementation (each Gram-Schmidt launch reads its j + 1 basis vectors and w, the "
            f"sweeps also write w): {impl / 1e9:.1f} GB, {impl / t / 1e12:.2f} TB/s (the first basis vectors of a "
            f"cycle are still in the 256 MB last-level cache);",
            f"- algorithmic bytes of the Krylov model (2 j vectors per step for the orthogonalisation, the "
            f"correction's linear combination): {alg / 1e9:.1f} GB, {alg / t / 1e12:.2f} TB/s, "
            f"{alg / t / 8e12:.3f} of the 8 TB/s peak.",
            "CGS2 reads the basis three times per step where the model's modified Gram-Schmidt reads it twice; "
            "it needs one host round trip per step instead of j + 1."]


def summarize_profiles(d, n=100, ns_n=256):
    """Turn the rocprofv3 CSVs under d (made by the PROFILE_CMDS commands) into the committed
    summaries under profiles/: per-kernel durations, HBM bytes per launch from FETCH_SIZE (KiB) and
    WRITE_SIZE (KiB) scaled by what the same counters read for a plain 1 GiB device copy, the
    algorithmic bytes of each residual kernel and SQ issue counters."""
    nc2, nf2 = 6 * n ** 3, 12 * n ** 3  # periodic Kuhn-split box: every face shared by two tets
    ncn = ns_n ** 3
    nfn, nbn = 3 * ncn - ns_n * ns_n, 2 * ns_n * ns_n  # periodic in x and z, slip walls at the y ends
    stats = {r["Name"]: r for r in _csv_rows(os.path.join(d, "trace", "a_kernel_stats.csv"))}
    copy_bytes = 1024 << 20

    def copy_factor(sub, counter):
        vals = [v * 1024 for k, vs in _pmc(os.path.join(d, sub, "copy_counter_collection.csv"), counter).items()
                if "opy" in k for v in vs if v * 1024 > copy_bytes / 4]
        return copy_bytes / _mean(vals)

    f_fac, w_fac = copy_factor("cfetch", "FETCH_SIZE"), copy_factor("cwrite", "WRITE_SIZE")
    fe = _pmc(os.path.join(d, "fetch", "a_counter_collection.csv"), "FETCH_SIZE")
    wr = _pmc(os.path.join(d, "write", "a_counter_collection.csv"), "WRITE_SIZE")
    sqf = os.path.join(d, "sq", "a_counter_collection.csv")
    sq = {c: _pmc(sqf, c) for c in ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "SQ_WAVE_CYCLES",
                                    "SQ_ACTIVE_INST_VALU")} if os.path.exists(sqf) else {}
    models = {  # kernel name prefix -> (label, algorithmic bytes per launch)
        "void ty::k_gradient<3, double, 4, 1, true, false>": ("6M tets K1 FP64 (LSQ + Venkatakrishnan, face states)",
                                                            byte_model(nc2, nf2, 0)["grad"]),
        "void ty::k_flux<3, double, true, false, 4, false, true, false>": ("6M tets K2 FP64 (HLLC, gather, RK4)",
                                                                   byte_model(nc2, nf2, 0)["flux"]),
        "void ty::k_gradient<3, float, 4, 1, true, false>": ("6M tets K1 FP32",
                                                           byte_model(nc2, nf2, 0, mixed=True)["grad"]),
        "void ty::k_flux<3, float, true, false, 4, false, true, false>": ("6M tets K2 FP32 (FP64 RK accumulation)",
                                                                  byte_model(nc2, nf2, 0, mixed=True)["flux"]),
        "void ty::k_gradient<3, double, 6, 0, true, true>": ("16.8M hexes NS K1 FP64 (LSQ, viscous records)",
                                                           byte_model(ncn, nfn, nbn)["grad"]),
        "void ty::k_flux<3, double, false, true, 6, false, true, false>": ("16.8M hexes NS K2 FP64 (HLLC + viscous, RK4)",
                                                                   byte_model(ncn, nfn, nbn)["flux"]),
        "void ty::k_gradient<3, float, 6, 0, true, true>": ("16.8M hexes NS K1 FP32",
                                                          byte_model(ncn, nfn, nbn, mixed=True)["grad"]),
        "void ty::k_flux<3, float, false, true, 6, false, true, false>": ("16.8M hexes NS K2 FP32 (FP64 RK accumulation)",
                                                                  byte_model(ncn, nfn, nbn, mixed=True)["flux"]),
    }
    adj_kernels = ("k_adj_face_geom<3>", "k_res_adjoint_face<3, true, false>", "k_res_adjoint_face<3, true, true>",
                   "k_res_adjoint_bnd_cells<3>", "k_res_adjoint_gather<3, true, 6>", "k_res_adjoint_tail<3>")

    def row(name):
        st = stats.get(name)
        us = float(st["AverageNs"]) / 1e3 if st else float("nan")
        f = _mean(fe.get(name, [])) * 1024 * f_fac
        w = _mean(wr.get(name, [])) * 1024 * w_fac
        return us, f, w

    def sq_cols(name):
        if not sq or name not in sq["SQ_WAVES"]:
            return "| | |"
        waves = _mean(sq["SQ_WAVES"][name])
        valu = _mean(sq["SQ_INSTS_VALU"][name]) / waves
        vmem = _mean(sq["SQ_INSTS_VMEM_RD"][name]) / waves
        busy = _mean(sq["SQ_ACTIVE_INST_VALU"][name]) / _mean(sq["SQ_WAVE_CYCLES"][name])
        return f"| {valu:.0f} / {vmem:.0f} | {100 * busy:.0f}% |"

    head = ["Commands (from the repository root, output directory D):", ""]
    head += ["    " + c.format(d="D") for c in PROFILE_CMDS.values()]
    head += ["    python3 bench.py --summarize D", "",
             f"Copy calibration (1 GiB device copy): FETCH_SIZE factor {f_fac:.3f}, WRITE_SIZE factor {w_fac:.3f} "
             "(bytes moved / counter bytes).", ""]
    lines = ["# Residual kernels: durations, HBM traffic and issue counters", "",
             "6M periodic tets (100^3 hexes x 6; MUSCL + LSQ + Venkatakrishnan, HLLC, RK4) and 16.8M hexes (256^3, "
             "NS first-order HLLC + cell-full viscous flux, LSQ gradients, RK4), FP64 and FP32 (mixed: FP64 U0 and "
             "RK accumulator).  Algorithmic bytes: the byte model of bench.py (each array once per pass, int32 "
             "indices).", ""] + head
    lines += ["| kernel | avg us | read GB | written GB | HBM GB | HBM TB/s | algorithmic GB | traffic / algorithmic "
              "| algorithmic TB/s (frac of 8) | VALU / VMEM-read instr per wave | VALU busy |",
              "|---|---|---|---|---|---|---|---|---|---|---|"]
    traffic = {}
    for prefix, (label, alg) in models.items():
        name = next((k for k in stats if k.startswith(prefix)), None)
        if name is None:
            continue
        us, f, w = row(name)
        tb = (f + w) / (us * 1e-6) / 1e12
        at = alg / (us * 1e-6) / 1e12
        lines.append(f"| {label} | {us:.1f} | {f / 1e9:.3f} | {w / 1e9:.3f} | {(f + w) / 1e9:.3f} | {tb:.2f} | "
                     f"{alg / 1e9:.3f} | {(f + w) / alg:.2f} | {at:.2f} ({at / 8:.3f}) {sq_cols(name)}")
        traffic[prefix] = dict(kernel=name, bytes_per_launch=int(f + w), algorithmic_bytes=int(alg),
                               fetch_factor=f_fac, write_factor=w_fac, avg_us=us)
    lines += ["", "## All kernels of the trace run (rocprofv3 --kernel-trace --stats)", "",
              "| kernel | calls | avg us | min us | max us | share % |", "|---|---|---|---|---|---|"]
    for name, r in stats.items():
        if "ty::" not in name:
            continue
        lines.append(f"| {name.replace('|', '/')[:110]} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", "residual_kernels.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    for prefix, name in ((list(models)[1], "traffic_flux_c2_fp64.json"), (list(models)[0], "traffic_grad_c2_fp64.json"),
                         (list(models)[5], "traffic_flux_ns16m_fp64.json"),
                         (list(models)[4], "traffic_grad_ns16m_fp64.json")):
        if prefix in traffic:
            with open(os.path.join(ROOT, "profiles", name), "w") as f:
                json.dump(traffic[prefix], f, indent=1)
    # adjoint kernels of the 16.8M-hex NS residual
    alg = adjoint_byte_model(ncn, nfn, nbn)
    al = ["# Adjoint kernels of the 16.8M-hex NS residual (FP64): durations and HBM traffic", "",
          "One adjoint evaluation (lambda -> (dR/dU)^T lambda) runs the face kernel (two-sided faces, then boundary "
          "faces), the boundary-cell kernel, the gather (with the LSQ-gradient transpose) and the tail gather; "
          "k_adj_face_geom builds the per-face viscous geometry on the first call after ty_set_scheme.  Algorithmic bytes of one "
          "evaluation: the primal residual's reads (gradient and flux passes) plus lambda and the result, "
          f"{alg / ncn:.0f} B per cell = {alg / 1e9:.2f} GB.", ""] + head
    al += ["| kernel | avg us | read GB | written GB | HBM TB/s |", "|---|---|---|---|---|"]
    tot_us = 0.0
    for k in adj_kernels:
        name = next((x for x in stats if k in x), None)
        if name is None:
            continue
        us, f, w = row(name)
        if "geom" not in k:
            tot_us += us
        al.append(f"| {name} | {us:.1f} | {f / 1e9:.3f} | {w / 1e9:.3f} | {(f + w) / (us * 1e-6) / 1e12:.2f} |")
    if tot_us > 0:
        at = alg / (tot_us * 1e-6) / 1e12
        al += ["", f"Evaluation (all but the geometry kernel): {tot_us:.0f} us; algorithmic {at:.2f} TB/s, "
                   f"{at / 8:.3f} of the 8 TB/s peak."]
    with open(os.path.join(ROOT, "profiles", "adjoint_kernels_ns16m.md"), "w") as f:
        f.write("\n".join(al) + "\n")
    imp_stats = os.path.join(d, "imp", "imp_kernel_stats.csv")
    if os.path.exists(imp_stats):
        il = ["# Implicit NS kernels, 160^3 = 4,096,000 periodic hexes (Taylor-Green, dt = 10 x explicit)", "",
              "Two backward-Euler steps of the matrix-free path: FP32 GMRES(30) with block-Jacobi inside FP64 "
              "refinement (FD Jacobian-vector products = FP32 residual kernels, multi-dot / CGS2 update / "
              "block-Jacobi apply), FP64 outer residuals and the block-Jacobi diagonal (JAC kernel).  The run "
              "also holds one FP64 step of the 6M-tet workload and the device mesh set-up kernels.", "",
              "Command (from the repository root, output directory D):", "",
              "    " + PROFILE_CMDS["imp"].format(d="D"), "    python3 bench.py --summarize D", "",
              "| kernel | calls | avg us | min us | max us | share % |", "|---|---|---|---|---|---|"]
        for r in _csv_rows(imp_stats):
            if "ty::" not in r["Name"] and "rocprim" not in r["Name"]:
                continue
            il.append(f"| {r['Name'].replace('|', '/')[:120]} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                      f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
        il += krylov_bytes_section(os.path.join(d, "imp", "imp_kernel_trace.csv"))
        with open(os.path.join(ROOT, "profiles", "implicit_kernels_c3.md"), "w") as f:
            f.write("\n".join(il) + "\n")
    print("\n".join(lines[:40]))


def adjoint_ratio(reps=20):
    """Adjoint/primal cost on the adjoint workload: 3D first-order HLLC residual on the bump channel
    (63 x 33 x 20 hexes x 6 = 249,480 tets).  Device time of one first-order residual kernel vs one
    adjoint kernel launch ((dR/dU)^T lambda plus the face-area, volume and inlet sensitivities),
    HIP events on the context's stream."""
    import cases as K
    from typhon_amd.solver import Solver
    m, s, g, u = K.bump_case()
    solver = Solver(m, s, g)
    solver.set_state(u)
    dt = solver.timestep()
    rng = np.random.default_rng(0)
    lam = rng.normal(size=(5, m.ncell_int))
    solver.calc_rhs(dt)
    solver.residual_adjoint(dt, lam)
    solver.timing(True)
    for _ in range(reps):
        solver.calc_rhs(dt)
    tp = solver.timing(True)
    for _ in range(reps):
        solver.residual_adjoint(dt, lam)
    ta = solver.timing(False)
    solver.close()
    prim = tp["flux_ms"] / max(1, tp["flux_n"])
    adj = ta["adjoint_ms"] / max(1, ta["adjoint_n"])
    return dict(workload=f"3D first-order HLLC residual, bump channel, {m.ncell_int} tets", primal_kernel_ms=round(prim, 4),
                adjoint_kernel_ms=round(adj, 4), ratio=round(adj / prim, 3))


def mesh_build_bench(m, phases, sample_n=64):
    """Mesh preprocessing (faces, face and cell geometry: create_face_connect + calc_meshgeom) of the
    NS box ``m`` built on the device, beside the host numpy build of the same box at sample_n^3
    hexes (the device build of that sample is checked equal to it, bit for bit)."""
    import cases as K
    from typhon_amd import mesh as M
    nc = m.ncell_int
    dev_s = phases["faces_and_geometry"]
    kern_s = phases.get("device_faces", 0.0) + phases.get("device_finish", 0.0)
    t0 = time.perf_counter()
    hs = K.tgv_case(sample_n, walls=True)[0]
    th = M.LAST_BUILD["faces_and_geometry"]
    t_case = time.perf_counter() - t0
    ds = K.tgv_case(sample_n, walls=True, device=0)[0]
    same = all(np.array_equal(getattr(hs, k), getattr(ds, k)) for k in
               ("ic1", "ic2", "fvtx", "fnorm", "fcent", "fsurf", "ccent", "cvol"))
    return dict(workload=f"faces + face/cell geometry of the {round(nc ** (1 / 3))}^3-hex box ({nc} cells, "
                         f"{m.nface} faces)",
                device_s=round(dev_s, 2), device_Mcells_per_s=round(nc / dev_s / 1e6, 2),
                device_calls_s=round(kern_s, 2),
                device_calls_note="ty_meshbuild_faces + ty_meshbuild_finish, host-device copies included",
                host_numpy=dict(sample=f"{sample_n}^3 hexes", s=round(th, 2),
                                Mcells_per_s=round(sample_n ** 3 / th / 1e6, 3), case_s=round(t_case, 2)),
                speedup_vs_host=round((nc / dev_s) / (sample_n ** 3 / th), 1),
                device_equals_host_on_sample=bool(same))


def ns_bench(n=256, steps=1000, warmup=2, adj_steps=200, snaps=20, f32sum=True):
    """The 16M-cell unsteady NS workload: compressible Taylor-Green field on n^3 hexes (Re 1e4,
    Sutherland, Pr 0.72, 1% seeded noise), periodic in x and z with slip walls at y = 0 and
    y = 2 pi L (the field's v vanishes there), first-order HLLC + cell-full viscous flux with LSQ
    gradients, RK4 at CFL 0.5, one uploaded mesh for everything below.

    - throughput: FP64, mixed (FP32 flux/gradient/stage state, FP64 U0 and RK accumulator) and, with
      ``f32sum``, the mixed form that keeps the stage-increment sum in FP32 advance the same initial
      field by ``steps`` steps; per-field errors of each mixed form vs FP64;
    - adjoint: J = pressure force normal to the y = 0 wall after ``adj_steps`` FP64 steps,
      dJ/dU0 by the device-resident checkpointed sweep (``snaps`` snapshots, FP64), checked against a central difference of two FP64 forward runs along a seeded
      direction."""
    import torch
    import cases as K
    from typhon_amd import adjoint as A
    from typhon_amd import scheme as S
    from typhon_amd.solver import Solver
    from typhon_amd import mesh as M
    t0 = time.perf_counter()
    m, s, g, u = K.tgv_case(n, walls=True, device=0)
    tmesh = time.perf_counter() - t0
    phases = dict(M.LAST_BUILD)
    t0 = time.perf_counter()
    solver = Solver(m, s, g)
    tup = time.perf_counter() - t0
    nc = m.ncell_int
    out = dict(workload=f"3D NS Taylor-Green, {n}^3 = {nc} hexes (periodic x/z, slip walls y), first-order HLLC + "
                        f"viscous (cell-full), LSQ gradients, RK4, CFL 0.5", cells=nc, steps=steps,
               host_setup_s=dict(case_with_device_mesh_build=round(tmesh, 1), upload=round(tup, 1)),
               mesh_build=mesh_build_bench(m, phases))
    states = {}
    dt = None
    forms = [("fp64", S.PREC_FP64), ("mixed", S.PREC_MIXED)] + ([("mixed_f32sum", S.PREC_MIXED_F32SUM)] if f32sum else [])
    for name, prec in forms:
        sc = dict(s)
        sc["precision"] = prec
        solver.set_scheme(sc, g)
        solver.set_state(u)
        if dt is None:
            dt = solver.timestep()
        for _ in range(warmup):
            solver.rk_step(dt)
        solver.set_state(u)
        solver.sync()
        solver.timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            solver.rk_step(dt)
        solver.sync()
        el = time.perf_counter() - t0
        tm = solver.timing(False)
        states[name] = solver.get_state()
        out[name] = dict(value=round(nc * 4 * steps / el / 1e6, 1), unit="Mcell-updates/s",
                         ms_per_step=round(el / steps * 1e3, 3),
                         grad_ms_avg=round(tm["grad_ms"] / max(1, tm["grad_n"]), 4),
                         flux_ms_avg=round(tm["flux_ms"] / max(1, tm["flux_n"]), 4))
        info = solver.mesh_info()
        bm = byte_model(nc, info["nface"], info["nbface"], mixed=prec != S.PREC_FP64, f32sum=prec == S.PREC_MIXED_F32SUM)
        for kern in ("grad", "flux"):
            out[name][f"{kern}_achieved_GBs"] = round(bm[kern] / (out[name][f"{kern}_ms_avg"] * 1e-3) / 1e9, 1)
        out[name]["bytes_per_cell_eval"] = round(bm["total"] / nc, 1)
        if name == "fp64":  # the dominant kernel of the FP64 form and its HBM roofline
            kern = "flux" if out[name]["flux_ms_avg"] >= out[name]["grad_ms_avg"] else "grad"
            ms_avg = out[name][f"{kern}_ms_avg"]
            ach = bm[kern] / (ms_avg * 1e-3) / 1e9
            traffic = load_traffic(os.path.join(ROOT, "profiles", f"traffic_{kern}_ns16m_fp64.json"))
            out["roofline"] = {"bound": "hbm", "kernel": f"k_{'flux' if kern == 'flux' else 'gradient'}",
                               "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(ach / HBM_PEAK_GBS, 4),
                               "traffic": traffic.get("bytes_per_launch") if traffic else None,
                               "traffic_achieved_GBs": (round(traffic["bytes_per_launch"] / (ms_avg * 1e-3) / 1e9, 1)
                                                        if traffic else None),
                               "algorithmic_bytes_per_launch": int(bm[kern]), "avg_launch_ms": ms_avg}
        print(f"# ns {name}: {out[name]}", file=sys.stderr, flush=True)
    # adjoint/primal cost of one NS residual: device time of the FP64 primal kernels (LSQ gradient +
    # flux) vs the adjoint kernel chain (face reverse sweeps, gather with the gradient transpose, tail)
    solver.set_scheme(dict(s, precision=S.PREC_FP64), g)
    solver.set_state(u)
    lam = np.random.default_rng(5).normal(size=(5, nc))
    solver.calc_rhs(dt)
    solver.residual_adjoint(dt, lam, area_vectors=False, volumes=False)
    solver.sync()
    solver.timing(True)
    for _ in range(5):
        solver.calc_rhs(dt)
    tp = solver.timing(True)
    for _ in range(5):
        solver.residual_adjoint(dt, lam, area_vectors=False, volumes=False)
    ta = solver.timing(False)
    prim = tp["grad_ms"] / max(1, tp["grad_n"]) + tp["flux_ms"] / max(1, tp["flux_n"])
    adj = ta["adjoint_ms"] / max(1, ta["adjoint_n"])
    out["adjoint_ratio"] = dict(primal_kernels_ms=round(prim, 4), adjoint_kernels_ms=round(adj, 4),
                                ratio=round(adj / prim, 3))
    print(f"# ns adjoint/primal: {out['adjoint_ratio']}", file=sys.stderr, flush=True)
    del lam
    r64, rmx = states.pop("fp64"), states.pop("mixed")
    out["speedup_mixed_vs_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    out["mixed_error_after_steps"] = dict(steps=steps, reference="the FP64 GPU form",
                                          fields=field_errors(rmx, r64))
    if "mixed_f32sum" in states:
        rfx = states.pop("mixed_f32sum")
        out["speedup_mixed_f32sum_vs_fp64"] = round(out["mixed_f32sum"]["value"] / out["fp64"]["value"], 3)
        out["mixed_f32sum_error_after_steps"] = dict(steps=steps, reference="the FP64 GPU form",
                                                     fields=field_errors(rfx, r64))
        del rfx
    out["dt"] = dt
    del r64, rmx
    if adj_steps > 0:
        dts = np.full(adj_steps, dt)
        obj = lambda U, Sx: A.wall_force(m, U, Sx, "JMIN", axis=1)  # noqa: E731
        solver.set_scheme(dict(s, precision=S.PREC_FP64), g)
        t0 = time.perf_counter()
        J, gU0, _, st = A.unsteady_gradient_device(solver, m, u, dts, obj, snaps=snaps)
        el = time.perf_counter() - t0
        # central difference of J along a seeded direction, FP64 forward runs
        solver.set_scheme(dict(s, precision=S.PREC_FP64), g)
        rng = np.random.default_rng(12)
        d = rng.normal(size=u.shape) * np.abs(u).max(axis=1, keepdims=True)
        h = 1e-6
        Sx = torch.as_tensor(m.fnorm * m.fsurf[:, None])
        Jpm = []
        t1 = time.perf_counter()
        for sgn in (1.0, -1.0):
            solver.set_state(u + sgn * h * d)
            for k in range(adj_steps):
                solver.rk_step(dt)
            Jpm.append(float(obj(torch.as_tensor(solver.get_state()), Sx)))
        tfd = time.perf_counter() - t1
        fd = (Jpm[0] - Jpm[1]) / (2 * h)
        ad = float((gU0 * d).sum())
        out["adjoint"] = dict(objective="pressure force normal to the y = 0 wall", steps=adj_steps, snapshots=snaps,
                              forward_precision="fp64", reverse_precision="fp64", J=J,
                              total_s=round(el, 2), device_sweep_s=round(st["ms"] / 1e3, 2),
                              forward_steps_in_sweep=st["forward_steps"], max_snapshots=st["max_snapshots"],
                              snapshots_allocated=st["snapshots_allocated"],
                              snapshot_gb=round(snaps * 5 * nc * 8 / 1e9, 1),
                              fd_check=dict(direction="seeded normal x field scale", h=h, dJ_adjoint=ad, dJ_fd=fd,
                                            rel_diff=abs(ad - fd) / max(abs(fd), 1e-300), fd_s=round(tfd, 1)))
        print(f"# ns adjoint: {out['adjoint']}", file=sys.stderr, flush=True)
    solver.close()
    return out


def implicit_bench(n=160, dt_mult=10.0, steps=2):
    """The implicit workload: compressible NS Taylor-Green field on n^3 periodic hexes (Re 1e4,
    Sutherland, Pr 0.72), first-order HLLC + cell-full viscous flux; one linearised backward-Euler
    step = FP32 right-preconditioned GMRES(30) (block-Jacobi 5x5 from the HLL Jacobians, one-sided
    FD Jacobian-vector products) inside FP64 refinement (central-difference JVP) to 1e-8."""
    import cases as K
    from typhon_amd.solver import Solver, krylov
    t0 = time.perf_counter()
    m, s, g, u = K.tgv_case(n, device=0)
    tmesh = time.perf_counter() - t0
    solver = Solver(m, s, g)
    solver.set_state(u)
    dt = dt_mult * solver.timestep()
    kry = krylov(nkrylov=30, max_restart=4, max_outer=8, inner_precision=1, inner_tol=1e-3, outer_tol=1e-8)
    solver.implicit_step(dt, kry)  # warm-up (allocation of the Krylov basis)
    solver.sync()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(solver.implicit_step(dt, kry))
    solver.sync()
    el = time.perf_counter() - t0
    # the residual kernels inside the Krylov loop (FP32 JVPs, FP64 outer residuals), event-timed on a
    # repeat of the same steps
    solver.set_state(u)
    solver.implicit_step(dt, kry)
    solver.sync()
    solver.timing(True)
    for _ in range(steps):
        solver.implicit_step(dt, kry)
    solver.sync()
    tm = solver.timing(False)
    info = solver.mesh_info()
    solver.close()
    inner = sum(st["inner_iters"] for st in stats)
    nc = m.ncell_int
    # one FP32 JVP = one residual evaluation (gradient + flux passes, no RK combine); the average
    # also holds the few slower FP64 launches of the outer passes, priced at the FP32 bytes, so the
    # achieved rate is understated
    bm = byte_model(nc, info["nface"], info["nbface"], mixed=True)
    jvp_ms = tm["grad_ms"] / max(1, tm["grad_n"]) + tm["flux_ms"] / max(1, tm["flux_n"])
    res_bytes = bm["residual_only"]
    ach = res_bytes / (jvp_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_gradient + k_flux (residual of a JVP)", "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "algorithmic_bytes_per_launch": int(res_bytes), "avg_launch_ms": round(jvp_ms, 4),
            "model": f"mixed residual-only, {res_bytes / nc:.1f} B/cell",
            "residual_kernel_share_of_step": round((tm["grad_ms"] + tm["flux_ms"]) / (el * 1e3), 3)}
    return dict(workload=f"3D NS Taylor-Green, {n}^3 = {m.ncell_int} periodic hexes, implicit backward Euler, "
                         f"dt = {dt_mult:g} x explicit", ms_per_implicit_step=round(el / steps * 1e3, 2),
                outer_iters=[st["outer_iters"] for st in stats], inner_iters=[st["inner_iters"] for st in stats],
                rel_res=[st["rel_res"] for st in stats], ms_per_arnoldi_step=round(el / max(1, inner) * 1e3, 3),
                mesh_build_s=round(tmesh, 1), roofline=roof)


def load_traffic(path):
    try:
        with open(path) as f:
            return json.load(f)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=100, help="hexes per box edge (6 n^3 tets)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=40)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-implicit", action="store_true")
    ap.add_argument("--no-ns", action="store_true", help="skip the 16M-cell NS FP64/mixed section")
    ap.add_argument("--ns-n", type=int, default=256, help="hexes per edge of the NS section's box")
    ap.add_argument("--ns-steps", type=int, default=1000)
    ap.add_argument("--ns-adj-steps", type=int, default=200, help="RK steps of the NS adjoint (0: skip)")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps of the FP64-vs-mixed per-field error run of the headline workload (0: skip)")
    ap.add_argument("--no-rcm", action="store_true", help="skip the reverse Cuthill-McKee order run")
    ap.add_argument("--no-f32sum", action="store_true",
                    help="skip the mixed form that keeps the RK stage-increment sum in FP32")
    ap.add_argument("--no-c1", action="store_true", help="skip the 2D channel workload")
    ap.add_argument("--c4-vertices", type=int, default=64,
                    help="wall vertices in the bump-channel gradient check (-1: skip the check)")
    ap.add_argument("--order", default="hilbert", choices=["hilbert", "rcm", "none"],
                    help="device cell order set at mesh upload")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (PMC byte-counter calibration)")
    ap.add_argument("--summarize", metavar="DIR", default=None,
                    help="write the profiles/ summaries from rocprofv3 output under DIR and exit")
    args = ap.parse_args()
    if args.summarize:
        summarize_profiles(args.summarize, args.n, args.ns_n)
        return
    if args.copy_calibration:
        import torch
        nbytes = args.copy_calibration << 20
        x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no inter-GPU path is built); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import cases as K
    from typhon_amd import scheme as S

    t0 = time.perf_counter()
    m, s, g, u = K.vortex_case(args.n, L=10.0, device=0)
    print(f"# mesh: {m.ncell_int} tets, {m.nface} faces (reference form), built on the device in "
          f"{time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    t64, tm64, st64, info = run_gpu(m, s, g, u, args.steps, args.warmup, S.PREC_FP64, args.order)
    print(f"# fp64: {t64:.3f} s for {args.steps} steps; kernels {tm64}", file=sys.stderr, flush=True)
    nc = m.ncell_int
    nf_dev = info["nface"]
    nb_dev = info["nbface"]
    evals = 4 * args.steps
    value = nc * evals / t64 / 1e6
    bm = byte_model(nc, nf_dev, nb_dev, mixed=False)
    # dominant kernel and its roofline
    kern = "flux" if tm64["flux_ms"] >= tm64["grad_ms"] else "grad"
    ms_avg = tm64[f"{kern}_ms"] / max(1, tm64[f"{kern}_n"])
    ach = bm[kern] / (ms_avg * 1e-3) / 1e9
    traffic = load_traffic(os.path.join(ROOT, "profiles", f"traffic_{kern}_c2_fp64.json"))
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t64 / args.steps * 1e3, 4),
        # the reference's normalized walltime (integration_cyclezone.f90:174-175): processes x
        # threads x walltime / iterations / cells, in microseconds; one GPU counts as one
        "us_per_cell_it": round(t64 / args.steps / nc * 1e6, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered periodic tet mesh, isentropic vortex initial field)",
        "config": {"workload": f"3D Euler MUSCL (LSQ gradients, Venkatakrishnan K=5) + HLLC + RK4, "
                               f"{args.n}^3 hexes -> {nc} periodic jittered tets, isentropic vortex",
                   "cells": nc, "faces": nf_dev, "evals_per_step": 4, "precision": "fp64",
                   "cell_order": args.order},
        "roofline": {"bound": "hbm", "kernel": f"k_{'flux' if kern == 'flux' else 'gradient'}",
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "traffic_achieved_GBs": (round(traffic["bytes_per_launch"] / (ms_avg * 1e-3) / 1e9, 1)
                                              if traffic else None),
                     "algorithmic_bytes_per_launch": int(bm[kern]),
                     "avg_launch_ms": round(ms_avg, 4)},
        "kernels_fp64": {"grad_ms_avg": round(tm64["grad_ms"] / max(1, tm64["grad_n"]), 4),
                         "flux_ms_avg": round(tm64["flux_ms"] / max(1, tm64["flux_n"]), 4),
                         "dt_ms_avg": round(tm64["dt_ms"] / max(1, tm64["dt_n"]), 4),
                         "bytes_per_cell_eval": round(bm["total"] / nc, 1)},
    }
    if not args.no_mixed:
        tmx, tmm, stmx, _ = run_gpu(m, s, g, u, args.steps, args.warmup, S.PREC_MIXED, args.order)
        bmm = byte_model(nc, nf_dev, nb_dev, mixed=True)
        fl = tmm["flux_ms"] / max(1, tmm["flux_n"])
        out["mixed"] = {"value": round(nc * evals / tmx / 1e6, 3), "unit": "Mcell-updates/s",
                        "dtype": "fp32 flux/gradient/stage state and face states; fp64 U0, RK accumulator, "
                                 "primitive recovery from total energy",
                        "speedup_vs_fp64": round(t64 / tmx, 3),
                        "error_after_timed_steps": dict(steps=args.steps + args.warmup,
                                                        fields=field_errors(stmx, st64)),
                        "flux_ms_avg": round(fl, 4),
                        "grad_ms_avg": round(tmm["grad_ms"] / max(1, tmm["grad_n"]), 4),
                        "flux_achieved_GBs": round(bmm["flux"] / (fl * 1e-3) / 1e9, 1)}
        print(f"# mixed: {out['mixed']['value']} Mcell-updates/s", file=sys.stderr, flush=True)
        forms = [("mixed", S.PREC_MIXED)]
        if not args.no_f32sum:
            tfx, tmf, stfx, _ = run_gpu(m, s, g, u, args.steps, args.warmup, S.PREC_MIXED_F32SUM, args.order)
            bmf = byte_model(nc, nf_dev, nb_dev, mixed=True, f32sum=True)
            fl = tmf["flux_ms"] / max(1, tmf["flux_n"])
            out["mixed_f32sum"] = {
                "value": round(nc * evals / tfx / 1e6, 3), "unit": "Mcell-updates/s",
                "dtype": "the mixed form with the RK stage-increment sum kept in fp32 between stages (fp64 U0, "
                         "its update and the primitive recovery)",
                "speedup_vs_fp64": round(t64 / tfx, 3),
                "error_after_timed_steps": dict(steps=args.steps + args.warmup, fields=field_errors(stfx, st64)),
                "flux_ms_avg": round(fl, 4), "grad_ms_avg": round(tmf["grad_ms"] / max(1, tmf["grad_n"]), 4),
                "flux_achieved_GBs": round(bmf["flux"] / (fl * 1e-3) / 1e9, 1)}
            print(f"# mixed_f32sum: {out['mixed_f32sum']['value']} Mcell-updates/s", file=sys.stderr, flush=True)
            forms.append(("mixed_f32sum", S.PREC_MIXED_F32SUM))
        if args.accuracy_steps > 0:
            errs = mixed_accuracy(m, s, g, u, args.accuracy_steps, args.order, tuple(p for _, p in forms))
            for (name, _), e in zip(forms, errs):
                out[name]["error_after_1000_steps" if args.accuracy_steps == 1000 else "error_after_steps"] = e
            print("# mixed accuracy done", file=sys.stderr, flush=True)
    if not args.no_rcm and args.order != "rcm":
        tr, tmr, _, _ = run_gpu(m, s, g, u, args.steps, args.warmup, S.PREC_FP64, "rcm")
        out["rcm_order"] = {"value": round(nc * evals / tr / 1e6, 3), "unit": "Mcell-updates/s",
                            "grad_ms_avg": round(tmr["grad_ms"] / max(1, tmr["grad_n"]), 4),
                            "flux_ms_avg": round(tmr["flux_ms"] / max(1, tmr["flux_n"]), 4),
                            "note": "same FP64 workload with the cells in reverse Cuthill-McKee order"}
        print(f"# rcm: {out['rcm_order']}", file=sys.stderr, flush=True)
    del m, u
    if not args.no_c1:
        out["c1"] = c1_bench(cpu=not args.no_cpu)
        print(f"# c1: {out['c1']['value']} Mcell-updates/s", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_ratio()
        if args.c4_vertices >= 0:
            log = lambda msg: print("# " + msg, file=sys.stderr, flush=True)  # noqa: E731
            out["c4"] = c4_bench(nvert=args.c4_vertices, log=log)
            g4 = out["c4"]["gradient"]
            print(f"# c4: primal {out['c4']['primal']['value']} Mcell-updates/s; gradient {g4['passed']}/{g4['checks']} "
                  f"within {g4['tol']:g}", file=sys.stderr, flush=True)
    if not args.no_implicit:
        try:
            out["implicit"] = implicit_bench()
        except Exception as e:  # reported, not fatal: the headline line must still print
            out["implicit"] = {"error": str(e)}
    if not args.no_ns:
        try:
            out["ns_16m"] = ns_bench(args.ns_n, args.ns_steps, adj_steps=args.ns_adj_steps, f32sum=not args.no_f32sum)
        except Exception as e:  # reported, not fatal
            out["ns_16m"] = {"error": str(e)}
        if not args.no_cpu and "error" not in out["ns_16m"]:
            out["ns_16m"]["cpu_baseline"] = ns_cpu_baseline()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
