"""Benchmark: residual Mcell-updates/s of the MI355X residual path.

Workload (one GPU): 3D Euler, 100^3 hexes split into 6,000,000 tets on a periodic
[0,10]^3 box (vertices jittered +-5% from seed 2, cells RCM-ordered on upload),
isentropic vortex (strength 5) advected at M=0.5, second-order MUSCL with LSQ
gradients and the Venkatakrishnan limiter (K=5), HLLC, classical RK4, CFL 0.8.
A step is one RK4 step = 4 residual evaluations (time step on the device
included); Mcell-updates/s = cells * 4 * steps / time.  `value` is the FP64
form; the FP32+FP64-accumulate form and its density error after the same steps
are reported beside it.

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


def byte_model(nc, nf, nb, dim=3, nv=5, mixed=False, rk4=True):
    """Algorithmic HBM bytes per launch of the two residual kernels (compulsory traffic, each
    array once per pass, int32 indices): gradient+limiter pass and flux+gather+RK-update pass."""
    ps = pg = px = 4 if mixed else 8
    E = 2 * nf - nb
    b_grad = nc * (nv * ps + 8 * dim + dim * (dim + 1) // 2 * pg + dim * nv * pg + 4) + 8 * nf + 4 * E
    b_flux = nc * (nv * ps + 8 + 4) + nf * (dim * px + 8) + 4 * E
    b_flux += nc * (dim * nv * pg + 8 * dim) + nf * dim * px
    b_rk = nc * nv * (22 + 0.75 * ps) if rk4 else 0.0
    return dict(grad=b_grad, flux=b_flux + b_rk, residual_only=b_grad + b_flux, total=b_grad + b_flux + b_rk)


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


PROFILE_CMDS = {
    "trace": "rocprofv3 --kernel-trace --stats -f csv -d {d}/trace -o c2 -- python3 bench.py --steps 20 --warmup 2 "
             "--no-cpu --no-implicit --no-ns",
    "fetch": "rocprofv3 --pmc FETCH_SIZE -f csv -d {d}/fetch -o c2 -- python3 bench.py --steps 2 --warmup 1 --no-cpu "
             "--no-adjoint --no-implicit --no-ns",
    "write": "rocprofv3 --pmc WRITE_SIZE -f csv -d {d}/write -o c2 -- python3 bench.py --steps 2 --warmup 1 --no-cpu "
             "--no-adjoint --no-implicit --no-ns",
    "ns": "rocprofv3 --kernel-trace --stats -f csv -d {d}/ns -o ns -- python3 bench.py --steps 1 --warmup 1 --no-cpu "
          "--no-mixed --no-adjoint --no-implicit --ns-steps 20 --ns-adj-steps 4",
    "imp": "rocprofv3 --kernel-trace --stats -f csv -d {d}/imp -o imp -- python3 bench.py --steps 1 --warmup 1 "
           "--no-cpu --no-mixed --no-adjoint --no-ns",
    "cfetch": "rocprofv3 --pmc FETCH_SIZE -f csv -d {d}/cfetch -o copy -- python3 bench.py --copy-calibration 1024",
    "cwrite": "rocprofv3 --pmc WRITE_SIZE -f csv -d {d}/cwrite -o copy -- python3 bench.py --copy-calibration 1024",
}


def _csv_rows(path):
    import csv
    with open(path) as f:
        return list(csv.DictReader(f))


def _kernel_key(name):
    """short kernel key from a demangled HIP kernel name: 'k_flux<3, double, ...>' -> ('flux', 'double')"""
    for key, tag in (("k_flux<", "flux"), ("k_gradient<", "grad"), ("k_dt_local<", "dt")):
        i = name.find(key)
        if i >= 0:
            args = name[i + len(key):].split(">")[0].split(",")
            return tag, args[1].strip() if len(args) > 1 else ""
    return None, None


def summarize_profiles(d, n=100):
    """Turn rocprofv3 CSVs under d (made by the PROFILE_CMDS commands) into the committed text
    summaries under profiles/: per-kernel durations, and HBM bytes per launch from FETCH_SIZE /
    WRITE_SIZE scaled by what the same counters read for a plain copy of known size."""
    nc = 6 * n ** 3
    nf = 2 * nc  # periodic Kuhn-split box: every face is shared by two tets
    bm = byte_model(nc, nf, 0, mixed=False)
    bmm = byte_model(nc, nf, 0, mixed=True)
    stats = _csv_rows(os.path.join(d, "trace", "c2_kernel_stats.csv"))

    def pmc(sub, prefix, counter):
        rows = _csv_rows(os.path.join(d, sub, f"{prefix}_counter_collection.csv"))
        out = {}
        for r in rows:
            if r["Counter_Name"] != counter:
                continue
            out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]) * 1024.0)
        return out

    cf, cw = pmc("cfetch", "copy", "FETCH_SIZE"), pmc("cwrite", "copy", "WRITE_SIZE")
    copy_bytes = 1024 << 20

    def copy_factor(tab):
        vals = [v for k, vs in tab.items() if "copy" in k.lower() for v in vs if v > copy_bytes / 4]
        return copy_bytes / (sum(vals) / len(vals)) if vals else None

    f_fac, w_fac = copy_factor(cf), copy_factor(cw)
    fe, wr = pmc("fetch", "c2", "FETCH_SIZE"), pmc("write", "c2", "WRITE_SIZE")
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    lines = ["# Residual kernels, 3D Euler MUSCL/LSQ/Venkatakrishnan + HLLC + RK4, 6,000,000 periodic tets (FP64 and FP32)",
             "", "The trace run also holds the adjoint workload's kernels (first-order HLLC residual and its adjoint on "
             "the 249,480-tet bump channel, 20 launches each).", "",
             "Commands (from the repository root, output directory D):", ""]
    lines += ["    " + c.format(d="D") for c in PROFILE_CMDS.values()]
    lines += ["    python3 bench.py --summarize D", "",
              "## Kernel durations (rocprofv3 --kernel-trace --stats)", "",
              "| kernel | calls | avg us | min us | max us | share % |", "|---|---|---|---|---|---|"]
    for r in stats:
        if "ty::" not in r["Name"]:
            continue
        lines.append(f"| {r['Name'].replace('|', '/')} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    lines += ["", "## HBM traffic per launch (FP64 and FP32 residual kernels of the 6M-tet workload)", "",
              f"Copy calibration (1 GiB device copy): FETCH_SIZE factor {f_fac:.3f}, WRITE_SIZE factor {w_fac:.3f} "
              "(bytes moved / counter bytes).", "",
              "| kernel | FETCH_SIZE GB (raw) | read GB (scaled) | WRITE_SIZE GB (raw) | written GB (scaled) | "
              "HBM GB/launch | algorithmic GB/launch | ratio |", "|---|---|---|---|---|---|---|---|"]
    traffic = {}
    for name, vals in fe.items():
        tag, prec = _kernel_key(name)
        if tag not in ("flux", "grad") or prec not in ("double", "float"):
            continue
        f_raw = sum(vals) / len(vals)
        wv = wr.get(name, [])
        w_raw = sum(wv) / len(wv) if wv else float("nan")
        tot = f_raw * f_fac + w_raw * w_fac
        alg = (bm if prec == "double" else bmm)[tag]
        key = tag if prec == "double" else f"{tag}_fp32"
        traffic[key] = dict(kernel=name, fetch_raw=f_raw, write_raw=w_raw, fetch_factor=f_fac, write_factor=w_fac,
                            bytes_per_launch=int(tot), algorithmic_bytes=int(alg), launches=len(vals))
        lines.append(f"| {name} | {f_raw / 1e9:.3f} | {f_raw * f_fac / 1e9:.3f} | {w_raw / 1e9:.3f} | "
                     f"{w_raw * w_fac / 1e9:.3f} | {tot / 1e9:.3f} | {alg / 1e9:.3f} | {tot / alg:.2f} |")
    with open(os.path.join(ROOT, "profiles", "residual_kernels_c2.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    ns_stats = os.path.join(d, "ns", "ns_kernel_stats.csv")
    if os.path.exists(ns_stats):
        nl = ["# NS kernels, 256^3 = 16,777,216 hexes (periodic x/z, slip walls y), first-order HLLC + cell-full "
              "viscous flux, LSQ gradients, RK4", "",
              "FP64 and mixed residual kernels (20 RK4 steps each) and the FP64 adjoint kernels of a 4-step "
              "checkpointed unsteady adjoint (viscous face reverse sweep, gather with the LSQ-gradient transpose, "
              "tail gather, FP64 stage residuals).  The run also holds one FP64 step of the 6M-tet workload.", "",
              "Command (from the repository root, output directory D):", "",
              "    " + PROFILE_CMDS["ns"].format(d="D"), "    python3 bench.py --summarize D", "",
              "| kernel | calls | avg us | min us | max us | share % |", "|---|---|---|---|---|---|"]
        for r in _csv_rows(ns_stats):
            if "ty::" not in r["Name"]:
                continue
            nl.append(f"| {r['Name'].replace('|', '/')} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                      f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
        with open(os.path.join(ROOT, "profiles", "ns_kernels_16m.md"), "w") as f:
            f.write("\n".join(nl) + "\n")
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
        with open(os.path.join(ROOT, "profiles", "implicit_kernels_c3.md"), "w") as f:
            f.write("\n".join(il) + "\n")
    for key, t in traffic.items():
        name = f"traffic_{key[:-5]}_c2_fp32.json" if key.endswith("_fp32") else f"traffic_{key}_c2_fp64.json"
        with open(os.path.join(ROOT, "profiles", name), "w") as f:
            json.dump(t, f, indent=1)
    print("\n".join(lines))


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


def ns_bench(n=256, steps=1000, warmup=2, adj_steps=200, snaps=20):
    """The 16M-cell unsteady NS workload: compressible Taylor-Green field on n^3 hexes (Re 1e4,
    Sutherland, Pr 0.72, 1% seeded noise), periodic in x and z with slip walls at y = 0 and
    y = 2 pi L (the field's v vanishes there), first-order HLLC + cell-full viscous flux with LSQ
    gradients, RK4 at CFL 0.5, one uploaded mesh for everything below.

    - throughput: FP64 and mixed (FP32 flux/gradient/stage state, FP64 U0 and RK accumulator)
      advance the same initial field by ``steps`` steps; density error of mixed vs FP64;
    - adjoint: J = pressure force normal to the y = 0 wall after ``adj_steps`` FP64 steps,
      dJ/dU0 by the device-resident checkpointed sweep (``snaps`` snapshots, FP64), checked against a central difference of two FP64 forward runs along a seeded
      direction."""
    import torch
    import cases as K
    from typhon_amd import adjoint as A
    from typhon_amd import scheme as S
    from typhon_amd.solver import Solver
    t0 = time.perf_counter()
    m, s, g, u = K.tgv_case(n, walls=True)
    tmesh = time.perf_counter() - t0
    t0 = time.perf_counter()
    solver = Solver(m, s, g)
    tup = time.perf_counter() - t0
    nc = m.ncell_int
    out = dict(workload=f"3D NS Taylor-Green, {n}^3 = {nc} hexes (periodic x/z, slip walls y), first-order HLLC + "
                        f"viscous (cell-full), LSQ gradients, RK4, CFL 0.5", cells=nc, steps=steps,
               host_setup_s=dict(mesh_build=round(tmesh, 1), upload=round(tup, 1)))
    states = {}
    dt = None
    for name, prec in (("fp64", S.PREC_FP64), ("mixed", S.PREC_MIXED)):
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
        states[name] = solver.get_state()[0]
        out[name] = dict(value=round(nc * 4 * steps / el / 1e6, 1), unit="Mcell-updates/s",
                         ms_per_step=round(el / steps * 1e3, 3),
                         grad_ms_avg=round(tm["grad_ms"] / max(1, tm["grad_n"]), 4),
                         flux_ms_avg=round(tm["flux_ms"] / max(1, tm["flux_n"]), 4))
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
    out["density_rel_err_l2"] = float(np.sqrt(np.mean((rmx - r64) ** 2)) / np.sqrt(np.mean(r64 ** 2)))
    out["density_rel_err_linf"] = float(np.abs(rmx - r64).max() / np.abs(r64).max())
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
    m, s, g, u = K.tgv_case(n)
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
    solver.close()
    inner = sum(st["inner_iters"] for st in stats)
    return dict(workload=f"3D NS Taylor-Green, {n}^3 = {m.ncell_int} periodic hexes, implicit backward Euler, "
                         f"dt = {dt_mult:g} x explicit", ms_per_implicit_step=round(el / steps * 1e3, 2),
                outer_iters=[st["outer_iters"] for st in stats], inner_iters=[st["inner_iters"] for st in stats],
                rel_res=[st["rel_res"] for st in stats], ms_per_arnoldi_step=round(el / max(1, inner) * 1e3, 3),
                mesh_build_s=round(tmesh, 1))


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
    ap.add_argument("--order", default="hilbert", choices=["hilbert", "rcm", "none"],
                    help="device cell order set at mesh upload")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (PMC byte-counter calibration)")
    ap.add_argument("--summarize", metavar="DIR", default=None,
                    help="write the profiles/ summaries from rocprofv3 output under DIR and exit")
    args = ap.parse_args()
    if args.summarize:
        summarize_profiles(args.summarize, args.n)
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
    m, s, g, u = K.vortex_case(args.n, L=10.0)
    print(f"# mesh: {m.ncell_int} tets, {m.nface} faces (reference form), built in {time.perf_counter() - t0:.1f} s",
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
        rho64, rhomx = st64[0], stmx[0]
        err_l2 = float(np.sqrt(np.mean((rhomx - rho64) ** 2)) / np.sqrt(np.mean(rho64 ** 2)))
        err_inf = float(np.abs(rhomx - rho64).max() / np.abs(rho64).max())
        fl = tmm["flux_ms"] / max(1, tmm["flux_n"])
        out["mixed"] = {"value": round(nc * evals / tmx / 1e6, 3), "unit": "Mcell-updates/s",
                        "dtype": "fp32 flux/gradient/stage state, fp64 U0/accumulator",
                        "speedup_vs_fp64": round(t64 / tmx, 3),
                        "density_rel_err_l2": err_l2, "density_rel_err_linf": err_inf,
                        "after_steps": args.steps + args.warmup,
                        "flux_ms_avg": round(fl, 4),
                        "grad_ms_avg": round(tmm["grad_ms"] / max(1, tmm["grad_n"]), 4),
                        "flux_achieved_GBs": round(bmm["flux"] / (fl * 1e-3) / 1e9, 1)}
    if not args.no_adjoint:
        out["adjoint"] = adjoint_ratio()
    if not args.no_implicit:
        try:
            out["implicit"] = implicit_bench()
        except Exception as e:  # reported, not fatal: the headline line must still print
            out["implicit"] = {"error": str(e)}
    if not args.no_ns:
        try:
            out["ns_16m"] = ns_bench(args.ns_n, args.ns_steps, adj_steps=args.ns_adj_steps)
        except Exception as e:  # reported, not fatal
            out["ns_16m"] = {"error": str(e)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
