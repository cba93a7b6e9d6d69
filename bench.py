"""Benchmark of the fused bond-based Velocity Verlet step on one MI355X.

Workload (default): 200 x 200 x 50 = 2,000,000-point lattice, spacing 1e-3 m, positions
jittered uniformly by +-0.05 spacing (SplitMix64, seed 0x5eed), horizon 3.015 spacing,
BBMaterial E = 70e9 Pa, rho = 2700, Gc = 100 J/m^2, +-0.6 m/s in y on the top and bottom
three point rows; FP64.  One step = one Velocity Verlet step over the whole body (one
fused launch).  Prints one JSON line with the throughput in bond updates/s, the roofline
of the step kernel and the CPU baseline (the oracle's OpenMP restatement of the
reference's threads path, on a bounded sample of the same lattice)."""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "peridynamics.jl_amd"))

METRIC = "bond updates/s + % HBM peak (VV); JFNK s to 1e-8 residual, learned vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_body(nx, ny, nz, dx=1e-3, jitter=0.05, seed=0x5EED):
    import peridynamics as pd
    pos, vol = pd.lattice_block(nx, ny, nz, dx, jitter_amplitude=jitter, seed=seed)
    body = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(body, horizon=3.015 * dx, rho=2700.0, E=70e9, Gc=100.0)
    y = pos[1]
    pd.point_set(body, "top", y > y.max() - 2.5 * dx)
    pd.point_set(body, "bottom", y < y.min() + 2.5 * dx)
    pd.velocity_bc(lambda t: 0.6, body, "top", "y")
    pd.velocity_bc(lambda t: -0.6, body, "bottom", "y")
    return body


def step_bytes(n_points, n_bonds, precision, layout):
    """Algorithmic HBM bytes of one fused bond-based step: every array the step needs is
    read or written once (neighbour data is re-used from LDS/cache, so a neighbour's
    record counts once, as its own point's).

    per point: read x|V (32), u (24), v_half (24), bond count (4), BC slot (4); write x|V
               (32), u (24), v_half (24)  -> 168 B, plus 8 B per 64-point slice pointer;
               the mixed LDS step also reads the reference record X|V (32) -> 200 B;
    per bond:  LDS-staged FP64: 16-bit bond word (2) + FP64 reference length (8);
               LDS-staged mixed: the 16-bit bond word only (2; L = |X_j - X_i|);
               global-gather layout: 32-bit bond word (4) + length (8 FP64 / 4 FP32);
    per neighbour-table entry (LDS layouts): its 4-byte point index.
    A bond word is written back only when the bond breaks (not counted)."""
    n_slices = (n_points + 63) // 64
    if layout["lds_mode"]:
        if precision == "fp64":
            return (10 * n_bonds + 168 * n_points + 8 * n_slices + 4 * layout["halo_entries"],
                    "10*N_b + 168*N_p + 8*N_slices + 4*N_halo")
        return (2 * n_bonds + 200 * n_points + 8 * n_slices + 4 * layout["halo_entries"],
                "2*N_b + 200*N_p + 8*N_slices + 4*N_halo")
    per_bond = 12 if precision == "fp64" else 8
    return (per_bond * n_bonds + 168 * n_points + 8 * n_slices,
            f"{per_bond}*N_b + 168*N_p + 8*N_slices")


def reference_layout_bytes(kind, n_points, n_bonds, precision="fp64"):
    """Compulsory HBM bytes of one step in the reference's own data layout (Vector{Bond}
    with Int32 neighbour + FP64 length, a Bool bond state, and the per-point Verlet
    vectors written every step), as a yardstick that does not depend on this library's
    layout: BB 13 B per bond + 216 B per point (mixed: FP32 length, 9 B per bond); OSB two
    bond passes + a reverse-bond index, 30 B per bond + 248 B per point (mixed: FP32
    lengths in both passes, 22 B per bond); CPD 30 B per bond + 576 B per point (mixed
    22 B per bond)."""
    if kind == "BB":
        return ((13 if precision == "fp64" else 9) * n_bonds + 216 * n_points,
                f"{13 if precision == 'fp64' else 9}*N_b + 216*N_p")
    if kind == "OSB":
        pb = 30 if precision == "fp64" else 22
        return pb * n_bonds + 248 * n_points, f"{pb}*N_b + 248*N_p"
    pb = 30 if precision == "fp64" else 22
    return pb * n_bonds + 576 * n_points, f"{pb}*N_b + 576*N_p"


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "bb_step_2M_hbm_traffic.md")


def copy_calibration(gib=1, reps=3):
    """A plain device-to-device copy of `gib` GiB, `reps` times (torch on the HIP device):
    the reference kernel for the HBM traffic counters (rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE of this copy against its known bytes calibrates the counters read for the
    step kernels) and the practical HBM ceiling (read + write bytes / copy time, HIP events
    on torch's stream, on which the copies run)."""
    import torch
    n = (gib << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)  # warm-up
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        dst.copy_(src)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    moved = 2 * n * 8
    return {"bytes_read_and_written": moved, "ms_per_copy": ms,
            "achieved_GBs": moved / (ms * 1e-3) / 1e9,
            "frac_of_8TBs": moved / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def measured_traffic(precision, n_points, n_bonds):
    """HBM bytes per launch of the step kernel for this workload from the committed counter
    summary (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by the copy-kernel
    calibration), or None if the summary does not cover this workload."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    tag = f"traffic {precision} points {n_points} bonds {n_bonds}:"
    for line in text.splitlines():
        line = line.strip().strip("`")
        if line.startswith(tag):
            return int(line[len(tag):].split()[0])
    return None


def jfnk_block(nx, ny, nz, dx=1e-3, jitter=0.05, seed=0x5EED, top=1e-4):
    """Quasi-static workload: bond-based block, bottom face fixed, top face displaced in y."""
    import peridynamics as pd
    pos, vol = pd.lattice_block(nx, ny, nz, dx, jitter_amplitude=jitter, seed=seed)
    body = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(body, horizon=3.015 * dx, rho=2700.0, E=70e9)
    y = pos[1]
    pd.point_set(body, "bottom", y < y.min() + 0.5 * dx)
    pd.point_set(body, "top", y > y.max() - 0.5 * dx)
    for d in "xyz":
        pd.displacement_bc(lambda p: 0.0, body, "bottom", d)
    pd.displacement_bc(lambda p: top, body, "top", "y")
    return body


def jfnk_warmup(precs=("jacobi", "none", "two_level", "learned")):
    """Two load steps of a 16x16x8 block with each preconditioner, untimed: the first solve of
    a process pays one-time costs (code objects, rocSOLVER handles, first allocations) that
    no timed solve should carry."""
    import peridynamics as pd
    for pre in precs:
        dh = pd.MI355XBodyDataHandler(jfnk_block(16, 16, 8))
        nk = pd.NewtonKrylov(steps=2, tol=1e-30, rel_tol=1e-8, maxiter=40, gmres_restart=30,
                             gmres_maxiter=200, precond=pre, learn_steps=1, train_steps=2)
        dh.solve_nk(nk)
        dh.close()


def jfnk_leg(shape, train_steps):
    """Wall time of the 10 load steps to a 1e-8 relative Newton residual with GMRES(30),
    200 Krylov iterations per Newton step: unpreconditioned (the reference's GMRES), block
    Jacobi ("plain"), the two-level preconditioner, and the learned one (trained on the error
    samples of load steps 1-3, which it solves with the two-level preconditioner; applied
    while it passes validation).  Displacements are compared with block Jacobi's.  A small
    untimed solve with each preconditioner comes first, so that no timed run carries the
    process's one-time start-up costs."""
    import numpy as np
    import peridynamics as pd
    res = {}
    u_ref = None
    jfnk_warmup()
    for pre in ("jacobi", "none", "two_level", "learned"):
        body = jfnk_block(*shape)
        dh = pd.MI355XBodyDataHandler(body)
        nk = pd.NewtonKrylov(steps=10, tol=1e-30, rel_tol=1e-8, maxiter=40, gmres_restart=30,
                             gmres_maxiter=200, precond=pre, learn_steps=3,
                             train_steps=train_steps)
        dh.ctx.synchronize()
        t0 = time.perf_counter()
        stats = dh.solve_nk(nk)
        dh.ctx.synchronize()
        wall = time.perf_counter() - t0
        u = dh.get("displacement")
        if pre == "jacobi":
            u_ref = u
        info = getattr(dh, "train_info", None)
        res[pre] = {"seconds": wall,
                    "seconds_steps_4_10": sum(s.seconds for s in stats[3:]),
                    "newton_iters": sum(s.newton_iters for s in stats),
                    "krylov_iters": sum(s.krylov_iters for s in stats),
                    "krylov_iters_steps_4_10": sum(s.krylov_iters for s in stats[3:]),
                    "fallbacks": sum(s.precond_fallbacks for s in stats),
                    "max_final_rel_residual": max(s.residual_final / s.residual_first for s in stats),
                    "displacement_rel_diff_vs_jacobi": float(np.linalg.norm(u - u_ref) /
                                                             np.linalg.norm(u_ref))}
        if info:
            res[pre]["train_seconds"] = info["seconds"]
            res[pre]["train_loss"] = info["loss"]
            res[pre]["validation_loss"] = info["val_loss"]
            res[pre]["network_applied_after_validation"] = info["network_applied"]
            res[pre]["train_pairs"] = list(info["pairs"])
        log(f"[bench] jfnk {pre}: {wall:.3f} s, krylov {res[pre]['krylov_iters']}")
        dh.close()
    plain = res["jacobi"]
    ref = res["none"]
    summary = {"plain_s": plain["seconds"], "two_level_s": res["two_level"]["seconds"],
               "learned_s": res["learned"]["seconds"],
               "learned_vs_plain": plain["seconds"] / res["learned"]["seconds"],
               "unpreconditioned_s": ref["seconds"],
               "learned_vs_unpreconditioned": ref["seconds"] / res["learned"]["seconds"]}
    return {"summary": summary, "workload": (f"bond-based block {shape[0]}x{shape[1]}x{shape[2]}, bottom fixed, top "
                         f"displaced 1e-4 m in y over 10 load steps, Newton rel tol 1e-8, GMRES(30)"),
            "points": shape[0] * shape[1] * shape[2], **res,
            "speedup_learned_vs_plain": plain["seconds"] / res["learned"]["seconds"],
            "speedup_learned_vs_unpreconditioned": ref["seconds"] / res["learned"]["seconds"],
            "speedup_learned_vs_plain_steps_4_10": (plain["seconds_steps_4_10"] /
                                                   res["learned"]["seconds_steps_4_10"]),
            "speedup_learned_vs_two_level_steps_4_10": (res["two_level"]["seconds_steps_4_10"] /
                                                       res["learned"]["seconds_steps_4_10"]),
            "speedup_two_level_vs_plain": plain["seconds"] / res["two_level"]["seconds"],
            "speedup_two_level_vs_plain_steps_4_10": (plain["seconds_steps_4_10"] /
                                                     res["two_level"]["seconds_steps_4_10"])}


def jfnk_profile_leg(shape, steps):
    """The first `steps` load steps of the quasi-static block with block Jacobi, then with the
    two-level preconditioner (the workload of the Krylov-iteration kernel trace): Krylov
    iterations and milliseconds per iteration of every load step."""
    import peridynamics as pd
    out = {}
    for pre in ("jacobi", "two_level"):
        dh = pd.MI355XBodyDataHandler(jfnk_block(*shape))
        nk = pd.NewtonKrylov(steps=10, tol=1e-30, rel_tol=1e-8, maxiter=40, gmres_restart=30,
                             gmres_maxiter=200, precond=pre)
        dh.init_newton_krylov(nk)
        rows = []
        for n in range(1, steps + 1):
            st = dh.nk_load_step(nk, n)
            rows.append({"step": n, "newton_iters": st.newton_iters, "krylov_iters": st.krylov_iters,
                         "ms": st.seconds * 1e3,
                         "ms_per_krylov_iteration": st.seconds * 1e3 / max(st.krylov_iters, 1)})
            log(f"[bench] jfnk profile {pre} step {n}: {rows[-1]['ms_per_krylov_iteration']:.4f} ms/it")
        out[pre] = rows
        dh.close()
    return out


def jfnk_learned_leg(shape, train_steps):
    """The 10 load steps of the quasi-static block with the learned preconditioner alone, after
    a small untimed warm-up solve (the workload of the training and learned-solve kernel
    trace): wall time, Krylov iterations and the training figures."""
    import peridynamics as pd
    jfnk_warmup(("learned",))
    dh = pd.MI355XBodyDataHandler(jfnk_block(*shape))
    nk = pd.NewtonKrylov(steps=10, tol=1e-30, rel_tol=1e-8, maxiter=40, gmres_restart=30,
                         gmres_maxiter=200, precond="learned", learn_steps=3, train_steps=train_steps)
    dh.ctx.synchronize()
    t0 = time.perf_counter()
    stats = dh.solve_nk(nk)
    dh.ctx.synchronize()
    out = {"seconds": time.perf_counter() - t0,
           "krylov_iters": sum(s.krylov_iters for s in stats),
           "krylov_iters_steps_4_10": sum(s.krylov_iters for s in stats[3:]),
           "train_seconds": dh.train_info["seconds"], "train_steps": train_steps,
           "validation_loss": dh.train_info["val_loss"]}
    dh.close()
    log(f"[bench] jfnk learned: {out['seconds']:.3f} s, krylov {out['krylov_iters']}, "
        f"training {out['train_seconds']:.3f} s")
    return out


def state_based_bytes(kind, n_points, n_bonds, layout, precision="fp64", zem="Silling"):
    """Algorithmic bytes of the two passes of an OSB or CPD step (one parameter set).  Per
    bond and pass the bond word and the FP64 length (LDS layout 2 + 8, global layout
    4 + 8; mixed OSB only the 2-byte word: the length comes from the reference positions).
    Per point: OSB, global layout, 236 B (both passes read x|V and the count, pass 1 writes
    c1, c2 and pass 2 reads them, plus the Verlet tail); OSB, LDS layout, 268 B (pass 1
    writes (c1 V, c2 V) and the own force part, 40 B, pass 2 reads the own part, 24 B);
    mixed OSB 300 B (X|V also read in both passes, (c1, c2) written and read back, no own
    part); mixed CPD 414 B (X|V in both passes and the nine numbers of the state
    Q' = P K^-1 - beta (F - I) written and read back); CPD 511 B (as mixed CPD with
    Q = P K^-1 - beta F, plus the own bond terms Q_i S1 + beta S2 and a_i = Q_i X_i + beta x_i,
    six numbers written by pass 1 and read by pass 2; 415 B in the global layout, which has
    no such split); CPD with ZEMWan 607 B (the state is Q = P K^-1 - C_1 F and the six
    entries of the symmetric C_1, 15 numbers, plus the six numbers of the own terms and
    a_i = Q_i X_i + C_1i x_i).  LDS layout: plus the 4-byte
    neighbour-table index per entry and pass."""
    if precision == "mixed" and layout["lds_mode"]:
        per_pt = 300 if kind == "OSB" else 414
        return (4 * n_bonds + per_pt * n_points + 8 * layout["halo_entries"],
                f"4*N_b + {per_pt}*N_p + 8*N_halo")
    if kind == "OSB":
        per_pt = 268 if layout["lds_mode"] else 236
    else:
        per_pt = (607 if zem == "Wan" else 511) if layout["lds_mode"] else 415
    if layout["lds_mode"]:
        return (20 * n_bonds + per_pt * n_points + 8 * layout["halo_entries"],
                f"20*N_b + {per_pt}*N_p + 8*N_halo")
    return 24 * n_bonds + per_pt * n_points, f"24*N_b + {per_pt}*N_p"


def torsion_bar(kind, shape, dx=1e-3, zem="Silling"):
    """Configuration 3's body: OSB or CMaterial (Saint Venant-Kirchhoff; ZEMSilling, or
    ZEMWan with zem="Wan") bar, torsion 10 rad/s on one end (three point layers), the other
    end held."""
    import peridynamics as pd
    pos, vol = pd.lattice_block(*shape, dx)
    cmat = (lambda: pd.CMaterial(zem=pd.ZEMWan())) if zem == "Wan" else pd.CMaterial
    mat = pd.OSBMaterial() if kind == "OSB" else cmat()
    body = pd.Body(mat, pos, vol)
    pd.material(body, horizon=3.015 * dx, rho=7850.0, E=210e9, nu=0.25)
    x = pos[0]
    pd.point_set(body, "end", x > x.max() - 2.5 * dx)
    pd.point_set(body, "fixed", x < x.min() + 2.5 * dx)
    pd.velocity_bc(lambda p, t: -10.0 * p[2], body, "end", "y")
    pd.velocity_bc(lambda p, t: 10.0 * p[1], body, "end", "z")
    for d in "xyz":
        pd.velocity_bc(lambda t: 0.0, body, "fixed", d)
    return body


def state_based_leg(kind, shape, steps, warmup=10, dx=1e-3, precision="fp64", zem="Silling"):
    """Configuration 3: OSB or CMaterial (Saint Venant-Kirchhoff; ZEMSilling, or ZEMWan
    with zem="Wan") bar, torsion 10 rad/s on one end (three point layers), the other end
    held."""
    import peridynamics as pd
    body = torsion_bar(kind, shape, dx, zem)
    dh = pd.MI355XBodyDataHandler(body, precision=precision)
    vv = pd.VelocityVerlet(steps=warmup + steps)
    dh.init_time_solver(vv)
    dh.initialize(vv)
    ctx = dh.ctx
    bcv = dh.bc_values(range(1, vv.n_steps + 1), vv.dt)
    ctx.vv_run(1, warmup, vv.dt, bcv[:warmup])
    ctx.synchronize()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    ctx.vv_run(1 + warmup, steps, vv.dt, bcv[warmup:])
    ctx.synchronize()
    el = time.perf_counter() - t0
    k1, _ = ctx.kernel_time("pass1")
    k2, _ = ctx.kernel_time("step")
    ctx.set_profiling(False)
    nb, npnt = dh.n_bonds, body.n_points
    byt, formula = state_based_bytes(kind, npnt, nb, ctx.layout(), precision, zem)
    rbyt, rformula = reference_layout_bytes(kind, npnt, nb, precision)
    dh.close()
    return {"workload": f"{kind} torsion bar {shape[0]}x{shape[1]}x{shape[2]}, 10 rad/s, "
                        f"{precision.upper()}" + (", ZEMWan" if zem == "Wan" else ""),
            "points": npnt, "bonds": nb, "steps": steps, "ms_per_step": el / steps * 1e3,
            "bond_updates_per_s": nb * steps / el, "pass1_ms": k1, "pass2_ms": k2,
            "bytes_formula": formula, "bytes_per_step": byt,
            "achieved_GBs": byt / ((k1 + k2) * 1e-3) / 1e9,
            "frac_hbm": byt / ((k1 + k2) * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "reference_layout_bytes_formula": rformula,
            "frac_hbm_reference_layout": rbyt / ((k1 + k2) * 1e-3) / 1e9 / HBM_PEAK_GBS}


def osb_2m_leg(shape, steps, warmup=10, precision="fp64"):
    """The 2 M-point Velocity Verlet workload with OSBMaterial (E 70 GPa, nu 1/4, Gc 100)."""
    import peridynamics as pd
    dx = 1e-3
    pos, vol = pd.lattice_block(*shape, dx, jitter_amplitude=0.05, seed=0x5EED)
    body = pd.Body(pd.OSBMaterial(), pos, vol)
    pd.material(body, horizon=3.015 * dx, rho=2700.0, E=70e9, nu=0.25, Gc=100.0)
    y = pos[1]
    pd.point_set(body, "top", y > y.max() - 2.5 * dx)
    pd.point_set(body, "bottom", y < y.min() + 2.5 * dx)
    pd.velocity_bc(lambda t: 0.6, body, "top", "y")
    pd.velocity_bc(lambda t: -0.6, body, "bottom", "y")
    dh = pd.MI355XBodyDataHandler(body, precision=precision)
    vv = pd.VelocityVerlet(steps=warmup + steps)
    dh.init_time_solver(vv)
    dh.initialize(vv)
    ctx = dh.ctx
    bcv = dh.bc_values(range(1, vv.n_steps + 1), vv.dt)
    ctx.vv_run(1, warmup, vv.dt, bcv[:warmup])
    ctx.synchronize()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    ctx.vv_run(1 + warmup, steps, vv.dt, bcv[warmup:])
    ctx.synchronize()
    el = time.perf_counter() - t0
    k1, _ = ctx.kernel_time("pass1")
    k2, _ = ctx.kernel_time("step")
    ctx.set_profiling(False)
    nb, npnt = dh.n_bonds, body.n_points
    byt, formula = state_based_bytes("OSB", npnt, nb, ctx.layout(), precision)
    rbyt, rformula = reference_layout_bytes("OSB", npnt, nb, precision)
    dh.close()
    return {"workload": f"OSB Velocity Verlet {shape[0]}x{shape[1]}x{shape[2]} jittered lattice, "
                        f"+-0.6 m/s y-BCs, {precision.upper()}",
            "points": npnt, "bonds": nb, "steps": steps, "ms_per_step": el / steps * 1e3,
            "bond_updates_per_s": nb * steps / el, "pass1_ms": k1, "pass2_ms": k2,
            "bytes_formula": formula, "bytes_per_step": byt,
            "achieved_GBs": byt / ((k1 + k2) * 1e-3) / 1e9,
            "frac_hbm": byt / ((k1 + k2) * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "reference_layout_bytes_formula": rformula,
            "frac_hbm_reference_layout": rbyt / ((k1 + k2) * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": measured_traffic("osb" if precision == "fp64" else "osb_mixed", npnt, nb)}


def mode_i_leg(npxy, nz, steps):
    """Configuration 5: pre-cracked plate in mode I, +-2 m/s on the edges, FP64 and mixed;
    time per step, and the mixed run's damage and crack tip against the FP64 run."""
    import numpy as np
    import peridynamics as pd
    res = {}
    perturbations = {"fp64_perturbed_2e-16": 2e-16, "fp64_perturbed_5e-14": 5e-14,
                     "fp64_perturbed_5e-12": 5e-12}
    for prec in ("fp64", "mixed", *perturbations):
        body, geo = pd.mode_i_plate(npxy=npxy, nz=nz, dx=2e-4, v=2.0)
        if prec in perturbations:
            # sensitivity baseline: the FP64 run from reference positions moved by +-eps m
            # (5e-14 m is the size of the FP32 rounding of a bond's displacement difference)
            rng = np.random.default_rng(1)
            body.position += perturbations[prec] * rng.choice([-1.0, 1.0], size=body.position.shape)
        dh = pd.MI355XBodyDataHandler(body, precision="mixed" if prec == "mixed" else "fp64")
        vv = pd.VelocityVerlet(steps=steps)
        dh.init_time_solver(vv)
        dh.initialize(vv)
        bcv = dh.bc_values(range(1, steps + 1), vv.dt)
        dh.ctx.synchronize()
        t0 = time.perf_counter()
        dh.ctx.vv_run(1, steps, vv.dt, bcv)
        dh.ctx.synchronize()
        el = time.perf_counter() - t0
        nact = dh.get("n_active_bonds")
        tip = pd.crack_tip(dh.get("position"), dh.get("damage"), xmin=geo["xmin"],
                           cluster_radius=geo["delta"])
        res[prec] = {"ms_per_step": el / steps * 1e3, "bond_updates_per_s": dh.n_bonds * steps / el,
                     "broken_bonds": int(dh.n_bonds - nact.sum()), "crack_tip": list(tip),
                     "_nact": nact, "_dmg": dh.get("damage")}
        npnt, nb = body.n_points, dh.n_bonds
        dh.close()
    a, b = res["fp64"], res["mixed"]
    out = {"workload": (f"pre-cracked plate {npxy}x{npxy}x{nz}, dx 2e-4 m, crack 0.25 l, "
                        f"+-2 m/s, {steps} steps"),
           "points": npnt, "bonds": nb,
           "fp64": {k: v for k, v in a.items() if not k.startswith("_")},
           "mixed": {k: v for k, v in b.items() if not k.startswith("_")},
           "mixed_speedup": a["ms_per_step"] / b["ms_per_step"],
           "mixed_broken_bond_count_difference": int(np.sum(np.abs(a["_nact"] - b["_nact"]))),
           "mixed_max_damage_difference": float(np.max(np.abs(a["_dmg"] - b["_dmg"]))),
           "mixed_crack_tip_difference_m": (abs(a["crack_tip"][0] - b["crack_tip"][0])
                                            if not np.isnan(a["crack_tip"][0]) else None),
           "fp64_sensitivity": {k: {"perturbation_m": v,
                                    "broken_bond_count_difference": int(np.sum(np.abs(a["_nact"] - res[k]["_nact"]))),
                                    "max_damage_difference": float(np.max(np.abs(a["_dmg"] - res[k]["_dmg"]))),
                                    "crack_tip": res[k]["crack_tip"]}
                                for k, v in perturbations.items()}}
    return out


def horizon_leg(shape, m, kind="BB", precision="fp64", steps=50, warmup=5, dx=1e-3):
    """The bond-based (or OSB) tension workload at a larger horizon, delta = m dx: the
    neighbour tables outgrow one LDS budget (1813 entries at m = 3.015; ~3000 at 4.015,
    ~4600 at 5.015), so the step stages them in tiles (BB FP64) or reads the entries past
    the budget from global memory (the other kernels)."""
    import peridynamics as pd
    pos, vol = pd.lattice_block(*shape, dx, jitter_amplitude=0.05, seed=0x5EED)
    body = pd.Body(pd.BBMaterial() if kind == "BB" else pd.OSBMaterial(), pos, vol)
    kw = dict(horizon=m * dx, rho=2700.0, E=70e9, Gc=100.0)
    if kind != "BB":
        kw["nu"] = 0.25
    pd.material(body, **kw)
    y = pos[1]
    pd.point_set(body, "top", y > y.max() - 2.5 * dx)
    pd.point_set(body, "bottom", y < y.min() + 2.5 * dx)
    pd.velocity_bc(lambda t: 0.6, body, "top", "y")
    pd.velocity_bc(lambda t: -0.6, body, "bottom", "y")
    dh = pd.MI355XBodyDataHandler(body, precision=precision)
    vv = pd.VelocityVerlet(steps=warmup + steps)
    dh.init_time_solver(vv)
    dh.initialize(vv)
    ctx = dh.ctx
    bcv = dh.bc_values(range(1, vv.n_steps + 1), vv.dt)
    ctx.vv_run(1, warmup, vv.dt, bcv[:warmup])
    ctx.synchronize()
    ctx.set_profiling(True)
    ctx.vv_run(1 + warmup, steps, vv.dt, bcv[warmup:])
    ctx.synchronize()
    lay = ctx.layout()
    nb, npnt = dh.n_bonds, body.n_points
    if kind == "BB":
        kt, _ = ctx.kernel_time("step")
        byt, formula = step_bytes(npnt, nb, precision, lay)
    else:
        k1, _ = ctx.kernel_time("pass1")
        k2, _ = ctx.kernel_time("step")
        kt = k1 + k2
        byt, formula = state_based_bytes(kind, npnt, nb, lay, precision)
    ctx.set_profiling(False)
    dh.close()
    return {"workload": f"{kind} Velocity Verlet {shape[0]}x{shape[1]}x{shape[2]} jittered lattice, "
                        f"horizon {m} dx, {precision.upper()}",
            "points": npnt, "bonds": nb, "halo_max": lay["halo_max"], "kernel_ms": kt,
            "bond_updates_per_s": nb / (kt * 1e-3), "bytes_formula": formula,
            "frac_hbm": byt / (kt * 1e-3) / 1e9 / HBM_PEAK_GBS}


def export_leg(shape, steps, freq, fields=("displacement", "damage")):
    """VTK export every `freq` steps (the reference's Job(...; freq, fields) output) on the
    bond-based workload: wall time per step without export, with the asynchronous export
    pipeline (device snapshot + pinned copy overlapping the next steps, files compressed by
    worker threads), and with a synchronous read-then-write export at the same steps."""
    import shutil
    import tempfile

    import numpy as np
    import peridynamics as pd
    from peridynamics import vtk

    body = make_body(*shape)
    res = {"shape": list(shape), "steps": steps, "freq": freq, "fields": list(fields)}
    root = tempfile.mkdtemp(prefix="pdx_export_")
    try:
        for mode in ("none", "async", "sync"):
            dh = pd.MI355XBodyDataHandler(body)
            vv = pd.VelocityVerlet(steps=steps)
            dh.init_time_solver(vv)
            dh.initialize(vv)
            dh.ctx.synchronize()
            out_dir = os.path.join(root, mode)
            t0 = time.perf_counter()
            if mode == "none":
                dh.solve(vv)
            elif mode == "async":
                ex = vtk.VTKExporter(dh, out_dir, fields)
                ex.capture(dh, 0, 0.0)
                dh.solve(vv, freq=freq, on_export=ex.capture)
                files = ex.close()
            else:
                base = vtk.vtk_filebase(out_dir)
                os.makedirs(os.path.dirname(base), exist_ok=True)
                parts = max(1, min(16, os.cpu_count() or 1, dh.n // 100_000 + 1))
                from concurrent.futures import ThreadPoolExecutor
                with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
                    def sync_export(dh_, n, t):
                        pos = dh_.get("position")
                        data = {f: dh_.get(f) for f in fields}
                        for fut in vtk.write_pvtu(base, n, pos, data, t, parts, None, pool):
                            fut.result()
                    sync_export(dh, 0, 0.0)
                    dh.solve(vv, freq=freq, on_export=sync_export)
            dh.ctx.synchronize()
            wall = time.perf_counter() - t0
            res[mode] = {"seconds": wall, "ms_per_step": wall / steps * 1e3}
            if mode == "async":
                nbytes = sum(os.path.getsize(os.path.join(dp, f))
                             for dp, _, fs in os.walk(out_dir) for f in fs)
                res["exports"] = len(files)
                res["MB_written"] = nbytes / 1e6
                r = vtk.read_vtk(files[-1])
                res["last_export_matches_final_state"] = bool(
                    np.array_equal(r["displacement"], dh.get("displacement")))
            dh.close()
            shutil.rmtree(out_dir, ignore_errors=True)
    finally:
        shutil.rmtree(root, ignore_errors=True)
    res["async_overhead_ms_per_export"] = (res["async"]["seconds"] - res["none"]["seconds"]) / \
        max(res.get("exports", 1), 1) * 1e3
    res["sync_overhead_ms_per_export"] = (res["sync"]["seconds"] - res["none"]["seconds"]) / \
        max(res.get("exports", 1), 1) * 1e3
    return res


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    return oracle


def _cpu_threads():
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    return min(threads, 16)


def _oracle_params(body, kind):
    import peridynamics as pd
    p = body.point_params[0]
    d = dict(delta=p.delta, rho=p.rho, E=p.E, nu=p.nu, G=p.G, K=p.K, bc=p.bc, epsc=p.epsc,
             Gc=p.Gc, mu=p.mu)
    d["lambda"] = p.lam
    if kind == "CPD":
        names = {pd.NeoHookePenalty: "NHP", pd.LinearElastic: "LE", pd.NeoHooke: "NH",
                 pd.SaintVenantKirchhoff: "SVK"}
        wan = isinstance(body.mat.zem, pd.ZEMWan)
        d.update(model=names[type(body.mat.constitutive_model)],
                 Cs=0.0 if wan else body.mat.zem.Cs, maxdmg=body.mat.maxdmg,
                 zem="Wan" if wan else "Silling")
    return d


def cpu_state_based(kind, shape, target_seconds=10.0, max_steps=500):
    """CPU baseline of the torsion-bar workload (the same body, material and velocity
    conditions as state_based_leg) on the oracle's OpenMP loops: per-point phase and the
    ordered gather in parallel, Velocity Verlet sweeps in numpy.  One untimed step, then
    steps until target_seconds."""
    import numpy as np
    oracle = _oracle()
    threads = _cpu_threads()
    body = torsion_bar(kind, shape)
    t0 = time.perf_counter()
    sim = oracle.OracleSim(body.xyz, body.volume, kind, _oracle_params(body, kind),
                           fail_permit=body.fail_permit.astype(np.uint8), nthreads=threads)
    setup = time.perf_counter() - t0
    dt = sim.stable_dt(0.7)
    end, fixed = body.point_sets["end"], body.point_sets["fixed"]
    bcs = [("sd", fixed, d, lambda t: 0.0) for d in range(3)]
    bcs += [("pdv", end, 1, lambda X, t: -10.0 * X[:, 2]),
            ("pdv", end, 2, lambda X, t: 10.0 * X[:, 1])]
    sim.force()
    sim.vv_run(1, dt, bcs, first_step=1)  # warm-up step
    steps, elapsed = 0, 0.0
    t0 = time.perf_counter()
    while elapsed < target_seconds and steps < max_steps:
        sim.vv_run(1, dt, bcs, first_step=2 + steps)
        steps += 1
        elapsed = time.perf_counter() - t0
    nb = int(sim.row_ptr[-1])
    return {"value": nb * steps / elapsed, "unit": "bond updates/s", "cores": threads,
            "kind": "port", "ms_per_step": elapsed / steps * 1e3,
            "sample": (f"{kind} torsion bar {shape[0]}x{shape[1]}x{shape[2]} ({body.n_points} "
                       f"points, {nb} bonds), FP64, {steps} Velocity Verlet steps in "
                       f"{elapsed:.1f} s after one untimed step (setup {setup:.1f} s)")}


def cpu_jfnk(shape, target_seconds=10.0):
    """CPU baseline of one Krylov iteration of the quasi-static block (the same body and
    conditions as jfnk_leg): the oracle's central-difference J v (two OpenMP bond-based
    force evaluations) and its restarted GMRES(30) with modified Gram-Schmidt in numpy,
    unpreconditioned as the reference's, on the first Newton iteration of load step 1;
    whole cycles until target_seconds."""
    import numpy as np
    oracle = _oracle()
    threads = _cpu_threads()
    body = jfnk_block(*shape)
    sim = oracle.OracleSim(body.xyz, body.volume, "BB", _oracle_params(body, "BB"),
                           fail_permit=body.fail_permit.astype(np.uint8), nthreads=threads)
    bcs = ([(body.point_sets["bottom"], d, lambda p: 0.0) for d in range(3)] +
           [(body.point_sets["top"], 1, lambda p: 1e-5)])  # load step 1 of 10 (1e-4 m)
    cons = sim.constrained_dofs(bcs)
    for pts, dim, f in bcs:
        sim.u[np.asarray(pts, np.int64), dim] = f(None)
    sim.x[...] = sim.X + sim.u
    sim.force()
    r = sim.residual(cons)
    n_jv = [0]

    def A(v):
        n_jv[0] += 1
        return sim.jv(v, cons)

    # whole GMRES(30) cycles (30 products and the restart residual's) until target_seconds
    el, cycles = 0.0, 0
    t0 = time.perf_counter()
    while el < target_seconds and cycles < 20:
        oracle.gmres(A, -r, maxiter=30, reltol=1e-30, abstol=0.0, restart=30)
        cycles += 1
        el = time.perf_counter() - t0
    return {"value": n_jv[0] / el, "unit": "Krylov iterations/s", "cores": threads,
            "kind": "port", "ms_per_iteration": el / n_jv[0] * 1e3,
            "sample": (f"bond-based block {shape[0]}x{shape[1]}x{shape[2]} ({body.n_points} "
                       f"points, {int(sim.row_ptr[-1])} bonds), {n_jv[0]} unpreconditioned "
                       f"GMRES(30) iterations (J v products) in {el:.1f} s")}


def cpu_baseline(sample_shape, target_seconds=12.0, max_steps=5000):
    import numpy as np
    oracle = _oracle()
    threads = _cpu_threads()
    body = make_body(*sample_shape)
    p = body.point_params[0]
    params = dict(delta=p.delta, rho=p.rho, bc=p.bc, epsc=p.epsc, G=p.G, K=p.K)
    sim = oracle.OracleSim(body.xyz, body.volume, "BB", params,
                           fail_permit=body.fail_permit.astype(np.uint8), nthreads=threads)
    dt = sim.stable_dt(0.7)
    bcs = [("sd", body.point_sets["top"], 1, lambda t: 0.6),
           ("sd", body.point_sets["bottom"], 1, lambda t: -0.6)]
    sim.force()
    sim.vv_run(1, dt, bcs, first_step=1)  # warm-up step
    steps, elapsed = 0, 0.0
    t0 = time.perf_counter()
    while elapsed < target_seconds and steps < max_steps:
        sim.vv_run(1, dt, bcs, first_step=2 + steps)
        steps += 1
        elapsed = time.perf_counter() - t0
    nb = int(sim.row_ptr[-1])
    return {"value": nb * steps / elapsed, "unit": "bond updates/s", "cores": threads,
            "kind": "port",
            "sample": (f"BB Velocity Verlet FP64, {sample_shape[0]}x{sample_shape[1]}x"
                       f"{sample_shape[2]} block of the same jittered lattice "
                       f"({body.n_points} points, {nb} bonds), {steps} steps in {elapsed:.1f} s")}


def summary_line(out):
    """One short line with the headline figures of every leg that ran (the JSON line is
    long; a log that keeps only its tail still keeps this)."""
    parts = [f"BB fp64 {out['ms_per_step']:.4f} ms/step {out['value']:.3e} bond/s "
             f"{100 * out['roofline']['frac']:.1f}% HBM"]
    m = out.get("mixed")
    if m:
        parts.append(f"BB mixed {m['ms_per_step']:.4f} ms/step x{m['speedup_vs_fp64_step']:.2f}")
    sb = out.get("state_based") or {}
    for k in ("OSB_2M", "OSB_2M_mixed", "OSB", "CPD", "CPD_mixed", "CPD_ZEMWan"):
        if k in sb:
            parts.append(f"{k} {sb[k]['ms_per_step']:.3f} ms")
    j = out.get("jfnk")
    if j:
        parts.append("JFNK " + ", ".join(f"{k} {v:.3g}" for k, v in j.get("summary", {}).items()))
    c = out.get("cpu_baseline")
    if c:
        parts.append(f"CPU {c['value']:.3e} bond/s on {c['cores']} cores")
    return "[bench] summary: " + "; ".join(parts)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--shape", type=str, default="200,200,50")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-sample", type=str, default="100,100,25")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--no-jfnk", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the state-based (512k-point bar) and mode-I plate legs")
    ap.add_argument("--jfnk-shape", type=str, default="100,100,10")
    ap.add_argument("--jfnk-train-steps", type=int, default=40)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB device copy that calibrates the HBM counters")
    ap.add_argument("--jfnk-profile", type=int, default=0, metavar="STEPS",
                    help="only run the first STEPS load steps of the JFNK block with block "
                         "Jacobi and two-level (the Krylov-iteration profile workload)")
    ap.add_argument("--jfnk-learned", action="store_true",
                    help="only run the JFNK block with the learned preconditioner (the training "
                         "and learned-solve profile workload)")
    args = ap.parse_args()
    if args.copy_calibration:
        print(json.dumps({"copy_calibration": copy_calibration()}), flush=True)
        return
    if args.jfnk_learned:
        shape = tuple(int(v) for v in args.jfnk_shape.split(","))
        print(json.dumps({"jfnk_learned": jfnk_learned_leg(shape, args.jfnk_train_steps)}), flush=True)
        return
    if args.jfnk_profile:
        shape = tuple(int(v) for v in args.jfnk_shape.split(","))
        print(json.dumps({"jfnk_profile": jfnk_profile_leg(shape, args.jfnk_profile)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of this path is a design note, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import peridynamics as pd

    shape = tuple(int(s) for s in args.shape.split(","))
    log(f"[bench] building body {shape}")
    t0 = time.perf_counter()
    body = make_body(*shape)
    dh = pd.MI355XBodyDataHandler(body, precision=args.precision)
    vv = pd.VelocityVerlet(steps=args.warmup + args.steps + args.profile_steps)
    dh.init_time_solver(vv)
    dh.initialize(vv)
    n_points, n_bonds = body.n_points, dh.n_bonds
    log(f"[bench] {n_points} points, {n_bonds} bonds, setup {time.perf_counter() - t0:.1f} s")

    ctx = dh.ctx
    dt = vv.dt
    step0 = 1
    bcv = dh.bc_values(range(1, vv.n_steps + 1), dt)

    def run(first, k):
        ctx.vv_run(first, k, dt, bcv[first - 1:first - 1 + k])

    run(step0, args.warmup)
    step0 += args.warmup
    ctx.synchronize()  # device-wide (hipDeviceSynchronize): one GPU, one rank, no barrier
    t0 = time.perf_counter()
    run(step0, args.steps)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    step0 += args.steps
    ms_per_step = elapsed / args.steps * 1e3
    value = n_bonds * args.steps / elapsed
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.4f} ms/step, {value:.4e} bond updates/s")

    # per-launch duration of the step kernel with HIP events on the library's stream
    ctx.set_profiling(True)
    run(step0, args.profile_steps)
    step0 += args.profile_steps
    kms, kcount = ctx.kernel_time("step")
    ctx.set_profiling(False)
    layout = ctx.layout()
    bytes_step, bytes_formula = step_bytes(n_points, n_bonds, args.precision, layout)
    achieved = bytes_step / (kms * 1e-3) / 1e9
    log(f"[bench] step kernel {kms:.4f} ms avg over {kcount} launches, {achieved:.1f} GB/s")
    u = dh.get("displacement")
    assert np.all(np.isfinite(u))
    nact = dh.get("n_active_bonds")
    total_steps = args.warmup + args.steps + args.profile_steps

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bond updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-force/f64-integration",
        "data": "synthetic",
        "config": {"workload": (f"bond-based Velocity Verlet, {shape[0]}x{shape[1]}x{shape[2]} "
                                f"jittered lattice, horizon 3.015 dx, +-0.6 m/s y-BCs"),
                   "points": n_points, "bonds": n_bonds, "precision": args.precision,
                   "dt": dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.precision, n_points, n_bonds),
                     "kernel": "k_bb_lds" if layout["lds_mode"] else "k_bb_step",
                     "kernel_ms": kms, "bytes_formula": bytes_formula,
                     "bytes_per_launch": bytes_step, "halo_entries": layout["halo_entries"],
                     "halo_max": layout.get("halo_max"), "bond_slots": layout["bond_slots"],
                     "reference_layout_bytes_formula": reference_layout_bytes(
                         "BB", n_points, n_bonds, args.precision)[1],
                     "frac_reference_layout": reference_layout_bytes(
                         "BB", n_points, n_bonds, args.precision)[0] / (kms * 1e-3) / 1e9
                     / HBM_PEAK_GBS},
    }
    dh.close()

    if not args.no_mixed and args.precision == "fp64":
        # the same run in mixed precision: throughput, step-kernel time, and its final state
        # against the FP64 run (displacement error, difference of the broken-bond counts)
        log("[bench] mixed-precision leg")
        dhm = pd.MI355XBodyDataHandler(body, precision="mixed")
        vvm = pd.VelocityVerlet(steps=total_steps)
        dhm.init_time_solver(vvm)
        dhm.initialize(vvm)
        cm = dhm.ctx
        assert vvm.dt == dt
        bcm = dhm.bc_values(range(1, vvm.n_steps + 1), dt)
        cm.vv_run(1, args.warmup, dt, bcm[:args.warmup])
        cm.synchronize()
        t1 = time.perf_counter()
        cm.vv_run(1 + args.warmup, args.steps, dt, bcm[args.warmup:args.warmup + args.steps])
        cm.synchronize()
        em = time.perf_counter() - t1
        cm.set_profiling(True)
        s0 = 1 + args.warmup + args.steps
        cm.vv_run(s0, args.profile_steps, dt, bcm[s0 - 1:])
        kmm, _ = cm.kernel_time("step")
        cm.set_profiling(False)
        um = dhm.get("displacement")
        nm = dhm.get("n_active_bonds")
        mbytes, mformula = step_bytes(n_points, n_bonds, "mixed", cm.layout())
        out["mixed"] = {"value": n_bonds * args.steps / em, "ms_per_step": em / args.steps * 1e3,
                        "kernel_ms": kmm, "speedup_vs_fp64_step": ms_per_step / (em / args.steps * 1e3),
                        "speedup_vs_fp64_kernel": kms / kmm if kmm else None,
                        "achieved_GBs": mbytes / (kmm * 1e-3) / 1e9,
                        "frac_hbm": mbytes / (kmm * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "frac_hbm_reference_layout": reference_layout_bytes(
                            "BB", n_points, n_bonds, "mixed")[0] / (kmm * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "bytes_formula": mformula,
                        "final_displacement_rel_error_vs_fp64": float(
                            np.linalg.norm(um - u) / np.linalg.norm(u)),
                        "broken_bond_count_difference_vs_fp64": int(np.sum(np.abs(nm - nact))),
                        "broken_bonds_fp64": int(n_bonds - np.sum(nact)),
                        "broken_bonds_mixed": int(n_bonds - np.sum(nm)),
                        "steps_compared": total_steps}
        dhm.close()

    if not args.no_extra:
        log("[bench] state-based legs")
        out["state_based"] = {k: state_based_leg(k, (160, 80, 40), 100) for k in ("OSB", "CPD")}
        out["state_based"]["OSB_2M"] = osb_2m_leg(shape, 100)
        sb = out["state_based"]
        sb["OSB_mixed"] = state_based_leg("OSB", (160, 80, 40), 100, precision="mixed")
        sb["OSB_2M_mixed"] = osb_2m_leg(shape, 100, precision="mixed")
        sb["CPD_mixed"] = state_based_leg("CPD", (160, 80, 40), 100, precision="mixed")
        sb["CPD_ZEMWan"] = state_based_leg("CPD", (160, 80, 40), 100, zem="Wan")
        for a, m in (("OSB", "OSB_mixed"), ("OSB_2M", "OSB_2M_mixed"), ("CPD", "CPD_mixed")):
            sb[m]["speedup_vs_fp64_step"] = sb[a]["ms_per_step"] / sb[m]["ms_per_step"]
            sb[m]["speedup_vs_fp64_kernels"] = ((sb[a]["pass1_ms"] + sb[a]["pass2_ms"]) /
                                                (sb[m]["pass1_ms"] + sb[m]["pass2_ms"]))
        log("[bench] mode-I plate leg")
        out["mode_i"] = mode_i_leg(500, 4, 2000)
        log("[bench] export leg")
        out["export"] = export_leg(shape, 1000, 200)
        log("[bench] larger-horizon legs")
        out["horizons"] = [horizon_leg((120, 120, 40), m, kind)
                           for m in (3.015, 4.015, 5.015) for kind in ("BB", "OSB")]

    if not args.no_jfnk:
        log("[bench] JFNK leg")
        out["jfnk"] = jfnk_leg(tuple(int(v) for v in args.jfnk_shape.split(",")),
                               args.jfnk_train_steps)

    if not args.no_cpu:
        # numpy's BLAS threads spin between calls and steal the cores of the oracle's OpenMP
        # loops: the CPU legs run BLAS single-threaded (their vector work is memory-bound)
        from threadpoolctl import threadpool_limits
        blas1 = threadpool_limits(1, user_api="blas")
        log("[bench] CPU baseline")
        out["cpu_baseline"] = cpu_baseline(tuple(int(s) for s in args.cpu_sample.split(",")))
        if not args.no_extra:
            for k in ("OSB", "CPD"):
                log(f"[bench] CPU baseline, {k} torsion bar")
                c = cpu_state_based(k, (160, 80, 40))
                out["state_based"][k]["cpu_baseline"] = c
                c["gpu_speedup"] = out["state_based"][k]["bond_updates_per_s"] / c["value"]
        if not args.no_jfnk:
            log("[bench] CPU baseline, JFNK")
            j = out["jfnk"]
            c = cpu_jfnk(tuple(int(v) for v in args.jfnk_shape.split(",")))
            gpu_it = j["jacobi"]["seconds"] / j["jacobi"]["krylov_iters"]
            c["gpu_ms_per_iteration_block_jacobi"] = gpu_it * 1e3
            c["gpu_speedup_per_iteration"] = c["ms_per_iteration"] / (gpu_it * 1e3)
            j["cpu_baseline"] = c
        blas1.unregister()
    log(summary_line(out))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
