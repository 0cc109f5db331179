#!/usr/bin/env python
"""Benchmark of the hom3 finite-volume hot path on one MI355X.

Headline workload: 3D compressible Navier-Stokes Taylor-Green vortex, Re
1600, Mach 0.1, Pr 0.71, uniform periodic 256^3 grid on [0, 2pi]^3, AUSM +
three-point viscous flux, SSP-RK3.  A step is one SSP-RK3 time step (3
stages) over the whole grid; inputs are resident in HBM before the timed
region.

Prints ONE JSON line: value = FP64 Mcell-stage updates/s (whole job), the
FP32-state rate, the roofline of the dominant (stage) kernel measured live
with HIP events on the library's stream, the CPU oracle timed on a bounded
sample on the host, and (unless --no-extra) the other workloads: the 2D
isentropic vortex (MUSCL), the refined 3D blast, the batch of 512 uneven
2D instances, and the conjugate-heat-transfer adjoint with its
adjoint/primal cost ratio.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# compulsory HBM bytes per cell-stage, SSP-RK3, SoA, nv = 5 (3D):
#   FP64: (8/3) * nv * 8 ; FP32 stage state + FP64 U^n: 16 * nv
BYTES_FP64 = (8.0 / 3.0) * 5 * 8
BYTES_FP32 = 16.0 * 5


def tgv_solver(H, level, precision):
    from hom3_amd import cases
    return cases.tgv(H, level, precision)


def time_gpu(H, level, precision, steps, warmup):
    g, s = tgv_solver(H, level, precision)
    if s.kernel() != "brick":
        raise RuntimeError("TGV benchmark expects the brick kernel")
    n = s.first_gc
    s.solve(warmup)
    s.ctx.synchronize()
    t0 = time.perf_counter()
    s.solve(steps)
    s.ctx.synchronize()
    t1 = time.perf_counter()
    # dominant kernel: the brick stage kernel, timed with HIP events on the
    # library stream (mean over all stage launches of a few steps)
    prof_steps = max(2, min(steps, 5))
    total_ms, stage_ms, nl = s.profile_solve(prof_steps)
    bad = s.check_state()
    info = dict(n=n, seconds=t1 - t0, stage_ms=stage_ms, launches=nl, bad=bad,
                mcs=n * 3 * steps / (t1 - t0) / 1e6)
    s.close()
    return info


def _cpu_tgv(Om, level, steps):
    """The CPU oracle front-end Om on the same TGV at 2^level cells per
    side: (Mcell-stage/s, seconds, cells, final state)."""
    og = Om.Grid(3, [0.0] * 3, [2 * math.pi] * 3).min_level(level)
    os_ = Om.Solver(og, Om.phys(physics=Om.PHYS_CNS, nd=3, ti=Om.TI_SSP_RK3, cfl=0.5, cfl_viscous=0.5,
                                Minfty=0.1, ReInfty=1600.0, Pr0=0.71))
    os_.append_cube_bcs([Om.BC_PERIODIC] * 6)
    qt = os_.quantities()
    os_.set_ic(Om.IC_TGV, [qt["u_inf"], qt["rho_inf"], qt["p_inf"]])
    os_.initialize()
    secs = os_.time_solve(steps)
    n = os_.first_gc
    return n * 3 * steps / secs / 1e6, secs, n, os_.get("lhs")[:, :n]


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return dict(nproc=os.cpu_count(), available=avail, model=model)


def fast_oracle():
    """The oracle's timing build: -O3 -march=native -ffast-math -fopenmp built
    here for this host (the reference's Release flags, CMakeLists.txt:50-54),
    or the shipped -march=x86-64-v3 build when that fails."""
    import subprocess
    import tempfile
    import oracle as O
    out = tempfile.mkdtemp(prefix="hom3_oracle_native_")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"OUT={out}"], check=True,
                       timeout=180, capture_output=True)
        return O.fast(os.path.join(out, "liboracle_native.so")), "-O3 -march=native -ffast-math -fopenmp"
    except Exception:
        return O.fast(), "-O3 -march=x86-64-v3 -ffast-math -fopenmp"


def cpu_baseline(level=6, steps=2, level_mt=7):
    """The CPU oracle (C++ restatement of the reference's path) on bounded
    samples of the same TGV: the strict parity build on one core, the
    Release-flags timing build on one core, and the timing build on every
    core the host gives this process (OpenMP over the cells)."""
    import oracle as O
    info = host_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["available"]
    strict, s_secs, n, q_strict = _cpu_tgv(O, level, steps)
    Of, flags = fast_oracle()
    Of.set_threads(1)
    fast1, f_secs, _, q_fast = _cpu_tgv(Of, level, steps)
    used = Of.set_threads(threads)
    fastn, m_secs, n_mt, _ = _cpu_tgv(Of, level_mt, steps)
    d = q_fast - q_strict
    rel = float(np.sqrt(np.sum(d * d)) / np.sqrt(np.sum(q_strict * q_strict)))
    variants = [
        dict(build="strict (-O2 -ffp-contract=off, the parity checker)", cores=1, value=round(strict, 4),
             sample=f"{2 ** level}^3, {steps} steps, {s_secs:.2f} s"),
        dict(build=flags, cores=1, value=round(fast1, 4), sample=f"{2 ** level}^3, {steps} steps, {f_secs:.2f} s"),
        dict(build=flags, cores=used, value=round(fastn, 4),
             sample=f"{2 ** level_mt}^3, {steps} steps, {m_secs:.2f} s"),
    ]
    return dict(value=round(fastn, 4), unit="Mcell-stage/s", cores=used, kind="port",
                sample=f"oracle timing build ({flags}, OpenMP over the cells, {used} threads) on the same TGV at "
                       f"{2 ** level_mt}^3, {steps} SSP-RK3 steps, {m_secs:.2f} s",
                host=info, variants=variants, fast_vs_strict_rel_l2=rel)


def profile_run(H, level):
    """Short run for rocprofv3: ten FP64 steps and four FP32 steps of the TGV
    (the first steps warm the clocks up; summaries use the later ones), then
    a plain 1 GiB device-to-device copy whose FETCH/WRITE sizes calibrate the
    counters."""
    g, s = tgv_solver(H, level, "fp64")
    s.solve(10)
    s.ctx.synchronize()
    s.close()
    g, s = tgv_solver(H, level, "fp32")
    s.solve(4)
    s.ctx.synchronize()
    s.close()
    ms = copy_calibration(H, 1 << 30)
    print(f"[bench] profile run done (copy of 1 GiB in {ms:.3f} ms)", file=sys.stderr, flush=True)


def profile_batch_run(H, n_inst=512, seed=2024):
    """Short run for rocprofv3 on the batch workload: the bench's batch with
    the launch shape its autotuner picks (block 512, tile 2048, grid 16384),
    3 FP64 and 3 FP32-state steps, then a 1 GiB device copy."""
    from hom3_amd.batch import build_instance
    solvers = [build_instance(H, p) for p in H.four_quadrant_instances(seed, n_inst)]
    for tag in ("fp64", "fp32"):
        B = H.Batch(solvers, tag)
        B.set_config(512, 2048, 16384)
        B.solve(3)
        B.close()
    for s in solvers:
        s.close()
    ms = copy_calibration(H, 1 << 30)
    print(f"[bench] batch profile run done (copy of 1 GiB in {ms:.3f} ms)", file=sys.stderr, flush=True)


def profile_vortex_run(H, level=12, steps=4):
    """Short run for rocprofv3 on the 2D vortex at 4096^2 (the 2D MUSCL brick
    kernel on a grid that fills the GPU): 4 FP64 and 4 FP32-state steps, then
    a 1 GiB device copy."""
    from hom3_amd import cases
    for prec in ("fp64", "fp32"):
        g, s = cases.vortex(H, level, precision=prec)
        s.solve(steps)
        bad = s.check_state()
        s.close()
        if bad:
            raise RuntimeError(f"vortex {prec}: {bad} non-physical cells")
    ms = copy_calibration(H, 1 << 30)
    print(f"[bench] vortex profile run done (copy of 1 GiB in {ms:.3f} ms)", file=sys.stderr, flush=True)


def copy_calibration(H, nbytes):
    import ctypes as C
    from hom3_amd import _lib
    ms = C.c_float()
    _lib.check(_lib.lib().hom3g_copy_calibration(H.default_context().h, nbytes, C.byref(ms)))
    return ms.value


def state_error(q, qref, nd):
    """Per conserved variable: relative L2 error (momentum components
    normalised by the norm of the whole momentum field) and relative max
    error (max |q - qref| / max |qref|) of q against qref."""
    q = np.asarray(q, np.float64)
    qref = np.asarray(qref, np.float64)
    diff = q - qref
    l2 = np.sqrt(np.sum(diff ** 2, axis=1))
    den = np.sqrt(np.sum(qref ** 2, axis=1))
    mx = np.max(np.abs(diff), axis=1)
    mden = np.max(np.abs(qref), axis=1)
    if nd:
        den[:nd] = np.sqrt(np.sum(qref[:nd] ** 2))
        mden[:nd] = np.max(np.abs(qref[:nd]))
    return [float(x) for x in l2 / den], [float(x) for x in mx / mden]


def tgv_fp32_error(H, level, steps):
    """FP32-state TGV against the FP64 path after the same number of steps
    (each with its own CFL dt; the dts agree to the FP32 state error)."""
    out = {}
    for prec in ("fp64", "fp32"):
        g, s = tgv_solver(H, level, prec)
        s.solve(steps)
        out[prec] = (s.lhs[:, :s.first_gc].copy(), s.time())
        s.close()
    l2, mx = state_error(out["fp32"][0], out["fp64"][0], 3)
    return dict(steps=steps, vars=["rho_u0", "rho_u1", "rho_u2", "rho", "rho_E"], rel_l2=l2, rel_max=mx,
                time_fp64=out["fp64"][1], time_fp32=out["fp32"][1])


def tgv_against_oracle(H, level=5, steps=20):
    """FP64 and FP32-state TGV against the FP64 oracle (the strict parity
    build) at 2^level cells per side after `steps` steps, from the same
    initial state: relative L2 and max error per conserved variable."""
    import oracle as O
    _, _, n, qo = _cpu_tgv(O, level, steps)
    out = dict(level=level, cells=n, steps=steps, vars=["rho_u0", "rho_u1", "rho_u2", "rho", "rho_E"])
    for prec in ("fp64", "fp32"):
        g, s = tgv_solver(H, level, prec)
        s.solve(steps)
        l2, mx = state_error(s.lhs[:, :s.first_gc], qo, 3)
        out[prec] = dict(rel_l2=l2, rel_max=mx)
        s.close()
    return out


def timed_solve(s, steps, warmup):
    s.solve(warmup)
    s.ctx.synchronize()
    t0 = time.perf_counter()
    s.solve(steps)
    s.ctx.synchronize()
    return time.perf_counter() - t0


def bench_vortex(H, steps):
    """2D Euler isentropic vortex 256^2, periodic, AUSM + MUSCL, SSP-RK3."""
    from hom3_amd import cases
    g, s = cases.vortex(H, 8)
    n = s.first_gc
    sec = timed_solve(s, steps, 3)
    out = dict(workload="2D Euler isentropic vortex 256^2 periodic, AUSM + MUSCL(minmod), SSP-RK3, CFL 0.5",
               cells=n, steps=steps, kernel=s.kernel(), mcs=round(n * 3 * steps / sec / 1e6, 1),
               ms_per_step=round(sec / steps * 1e3, 4), bad_cells=s.check_state(),
               note="working set of a few MB stays in the caches: launch and latency bound, not HBM")
    s.close()
    return out


def bench_vortex_large(H, steps, level=12):
    """The same 2D vortex at 4096^2: the 2D brick kernel when the grid fills
    the GPU (256^2 gives one 256-cell brick per CU), FP64 and FP32 state."""
    from hom3_amd import cases
    side = 2 ** level
    out = dict(workload=f"2D Euler isentropic vortex {side}^2 periodic, AUSM + MUSCL(minmod), SSP-RK3, CFL 0.5",
               steps=steps)
    for prec in ("fp64", "fp32"):
        g, s = cases.vortex(H, level, precision=prec)
        n = s.first_gc
        sec = timed_solve(s, steps, 3)
        out["cells"] = n
        out[prec] = dict(mcs=round(n * 3 * steps / sec / 1e6, 1), ms_per_step=round(sec / steps * 1e3, 4),
                         bad_cells=s.check_state())
        s.close()
    # (8/3) nv 8 B per cell-stage, nv = 4 (SSP-RK3 mean, FP64)
    out["fp64"]["achieved_gbs"] = round(out["fp64"]["mcs"] * 1e6 * (8.0 / 3.0) * 4 * 8 / 1e9, 1)
    return out


def bench_blast(H, steps, band=0.025):
    """3D Euler blast, base 64^3 + 2 levels where |r - 0.2| <= band, FP64
    (table-driven kernel and the persistent kernel) and FP32 state (the
    table-driven kernel with FP32 stage states, and the persistent kernel).  band 0.025 is the configured shell of width 0.05
    (0.82 M leaves); band 0.2 refines the ball r <= 0.4, which gives the
    configured 4-6 M leaves (4.97 M) and a state beyond the 256 MB MALL."""
    from hom3_amd import cases
    t0 = time.perf_counter()
    g, s, (nref, nbal) = cases.blast(H, base=6, extra=2, band=band)
    build_s = time.perf_counter() - t0
    n = s.first_gc
    out = dict(workload=f"3D Euler blast, unit cube 64^3 + 2 levels where |r - 0.2| <= {band}, 2:1 balanced, "
                        "rho 1/0.125, p 10/0.1, 1% seeded density perturbation, outflow, AUSM, SSP-RK3, CFL 0.1",
               leaf_cells=n, grid_nodes=g.size(), refined_nodes=nref, balance_refined=nbal, steps=steps,
               build_seconds=round(build_s, 2))
    B64 = H.Batch([s], "fp64")
    B32 = H.Batch([s], "fp32")
    sec = timed_solve(s, steps, 2)
    out["fp64_table_driven"] = dict(mcs=round(n * 3 * steps / sec / 1e6, 1), ms_per_step=round(sec / steps * 1e3, 4))
    _, s32, _ = cases.blast(H, base=6, extra=2, band=band, precision="fp32")
    sec = timed_solve(s32, steps, 2)
    l2, mx = state_error(s32.lhs[:, :n], s.lhs[:, :n], 3)
    out["fp32_state_table_driven"] = dict(mcs=round(n * 3 * steps / sec / 1e6, 1),
                                          ms_per_step=round(sec / steps * 1e3, 4),
                                          vs_fp64_table_driven=dict(rel_l2=l2, rel_max=mx))
    s32.close()
    for tag, B in (("fp64_persistent", B64), ("fp32_state_persistent", B32)):
        cfg = B.autotune()
        B.solve(2)
        ms = B.time_solve(steps, graph=True)
        out[tag] = dict(mcs=round(n * 3 / (ms * 1e-3) / 1e6, 1), ms_per_step=round(ms, 4),
                        config={k: cfg[k] for k in ("block", "tile", "grid")})
    q64, q32 = B64.state(0), B32.state(0)
    l2, mx = state_error(q32, q64, 3)
    out["fp32_vs_fp64"] = dict(vars=["rho_u0", "rho_u1", "rho_u2", "rho", "rho_E"], rel_l2=l2, rel_max=mx)
    B64.close()
    B32.close()
    s.close()
    return out


def bench_batch(H, n_inst, steps, seed=2024):
    """512 independent 2D four-quadrant instances, sides 64..512 (level
    uniform over 6..9), 0-2 refinement levels inside a seeded disc, one
    persistent kernel per stage, hipGraph steps, autotuned launch shape."""
    from hom3_amd.batch import build_instance
    t0 = time.perf_counter()
    prms = H.four_quadrant_instances(seed, n_inst)
    solvers = [build_instance(H, p) for p in prms]
    build_s = time.perf_counter() - t0
    out = dict(workload=f"{n_inst} x 2D Euler four-quadrant Riemann, base 64^2..512^2 (level uniform in 6..9), "
                        "0-2 refinement levels in a seeded disc (radius U[0.1,0.3]), rho,p ~ U[0.1,2], "
                        "u,v ~ U[-1,1], AUSM, SSP-RK3, CFL 0.4",
               instances=n_inst, steps=steps, build_seconds=round(build_s, 1),
               refined_instances=int(sum(p["extra"] > 0 for p in prms)))
    batches = {}
    for tag in ("fp64", "fp32"):
        B = H.Batch(solvers, tag)
        batches[tag] = B
        _, ncell = B.counts()
        out["cells"] = ncell
        cfg = B.autotune()
        B.solve(2)
        ms_g = B.time_solve(steps, graph=True)
        ms_s = B.time_solve(10, graph=False)
        bytes_cs = ((8.0 / 3.0) * 4 * 8 if tag == "fp64" else 16.0 * 4) + 4 * 4
        out[tag] = dict(mcs=round(ncell * 3 / (ms_g * 1e-3) / 1e6, 1), ms_per_step_graph=round(ms_g, 4),
                        ms_per_step_stream=round(ms_s, 4), config={k: cfg[k] for k in ("block", "tile", "grid")},
                        autotune_stage_ms=round(cfg["stage_ms"], 4),
                        achieved_gbs=round(bytes_cs * ncell * 3 / (ms_g * 1e-3) / 1e9, 1))
    # both batches have taken the same steps: FP32 state against FP64, worst
    # instance per variable
    l2w, mxw = np.zeros(4), np.zeros(4)
    for i in range(n_inst):
        l2, mx = state_error(batches["fp32"].state(i), batches["fp64"].state(i), 2)
        l2w, mxw = np.maximum(l2w, l2), np.maximum(mxw, mx)
    out["fp32_vs_fp64"] = dict(steps=2 + steps + 10, vars=["rho_u0", "rho_u1", "rho", "rho_E"],
                               worst_instance_rel_l2=[float(x) for x in l2w],
                               worst_instance_rel_max=[float(x) for x in mxw])
    for B in batches.values():
        B.close()
    for s in solvers:
        s.close()
    return out


def bench_cht_adjoint(H, level, steps, ckpt):
    """Channel over a heated block (fluid 128x64x64 CNS, solid 32x16x32 heat
    at level 7), conductivity ratio 100, seeded source; adjoint of the
    time-integrated wall heat flux w.r.t. the source.  Three sweeps: every
    step's stage inputs kept in HBM (no recomputation; the headline ratio),
    checkpoints every `ckpt` steps (one recomputation per segment), and
    binomial checkpointing with at most 10 saved states."""
    from hom3_amd import adjoint as A, cases
    g, c, h, src = cases.channel_cht(H, level=level)
    snap = A.Snapshot(c, h)
    J, dts = A.cht_objective(c, h, 0.01, steps)
    snap.restore()
    t0 = time.perf_counter()
    J2, _ = A.cht_objective(c, h, 0.01, steps, dts)
    prim_s = time.perf_counter() - t0
    prim_ms = prim_s * 1e3 / steps
    sweeps = {}
    grads = {}
    for name, K in (("stored", 0), (f"checkpoint_every_{ckpt}", ckpt), ("binomial_10_states", -10)):
        snap.restore()
        Ja, grad, ms_f, ms_r = A.cht_adjoint(c, h, 0.01, dts, K)
        grads[name] = grad
        sweeps[name] = dict(J=Ja, forward_ms_per_step=round(ms_f / steps, 4),
                            reverse_ms_per_step=round(ms_r / steps, 4),
                            adjoint_primal_cost_ratio=round((ms_r / steps) / prim_ms, 2),
                            gradient_total_cost_ratio=round(((ms_f + ms_r) / steps) / prim_ms, 2))
    g0, g1 = grads["stored"], grads[f"checkpoint_every_{ckpt}"]
    g2 = grads["binomial_10_states"]
    out = dict(workload=f"channel CHT, fluid {c.first_gc} CNS cells, solid {h.first_gc} heat cells, Mach 0.2, "
                        "Re 500 on the channel height, Pr 0.71, kratio 100, seeded source U[0,1], SSP-RK3",
               steps=steps, J=J2, primal_ms_per_step=round(prim_ms, 4), sweeps=sweeps,
               adjoint_primal_cost_ratio=sweeps["stored"]["adjoint_primal_cost_ratio"],
               grad_norm=float(np.linalg.norm(g0)),
               stored_vs_checkpointed_rel=float(np.linalg.norm(g0 - g1) / np.linalg.norm(g1)),
               stored_vs_binomial_rel=float(np.linalg.norm(g0 - g2) / np.linalg.norm(g2)),
               note="adjoint_primal_cost_ratio = reverse sweep per step / primal step; "
                    "gradient_total_cost_ratio = (forward sweep + reverse sweep) per step / primal step")
    c.close()
    h.close()
    return out


def bench_channel(H, level, steps, warmup=3):
    """The conjugate-heat-transfer channel (fluid 128x64x64 CNS with inflow,
    outflow, no-slip walls and the conjugate patch; solid 32x16x32 heat at
    level 7), coupled SSP-RK3 steps: the fluid on tiled bricks in FP64 and
    in FP32 state, and on the table-driven kernel; the FP32 state's error
    against the FP64 tiled run after the timed steps."""
    import torch
    from hom3_amd import cases
    out, finals = {}, {}
    for name, kernel, prec in (("fp64_tiled_bricks", "auto", "fp64"), ("fp32_state_tiled_bricks", "auto", "fp32"),
                               ("fp64_table_driven", "generic", "fp64")):
        g, c, h, _ = cases.channel_cht(H, level=level, kernel=kernel, precision=prec)
        H.coupling.apply_and_step(0.01, c, h, warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        H.coupling.apply_and_step(0.01, c, h, steps)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        ms = sec * 1e3 / steps
        out[name] = dict(kernel=c.kernel(), ms_per_step=round(ms, 4),
                         fluid_mcs=round(c.first_gc * 3 * steps / sec / 1e6, 1))
        finals[name] = (c.lhs[:, :c.first_gc].copy(), h.lhs[:, :h.first_gc].copy())
        c.close()
        h.close()
    a, b = finals["fp32_state_tiled_bricks"], finals["fp64_tiled_bricks"]
    out["fp32_state_tiled_bricks"]["error_vs_fp64"] = dict(
        steps=warmup + steps, fluid=state_error(a[0], b[0], 3), solid=state_error(a[1], b[1], 0))
    out["tiled_speedup_vs_table_driven"] = round(out["fp64_table_driven"]["ms_per_step"] /
                                                 out["fp64_tiled_bricks"]["ms_per_step"], 2)
    out["fp32_speedup_vs_fp64"] = round(out["fp64_tiled_bricks"]["ms_per_step"] /
                                        out["fp32_state_tiled_bricks"]["ms_per_step"], 3)
    out["workload"] = (f"channel CHT level {level}, fluid and solid coupled, SSP-RK3; fluid_mcs = fluid cells x 3 "
                       "stages / coupled step time (the step includes the solid solver and the coupling)")
    return out


def profile_channel_run(H, level=7, steps=5):
    """Profiling run of the channel: `steps` coupled steps each of the tiled
    FP64, tiled FP32-state and table-driven fluid, then a 1 GiB device copy
    (the calibration of the byte counters)."""
    from hom3_amd import cases
    for kernel, prec in (("auto", "fp64"), ("auto", "fp32"), ("generic", "fp64")):
        g, c, h, _ = cases.channel_cht(H, level=level, kernel=kernel, precision=prec)
        H.coupling.apply_and_step(0.01, c, h, steps)
        c.close()
        h.close()
        print(f"[profile] channel {kernel} {prec} {steps} steps", flush=True)
    ms = copy_calibration(H, 1 << 30)
    print(f"[profile] copy of 1 GiB in {ms:.3f} ms", flush=True)


# the sources the FP64 TGV stage kernel is compiled from: the traffic
# figure in profiles/ is valid for these bytes only
STAGE_KERNEL_SOURCES = ("hom3_amd/csrc/brick.hip", "hom3_amd/csrc/brick_geo.hpp", "hom3_amd/csrc/physics.hpp",
                        "hom3_amd/csrc/internal.hpp", "include/hom3g.h")
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "tgv256_fp64_stage_traffic.json")


def stage_kernel_sha16():
    import hashlib
    h = hashlib.sha256()
    for f in STAGE_KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def git_head():
    import subprocess
    try:
        return subprocess.run(["git", "-C", ROOT, "rev-parse", "--short=12", "HEAD"], capture_output=True,
                              text=True, timeout=10).stdout.strip() or None
    except Exception:
        return None


def load_traffic():
    """(bytes per stage launch, provenance) from profiles/; the provenance
    says whether the counters were taken on the kernel sources of this tree."""
    if not os.path.exists(TRAFFIC_JSON):
        return None, None
    with open(TRAFFIC_JSON) as f:
        d = json.load(f)
    sha = d.get("stage_kernel_sha16")
    prov = dict(commit=d.get("commit"), stage_kernel_sha16=sha, matches_current_sources=sha == stage_kernel_sha16())
    return d.get("traffic_bytes_per_launch"), prov


def _pmc_rows(path):
    import csv
    with open(path) as f:
        for r in csv.DictReader(f):
            yield r["Kernel_Name"], r["Counter_Name"], float(r["Counter_Value"]), int(r["Dispatch_Id"])


def write_traffic(fetch_csv, write_csv, level=8, out_path=TRAFFIC_JSON):
    """profiles/tgv256_fp64_stage_traffic.json from the counter files of two
    `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of `bench.py
    --profile`: HBM bytes per FP64 stage launch (mean over the launches after
    the first step), each counter scaled by what it reads for the 1 GiB copy
    of the same run."""
    def per_kernel(path, counter):
        acc = {}
        for name, cname, val, disp in _pmc_rows(path):
            if cname != counter:
                continue
            acc.setdefault((name, disp), 0.0)
            acc[(name, disp)] += val
        return acc
    stage_key = "k_brick<3, 1, false, double, double, double, false>"
    out = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        acc = per_kernel(path, counter)
        copy = [v for (n, d), v in acc.items() if "k_copy(" in n]
        stage = [v for (n, d), v in sorted(acc.items(), key=lambda kv: kv[0][1]) if stage_key in n]
        factor = (float(1 << 30) / 1024.0) / (sum(copy) / len(copy))
        out[counter] = dict(factor=factor, stage_kib=stage[3:])
    n = 1 << (3 * level)
    fetch = [k * out["FETCH_SIZE"]["factor"] * 1024 for k in out["FETCH_SIZE"]["stage_kib"]]
    write = [k * out["WRITE_SIZE"]["factor"] * 1024 for k in out["WRITE_SIZE"]["stage_kib"]]
    traffic = (sum(fetch) / len(fetch)) + (sum(write) / len(write))
    d = dict(kernel="k_brick<3,CNS,fp64> stage (256^3 TGV, SSP-RK3)",
             traffic_bytes_per_launch=int(round(traffic)),
             algorithmic_bytes_per_launch=int(round(BYTES_FP64 * n)),
             fetch_factor=round(out["FETCH_SIZE"]["factor"], 4), write_factor=round(out["WRITE_SIZE"]["factor"], 4),
             launches=len(fetch), commit=git_head(), stage_kernel_sha16=stage_kernel_sha16(),
             source="rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE -f csv -- python3 bench.py --profile; "
                    "python3 bench.py --write-traffic <FETCH csv> <WRITE csv>")
    with open(out_path, "w") as f:
        json.dump(d, f, indent=1)
        f.write("\n")
    return d


def copy_sweep(H, nbytes=1 << 30, reps=5):
    """Device copy of nbytes over copy variants and launch shapes (the
    achievable-bandwidth reference of the roofline report)."""
    import ctypes as C
    from hom3_amd import _lib
    names = {0: "grid-stride", 1: "grid-stride non-temporal", 2: "grid-stride x4 in flight, non-temporal stores",
             3: "one element per thread, non-temporal store"}
    rows = []
    for variant in (0, 1, 2, 3):
        for block in (256, 512, 1024):
            for grid in ((0,) if variant == 3 else (1024, 2048, 4096, 8192, 16384)):
                ms = C.c_float()
                _lib.check(_lib.lib().hom3g_copy_bandwidth(H.default_context().h, nbytes, variant, max(grid, 1),
                                                           block, reps, C.byref(ms)))
                rows.append(dict(variant=names[variant], grid=grid or None, block=block,
                                 gbs=round(2.0 * nbytes / (ms.value * 1e-3) / 1e9, 1)))
    best = max(rows, key=lambda r: r["gbs"])
    return dict(bytes=nbytes, best=best, sweep=rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--level", type=int, default=8, help="grid level (8 -> 256^3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--cpu-level", type=int, default=6)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-extra", action="store_true", help="only the headline TGV workload")
    ap.add_argument("--error-steps", type=int, default=200, help="steps of the FP32-vs-FP64 TGV comparison")
    ap.add_argument("--batch-n", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-level", type=int, default=7)
    ap.add_argument("--profile", action="store_true",
                    help="profiling run: 10 FP64 + 4 FP32-state steps + a 1 GiB device copy for counter calibration")
    ap.add_argument("--profile-adjoint", action="store_true",
                    help="profiling run of the CHT adjoint workload alone (--adjoint-steps steps)")
    ap.add_argument("--profile-vortex", action="store_true",
                    help="profiling run of the 4096^2 2D vortex: 4 FP64 + 4 FP32-state steps + a 1 GiB copy")
    ap.add_argument("--profile-channel", action="store_true",
                    help="profiling run of the CHT channel: 5 coupled steps per fluid kernel variant")
    ap.add_argument("--channel-steps", type=int, default=50)
    ap.add_argument("--copy-sweep", action="store_true", help="only the device-copy bandwidth sweep")
    ap.add_argument("--write-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/tgv256_fp64_stage_traffic.json from two rocprofv3 --pmc runs")
    ap.add_argument("--profile-batch", action="store_true",
                    help="profiling run of the 512-instance batch: 3 FP64 + 3 FP32-state steps + a 1 GiB copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is designed but not built).", file=sys.stderr)
        sys.exit(2)

    if args.write_traffic:
        print(json.dumps(write_traffic(*args.write_traffic, level=args.level)))
        return
    import hom3_amd as H
    if args.copy_sweep:
        print(json.dumps(copy_sweep(H)))
        return
    if args.profile:
        profile_run(H, args.level)
        return
    if args.profile_vortex:
        profile_vortex_run(H)
        return
    if args.profile_batch:
        profile_batch_run(H, args.batch_n)
        return
    if args.profile_channel:
        profile_channel_run(H)
        return
    if args.profile_adjoint:
        print(json.dumps(bench_cht_adjoint(H, args.adjoint_level, args.adjoint_steps, 20)))
        return
    print(f"[bench] TGV level {args.level} FP64 warmup {args.warmup} steps {args.steps}", file=sys.stderr, flush=True)
    r64 = time_gpu(H, args.level, "fp64", args.steps, args.warmup)
    print(f"[bench] fp64 {r64}", file=sys.stderr, flush=True)
    r32 = None
    if not args.no_fp32:
        r32 = time_gpu(H, args.level, "fp32", args.steps, args.warmup)
        print(f"[bench] fp32 {r32}", file=sys.stderr, flush=True)
    n = r64["n"]
    sweep = copy_sweep(H)
    achieved = BYTES_FP64 * n / (r64["stage_ms"] * 1e-3) / 1e9
    traffic, prov = load_traffic()
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_provenance=prov,
                kernel="k_brick<3,CNS,fp64> stage", bytes_per_cell_stage=round(BYTES_FP64, 2),
                stage_ms=round(r64["stage_ms"], 4), measured_copy_gbs=sweep["best"]["gbs"],
                measured_copy=sweep["best"])
    if prov is not None and not prov["matches_current_sources"]:
        roof["traffic_note"] = "counters taken on other stage-kernel sources than this tree's"
    cpu = None
    if not args.no_cpu:
        print(f"[bench] cpu oracle level {args.cpu_level} steps {args.cpu_steps}", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_level, args.cpu_steps, args.cpu_level + 1)
        print(f"[bench] cpu {cpu}", file=sys.stderr, flush=True)
    out = {
        "metric": "Mcell-stage updates/s + % HBM roofline, FP64 vs FP32; adjoint/primal cost ratio",
        "value": round(r64["mcs"], 2),
        "unit": "Mcell-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["seconds"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic Taylor-Green initial field)",
        "config": {"workload": f"CNS Taylor-Green vortex Re1600 M0.1 {2 ** args.level}^3 periodic, AUSM + "
                               "3-point viscous, SSP-RK3", "cells": n, "stages_per_step": 3},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if r32 is not None:
        a32 = BYTES_FP32 * n / (r32["stage_ms"] * 1e-3) / 1e9
        print(f"[bench] fp32 vs fp64 error after {args.error_steps} steps", file=sys.stderr, flush=True)
        out["fp32_state"] = dict(value=round(r32["mcs"], 2), unit="Mcell-stage/s",
                                 speedup_vs_fp64=round(r32["mcs"] / r64["mcs"], 3),
                                 roofline_achieved_gbs=round(a32, 1), frac=round(a32 / HBM_PEAK_GBS, 4),
                                 error_vs_fp64=tgv_fp32_error(H, args.level, args.error_steps))
        print("[bench] fp64 / fp32 against the oracle at 32^3", file=sys.stderr, flush=True)
        out["error_vs_oracle"] = tgv_against_oracle(H)
    if not args.no_extra:
        extra = {}
        for name, fn in (("vortex_2d", lambda: bench_vortex(H, 100)),
                         ("vortex_2d_4096", lambda: bench_vortex_large(H, 20)),
                         ("blast_3d_refined", lambda: bench_blast(H, 50)),
                         ("blast_3d_refined_5m", lambda: bench_blast(H, 50, band=0.2)),
                         ("batch_2d", lambda: bench_batch(H, args.batch_n, args.batch_steps)),
                         ("cht_channel", lambda: bench_channel(H, args.adjoint_level, args.channel_steps)),
                         ("cht_adjoint", lambda: bench_cht_adjoint(H, args.adjoint_level, args.adjoint_steps, 20))):
            print(f"[bench] {name}", file=sys.stderr, flush=True)
            try:
                extra[name] = fn()
            except Exception as e:  # report, keep the headline line
                extra[name] = dict(error=f"{type(e).__name__}: {e}")
            print(f"[bench] {name} {extra[name]}", file=sys.stderr, flush=True)
        out["extra"] = extra
        if "adjoint_primal_cost_ratio" in extra.get("cht_adjoint", {}):
            out["adjoint_primal_cost_ratio"] = extra["cht_adjoint"]["adjoint_primal_cost_ratio"]
    if r64["bad"] or (r32 and r32["bad"]):
        out["state_check"] = dict(fp64_bad_cells=r64["bad"], fp32_bad_cells=r32["bad"] if r32 else None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
