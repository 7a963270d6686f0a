"""Benchmark: implicit cloth step on one MI355X, mixed precision vs FP64.

Workload (default): self-contact fold, 512x512-vertex sheet (522,242
triangles, 1.5 m side) with a seeded sinusoidal crumple dropped onto the
floor; d_hat 5e-4, mu 0.4, h = 1/60, additive Schwarz on 64-vertex domains.
A "step" is one ClothSolver.step(h) of the whole scene.

Prints one JSON line:
  value        mixed-precision HipSolver steps/s (state resident in HBM;
               host copy-out of positions is excluded and reported apart)
  fp64         the FP64 HipSolver (reference operator) on the same scene
  roofline     HVP kernel (k_mx_hvp_tile): algorithmic (compulsory) bytes per
               launch, 40n + 32m + 64C for n vertices, m triangles and C
               active contacts, / mean launch time of back-to-back launches on
               the operands the timed steps left resident (HIP events on the
               solver's stream), vs the 8 TB/s HBM peak; the bytes the kernel
               actually moves (its index maps and prepared triangle records
               included) are kernel_bytes_per_launch, the counter-measured
               HBM traffic is traffic
  lbvh, ccd    the standalone broadphase (Karras LBVH) on the final state:
               Mprims/s = triangles / hierarchy build time; CCD queries/s =
               VT + EE primitives of its candidate pairs / kernel time
  accuracy_converged  mixed vs FP64 where both reach their Newton tolerance
  cpu_baseline the CPU oracle (FP64, one thread) timed on a bounded sample,
               plus whole strip-pull steps timed beside the GPU
  drape        the 256x256 sphere + floor drape (block-Jacobi): mixed and
               FP64 steps/s and its own HVP roofline
  batch        64 uneven drapes stepped together in mixed precision
               (cs_batch_step) vs one instance at a time in FP64
  adjoint      the 128x128 material adjoint: loss gradient by checkpointed
               step adjoints (FP64 and mixed forwards), FD check in shear, kb
  coarse       the two-level Schwarz preconditioner on the workload: steps/s
               and PCG iterations per step, one-level vs each coarse level
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def hip_runtime():
    """The HIP runtime the cloth library is linked against (torch may map
    another copy into the process)."""
    from closim_amd.solver import load_library
    load_library()
    path = next((ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln), "libamdhip64.so.7")
    return C.CDLL(path)


def hip_sync():
    hip_runtime().hipDeviceSynchronize()


def make_scene(name):
    from closim_amd import scenes as S
    if name == "sphere":
        return S.sphere_drape(255)
    if name == "fold":
        return S.crumple_fold(511)
    if name == "strip":
        return S.strip_pull()
    raise SystemExit(f"unknown workload {name}")


PROFILE_NAME = {"sphere": "sphere_drape_256", "fold": "crumple_fold_512", "strip": "strip_pull"}


def profile_traffic(workload):
    """HBM bytes per k_mx_hvp_tile launch from the committed counter summary of
    this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated), if there is one."""
    prof = os.path.join(ROOT, "profiles", f"k_mx_hvp_tile_{PROFILE_NAME.get(workload, workload)}.json")
    if not os.path.exists(prof):
        return None, None
    with open(prof) as f:
        pj = json.load(f)
    return pj["traffic_bytes_per_launch"], f"profiles/{pj['summary']} (FETCH_SIZE + WRITE_SIZE per launch, copy-calibrated)"

# Accuracy at convergence.  On the drape and fold sheets the reference's own
# Newton loop does not converge at any cap (its FD Hessian-vector product
# stalls at |g| ~ 5e-5, and with the analytic one the line search stalls on
# the compressed, indefinite membrane), so a mixed-vs-FP64 position error
# there measures where two unconverged iterations stopped.  The converged
# comparison runs the small scenes on which both precisions reach
# |g| < 1e-11: a 13x7 strip and a 7x7 patch on the floor without friction
# (lagged friction is not the gradient of the merit energy).
CONVERGED = dict(newton_iters=60, pcg_max_iters=400, grad_tol=1e-11)


def converged_comparison(steps, h):
    from closim_amd import _abi, scenes as S
    from closim_amd.solver import HipSolver
    out = []
    for name, make, popts in (("strip 13x7, Jacobi", S.small_strip, dict(precond=_abi.CS_PRECOND_JACOBI)),
                              ("floor patch 7x7 frictionless, block-Jacobi", lambda: S.floor_patch(mu=0.0),
                               dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI))):
        gm = HipSolver("mixed", **popts, **CONVERGED)
        gm.initialize(make())
        gd = HipSolver("fp64", **popts, hvp_mode=_abi.CS_HVP_ANALYTIC, **CONVERGED)
        gd.initialize(make())
        worst, gmax = 0.0, 0.0
        for _ in range(steps):
            for g in (gm, gd):
                raw_step(g, h)
                g._copy_out()
            gmax = max(gmax, gm.lastStats["grad_norm"], gd.lastStats["grad_norm"])
            pm, pd = gm.getPositions(), gd.getPositions()
            worst = max(worst, float(np.max(np.abs(pm - pd)) / np.max(np.abs(pd))))
        out.append(dict(scene=name, steps=steps, grad_tol=CONVERGED["grad_tol"], max_grad_norm=gmax,
                        pos_rel_err_mixed_vs_fp64_max_over_steps=worst))
        gm.close()
        gd.close()
    return out


WORKLOADS = {
    "fold": dict(name="self_contact_fold_512x512", precond="additive-schwarz 64/1/64"),
    "sphere": dict(name="sphere_floor_drape_256x256", precond="block-jacobi"),
    "strip": dict(name="strip_pull_20x100", precond="jacobi"),
}


def precond_options(workload):
    from closim_amd import _abi
    return {"sphere": dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI),
            "fold": dict(precond=_abi.CS_PRECOND_SCHWARZ, schwarz=(64, 1, 64)),
            "strip": dict(precond=_abi.CS_PRECOND_JACOBI)}[workload]


def coarse_section(workload, steps, warmup, h):
    """The two-level Schwarz preconditioner on the workload (mixed
    precision): steps/s, PCG iterations and Newton iterations per step, and
    the last step's gradient norm, one-level against the coarse corrections
    (damped Jacobi, block Jacobi, 8 coarse PCG iterations), each from the
    same initial state after `warmup` steps."""
    from closim_amd import _abi
    popts = precond_options(workload)
    if popts.get("precond") != _abi.CS_PRECOND_SCHWARZ:
        return None
    res = {}
    for name, mode in (("one_level", _abi.CS_COARSE_NONE), ("mas2", _abi.CS_COARSE_JACOBI),
                       ("c0", _abi.CS_COARSE_BLOCK), ("c1_8", _abi.CS_COARSE_PCG)):
        sc = make_scene(workload)
        g = solver_for(sc, "mixed", dict(popts, coarse=mode), timing=False)
        for _ in range(warmup):
            raw_step(g, h)
        pcg, newton = 0, 0
        t = 0.0
        for _ in range(steps):
            t += timed_steps(g, h, 1)
            pcg += g.lastStats["pcg_iters"]
            newton += g.lastStats["newton_used"]
        res[name] = dict(steps_per_s=steps / t, pcg_iters_per_step=pcg / steps, newton_per_step=newton / steps,
                         grad_norm_last_step=g.lastStats["grad_norm"])
        print(f"[bench] coarse {name}: {steps} steps in {t:.3f} s, {pcg / steps:.1f} PCG iterations per step",
              file=sys.stderr, flush=True)
        g.close()
    res["steps"], res["warmup"], res["precision"] = steps, warmup, "mixed"
    return res


def pair_section(workload, steps, warmup, h):
    """A second workload: mixed and FP64 steps/s from the same initial state
    after `warmup` steps each, the position difference, and the mixed HVP
    roofline on that mesh."""
    popts = precond_options(workload)
    res = {}
    for prec in ("mixed", "fp64"):
        sc = make_scene(workload)
        g = solver_for(sc, prec, popts, timing=False)
        for _ in range(warmup):
            raw_step(g, h)
        t = timed_steps(g, h, steps)
        st = g.lastStats
        ms, nbytes = g.time_hvp(200)
        _, rbytes = g.hvp_bytes()
        tuned = g.tuned_launch()
        g._copy_out()
        res[prec] = dict(steps_per_s=steps / t, ms_per_step=1000 * t / steps, newton_used_last_step=st["newton_used"],
                         pcg_iters_last_step=st["pcg_iters"], hvp_mean_us=1000 * ms,
                         hvp_algorithmic_bytes=nbytes, hvp_kernel_bytes=rbytes, tuned_launch=tuned,
                         hvp_gbs=nbytes / (ms / 1000) / 1e9 if ms > 0 else 0.0,
                         positions=g.getPositions().copy())
        print(f"[bench] {workload} {prec}: {steps} steps in {t:.3f} s", file=sys.stderr, flush=True)
        g.close()
    pm, pd = res["mixed"].pop("positions"), res["fp64"].pop("positions")
    m = res["mixed"]
    mesh = make_scene(workload).mesh
    tr = profile_traffic(workload)
    return dict(workload=WORKLOADS[workload]["name"], vertices=mesh.count, triangles=mesh.tri_count,
                preconditioner=WORKLOADS[workload]["precond"], warmup=warmup, steps=steps,
                mixed=m, fp64=res["fp64"], mixed_over_fp64=res["fp64"]["ms_per_step"] / m["ms_per_step"],
                pos_rel_err_mixed_vs_fp64=float(np.max(np.abs(pm - pd)) / np.max(np.abs(pd))),
                roofline_hvp={"bound": "hbm", "achieved": m["hvp_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": m["hvp_gbs"] / HBM_PEAK_GBS, "kernel": "k_mx_hvp_tile",
                              "bytes_per_launch": m["hvp_algorithmic_bytes"],
                              "kernel_bytes_per_launch": m["hvp_kernel_bytes"],
                              "traffic": tr[0], "traffic_source": tr[1]})


def gate_section(workload, steps, warmup, h, base_flags):
    """The workload with CpuSolver's descent gate (only a non-descent PCG
    direction falls back to steepest descent; newton.ts:154-160), both
    precisions from the same initial state after `warmup` steps: steps/s and
    the timed Newton / PCG / line-search work, with the steps whose line
    search failed and those in which the sheet did not move."""
    popts = precond_options(workload)
    res = {}
    for prec in ("mixed", "fp64"):
        g = solver_for(make_scene(workload), prec, popts, timing=False, flags=base_flags)
        for _ in range(warmup):
            raw_step(g, h)
        work = {}
        t = timed_steps(g, h, steps, work)
        res[prec] = dict(steps_per_s=steps / t, ms_per_step=1000 * t / steps, timed_work=work)
        print(f"[bench] {workload} {prec}, CpuSolver descent gate: {steps} steps in {t:.3f} s, {work}",
              file=sys.stderr, flush=True)
        g.close()
    res.update(steps=steps, warmup=warmup, mixed_over_fp64=res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"])
    return res


def lbvh_ccd_section(scene, h, reps=5):
    """The standalone broadphase (cs_broadphase_self_pairs: Morton codes,
    radix sort, Karras hierarchy, refit; then the pair query) over the swept
    boxes of the scene's current state -> its predictor, padded by d_hat;
    median of `reps` timed calls after one untimed call.  Then cs_ccd_queries
    on every VT / EE primitive of those pairs (expandTriPair), thickness
    d_min, kernel time of the second call."""
    from closim_amd.solver import broadphase_self_pairs, ccd_queries, expand_pairs
    x0 = scene.positions.copy()
    g = np.asarray(scene.gravity, np.float64)
    x1 = (x0.reshape(-1, 3) + h * scene.velocities.reshape(-1, 3) + h * h * g).reshape(-1)
    idx = np.ascontiguousarray(scene.mesh.indices, np.uint32)
    m = len(idx) // 3
    pad = scene.contact.params.dHatM
    pairs = broadphase_self_pairs(x0, x1, idx, pad)
    builds, queries = [], []
    for _ in range(reps):
        _, b, q = broadphase_self_pairs(x0, x1, idx, pad, cap=len(pairs) + 1, timing=True)
        builds.append(b)
        queries.append(q)
    bms, qms = float(np.median(builds)), float(np.median(queries))
    kinds, ids = expand_pairs(pairs, idx)
    thick = scene.contact.params.dMinM
    ccd_queries(x0, x1, kinds, ids, thick)
    toi, cms = ccd_queries(x0, x1, kinds, ids, thick, timing=True)
    print(f"[bench] lbvh: {m} triangles, build {bms:.3f} ms, query {qms:.3f} ms, {len(pairs)} pairs; "
          f"ccd: {len(kinds)} prims in {cms:.3f} ms", file=sys.stderr, flush=True)
    lbvh = dict(triangles=m, pad=pad, build_ms=bms, query_ms=qms, candidate_pairs=int(len(pairs)),
                mprims_per_s=m / (bms / 1000) / 1e6, reps=reps,
                note="build = bounds, Morton codes, radix sort, leaves, Karras hierarchy, refit; query = "
                     "count, scan, emit, per-triangle sort of the pairs (HIP events, null stream)")
    ccd = dict(queries=int(len(kinds)), vt=int((kinds == 0).sum()), ee=int((kinds == 1).sum()),
               thickness=thick, kernel_ms=cms, queries_per_s=len(kinds) / (cms / 1000),
               impacts=int(np.isfinite(toi).sum()),
               note="every VT / EE primitive of the broadphase pairs, x0 = state, x1 = predictor")
    return lbvh, ccd


def batch_section(steps, h, threads=8, count=64):
    """The batched uneven drape (`scenes.batch_drapes`: 64 sheets of 1e3..1e5
    vertices, each on its own sphere and floor): all instances stepped
    together in mixed precision (cs_batch_step, `threads` host threads, one
    HIP stream per instance), the same with the inner solves packed
    (cs_batch_step_packed, identical trajectories), and one instance at a time
    in FP64, all from the same initial states after one warmup step,
    block-Jacobi."""
    from closim_amd import _abi, scenes as S
    from closim_amd.solver import HipBatch
    popts = dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI)
    t0 = time.perf_counter()
    scenes_m, scenes_d = S.batch_drapes(count), S.batch_drapes(count)
    verts = int(sum(sc.mesh.count for sc in scenes_m))
    gm = [solver_for(sc, "mixed", popts, timing=False) for sc in scenes_m]
    gd = [solver_for(sc, "fp64", popts, timing=False) for sc in scenes_d]
    print(f"[bench] batch: {count} instances, {verts} vertices, set up in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    B = HipBatch(gm, threads=threads)
    B.step(h)
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        B.step(h)
    hip_sync()
    t_b = time.perf_counter() - t0
    print(f"[bench] batch mixed: {steps} steps of {count} instances in {t_b:.3f} s", file=sys.stderr, flush=True)
    scenes_p = S.batch_drapes(count)
    gp = [solver_for(sc, "mixed", popts, timing=False) for sc in scenes_p]
    P = HipBatch(gp, packed=True)
    P.step(h)
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        P.step(h)
    hip_sync()
    t_p = time.perf_counter() - t0
    same = all(np.array_equal(a.positions, b.positions) for a, b in zip(scenes_p, scenes_m))
    print(f"[bench] batch mixed packed: {steps} steps in {t_p:.3f} s (trajectories identical: {same})",
          file=sys.stderr, flush=True)
    for g in gp:
        g.close()
    for g in gd:
        raw_step(g, h)
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for g in gd:
            raw_step(g, h)
    hip_sync()
    t_s = time.perf_counter() - t0
    print(f"[bench] batch fp64 one at a time: {t_s:.3f} s", file=sys.stderr, flush=True)
    for g in gd:
        g._copy_out()
    err = max(float(np.max(np.abs(a.positions - b.positions)) / np.max(np.abs(b.positions)))
              for a, b in zip(scenes_m, scenes_d))
    for g in gm + gd:
        g.close()
    return dict(workload="batched_uneven_drape_64", instances=count, vertices=verts, threads=threads,
                preconditioner="block-jacobi", warmup=1, steps=steps,
                mixed_batched=dict(batch_steps_per_s=steps / t_b, instance_steps_per_s=steps * count / t_b,
                                   vertex_steps_per_s=steps * verts / t_b),
                mixed_packed=dict(batch_steps_per_s=steps / t_p, instance_steps_per_s=steps * count / t_p,
                                  identical_to_batched=same),
                fp64_one_at_a_time=dict(batch_steps_per_s=steps / t_s, instance_steps_per_s=steps * count / t_s),
                speedup=t_s / t_b, pos_rel_err_mixed_vs_fp64_max=err)


def adjoint_section(steps, h, fd_eps=1e-3):
    """The material-parameter adjoint on the 128x128 sheet pinned at two
    corners (`adjoint.material_problem`, seed 0): the target trajectory from
    the drawn parameters, then the loss and its gradient at the perturbed
    guess (forward, checkpoints every 5 steps, backward through
    cs_step_adjoint) with FP64 and with mixed forwards, and the central
    differences of the loss in shear and kb (relative step `fd_eps`) beside
    the FP64 gradient.  Forwards converge to the root (root polish, up to 400
    Newton iterations, Jacobi PCG capped at 100 iterations per Newton
    iteration); the largest final gradient norm of the steps is reported."""
    from closim_amd import adjoint as A
    out = dict(workload="material_adjoint_128x128", steps=steps, checkpoint_every=5, side_m=1.0)
    grads = {}
    for prec in ("fp64", "mixed"):
        sc, th_true, th_guess = A.material_problem(127)
        g = A.converged_solver(sc, prec, newton_iters=400, pcg_max_iters=100)
        r = A.MaterialRollout(g, steps=steps, h=h, checkpoint_every=5, strict=False,
                              progress=lambda m: print(f"[bench] adjoint {prec}: {m}", file=sys.stderr, flush=True))
        t0 = time.perf_counter()
        targets = r.run(th_true)
        t_fwd = time.perf_counter() - t0
        print(f"[bench] adjoint {prec}: target rollout {t_fwd:.1f} s", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        L, grad = A.loss_and_grad(r, th_guess, targets)
        t_fb = time.perf_counter() - t0
        its = [it for it, _ in r.adjoint_iters]
        res = dict(loss=L, gradient=dict(zip(A.PARAMS, map(float, grad))), max_final_grad_norm=r.max_grad_norm,
                   forward_s_per_step=t_fwd / steps,
                   forward_backward_s=t_fb, backward_s_per_step=(t_fb - t_fwd) / steps,
                   adjoint_pcg_iters_mean=float(np.mean(its)),
                   adjoint_rel_residual_max=float(max(rr for _, rr in r.adjoint_iters)))
        if prec == "fp64":
            out["theta_true"] = dict(zip(A.PARAMS, map(float, th_true)))
            out["theta_guess"] = dict(zip(A.PARAMS, map(float, th_guess)))
            res["fd"] = {}
            for j in (3, 4):  # shear (the largest relative signal) and kb
                e = A.fd_step(th_guess, j, fd_eps)
                tp, tm = th_guess.copy(), th_guess.copy()
                tp[j] += e
                tm[j] -= e
                lp = A.trajectory_loss(r.run(tp), targets)
                print(f"[bench] adjoint fd {A.PARAMS[j]} +", file=sys.stderr, flush=True)
                lm = A.trajectory_loss(r.run(tm), targets)
                print(f"[bench] adjoint fd {A.PARAMS[j]} -", file=sys.stderr, flush=True)
                fd = (lp - lm) / (2 * e)
                res["fd"][A.PARAMS[j]] = dict(rel_step=fd_eps, fd=fd, adjoint=float(grad[j]),
                                              rel_err=abs(fd - grad[j]) / abs(fd))
            res["max_final_grad_norm_fd_rollouts"] = r.max_grad_norm
        grads[prec] = grad
        out[prec] = res
        print(f"[bench] adjoint {prec}: {steps} steps forward {t_fwd:.1f} s, forward+backward {t_fb:.1f} s",
              file=sys.stderr, flush=True)
        g.close()
    out["gradient_rel_diff_mixed_vs_fp64"] = float(np.max(np.abs(grads["mixed"] - grads["fp64"]) /
                                                        np.maximum(np.abs(grads["fp64"]), 1e-300)))
    return out


def solver_for(scene, precision, popts, timing, flags=0):
    from closim_amd import _abi
    from closim_amd.solver import HipSolver
    flags = flags | (_abi.CS_FLAG_TIMING if timing else 0)
    g = HipSolver(precision, **popts, flags=flags)
    g.initialize(scene)
    return g


def raw_step(g, h):
    from closim_amd.solver import _check
    _check(g._lib.cs_step(g._h, h))


def timed_steps(g, h, steps, work=None):
    """Wall time of `steps` steps.  With a dict `work`, also the Newton, PCG
    and line-search totals of those steps (cs_last_stats after each step, a
    host-side read), the steps whose Newton loop stopped before its cap
    without converging (a failed line search), and the steps whose energy
    equals the previous step's (the sheet did not move)."""
    hip_sync()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        raw_step(g, h)
        if work is not None:
            stats.append(last_stats_raw(g))
    hip_sync()
    t = time.perf_counter() - t0
    if work is not None:
        work.update(newton_iters=sum(s.newton_used for s in stats), pcg_iters=sum(s.pcg_iters for s in stats),
                    line_search_trials=sum(s.line_search_trials for s in stats),
                    steps_stopped_early=sum(1 for s in stats if s.newton_used < s.newton_iters and not s.converged),
                    steps_not_moved=sum(1 for a, b in zip(stats, stats[1:]) if a.energy == b.energy))
    return t


def last_stats_raw(g):
    import ctypes as C
    from closim_amd import _abi
    st = _abi.CsStepStats()
    g._lib.cs_last_stats(g._h, C.byref(st))
    return st


def cpu_baseline(name, popts, budget_s, newton_per_step, h):
    """CPU oracle (FP64, one host thread) on a bounded sample of the same
    workload.  Scenes whose steps fit the budget run whole steps until
    `budget_s` has elapsed.  Otherwise (the oracle needs ~50 s per Newton
    iteration of the 512x512 fold, ~10 per step) the sample is step 0 with
    the Newton cap at 1 (predictor and CCD bisection, one Newton iteration
    with its whole PCG solve and line search, velocity update and filter),
    timed whole, and the rate is scaled to whole steps by the Newton
    iterations per step the GPU FP64 run used (stated in `sample`)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ffi as O
    from closim_amd import _abi
    scene = make_scene(name)
    if scene.mesh.count <= 20000:
        o = O.OracleSolver(scene, _abi.options(**popts))
        n = 0
        t0 = time.perf_counter()
        while True:
            o.step(h)
            n += 1
            if time.perf_counter() - t0 >= budget_s:
                break
        dt = time.perf_counter() - t0
        return dict(value=n / dt, unit="steps/s", cores=1, kind="port",
                    sample=f"first {n} whole step(s) of the same workload from its initial state "
                           f"(CPU oracle, FP64, one host thread), {dt:.1f} s")
    o = O.OracleSolver(scene, _abi.options(**popts, newton_iters=1))
    t0 = time.perf_counter()
    o.step(h)
    dt = time.perf_counter() - t0
    st = o.lastStats
    used = max(st["newton_used"], 1)
    steps_per_s = 1.0 / (dt / used * newton_per_step)
    return dict(value=steps_per_s, unit="steps/s", cores=1, kind="port",
                sample=f"step 0 of the same workload with the Newton cap at 1 ({used} Newton iteration, "
                       f"{st['pcg_iters']} PCG iterations, {st['line_search_trials']} line-search trial(s), with "
                       f"predictor, CCD bisection, velocity update and filter) timed whole on one host thread: "
                       f"{dt:.1f} s; scaled by the {newton_per_step} Newton iterations per step of the GPU FP64 "
                       f"run")


def strip_whole_steps(h, budget_s):
    """Whole steps of the strip-pull workload (bottom row driven 5 cm down
    over 60 steps, Newton tol 1e-6, Jacobi): the CPU oracle on one host thread
    for up to `budget_s`, and the FP64 and mixed HipSolver over the same
    steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ffi as O
    from closim_amd import _abi, scenes as S
    from closim_amd.solver import HipSolver
    kw = dict(precond=_abi.CS_PRECOND_JACOBI, grad_tol=1e-6)
    sc = S.strip_pull()
    o = O.OracleSolver(sc, _abi.options(**kw))
    n = 0
    t0 = time.perf_counter()
    while n < 60:
        for i, p in S.strip_pull_targets(sc, n):
            o.pinVertex(i, p)
        o.step(h)
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    t_cpu = time.perf_counter() - t0
    ref = o.getPositions().copy()
    out = dict(workload="strip_pull_20x100", steps=n, cpu_oracle_steps_per_s=n / t_cpu, cpu_cores=1)
    for prec in ("fp64", "mixed"):
        sc = S.strip_pull()
        g = HipSolver(prec, **kw)
        g.initialize(sc)
        t = 0.0
        for k in range(n):
            for i, p in S.strip_pull_targets(sc, k):
                g.pinVertex(i, p)
            t0 = time.perf_counter()
            raw_step(g, h)
            t += time.perf_counter() - t0
        g._copy_out()
        out[f"gpu_{prec}_steps_per_s"] = n / t
        out[f"pos_rel_err_{prec}_vs_oracle"] = float(np.max(np.abs(g.getPositions() - ref)) / np.max(np.abs(ref)))
        g.close()
    out["note"] = "whole steps, pins driven between steps (not timed); the GPU is latency-bound on 2,000 vertices"
    return out


def copy_calibration(nbytes=256 << 20, reps=8):
    """Device-to-device copies of a known size (hipMemcpyAsync), so that the
    FETCH_SIZE / WRITE_SIZE counters of the same run can be scaled to bytes."""
    lib = hip_runtime()
    a, b = C.c_void_p(), C.c_void_p()
    assert lib.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert lib.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert lib.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert lib.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert lib.hipDeviceSynchronize() == 0
    lib.hipFree(a)
    lib.hipFree(b)
    print(f"[bench] copy calibration: {reps} device copies of {nbytes} bytes", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="fold", choices=["fold", "sphere", "strip"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--hvp-reps", type=int, default=200)
    ap.add_argument("--host-newton", action="store_true",
                    help="decide the descent safeguard, trust region and each Armijo trial on the host "
                         "(default: CS_FLAG_DEVICE_NEWTON, the same results with one host read per trial)")
    ap.add_argument("--speculate", action="store_true",
                    help="with the device Newton control, evaluate Armijo trials in speculative batches of "
                         "1, 2, 4, 8 (CS_FLAG_ARMIJO_SPECULATE; the same results)")
    ap.add_argument("--descent-gate", choices=["webgpu", "cpusolver"], default="webgpu",
                    help="Newton's descent gate: the reference WebGPU Newton's (steepest descent also after a PCG "
                         "breakdown, CS_FLAG_BREAKDOWN_DESCENT) or CpuSolver's (only for a non-descent direction)")
    ap.add_argument("--cpusolver-gate-steps", type=int, default=None,
                    help="timed steps of the fold with CpuSolver's descent gate, both precisions (default --steps, "
                         "0 = skip)")
    ap.add_argument("--eager", action="store_true",
                    help="launch the mixed PCG kernels one by one instead of replaying captured graphs")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per HVP launch from the FETCH_SIZE/WRITE_SIZE counter runs")
    ap.add_argument("--accuracy-steps", type=int, default=None,
                    help="steps of the converged small-scene comparison (default 20, 0 = skip)")
    ap.add_argument("--drape-steps", type=int, default=5,
                    help="timed steps of the 256x256 sphere drape section (0 = skip)")
    ap.add_argument("--lbvh-reps", type=int, default=5, help="timed broadphase calls (0 = skip lbvh and ccd)")
    ap.add_argument("--coarse-steps", type=int, default=5,
                    help="steps per preconditioner of the two-level Schwarz comparison (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=2,
                    help="timed steps of the 64-instance batched drape section (0 = skip)")
    ap.add_argument("--batch-threads", type=int, default=16)
    ap.add_argument("--adjoint-steps", type=int, default=5,
                    help="rollout steps of the 128x128 material-adjoint section (0 = skip; the workload has 30)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run device copies of known size (for counter calibration runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the cloth step is single-device); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    from closim_amd import _abi
    from closim_amd.solver import device_info
    h = 1.0 / 60.0
    popts = precond_options(args.workload)
    base_flags = (_abi.CS_FLAG_EAGER if args.eager else 0) | (0 if args.host_newton else _abi.CS_FLAG_DEVICE_NEWTON)
    if args.speculate and not args.host_newton:
        base_flags |= _abi.CS_FLAG_ARMIJO_SPECULATE
    flags = base_flags | (_abi.CS_FLAG_BREAKDOWN_DESCENT if args.descent_gate == "webgpu" else 0)
    print(f"[bench] device: {device_info(0)}", file=sys.stderr, flush=True)
    sc_m = make_scene(args.workload)
    mesh = sc_m.mesh
    print(f"[bench] workload {args.workload}: {mesh.count} vertices, {mesh.tri_count} triangles", file=sys.stderr,
          flush=True)
    gm = solver_for(sc_m, "mixed", popts, timing=False, flags=flags)
    gd = None
    if not args.no_fp64:
        sc_d = make_scene(args.workload)
        gd = solver_for(sc_d, "fp64", popts, timing=False, flags=flags)
    # warmup (both solvers advance from the same initial state)
    for k in range(args.warmup):
        raw_step(gm, h)
        if gd:
            raw_step(gd, h)
        print(f"[bench] warmup step {k} done", file=sys.stderr, flush=True)
    work_m = {}
    t_m = timed_steps(gm, h, args.steps, work_m)
    print(f"[bench] mixed: {args.steps} steps in {t_m:.3f} s", file=sys.stderr, flush=True)
    st_m = gm.lastStats
    hvp_ms, hvp_bytes = gm.time_hvp(args.hvp_reps)
    print(f"[bench] k_mx_hvp_tile: {1000 * hvp_ms:.2f} us x {args.hvp_reps} launches ({gm.hvp_layout()} tiles)",
          file=sys.stderr, flush=True)
    res = {}
    if gd:
        work_d = {}
        t_d = timed_steps(gd, h, args.steps, work_d)
        print(f"[bench] fp64: {args.steps} steps in {t_d:.3f} s", file=sys.stderr, flush=True)
        st_d = gd.lastStats
        hvp_ms_d, hvp_bytes_d = gd.time_hvp(args.hvp_reps)
        gm._copy_out()
        gd._copy_out()
        pm, pd = gm.getPositions(), gd.getPositions()
        err = float(np.max(np.abs(pm - pd)) / np.max(np.abs(pd)))
        res["fp64"] = dict(steps_per_s=args.steps / t_d, ms_per_step=1000 * t_d / args.steps,
                           hvp_kernel="k_pcg_hvp (FP64 finite-difference membrane HVP, the CpuSolver operator)",
                           hvp_mean_us=1000 * hvp_ms_d, hvp_gbs=hvp_bytes_d / (hvp_ms_d / 1000) / 1e9,
                           newton_used_last_step=st_d["newton_used"], pcg_iters_last_step=st_d["pcg_iters"],
                           timed_work=work_d)
        res["mixed_over_fp64"] = t_d / t_m
        res["pos_rel_err_mixed_vs_fp64"] = err
        # sensitivity of the same capped solve to the Hessian operator alone:
        # FP64 with the analytic HVP instead of the FD one
        sc_a = make_scene(args.workload)
        from closim_amd.solver import HipSolver
        ga = HipSolver("fp64", **popts, hvp_mode=_abi.CS_HVP_ANALYTIC, flags=flags)
        ga.initialize(sc_a)
        for _ in range(args.warmup + args.steps):
            raw_step(ga, h)
        ga._copy_out()
        pa = ga.getPositions()
        res["pos_rel_err_fp64_analytic_vs_fp64"] = float(np.max(np.abs(pa - pd)) / np.max(np.abs(pd)))
        res["pos_rel_err_note"] = ("at the default caps (10 Newton, 60 PCG iterations) neither precision converges "
                                   "on this sheet; the FP64 analytic-vs-FD figure is the spread from changing the "
                                   "Hessian operator alone; accuracy_converged holds the converged comparison")
        ga.close()
    acc_steps = 20 if args.accuracy_steps is None else args.accuracy_steps
    if gd and acc_steps > 0:
        res["accuracy_converged"] = converged_comparison(acc_steps, h)
    mean_s = hvp_ms / 1000.0
    achieved = hvp_bytes / mean_s / 1e9 if mean_s > 0 else 0.0
    _, hvp_kernel_bytes = gm.hvp_bytes()
    traffic = args.traffic
    traffic_src = "--traffic" if traffic is not None else None
    if traffic is None:
        traffic, traffic_src = profile_traffic(args.workload)
    # PCIe-inclusive rate (positions + velocities copied out each step)
    t0 = time.perf_counter()
    for _ in range(2):
        raw_step(gm, h)
        gm._copy_out()
    pcie_rate = 2 / (time.perf_counter() - t0)
    out = {
        "metric": "sim steps/s + HVP HBM GB/s, mixed vs FP64 GPU at <=1e-5 rel pos err",
        "value": args.steps / t_m,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000 * t_m / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner solve / f64 Newton",
        "data": "synthetic",
        "config": {"workload": sc_m.meta.get("name", WORKLOADS[args.workload]["name"]),
                   "vertices": mesh.count, "triangles": mesh.tri_count, "h": h,
                   "preconditioner": WORKLOADS[args.workload]["precond"],
                   "pcg_launch": "eager" if args.eager else "graph replay",
                   "newton_control": "host" if args.host_newton else (
                       "device (speculative Armijo batches)" if args.speculate else "device (one Armijo trial per host read)"),
                   "descent_gate": ("reference WebGPU Newton: steepest descent after a non-descent direction or a "
                                    "PCG breakdown" if args.descent_gate == "webgpu" else
                                    "CpuSolver: steepest descent after a non-descent direction"),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_mx_hvp_tile (FP32 analytic membrane HVP + mass + frozen-contact HVP over "
                               "vertex-owner tiles)",
                     "bytes_per_launch": hvp_bytes, "kernel_bytes_per_launch": hvp_kernel_bytes,
                     "bytes_model": "40n + 32m + 64C (compulsory: x, p, y, mass per vertex; indices, Dm^-1, area "
                                    "per triangle; one record per active contact)",
                     "mean_launch_us": mean_s * 1e6, "launches": args.hvp_reps,
                     "traffic_source": traffic_src},
        "timed_work": work_m,
        "newton_used_last_step": st_m["newton_used"],
        "pcg_iters_last_step": st_m["pcg_iters"],
        "active_contacts_last_step": st_m["contact"]["active_pairs"] if st_m["contact"] else 0,
        "candidate_prims_last_step": st_m["contact"]["candidate_pairs"] if st_m["contact"] else 0,
        "velocity_filter_rounds_last_step": st_m["contact"]["filter_rounds"] if st_m["contact"] else 0,
        "pcie_inclusive_steps_per_s": pcie_rate,
        "tuned_launch": gm.tuned_launch(),
    }
    out.update(res)
    if args.lbvh_reps > 0 and sc_m.contact is not None:
        gm._copy_out()
        out["lbvh"], out["ccd"] = lbvh_ccd_section(sc_m, h, reps=args.lbvh_reps)
    gm.close()
    if gd:
        gd.close()
    gate_steps = args.steps if args.cpusolver_gate_steps is None else args.cpusolver_gate_steps
    if args.workload == "fold" and args.descent_gate == "webgpu" and gate_steps > 0 and gd:
        out["cpusolver_gate"] = gate_section(args.workload, gate_steps, args.warmup, h, base_flags)
    if args.coarse_steps > 0:
        out["coarse"] = coarse_section(args.workload, args.coarse_steps, 2, h)
    if args.workload == "fold" and args.drape_steps > 0 and gd:
        out["drape"] = pair_section("sphere", args.drape_steps, 2, h)
    if args.batch_steps > 0 and gd:
        out["batch"] = batch_section(args.batch_steps, h, threads=args.batch_threads)
    if args.adjoint_steps > 0 and gd:
        out["adjoint"] = adjoint_section(args.adjoint_steps, h)
    if args.copy_calibration:
        copy_calibration()
    if not args.no_cpu:
        print("[bench] timing the CPU oracle sample", file=sys.stderr, flush=True)
        nps = res.get("fp64", {}).get("newton_used_last_step", st_m["newton_used"])
        out["cpu_baseline"] = cpu_baseline(args.workload, popts, args.cpu_budget, nps, h)
        out["cpu_baseline"]["whole_steps_strip_pull"] = strip_whole_steps(h, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
