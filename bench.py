"""Benchmark: 3D acoustic FD modelling (201^3 model + 30-cell absorbing boundary, nt=1500) on one
MI355X, reported as grid-cell updates per second and as a fraction of the HBM roofline, with the
reference CPU path timed on this host's cores beside it.

A "step" is one complete modelling run (all nt time steps) over the resident model.
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--variant V] [--ny-chunk C]
                       [--cpu-steps S] [--no-cpu] [--no-fwi] [--no-tri] [--no-batch] [--no-tune]
Beside the headline 3D line it reports, as sub-objects, the FWI gradient (batched through
torch.autograd vs the plain per-shot loop), time-reversal imaging and batched ragged 2D modelling.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E
CPU_PROCS = 16  # worker processes of the multi-process CPU baselines: the host-core share of one GPU


def cpu_baseline(w, nt_cpu):
    """Reference C (oracle/_ref/pwref, single thread, reference flags) on a truncated run of the
    same workload; the per-step cost does not depend on the step index, so Gcell/s carries over.
    Falls back to the oracle restatement when the reference harness is not present."""
    vel = w["vel"]
    nz, nx, ny = vel.shape
    nb = w["nbt"]
    cells = (nz + 2 * nb) * (nx + 2 * nb) * (ny + 2 * nb)
    c = nz // 2
    # onset at step 0 so the truncated run injects a non-trivial source
    src = np.array([c, c, c, 15.0, 0.0, 1.0], np.float32)
    try:
        from oracle import oracle as O
        from oracle import refrun
        rick = O.ricker_bank(nt_cpu, w["dt"], [15.0], [0.0], [1.0])
        t = time.perf_counter()
        O.afd3d(vel, nt_cpu, w["dt"], 10, 10, 10, nb, w["ct"], [c], [c], [c], rick)
        el_port = time.perf_counter() - t
        port = dict(value=cells * nt_cpu / el_port / 1e9, unit="Gcell-updates/s", cores=1, kind="port",
                    sample="the oracle's bit-exact C restatement of afd3d (oracle/orc_afd.c, z-innermost loops) on "
                           "the same 201^3 model, %d of 1500 steps, %.1f s, single thread" % (nt_cpu, el_port))
        if not refrun.available():
            return port
        _, el = refrun.afd3d(vel, nt_cpu, w["dt"], *w["d"][1:], w["d"][0], nb, w["ct"], src, 1)
        out = dict(value=cells * nt_cpu / el / 1e9, unit="Gcell-updates/s", cores=1, kind="reference",
                   sample="reference C afd3d (fdm, pywave/src/wave_fdm.c:28) compiled from its sources "
                          "(oracle/Makefile), on the same 201^3 model, %d of 1500 steps, %.1f s, single thread, "
                          "-O2 -fwrapv" % (nt_cpu, el))
        port["reference_over_port_time"] = round(el / el_port, 3)
        out["port"] = port
        return out
    except Exception as e:  # the baseline is reported, never required
        return dict(value=None, unit="Gcell-updates/s", cores=1, kind="reference", sample="failed: %s" % e)


def fwi_bench(plain_shots, widths=None):
    """FWI gradient throughput on the 400x800, 32-shot, nt=3000 workload (8 Hz Ricker, 60-cell
    sponge): the batched gradient -- all 32 shots per launch, checkpoints every 50 steps --
    taken through torch.autograd (fwi_traces + L2 misfit + backward, i.e. the hand-written
    adjoint kernel), against the plain per-shot loop (one shot per launch, same kernels, direct
    C-ABI gradient), in shots per second."""
    import torch
    from pywave_amd import FwiProblem, fwi_traces
    from pywave_amd.workloads import fwi2d

    w = fwi2d()
    par = w["par"]
    nt, ns = par["nt"], par["ns"]
    wav = (w["amp"] * ricker_wavelet(nt, par["dt"], w["f0"], w["t0"])).astype(np.float32)
    gen = FwiProblem(par, wav, batch=ns)
    obs = gen.forward(w["vtrue"])  # (nt, nr, ns)
    gen.close()
    out = {"workload": "fwi2d 400x800 at 10 m (+60-cell sponge: 520x920), 32 shots, nt=3000, dt=1 ms, 8 Hz, "
                       "exact-adjoint L2 gradient; the forward fields of every step kept in HBM when they fit "
                       "(no recomputation), else checkpoints every 50 steps"}
    dev = torch.device("cuda:0")
    d = torch.tensor(np.ascontiguousarray(obs.transpose(2, 0, 1)), device=dev)  # (ns, nt, nr)

    def grad_autograd(prob):
        v = torch.tensor(w["v0"], device=dev, requires_grad=True)
        tr = fwi_traces(v, prob)
        loss = 0.5 * ((tr.double() - d.double()) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        return v.grad, loss.item()

    def timed_gradient(checkpoint):
        prob = FwiProblem(par, wav, batch=ns, checkpoint=checkpoint)
        fwi_set_widths(prob, widths)
        grad_autograd(prob)  # warm-up (also times the step kernels' tile widths)
        t = time.perf_counter()
        g, f = grad_autograd(prob)
        el = time.perf_counter() - t
        fw, bw = prob.timing()
        info = prob.info()
        return prob, g, f, el, fw, bw, info

    prob, g, f, tb, fw, bw, info = timed_gradient(0)
    out["batched_checkpoint_steps"] = info["checkpoint_steps"]
    out["batched_segments"] = info["segments"]
    out["batched_tile_widths"] = prob.widths()
    out["batched_tuning_ms_per_launch"] = prob.tuning_log()
    prob.close()
    out["batched_shots_per_s"] = round(ns / tb, 3)
    out["batched_s_per_gradient"] = round(tb, 4)
    out["batched_forward_ms"] = round(fw, 1)
    out["batched_backward_ms"] = round(bw, 1)
    out["misfit"] = f
    out["gradient_norm"] = float(torch.linalg.vector_norm(g.double()).item())
    # the same gradient with checkpoints every 50 steps (59 of 60 segments recomputed)
    prob, g50, _, t50, fw50, bw50, _ = timed_gradient(50)
    prob.close()
    out["checkpointed_50_shots_per_s"] = round(ns / t50, 3)
    out["checkpointed_50_forward_ms"] = round(fw50, 1)
    out["checkpointed_50_backward_ms"] = round(bw50, 1)
    # the exact adjoint's FMA contraction can differ between tile loops, so the two gradients agree
    # to f32 rounding rather than bit for bit
    out["checkpointed_50_rel_l2_vs_stored"] = float(torch.linalg.vector_norm((g - g50).double()).item() /
                                                    torch.linalg.vector_norm(g.double()).item())
    # the reference's own gradient (gradient_av: imaging condition, windowed residual, first-call
    # scaling, smoothing) that pfwi mode 2 (fwic) iterates on, every shot in one batch
    ref = FwiProblem(par, wav, data=obs, batch=ns)
    ref.gradient(w["v0"], "reference")  # warm-up (sets the first-call scaling)
    t = time.perf_counter()
    ref.gradient(w["v0"], "reference")
    tr = time.perf_counter() - t
    rf, rb = ref.timing()
    ref.close()
    out["reference_gradient_shots_per_s"] = round(ns / tr, 3)
    out["reference_gradient_s"] = round(tr, 4)
    out["reference_gradient_forward_ms"] = round(rf, 1)
    out["reference_gradient_backward_ms"] = round(rb, 1)
    # fwi() (pfwi mode 2's loop, pywave/src/wave_fwi.c:24-119): 5 L-BFGS iterations of the reference
    # gradient with Wolfe line searches, every shot in one batch; misfit of each stored model (f64)
    lb = FwiProblem(par, wav, data=obs, batch=ns)
    t = time.perf_counter()
    vinv, k = lb.lbfgs(w["v0"], 5)
    tl = time.perf_counter() - t
    mis = [lb.misfit(w["v0"])] + [lb.misfit(vinv[:, :, i]) for i in range(k)]
    lb.close()
    out["lbfgs"] = {"iterations": k, "s": round(tl, 3), "s_per_iteration": round(tl / max(k, 1), 3),
                    "misfit_start_then_per_iterate": [round(m, 6) for m in mis]}
    out["cpu_baseline"] = fwi_cpu_baseline(par, w)
    if plain_shots <= 0:
        return out
    # plain loop: one shot per launch, timed on the first `plain_shots` shots
    par1 = dict(par)
    par1["ns"] = plain_shots
    obs1 = np.ascontiguousarray(obs[:, :, :plain_shots])
    plain = FwiProblem(par1, wav, data=obs1, batch=1)
    plain.set_steps_per_launch(1)  # the plain method: one kernel per time step
    plain.gradient(w["v0"], "exact")
    t = time.perf_counter()
    plain.gradient(w["v0"], "exact")
    tp = time.perf_counter() - t
    out["plain_tile_widths"] = plain.widths()
    plain.close()
    out["plain_shots_per_s"] = round(plain_shots / tp, 3)
    out["plain_sample_shots"] = plain_shots
    out["batched_over_plain"] = round((ns / tb) / (plain_shots / tp), 3)
    return out


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_concurrently(fn, procs):
    """Run fn() procs times at once (each call runs one reference process and returns its engine
    seconds); returns the slowest engine time."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(procs) as ex:
        return max(ex.map(lambda _: fn(), range(procs)))


def fwi_cpu_baseline(par, w, nt_cpu=300):
    """Reference C gradient_av (pywave/src/wave_fwigradient.c:93-262, single thread) for one shot
    of the same 400x800 geometry over the first nt_cpu of 3000 steps (its per-step cost does not
    depend on the step), extrapolated to shots/s at nt=3000.  It is the reference's own gradient
    (imaging condition, every forward snapshot kept), the CPU path this engine replaces."""
    try:
        from oracle import refrun
        if not refrun.available():
            return None
        from pywave_amd.fwi import make_pararray
        p1 = dict(par)
        p1["ns"], p1["nt"] = 1, nt_cpu
        pa = make_pararray(p1)
        wav = (w["amp"] * ricker_wavelet(nt_cpu, par["dt"], w["f0"], w["t0"])).astype(np.float32)
        data = np.zeros((nt_cpu, par["nx"], 1), np.float32)
        q = np.ones_like(w["v0"])
        _, _, el = refrun.fwigrad(pa, w["v0"], q, wav, data)
        per_shot = el * par["nt"] / nt_cpu
        out = dict(value=round(1.0 / per_shot, 5), unit="shots/s", cores=1, kind="reference",
                   sample="reference C gradient_av, one shot of the 400x800 (+60 sponge) geometry, %d of %d steps "
                          "in %.1f s, extrapolated to %.1f s per shot; single thread, -O2 -fwrapv"
                          % (nt_cpu, par["nt"], el, per_shot))
        # shots are independent: the same sample in CPU_PROCS processes at once
        procs = min(CPU_PROCS, os.cpu_count() or 1)
        elp = run_concurrently(lambda: refrun.fwigrad(pa, w["v0"], q, wav, data)[2], procs)
        out["parallel"] = dict(value=round(procs / (elp * par["nt"] / nt_cpu), 5), unit="shots/s", procs=procs,
                               cpu_model=cpu_model(),
                               sample="%d concurrent reference processes, each the same one-shot sample; slowest "
                                      "%.1f s" % (procs, elp))
        return out
    except Exception as e:  # reported, never required
        return dict(value=None, unit="shots/s", cores=1, kind="reference", sample="failed: %s" % e)


def batch_cpu_baseline(shots, total_cu, nt_cpu=400):
    """Reference C afd2d (fdm2d, pywave/src/wave_fdm.c:397-686, single thread) on the first shot of
    the batch over nt_cpu steps, as cell-updates/s, extrapolated to the whole 96-shot workload."""
    try:
        from oracle import refrun
        if not refrun.available():
            return None
        s0 = shots[0]
        nz, nx = s0["vel"].shape
        src = s0["src"].copy()
        _, el = refrun.afd2d(s0["vel"], nt_cpu, s0["dt"], s0["d"][1], s0["d"][0], 30, 0.01, src, 1)
        cells = (nz + 60) * (nx + 60) * nt_cpu
        rate = cells / el
        out = dict(value=round(len(shots) / (total_cu / rate), 5), unit="shots/s", cores=1, kind="reference",
                   sample="reference C afd2d on shot 0 (%dx%d +30 ABC), %d steps in %.2f s = %.3f Gcell/s, "
                          "extrapolated to the 96 shots; single thread" % (nz, nx, nt_cpu, el, rate / 1e9))
        # shots are independent: the same sample in CPU_PROCS processes at once
        procs = min(CPU_PROCS, os.cpu_count() or 1)
        elp = run_concurrently(
            lambda: refrun.afd2d(s0["vel"], nt_cpu, s0["dt"], s0["d"][1], s0["d"][0], 30, 0.01, src, 1)[1], procs)
        prate = procs * cells / elp
        out["parallel"] = dict(value=round(len(shots) / (total_cu / prate), 5), unit="shots/s", procs=procs,
                               cpu_model=cpu_model(),
                               sample="%d concurrent reference processes, each the same shot-0 sample; slowest %.2f s "
                                      "= %.3f Gcell/s aggregate" % (procs, elp, prate / 1e9))
        return out
    except Exception as e:
        return dict(value=None, unit="shots/s", cores=1, kind="reference", sample="failed: %s" % e)


def afd2d_small_bench():
    """The reference's own CPU-sized 2D case (101x101 three-layer model + 30-cell ABC = 161x161,
    nt=1000, 20 Hz source, jsnap=999): launch-bound on a GPU, so reported as the latency of the
    whole modelling run (kernels only, fields resident) and of the drop-in afd2d call (host arrays
    in and out), beside the reference C on one core."""
    import pywave_amd
    from pywave_amd import AfdPlan
    from pywave_amd.workloads import plain2d

    w = plain2d()
    vel = w["vel"]
    plan = AfdPlan(vel, w["d"], w["nt"], w["dt"], w["nbt"], w["ct"], w["jsnap"])
    plan.set_sources(w["src"], w["ns"])
    plan.run()
    total_ms, _ = plan.run_timed()  # plain launches
    _, step_ms = plan.run_timed(per_step=True)  # per-step timing events (slower run)
    plan.close()
    src = w["src"]
    kw = dict(nt=w["nt"], dt=w["dt"], ax=[0, 10, 101], az=[0, 10, 101], ns=1, sx=[src[0]], sz=[src[1]], f=[src[2]],
              t=[src[3]], A=[src[4]], nbt=w["nbt"], ct=w["ct"], jsnap=w["jsnap"], verb=False)
    pywave_amd.afd2d(vel, **kw)
    t = time.perf_counter()
    pywave_amd.afd2d(vel, **kw)
    call_ms = (time.perf_counter() - t) * 1e3
    cells = 161 * 161 * w["nt"]
    out = {"workload": "afd2d 101x101 three layers (+30 ABC: 161x161), nt=1000, jsnap=999",
           "run_ms": round(total_ms, 3), "kernel_us_per_step": round(step_ms * 1e3, 3),
           "dropin_call_ms": round(call_ms, 3), "Gcell_per_s_run": round(cells / (total_ms * 1e-3) / 1e9, 3)}
    try:
        from oracle import refrun
        if refrun.available():
            pack = np.array(w["src"], np.float32)
            _, el = refrun.afd2d(vel, w["nt"], w["dt"], 10.0, 10.0, w["nbt"], w["ct"], pack, 1, jsnap=w["jsnap"])
            out["cpu_baseline"] = dict(value=round(el * 1e3, 3), unit="ms per run", cores=1, kind="reference",
                                       sample="reference C afd2d (fdm2d), the whole 1000-step run, single thread")
    except Exception as e:
        out["cpu_baseline"] = dict(value=None, unit="ms per run", cores=1, kind="reference", sample="failed: %s" % e)
    return out


def lstri_bench():
    """Passive modelling and least-squares time-reversal source inversion (pfwi mode 4 lstric) on
    the geometry of the reference's passive demos: 130x334 (+60-cell decay: 250x454), nt=3001,
    3 events.  Modelling advances the 3 shots in one launch per step; the inversion runs 30 CG
    iterations per shot (an adjoint and a modelling sweep each, replayed from hipGraphs), shots
    one after the other as the reference's carried CG state requires.  Bytes per step and shot:
    16 B per padded cell (two levels and v^2dt^2 read, one level written) + 4 B per model cell
    (source plane read, or image plane written) + 4 B per receiver."""
    from pywave_amd import LstriProblem
    from pywave_amd.workloads import lstri2d

    w = lstri2d()
    par = w["par"]
    nz, nx, nt, ns, nb, niter = par["nz"], par["nx"], par["nt"], par["ns"], par["nb"], par["niter"]
    cells = (nz + 2 * nb) * (nx + 2 * nb)
    step_bytes = 16.0 * cells + 4.0 * nz * nx + 4.0 * nx
    p = LstriProblem(par, w["vel"])
    p.model(w["src"])
    data = p.model(w["src"])
    model_ms = p.timing()
    model_step_us = model_ms * 1e3 / nt
    p.invert(data, niter=1)  # captures the sweeps' hipGraphs
    t = time.perf_counter()
    src_inv, its = p.invert(data, return_iters=True)
    inv_s = time.perf_counter() - t
    inv_dev_ms = p.timing()
    (inv_w, inv_spl, inv_cols), inv_log = p.launch_choice()
    fit = p.model(src_inv)
    (mod_w, mod_spl, mod_cols), mod_log = p.launch_choice()
    resid = float(np.linalg.norm(fit - data) / np.linalg.norm(data))
    n_sweep_steps = 2 * nt * int(np.sum(its))
    out = {"workload": "lstri2d 130x334 (+60 decay: 250x454), nt=3001, 3 events x 6 buried sources, "
                       "30 CG iterations per shot",
           "model_ms": round(model_ms, 3), "model_step_us": round(model_step_us, 3),
           # the shots of one launch share v^2dt^2: 12 B per shot and padded cell + 4 B per cell
           "model_achieved_GBs": round((ns * (step_bytes - 4.0 * cells) + 4.0 * cells) / (model_step_us * 1e-6) / 1e9, 1),
           "invert_s": round(inv_s, 3), "invert_device_ms": round(inv_dev_ms, 1),
           "invert_ms_per_cg_iteration": round(inv_dev_ms / max(int(np.sum(its)), 1), 2),
           "invert_us_per_sweep_step": round(inv_dev_ms * 1e3 / max(n_sweep_steps, 1), 3),
           "cg_iterations": [int(x) for x in its], "relative_residual": round(resid, 4),
           "launch_choice": {"invert_one_shot": {"tile_width": inv_w, "steps_per_launch": inv_spl,
                                                 "four_step_columns": inv_cols, "timing_ms_per_launch": inv_log},
                             "model_all_shots": {"tile_width": mod_w, "steps_per_launch": mod_spl,
                                                 "four_step_columns": mod_cols, "timing_ms_per_launch": mod_log}}}
    p.close()
    try:
        from oracle import refrun
        if refrun.available():
            pr = np.array([par[k] for k in ["nz", "nx", "dz", "dx", "oz", "ox", "nt", "dt", "ot", "inv", "ns", "ds",
                                            "os", "sz", "nb", "coef", "f0", "acqui_type", "interval", "niter"]],
                          np.float32)
            pr[10] = 1  # one shot
            q = np.full((nz, nx), 1e4, np.float32)
            wav = np.zeros(nt, np.float32)
            _, el_mod = refrun.lstri(pr, w["vel"], q, wav, w["src"][..., :1])
            pr_i = pr.copy()
            pr_i[9], pr_i[19] = 1, 1
            _, el_it = refrun.lstri(pr_i, w["vel"], q, wav, data[..., :1])
            cpu_inv = el_it * niter * ns
            out["cpu_baseline"] = dict(
                value=round(cpu_inv, 1), unit="s per inversion (3 shots x 30 CG iterations)", cores=1,
                kind="reference",
                sample="reference C lstri: one shot's modelling (%.2f s) and one CG iteration of one shot "
                       "(%.2f s, adjoint + modelling sweep + CG step), extrapolated x30 iterations x3 shots; "
                       "single thread, -O2 -fwrapv" % (el_mod, el_it),
                model_s_per_shot=round(el_mod, 3))
            out["invert_speedup_vs_cpu"] = round(cpu_inv / inv_s, 1)
    except Exception as e:
        out["cpu_baseline"] = dict(value=None, unit="s per inversion", cores=1, kind="reference",
                                   sample="failed: %s" % e)
    return out


def aps3d_bench(cpu_steps=20):
    """Pseudo-spectral 3D modelling on the reference demo's configuration (demos/test_first.py:
    81^3 at 20 m, v = 1500 + 24 z m/s, nbt=30 -> 141^3 updated, 144^3 padded FFT grid, nt=1501,
    aps3d's default three sources).  Per step: real z transforms, x and y transforms (the y pass
    fused with the wavenumber multiply and its inverse), inverse x and z, then the leapfrog update.
    Bytes per step and padded cell as the passes move them: z forward 8, x forward 16, y fused 20,
    x inverse 16, z inverse 12 (f32 in, complex half-spectrum out and back), update 20 per updated
    cell (u_n, u_{n-1}, vv, Laplacian read; u_{n+1} written)."""
    from pywave_amd import AfdPlan

    n, nbt, nt, dt, d = 81, 30, 1501, 0.001, 20.0
    v = (np.arange(n) * 20 * 1.2 + 1500).astype(np.float32)
    vel = np.ascontiguousarray(np.broadcast_to(v[:, None, None], (n, n, n)).astype(np.float32))
    pack = np.array([30, 40, 50, 30, 40, 50, 30, 40, 50, 10, 10, 10, 0.2, 0.35, 0.5, 1, 2, 2], np.float32)
    plan = AfdPlan(vel, (d, d, d), nt, dt, nbt, 0.01, 0, spectral=True)
    plan.set_sources(pack, 3)
    plan.run()
    ms, _ = plan.run_timed()
    out = plan.output()
    assert np.isfinite(out).all()
    plan.close()
    m = n + 2 * nbt
    n1 = n2 = n3 = 144
    nk = n1 // 2 + 1
    cells_p = n1 * n2 * n3
    spec = nk * n2 * n3 * 8.0  # bytes of the half spectrum
    step_bytes = (4.0 * cells_p + spec) + 2 * spec + (2 * spec + 4.0 * nk * n2 * n3) + 2 * spec + \
                 (spec + 4.0 * cells_p) + 20.0 * m ** 3
    step_us = ms * 1e3 / nt
    res = {"workload": "aps3d 81^3 at 20 m (+30 ABC: 141^3, FFT grid 144^3), nt=1501, 3 sources (demos/test_first.py)",
           "run_ms": round(ms, 2), "step_us": round(step_us, 2),
           "Gcell_per_s": round(m ** 3 * nt / (ms * 1e-3) / 1e9, 2),
           "algorithmic_bytes_per_step": step_bytes,
           "achieved_GBs": round(step_bytes / (step_us * 1e-6) / 1e9, 1)}
    res["frac_hbm"] = round(res["achieved_GBs"] / PEAK_HBM_GBS, 4)
    try:
        from oracle import refrun
        if cpu_steps > 0 and refrun.available():
            _, el = refrun.afd3d(vel, cpu_steps, dt, d, d, d, nbt, 0.01, pack, 3, ps=True)
            per = el / cpu_steps
            res["cpu_baseline"] = dict(value=round(per * nt, 1), unit="s per 1501-step run", cores=1,
                                       kind="reference",
                                       sample="reference C psm (pywave/src/wave_psp.c:25) on the same model, %d of "
                                              "1501 steps in %.2f s, extrapolated; single thread" % (cpu_steps, el))
            res["speedup_vs_cpu"] = round(per * nt / (ms * 1e-3), 1)
    except Exception as e:
        res["cpu_baseline"] = dict(value=None, unit="s per run", cores=1, kind="reference", sample="failed: %s" % e)
    return res


def fwi_set_widths(prob, widths):
    """widths: None (the plan's timing sweep) or (forward, exact adjoint, reference adjoint)."""
    if widths:
        prob.set_widths(forward=widths[0], exact_adjoint=widths[1], reference_adjoint=widths[2])


def fwi_forward_bench(nt=300, widths=None):
    """The acoustic forward of the FWI workload (400x800 +60 sponge, 32 shots in one launch per
    step) over its first nt steps, twice: a short run of the FWI step kernel for counter passes."""
    from pywave_amd import FwiProblem
    from pywave_amd.workloads import fwi2d

    w = fwi2d(nt=nt)
    par = w["par"]
    wav = (w["amp"] * ricker_wavelet(nt, par["dt"], w["f0"], w["t0"])).astype(np.float32)
    p = FwiProblem(par, wav, batch=par["ns"])
    fwi_set_widths(p, widths)
    p.forward(w["vtrue"])
    p.forward(w["vtrue"])
    ms = p.timing()[0]
    p.close()
    grid = (par["nz"] + 2 * par["nb"]) * (par["nx"] + 2 * par["nb"])
    # per shot and padded cell: p1, p0 read, p2 written (12 B); v^2dt^2 is one field for all shots
    alg = 12.0 * grid * par["ns"] + 4.0 * grid + 4.0 * par["nx"] * par["ns"]
    return {"workload": "fwi2d acoustic forward, 32 shots, first %d steps" % nt, "device_ms": round(ms, 2),
            "step_us": round(ms * 1e3 / nt, 2), "algorithmic_bytes_per_step": alg,
            "achieved_GBs": round(alg / (ms * 1e-3 / nt) / 1e9, 1)}


def fwi_gradient_bench(nt=300, widths=None):
    """The FWI workload's 32 shots (400x800 +60 sponge, all shots per launch, every step's forward
    fields kept in HBM as in the full workload) over its first nt steps: the exact-adjoint gradient
    and the reference gradient (gradient_av), two of each -- short runs of the adjoint kernels for
    counter passes."""
    from pywave_amd import FwiProblem
    from pywave_amd.workloads import fwi2d

    w = fwi2d(nt=nt)
    par = w["par"]
    wav = (w["amp"] * ricker_wavelet(nt, par["dt"], w["f0"], w["t0"])).astype(np.float32)
    gen = FwiProblem(par, wav, batch=par["ns"])
    obs = gen.forward(w["vtrue"])
    gen.close()
    p = FwiProblem(par, wav, data=obs, batch=par["ns"])
    fwi_set_widths(p, widths)
    out = {"workload": "fwi2d gradients, 32 shots, first %d steps, forward fields kept in HBM" % nt,
           "widths": p.widths()}
    for mode in ("exact", "reference"):
        for _ in range(2):
            p.gradient(w["v0"], mode)
        f, b = p.timing()
        out[mode] = {"forward_ms": round(f, 2), "backward_ms": round(b, 2)}
    p.close()
    return out


def visco_bench(cpu_steps=300):
    """Visco-acoustic forward modelling (forward_modeling, pywave/src/wave_fwimodeling.c:116-220,
    the media=2 modelling of pfwi) on the FWI workload's 400x800, 32-shot, nt=3000 acquisition
    with a seeded Q of 20-100: all 32 shots per launch.  Per shot, padded cell and step the kernel
    reads p_n, p_{n-1}, r_n and writes p_{n+1}, r_{n+1} (20 algorithmic bytes); v^2 dt^2, tau and
    taus are one field each for all shots (12 B per padded cell per step).
    The acoustic forward of the same plan is timed beside it, and the visco-acoustic imaging
    gradient (gradient_v, pywave/src/wave_fwigradient.c:264-452) of all 32 shots with checkpoints
    every 50 steps."""
    from pywave_amd import FwiProblem
    from pywave_amd.workloads import fwi2d

    w = fwi2d()
    par = w["par"]
    nt, ns, nz, nx, nb = par["nt"], par["ns"], par["nz"], par["nx"], par["nb"]
    rng = np.random.default_rng(12)
    q = (20.0 + 80.0 * rng.random((nz, nx))).astype(np.float32)
    wav = (w["amp"] * ricker_wavelet(nt, par["dt"], w["f0"], w["t0"])).astype(np.float32)
    prob = FwiProblem(par, wav, batch=ns)
    prob.forward(w["vtrue"])  # warm-up, acoustic
    t = time.perf_counter()
    prob.forward(w["vtrue"])
    ac_s = time.perf_counter() - t
    ac_dev_ms = prob.timing()[0]
    prob.set_visco(q)
    prob.forward(w["vtrue"])  # warm-up, visco
    t = time.perf_counter()
    d = prob.forward(w["vtrue"])
    vs_s = time.perf_counter() - t
    vs_dev_ms = prob.timing()[0]
    prob.close()
    assert np.isfinite(d).all()
    # gradient_v of the smooth start model against the visco data, every shot in one batch
    gp = FwiProblem(par, wav, data=d, batch=ns, checkpoint=50)
    gp.set_visco(q)
    gp.gradient(w["v0"], "reference")  # warm-up (sets the first-call scaling)
    t = time.perf_counter()
    gv, fv, _ = gp.gradient(w["v0"], "reference")
    gr_s = time.perf_counter() - t
    gr_f, gr_b = gp.timing()
    gp.close()
    assert np.isfinite(gv).all()
    grid = (nz + 2 * nb) * (nx + 2 * nb)
    cells = grid * ns
    step_us = vs_dev_ms * 1e3 / nt
    step_bytes = 20.0 * cells + 12.0 * grid + 4.0 * nx * ns
    achieved = step_bytes / (step_us * 1e-6) / 1e9
    out = {"workload": "visco-acoustic forward modelling on fwi2d 400x800 at 10 m (+60-cell sponge: 520x920), "
                       "32 shots in one launch, nt=3000, dt=1 ms, 8 Hz, Q 20-100 seeded",
           "shots_per_s": round(ns / vs_s, 3), "forward_s": round(vs_s, 4), "device_ms": round(vs_dev_ms, 2),
           "step_us": round(step_us, 2), "Gcell_per_s": round(cells * nt / (vs_dev_ms * 1e-3) / 1e9, 2),
           "algorithmic_bytes_per_step": step_bytes, "achieved_GBs": round(achieved, 1),
           "frac_hbm": round(achieved / PEAK_HBM_GBS, 4),
           "acoustic_shots_per_s": round(ns / ac_s, 3), "acoustic_device_ms": round(ac_dev_ms, 2),
           "gradient_v_shots_per_s": round(ns / gr_s, 3), "gradient_v_s": round(gr_s, 4),
           "gradient_v_forward_ms": round(gr_f, 1), "gradient_v_backward_ms": round(gr_b, 1), "gradient_v_misfit": fv}
    out["cpu_baseline"] = None
    if cpu_steps > 0:
        try:
            from oracle import refrun
            if refrun.available():
                from pywave_amd.fwi import make_pararray
                p1 = dict(par)
                p1["ns"], p1["nt"] = 1, cpu_steps
                wav1 = (w["amp"] * ricker_wavelet(cpu_steps, par["dt"], w["f0"], w["t0"])).astype(np.float32)
                _, el = refrun.fwimod(make_pararray(p1), w["vtrue"], q, wav1, visco=True)
                per_shot = el * nt / cpu_steps
                out["cpu_baseline"] = dict(
                    value=round(1.0 / per_shot, 5), unit="shots/s", cores=1, kind="reference",
                    sample="reference C forward_modeling (visco), one shot of the same geometry, %d of %d steps in "
                           "%.2f s, extrapolated to %.1f s per shot; single thread, -O2 -fwrapv"
                           % (cpu_steps, nt, el, per_shot))
                out["speedup_vs_cpu"] = round(out["shots_per_s"] * per_shot, 1)
                d1 = np.zeros((cpu_steps, nx, 1), np.float32)
                _, _, elg = refrun.fwigrad(make_pararray(p1), w["v0"], q, wav1, d1, visco=True)
                per_shot_g = elg * nt / cpu_steps
                out["gradient_v_cpu_baseline"] = dict(
                    value=round(1.0 / per_shot_g, 5), unit="shots/s", cores=1, kind="reference",
                    sample="reference C gradient_v, one shot of the same geometry, %d of %d steps in %.2f s, "
                           "extrapolated to %.1f s per shot; single thread, -O2 -fwrapv"
                           % (cpu_steps, nt, elg, per_shot_g))
        except Exception as e:  # reported, never required
            out["cpu_baseline"] = dict(value=None, unit="shots/s", cores=1, kind="reference", sample="failed: %s" % e)
    return out


def tri_bench(cpu_steps=12):
    """Time-reversal imaging on the 128^3 workload (+30-cell ABC: 188^3 updated, nt=2000): the
    forward modelling of three sources, then the four interleaved receiver groups back-propagated
    together with the image sum_snap prod_g u_g^2 (jsnap=4) accumulated on the GPU.  Bytes per
    step: G fields x (read u_n, u_{n-1}, write u_{n+1}) + one vv read = (12 G + 4) B per cell,
    plus the f64 image read-modify-write (16 B per interior cell) at every snapshot step."""
    from pywave_amd import AfdPlan, tri3d_image
    from pywave_amd.workloads import tri3d

    w = tri3d()
    vel = w["vel"]
    n = vel.shape[0]
    nb = w["nbt"]
    plan = AfdPlan(vel, w["d"], w["nt"], w["dt"], nb, w["ct"], 0)
    plan.set_sources(w["src"], w["ns"])
    fwd_variant = list(plan.autotune(steps=5))
    fwd_ms, _ = plan.run_timed()
    dat = plan.output().reshape(w["nt"], n, n, order="F")
    plan.close()
    ax = [0, w["d"][1], n]
    launch = {}
    _, ms = tri3d_image(vel, dat, w["groups"], dt=w["dt"], ax=ax, ay=ax, az=ax, nbt=nb, ct=w["ct"],
                        jsnap=w["jsnap"], return_plan_time=True, launch_info=launch)
    G = len(w["groups"])
    cells = (n + 2 * nb) ** 3
    nsnap = (w["nt"] - 1) // w["jsnap"] + 1
    alg = w["nt"] * cells * (12.0 * G + 4.0) + nsnap * 16.0 * n ** 3
    out = {"workload": "tri3d 128^3 (+30 ABC: 188^3), nt=2000, 3 sources, 4 interleaved receiver groups, jsnap=4",
           "forward_ms": round(fwd_ms, 2), "forward_kernel_variant": fwd_variant, "backprop_image_ms": round(ms, 2),
           "field_cell_updates_per_s_G": round(G * cells * w["nt"] / (ms * 1e-3) / 1e9, 2),
           "achieved_GBs": round(alg / (ms * 1e-3) / 1e9, 1),
           "frac_hbm": round(alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
           "backprop_rows_per_block": launch.get("rows_per_block"), "backprop_y_chunk": launch.get("y_chunk"),
           "backprop_non_temporal": launch.get("non_temporal"),
           "backprop_timing_ms_per_launch": launch.get("timing_ms_per_launch")}
    if cpu_steps > 0:
        out["cpu_baseline"] = tri_cpu_baseline(w, dat, cpu_steps, ms)
    return out


def tri_cpu_baseline(w, dat, nt_cpu, gpu_ms):
    """The reference recipe (demos/test_third_tri.py:57-77) on the CPU: one group's time reversal
    through the reference C afd3d (tri=True, jsnap=4; fdm, pywave/src/wave_fdm.c:28-381) over
    nt_cpu of the 2000 steps, and the image formula sum_snap prod_g u_g^2 in numpy on that run's
    snapshots; both extrapolated to 4 groups x 2000 steps (the per-step cost does not depend on
    the step).  Single thread."""
    try:
        from oracle import refrun
        if not refrun.available():
            return None
        vel, n, nb, js = w["vel"], w["vel"].shape[0], w["nbt"], w["jsnap"]
        G, nt = len(w["groups"]), w["nt"]
        g0 = np.zeros((nt_cpu, n, n), np.float32)
        g0[:, w["groups"][0], :] = dat[nt - nt_cpu:, w["groups"][0], :]  # the group's last samples
        out, el = refrun.afd3d(vel, nt_cpu, w["dt"], *w["d"][1:], w["d"][0], nb, w["ct"], tri=True, dat=g0, jsnap=js)
        wfd = out[nt_cpu * n * n:].reshape(n, n, n, -1, order="F")
        t = time.perf_counter()
        img = np.zeros((n, n, n), np.float64)
        for k in range(wfd.shape[3]):
            prod = np.ones((n, n, n), np.float64)
            for _ in range(G):  # the same group's field stands in for each of the G groups
                prod *= wfd[:, :, :, k].astype(np.float64) ** 2
            img += prod
        el_img = time.perf_counter() - t
        nsnap_cpu = wfd.shape[3]
        total = G * el * nt / nt_cpu + el_img * ((nt - 1) // js + 1) / nsnap_cpu
        return dict(value=round(total, 1), unit="s per image (4 group back-propagations + image)", cores=1,
                    kind="reference",
                    sample="reference C afd3d(tri=True, jsnap=4) of one group on the 128^3 model, %d of %d steps in "
                           "%.2f s, and the image formula on its %d snapshots in %.3f s (numpy), extrapolated to %d "
                           "groups x %d steps; single thread, -O2 -fwrapv" % (nt_cpu, nt, el, nsnap_cpu, el_img, G, nt),
                    speedup_vs_gpu_backprop_image=round(total / (gpu_ms * 1e-3), 1))
    except Exception as e:  # reported, never required
        return dict(value=None, unit="s per image", cores=1, kind="reference", sample="failed: %s" % e)


def batch_bench():
    """96 ragged 2D shots (nz 200-600, nx 400-1600, nt 2000-4000) advanced in one launch per step,
    the step loop replayed from a hipGraph; 16 B per cell-update (+ the trace sample)."""
    from pywave_amd import Afd2dBatch
    from pywave_amd.workloads import batch2d

    shots = batch2d()
    b = Afd2dBatch(shots)
    b.run(use_graph=True)  # warm-up (graph capture)
    ms = b.run(use_graph=True)
    cu = b.cell_updates()
    width, wlog = b.width()
    b.close()
    nrec = sum(s["vel"].shape[1] * s["nt"] for s in shots)
    alg = 16.0 * cu + 4.0 * nrec
    cpu = batch_cpu_baseline(shots, cu)
    return {"workload": "batch2d 96 shots, nz 200-600, nx 400-1600 (+30 ABC), nt 2000-4000, 12 Hz, hipGraph",
            "ms": round(ms, 2), "shots_per_s": round(len(shots) / (ms * 1e-3), 2),
            "cell_updates": cu, "Gcell_per_s": round(cu / (ms * 1e-3) / 1e9, 2),
            "achieved_GBs": round(alg / (ms * 1e-3) / 1e9, 1),
            "frac_hbm": round(alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "tile_width": width, "width_timing_ms_per_launch": {str(w): t for w, t in wlog},
            "cpu_baseline": cpu}


def big3d_bench(n=261, nt=200):
    """The 3D step on a grid whose fields do not fit the 256 MB Infinity Cache: the headline model
    recipe at 261^3 + 30-cell ABC = 321^3 updated (324^3 padded, 3 fields = 408 MB), a 15 Hz source
    at the centre, surface receivers, nt steps.  The autotuned kernel, the fastest single-step and
    each two-step (temporally blocked) variant, by dispatch-stamped HIP events per launch."""
    from pywave_amd import AfdPlan
    from pywave_amd import _native as N
    from pywave_amd.workloads import plain3d_model

    vel = plain3d_model(n)
    c = n // 2
    nb = 30
    plan = AfdPlan(vel, (10.0, 10.0, 10.0), nt, 0.001, nb, 0.01, 0)
    plan.set_sources(np.array([c, c, c, 15.0, 0.05, 1.0], np.float32), 1)
    L = N.lib()
    tuned = plan.autotune(steps=6)
    cells = (n + 2 * nb) ** 3
    alg_step = 16.0 * cells + 4.0 * n * n
    res = {"workload": "afd3d %d^3 model + 30-cell ABC = %d^3 updated (324^3 padded: 3 fields of more than "
                       "136 MB, beyond the 256 MB Infinity Cache), nt=%d" % (n, n + 2 * nb, nt),
           "autotuned": list(tuned), "variants": []}
    info = np.zeros(4, np.int32)
    for v in range(L.pw_afd_variant_count()):
        N.check(L.pw_afd_variant_info(v, N.ptr(info)), None, "variant_info")
        spl = int(info[3])
        if v != tuned[0] and spl == 1:
            continue  # single-step variants: the tuned one only
        try:
            plan.set_variant(v, tuned[1] if v == tuned[0] else 0)
            plan.run()
            total_ms, step_ms = plan.run_timed(per_step=True)
        except (ValueError, N.PywaveError):
            continue  # not applicable to this grid
        ach = alg_step / (step_ms * 1e-3) / 1e9
        res["variants"].append({"variant": v, "time_steps_per_launch": spl, "us_per_time_step": round(step_ms * 1e3, 2),
                                "run_ms": round(total_ms, 2), "Gcell_per_s": round(cells / (step_ms * 1e-3) / 1e9, 1),
                                "achieved_GBs": round(ach, 1), "frac_hbm": round(ach / PEAK_HBM_GBS, 4)})
    plan.close()
    return res


def ricker_wavelet(nt, dt, f0, t0):
    from pywave_amd import _native as N
    wv = np.zeros(nt, np.float32)
    N.check(N.lib().pw_ricker(nt, dt, f0, t0, 1.0, N.ptr(wv)), None, "ricker")
    return wv


def calibrate_copy():
    """Copies of 2^28 float32 (1 GiB read + 1 GiB written): the known traffic against which
    FETCH_SIZE / WRITE_SIZE counter readings are scaled, and the device's practical copy
    bandwidth (read + write bytes / time, 10 timed copies)."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 10
    print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4,
                      "copy_ms": round(ms, 4), "copy_GBs": round(2 * a.numel() * 4 / (ms * 1e-3) / 1e9, 1),
                      "stream": stream_bandwidth()}))


def sizes_bench(sizes=(101, 141, 181, 201, 221, 241, 261), nt=200):
    """The autotuned 3D step against the grid size (the headline model recipe at n^3 + 30-cell
    ABC, a centred 15 Hz source, surface receivers, nt steps): where the three fields stop fitting
    the 256 MB Infinity Cache, the rate falls towards the device's streaming rate."""
    import ctypes
    from pywave_amd import AfdPlan
    from pywave_amd import _native as N
    from pywave_amd.workloads import plain3d_model

    out = []
    for n in sizes:
        nb = 30
        m = n + 2 * nb
        z2, x2, y2 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        N.lib().pw_padded_sizes(m, m, m, ctypes.byref(z2), ctypes.byref(x2), ctypes.byref(y2))
        c = n // 2
        plan = AfdPlan(plain3d_model(n), (10.0, 10.0, 10.0), nt, 0.001, nb, 0.01, 0)
        plan.set_sources(np.array([c, c, c, 15.0, 0.05, 1.0], np.float32), 1)
        tuned = plan.autotune(steps=6)
        plan.run()
        _, step_ms = plan.run_timed(per_step=True)
        plan.close()
        cells = m ** 3
        ach = (16.0 * cells + 4.0 * n * n) / (step_ms * 1e-3) / 1e9
        out.append({"model": n, "updated": m, "padded": [z2.value, x2.value, y2.value],
                    "three_fields_MB": round(3 * 4.0 * z2.value * x2.value * y2.value / 1e6, 1),
                    "kernel_variant": list(tuned), "us_per_time_step": round(step_ms * 1e3, 2),
                    "Gcell_per_s": round(cells / (step_ms * 1e-3) / 1e9, 1), "achieved_GBs": round(ach, 1),
                    "frac_hbm": round(ach / PEAK_HBM_GBS, 4)})
    return {"workload": "afd3d n^3 models + 30-cell ABC, nt=%d, autotuned kernel per grid" % nt, "grids": out}


def stream_bandwidth(reps=10, n=1 << 27, n_small=12_500_000):
    """The device's plain streaming rate (pw_stream_bandwidth: 1, 2 or 3 input streams and one
    output stream of 2^27 floats, 512 MB each, more than the 256 MB Infinity Cache holds, with
    non-temporal accesses): the practical HBM ceiling against which the step kernels' actual
    traffic is read.  The same mixes with the default cache policy at 512 MB a stream (`cached`)
    and at 50 MB a stream (100-200 MB in all, which the Infinity Cache holds between launches:
    `in_cache`) give the ceiling of kernels whose fields fit that cache."""
    import ctypes
    from pywave_amd import _native as N
    L, ctx = N.lib(), N.context()
    out = {}
    for key, size, cached in (("", n, 0), ("_cached", n, 1), ("_in_cache", n_small, 1)):
        for reads in (1, 2, 3):
            g = ctypes.c_double()
            N.check(L.pw_stream_bandwidth(ctx, size, reads, reps, cached, ctypes.byref(g)), ctx, "stream_bandwidth")
            out["read%d_write1%s_GBs" % (reads, key)] = round(g.value, 1)
    return out


def load_traffic(name, variant):
    """HBM bytes per launch of the step kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    runs (scaled by the copy calibration), if they were taken on the kernel variant in use."""
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        for e in d.get("variants", []):
            if variant is not None and list(e["variant"]) == list(variant):
                return e["hbm_bytes_per_launch"]
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--ny-chunk", type=int, default=0)
    ap.add_argument("--nt", type=int, default=1500)
    ap.add_argument("--cpu-steps", type=int, default=25)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fwi", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-tri", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--no-lstri", action="store_true")
    ap.add_argument("--no-aps", action="store_true")
    ap.add_argument("--no-visco", action="store_true")
    ap.add_argument("--no-big", action="store_true")
    ap.add_argument("--fwi-plain-shots", type=int, default=4)
    ap.add_argument("--fwi-widths", type=lambda t: tuple(int(x) for x in t.split(",")), default=None,
                    help="forward,exact-adjoint,reference-adjoint tile widths of the batched FWI plans "
                         "(default: each plan's timing sweep); for counter runs, where the sweep sees "
                         "serialised launches")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (HBM counter calibration)")
    ap.add_argument("--only", choices=["fwi", "tri", "batch", "small", "lstri", "aps", "visco", "fwifwd", "fwigrad",
                                       "big3d", "sizes"],
                    help="run one secondary workload alone and print its JSON object (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: pywave_amd has no multi-GPU path (requested --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.only:
        leg = {"fwi": lambda: fwi_bench(args.fwi_plain_shots, args.fwi_widths), "tri": lambda: tri_bench(0 if args.no_cpu else 12),
               "batch": batch_bench,
               "small": afd2d_small_bench, "lstri": lstri_bench,
               "aps": lambda: aps3d_bench(0 if args.no_cpu else 20),
               "visco": lambda: visco_bench(0 if args.no_cpu else 300),
               "fwifwd": lambda: fwi_forward_bench(widths=args.fwi_widths),
               "fwigrad": lambda: fwi_gradient_bench(widths=args.fwi_widths), "big3d": big3d_bench,
               "sizes": sizes_bench}[args.only]
        print(json.dumps({args.only: leg()}))
        return

    from pywave_amd import AfdPlan
    from pywave_amd import _native as N
    from pywave_amd.workloads import plain3d

    w = plain3d(nt=args.nt)
    nz, nx, ny = w["vel"].shape
    nb = w["nbt"]
    cells = (nz + 2 * nb) * (nx + 2 * nb) * (ny + 2 * nb)
    plan = AfdPlan(w["vel"], w["d"], w["nt"], w["dt"], nb, w["ct"], 0)
    plan.set_sources(w["src"], w["ns"])
    tuned = None
    if args.variant >= 0:
        plan.set_variant(args.variant, args.ny_chunk)
    elif not args.no_tune:
        tuned = plan.autotune(steps=10)
    ctx = N.context()
    L = N.lib()
    used = [args.variant, args.ny_chunk] if args.variant >= 0 else (list(tuned) if tuned else None)

    for _ in range(args.warmup):
        plan.run()
    N.check(L.pw_synchronize(ctx), ctx, "sync")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run()
    N.check(L.pw_synchronize(ctx), ctx, "sync")
    t1 = time.perf_counter()
    sec = (t1 - t0) / args.steps
    value = cells * w["nt"] / sec / 1e9

    # dominant kernel: the fused step; per-launch duration from HIP events around every step
    # launch on the library stream (one extra run)
    total_ms, step_ms = plan.run_timed(per_step=True)
    nrec = nx * ny
    # per time step: read u_n, u_{n-1}, vv, write u_{n+1}; one trace sample per receiver
    alg_step = 16.0 * cells + 4.0 * nrec
    spl = 1  # time steps per launch of the step kernel in use
    if used is not None:
        info = np.zeros(4, np.int32)
        N.check(L.pw_afd_variant_info(int(used[0]), N.ptr(info)), ctx, "variant_info")
        spl = int(info[3])
    launch_ms = step_ms * spl  # step_ms is the dispatch-stamped kernel time per time step
    alg_bytes = alg_step * spl
    achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
    out = plan.output()
    assert np.isfinite(out).all()

    cpu = None if args.no_cpu else cpu_baseline(w, args.cpu_steps)
    fwi = None if args.no_fwi else fwi_bench(args.fwi_plain_shots)
    tri = None if args.no_tri else tri_bench(0 if args.no_cpu else 12)
    batch = None if args.no_batch else batch_bench()
    small = None if args.no_small else afd2d_small_bench()
    lstri = None if args.no_lstri else lstri_bench()
    aps = None if args.no_aps else aps3d_bench()
    visco = None if args.no_visco else visco_bench()
    big = None if args.no_big else big3d_bench()
    line = {
        "metric": "Gcell-updates/s 3D FP32 FD (% HBM roofline); FWI grad shots/s batched vs plain",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3),  # one step = one whole nt-step modelling run
        "ms_per_time_step": round(sec * 1e3 / w["nt"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded model, no external data)",
        "config": {
            "workload": "afd3d plain 3D: model 201^3 at 10 m + 30-cell ABC = 261^3 updated (270^3 padded), "
                        "v 1500->3500 m/s linear in depth x 8 seeded Gaussian perturbations, nt=1500, dt=1 ms, "
                        "15 Hz Ricker at the centre, 201x201 surface receivers",
            "cells_per_time_step": cells,
            "time_steps": w["nt"],
            "kernel_variant": used,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": load_traffic("afd3d_step_traffic.json", used),
            "kernel_avg_ms": round(launch_ms, 5),
            "time_steps_per_launch": spl,
            "algorithmic_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": cpu,
        # the plain streaming ceiling of this device (three reads and one write, like the step
        # kernel): the step kernel's actual traffic rate is read against it
        "stream_GBs": stream_bandwidth(),
        "fwi": fwi,
        "tri3d": tri,
        "batch2d": batch,
        "afd2d_small": small,
        "lstri2d": lstri,
        "aps3d": aps,
        "visco2d": visco,
        "plain3d_321": big,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
