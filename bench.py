"""Benchmark of the MI355X KORAL hot path.  Prints ONE JSON line.

Default workload (--workload grmhd): the 3-D GRMHD RK2IMEX step of KORAL on
the MKS1 Kerr (a = 0.9) magnetised torus at 128x128x64 (MP5, HLL, flux-CT,
Noble u2p), FP64, device-resident, in zone-updates/s.  The field + radiation
step is refused by this build (see the C-ABI's KORAL_E_STATE), so the
GRRMHD+M1 zone-update rate is not measured.  Other workloads:
  implicit  -- batched implicit radiation-source solve, 2^22 seeded cells,
               plain guess (implicit solves/s)
  learned   -- the same solve with the learned initial guess, trained here
               on pairs from the plain solver, against the plain solve
  grrmhd    -- the field + radiation torus step (refused)

The line carries:
  roofline      -- for the dominant kernel: algorithmic bytes (or FP64
                   operations) per launch over its HIP-event-measured duration
  cpu_baseline  -- the CPU oracle timed on a bounded sample on the host cores
                   (rank 0, N=1 only)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s (3D GRRMHD+M1, FP64); implicit solves/s learned vs plain"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)

# HBM bytes per launch of the grid kernels on the 128x128x64 torus:
# FETCH_SIZE and WRITE_SIZE per dispatch scaled by the same counters on a
# streaming kernel of known size (profiles/hbm_traffic_grmhd_torus_128x128x64.md).
# The face kernels per sweep direction (x1, x2, x3) and schedule (None: not
# measured); the update per variant (tune.update_pre).
GRID_TRAFFIC_128x128x64 = {
    "u2p": 315e6,
    "update": {1: 581e6},
    "flux": {"split": (None, None, None), "fused": (448e6, 880e6, 398e6)},
    "recon": {"split": (None, None, None)},
}

# FP64 operations per zone-update of the GRMHD RK2IMEX step on the
# 128x128x64 torus, counted by the operation-counting build of the oracle
# (oracle/flops, each add, mul, div, sqrt and transcendental one operation;
# tests/test_oracle_flops.py re-counts a smaller grid): the three face sweeps
# (reconstruction, floors, fluxes, HLL) of both explicit stages, u2p with
# its fixups and ghost fill (four calls), update with metric sources, wave
# speeds and the rest of the step.
GRMHD_FLOPS_PER_ZONE = {"sweeps": 16581.7, "u2p": 5243.4, "update": 1356.0, "wavespeeds": 1240.6,
                        "other": 392.0, "step": 24813.8}
# Algorithmic HBM bytes per zone-update of the GRMHD step in the fused
# schedule of the scope table (K1: explicit1 + u2p, K2: explicit2 + u2p; 9
# NV-vectors of FP64 per zone-update and 56 B of flags): the metric counted
# once per (x1, x2) plane as this layout stores it (amortised over NZ = 64,
# about 760 B), and counted per cell in 3-D as KORAL tabulates it (4288 B).
GRMHD_BYTES_PER_ZONE = {"axisymmetric_metric": 760.0, "metric_3d": 4288.0}
# The same count per operator and launch: the x1, x2, x3 sweeps (ghost
# fill excluded; reconstruction, floors, fluxes and HLL of one stage), the
# update, u2p and wave speeds, in FP64 operations per zone of one launch
# (per zone-update divided by the launches of a step: 2 explicit stages, 4
# u2p passes).
GRMHD_FLOPS_PER_LAUNCH_ZONE = {"sweep": (5505.1 / 2, 5517.8 / 2, 5558.8 / 2), "update": 1356.0 / 2,
                               "u2p": 5243.4 / 4, "wavespeeds": 1240.6 / 2}

# Algorithmic FP64 operations per implicit solve of the seeded batch workload
# (all six fallback attempts included), counted by the operation-counting
# build of the CPU restatement (oracle/flops) over 20000 cells of the same
# generator: adds 229464, muls 357533, divs 31007, sqrts 7773, pow/log/exp
# 2163 (each counted as one).  tests/test_oracle_flops.py re-counts a sample.
IMPLICIT_FLOPS_PER_SOLVE = 627940.0


def bench_implicit(args):
    import numpy as np
    import torch
    from koral_amd import params as KP, problems as PR, cell

    P = KP.params_implicit_batch()
    n = args.cells
    pp, g20, G20, dt = PR.implicit_batch_states(P, n, seed=args.seed)
    dev = torch.device("cuda:0")
    Pp0 = cell.to_dev(pp)
    g = cell.to_dev(g20)
    G = cell.to_dev(G20)
    dtd = torch.from_numpy(dt).to(dev)
    U0 = cell.p2u(P, Pp0, g, G)
    U = torch.empty_like(U0)
    Pp = torch.empty_like(Pp0)
    stream = torch.cuda.current_stream()

    def step():
        U.copy_(U0)
        Pp.copy_(Pp0)
        return cell.implicit(P, U, Pp, g, G, dtd)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kt = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        U.copy_(U0)
        Pp.copy_(Pp0)
        ev0.record(stream)
        fl, info = cell.implicit(P, U, Pp, g, G, dtd)
        ev1.record(stream)
        ev1.synchronize()
        kt.append(ev0.elapsed_time(ev1) / 1e3)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    info = info.cpu().numpy()
    ok = info[1] >= 0
    kern = sum(kt) / len(kt)
    # algorithmic bytes per solve: read U(nv) P(nv) g(20) G(20) dt, write U P,
    # flags (3 int) and info (6 int)
    nv = P.nv
    bytes_per = (2 * nv + 40 + 1) * 8 + 2 * nv * 8 + 9 * 4
    achieved_bw = bytes_per * n / kern / 1e9
    achieved = IMPLICIT_FLOPS_PER_SOLVE * n / kern / 1e12
    res = {
        "metric": METRIC,
        "value": n * args.steps / wall,
        "unit": "implicit solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": "batched implicit radiation-source solve, plain guess",
                   "cells": n, "nv": nv, "opacity": "Kramers ff + Thomson",
                   "success_frac": float(ok.mean()),
                   "mean_newton_iters": float(info[0][ok].mean()),
                   "mean_total_iters": float(info[5].mean())},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "compute": "FP64 vector ALU (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
                     "kernel": "implicit solve: k_imp_classify + 12 k_imp_pass launches (one batch)",
                     "kernel_ms": kern * 1e3,
                     "algorithmic_flops_per_launch": IMPLICIT_FLOPS_PER_SOLVE * n,
                     "hbm": {"achieved_GBps": achieved_bw, "frac": achieved_bw / HBM_PEAK_GBS,
                             "algorithmic_bytes_per_launch": bytes_per * n}},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_implicit(P, pp, g20, G20, dt, args)
    return res


def _timed_solves(solve, reset, steps, warmup):
    """Times `steps` solves of one batch (state restored before each by
    `reset`, outside the event pair).  Returns (wall s per solve, event s
    per solve, last (flags, info))."""
    import torch
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        reset()
        solve()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kt, wall, out = [], 0.0, None
    for _ in range(steps):
        reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(stream)
        out = solve()
        ev1.record(stream)
        ev1.synchronize()
        wall += time.perf_counter() - t0
        kt.append(ev0.elapsed_time(ev1) / 1e3)
    return wall / steps, sum(kt) / len(kt), out


def bench_learned(args):
    """Learned-guess implicit solve: train the network on pairs from the plain
    FP64 solver over a seeded training batch, then time the plain and the
    learned solve of a fresh batch of the same generator."""
    import numpy as np
    import torch
    from koral_amd import params as KP, problems as PR, cell
    from koral_amd import learn as LN

    P = KP.params_implicit_batch()
    dev = torch.device("cuda:0")
    say = lambda *a: print("[learned]", *a, file=sys.stderr, flush=True)
    # ---- training pairs from the plain solver ----
    t0 = time.perf_counter()
    pp, g20, G20, dt = PR.implicit_batch_states(P, args.train_cells, seed=args.seed + 1000)
    Pp, g, G, dtd = cell.to_dev(pp), cell.to_dev(g20), cell.to_dev(G20), torch.from_numpy(dt).to(dev)
    del pp, g20, G20, dt
    U = cell.p2u(P, Pp, g, G)
    P0 = Pp.clone()
    _, info = cell.implicit(P, U, Pp, g, G, dtd)
    F, Tg, Tr = LN.training_pairs(P, P0, Pp, dtd, info)
    del U, Pp, info
    torch.cuda.empty_cache()
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t0
    n_pairs = F.shape[1]
    say("%d training pairs from %d cells in %.1f s" % (n_pairs, args.train_cells, t_pairs))
    sched = [(1e-3, args.epochs), (3e-4, max(1, args.epochs // 2)), (1e-4, max(1, args.epochs // 4))]
    t0 = time.perf_counter()
    gas, hist_g = LN.train(P, F, Tg, LN.GAS, schedule=sched)
    rad, hist_r = LN.train(P, F, Tr, LN.RAD, schedule=sched)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    say("trained both networks (%s) in %.1f s, final loss gas %.4f rad %.4f" %
        (sched, t_train, hist_g[-1], hist_r[-1]))
    del F, Tg, Tr
    # the gate: trained on whether the learned tries solve each training cell
    t0 = time.perf_counter()
    U, Pp = cell.p2u(P, P0, g, G), P0.clone()
    _, info_l = LN.implicit_learned(P, gas, rad, U, Pp, g, G, dtd, maxiter=args.learned_maxiter)
    Fg, Tgt = LN.gate_pairs(P, P0, dtd, info_l)
    gate_pos = float((Tgt[0] > 0).float().mean())
    # a slice of the training cells, kept to choose the gate's threshold
    nt = min(1 << 20, P0.shape[1])
    tune_set = (P0[:, :nt].contiguous(), g[:, :nt].contiguous(), G[:, :nt].contiguous(), dtd[:nt].contiguous())
    del U, Pp, P0, g, G, dtd, info_l
    torch.cuda.empty_cache()
    gate, hist_gate = LN.train(P, Fg, Tgt, LN.GATE, schedule=sched)
    torch.cuda.synchronize()
    del Fg, Tgt
    torch.cuda.empty_cache()
    # the threshold: the fastest gated solve of the training slice over a
    # few candidates (the evaluation batch is not looked at)
    thr_log = []
    if args.gate_thr is None:
        P0t, gt, Gt, dtt = tune_set
        U0t = cell.p2u(P, P0t, gt, Gt)
        Ut, Pt = torch.empty_like(U0t), torch.empty_like(P0t)

        def reset_t():
            Ut.copy_(U0t)
            Pt.copy_(P0t)

        for thr in (-0.4, -0.2, 0.0, 0.2, 0.4, 0.6):
            w, _, _ = _timed_solves(lambda: LN.implicit_learned_gated(P, gas, rad, gate, Ut, Pt, gt, Gt, dtt,
                                                                      maxiter=args.learned_maxiter, thr=thr),
                                    reset_t, 2, 1)
            thr_log.append((thr, w * 1e3))
        gate_thr = min(thr_log, key=lambda x: x[1])[0]
        del U0t, Ut, Pt
    else:
        gate_thr = args.gate_thr
    del tune_set
    torch.cuda.empty_cache()
    t_gate = time.perf_counter() - t0
    say("trained the gate (%.3f of the training cells solved by the learned tries) in %.1f s, final loss %.4f; "
        "threshold %.2f from %s" % (gate_pos, t_gate, hist_gate[-1], gate_thr, thr_log))
    # ---- fresh evaluation batch ----
    n = args.cells
    pp, g20, G20, dt = PR.implicit_batch_states(P, n, seed=args.seed)
    Pp0, g, G, dtd = cell.to_dev(pp), cell.to_dev(g20), cell.to_dev(G20), torch.from_numpy(dt).to(dev)
    U0 = cell.p2u(P, Pp0, g, G)
    U, Pp = torch.empty_like(U0), torch.empty_like(Pp0)

    def reset():
        U.copy_(U0)
        Pp.copy_(Pp0)

    w_plain, k_plain, (_, ia) = _timed_solves(lambda: cell.implicit(P, U, Pp, g, G, dtd), reset,
                                              args.steps, args.warmup)
    pa = Pp.clone()
    say("plain solve %.2f ms" % (w_plain * 1e3))
    w_learn, k_learn, (_, ib) = _timed_solves(lambda: LN.implicit_learned(P, gas, rad, U, Pp, g, G, dtd,
                                                                          maxiter=args.learned_maxiter), reset,
                                              args.steps, args.warmup)
    pb = Pp.clone()
    say("learned solve %.2f ms" % (w_learn * 1e3))
    w_gated, k_gated, (_, ic) = _timed_solves(lambda: LN.implicit_learned_gated(P, gas, rad, gate, U, Pp, g, G, dtd,
                                                                                maxiter=args.learned_maxiter,
                                                                                thr=gate_thr), reset,
                                              args.steps, args.warmup)
    pc = Pp.clone()
    say("gated learned solve %.2f ms" % (w_gated * 1e3))
    # network stages alone (features, BF16 inference, guess) on the same batch
    reset()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(stream)
    Fe = gas.features(P, Pp, dtd)
    ev[1].record(stream)
    Ye = gas.infer(Fe)
    ev[2].record(stream)
    gas.guess(Pp, Ye)
    ev[3].record(stream)
    ev[3].synchronize()
    t_feat, t_inf, t_guess = (ev[i].elapsed_time(ev[i + 1]) / 1e3 for i in range(3))
    ia, ib, ic = ia.cpu().numpy(), ib.cpu().numpy(), ic.cpu().numpy()
    ok_a, ok_b, ok_c = ia[1] >= 0, ib[1] >= 0, ic[1] >= 0
    pa, pb, pc = pa.cpu().numpy(), pb.cpu().numpy(), pc.cpu().numpy()
    # cells the gate sends straight to the chain: the plain solve's result
    # bit for bit (same attempt, same iterations, same primitives)
    gated_out = (ic[1] != 6) & (ic[5] == ia[5])
    same_as_plain = bool(np.array_equal(pc[:, gated_out], pa[:, gated_out], equal_nan=True))
    it_c = float(ic[5].mean())
    both = (ib[1] == 6) & (ia[1] == 0) & (ia[2] == 2) & (ib[2] == 2)
    du = np.abs(pa[1] - pb[1]) / np.abs(pa[1])
    it_a, it_b = float(ia[5].mean()), float(ib[5].mean())
    # algorithmic flops: the counted FP64 operations of the plain solve scaled
    # per Newton iteration to the learned solve's iterations, plus the network
    flops_it = IMPLICIT_FLOPS_PER_SOLVE / it_a
    mlp_flops = 2.0 * (16 * 128 + 2 * 128 * 128 + 128 * 16)   # padded MFMA tiles, per network
    learned_flops = flops_it * it_c + 3 * mlp_flops
    achieved = learned_flops * n / k_gated / 1e12
    inf_tf = mlp_flops * n / t_inf / 1e12
    res = {
        "metric": METRIC,
        "value": n / w_gated,
        "unit": "implicit solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": w_gated * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": "learned-guess implicit radiation-source solve (gated) vs plain guess",
                   "gated": {"solves_per_s": n / w_gated, "speedup": w_plain / w_gated,
                             "mean_total_iters": it_c, "iteration_ratio": it_a / it_c,
                             "solved_frac": float(ok_c.mean()), "learned_attempt_frac": float((ic[1] == 6).mean()),
                             "gate_thr": gate_thr, "gate_thr_search_ms": thr_log,
                             "gate_train_positive_frac": gate_pos,
                             "gate_train_s": t_gate, "train_loss_gate": hist_gate,
                             "chain_cells_equal_plain": same_as_plain},
                   "cells": n, "train_cells": args.train_cells, "train_pairs": int(n_pairs),
                   "schedule_lr_epochs": sched, "batch": LN.BATCH, "networks": "gas + radiation",
                   "learned_maxiter": args.learned_maxiter or P.radimpmaxiter,
                   "train_loss_gas": hist_g, "train_loss_rad": hist_r, "pairs_s": t_pairs, "train_s": t_train,
                   "plain_solves_per_s": n / w_plain, "learned_solves_per_s": n / w_learn,
                   "speedup": w_plain / w_learn,
                   "mean_total_iters_plain": it_a, "mean_total_iters_learned": it_b,
                   "iteration_ratio": it_a / it_b,
                   "solved_frac_plain": float(ok_a.mean()), "solved_frac_learned": float(ok_b.mean()),
                   "learned_attempt_frac": float((ib[1] == 6).mean()),
                   "learned_rad_prims_frac": float(((ib[1] == 6) & (ib[2] == 1)).mean()),
                   "same_system_cells": int(both.sum()),
                   "same_system_du_median": float(np.median(du[both])) if both.any() else None,
                   "same_system_du_max": float(du[both].max()) if both.any() else None,
                   "network_ms": {"features": t_feat * 1e3, "infer_bf16": t_inf * 1e3,
                                  "guess": t_guess * 1e3}},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "gated learned implicit solve: 3 x (features + BF16 inference) + guesses + "
                               "gate + classify + 16 k_imp_pass launches (one batch)",
                     "kernel_ms": k_gated * 1e3,
                     "algorithmic_flops_per_launch": learned_flops * n,
                     "flops_basis": "counted FP64 operations per plain solve / plain mean iterations x "
                                    "learned mean iterations + network MACs x 2",
                     "infer": {"achieved_TFLOPs": inf_tf, "peak_TFLOPs": 2500.0, "frac": inf_tf / 2500.0,
                               "bound": "mfma (bf16)"}},
    }
    if not args.no_cpu_baseline:
        # the reference's CPU path is the plain solve (it has no learned guess)
        res["cpu_baseline"] = cpu_baseline_implicit(P, pp, g20, G20, dt, args)
    return res


def cpu_baseline_implicit(P, pp, g20, G20, dt, args):
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    ns = min(args.cpu_sample, len(dt))
    threads = cpu_threads(args)
    uu = pyoracle.p2u(P, pp[:ns], g20[:ns], G20[:ns])
    t = time.perf_counter()
    pyoracle.implicit(P, uu, pp[:ns], g20[:ns], G20[:ns], dt[:ns], nthreads=threads)
    el = time.perf_counter() - t
    return {"value": ns / el, "unit": "implicit solves/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": "first %d cells of the same seeded batch, C oracle (gcc -O3, OpenMP, %d threads)"
                      % (ns, threads)}


# Algorithmic HBM bytes per thread-unit of each grid operator: the
# compulsory traffic of that operator as this build lays the data out.  Each
# primitive/conserved vector is one NV-vector of FP64; stencil neighbours are
# shared through the caches; the metric lives in per-(x1, x2) tables read
# through the caches by every x3 column (no per-unit HBM bytes).
def grid_alg_bytes(kind, nv, fused=False):
    return {
        "recon": (nv + 2 * nv) * 8,                # per face: prims of its cell, both face states
        # per face: face states (split) or cell prims (fused), 3 wave-speed
        # values (each cell's shared by its two faces), the flux
        "flux": ((nv if fused else 2 * nv) + 3 + nv) * 8,
        "u2p": (4 * nv) * 8 + 4,                   # per cell: u, p in and out, 4 flag bytes
        "update": (nv + nv + 3 * nv + nv) * 8,     # per cell: p, u, 3 face fluxes, u
        "wavespeeds": (nv + 18) * 8,               # per cell: prims, 18 speeds
        "emf": 9 * 8, "ct_faces": 9 * 8,           # per edge
        "comb": 3 * 8,                             # per (cell, variable)
        "bc_faces": (3 * nv) * 8,                  # per ghost cell (the whole fill): source prims, p and u
        "p2u": (2 * nv) * 8,
        "fixup": 8, "fixup_apply": 8,
        "implicit": (4 * nv + 40 + 1) * 8 + 3,
    }.get(kind, 0)


def bench_grid(args, kind):
    import numpy as np
    import torch
    from koral_amd import params as KP, problems as PR, grid as GR
    from koral_amd.gridrun import DeviceGrid

    if kind == "grmhd":
        P = KP.params_torus_mhd()
        n = tuple(args.grid) if args.grid else (128, 128, 64)
        label = "3D GRMHD RK2IMEX step: MKS1 Kerr a=0.9 magnetised torus, MP5+HLL+flux-CT, u2p"
    else:
        P = KP.params_torus_rad()
        n = tuple(args.grid) if args.grid else (192, 192, 96)
        label = "3D GRRMHD+M1 RK2IMEX step: MKS1 Kerr a=0.9 torus with radiation, MP5+HLL+flux-CT, u2p, u2prad, implicit"
    spec, p = PR.torus_setup(P, n=n, seed=args.seed)
    stream = torch.cuda.Stream()
    dg = DeviceGrid(P, spec, stream=stream.cuda_stream)
    dg.upload(GR.to_aos(p))
    dg.init()
    if args.advance > 0:
        # evolve the torus before anything is tuned or timed (later steps
        # carry more cells and faces on the rare deferred paths)
        dg.run(args.advance)
    for _ in range(args.warmup):
        dg.step()
    zones = n[0] * n[1] * n[2]
    # launch choices (block sizes, fused or split sweeps, deferred-pass
    # mapping) searched on this state before the timed steps
    tune_info = None
    if args.tune_json:
        import json
        want = json.loads(args.tune_json)
        tn = dg.get_tune()
        names = {dg.L.koral_hip_grid_op_name(k).decode(): k for k in range(32)}
        for d, v in enumerate(want.get("fused", list(tn.fused))):
            tn.fused[d] = int(v)
        for key in ("floor_wave", "update_pre", "comb_nv", "u2p_budget", "u2p_group", "floor_nw", "u2p_nw"):
            if key in want:
                setattr(tn, key, int(want[key]))
        for nm, b in want.get("block", {}).items():
            tn.block[names[nm]] = int(b)
        dg.set_tune(tn)
        tune_info = {"given": want}
    elif not args.no_autotune:
        t0 = time.perf_counter()
        rows = dg.autotune(steps=args.tune_steps)
        if args.tune_log:
            with open(args.tune_log, "w") as f:
                for name, val, ms in rows:
                    f.write("%-28s %8s %8.4f\n" % (name, "" if val is None else val, ms))
        tn = dg.get_tune()
        names = {k: dg.L.koral_hip_grid_op_name(k).decode() for k in range(32)}
        chosen = [r for r in rows if r[0].startswith("chose")]
        tune_info = {"seconds": time.perf_counter() - t0, "trials": len(rows) - len(chosen),
                     "start_ms_per_step": rows[0][2], "end_ms_per_step": chosen[-1][2] if chosen else rows[0][2],
                     "fused": list(tn.fused), "floor_wave": tn.floor_wave, "update_pre": tn.update_pre,
                     "comb_nv": tn.comb_nv, "u2p_budget": tn.u2p_budget, "u2p_group": tn.u2p_group,
                     "floor_nw": tn.floor_nw,
                     "u2p_nw": tn.u2p_nw,
                     "block": {names[k]: tn.block[k] for k in range(32) if names[k] and tn.block[k] != 256}}
    # pilot: three steps with HIP events around every launch give the
    # per-operator split and the dominant operator
    dg.timing(enable=True, read=False)
    for _ in range(3):
        dg.step()
    torch.cuda.synchronize()
    split = dg.timing(enable=False)
    # the floor passes run on side streams beside the main work (flux) or
    # on a few cells (u2p): not candidates for the dominant kernel
    dom = max((k for k in split if not k.endswith("_floor")), key=lambda k: split[k][0])
    # timed steps, with the time kept on the device (no host round trip
    # between steps; dt of each step from the previous one, as ko.c does):
    # all but the last replay one captured HIP graph per step (the graph is
    # captured and replayed once before the timed region); the last runs as
    # stream launches with HIP events around the dominant operator's
    # launches, which give its duration.  --eager times every step that way.
    graph = not args.eager and args.steps > 1
    if graph:
        dg.use_graph(True)
        dg.run(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    counters = np.zeros(6, np.int64)
    if graph:
        dg.run(args.steps - 1)
        counters += np.asarray(dg.counters, np.int64)
    dg.timing_only([dom])
    dg.timing(enable=True, read=False)
    dg.run(1 if graph else args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    counters += np.asarray(dg.counters, np.int64)
    tm = dg.timing(enable=False)
    dg.timing_only(None)
    if dom not in tm:
        tm = split
    dom = max((k for k in tm if not k.endswith("_floor")), key=lambda k: tm[k][0])
    ms, units, nl = tm[dom]
    per_launch = ms / nl / 1e3
    tn = dg.get_tune()
    dims = [d for d in range(3) if n[d] > 1]
    nfused = sum(1 for d in dims if tn.fused[d])
    # face kernels: the bytes of each direction's schedule, averaged
    if dom == "flux":
        bpu = sum(grid_alg_bytes("flux", P.nv, bool(tn.fused[d])) for d in dims) / len(dims)
    else:
        bpu = grid_alg_bytes(dom, P.nv)
    bytes_launch = bpu * units / nl
    achieved = bytes_launch / per_launch / 1e9
    traffic = None
    if kind == "grmhd" and n == (128, 128, 64):
        tr = GRID_TRAFFIC_128x128x64.get(dom)
        if dom == "flux":
            v = [tr["fused" if tn.fused[d] else "split"][d] for d in dims]
            traffic = sum(v) / len(v) if None not in v else None
        elif dom == "update":
            traffic = tr.get(tn.update_pre)
        elif dom == "recon":
            v = [tr["split"][d] for d in dims if not tn.fused[d]]
            traffic = sum(v) / len(v) if v and None not in v else None
        elif isinstance(tr, float):
            traffic = tr
    # FP64 rate of the dominant kernel from the operation count (the face
    # kernels only when every sweep is fused: a split sweep's flux kernel
    # does not do the reconstruction the count includes)
    fp64 = None
    if kind == "grmhd":
        fz = GRMHD_FLOPS_PER_LAUNCH_ZONE
        if dom == "flux" and nfused == len(dims):
            fp64 = sum(fz["sweep"][d] for d in dims) / len(dims) * zones
        elif dom in fz and dom != "sweep":
            fp64 = fz[dom] * zones
    step_s = wall / args.steps
    zups = zones * args.steps / wall
    step_info = None
    if kind == "grmhd":
        fl = GRMHD_FLOPS_PER_ZONE["step"] * zups / 1e12
        step_info = {"fp64_flops_per_zone_update": GRMHD_FLOPS_PER_ZONE["step"],
                     "fp64_achieved_TFLOPs": fl, "fp64_frac": fl / FP64_PEAK_TFLOPS}
        for k, b in GRMHD_BYTES_PER_ZONE.items():
            step_info["hbm_bytes_per_zone_update_" + k] = b
            step_info["hbm_frac_" + k] = b * zups / 1e9 / HBM_PEAK_GBS
        # the face sweeps (reconstruction + fluxes) against their FP64 count
        sw_ms = (split.get("flux", (0.,))[0] + split.get("recon", (0.,))[0]) / 3.
        if sw_ms > 0:
            sw = GRMHD_FLOPS_PER_ZONE["sweeps"] * zones / (sw_ms / 1e3) / 1e12
            step_info["sweeps_ms_per_step"] = sw_ms
            step_info["sweeps_fp64_TFLOPs"] = sw
            step_info["sweeps_fp64_frac"] = sw / FP64_PEAK_TFLOPS
    # the roof the dominant kernel is closer to: HBM by this layout's
    # compulsory bytes, or FP64 by the operation count; the other alongside
    hbm = {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "algorithmic_bytes_per_launch": bytes_launch,
           "algorithmic_bytes_basis": "compulsory bytes of this layout (metric in per-(x1,x2) tables, "
                                      "served by the caches)"}
    comp = None
    if fp64:
        fa = fp64 / per_launch / 1e12
        comp = {"achieved": fa, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": fa / FP64_PEAK_TFLOPS,
                "algorithmic_flops_per_launch": fp64,
                "compute": "FP64 vector ALU (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)"}
    if comp and comp["frac"] > hbm["frac"]:
        roof = dict(bound="mfma", **comp)
        roof["hbm"] = hbm
    else:
        roof = dict(bound="hbm", **hbm)
        roof["fp64"] = comp
    roof.update({"traffic": traffic, "traffic_GBps": traffic / per_launch / 1e9 if traffic else None,
                 "kernel": dom, "kernel_ms": per_launch * 1e3, "launches": nl, "fused_sweeps": nfused,
                 "step": step_info})
    res = {
        "metric": METRIC,
        "value": zups,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded torus initial data)",
        "config": {"workload": label, "grid": list(n), "nv": P.nv, "ng": spec.ng, "autotune": tune_info,
                   "advance_steps": args.advance,
                   "headline_note": ("the GRRMHD+M1 step the metric names is refused by this build (unresolved "
                                     "device fault of an earlier build); this line is the GRMHD step")
                   if kind == "grmhd" else None,
                   "counters_per_step": (counters / args.steps).tolist(),
                   "timing": ("every step as stream launches, HIP events around the dominant operator" if not graph
                              else "steps 1..K-1 replay one captured HIP graph per step; step K as stream launches "
                                   "with HIP events around the dominant operator's launches"),
                   "operator_ms_per_step_pilot": {k: v[0] / 3 for k, v in sorted(split.items(), key=lambda kv: -kv[1][0])}},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_grid(P, kind, args, n)
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads(args):
    # the GPU box's host share is 16 cores (whatever os.cpu_count() shows)
    return args.cpu_threads or min(16, os.cpu_count() or 1)


def cpu_baseline_grid(P, kind, args, n_default):
    """The oracle's RK2IMEX step (the C restatement of ko.c:444-570 and the
    operators it calls; gcc -O3, OpenMP over the per-cell loops, results
    independent of the thread count) on the same torus set-up, timed for
    about --cpu-seconds on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from koral_amd import problems as PR, grid as GR
    n = tuple(args.cpu_grid) if args.cpu_grid else n_default
    threads = cpu_threads(args)
    spec, p = PR.torus_setup(P, n=n, seed=args.seed)
    G = pyoracle.Grid(spec)
    S = pyoracle.GridState(P, G, GR.to_aos(p))
    pyoracle.grid_set_threads(threads)
    try:
        S.init_u()
        S.set_bc()
        dt = 1. / S.tstepden()
        t, steps, el = 0., 0, 0.
        while steps < 1 or (el < args.cpu_seconds and steps < 50):
            t0 = time.perf_counter()
            tm, _ = S.step(t, dt)
            el += time.perf_counter() - t0
            t += dt
            dt = 1. / tm
            steps += 1
    finally:
        pyoracle.grid_set_threads(1)
    zones = n[0] * n[1] * n[2]
    return {"value": zones * steps / el, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": "%d RK2IMEX steps of the same torus set-up at %dx%dx%d, C oracle (gcc -O3, OpenMP, "
                      "%d threads)" % (steps, n[0], n[1], n[2], threads)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="grmhd", choices=["implicit", "learned", "grmhd", "grrmhd"])
    ap.add_argument("--grid", type=int, nargs=3, default=None)
    ap.add_argument("--cpu-grid", type=int, nargs=3, default=None,
                    help="grid of the CPU baseline (default: the benchmark's own grid)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cells", type=int, default=1 << 22)
    ap.add_argument("--train-cells", type=int, default=1 << 23)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--learned-maxiter", type=int, default=16)
    ap.add_argument("--gate-thr", type=float, default=None,
                    help="gate threshold (default: chosen on a slice of the training cells)")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=1 << 17)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--eager", action="store_true",
                    help="grid workloads: every timed step as stream launches (default: all but the last replay a "
                         "captured HIP graph)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch choices")
    ap.add_argument("--tune-json", default=None,
                    help="launch choices to use instead of the search, as the line's config.autotune reports "
                         "them (fused, floor_wave, update_pre, comb_nv, u2p_budget, u2p_group, floor_nw, u2p_nw, "
                         "block)")
    ap.add_argument("--tune-steps", type=int, default=3)
    ap.add_argument("--advance", type=int, default=0,
                    help="grid workloads: untimed steps before the tuning and the timed steps")
    ap.add_argument("--tune-log", default=None, help="write the autotuner's search (choice, value, ms per step) here")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the single-GPU build holds the whole domain "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "implicit":
        res = bench_implicit(args)
    elif args.workload == "learned":
        res = bench_learned(args)
    else:
        res = bench_grid(args, args.workload)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
