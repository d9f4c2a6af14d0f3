"""Benchmark of the immersed-boundary transfer hot path on one MI355X.

One step = one interpolation (eulr2lagr) plus one spread (lagr2eulr) over the whole batch,
with inputs resident in HBM.  The step is captured as one hipGraph; K replays between two
synchronisations give `value` and `ms_per_step`.  The same step is then launched eagerly K
times with HIP events between its stages (`eager_ms_per_step`); those events give the
per-kernel durations of the roofline.  Default workload: the
512 x 256 x 512 channel (h = 1/64, x,z periodic, y walls) with 2000 seeded non-overlapping
capsules of radius U(4h,10h) and degree 8..24 by radius (1.16 M Lagrangian points).
`--workload dense-full`: the dense coupled channel at full size, 1024 x 512 x 256 cells with
8000 capsules of radius 10.6h and degree 16 at 30 % volume fraction (4 624 000 points);
`--workload dense`: its 1/8 share, 128 x 512 x 256 cells with 1000 of those capsules.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5]
                    [--workload uneven|dense-full|dense|moderate|small]

Prints one JSON line.  `value` = Lagrangian points per second over interp + spread.
`roofline` is for the dominant kernel, the spread's main kernel (k_obj_spread for the
capsule-object configurations, k_spread for the tile ones), which writes the whole force
field; `path_roofline` is the same figure for the whole interp + spread step.  `surrogate`
compares the step time with the learned membrane traction against the plain FP64 law
(`--no-surrogate` skips that leg), and `cpu_baseline` times the reference's transfer loops on
the host cores (a Fortran restatement at -O0 and -O2, serial and OpenMP) on a bounded sample
of the same workload (`--no-cpu` skips it).  `capsule_step` times the whole capsule step
without the fluid update (interpolation, FP64 membrane law, spread, position update;
`--no-capsule-step` skips it).  `--mlp-only` prints the surrogate network's BF16 inference
throughput at `--mlp-batch` rows (default 2048) and nothing else.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
_T0 = time.perf_counter()


def progress(msg: str) -> None:
    """A progress line on stderr (the JSON result line alone goes to stdout)."""
    print(f"bench [{time.perf_counter() - _T0:7.1f} s]: {msg}", file=sys.stderr, flush=True)


def support_cells(g, caps):
    """U: distinct grid cells inside any point's support window [f-1, f+2]^3 (slot-0 copy,
    x and y wrapped periodically, z clamped to the walls)."""
    mark = np.zeros(g.itot * g.jtot * g.ktot, dtype=bool)
    n = caps.npts
    step = 1 << 16
    off = np.arange(4)
    for a in range(0, n, step):
        b = min(n, a + step)
        fi = np.floor(caps.x[a:b] * g.dxi).astype(np.int64) - 1
        fj = np.floor(caps.y[a:b] * g.dyi).astype(np.int64) - 1
        fk = np.floor(caps.z[a:b] * g.dzi).astype(np.int64) - 1
        I = (fi[:, None] + off[None, :] - 1) % g.itot
        J = (fj[:, None] + off[None, :] - 1) % g.jtot
        K = fk[:, None] + off[None, :]
        Kok = (K >= 1) & (K <= g.ktot)
        K = np.clip(K, 1, g.ktot) - 1
        idx = (I[:, :, None, None] + g.itot * (J[:, None, :, None] + g.jtot * K[:, None, None, :]))
        ok = np.broadcast_to(Kok[:, None, None, :], idx.shape)
        mark[idx[ok]] = True
    return int(mark.sum())


def algorithmic_bytes(g, caps, U):
    n = caps.npts
    interp = n * 48 + U * 24
    spread = n * 56 + 24 * g.ncells_halo
    return interp, spread


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl):
    """The reference's transfer loops on the host: the Fortran restatement oracle/ibx_port.f90
    (bit for bit the C oracle when serial; tests/test_cpu_port.py) built as the reference ships
    it (-O0, OpenMP off; Makefile:7,14), at -O2, and both with the reference's own OpenMP
    directives on (parallel loop over capsules, atomic adds in the spread) on every host core
    this process may use.  Bounded sample: every other capsule of the workload (about half of
    its points) on the full grid, one eulr2lagr + one lagr2eulr (including the zeroing of the
    three force fields and the halos) per variant, after one untimed pass that touches the
    output pages.  `value` is the fastest variant (-O2 with OpenMP); all four are listed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from suguru_amd import workload as W

    g, caps = wl.grid, wl.caps
    sel = np.arange(0, caps.ncap, 2)
    idx = np.concatenate([np.arange(caps.cap_off[p], caps.cap_off[p + 1]) for p in sel])
    off = np.concatenate([[0], np.cumsum(np.diff(caps.cap_off)[sel])]).astype(np.int64)
    x, y, z, area = (np.ascontiguousarray(getattr(caps, k)[idx]) for k in ("x", "y", "z", "area"))
    fx, fy, fz = (np.ascontiguousarray(a[idx]) for a in (wl.fx, wl.fy, wl.fz))
    n = idx.size
    nthreads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    out = [g.empty_field() for _ in range(3)]
    variants = {}
    for v, th in (("O0", 1), ("O2", 1), ("O0_omp", nthreads), ("O2_omp", nthreads)):
        if O.port_lib(v) is None:
            continue
        if not variants:  # touch the output pages once (not timed)
            O.port_lagr2eulr(v, g, off, x, y, z, fx, fy, fz, area, threads=th, out=out)
        t0 = time.perf_counter()
        O.port_eulr2lagr(v, g, wl.u, wl.v, wl.w, off, x, y, z, threads=th)
        t1 = time.perf_counter()
        O.port_lagr2eulr(v, g, off, x, y, z, fx, fy, fz, area, threads=th, out=out)
        t2 = time.perf_counter()
        variants[f"{v.split('_')[0]}_{th}t"] = {"value": n / (t2 - t0), "interp_s": t1 - t0,
                                                "spread_s": t2 - t1, "threads": th,
                                                "openmp": v.endswith("omp")}
    if not variants:
        return {"error": "oracle/_port not built (make -C oracle port)"}
    best_key = max(variants, key=lambda k: variants[k]["value"])
    best = variants[best_key]
    return {
        "value": best["value"],
        "unit": "Lag-pts/s",
        "cores": best["threads"],
        "kind": "port",
        "variant": best_key,
        "cpu_model": _cpu_model(),
        "host_cores_usable": nthreads,
        "sample": (f"every other capsule ({sel.size} of {caps.ncap}, {n} points) on the full "
                   f"{g.itot}x{g.jtot}x{g.ktot} grid: one eulr2lagr + one lagr2eulr per variant "
                   "(oracle/ibx_port.f90 built by flang: -O0 as shipped and -O2, serial and "
                   "with the reference's OpenMP directives)"),
        "variants": variants,
    }


def surrogate_leg():
    """The learned membrane traction at degree 16 on its own (suguru_amd.surrogate): the
    Jacobian of the FP64 law at the unit sphere, the second-order network (289 SH coefficients
    x 3 -> 578 points x 3; 4 hidden layers of 512, SiLU; BF16 MFMA, FP32 accumulation) trained
    here on 200 000 flow-like shapes labelled by the GPU law, then measured on 4096 held-out
    shapes against the law: relative traction error by deformation amplitude with and without
    the learned term, the gate (error ~ alpha x amplitude) and, at tol = 1e-3, the accepted
    error and fallback share of the stateless predictor.  The time-stepping (anchored) form
    is measured inside the coupled-step leg."""
    from suguru_amd.surrogate import run_experiment

    r = run_experiment(n_train=200_000, n_test=4096, steps=20_000, batch=512, lr=1e-3, seed=1,
                       tol=1e-3, kind="flow")
    keep = ("n_train", "steps", "train_s", "loss_first", "loss_last", "rel_err_linear_median",
            "rel_err_median", "rel_err_p90", "by_amplitude", "gate_alpha", "gate_spearman", "tol",
            "fallback_fraction", "accepted_rel_err_median", "accepted_rel_err_max",
            "t_law_ms", "t_surrogate_ms")
    return {k: r[k] for k in keep}


def mlp_inference_leg(torch, batch=2048, reps=50, warmup=5):
    """Inference throughput of the surrogate network alone (suguru_amd.membrane.SurrogateMLP):
    3 x 289 SH coefficients -> 4 hidden layers of 512 (SiLU) -> 578 x 3 points, BF16 MFMA with
    FP32 accumulation, random weights and inputs, `batch` rows per forward pass on the
    network's own stream.  FLOPs are the unpadded products, 2 x batch x sum(in x out) per pass;
    the fraction is of MI355X's dense BF16 peak (2.5 PFLOP/s)."""
    from suguru_amd.surrogate import Surrogate

    sur = Surrogate(L=16, max_batch=batch)
    sur.mlp.init(3)
    dims = [sur.din] + [512] * 4 + [sur.dout]
    flop = 2.0 * batch * sum(a * b for a, b in zip(dims[:-1], dims[1:]))
    X = torch.randn(batch, sur.din, dtype=torch.float32, device=sur.dev)
    for _ in range(warmup):
        sur.predict(X)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sur.predict(X)
    torch.cuda.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    return {"batch": batch, "layers": dims, "gflop_per_pass": flop / 1e9, "ms_per_pass": ms,
            "tflops": flop / (ms * 1e-3) / 1e12, "frac_of_bf16_peak": flop / (ms * 1e-3) / 2.5e15,
            "timing": "wall clock over %d passes between two synchronisations" % reps}


def capsule_step_leg(torch, steps=20, warmup=3):
    """The capsule step without the fluid update (suguru_amd.capsule_step): interpolation,
    FP64 membrane law, spread and position update in the reference's order (main.f90:225-241)
    on one stream, on the dense share with unit-radius capsules (h = 1/10.6; 1000 capsules of
    degree 16, 578 000 points), gs = 6250, eb = 0.3 gs, dt = 1e-4.  The velocity field stays
    frozen: there is no flow solver here.  HIP events between the stages of each timed step."""
    from suguru_amd import workload as W
    from suguru_amd.capsule_step import CapsuleStep

    wl = W.config_dense_share(h=1.0 / 10.6)
    g, caps = wl.grid, wl.caps
    dev = torch.device("cuda:0")

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to(dev)

    u, v, w = up(wl.u), up(wl.v), up(wl.w)
    x, y, z = up(caps.x), up(caps.y), up(caps.z)
    F = [torch.empty(g.ncells_halo, dtype=torch.float64, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    st = CapsuleStep(g, caps.cap_off, 16, 6250.0, 0.3 * 6250.0, 1e-4, spread_config="objects-w8")
    try:
        st.bind(u, v, w, x, y, z, *F)
        for _ in range(warmup):
            st.step()
        st.sync()
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(steps)]
        for k in range(steps):
            st.step(ev[k])
        st.sync()
        stage = np.array([[ev[k][i].elapsed_time(ev[k][i + 1]) for i in range(4)]
                          for k in range(steps)]) * 1e3
        total_ms = ev[0][0].elapsed_time(ev[-1][4]) / steps
    finally:
        st.close()
    names = ("interp", "membrane_law", "spread", "update_pos")
    return {"workload": wl.name + " at h = 1/10.6 (unit-radius capsules)", "capsules": caps.ncap,
            "points": caps.npts, "steps": steps, "ms_per_step": total_ms,
            "stages_us_mean": {n: float(stage[:, i].mean()) for i, n in enumerate(names)},
            "fluid": "not included: frozen velocity field (the flow update is in the coupled_step leg)",
            "spread_config": "objects-w8 (what the autotuner picks for this workload)"}


def coupled_leg(torch, steps=100, warmup=2, surrogate=True):
    """The coupled step at the dense channel (1024 x 512 x 256 cells, 8000 unit-radius capsules
    of degree 16 at 30 %, h = 1/10.6; suguru_amd.coupled_step): interpolation, membrane
    traction, spread, Update_Pos and the flow update (FFT momentum and pressure solves),
    Rep = 0.0016, dt = 1e-4, gs = 6250, eb = 0.3 gs.  Per traction stage (the FP64 law, then
    the anchored surrogate with and without its learned term): `warmup` eager steps with HIP
    events between the stages give the stage times; then `steps` replays of the three captured
    hipGraphs from the same initial state give the step time and the state after `steps`.
    The surrogate runs are compared with the FP64 run after `steps`: capsule centroid
    displacements and mean point velocities, and the velocity field; at the final shapes the
    surrogate's prediction is also compared with the law capsule by capsule."""
    from suguru_amd import workload as W
    from suguru_amd.coupled_step import CoupledStep

    t0 = time.perf_counter()
    wl = W.config_dense_full()
    g, caps = wl.grid, wl.caps
    names = ("interp", "traction", "spread", "update_pos", "fluid")
    out = {"workload": wl.name + " at h = 1/10.6 (unit-radius capsules)",
           "grid_itot_jtot_ktot": [g.itot, g.jtot, g.ktot], "capsules": caps.ncap,
           "points": caps.npts, "steps": steps}
    sur = None
    if surrogate:
        from suguru_amd.surrogate import AnchoredTraction, Surrogate, calibrate_anchored

        a = time.perf_counter()
        sur = Surrogate(L=16, max_batch=8192)
        sur.build_jacobian()
        progress("coupled: surrogate dataset and training")
        X, Y, _, _, _ = sur.dataset(200_000, 1, kind="flow")
        sur.fit_norm(X, Y)
        sur.mlp.init(1)
        losses = sur.train(X, Y, 20_000, batch=512, lr=1e-3, seed=1)
        del X, Y
        beta_l, _, _ = calibrate_anchored(sur, learned=True, along=0.8)
        beta_0, _, _ = calibrate_anchored(sur, learned=False, along=0.8)
        out["surrogate_training"] = {"shapes": 200_000, "adam_steps": 20_000,
                                     "loss_first": losses[0], "loss_last": float(np.mean(losses[-50:])),
                                     "beta_learned": beta_l, "beta_linear": beta_0,
                                     "seconds": round(time.perf_counter() - a, 1)}
    out["setup_s"] = round(time.perf_counter() - t0, 1)

    def run(kind):
        traction, at = None, None
        st = CoupledStep(g, caps.cap_off, degree=16)
        try:
            if kind != "law":
                sur.set_stream(st.stream.cuda_stream)
                learned = kind == "surrogate"
                at = AnchoredTraction(sur, caps.ncap, beta_l if learned else beta_0, tol=1e-3,
                                      learned=learned)
                st.traction = at
            st.load(wl.u, wl.v, wl.w, caps.x, caps.y, caps.z)
            cen0 = None
            ev = [[torch.cuda.Event(enable_timing=True) for _ in range(6)] for _ in range(warmup)]
            for k in range(warmup):
                st.step(ev[k])
            st.sync()
            stage = np.array([[ev[k][i].elapsed_time(ev[k][i + 1]) for i in range(5)]
                              for k in range(warmup)]) * 1e3
            st.load(wl.u, wl.v, wl.w, caps.x, caps.y, caps.z)
            if at is not None:
                at.reset()
            st.capture()
            st.sync()
            a = time.perf_counter()
            for _ in range(steps):
                st.run()
            st.sync()
            ms = 1e3 * (time.perf_counter() - a) / steps
            res = {"graph_ms_per_step": ms,
                   "stages_us_eager": {n: float(stage[-1, i]) for i, n in enumerate(names)},
                   "centroid": st.centroid.view(-1, 3).cpu().numpy(),
                   "umean": st.umean.view(-1, 3).cpu().numpy(),
                   "vel": [u.clone() for u in st.current()]}
            if at is not None:
                res["law_fraction"] = float(at.refreshed.double().sum()) / (caps.ncap * (steps + 0))
                # the prediction at the final shapes against the law, capsule by capsule
                f, drift = at.predict(st.x, st.y, st.z)
                fl, _, _ = sur.law(st.x, st.y, st.z)
                ref = fl.view(3, caps.ncap, -1).permute(1, 0, 2).reshape(caps.ncap, -1)
                err = Surrogate.rel_error(f.reshape(caps.ncap, -1), ref)
                acc = (at.beta * drift <= at.tol)
                res["final_pred"] = {
                    "accepted_fraction": float(acc.double().mean()),
                    "accepted_rel_err_max": float(err[acc].max()) if bool(acc.any()) else 0.0,
                    "accepted_rel_err_median": float(err[acc].median()) if bool(acc.any()) else 0.0,
                    "all_rel_err_median": float(err.median())}
            return res
        finally:
            st.traction = None
            st.close()

    progress("coupled: FP64 law run")
    plain = run("law")
    out["plain"] = {k: v for k, v in plain.items() if k in ("graph_ms_per_step", "stages_us_eager")}
    x0 = np.stack([caps.centers[:, 0], caps.centers[:, 1], caps.centers[:, 2]], 1)
    if surrogate:
        for kind in ("surrogate", "linear_anchored"):
            progress(f"coupled: {kind} run")
            r = run(kind)
            disp_p = plain["centroid"] - x0
            disp_s = r["centroid"] - x0
            cmp = {"centroid_displacement_rel_err": float(np.linalg.norm(disp_s - disp_p) / np.linalg.norm(disp_p)),
                   "umean_rel_err": float(np.linalg.norm(r["umean"] - plain["umean"]) / np.linalg.norm(plain["umean"])),
                   "velocity_field_rel_err": float(max(
                       float((a - b).abs().max()) for a, b in zip(r["vel"], plain["vel"])) /
                       max(float(b.abs().max()) for b in plain["vel"]))}
            out[kind] = {"graph_ms_per_step": r["graph_ms_per_step"],
                         "stages_us_eager": r["stages_us_eager"], "law_fraction": r["law_fraction"],
                         "final_prediction_vs_law": r["final_pred"], "vs_plain_after_steps": cmp,
                         "step_speedup": plain["graph_ms_per_step"] / r["graph_ms_per_step"]}
        tr = plain["stages_us_eager"]["traction"]
        tot = sum(plain["stages_us_eager"].values())
        out["amdahl_ceiling_step"] = tot / (tot - tr)
        cap = sum(v for k, v in plain["stages_us_eager"].items() if k != "fluid")
        out["amdahl_ceiling_capsule_stages"] = cap / (cap - tr)
    return out


def copy_peak_gbs(torch, reps=10):
    """Practical HBM bandwidth: a 1 GiB device-to-device copy (read + write = 2 GiB per copy),
    timed with events on torch's stream."""
    a = torch.ones(2 ** 27, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    s = e0.elapsed_time(e1) * 1e-3 / reps
    del a, b
    return 2.0 * 2 ** 30 / s / 1e9


def read_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), committed
    under profiles/ as JSON; None if absent."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="uneven",
                    choices=["uneven", "dense", "dense-full", "moderate", "small"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of a graph")
    ap.add_argument("--config", type=int, default=-1,
                    help="tile configuration id (default: autotune)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="interpolation on the same stream as the spread (default: concurrent)")
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup + steps only (for rocprofv3), no extras")
    ap.add_argument("--no-capsule-step", action="store_true",
                    help="skip the capsule-step leg (interp, membrane law, spread, Update_Pos)")
    ap.add_argument("--no-coupled", action="store_true",
                    help="skip the coupled-step leg (dense 8000-capsule channel with the flow update)")
    ap.add_argument("--mlp-only", action="store_true",
                    help="only the surrogate network's inference throughput")
    ap.add_argument("--mlp-batch", type=int, default=2048, help="rows per pass for --mlp-only")
    ap.add_argument("--coupled-only", action="store_true",
                    help="run the coupled-step leg alone and print it")
    ap.add_argument("--coupled-steps", type=int, default=100,
                    help="graph replays of the coupled-step leg (default 100)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the learned-traction leg (training + step-time comparison)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (the reference kernel "
                         "for scaling rocprofv3's FETCH_SIZE and WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; the path's sharding over several GPUs (slabs in "
              "x and y with a ghost-capsule exchange) is a design and is not run.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        a = torch.ones(2 ** 27, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": a.numel() * 8, "copies": 3}))
        return

    if args.mlp_only:
        print(json.dumps({"mlp_inference": mlp_inference_leg(torch, batch=args.mlp_batch)}))
        return

    if args.coupled_only:
        print(json.dumps({"coupled_step": coupled_leg(torch, steps=args.coupled_steps,
                                                      surrogate=not args.no_surrogate)}))
        return

    from suguru_amd import Batch, IBTransfer, workload as W

    t_setup = time.perf_counter()
    maker = {"uneven": W.config_uneven, "dense": W.config_dense_share, "dense-full": W.config_dense_full,
             "moderate": W.config_moderate, "small": W.config_small}
    wl = maker[args.workload]()
    g, caps = wl.grid, wl.caps
    dev = torch.device("cuda:0")

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to(dev)

    t = {k: up(getattr(wl, k)) for k in ("u", "v", "w", "fx", "fy", "fz")}
    for k in ("x", "y", "z", "area"):
        t[k] = up(getattr(caps, k))
    n = caps.npts
    for k in ("ul", "vl", "wl"):
        t[k] = torch.empty(n, dtype=torch.float64, device=dev)
    for k in ("forcex", "forcey", "forcez"):
        t[k] = torch.empty(g.ncells_halo, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    ctx = IBTransfer(g, 0)
    batch = Batch(ctx, caps.cap_off)
    batch.bind(**t)
    if args.config >= 0:
        batch.set_config(args.config)
        choice = args.config
    else:
        cache = os.path.join(os.environ.get("TMPDIR", "/tmp"), "ibx_autotune.txt")
        choice = batch.autotune(cache)
    batch.set_overlap(not args.no_overlap)
    # the production path: one captured hipGraph per step
    batch.capture(use_graph=not args.no_graph)
    t_setup = time.perf_counter() - t_setup
    for _ in range(args.warmup):
        batch.run()
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    ctx.sync()
    torch.cuda.synchronize()
    graph_ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    # per-replay times (each replay synchronised on its own), for the median
    replay_ms = []
    for _ in range(args.steps):
        a = time.perf_counter()
        batch.run()
        ctx.sync()
        replay_ms.append(1000.0 * (time.perf_counter() - a))

    # the timed region: the same step, launched eagerly with HIP events on the batch stream
    # between the stages, so that the per-kernel durations come from the timed steps
    # themselves (a graph's event-record nodes cannot be read back as timings)
    batch.capture_timed(args.steps)
    for _ in range(args.warmup):
        batch.run()
    batch.reset_slots()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    ctx.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = 1000.0 * elapsed / args.steps
    slots = batch.slot_times()
    if args.profile_only:
        out = {"ms_per_step": graph_ms, "eager_ms_per_step": ms, "graph_ms_per_step": graph_ms,
               "config": Batch.config_names()[choice],
               "tile_spread_us": float(np.mean([s["tile_spread_us"] for s in slots]))}
        print(json.dumps(out))
        return

    prof = batch.profile(reps=max(5, args.steps))
    U = support_cells(g, caps)
    b_interp, b_spread = algorithmic_bytes(g, caps, U)
    stage = {k: float(np.mean([s[k] for s in slots])) for k in Batch.SLOT_KEYS}
    tile_s = stage["tile_spread_us"] * 1e-6
    achieved = b_spread / tile_s / 1e9
    cfg_name = Batch.config_names()[choice]
    objects = cfg_name.startswith("objects")
    traffic = read_traffic(f"traffic_{cfg_name}_{args.workload}.json")
    practical = copy_peak_gbs(torch)
    out = {
        "metric": "IB interp+spread Lag-pts/s & HBM GB/s; step time surrogate vs plain @1e-3",
        "value": n / (graph_ms * 1e-3),
        "unit": "Lag-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": graph_ms,
        "graph_ms_per_step": graph_ms,
        "graph_ms_per_step_median": float(np.median(replay_ms)),
        "eager_ms_per_step": ms,
        "timing": ("value and ms_per_step: K graph replays between two synchronisations; "
                   "eager_ms_per_step: K eager launches with HIP events between the stages, "
                   "which give the roofline's kernel durations"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Poiseuille + divergence-free perturbation; N(0,1) forces; "
                "seeded non-overlapping spherical capsules)",
        "config": {
            "workload": wl.name,
            "grid_itot_jtot_ktot": [g.itot, g.jtot, g.ktot],
            "capsules": caps.ncap,
            "lagrangian_points": n,
            "volume_fraction": round(W.volume_fraction(g, caps), 4),
            "tile_config": Batch.config_names()[choice],
            "graph": not args.no_graph,
            "interp_spread_overlap": not args.no_overlap,
        },
        "roofline": {
            "kernel": ("k_obj_spread (capsule-object spread: ordered per-cell sums of the "
                       "work items plus the zero units; writes the full 3-component force "
                       "field)") if objects else
                      "k_spread (tile spread, writes the full 3-component force field)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if traffic is None else traffic.get("bytes_per_launch"),
            "practical_peak": practical,
            "frac_of_practical": achieved / practical,
            "practical_peak_from": "1 GiB device-to-device copy, read + write bytes / time",
            "algorithmic_bytes": b_spread,
            "avg_launch_us": stage["tile_spread_us"],
            "timing": "HIP events on the batch stream around the kernel in each timed step",
        },
        "path_roofline": {
            "achieved": (b_interp + b_spread) / (graph_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (b_interp + b_spread) / (graph_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "algorithmic_bytes_interp": b_interp,
            "algorithmic_bytes_spread": b_spread,
            "support_cells_U": U,
            "stage_us_timed": stage,
            "stage_us_eager_profile": prof,
        },
        "setup_s": round(t_setup, 2),
    }
    progress(f"transfer line done ({out['value'] / 1e9:.3f} G pts/s)")
    if not args.no_surrogate:
        progress("surrogate leg")
        try:
            out["surrogate"] = surrogate_leg()
        except Exception as e:  # reported, never hides the transfer line
            out["surrogate"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_capsule_step:
        progress("capsule-step leg")
        try:
            out["capsule_step"] = capsule_step_leg(torch)
        except Exception as e:  # reported, never hides the transfer line
            out["capsule_step"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_coupled:
        progress("coupled-step leg")
        try:
            out["coupled_step"] = coupled_leg(torch, steps=args.coupled_steps)
        except Exception as e:  # reported, never hides the transfer line
            out["coupled_step"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(wl)
        if "value" in out["cpu_baseline"]:
            out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    batch.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
