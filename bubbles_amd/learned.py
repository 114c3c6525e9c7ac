"""Learned PCISPH pressure warm start: harvest, train, evaluate.

Data: (16 features at the start of a substep -> converged pressure / P) pairs
harvested from plain TEXTBOOK PCISPH runs on seeded dam-break and
block-drop scenes (P = delta * rho0, the pressure one correction assigns to a
100% density error).  The network (16 -> 64 -> 64 -> 1, GELU, BF16 MFMA,
FP32 accumulation) is trained on the device with Adam by the library's own
forward/backward kernels on the residual t - p_prev (p_prev / P is feature
10, the particle's pressure at the end of the previous substep).  At run time
the solver starts its predict-correct loop from
    p0 = P * alpha * clamp(net(features) * out_scale + p_prev / P, 0, pcap)
(alpha < 1: an over-estimate cannot be corrected because negative pressure
increments are scaled by 0) and the unchanged loop still iterates to the same
tolerance.

CLI:
  python -m bubbles_amd.learned train  [--out PATH] [--quick]
  python -m bubbles_amd.learned eval   [--weights PATH] [--steps N]
  python -m bubbles_amd.learned caps   [--caps 32,64,128,256] [--seed 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import abi, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "warmstart_mlp.npz")

TRAIN_SEEDS = list(range(1000, 1032))
HELDOUT_SEEDS = list(range(2000, 2008))


def training_scenes(seeds=TRAIN_SEEDS, scale=1.0):
    """Half dam breaks, half block drops, sizes varied by seed; spacing 0.01
    and dt 5e-4 like the benchmark scene."""
    out = []
    rng = np.random.default_rng(12345)
    for k, seed in enumerate(seeds):
        if k % 2 == 0:
            nx = int(round(rng.integers(48, 97) * scale))
            ny = int(round(rng.integers(24, 65) * scale))
            nz = int(round(rng.integers(48, 97) * scale))
            out.append(scenes.dam_break(max(nx, 8), max(ny, 8), max(nz, 8), spacing=0.01,
                                        tank=(2.0, 1.0, 2.0), seed=seed))
        else:
            out.append(scenes.block_drop(scale=0.25 * scale, seed=seed, spacing=0.01, dt=5e-4))
    return out


def harvest(scene_list, steps=400, every=25, start=25, per_snapshot=8192, max_it=8, tol=0.01,
            seed=0, log=print, drop_unconverged=False):
    """Run plain TEXTBOOK PCISPH on each scene; every `every` substeps record a
    random subset of (features, converged p / P).  drop_unconverged: leave
    out snapshots whose loop stopped at the cap above the tolerance (their
    pressures are what a stalled loop built, not a converged field)."""
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    for si, sc in enumerate(scene_list):
        s = abi.Solver(sc)
        iters = []
        for step in range(steps):
            if step >= start and (step - start) % every == 0:
                f, t, st = s.harvest(sc.dt, max_it, tol)
                k = min(per_snapshot, sc.n)
                sel = rng.choice(sc.n, size=k, replace=False)
                if not (drop_unconverged and st.max_density_error >= tol):
                    xs.append(f[sel])
                    ts.append(t[sel])
            else:
                st = s.pcisph_substep(sc.dt, max_it, tol, abi.MODE_TEXTBOOK)
            iters.append(st.iterations)
        s.close()
        log(f"harvest scene {si} {sc.name} n={sc.n} mean_it={np.mean(iters):.2f}")
    return np.concatenate(xs).astype(np.float32), np.concatenate(ts).astype(np.float32)


F_PPREV = 10        # feature: previous substep's converged pressure / P
N_TRAIN = 5313      # trainable parameters; then mean, inv_std, out_scale, alpha, pcap, resid
OFF_MEAN = N_TRAIN
OFF_ISTD = OFF_MEAN + 16
OFF_OSCALE = OFF_ISTD + 16
OFF_ALPHA, OFF_PCAP, OFF_RESID = OFF_OSCALE + 1, OFF_OSCALE + 2, OFF_OSCALE + 3


def fit(x, t, steps=20000, batch=4096, lr=1e-3, seed=0, alpha=0.7, expectile=0.5):
    """Train the network on the residual t - p_prev (the previous substep's
    converged pressure is the strongest single predictor); the solver then
    starts from alpha * clamp(p_prev + net, 0, pcap).  Under-estimating is
    safe -- the unchanged loop adds pressure -- while an over-estimate can
    never be removed (negative increments are scaled by 0), hence alpha < 1
    and/or an expectile below 0.5 (a loss that weighs over-estimates more)."""
    net = abi.Mlp(seed)
    if expectile != 0.5:
        net.set_expectile(expectile)
    p = net.get_params()
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    std[std < 1e-12] = 1.0
    res = (t - x[:, F_PPREV]).astype(np.float32)
    p[OFF_MEAN:OFF_MEAN + 16] = mean
    p[OFF_ISTD:OFF_ISTD + 16] = 1.0 / std
    p[OFF_OSCALE] = max(float(res.std()), 1e-6)
    net.set_params(p)
    loss = net.train(x, res, steps=steps, batch=batch, lr=lr, seed=seed)
    p = net.get_params()
    p[OFF_ALPHA] = alpha
    p[OFF_PCAP] = 1.5 * float(t.max())
    p[OFF_RESID] = 1.0
    net.set_params(p)
    return net, loss


def save(net, path=DEFAULT_WEIGHTS, meta=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=net.get_params(), meta=json.dumps(meta or {}))


def load(path=DEFAULT_WEIGHTS):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path}: train it with `python -m bubbles_amd.learned train`")
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
    net = abi.Mlp(0)
    net.set_params(params)
    return net


def load_or_train(solver=None, scene=None, path=DEFAULT_WEIGHTS):
    """Trained weights from `path` (the benchmark never trains inside timing)."""
    return load(path)


def compare(scene, net, steps=200, warmup=0, max_it=8, tol=0.01):
    """Plain vs warm-started TEXTBOOK PCISPH on the same scene: iterations,
    errors, time and the final state difference."""
    res = {}
    finals = {}
    for label, warm in (("plain", abi.WARM_NONE), ("learned", abi.WARM_MLP)):
        s = abi.Solver(scene)
        if warm == abi.WARM_MLP:
            s.attach_mlp(net)
        for _ in range(warmup):
            s.pcisph_substep(scene.dt, max_it, tol, abi.MODE_TEXTBOOK, abi.WARM_NONE)
        s.synchronize()
        it, err, ms = [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            st = s.pcisph_substep(scene.dt, max_it, tol, abi.MODE_TEXTBOOK, warm)
            it.append(st.iterations)
            err.append(st.max_density_error)
            ms.append(st.gpu_ms)
        s.synchronize()
        el = time.perf_counter() - t0
        res[label] = {"mean_iterations": float(np.mean(it)), "max_iterations": int(np.max(it)),
                      "max_density_error": float(np.max(err)),
                      "converged_fraction": float(np.mean(np.array(err) < tol)),
                      "particle_steps_per_s": scene.n * steps / el, "gpu_ms_per_step": float(np.mean(ms))}
        rho = s.get("density")
        res[label]["mean_abs_density_dev"] = float(np.mean(np.abs(rho / scene.target_density - 1.0)))
        finals[label] = s.get("position")
        s.close()
    res["speedup"] = res["learned"]["particle_steps_per_s"] / res["plain"]["particle_steps_per_s"]
    res["max_position_diff_over_h"] = float(np.abs(finals["plain"] - finals["learned"]).max() / scene.kernel_radius)
    return res


def run_variant(net, alpha, max_it, seed, steps=480, warmup=100, dims=(128, 64, 128), tol=0.01):
    """Iterations, error and throughput of one (network, alpha, cap) on a dam
    break; net None = plain PCISPH.  A run whose density error exceeds 50%
    is stopped and reported as exploded."""
    sc = scenes.dam_break(*dims, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=seed)
    s = abi.Solver(sc)
    warm = abi.WARM_NONE
    if net is not None:
        p = net.get_params()
        p[OFF_ALPHA] = alpha
        variant = abi.Mlp(0)
        variant.set_params(p)
        s.attach_mlp(variant)
        warm = abi.WARM_MLP
    exploded = False
    for _ in range(warmup):
        if s.pcisph_substep(sc.dt, max_it, tol, abi.MODE_TEXTBOOK, warm).max_density_error > 0.5:
            exploded = True
            break
    its, errs = [], []
    s.synchronize()
    t0 = time.perf_counter()
    if not exploded:
        for _ in range(steps):
            st = s.pcisph_substep(sc.dt, max_it, tol, abi.MODE_TEXTBOOK, warm)
            its.append(st.iterations)
            errs.append(st.max_density_error)
            if st.max_density_error > 0.5:
                exploded = True
                break
    s.synchronize()
    el = time.perf_counter() - t0
    s.close()
    if exploded:
        return {"exploded": True}
    return {"mean_it": round(float(np.mean(its)), 3), "max_err": round(float(np.max(errs)), 4),
            "conv": round(float(np.mean(np.array(errs) < tol)), 3),
            "psps": round(sc.n * steps / el / 1e6, 1)}


def cap_sweep(caps, seed=2, steps=500, dims=(128, 64, 128), tol=0.01, log=print):
    """The whole dam break from rest at each iteration cap: plain, the AUTO
    warm start and the warm start on every substep (TEXTBOOK, tolerance
    `tol`): mean and max iterations, max error, converged fraction,
    throughput, and the iterations per 100-substep block."""
    sc = scenes.dam_break(*dims, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=seed)
    net = load()
    out = []
    for cap in caps:
        for label, warm in (("plain", abi.WARM_NONE), ("learned auto", abi.WARM_AUTO),
                            ("learned every substep", abi.WARM_MLP)):
            s = abi.Solver(sc)
            if warm != abi.WARM_NONE:
                s.attach_mlp(net)
            s.synchronize()
            t0 = time.perf_counter()
            st = s.pcisph_run(sc.dt, steps, cap, tol, abi.MODE_TEXTBOOK, warm)
            s.synchronize()
            el = time.perf_counter() - t0
            s.close()
            it = np.array([x.iterations for x in st])
            er = np.array([x.max_density_error for x in st])
            rec = {"cap": cap, "run": label, "mean_it": round(float(it.mean()), 3), "max_it": int(it.max()),
                   "max_err": round(float(er.max()), 5), "conv": round(float((er < tol).mean()), 4),
                   "first_over_10pct": int(np.argmax(er > 0.1)) if (er > 0.1).any() else None,
                   "mpsps": round(sc.n * steps / el / 1e6, 1),
                   "it_per_100": [round(float(it[a:a + 100].mean()), 2) for a in range(0, steps, 100)]}
            log(json.dumps(rec))
            out.append(rec)
    return out


def select_whole(alphas, seeds, cap=128, steps=500, dims=(128, 64, 128), tol=0.01, log=print,
                 weights=DEFAULT_WEIGHTS, plain=True):
    """Whole runs from rest for each (alpha, seed, policy): the shipped
    network with its warm-start scale alpha replaced, the AUTO policy and
    the network on every substep, and plain for reference."""
    net = load(weights)
    out = []
    for seed in seeds:
        sc = scenes.dam_break(*dims, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=seed)
        variants = [("plain", None, abi.WARM_NONE)] if plain else []
        for a in alphas:
            variants += [("auto", a, abi.WARM_AUTO), ("every", a, abi.WARM_MLP)]
        for label, a, warm in variants:
            s = abi.Solver(sc)
            if a is not None:
                p = net.get_params()
                p[OFF_ALPHA] = a
                v = abi.Mlp(0)
                v.set_params(p)
                s.attach_mlp(v)
            s.synchronize()
            t0 = time.perf_counter()
            st = s.pcisph_run(sc.dt, steps, cap, tol, abi.MODE_TEXTBOOK, warm)
            s.synchronize()
            el = time.perf_counter() - t0
            s.close()
            it = np.array([x.iterations for x in st])
            er = np.array([x.max_density_error for x in st])
            rec = {"seed": seed, "run": label, "alpha": a, "cap": cap, "mean_it": round(float(it.mean()), 3),
                   "max_it": int(it.max()), "max_err": round(float(er.max()), 5),
                   "conv": round(float((er < tol).mean()), 4), "mpsps": round(sc.n * steps / el / 1e6, 1)}
            log(json.dumps(rec))
            out.append(rec)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--out", default=DEFAULT_WEIGHTS)
    tr.add_argument("--quick", action="store_true")
    tr.add_argument("--steps", type=int, default=20000)
    tr.add_argument("--batch", type=int, default=4096)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--harvest-max-it", type=int, default=32,
                    help="iteration cap of the plain runs that produce the targets")
    tr.add_argument("--alpha", type=float, default=0.7)
    tr.add_argument("--expectile", type=float, default=0.5,
                    help="training loss expectile (0.5 = mean squared error)")
    tr.add_argument("--drop-unconverged", action="store_true",
                    help="leave out snapshots whose loop stopped at the cap above the tolerance")
    ev = sub.add_parser("eval")
    ev.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ev.add_argument("--steps", type=int, default=480)
    ev.add_argument("--warmup", type=int, default=20)
    ev.add_argument("--nx", type=int, default=128)
    ev.add_argument("--ny", type=int, default=64)
    ev.add_argument("--nz", type=int, default=128)
    ev.add_argument("--seed", type=int, default=HELDOUT_SEEDS[0])
    se = sub.add_parser("select", help="sweep alpha and the iteration cap on given seeds")
    se.add_argument("--weights", default=DEFAULT_WEIGHTS)
    se.add_argument("--alphas", default="0.6,0.7,0.8")
    se.add_argument("--caps", default="8,64")
    se.add_argument("--seeds", default="2000,2")
    sw = sub.add_parser("select-whole", help="warm-start scale alpha over whole runs from rest")
    sw.add_argument("--alphas", default="0.7,0.75,0.8,0.85")
    sw.add_argument("--seeds", default="2000,2001")
    sw.add_argument("--cap", type=int, default=128)
    sw.add_argument("--weights", default=DEFAULT_WEIGHTS)
    sw.add_argument("--no-plain", action="store_true")
    cs = sub.add_parser("caps", help="whole-run plain vs learned at several iteration caps")
    cs.add_argument("--caps", default="32,64,128,256")
    cs.add_argument("--seed", type=int, default=2)
    cs.add_argument("--steps", type=int, default=500)
    args = ap.parse_args(argv)
    if args.cmd == "select-whole":
        select_whole([float(a) for a in args.alphas.split(",")], [int(x) for x in args.seeds.split(",")],
                     cap=args.cap, log=lambda m: print(m, flush=True), weights=args.weights,
                     plain=not args.no_plain)
        return 0
    if args.cmd == "caps":
        cap_sweep([int(c) for c in args.caps.split(",")], seed=args.seed, steps=args.steps,
                  log=lambda m: print(m, flush=True))
        return 0
    if args.cmd == "select":
        net = load(args.weights)
        caps = [int(c) for c in args.caps.split(",")]
        for seed in [int(x) for x in args.seeds.split(",")]:
            for cap in caps:
                print(json.dumps({"run": "plain", "seed": seed, "cap": cap, **run_variant(None, 0.0, cap, seed)}),
                      flush=True)
            for alpha in [float(a) for a in args.alphas.split(",")]:
                for cap in caps:
                    r = run_variant(net, alpha, cap, seed)
                    print(json.dumps({"run": "learned", "alpha": alpha, "seed": seed, "cap": cap, **r}),
                          flush=True)
        return 0
    if args.cmd == "train":
        t0 = time.perf_counter()
        seeds = TRAIN_SEEDS[:4] if args.quick else TRAIN_SEEDS
        scale = 0.5 if args.quick else 1.0
        x, t = harvest(training_scenes(seeds, scale), steps=200 if args.quick else 400,
                       per_snapshot=2048 if args.quick else 8192, max_it=args.harvest_max_it,
                       drop_unconverged=args.drop_unconverged)
        t1 = time.perf_counter()
        print(f"harvested {x.shape[0]} samples in {t1 - t0:.1f} s; target mean {t.mean():.4f} std {t.std():.4f}")
        net, loss = fit(x, t, steps=200 if args.quick else args.steps, batch=args.batch, lr=args.lr,
                        alpha=args.alpha, expectile=args.expectile)
        t2 = time.perf_counter()
        print(f"trained {args.steps} Adam steps in {t2 - t1:.1f} s, final batch loss {loss:.5f}")
        save(net, args.out, {"samples": int(x.shape[0]), "steps": args.steps, "batch": args.batch,
                             "lr": args.lr, "seeds": seeds, "final_loss": loss,
                             "harvest_max_it": args.harvest_max_it, "alpha": args.alpha,
                             "expectile": args.expectile, "drop_unconverged": args.drop_unconverged})
        print("saved", args.out)
    else:
        net = load(args.weights)
        sc = scenes.dam_break(args.nx, args.ny, args.nz, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=args.seed)
        print(json.dumps(compare(sc, net, steps=args.steps, warmup=args.warmup)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
