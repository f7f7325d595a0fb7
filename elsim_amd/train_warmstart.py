"""Train the learned CG warm start on solver-generated pairs, on the GPU.

    python -m elsim_amd.train_warmstart --n 64 --train-seeds 40 --test-seeds 10 \\
        --steps 50 --epochs 3 --out weights.npz

Pairs: for each seeded sphere pack and each step t of a run with the fast
solver, the input is the network features of the state at the start of the
step (psi_P, phi_P, c_P, the through-thickness coordinate and phi_P's last
change, see elsim_sbm_nn_features) and the target is this step's change of
phi_P in units of the last change's rms.  Samples stay in device memory.
Training: Adam on the mean squared error, one volume per update, with the
project's own forward/backward kernels.  Evaluation: held-out seeds run the
same steps with and without the warm start; phP CG iterations and time are
reported.
"""
import argparse
import json
import sys
import time

import numpy as np

import elsim_amd as E
from elsim_amd import nn as NN
from elsim_amd._lib import F_CNP, F_PHP, DeviceArray


def log(*a):
    print(*a, flush=True)


def make_state(n, seed, rmin, rmax):
    rng = np.random.default_rng(10_000 + seed)
    target = float(rng.uniform(0.40, 0.60))
    sdf, ns, frac = E.sphere_pack(n, seed, rmin, rmax, target=target, hardcore=0.85, connect=True)
    s = E.Sbm(sdf, -sdf, E.synthetic_params(), E.tables())
    s.set(F_CNP, E.initial_cnp(n, seed))
    return s, ns, frac


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def generate(n, seeds, steps, rmin, rmax, mode):
    """Device-resident samples [(x bf16 [N,16], target f32 [N])] and stats."""
    N = n ** 3
    samples, its = [], []
    t0 = time.perf_counter()
    for k, seed in enumerate(seeds):
        s, ns, frac = make_state(n, seed, rmin, rmax)
        for t in range(steps):
            x = DeviceArray((N * NN.IN_CH,), np.uint16)
            scale = s.nn_features(x)
            before = interior(s.get(F_PHP))
            st = s.step(mode)
            its.append(st.it_php)
            after = interior(s.get(F_PHP))
            if scale > 0.0:
                # voxel = i + n1 (j + n2 k): Fortran order of the interior
                tgt = ((after - before) / scale).ravel(order="F").astype(np.float32)
                samples.append((x, DeviceArray.from_host(tgt)))
            else:
                x.free()
        s.close()
        log(f"  seed {seed}: {ns} spheres, solid {frac:.3f}; {len(samples)} samples, "
            f"{time.perf_counter() - t0:.0f} s")
    return samples, its


def train(net, samples, n, epochs, lr, seed=0):
    rng = np.random.default_rng(seed)
    dims = (n, n, n)
    pred = DeviceArray((n ** 3,), np.float32)
    total = epochs * len(samples)
    hist, k, t0 = [], 0, time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        run = []
        for i in order:
            x, y = samples[i]
            # cosine decay over the whole run
            cur = lr * 0.5 * (1.0 + np.cos(np.pi * k / max(total, 1)))
            loss = net.loss_grad(dims, x, y, pred)
            net.adam(max(cur, 1e-6))
            run.append(loss)
            k += 1
        hist.append(float(np.mean(run)))
        log(f"  epoch {ep}: mean loss {hist[-1]:.4f} ({time.perf_counter() - t0:.0f} s)")
    return hist


def evaluate(net, n, seeds, steps, rmin, rmax, mode, warmup=2):
    """phP CG iterations and seconds per step with and without the warm
    start, on the same seeds from the same initial states (the first
    `warmup` steps, which cannot use it, are excluded)."""
    res = {"plain": {"it_php": 0, "seconds": 0.0}, "learned": {"it_php": 0, "seconds": 0.0, "used": 0, "steps": 0}}
    for seed in seeds:
        for key in ("plain", "learned"):
            s, _, _ = make_state(n, seed, rmin, rmax)
            if key == "learned":
                s.set_warm_start(net)
            for t in range(steps):
                st = s.step(mode)
                if t >= warmup:
                    res[key]["it_php"] += st.it_php
                    res[key]["seconds"] += st.seconds
                    if key == "learned":
                        res[key]["used"] += st.ws_used
                        res[key]["steps"] += 1
            s.close()
        log(f"  eval seed {seed}: plain {res['plain']['it_php']} learned {res['learned']['it_php']} phP iterations")
    p, l = res["plain"], res["learned"]
    res["iterations_saved"] = 1.0 - l["it_php"] / max(p["it_php"], 1)
    res["time_saved"] = 1.0 - l["seconds"] / max(p["seconds"], 1e-30)
    return res


def parse_seeds(spec, default_start):
    if "-" in spec:
        a, b = spec.split("-")
        return list(range(int(a), int(b) + 1))
    return list(range(default_start, default_start + int(spec)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--train-seeds", default="40")
    ap.add_argument("--test-seeds", default="10")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--eval-steps", type=int, default=12)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--rmin", type=float, default=4.0)
    ap.add_argument("--rmax", type=float, default=8.0)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    E.check(E.lib().elsim_init(0), "elsim_init")
    train_seeds = parse_seeds(a.train_seeds, 0)
    test_seeds = parse_seeds(a.test_seeds, 1000)
    mode = E.MODE_MIXED
    log(f"generating: {len(train_seeds)} microstructures x {a.steps} steps at {a.n}^3")
    t0 = time.perf_counter()
    samples, its = generate(a.n, train_seeds, a.steps, a.rmin, a.rmax, mode)
    t_gen = time.perf_counter() - t0
    log(f"{len(samples)} samples in {t_gen:.0f} s; training {a.epochs} epochs")
    net = NN.WarmStartNet(seed=a.init_seed)
    t0 = time.perf_counter()
    hist = train(net, samples, a.n, a.epochs, a.lr)
    t_train = time.perf_counter() - t0
    for x, y in samples:
        x.free()
        y.free()
    log(f"evaluating on {len(test_seeds)} held-out microstructures x {a.eval_steps} steps")
    ev = evaluate(net, a.n, test_seeds, a.eval_steps, a.rmin, a.rmax, mode)
    report = {"n": a.n, "train_seeds": [train_seeds[0], train_seeds[-1]], "test_seeds": [test_seeds[0], test_seeds[-1]],
              "steps": a.steps, "samples": len(samples), "epochs": a.epochs, "lr": a.lr, "loss_per_epoch": hist,
              "generate_s": t_gen, "train_s": t_train, "updates_per_s": a.epochs * len(samples) / max(t_train, 1e-9),
              "radius": [a.rmin, a.rmax], "eval": ev}
    log(json.dumps(report))
    if a.out:
        net.meta = {"n": a.n, "features": list(NN.FEATURES), "target": "dphi_P / rms(last dphi_P)",
                    "train": report}
        net.save(a.out)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
