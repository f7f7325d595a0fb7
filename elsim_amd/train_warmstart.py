"""Train the learned CG warm start on solver-generated pairs, on the GPU.

    python -m elsim_amd.train_warmstart --n 128 --train-seeds 40 --test-seeds 10 \\
        --steps 50 --epochs 3 --eval-n 256 --out weights.npz

Pairs: for each seeded sphere pack and each step (from the third) of a run
with the fast solver, the input is the network features of the extrapolated
start of the step's first phP solve (psi_P, its FP64 residual, the
through-thickness coordinate, log10 of the coefficient and phi_P's last
change; see elsim_sbm_ws_record) and the target is the converged phi_P minus
that start, in units of its rms.  Samples stay in device memory.
Training: Adam on the squared A-norm error of the prediction (the norm CG
minimises, A the phP operator of the sample's microstructure) or on the
mean squared error, one volume per update, with the project's own
forward/backward kernels.  Evaluation: held-out seeds run the
same steps with the plain, the extrapolated and the learned warm start; phP
CG iterations and time are reported.
"""
import argparse
import json
import sys
import time

import numpy as np

import elsim_amd as E
from elsim_amd import nn as NN
from elsim_amd._lib import F_CNP, F_KAPP, F_PHP, DeviceArray


def log(*a):
    print(*a, flush=True)


def make_state(n, seed, rmin, rmax):
    rng = np.random.default_rng(10_000 + seed)
    target = float(rng.uniform(0.40, 0.60))
    sdf, ns, frac = E.sphere_pack(n, seed, rmin, rmax, target=target, hardcore=0.85, connect=True)
    # the benchmark's workload: the CnE solve resolves the step's change
    s = E.Sbm(sdf, -sdf, E.synthetic_params(cne_rule=1), E.tables())
    s.set(F_CNP, E.initial_cnp(n, seed))
    return s, ns, frac


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def generate(n, seeds, steps, rmin, rmax, mode):
    """Device-resident samples [(x bf16 [N,16], target f32 [N], kap f32
    ghosted)].  At every step from the third on, the solver records the
    network input of its extrapolated phP start (elsim_sbm_ws_record); the
    target is the converged phi_P minus that start, in units of its rms (the
    network predicts the shape of the remaining error, the solver's Galerkin
    step its size).  kap: the phP coefficient of the microstructure."""
    N = n ** 3
    samples, its = [], []
    t0 = time.perf_counter()
    for k, seed in enumerate(seeds):
        s, ns, frac = make_state(n, seed, rmin, rmax)
        s.set_params(s.params.copy(ws_mode=E.WS_LEARNED))
        kap = DeviceArray.from_host(s.get(F_KAPP).ravel(order="F").astype(np.float32))
        for t in range(steps):
            x = start = None
            if t >= 2:
                x = DeviceArray((N * NN.IN_CH,), np.uint16)
                start = DeviceArray((N,), np.float64)
                s.ws_record(x, start)
            st = s.step(mode)
            its.append(st.it_php)
            if x is not None:
                # voxel = i + n1 (j + n2 k): Fortran order of the interior
                e = interior(s.get(F_PHP)).ravel(order="F") - start.to_host()
                start.free()
                rms = float(np.sqrt(np.mean(e ** 2)))
                if rms > 0.0 and np.isfinite(rms):
                    samples.append((x, DeviceArray.from_host((e / rms).astype(np.float32)), kap))
                else:
                    x.free()
        s.close()
        log(f"  seed {seed}: {ns} spheres, solid {frac:.3f}; {len(samples)} samples, "
            f"{time.perf_counter() - t0:.0f} s")
    return samples, its


def train(net, samples, n, epochs, lr, loss_kind="energy", seed=0):
    rng = np.random.default_rng(seed)
    dims = (n, n, n)
    pred = DeviceArray((n ** 3,), np.float32)
    total = epochs * len(samples)
    hist, k, t0 = [], 0, time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        run = []
        for i in order:
            x, y, kap = samples[i]
            # cosine decay over the whole run
            cur = lr * 0.5 * (1.0 + np.cos(np.pi * k / max(total, 1)))
            if loss_kind == "energy":
                net.train_forward(dims, x, pred)
                loss = net.energy_loss(pred, y, kap)
                net.backward(x)
            else:
                loss = net.loss_grad(dims, x, y, pred)
            net.adam(max(cur, 1e-6))
            run.append(loss)
            k += 1
            if k % 400 == 0:
                log(f"    update {k}/{total}: loss {np.mean(run[-400:]):.4f} ({time.perf_counter() - t0:.0f} s)")
        hist.append(float(np.mean(run)))
        log(f"  epoch {ep}: mean loss {hist[-1]:.4f} ({time.perf_counter() - t0:.0f} s)")
    return hist


def evaluate(net, n, seeds, steps, rmin, rmax, mode, warmup=2, modes=(E.WS_PLAIN, E.WS_EXTRAP, E.WS_LEARNED)):
    """phP CG iterations and seconds per step of the same runs (same seeds,
    same initial states) for each warm-start mode; the first `warmup` steps,
    which no warm start applies to, are excluded."""
    res = {int(m): {"it_php": 0, "seconds": 0.0, "used": 0, "steps": 0} for m in modes}
    for seed in seeds:
        for m in modes:
            s, _, _ = make_state(n, seed, rmin, rmax)
            s.set_params(s.params.copy(ws_mode=int(m)))
            if m == E.WS_LEARNED:
                s.set_warm_start(net)
            for t in range(steps):
                st = s.step(mode)
                if t >= warmup:
                    r = res[int(m)]
                    r["it_php"] += st.it_php
                    r["seconds"] += st.seconds
                    r["used"] += int(st.ws_used == (2 if m == E.WS_LEARNED else 1))
                    r["steps"] += 1
            s.close()
        log(f"  eval seed {seed}: " + ", ".join(f"mode {m} {res[int(m)]['it_php']}" for m in modes))
    base = res[int(modes[0])]
    for m in modes:
        r = res[int(m)]
        r["iterations_saved_vs_plain"] = 1.0 - r["it_php"] / max(base["it_php"], 1)
        r["time_saved_vs_plain"] = 1.0 - r["seconds"] / max(base["seconds"], 1e-30)
    return res


def parse_seeds(spec, default_start):
    if "-" in spec:
        a, b = spec.split("-")
        return list(range(int(a), int(b) + 1))
    return list(range(default_start, default_start + int(spec)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--eval-n", type=int, default=0, help="also evaluate at this size (held-out seeds)")
    ap.add_argument("--eval-n-seeds", type=int, default=3)
    ap.add_argument("--train-seeds", default="40")
    ap.add_argument("--test-seeds", default="10")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--eval-steps", type=int, default=12)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--rmin", type=float, default=6.0)
    ap.add_argument("--rmax", type=float, default=16.0)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--loss", choices=["energy", "mse"], default="mse")
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
    hist = train(net, samples, a.n, a.epochs, a.lr, a.loss)
    t_train = time.perf_counter() - t0
    for x, y, _ in samples:
        x.free()
        y.free()
    log(f"evaluating on {len(test_seeds)} held-out microstructures x {a.eval_steps} steps")
    ev = evaluate(net, a.n, test_seeds, a.eval_steps, a.rmin, a.rmax, mode)
    ev_big = None
    if a.eval_n:
        log(f"evaluating at {a.eval_n}^3 on {a.eval_n_seeds} held-out microstructures x {a.eval_steps} steps")
        ev_big = evaluate(net, a.eval_n, test_seeds[:a.eval_n_seeds], a.eval_steps, a.rmin, a.rmax, mode,
                          warmup=min(6, a.eval_steps - 2))
    report = {"n": a.n, "train_seeds": [train_seeds[0], train_seeds[-1]], "test_seeds": [test_seeds[0], test_seeds[-1]],
              "steps": a.steps, "samples": len(samples), "epochs": a.epochs, "loss": a.loss, "lr": a.lr, "loss_per_epoch": hist,
              "generate_s": t_gen, "train_s": t_train, "updates_per_s": a.epochs * len(samples) / max(t_train, 1e-9),
              "radius": [a.rmin, a.rmax], "eval": ev,
              "eval_large": {"n": a.eval_n, "seeds": test_seeds[:a.eval_n_seeds], "eval": ev_big} if ev_big else None}
    log(json.dumps(report))
    if a.out:
        net.meta = {"n": a.n, "features": list(NN.FEATURES),
                    "target": "(converged phi_P - extrapolated start) / rms",
                    "train": report}
        net.save(a.out)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
