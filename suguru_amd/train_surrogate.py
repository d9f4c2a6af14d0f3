"""Train the membrane-traction surrogate on one GPU and report what it gains (one JSON line).

    python -m suguru_amd.train_surrogate [--n-train 200000] [--steps 20000] [--save w.npz]
"""
import argparse
import json

from .surrogate import run_experiment


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=200_000)
    ap.add_argument("--n-test", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=20_000)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--log-every", type=int, default=1000)
    ap.add_argument("--save", default=None)
    ap.add_argument("--shapes", default="flow", choices=["flow", "field"])
    a = ap.parse_args()
    r = run_experiment(n_train=a.n_train, n_test=a.n_test, steps=a.steps, batch=a.batch,
                       lr=a.lr, seed=a.seed, tol=a.tol, log_every=a.log_every, save=a.save,
                       kind=a.shapes)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
