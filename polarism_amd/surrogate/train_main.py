"""Generate surrogate training data with the solver and train the FNO on the GPU.

python -m polarism_amd.surrogate.train_main --members 4096 --epochs 40 --out weights.npz
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=20240)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--max-time", type=float, default=4000.0)
    ap.add_argument("--train-seconds", type=float, default=None)
    ap.add_argument("--init", default=None, help="continue from these weights (.npz)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    from . import warmstart as ws
    log = lambda m: print(m, flush=True)
    t0 = time.perf_counter()
    members = ws.make_members(args.members, args.seed)
    log(f"built {len(members)} members in {time.perf_counter() - t0:.1f} s")
    x, y, conv, res = ws.generate(members, rtol=args.rtol, max_time=args.max_time, log=log)
    keep = np.flatnonzero(conv)
    log(f"training on {keep.size} converged steady states")
    init, cfg0 = None, None
    if args.init:
        init, cfg0, _ = ws.load_weights(args.init)
        log(f"continuing from {args.init}")
    params, hist, cfg = ws.train(x[:, keep], y[:, keep], epochs=args.epochs, batch=args.batch, lr=args.lr,
                                 log=log, time_limit=args.train_seconds, init=init, cfg=cfg0)
    meta = {"members": args.members, "seed": args.seed, "converged": int(keep.size), "epochs_run": len(hist),
            "init": args.init or "",
            "final_loss": hist[-1], "data_rtol": args.rtol, "data_max_time": args.max_time}
    ws.save_weights(args.out, params, cfg, ws.default_normaliser(), meta)
    log(json.dumps({"weights": args.out, **meta, "loss_history": hist}))


if __name__ == "__main__":
    sys.exit(main())
