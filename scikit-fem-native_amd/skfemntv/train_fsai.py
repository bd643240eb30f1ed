"""Train the learned-FSAI network on systems produced by the batched assembler.

    python -m skfemntv.train_fsai [--systems 2000] [--epochs 40] [--out PATH]

Training set: ``--systems`` P1 anisotropic-diffusion systems (seed 1000)
with per-axis cells i.i.d. in {8..16}; validation: 64 systems (seed 2000).
Both are assembled on the GPU in batches of 250 systems (one plan and one
assembly per batch) and condensed; each minibatch is a group of
``--group`` consecutive systems.  Forward, objective, backward and Adam all
run in libskfem_amd kernels.  The parameters and the training record are
written to ``--out`` (default: the package's pretrained weights file).
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(msg):
    print(f"[train_fsai] {msg}", flush=True)


def build_groups(count, lo, hi, seed, group, chunk=250):
    from .batch import AnisotropicBatch, batch_shapes
    from .fsai import split_groups
    shapes = batch_shapes(count, lo, hi, seed)
    groups = []
    rows = 0
    for c0 in range(0, count, chunk):
        b = AnisotropicBatch(shapes[c0:c0 + chunk], seed=seed + c0)
        A, _ = b.assemble()
        groups += split_groups(A, b.sys_ptr_host, b.xyz, group)
        rows += A.shape[0]
        del b
    return groups, shapes, rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=2000)
    ap.add_argument("--valid", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--group", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lo", type=int, default=8)
    ap.add_argument("--hi", type=int, default=16)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    import torch
    from . import fsai
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    train_groups, shapes, rows = build_groups(args.systems, args.lo, args.hi, 1000, args.group)
    valid_groups, _, vrows = build_groups(args.valid, args.lo, args.hi, 2000, args.group)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    nnz = sum(g.nnz for g in train_groups)
    log(f"{args.systems} training systems ({rows} unknowns, {nnz} lower nonzeros), "
        f"{args.valid} validation systems; setup {setup:.1f}s")
    net = fsai.FSAINetwork(seed=args.seed)
    v0 = fsai.mean_loss(net, valid_groups)
    log(f"validation loss before training (Jacobi): {v0:.5f}")
    hist = fsai.train(net, train_groups, epochs=args.epochs, lr=args.lr, valid=valid_groups, seed=args.seed,
                      log=log)
    log(f"trained {hist.steps} steps in {hist.seconds:.1f}s; validation loss {hist.valid_loss[-1]:.5f}")
    out = args.out or fsai.WEIGHTS_PATH
    record = {"systems": args.systems, "valid_systems": args.valid, "cells_per_axis": [args.lo, args.hi],
              "train_seed": 1000, "valid_seed": 2000, "init_seed": args.seed, "epochs": args.epochs,
              "group": args.group, "lr": args.lr, "steps": hist.steps, "train_seconds": hist.seconds,
              "setup_seconds": setup, "valid_loss_initial": v0, "train_loss": hist.train_loss,
              "valid_loss": hist.valid_loss, "lower_nonzeros": nnz, "unknowns": rows}
    net.save(out, record=json.dumps(record))
    log(f"wrote {out}")
    print(json.dumps({k: record[k] for k in ("systems", "epochs", "steps", "train_seconds", "valid_loss_initial")}
                     | {"valid_loss_final": hist.valid_loss[-1]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
