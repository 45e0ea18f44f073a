"""Training of the learned preconditioner across angles of attack (no reference counterpart).

The network works per brick (its input is the post-V-cycle residual of a brick and its
neighbours, scaled by the brick's RMS, and one weight set per AMR level), so weights trained on
one wing apply to any other.  `train_across_angles` builds wings at several angles of attack
(topologies built on the device), runs each with the learned preconditioner, records
post-V-cycle residuals of its own MGPCG solves and continues training the same weights on
them, instance after instance.  The result seeds the solver of the target case, which then
fine-tunes on its own residuals and validates the network against plain MGPCG.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from .core import ExternalAeroCore


def default_angles(count, lo=0.0, hi=12.0, seed=2412):
    """`count` angles of attack drawn uniformly from [lo, hi] degrees (fixed seed), sorted."""
    rng = np.random.default_rng(seed)
    return sorted(float(a) for a in rng.uniform(lo, hi, size=count))


def train_across_angles(make_topology, angles, options, collects=4, pairs=16, train_steps=200, batch=4, lr=1e-3,
                        weights=None, device=0, log=None):
    """Train one network on the instances make_topology(alpha, device) for alpha in `angles`.

    options: solver options with preconditioner PC_LEARNED (learned_capacity >= collects * pairs).
    Per instance: initialize, two steps, then `collects` times (record up to `pairs` residuals of
    the next projection, step), then `train_steps` Adam steps from the weights so far.
    Returns (weights, biases, stats) with per-instance statistics and the totals of residuals and
    brick samples (a residual contributes one sample per network brick).
    """
    if options.preconditioner != L.PC_LEARNED:
        raise ValueError("options must select the learned preconditioner")
    stats = {"instances": [], "residuals": 0, "brick_samples": 0, "seconds": 0.0}
    t_all = time.perf_counter()
    for alpha in angles:
        t0 = time.perf_counter()
        topo = make_topology(alpha, device)
        core = ExternalAeroCore(topo, options)
        try:
            core.set_reference(topo.info.reference_area)
            core.initialize()
            if weights is not None:
                core.set_learned_weights(*weights)
            for _ in range(2):
                core.step()
            st = None
            for _ in range(collects):
                st = core.learned_collect(max_pairs=pairs)
                core.step()
            tr = core.learned_train(steps=train_steps, batch=batch, lr=lr)
            weights = core.learned_weights()
        finally:
            core.close()
        rec = {"alpha": alpha, "cells": topo.n_active, "bricks": tr["bricks"], "residuals": st["pairs"],
               "loss_first": tr["loss_first"], "loss_last": tr["loss_last"], "seconds": time.perf_counter() - t0}
        stats["instances"].append(rec)
        stats["residuals"] += rec["residuals"]
        stats["brick_samples"] += rec["residuals"] * rec["bricks"]
        if log:
            log(f"pretrain alpha {alpha:.2f}: {rec['cells']} cells, {rec['residuals']} residuals on {rec['bricks']} "
                f"bricks, loss {rec['loss_first']:.4f} -> {rec['loss_last']:.4f}, {rec['seconds']:.1f}s")
    stats["seconds"] = time.perf_counter() - t_all
    return weights[0], weights[1], stats
