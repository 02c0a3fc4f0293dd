"""Training of the learned AAR Poisson preconditioner (sgpu_unet_*) on pairs
the solver itself produces.

The right-hand sides are the Poisson problems of the first SCF iterations of
psp8 systems (rhs = 4 pi (b + rho) for the guess density and the densities
after 1, 2 and 4 iterations, Calculate_elecstPotential,
src/electrostatics.c:1494-1699); every second iterate of a Jacobi-AAR solve of
each gives a (residual, error) pair (sgpu_unet_harvest).  Pairs of several
grids train one set of weights: Adam runs in blocks of steps on each grid's
network in turn, the weights carried from one to the next
(gpu.train_multigrid).  After every block the current weights solve a
held-out problem per system (the density after 3 SCF iterations) with AAR to
the same tolerance, and the earliest weights with the fewest AAR iterations
summed over the systems are kept (early stopping): a block after which the
network stops preconditioning (which longer training can reach) is never the
one returned, and of equally good checkpoints the one trained least.
"""
from __future__ import annotations

import time

import numpy as np

from . import gpu
from .systems import scf_system


def scf_densities(ctx, st, p, iterations=(1, 2, 4)):
    """The guess density and the mixed densities after the given numbers of
    SCF iterations (mixed precision)."""
    rhos = [st.rho_at]
    for k in iterations:
        r = ctx.scf_run(st.b, st.rho_at, dict(p, max_scf=k, dtype=gpu.F32), rho_core=st.rho_core)
        rhos.append(r["rho"])
    return rhos


def train_poisson_unet(systems=("si8", "si64"), steps: int = 400, block: int = 50,
                       lr: float = 2e-3, batch: int = 4, seed: int = 5, max_pairs: int = 120,
                       validate: bool = True):
    """Returns (weights, info): the FP32 weights and a record of the harvest,
    the training and the validation solves."""
    t0 = time.perf_counter()
    ctxs, nets, pairs = [], [], {}
    vals = []
    try:
        for name in systems:
            st, p = scf_system(name)
            ctx = gpu.Context(st.grid)
            ctxs.append(ctx)
            ctx.set_nonlocal(st.nloc)
            rhos = scf_densities(ctx, st, p)
            net = ctx.unet(seed=1)
            nets.append(net)
            for rho in rhos:
                net.harvest(4.0 * np.pi * (st.b + rho), np.zeros(st.grid.nd), tol=1e-8,
                            max_iter=3000, stride=2, max_pairs=max_pairs)
            pairs[name] = net.npairs
            if validate:
                rho3 = scf_densities(ctx, st, p, (3,))[1]
                vals.append((ctx, net, 4.0 * np.pi * (st.b + rho3), st.grid.nd))
        t_harvest = time.perf_counter() - t0
        for net in nets[1:]:
            net.set_weights(nets[0].get_weights())

        def val_iters(w):
            tot = 0
            for ctx, net, rhs, nd in vals:
                net.set_weights(w)
                _, it, _, stt = ctx.AAR_unet(net, np.zeros(nd), rhs, tol=1e-8, max_iter=3000)
                if stt != 0:
                    return 10 ** 9
                tot += it
            return tot

        t1 = time.perf_counter()
        hist = []
        w = nets[0].get_weights()
        best = (val_iters(w) if vals else 0, 0, w)  # (iterations, steps, weights): U = 0 first
        val_log = [(0, best[0])]
        done = 0
        while done < steps:
            # one block on every grid, the weights carried along
            n = min(block * len(nets), steps - done)
            h, w = gpu.train_multigrid(nets, n, batch=batch, lr=lr, seed=seed + done, block=block)
            hist.append(h)
            done += n
            if vals:
                it = val_iters(w)
                val_log.append((done, it))
                if it < best[0]:
                    best = (it, done, w)
            else:
                best = (0, done, w)
        t_train = time.perf_counter() - t1
        hist = np.concatenate(hist) if hist else np.zeros(0)
        info = {"systems": list(systems), "pairs": pairs, "steps": steps, "block": block,
                "batch": batch, "lr": lr, "harvest_s": round(t_harvest, 2),
                "train_s": round(t_train, 2),
                "loss_first": float(hist[:5].mean()) if hist.size else None,
                "loss_last": float(hist[-20:].mean()) if hist.size else None,
                "validation_aar_iterations": val_log, "kept_after_steps": best[1],
                "source": "Jacobi-AAR Poisson solves (tol 1e-8) of the SCF densities (guess and "
                          "after 1, 2, 4 iterations) of " + ", ".join(systems)
                          + "; validation: the densities after 3 iterations of " + ", ".join(systems)
                          + " (AAR iterations summed, the earliest of the fewest kept)"}
        return best[2], info
    finally:
        for net in nets:
            net.close()
        for ctx in ctxs:
            ctx.close()
