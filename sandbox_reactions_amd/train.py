"""Training of the learned FLD preconditioner's network on the GPU, by the library's own forward, backward and Adam
kernels (sr_net_loss_grad / sr_net_adam).

The data are stiff FLD systems drawn like the 2048^2 benchmark system (rho lognormal sigma_ln 1.5, T lognormal 0.8,
P(k) ~ k^-3, Kramers opacity, diffusion number 1e3) at 256^2 and 512^2, each with several right-hand sides. For a
right-hand side r the network sees what one multigrid V-cycle leaves, r1 = r - A V r, and learns the correction
e1 = A^-1 r1, which the plain MG-PCG solver computes (to 1e-6); the loss is the A-norm of the error,
(N(r1) - e1)^T A (N(r1) - e1) / e1^T A e1. So the learned preconditioner V r + N(r - A V r) is trained for exactly
the part plain multigrid misses.

    python -m sandbox_reactions_amd.train --out sandbox_reactions_amd/weights/fld_unet.npz [--solves 4000]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "fld_unet.npz")
TRAIN_SEED = 0x7EA1000


def system_tensors(n, seed):
    import torch
    from . import scenes
    c, kx, ky, _, _ = scenes.fld_system(n, seed=seed)
    t = lambda a: torch.tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    return t(c), t(kx), t(ky)


def make_sample(f, own, kx, ky, gen, tol=1e-6, cg_steps=0):
    """(r1, e1) for one right-hand side: white noise in the diagonally scaled space (b = d^1/2 xi); with
    cg_steps > 0, the residual plain MG-PCG leaves after that many iterations on b (the residuals the
    preconditioner meets inside the solve), else b itself."""
    import torch
    from . import PRECOND_MULTIGRID
    h, w = own.shape
    d = own.clone()
    d += kx[:, :w] + kx[:, 1:] + ky[:h, :] + ky[1:, :]
    r = (torch.randn((h, w), generator=gen, device="cuda") * d.sqrt()).contiguous()
    if cg_steps > 0:
        b = r.double().contiguous()
        x = torch.zeros_like(b)
        f.solve(b, x, tol=1e-14, max_iter=cg_steps)
        ax = torch.empty_like(b)
        f.matvec(x, ax)
        r = (b - ax).float().contiguous()
    z = torch.empty_like(r)
    f.apply_precond(r, z, PRECOND_MULTIGRID)
    az = torch.empty((h, w), dtype=torch.float64, device="cuda")
    f.matvec(z.double().contiguous(), az)
    r1 = (r.double() - az).contiguous()
    e1 = torch.zeros_like(r1)
    f.solve(r1, e1, tol=tol)
    return r1.float().contiguous(), e1.float().contiguous()


def train(solves=4000, sizes=(256, 512), systems=24, per_system_batch=8, lr=2e-3, lr_final=2e-4, seed=0,
          log_every=200, params=None, verbose=True, cg_max=40):
    """Returns (flat parameters, loss history). `solves` right-hand sides in all, spread over `systems` systems
    per size; each system's samples are drawn in batches of `per_system_batch` and reused twice."""
    import torch
    import sandbox_reactions_amd as sr
    from .net import Net, init_params
    nmax = max(sizes)
    net = Net(nmax, nmax, params if params is not None else init_params(seed))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed + 1)
    solvers = {n: sr.Fld(n, n) for n in sizes}
    history = []
    done = 0
    step = 0
    t0 = time.perf_counter()
    k = 0
    while done < solves:
        n = sizes[k % len(sizes)]
        sys_seed = TRAIN_SEED + 97 * (k % (systems * len(sizes)))
        k += 1
        own, kx, ky = system_tensors(n, sys_seed)
        f = solvers[n]
        f.set_system(own, kx, ky)
        # right-hand sides as plain MG-PCG meets them: after 0 .. cg_max iterations
        steps = np.random.default_rng(seed + k).integers(0, cg_max + 1, per_system_batch) if cg_max > 0 \
            else np.zeros(per_system_batch, int)
        batch = [make_sample(f, own, kx, ky, gen, cg_steps=int(c)) for c in steps]
        done += len(batch)
        for rep in range(2):
            for r1, e1 in batch:
                frac = min(1.0, step / max(1, 2 * solves))
                rate = lr * (lr_final / lr) ** frac  # geometric decay
                loss = net.loss_grad(own, kx, ky, r1, e1)
                net.adam(rate)
                history.append(loss)
                step += 1
                if verbose and step % log_every == 0:
                    print(f"train step {step}: solves {done}, size {n}, loss {np.mean(history[-log_every:]):.4f}, "
                          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    return net.params(), np.array(history)


def save(path, flat, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=flat, **{k: np.asarray(v) for k, v in meta.items()})


def load(path=DEFAULT_WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--solves", type=int, default=4000)
    ap.add_argument("--systems", type=int, default=24)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    import sandbox_reactions_amd as sr
    sr.require_device()
    t0 = time.perf_counter()
    flat, hist = train(solves=a.solves, systems=a.systems, lr=a.lr, seed=a.seed)
    el = time.perf_counter() - t0
    save(a.out, flat, dict(loss_first=hist[:100].mean(), loss_last=hist[-200:].mean(), solves=a.solves,
                           seconds=el))
    print(f"trained on {a.solves} solves in {el:.1f} s: loss {hist[:100].mean():.4f} -> {hist[-200:].mean():.4f}; "
          f"saved {a.out}", flush=True)


if __name__ == "__main__":
    main()
