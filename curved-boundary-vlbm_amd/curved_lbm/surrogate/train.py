"""Training of the warm-start network on solver equilibria, and its use.

    python -m curved_lbm.surrogate.train --family shell --samples 512 --out shell.npz

draws ``samples`` loads of the family from numpy's default_rng(seed),
solves each to the displacement-rate residual on the family's coarse
lattice (datagen), trains the network on them with Adam (every kernel in
libvlbm.so) and writes the parameters, the training history and the held-out
errors to ``out``.  ``WarmStartModel`` loads such a file, predicts u and F
for a load and starts a solver from them (warm_start.run_warm).
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np
import torch

from .. import warm_start as ws
from . import datagen
from .families import FAMILIES
from .network import CH, TARGET_CH, ConvOperator


def _pad(a: np.ndarray) -> np.ndarray:
    """(..., X, Y, Z, C) -> zero halo on the three lattice axes."""
    pad = [(0, 0)] * (a.ndim - 4) + [(1, 1)] * 3 + [(0, 0)]
    return np.pad(a, pad)


def inputs(family, n, p, device) -> torch.Tensor:
    """Network input (1, X+2, Y+2, Z+2, 32) BF16 for load p."""
    f = family.features(n, p)
    x = np.zeros(f.shape[:3] + (CH,), np.float32)
    x[..., :f.shape[-1]] = f
    return torch.from_numpy(_pad(x)[None]).to(device=device, dtype=torch.bfloat16)


def targets(family, n, p, u, F, device) -> torch.Tensor:
    """Target (1, X+2, Y+2, Z+2, 16) FP32: (u - u_prior) / scale,
    (F - F_prior) / scale, channel 15 the active mask."""
    up, Fp = family.prior(n, p)
    s = family.scale(p)
    t = np.zeros(family.lattice(n) + (TARGET_CH,), np.float32)
    t[..., 0:3] = np.moveaxis((u - up) / s, 0, -1)
    t[..., 3:12] = np.moveaxis((F - Fp) / s, 0, -1)
    t[..., 15] = family.geometry(n).active
    return torch.from_numpy(_pad(t)[None]).to(device)


def decode(family, n, p, y: torch.Tensor):
    """(u (3, X, Y, Z), F (9, X, Y, Z)) float64 from the network output
    y (1, X+2, Y+2, Z+2, 32)."""
    up, Fp = family.prior(n, p)
    s = family.scale(p)
    core = y[0, 1:-1, 1:-1, 1:-1].permute(3, 0, 1, 2).double()
    dev = y.device
    u = torch.from_numpy(up).to(dev) + s * core[0:3]
    F = torch.from_numpy(Fp).to(dev) + s * core[3:12]
    return u, F


def train(family, n, loads, U, F, *, epochs=40, batch=8, lr=2e-3, seed=0, val_fraction=0.1, device="cuda:0",
          log=print):
    N = len(loads)
    n_val = max(1, int(round(val_fraction * N))) if N > 1 else 0
    order = np.random.default_rng(seed).permutation(N)
    val, trn = order[:n_val], order[n_val:]
    X = torch.cat([inputs(family, n, loads[k], device) for k in range(N)])
    T = torch.cat([targets(family, n, loads[k], U[k].astype(np.float64), F[k].astype(np.float64), device)
                   for k in range(N)])
    op, rec = fit(X, T, trn, family.lattice(n), n_in=family.n_in, active=float(family.geometry(n).active.sum()),
                  name=family.name, epochs=epochs, batch=batch, lr=lr, seed=seed, device=device, log=log)
    errs = evaluate(op, family, n, [loads[k] for k in val], U[val], F[val], device) if n_val else {}
    rec.update(val_index=val.tolist(), val_errors=errs)
    return op, rec


def fit(X, T, trn, shape, *, n_in, active, name, epochs=40, batch=8, lr=2e-3, seed=0, device="cuda:0", log=print):
    """Trains a ConvOperator on inputs X (N, padded, 32) BF16 and targets T
    (N, padded, 16) FP32 (channels 0-11 the prediction, 15 the voxel weight)
    over the samples ``trn``; ``active`` is the number of weighted voxels per
    sample.  Returns the operator with the best epoch's parameters and the
    training record."""
    batch = min(batch, len(trn))
    op = ConvOperator(shape, batch=batch, n_in=n_in, n_out=12, device=device, seed=seed)
    # every output channel weighted by the inverse of its mean square over
    # the training targets, so displacement and gradient rows count alike;
    # a channel the prior already fits (mean square near zero) is floored at
    # 1% of the channel mean, since the inverse of a near-zero mean square
    # makes the loss follow that channel's noise and the training unstable
    sq = torch.zeros(12, dtype=torch.float64, device=device)
    cnt = 0.0
    for k in trn:  # sample by sample: a fine-lattice target set is several GB
        act = T[int(k), ..., 15] > 0
        sq += (T[int(k)][act][:, :12].double() ** 2).sum(0)
        cnt += float(act.sum())
    ms = (sq / max(cnt, 1.0)).float()
    ms = ms.clamp(min=max(1e-2 * float(ms.mean()), 1e-12))
    chw = torch.zeros(CH, dtype=torch.float32, device=device)
    chw[:12] = (1.0 / ms).float()
    inv = 1.0 / (batch * active * 12)
    tbuf = torch.zeros((batch,) + T.shape[1:], dtype=T.dtype, device=device)
    rng = np.random.default_rng(seed + 1)
    steps_per_epoch = max(1, len(trn) // batch)
    total = epochs * steps_per_epoch
    hist = []
    t0 = time.perf_counter()
    it = 0
    # A single oversized Adam step can switch off most ReLUs at once, after
    # which the network only outputs its last bias.  Guard: a batch loss that
    # is not finite or exceeds ten times the running mean (after a short
    # warm-up) rolls the parameters back to the best epoch so far, clears
    # the Adam moments and halves the learning rate; the best epoch's
    # parameters are the result.
    best_loss, best = math.inf, op.params.detach().clone()
    ema, lr_scale, rollbacks = None, 1.0, 0
    for ep in range(epochs):
        perm = rng.permutation(trn)
        acc, used = 0.0, 0
        for b in range(steps_per_epoch):
            idx = torch.from_numpy(perm[b * batch:(b + 1) * batch]).to(device)
            torch.index_select(X, 0, idx, out=op.acts[0])
            torch.index_select(T, 0, idx, out=tbuf)
            op.forward()
            loss = op.loss_and_backward(tbuf, chw, inv)
            op.stream.synchronize()
            lv = float(loss[0].item())
            it += 1
            if not math.isfinite(lv) or (ema is not None and it > 50 and lv > 10.0 * ema):
                op.params.copy_(best)
                op.adam_m.zero_()
                op.adam_v.zero_()
                op.adam_t = 0
                op._pack()
                lr_scale *= 0.5
                rollbacks += 1
                log(f"[train] {name} epoch {ep + 1}: batch loss {lv:.3e} against a running "
                    f"{ema:.3e}; back to the best epoch, learning rate x{lr_scale:g}")
                continue
            ema = lv if ema is None else 0.95 * ema + 0.05 * lv
            acc += lv
            used += 1
            rate = lr_scale * lr * 0.5 * (1.0 + math.cos(math.pi * (it - 1) / total))
            op.adam_step(max(rate, 1e-5 * lr_scale))
        loss = acc / max(used, 1)
        hist.append(loss)
        if used and loss < best_loss:
            best_loss = loss
            best.copy_(op.params)
        if ep % max(1, epochs // 10) == 0 or ep == epochs - 1:
            log(f"[train] {name} epoch {ep + 1}/{epochs} loss {loss:.4e} ({time.perf_counter() - t0:.1f} s)")
    op.params.copy_(best)
    op._pack()
    train_s = time.perf_counter() - t0
    return op, {"loss_history": hist, "train_s": train_s, "iterations": total, "batch": batch,
                "channel_weights": chw[:12].tolist(), "best_loss": best_loss, "rollbacks": rollbacks}


def evaluate(op, family, n, loads, U, F, device):
    """Relative l2 errors of the decoded prediction against solver
    equilibria over the active cells (u, F - I), and of the prior alone."""
    op.resize_batch(1)
    act = torch.from_numpy(family.geometry(n).active.astype(bool)).to(device)
    eye = torch.eye(3, dtype=torch.float64, device=device).reshape(9, 1, 1, 1)
    out = {"u": [], "F_minus_I": [], "prior_u": [], "prior_F_minus_I": []}
    for k, p in enumerate(loads):
        op.set_input(inputs(family, n, p, device))
        u, Fp = decode(family, n, p, op.forward())
        ut = torch.from_numpy(U[k].astype(np.float64)).to(device)
        Ft = torch.from_numpy(F[k].astype(np.float64)).to(device)
        up, Fpr = (torch.from_numpy(a).to(device) for a in family.prior(n, p))

        def rel(a, b):
            return float(torch.linalg.norm((a - b)[:, act]) / torch.linalg.norm(b[:, act]))

        out["u"].append(rel(u, ut))
        out["F_minus_I"].append(rel(Fp - eye, Ft - eye))
        out["prior_u"].append(rel(up, ut))
        out["prior_F_minus_I"].append(rel(Fpr - eye, Ft - eye))
    return {k: {"mean": float(np.mean(v)), "max": float(np.max(v))} for k, v in out.items()}


class WarmStartModel:
    """A trained network for one family: predict(p) gives the network's u
    and F on the family's coarse lattice; prepare(p) makes them invariant
    under the family's symmetry group and re-integrates u from F
    (warm_start.symmetrize, warm_start.integrate_displacement); start(solver,
    p, ...) carries the prepared fields to the solver's lattice and runs it
    to the residual with the fallback."""

    def __init__(self, path, device="cuda:0"):
        z = np.load(path, allow_pickle=False)
        meta = json.loads(str(z["meta"]))
        self.family = FAMILIES[meta["family"]]
        self.n = int(meta["n"])
        self.meta = meta
        self.device = device
        self.op = ConvOperator(self.family.lattice(self.n), batch=1, layers=int(meta["layers"]),
                               n_in=self.family.n_in, n_out=12, device=device)
        self.op.load_state(z["params"])
        self._mask = torch.from_numpy(self.family.geometry(self.n).active.astype(bool)).to(device)

    def predict(self, p):
        self.op.set_input(inputs(self.family, self.n, p, self.device))
        return decode(self.family, self.n, p, self.op.forward())

    def prepare(self, p, *, symmetric=True, integrate=None, integrate_tol=1e-6):
        """(u, F, info) on the coarse lattice.  The network's error breaks the
        load's symmetry and so excites the lowest antisymmetric modes
        (rocking and torsion of the shell, bending of the tube), which the
        damped dynamics remove at omega^2 / gamma, many times slower than the
        symmetric error; the group average removes those components.  u is
        then re-integrated from F (least squares pinned at the Dirichlet
        walls) starting from the network's u, which leaves u the displacement
        of the predicted F (by default where the family asks for it)."""
        u, F = self.predict(p)
        info = {}
        if integrate is None:
            integrate = self.family.integrate_u
        if symmetric:
            u, F = ws.symmetrize(u, F, self._mask, self.family.symmetry_group())
        if integrate:
            g = self.family.geometry(self.n)
            u, info = ws.integrate_displacement(F, self._mask, (g.dx, g.dy, g.dz), self.family.anchors(self.n, p),
                                                u0=u, tol=integrate_tol)
        return u, F, info

    def warm_up(self, fine_shape) -> float:
        """Runs the prediction path once (a training load, carried to
        fine_shape) so that the first real prediction does not pay the
        one-time loading of the torch kernels it uses; returns the seconds
        taken."""
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        u, F, _ = self.prepare(self.meta["loads"][0])
        ws.carry_to_lattice(u, F, self._mask, tuple(fine_shape))
        torch.cuda.synchronize(self.device)
        return time.perf_counter() - t0

    def start(self, solver, p, *, tol=1e-9, check_every=200, max_steps=10**6, symmetric=True, integrate=None,
              **kw):
        """Warm-started run of a solver of the same family (any lattice
        side that refines the coarse one); returns the run record with the
        preparation time included in wall_s."""
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        u, F, info = self.prepare(p, symmetric=symmetric, integrate=integrate)
        uf, Ff = ws.carry_to_lattice(u, F, self._mask, (solver.nx, solver.ny, solver.nz))
        torch.cuda.synchronize(self.device)
        t_pred = time.perf_counter() - t0
        ramp_time = self.family.ramp_steps(solver.nx if self.family.name == "shell" else solver.ny) * solver.dt
        kw.setdefault("max_initial_residual", self.family.max_initial_residual)
        out = ws.run_warm(solver, uf, Ff, t_start=ramp_time, tol=tol, check_every=check_every, max_steps=max_steps,
                          **kw)
        out["predict_s"] = t_pred
        out["integration"] = info
        out["wall_s"] += t_pred
        return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--family", choices=sorted(FAMILIES), default="shell")
    ap.add_argument("--n", type=int, default=0, help="lattice side of the training lattice (family default)")
    ap.add_argument("--samples", type=int, default=512)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--pool", type=int, default=8, help="concurrent solvers of the data generator")
    ap.add_argument("--ramp-steps", type=int, default=0, help="load ramp of the data generator (family default)")
    ap.add_argument("--data", default="", help="equilibria file: read if present, else written after solving")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    fam = FAMILIES[args.family]
    n = args.n or fam.coarse_n
    rng = np.random.default_rng(args.seed)
    loads = [fam.draw(rng) for _ in range(args.samples)]
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.data and Path(args.data).exists():
        z = np.load(args.data, allow_pickle=False)
        dm = json.loads(str(z["meta"]))
        if dm["family"] != fam.name or dm["n"] != n or dm["loads"] != loads:
            raise ValueError(f"{args.data} holds other equilibria than --family/--n/--samples/--seed ask for")
        act = fam.geometry(n).active.astype(bool)
        U = np.zeros((len(loads), 3) + act.shape, np.float32)
        F = np.zeros((len(loads), 9) + act.shape, np.float32)
        U[:, :, act], F[:, :, act] = z["U_active"], z["F_active"]
        info, gen_s = dm["info"], float(dm["datagen_s"])
        log(f"[train] read {len(loads)} equilibria from {args.data}")
    else:
        t0 = time.perf_counter()
        U, F, info = datagen.solve_equilibria(fam, n, loads, pool=args.pool, ramp_steps=args.ramp_steps or None,
                                              log=log)
        gen_s = time.perf_counter() - t0
        if args.data:
            Path(args.data).parent.mkdir(parents=True, exist_ok=True)
            act = fam.geometry(n).active.astype(bool)
            np.savez(args.data, U_active=U[:, :, act], F_active=F[:, :, act], meta=np.array(json.dumps(
                {"family": fam.name, "n": n, "loads": loads, "info": info, "datagen_s": gen_s})))
    ok = [k for k, i in enumerate(info) if i["converged"]]
    log(f"[train] {len(ok)}/{len(loads)} equilibria converged in {gen_s:.1f} s")
    keep = np.array(ok)
    op, rec = train(fam, n, [loads[k] for k in keep], U[keep], F[keep], epochs=args.epochs, batch=args.batch,
                    lr=args.lr, seed=args.seed, log=log)
    meta = {"family": fam.name, "n": n, "layers": op.layers, "samples": len(loads), "converged": len(ok),
            "datagen_s": gen_s, "datagen_steps": [i["steps"] for i in info], "loads": loads, "seed": args.seed,
            "epochs": args.epochs, **rec}
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    np.savez(args.out, params=op.params.cpu().numpy(), meta=np.array(json.dumps(meta)))
    print(json.dumps({k: meta[k] for k in ("family", "n", "samples", "converged", "datagen_s", "train_s",
                                             "val_errors")}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
