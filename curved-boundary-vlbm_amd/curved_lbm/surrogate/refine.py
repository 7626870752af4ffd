"""Second stage of the learned warm start: a network on the solver's own
lattice that corrects the carried coarse prediction.

    python -m curved_lbm.surrogate.refine --base shell64.npz --samples 48 --out shell128_refine.npz

The coarse stage (``train.WarmStartModel``) predicts u and F on the family's
coarse lattice.  Carried to the solver lattice its prediction keeps the
coarse lattice's discretisation gap: for the 128^3 shell the 64^3 solver's
own equilibrium, carried up, is 0.5% off in u and 1% in F - I, mostly a
smooth radial profile, and a warm run from it takes 3200 steps against 1800
from the exact 128^3 (u, F).  The refinement network is the same operator
(``network.ConvOperator``: six 3x3x3 layers of 32 channels, BF16 on MFMA,
FP32 accumulation) on the solver lattice.

* Inputs: the family's features on that lattice followed by the carried,
  symmetric coarse prediction, u_c / s and (F_c - I) / s with
  s = family.scale(p).
* Output: the correction (u - u_c) / (REL s) and (F - F_c) / (REL s); the
  correction is about a percent of the field, REL scales it to order one.
* Training data: equilibria the solver itself reaches on that lattice from
  the coarse warm start (``solve_fine``), FP32 populations.
* Use: the correction is averaged over the family's symmetry group (the
  carried coarse fields already are symmetric), added to the carried fields
  and installed as the coarse stage's prediction is (``warm_start.run_warm``
  with its fallback).

A refinement stage can also stack on an earlier one on the same lattice
(``--base`` a refinement file): it then corrects that stage's prediction.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np
import torch

from .. import warm_start as ws
from ..three_d import curved_boundary_warp3d as cb
from .network import CH, TARGET_CH, ConvOperator
from .train import WarmStartModel, _pad, fit

REL = 0.02  # output scale of the correction, relative to family.scale(p)
LOAD_CHANNEL = {"shell": 8}  # the feature channel that carries the load


class _Inputs:
    """Device-side inputs of the refinement network on one lattice: the
    family's load-independent feature channels cached once; the load channel
    and the carried coarse prediction filled per call."""

    def __init__(self, family, n, device):
        if family.name not in LOAD_CHANNEL:
            raise ValueError(f"no refinement stage for the {family.name} family")
        self.family, self.n, self.device = family, int(n), device
        p0 = family.draw(np.random.default_rng(0))
        f = family.features(self.n, p0)
        x = np.zeros(f.shape[:3] + (CH,), np.float32)
        x[..., :f.shape[-1]] = f
        x[..., LOAD_CHANNEL[family.name]] = 0.0
        self.static = torch.from_numpy(_pad(x)[None]).to(device=device, dtype=torch.bfloat16)
        self.n_in = family.n_in + 12

    def load_feature(self, p) -> float:
        """The load channel's value (as the family's features() sets it)."""
        return float(p["traction"]) / 0.05

    def build(self, p, uc: torch.Tensor, Fc: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """(1, X+2, Y+2, Z+2, 32) BF16 input for load p and carried coarse
        fields uc (3, X, Y, Z), Fc (9, X, Y, Z)."""
        s = self.family.scale(p)
        x = self.static.clone() if out is None else out.copy_(self.static)
        core = x[0, 1:-1, 1:-1, 1:-1]
        core[..., LOAD_CHANNEL[self.family.name]] = self.load_feature(p)
        k = self.family.n_in
        eye = torch.eye(3, dtype=Fc.dtype, device=Fc.device).reshape(9, 1, 1, 1)
        core[..., k:k + 3] = (uc / s).permute(1, 2, 3, 0).to(torch.bfloat16)
        core[..., k + 3:k + 12] = ((Fc - eye) / s).permute(1, 2, 3, 0).to(torch.bfloat16)
        return x


def targets(family, n, p, uc, Fc, u, F) -> torch.Tensor:
    """(1, X+2, Y+2, Z+2, 16) FP32: the correction (u - u_c, F - F_c) / (REL s)
    in channels 0-11, the active mask in channel 15."""
    s = REL * family.scale(p)
    act = torch.from_numpy(family.geometry(n).active.astype(np.float32)).to(u.device)
    t = torch.zeros(family.lattice(n) + (TARGET_CH,), dtype=torch.float32, device=u.device)
    t[..., 0:3] = ((u.double() - uc.double()) / s).permute(1, 2, 3, 0).float()
    t[..., 3:12] = ((F.double() - Fc.double()) / s).permute(1, 2, 3, 0).float()
    t[..., 15] = act
    return torch.nn.functional.pad(t, (0, 0, 1, 1, 1, 1, 1, 1))[None]


def load_stage(path, device="cuda:0"):
    """The warm-start model in a weights file: a refinement stage (which
    loads the stages below it from the same directory) or a coarse one."""
    z = np.load(path, allow_pickle=False)
    if json.loads(str(z["meta"])).get("stage") == "refine":
        return RefinedWarmStartModel(path, device=device)
    return WarmStartModel(path, device=device)


def base_fields(base, p, shape):
    """(u, F, info) of a stage's prediction on the lattice ``shape``: a
    refinement stage's own, or a coarse stage's carried there."""
    u, F, info = base.prepare(p)
    if isinstance(base, RefinedWarmStartModel):
        base._check_shape(shape)
        return u, F, info
    uc, Fc = ws.carry_to_lattice(u, F, base._mask, shape)
    return uc, Fc, info


def solve_fine(base, n, loads, *, tol=1e-9, precision="fp32", device="cuda:0", log=None):
    """For each load: the base stage's prediction (uc, Fc) on lattice side n
    and the equilibrium (u, F) that the solver reaches there from it
    (float32 device tensors, (N, 3|9, X, Y, Z)), and each run's record."""
    fam = base.family
    g = fam.geometry(n)
    shape = fam.lattice(n)
    N = len(loads)
    UC = torch.zeros((N, 3) + shape, dtype=torch.float32, device=device)
    FC = torch.zeros((N, 9) + shape, dtype=torch.float32, device=device)
    U = torch.zeros_like(UC)
    F = torch.zeros_like(FC)
    info = []
    s = cb.WarpCurvedBoundaryLBM3D(g, **fam.solver_kwargs(n, loads[0]), precision=precision, device=device)
    ramp = fam.ramp_steps(n)
    every = max(20, 200 * n // fam.fine_n)
    max_steps = fam.max_steps_fine * n // fam.fine_n
    t0 = time.perf_counter()
    for k, p in enumerate(loads):
        s.set_boundary_params(fam.load_id(), fam.params_row(p, ramp * s.dt))
        uc, Fcf, _ = base_fields(base, p, shape)
        out = ws.run_warm(s, uc, Fcf, t_start=ramp * s.dt, tol=tol, check_every=every, max_steps=max_steps,
                          max_initial_residual=fam.max_initial_residual)
        uu, FF, _ = ws.dense_state(s)
        UC[k], FC[k], U[k], F[k] = uc.float(), Fcf.float(), uu.float(), FF.float()
        info.append({"steps": int(out["steps"]), "converged": bool(out["converged"]), "warm": bool(out["warm"])})
        if log is not None and (k % 8 == 0 or k == N - 1):
            log(f"[refine] {fam.name} n={n}: {k + 1}/{N} equilibria, {time.perf_counter() - t0:.1f} s "
                f"(last {out['steps']} steps)")
    del s
    return UC, FC, U, F, info


class RefinedWarmStartModel:
    """The staged warm start: the base stage's symmetric prediction on the
    refinement lattice (the coarse WarmStartModel's carried there, or an
    earlier refinement's), plus this stage's symmetric correction.
    ``prepare``, ``warm_up`` and ``start`` mirror WarmStartModel's (the
    fields are on the refinement lattice, the lattice ``start``'s solver
    must have)."""

    def __init__(self, path, device="cuda:0", base=None):
        path = Path(path)
        z = np.load(path, allow_pickle=False)
        meta = json.loads(str(z["meta"]))
        self.meta = meta
        # the base stage: from the file named in meta, beside this one, unless given
        self.base = base if base is not None else load_stage(path.parent / meta["base"], device=device)
        self.coarse = self.base.coarse if isinstance(self.base, RefinedWarmStartModel) else self.base
        self.family = self.base.family
        self.n = int(meta["n"])
        self.device = device
        self.inputs = _Inputs(self.family, self.n, device)
        self.op = ConvOperator(self.family.lattice(self.n), batch=1, layers=int(meta["layers"]),
                               n_in=self.inputs.n_in, n_out=12, device=device)
        self.op.load_state(z["params"])
        self._mask = torch.from_numpy(self.family.geometry(self.n).active.astype(bool)).to(device)

    def prepare(self, p, *, symmetric=True):
        """(u, F, info) on the refinement lattice."""
        uc, Fc, info = base_fields(self.base, p, self.family.lattice(self.n))
        self.inputs.build(p, uc, Fc, out=self.op.acts[0])
        y = self.op.forward()
        s = REL * self.family.scale(p)
        core = y[0, 1:-1, 1:-1, 1:-1].permute(3, 0, 1, 2)
        du, dF = s * core[0:3], s * core[3:12]
        if symmetric:
            du, dF = ws.symmetrize(du, dF, self._mask, self.family.symmetry_group())
        return uc + du.double(), Fc + dF.double(), info

    def warm_up(self, fine_shape) -> float:
        """One untimed pass of the prediction path (see
        WarmStartModel.warm_up); returns the seconds taken."""
        self._check_shape(fine_shape)
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        self.prepare(self.coarse.meta["loads"][0])
        torch.cuda.synchronize(self.device)
        return time.perf_counter() - t0

    def _check_shape(self, shape):
        if tuple(shape) != tuple(self.family.lattice(self.n)):
            raise ValueError(f"the refinement stage works on {self.family.lattice(self.n)}, not {tuple(shape)}")

    def start(self, solver, p, *, tol=1e-9, check_every=200, max_steps=10**6, **kw):
        """Warm-started run of a solver on the refinement lattice; the run
        record has the preparation time included in wall_s."""
        self._check_shape((solver.nx, solver.ny, solver.nz))
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        u, F, info = self.prepare(p)
        torch.cuda.synchronize(self.device)
        t_pred = time.perf_counter() - t0
        ramp_time = self.family.ramp_steps(self.n) * solver.dt
        kw.setdefault("max_initial_residual", self.family.max_initial_residual)
        out = ws.run_warm(solver, u, F, t_start=ramp_time, tol=tol, check_every=check_every, max_steps=max_steps,
                          **kw)
        out["predict_s"] = t_pred
        out["integration"] = info
        out["wall_s"] += t_pred
        return out


def evaluate(model: RefinedWarmStartModel, UC, FC, U, F, idx):
    """Relative l2 errors over the active cells of the refined prediction and
    of the base stage's against the solver equilibria idx."""
    act = model._mask
    eye = torch.eye(3, dtype=torch.float64, device=act.device).reshape(9, 1, 1, 1)
    out = {"u": [], "F_minus_I": [], "base_u": [], "base_F_minus_I": []}

    def rel(a, b):
        return float(torch.linalg.norm((a - b)[:, act]) / torch.linalg.norm(b[:, act]))

    for k in idx:
        p = model.meta["loads"][int(k)]
        u, Fp, _ = model.prepare(p)
        ut, Ft = U[int(k)].double(), F[int(k)].double()
        out["u"].append(rel(u, ut))
        out["F_minus_I"].append(rel(Fp - eye, Ft - eye))
        out["base_u"].append(rel(UC[int(k)].double(), ut))
        out["base_F_minus_I"].append(rel(FC[int(k)].double() - eye, Ft - eye))
    return {k: {"mean": float(np.mean(v)), "max": float(np.max(v))} for k, v in out.items()}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--base", required=True, help="the stage to refine: coarse weights or a refinement file "
                    "(the stages below it are loaded from its directory)")
    ap.add_argument("--n", type=int, default=0, help="refinement lattice side (the family's full lattice)")
    ap.add_argument("--samples", type=int, default=48)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--epochs", type=int, default=160)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--val", type=int, default=4, help="samples held out of the training for the errors")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    base = load_stage(args.base)
    fam = base.family
    n = args.n or fam.fine_n
    rng = np.random.default_rng(args.seed)
    loads = [fam.draw(rng) for _ in range(args.samples)]
    t0 = time.perf_counter()
    UC, FC, U, F, info = solve_fine(base, n, loads, log=log)
    gen_s = time.perf_counter() - t0
    ok = [k for k, i in enumerate(info) if i["converged"] and i["warm"]]
    log(f"[refine] {len(ok)}/{len(loads)} equilibria reached from the base prediction in {gen_s:.1f} s")
    inp = _Inputs(fam, n, base.device)
    X = torch.cat([inp.build(loads[k], UC[k].double(), FC[k].double()) for k in range(len(loads))])
    T = torch.cat([targets(fam, n, loads[k], UC[k], FC[k], U[k], F[k]) for k in range(len(loads))])
    order = np.random.default_rng(args.seed).permutation(np.array(ok))
    val, trn = order[:args.val], order[args.val:]
    op, rec = fit(X, T, trn, fam.lattice(n), n_in=inp.n_in, active=float(fam.geometry(n).active.sum()),
                  name=f"{fam.name}-refine", epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed,
                  device=base.device, log=log)
    del X, T
    meta = {"family": fam.name, "stage": "refine", "base": Path(args.base).name, "n": n, "layers": op.layers,
            "n_in": inp.n_in, "rel": REL, "samples": len(loads), "used": len(ok), "datagen_s": gen_s,
            "datagen_steps": [i["steps"] for i in info], "loads": loads, "seed": args.seed, "epochs": args.epochs,
            "val_index": val.tolist(), **rec}
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    np.savez(out, params=op.params.cpu().numpy(), meta=np.array(json.dumps(meta)))
    model = RefinedWarmStartModel(out, device=base.device, base=base)
    meta["val_errors"] = evaluate(model, UC, FC, U, F, val)
    np.savez(out, params=op.params.cpu().numpy(), meta=np.array(json.dumps(meta)))
    print(json.dumps({k: meta[k] for k in ("family", "n", "samples", "used", "datagen_s", "train_s",
                                             "val_errors")}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
