"""Learned warm start for SCFT: a network maps (chiN, f, space group, cell size)
to the first star coefficients of w_A and w_B; the fields decoded from them
start the solver, and plain field iteration from the seeded cold start is the
fallback when the guess does not converge.  No counterpart in the reference.

Everything that touches the fields runs on the GPU through the library: the
solves (``SCFTSolver``), the star coefficients (KRFFT forward / backward on the
subgrid) and the network (``MLP``: BF16 matrix-core GEMMs, FP32 accumulation).

Run as a module to generate the training set, train and evaluate, e.g.
    python -m cfft_hip.warmstart --n-train 2000 --n-eval 200 --out model.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor, as_completed
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import stars
from .mlp import MLP
from .plans import DevicePlan, auto_L, find_optimal_shift
from .scft import SCFTSolver, double_gyroid_211, first_star_bcc, symmetrize_subgrid
from .symmetry_ops import get_ops

GROUPS = (229, 230, 70)  # Im-3m, Ia-3d, Fddd
N_STARS = 64
NS = 100


@dataclass
class Params:
    chiN: float
    f: float
    sg: int
    cell: float
    seed: int


def draw(n, seed, chiN=(12.0, 40.0), f=(0.25, 0.45), cell=(3.6, 4.4)):
    rng = np.random.default_rng(seed)
    return [Params(float(rng.uniform(*chiN)), float(rng.uniform(*f)), int(rng.choice(GROUPS)),
                   float(rng.uniform(*cell)), int(rng.integers(1 << 30))) for _ in range(n)]


def features(ps):
    out = np.zeros((len(ps), 6), dtype=np.float32)
    for i, p in enumerate(ps):
        out[i, 0] = (p.chiN - 26.0) / 14.0
        out[i, 1] = (p.f - 0.35) / 0.1
        out[i, 2 + GROUPS.index(p.sg)] = 1.0
        out[i, 5] = (p.cell - 4.0) / 0.4
    return out


class Group:
    """Per (group, grid) setup shared by every instance: shifted ops, subgrid,
    a plan for the star coefficients and their order."""

    def __init__(self, sg, N):
        self.sg, self.N = sg, tuple(N)
        ops = get_ops(sg, 3, self.N)
        self.shift, self.ops_s = find_optimal_shift(ops, self.N)
        self.Lv = auto_L(self.ops_s)
        self.plan = DevicePlan(self.N, self.ops_s, 0.0, 4.0 * np.eye(3), L.PATH_AUTO)
        self.order, self.k2 = stars.star_order(self.plan)
        self.lock = threading.Lock()  # the plan's buffers serve one call at a time

    def cold_fields(self, seed, amp=3.0, noise=0.05):
        """Seeded cold start: the group's lowest harmonic (BCC {110} star for
        Im-3m, the double-gyroid {211} star for Ia-3d with w_A low on the
        networks, the {111} star for Fddd) +- amp, plus noise."""
        b = [s * n for s, n in zip(self.shift, self.N)]
        if self.sg == 229:
            base = first_star_bcc(self.N, b, 1.0)
        elif self.sg == 230:
            base = -double_gyroid_211(self.N, b, 1.0)
        else:
            x = [(np.arange(n) + s) / n for n, s in zip(self.N, b)]
            X, Y, Z = np.meshgrid(*x, indexing="ij")
            base = np.cos(2 * np.pi * (X + Y + Z))
        rng = np.random.default_rng(seed)
        wA = symmetrize_subgrid(amp * base + noise * rng.standard_normal(self.N), self.ops_s, self.Lv)
        wB = symmetrize_subgrid(-amp * base + noise * rng.standard_normal(self.N), self.ops_s, self.Lv)
        return wA, wB

    def encode(self, wA, wB, n=N_STARS):
        """Real vector (4 n): Re, Im of the first n star coefficients of wA, wB."""
        with self.lock:
            ca = stars.encode(self.plan, wA, n, self.order)
            cb = stars.encode(self.plan, wB, n, self.order)
        # grids with fewer than n stars: the missing coefficients are zero
        ca = np.pad(ca, (0, n - ca.size))
        cb = np.pad(cb, (0, n - cb.size))
        return np.concatenate([ca.real, ca.imag, cb.real, cb.imag]).astype(np.float32)

    def decode(self, v, n=N_STARS):
        v = np.asarray(v, dtype=np.float64)
        m = min(n, self.plan.n_spec)
        ca = v[0:m] + 1j * v[n:n + m]
        cb = v[2 * n:2 * n + m] + 1j * v[3 * n:3 * n + m]
        with self.lock:
            return stars.decode(self.plan, ca, self.order), stars.decode(self.plan, cb, self.order)


_tls = threading.local()


def _run_all(fn, items, workers, stage, every=100):
    """fn over items on a thread pool, in order; prints a progress line to
    stderr each time another `every` items have finished."""
    out = [None] * len(items)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        futs = {ex.submit(fn, it): i for i, it in enumerate(items)}
        for k, fut in enumerate(as_completed(futs), 1):
            out[futs[fut]] = fut.result()
            if k % every == 0 or k == len(items):
                print(json.dumps({"stage": stage, "done": k, "of": len(items),
                                  "wall_s": round(time.perf_counter() - t0, 2)}), file=sys.stderr, flush=True)
    return out


def _solver(g: Group, p: Params, precision):
    """One solver (plan + context) per thread and group, re-pointed at each
    instance's cell, f and chiN (plan creation is not repeated per instance)."""
    cache = getattr(_tls, "solvers", None)
    if cache is None:
        cache = _tls.solvers = {}
    key = (g.sg, g.N, precision)
    s = cache.get(key)
    if s is None:
        s = cache[key] = SCFTSolver(g.N, g.ops_s, p.cell * np.eye(3), p.f, p.chiN, NS, precision, m_hist=10)
    else:
        s.update(lattice=p.cell * np.eye(3), f=p.f, chiN=p.chiN)
    return s


def solve(p: Params, g: Group, fields=None, tol=1e-8, max_iter=600, precision=L.FP64):
    """Solve one instance; fields=None: the seeded cold start.  Returns
    (iterations, wall seconds, result dict, final fields)."""
    s = _solver(g, p, precision)
    wA, wB = g.cold_fields(p.seed) if fields is None else fields
    t0 = time.perf_counter()
    s.set_fields(wA, wB)
    it, out = s.solve(tol=tol, max_iter=max_iter)
    el = time.perf_counter() - t0
    return it, el, out, s.fields()


def generate(ps, N, workers=8, tol=1e-8, max_iter=600):
    """Converged fields of every instance (cold start) -> (features, targets,
    converged mask, iterations, wall seconds)."""
    groups = {sg: Group(sg, N) for sg in sorted({p.sg for p in ps})}

    def one(p):
        g = groups[p.sg]
        it, el, out, (wA, wB) = solve(p, g, None, tol, max_iter)
        ok = out["res"] < tol and np.isfinite(out["F"])
        return g.encode(wA, wB), ok, it, el

    res = _run_all(one, ps, workers, "generate")
    T = np.stack([r[0] for r in res])
    ok = np.array([r[1] for r in res])
    its = np.array([r[2] for r in res])
    els = np.array([r[3] for r in res])
    return features(ps), T, ok, its, els, groups


class WarmStart:
    """Trained network + target normalisation."""

    def __init__(self, dims=(6, 512, 512, 512, 4 * N_STARS), seed=0):
        self.net = MLP(list(dims), seed=seed)
        self.mu = np.zeros(dims[-1], dtype=np.float32)
        self.sd = np.ones(dims[-1], dtype=np.float32)

    def fit(self, X, T, epochs=300, batch=64, lr=1e-3, seed=0, chunk=25, lr_min=0.01):
        """Adam on the normalised targets; the learning rate follows a cosine
        from lr down to lr * lr_min, constant over chunks of `chunk` epochs."""
        self.mu = T.mean(axis=0).astype(np.float32)
        self.sd = (T.std(axis=0) + 1e-6).astype(np.float32)
        Tn = (T - self.mu) / self.sd
        hist = []
        n_chunks = max(1, -(-epochs // chunk))
        for k in range(n_chunks):
            e = min(chunk, epochs - k * chunk)
            lr_k = lr * (lr_min + (1.0 - lr_min) * 0.5 * (1.0 + np.cos(np.pi * k / n_chunks)))
            hist.append(self.net.train(X, Tn, epochs=e, batch=batch, lr=float(lr_k), seed=seed + k))
        return np.concatenate(hist)

    def predict(self, X):
        return self.net.forward(X) * self.sd + self.mu

    def save(self, path, meta=None):
        np.savez(path, params=self.net.params(), dims=np.array(self.net.dims), mu=self.mu, sd=self.sd,
                 meta=json.dumps(meta or {}))

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        ws = cls(tuple(int(v) for v in d["dims"]))
        ws.net.set_params(d["params"])
        ws.mu, ws.sd = d["mu"], d["sd"]
        ws.meta = json.loads(str(d["meta"]))
        return ws


def evaluate(ws: WarmStart, ps, groups, tol=1e-8, max_iter=600, workers=8):
    """Cold start vs network start on every instance; a network start that does
    not reach `tol` is finished from the cold start (fallback), and its
    iterations and time are both counted.  The network start's time includes
    its share of the inference and the decoding of its star coefficients."""
    t0 = time.perf_counter()
    V = ws.predict(features(ps))
    t_pred = (time.perf_counter() - t0) / max(len(ps), 1)  # inference, per instance

    def one(i):
        p = ps[i]
        g = groups[p.sg]
        ic, tc, oc, (ca, cb) = solve(p, g, None, tol, max_iter)
        target = g.encode(ca, cb)
        coef_err = float(np.linalg.norm(V[i] - target) / max(np.linalg.norm(target), 1e-300))
        t1 = time.perf_counter()
        guess = g.decode(V[i])
        t_dec = time.perf_counter() - t1
        iw, tw, ow, _ = solve(p, g, guess, tol, max_iter // 2)
        tw += t_pred + t_dec
        fell_back = not (ow["res"] < tol)
        if fell_back:
            iw += ic
            tw += tc
        return dict(sg=p.sg, coef_err=coef_err, cold_iters=ic, cold_s=tc, warm_iters=iw, warm_s=tw,
                    fallback=fell_back,
                    dF=abs(ow["F"] - oc["F"]) if not fell_back else 0.0, cold_ok=oc["res"] < tol)

    return _run_all(one, list(range(len(ps))), workers, "evaluate", every=25)


def summary(rows):
    ci = np.array([r["cold_iters"] for r in rows], dtype=float)
    wi = np.array([r["warm_iters"] for r in rows], dtype=float)
    cs = np.array([r["cold_s"] for r in rows])
    wsec = np.array([r["warm_s"] for r in rows])
    return {
        "n": len(rows),
        "cold_iters_mean": float(ci.mean()), "warm_iters_mean": float(wi.mean()),
        "cold_s_total": float(cs.sum()), "warm_s_total": float(wsec.sum()),
        "time_ratio": float(wsec.sum() / cs.sum()), "iter_ratio": float(wi.sum() / ci.sum()),
        "fallbacks": int(sum(r["fallback"] for r in rows)),
        "max_abs_dF": float(max(r["dF"] for r in rows)),
        # warm starts that converged to another stationary point than the cold start
        "other_solution": int(sum(r["dF"] > 1e-6 for r in rows)),
        "cold_not_converged": int(sum(not r["cold_ok"] for r in rows)),
        "coef_relerr_median": float(np.median([r["coef_err"] for r in rows])),
        "per_group": {str(sg): {"n": len(g), "cold_iters_mean": float(np.mean([r["cold_iters"] for r in g])),
                                "warm_iters_mean": float(np.mean([r["warm_iters"] for r in g])),
                                "time_ratio": float(sum(r["warm_s"] for r in g) / sum(r["cold_s"] for r in g))}
                      for sg in sorted({r["sg"] for r in rows})
                      for g in [[r for r in rows if r["sg"] == sg]]},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=2000)
    ap.add_argument("--n-eval", type=int, default=200)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--out", default="warmstart.npz")
    ap.add_argument("--data", default=None,
                    help="training set (.npz): loaded when present, else generated and saved there")
    args = ap.parse_args()
    N = (args.grid,) * 3
    t0 = time.perf_counter()
    ps = draw(args.n_train, seed=1)
    if args.data and os.path.exists(args.data):
        d = np.load(args.data, allow_pickle=False)
        X, T, ok, its = d["X"], d["T"], d["ok"], d["iters"]
        groups = {sg: Group(sg, N) for sg in GROUPS}
    else:
        X, T, ok, its, els, groups = generate(ps, N, workers=args.workers)
        if args.data:
            np.savez(args.data, X=X, T=T, ok=ok, iters=its, wall_s=els)
    t1 = time.perf_counter()
    print(json.dumps({"stage": "generate", "n": len(ps), "converged": int(ok.sum()),
                      "iters_mean": float(its.mean()), "wall_s": t1 - t0}), flush=True)
    ws = WarmStart()
    hist = ws.fit(X[ok], T[ok], epochs=args.epochs, batch=args.batch, lr=args.lr)
    t2 = time.perf_counter()
    print(json.dumps({"stage": "train", "epochs": args.epochs, "loss_first": float(hist[0]),
                      "loss_last": float(hist[-1]), "wall_s": t2 - t1}), flush=True)
    pe = draw(args.n_eval, seed=2)
    rows = evaluate(ws, pe, groups, workers=args.workers)
    t3 = time.perf_counter()
    summ = summary(rows)
    summ["wall_s"] = t3 - t2
    print(json.dumps({"stage": "evaluate", **summ}), flush=True)
    ws.save(args.out, {"grid": args.grid, "n_train": int(ok.sum()), "groups": list(GROUPS),
                       "n_stars": N_STARS, "eval": summ})


if __name__ == "__main__":
    main()
