"""Offline autotuner of the GEMM plans (tile and split-K depth) of the sweeps.

For each workload, a (D, w, d) configuration at the lockstep-copy count its
pool runs with, one RK4 step is launched eagerly with GEMM recording on
(``mps_ctx_gemm_record``).  Every distinct contraction shape of that step is
then re-launched on its own operands with each candidate plan -- tile 32 or
64, split-K depths 1..32 that fit the call's partials workspace -- and timed
with HIP events (``mps_ctx_gemm_time``).  Where the fastest candidate beats
the plan the fixed rule picks by more than ``margin``, the shape gets a table
entry.  The table is written as ``csrc/gemm_tuned.inc`` and compiled into
the library, so a build's plans, and with them the summation order of every
contraction, are a fixed function of the shape: results do not depend on
timing noise at run time.

Run on the target GPU, then rebuild:

    python -m mpsamd.autotune --out matrixproductstates.jl_amd/csrc/gemm_tuned.inc
"""
from __future__ import annotations

import argparse
import math
import sys
import time

from . import core as M
from ._lib import MPSError
from . import models as MD

BK = 16
SPLITS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)


def _effective_splits(K, s):
    """The split count launch_gemm runs for a requested depth s: chunks of
    ceil(K / s) rounded up to the K step."""
    chunk = -(-K // s)
    chunk = max(BK, -(-chunk // BK) * BK)
    return max(1, -(-K // chunk))


# ---------------------------------------------------------------- workloads
# Each builder returns (context, one-step function, close function).

def rydberg_single(N=100, D=128, n_exp=4, prec="fp64", dt=0.005):
    """The density-matrix benchmark workload on a plain context."""
    ctx = M.Context(0)
    ctx.set_qr_rank_tol(1e-13)
    ctx.set_precision(prec)
    p = MD.RydbergParams(n_exp=n_exp)
    z = MD.positions(N, float(N), 1)
    L = [M.MPO(h, ctx) for h in MD.rydberg_rk4_mpos(p, z, dt)]
    rho, dims = M.mpsgroundstate(N, D, 9, ctx)
    rs = [rho.similar().copy_from(rho) for _ in range(3)]
    envs = [M.build_env(dims, dims, ctx=ctx) for _ in range(3)]
    envop = M.build_env(dims, dims, [1] + [p.w] * (N - 1) + [1], ctx=ctx)

    def step():
        M.RK4stp_apply_H_half_(dt, L, rho, *rs, *envs, envop)

    def close():
        for o in (rho, *rs, *envs, envop, *L):
            o.__del__()
        ctx.close()
    return ctx, step, close


def spin_pool(N=60, D=32, B=96, prec="fp64", dt=0.01):
    """A lockstep pool of B spin-chain trajectories."""
    from .trajectory import SpinEnsemble
    ens = SpinEnsemble(N, D, 1, list(range(B)), dt=dt, qr_rank_tol=1e-13)
    ens.ctx.set_precision(prec)
    return ens.ctx, ens.step, ens.close


def density_pool(N=40, D=32, B=64, dt=0.005):
    """A lockstep pool of B Rydberg density matrices (d = 9)."""
    from .serving import DensityMatrixPool
    pool = DensityMatrixPool(N, D, slots=B, dt=dt)

    def step():
        pool.launch()
        pool.ctx.synchronize()
    return pool.ctx, step, pool.close


def default_workloads():
    """The configurations the bench and the pools run: (label, builder, kwargs)."""
    out = [("rydberg N=100 D=128 w=14 d=9 nb=1 fp64", rydberg_single, dict(prec="fp64")),
           ("rydberg N=100 D=128 w=14 d=9 nb=1 fp32", rydberg_single, dict(prec="fp32"))]
    for D, B in ((16, 96), (32, 96), (64, 96), (128, 96), (256, 48)):
        out.append((f"spin D={D} w=4 d=2 nb={B} fp32", spin_pool, dict(D=D, B=B, prec="fp32")))
    for D, B in ((16, 64), (32, 64), (64, 64)):
        out.append((f"spin D={D} w=4 d=2 nb={B} fp64", spin_pool, dict(D=D, B=B, prec="fp64")))
    out.append(("spin D=32 w=4 d=2 nb=256 fp64", spin_pool, dict(D=32, B=256, prec="fp64")))
    out.append(("rydberg D=32 w=10 d=9 nb=64 fp64", density_pool, dict(N=40, D=32, B=64)))
    return out


# -------------------------------------------------------------------- tuning

def tune_workload(label, build, kw, margin=0.03, budget_us=3000.0, log=print):
    ctx, step, close = build(**kw)
    try:
        step()
        step()
        ctx.synchronize()
        ctx.set_graphs(False)
        M.gemm_record(ctx, True)
        step()
        ctx.synchronize()
        M.gemm_record(ctx, False)
        shapes = M.gemm_recorded(ctx)
        entries = []
        for r, s in enumerate(shapes):
            t_def = M.gemm_time(ctx, r, s["tm"], s["splits"], 3)
            iters = max(3, min(200, int(math.ceil(budget_us / max(t_def, 1.0)))))
            t_def = M.gemm_time(ctx, r, s["tm"], s["splits"], iters)
            best = (t_def, s["tm"], s["splits"])
            tried = set()
            for tm in (32, 64):
                for sp in SPLITS:
                    e = _effective_splits(s["K"], sp)
                    if (tm, e) in tried or (tm, e) == (s["tm"], s["splits"]):
                        continue
                    tried.add((tm, e))
                    try:
                        t = M.gemm_time(ctx, r, tm, e, iters)
                    except MPSError:
                        continue           # partials exceed the call's workspace
                    if t < best[0]:
                        best = (t, tm, e)
            gain = 1.0 - best[0] / t_def
            log(f"{label}: {s['M']}x{s['N']}x{s['K']} batch {s['batch']} nb {s['nb']}: rule tm {s['tm']} "
                f"splits {s['splits']} {t_def:.2f} us; best tm {best[1]} splits {best[2]} {best[0]:.2f} us "
                f"({100 * gain:.1f}%)")
            if gain > margin:
                entries.append(dict(s, tm=best[1], splits=best[2], t_rule=t_def, t_best=best[0],
                                    rule=(s["tm"], s["splits"]), label=label))
        return entries
    finally:
        close()


def write_table(entries, path):
    seen = {}
    for e in entries:
        k = (e["M"], e["N"], e["K"], e["batch"], e["nb_class"], e["cmp"])
        if k not in seen or e["t_rule"] - e["t_best"] > seen[k]["t_rule"] - seen[k]["t_best"]:
            seen[k] = e
    lines = ["// Autotuned GEMM plans: {M, N, K, batch, nb_class, cmp, tm, splits},",
             "// written by mpsamd/autotune.py from timings on an MI355X; each entry's",
             "// comment gives its workload, the fixed rule's plan and both timings."]
    for k, e in seen.items():
        lines.append(f"    {{{e['M']}, {e['N']}, {e['K']}, {e['batch']}, {e['nb_class']}, {e['cmp']}, {e['tm']}, "
                     f"{e['splits']}}},  // {e['label']}: rule {e['rule'][0]}/{e['rule'][1]} "
                     f"{e['t_rule']:.1f} us -> {e['t_best']:.1f} us")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return len(seen)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", required=True, help="path of the generated gemm_tuned.inc")
    ap.add_argument("--margin", type=float, default=0.03,
                    help="minimum fractional gain over the fixed rule for a table entry")
    ap.add_argument("--only", default=None, help="substring of the workload labels to tune")
    a = ap.parse_args(argv)
    entries = []
    t0 = time.perf_counter()
    for label, build, kw in default_workloads():
        if a.only and a.only not in label:
            continue
        print(f"== {label}", flush=True)
        entries += tune_workload(label, build, kw, margin=a.margin, log=lambda s: print(s, flush=True))
    n = write_table(entries, a.out)
    print(f"{n} table entries written to {a.out} in {time.perf_counter() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    sys.exit(main())
