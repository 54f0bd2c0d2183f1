"""Launch-shape autotuner for the device MG-PCG (the role of the reference's Tuning helper,
lafem/tuning.hpp:72-163, which times Arch kernels for block sizes and stores them in globals).

For one hierarchy and inner precision it times candidate configurations on the GPU and returns
the fastest:
  * the level-matrix format: banded (where the matrix has few diagonals), CSR with 4/8/16 lanes
    per row, or SELL-C-sigma with chunk height C in {32, 64, 128, 256} and sort window sigma in
    {1, 64};
  * rows per workgroup of the row kernels (0 = derived from the row count, or 128..2048);
  * rows per workgroup of the transfer kernels (restriction, prolongation);
  * for the batched solver, the locality tile of the row order.
The objective is the measured time of one full V-cycle (feat3_*_time_kernel(3)), which covers
every level, the transfers and the coarse solve, i.e. what a PCG iteration spends in the
preconditioner.  Stage 1 searches format and chunk shape at the default grid, stage 2 the rows
per workgroup for the winner, stage 3 the rows per workgroup of its transfer kernels.  Results can be cached in a JSON file keyed by the problem shape.
"""
from __future__ import annotations

import json
import os
from typing import Dict, List, Optional

from .problems import Hierarchy

FORMATS: List[dict] = (
    [dict(fmt="banded"), dict(fmt="csr")]
    + [dict(fmt="csr", lanes=L) for L in (4, 8, 16)]
    + [dict(fmt="sell", sell_C=C, sell_sigma=s) for C in (32, 64, 128, 256) for s in (1, 64)]
)
ROWS_PER_BLOCK = (0, 128, 256, 512, 1024, 2048)
TRANSFER_ROWS_PER_BLOCK = (0, 256, 512)   # the solver's default, 128, is the reference point
_OPTION_KEYS = ("rows_per_block", "transfer_rows_per_block")
TILE_ROWS = (0, 512, 4096)


def _key(H: Hierarchy, inner: str, extra: str = "") -> str:
    L0 = H.levels[0].A
    return f"d{H.dim}q{H.degree}l{H.level_fine}-{H.level_coarse}_n{L0.rows}_nnz{L0.nnz}_{inner}{extra}"


def _load(cache: Optional[str]) -> Dict[str, dict]:
    if cache and os.path.exists(cache):
        with open(cache, "r", encoding="utf-8") as f:
            return json.load(f)
    return {}


def _store(cache: Optional[str], table: Dict[str, dict]) -> None:
    if cache:
        with open(cache, "w", encoding="utf-8") as f:
            json.dump(table, f, indent=1, sort_keys=True)


def time_vcycle(H: Hierarchy, inner: str, reps: int = 10, **cfg) -> float:
    """Average milliseconds of one V-cycle of a solver built with the given configuration."""
    from .solver import MGPCG
    kw = solver_kwargs({"config": cfg})
    s = MGPCG(H, inner=inner, **kw).init()
    try:
        s.time_kernel(3, 2)
        ms, _ = s.time_kernel(3, reps)
    finally:
        s.done()
    return ms


def tune_mgpcg(H: Hierarchy, inner: str = "f32", reps: int = 10, cache: Optional[str] = None,
               formats: Optional[List[dict]] = None, rows_per_block=ROWS_PER_BLOCK,
               transfer_rows_per_block=TRANSFER_ROWS_PER_BLOCK, log=None) -> dict:
    """Returns {"config": best keyword arguments for solver.MGPCG, "ms": V-cycle time,
    "trials": [(config, ms), ...]}."""
    key = _key(H, inner)
    table = _load(cache)
    if key in table:
        return table[key]
    trials = []
    best, best_ms = None, float("inf")
    for cfg in formats or FORMATS:
        ms = time_vcycle(H, inner, reps, **cfg)
        trials.append((cfg, ms))
        if log:
            log(f"tune {cfg}: {ms:.3f} ms")
        if ms < best_ms:
            best, best_ms = dict(cfg), ms
    for rpb in rows_per_block:
        if rpb == 0:
            continue
        cfg = dict(best, rows_per_block=rpb)
        ms = time_vcycle(H, inner, reps, **cfg)
        trials.append((cfg, ms))
        if log:
            log(f"tune {cfg}: {ms:.3f} ms")
        if ms < best_ms:
            best, best_ms = cfg, ms
    for t in transfer_rows_per_block:
        cfg = dict(best, transfer_rows_per_block=t)
        ms = time_vcycle(H, inner, reps, **cfg)
        trials.append((cfg, ms))
        if log:
            log(f"tune {cfg}: {ms:.3f} ms")
        if ms < best_ms:
            best, best_ms = cfg, ms
    out = {"config": best, "ms": best_ms, "trials": trials}
    table[key] = out
    _store(cache, table)
    return out


def solver_kwargs(result: dict) -> dict:
    """Turns a tuning result into keyword arguments of solver.MGPCG."""
    cfg = dict(result["config"])
    opts = dict(cfg.pop("options", {}) or {})
    for k in _OPTION_KEYS:
        if k in cfg:
            v = cfg.pop(k)
            if v or k == "transfer_rows_per_block":   # 0 there selects the derived grid
                opts[k] = v
    if opts:
        cfg["options"] = opts
    return cfg


def tune_batch(H: Hierarchy, nsys: int, values=None, block: int = 1, shared: bool = True, inner: str = "f32",
               reps: int = 5, tiles=TILE_ROWS, log=None) -> dict:
    """The batched solver's locality tile (rows per tile of the row order), by V-cycle time."""
    from .batch import BatchMGPCG
    trials = []
    best, best_ms = None, float("inf")
    for t in tiles:
        s = BatchMGPCG(H, nsys, values=values, block=block, shared=shared, inner=inner,
                       options={"tile_rows": t}).init()
        try:
            s.time_kernel(3, 1)
            ms, _ = s.time_kernel(3, reps)
        finally:
            s.done()
        trials.append(({"tile_rows": t}, ms))
        if log:
            log(f"tune batch tile_rows={t}: {ms:.3f} ms")
        if ms < best_ms:
            best, best_ms = {"tile_rows": t}, ms
    return {"config": best, "ms": best_ms, "trials": trials}
