"""Autotuner for batched scans (no reference counterpart).

A batch's speed depends on choices that never change its results: the tile of each msc
lattice size (queue compaction over 512 words, or lockstep over 256 words; small lattices
have too few words to fill waves after compaction), the sites per lane of each int8 lattice
size (16, 8 or 4: wide vectors amortise per-lane work, narrow ones fill the lanes of a small
lattice's tile), the launch form (one launch per half sweep, or dependency-ordered launches of many sweeps)
together with the tiles each workgroup runs, and the number of sweeps per captured graph.
:func:`tune_batch` times candidates on scratch lattices of the same shapes, temperatures and
algorithms (hot starts from other seeds, so the callers' lattices are untouched), by
coordinate descent: the msc tile per shape, the int8 width per shape, then the launch form and
tiles per workgroup, then the graph length.
Results are cached per batch signature in a JSON file outside the repository
(``$PYISING_AMD_TUNE_CACHE`` or ``~/.cache/pyising_amd/batch_tune.json``).
"""

from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass, field
from pathlib import Path

GRAPH_CHOICES = (0, 10, 50)
TILES_PER_WG_CHOICES = (1, 2, 4, 8)  # tiles (lattices or parts of one) per workgroup, either form


@dataclass
class BatchTuning:
    msc_tiles: dict = field(default_factory=dict)  # shape tuple -> "queue" | "lockstep"
    int8_vecs: dict = field(default_factory=dict)  # shape tuple -> 16 | 8 | 4 (absent: default)
    persistent: bool = False
    graph_sweeps: int = 50
    tiles_per_workgroup: int = 1
    flips_per_ns: float = 0.0
    trials: list = field(default_factory=list)

    def to_json(self) -> dict:
        d = asdict(self)
        d["msc_tiles"] = {"x".join(map(str, k)): v for k, v in self.msc_tiles.items()}
        d["int8_vecs"] = {"x".join(map(str, k)): v for k, v in self.int8_vecs.items()}
        return d

    @classmethod
    def from_json(cls, d: dict) -> "BatchTuning":
        tiles = {tuple(int(x) for x in k.split("x")): v for k, v in d.get("msc_tiles", {}).items()}
        vecs = {tuple(int(x) for x in k.split("x")): int(v)
                for k, v in d.get("int8_vecs", {}).items()}
        return cls(msc_tiles=tiles, int8_vecs=vecs, persistent=bool(d.get("persistent", False)),
                   graph_sweeps=int(d.get("graph_sweeps", 50)),
                   tiles_per_workgroup=int(d.get("tiles_per_workgroup", 1)),
                   flips_per_ns=float(d.get("flips_per_ns", 0.0)), trials=d.get("trials", []))


def signature(sims) -> str:
    """Order-independent key: the multiset of (variant, shape, algorithm)."""
    items = sorted(f"{s.variant}:{'x'.join(map(str, s.shape))}:{s.algorithm}" for s in sims)
    out, prev, count = [], None, 0
    for it in items + [None]:
        if it == prev:
            count += 1
            continue
        if prev is not None:
            out.append(f"{prev}*{count}")
        prev, count = it, 1
    return ";".join(out)


def _cache_path() -> Path:
    env = os.environ.get("PYISING_AMD_TUNE_CACHE")
    return Path(env) if env else Path.home() / ".cache" / "pyising_amd" / "batch_tune.json"


def _load_cache() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _save_cache(cache: dict) -> None:
    path = _cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: tuning still works, it is just not remembered


def _scratch(sims):
    out = []
    for i, s in enumerate(sims):
        out.append(type(s)(s.shape, s.beta, seed=(s.seed ^ 0x9E3779B97F4A7C15) + i,
                           algorithm=s.algorithm, init="hot"))
    return out


def _rate(scratch, t: BatchTuning, sweeps: int) -> float:
    from . import _lib
    from .batch import BatchScan

    graph_sweeps = t.graph_sweeps
    scan = BatchScan(scratch, msc_tiles=t.msc_tiles, int8_vecs=t.int8_vecs,
                     persistent=t.persistent, tiles_per_workgroup=t.tiles_per_workgroup)
    try:
        n_spins = int(scan.n_spins.sum())
        e0, e1 = _lib.Event(), _lib.Event()
        graph = bool(graph_sweeps) and scan.capture(graph_sweeps)
        reps = max(1, sweeps // graph_sweeps) if graph else 1
        done = reps * graph_sweeps if graph else sweeps

        def run() -> None:
            if graph:
                scan.replay(reps)
            else:
                scan.sweep(sweeps)

        run()  # warm-up: the same work as a timed run
        scan.synchronize()
        best = 0.0
        for _ in range(2):  # the better of two timed runs
            e0.record(scan.stream.ptr)
            run()
            e1.record(scan.stream.ptr)
            best = max(best, done * n_spins / (e0.elapsed_ms(e1) * 1e-3) / 1e9)
        scan.synchronize()
        for ev in (e0, e1):
            ev.close()
        return best
    finally:
        scan.close()


def _settle(scratch, min_s: float = 0.3) -> None:
    """Sweep the batch in its default form for at least ``min_s`` seconds of wall time."""
    import time

    from .batch import BatchScan

    scan = BatchScan(scratch)
    try:
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < min_s:
            scan.sweep(10)
            scan.synchronize()
    finally:
        scan.close()


def _vec_choices(shape) -> list[int]:
    lh = shape[-1] // 2
    return [v for v in (16, 8, 4) if lh % v == 0]


def tune_batch(sims, sweeps: int = 50, use_cache: bool = True) -> BatchTuning:
    """Pick the msc tile and int8 width per shape, the launch form, the tiles each workgroup
    runs in sequence (launch form) and the graph length."""
    key = signature(sims)
    cache = _load_cache() if use_cache else {}
    if key in cache:
        return BatchTuning.from_json(cache[key])
    scratch = _scratch(sims)
    msc_shapes = sorted({tuple(s.shape) for s in sims if s.variant == "msc"})
    int8_shapes = sorted({tuple(s.shape) for s in sims if s.variant == "int8"})
    best = BatchTuning(msc_tiles={sh: "queue" for sh in msc_shapes})
    trials = []

    def copy(**kw) -> BatchTuning:
        base = {"msc_tiles": dict(best.msc_tiles), "int8_vecs": dict(best.int8_vecs),
                "persistent": best.persistent, "graph_sweeps": best.graph_sweeps,
                "tiles_per_workgroup": best.tiles_per_workgroup}
        base.update(kw)
        return BatchTuning(**base)

    def measure(t: BatchTuning) -> float:
        r = _rate(scratch, t, sweeps)
        trials.append({"tiles": {"x".join(map(str, k)): v for k, v in t.msc_tiles.items()},
                       "int8_vecs": {"x".join(map(str, k)): v for k, v in t.int8_vecs.items()},
                       "persistent": t.persistent, "graph_sweeps": t.graph_sweeps,
                       "tiles_per_workgroup": t.tiles_per_workgroup,
                       "flips_per_ns": round(r, 1)})
        return r

    def consider(cand: BatchTuning) -> None:
        nonlocal best
        r = measure(cand)
        if r > best.flips_per_ns:
            best, best.flips_per_ns = cand, r

    # the GPU's clocks need a few hundred milliseconds of load after idling: trials timed
    # before that would favour whichever ran last
    _settle(scratch)
    best.flips_per_ns = measure(best)
    for sh in msc_shapes:
        tiles = dict(best.msc_tiles)
        tiles[sh] = "lockstep"
        consider(copy(msc_tiles=tiles))
    for sh in int8_shapes:
        for v in _vec_choices(sh):
            vecs = dict(best.int8_vecs)
            vecs[sh] = v
            consider(copy(int8_vecs=vecs))
    start = (best.persistent, best.tiles_per_workgroup)
    for persistent in (False, True):
        for k in TILES_PER_WG_CHOICES:
            if (persistent, k) != start:
                consider(copy(persistent=persistent, tiles_per_workgroup=k))
    # the msc tile choice again, under the chosen form (a tile's cost depends on the form)
    for sh in msc_shapes:
        tiles = dict(best.msc_tiles)
        tiles[sh] = "queue" if tiles[sh] == "lockstep" else "lockstep"
        consider(copy(msc_tiles=tiles))
    for g in GRAPH_CHOICES:
        if g != best.graph_sweeps:
            consider(copy(graph_sweeps=g))
    best.trials = trials
    if use_cache:
        cache[key] = best.to_json()
        _save_cache(cache)
    del scratch
    return best
