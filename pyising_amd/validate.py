"""Physics validation suite on the GPU (src/pyising/validate.py:68-215, with this backend's own
samplers) -> ``results/validation.json``.

1. Identity across kernels: the default msc kernel, the lockstep kernel, the eager ablation and
   the batched scan leave the same lattice after the same sweeps; int8 alone and batched too.
2. 2D energy per spin vs the exact finite-torus value (Kaufman), |m| vs Onsager-Yang, for
   Metropolis int8, Metropolis msc and heat-bath int8.
3. 2D Binder-cumulant crossing of a batched temperature scan vs Onsager's T_c.
4. 3D Binder crossing vs the simple-cubic critical coupling (Ferrenberg, Xu & Landau 2018).
5. 3D Metropolis vs heat-bath energies at the same temperature (two different Markov chains
   with the same stationary distribution).

The reference equilibrates and measures the Binder curves with Wolff cluster updates on the
CPU; here every chain starts cold (T < T_c) or hot and is thermalised with local sweeps, and
every measurement comes from the GPU kernels (the Binder moments from the batched scan's
device sums). Bitwise identity with the reference's algorithm itself is checked by the test
suite against its restatement, not by this driver.
"""

from __future__ import annotations

import json
import time
from pathlib import Path

import numpy as np

from . import backends as B
from .exact import energy_torus, magnetization_2d
from .lattice import TC_2D, TC_3D
from .observables import crossing, sample, scan_binder

QUICK = {
    "L2": 64, "T2": [1.8, 2.0, TC_2D, 2.5, 3.0], "therm": 20000, "meas": 2000, "every": 20,
    "binder2_L": [16, 32], "binder2_T": list(np.linspace(2.22, 2.32, 6)),
    "binder3_L": [8, 16], "binder3_T": list(np.linspace(4.46, 4.56, 6)),
    "binder_therm": 20000, "binder_blocks": 32, "binder_block_sweeps": 2000,
    "L3": 16, "T3": [4.0, TC_3D, 5.0],
}
FULL = {
    "L2": 128, "T2": sorted(list(np.round(np.linspace(1.6, 3.4, 19), 3)) + [TC_2D]),
    "therm": 50000, "meas": 5000, "every": 40,
    "binder2_L": [16, 32, 64], "binder2_T": list(np.round(np.linspace(2.20, 2.34, 15), 4)),
    "binder3_L": [8, 16, 32], "binder3_T": list(np.round(np.linspace(4.47, 4.55, 9), 4)),
    "binder_therm": 100000, "binder_blocks": 40, "binder_block_sweeps": 10000,
    "L3": 32, "T3": [3.5, 4.2, TC_3D, 5.0, 6.0],
}
#: pass criteria: |z| of an energy, |m| error vs Onsager-Yang away from T_c, relative T_c error
E_Z_MAX, M_ABS_MAX, TC_REL_MAX = 5.0, 0.01, 0.02


def _n_meas(cfg: dict, T: float, tc: float) -> int:
    """Four times more measurements near T_c, where autocorrelations are longest
    (src/pyising/validate.py:53-55)."""
    return cfg["meas"] * (4 if abs(T - tc) < 0.15 * tc / 2.27 else 1)


def _init(T: float, tc: float) -> str:
    return "cold" if T < tc else "hot"


def check_identity(log=print, sweeps: int = 7) -> list[dict]:
    """Every kernel path of a lattice gives the same configuration (the cross-backend identity
    of src/pyising/validate.py:68-93, across this backend's kernels)."""
    from .backends.hip_backend import MSC_LOCKSTEP, HipMultiSpin
    from .batch import BatchScan

    lockstep = type("HipMultiSpinLockstep", (HipMultiSpin,), {"mode": MSC_LOCKSTEP})
    cases = [((32, 64), "metropolis", "int8"), ((32, 64), "heatbath", "int8"),
             ((8, 8, 16), "metropolis", "int8"), ((8, 8, 16), "heatbath", "int8"),
             ((128, 256), "metropolis", "msc")]
    out = []
    for shape, alg, variant in cases:
        beta = 1.0 / (TC_2D if len(shape) == 2 else TC_3D)
        kw = {"shape": shape, "beta": beta, "seed": 42, "algorithm": alg, "init": "hot"}
        ref = B.create("hip", variant, **kw)
        ref.sweep(sweeps)
        want = ref.get_lattice()
        paths = {"batched": B.create("hip", variant, **kw)}
        if variant == "msc":
            paths["lockstep"] = lockstep(**kw)
            paths["eager32"] = B.create("hip", "msc-eager32", **kw)
        for name, sim in paths.items():
            if name == "batched":
                scan = BatchScan([sim])
                scan.sweep(sweeps)
                scan.synchronize()
                scan.close()
            else:
                sim.sweep(sweeps)
            same = bool(np.array_equal(sim.get_lattice(), want))
            out.append({"variant": variant, "path": name, "shape": list(shape),
                        "algorithm": alg, "identical_to_default": same})
            log(f"  identity hip/{variant:<5} {name:<9} {alg:<10} {shape}: {same}")
    return out


def energy_scan_2d(cfg: dict, log=print) -> list[dict]:
    """2D |m|(T), e(T) for Metropolis int8, Metropolis msc, heat-bath int8
    (src/pyising/validate.py:96-119)."""
    rows = []
    for k, (variant, alg) in enumerate((("int8", "metropolis"), ("msc", "metropolis"),
                                         ("int8", "heatbath"))):
        L = -(-cfg["L2"] // 128) * 128 if variant == "msc" else cfg["L2"]  # msc rows: 128 bits
        for i, T in enumerate(cfg["T2"]):
            # an independent Philox key per (variant, T): with one key the uniforms would be
            # the same at every T and the errors would correlate
            sim = B.create("hip", variant, shape=(L, L), beta=1.0 / T, seed=7 + 1000 * k + i,
                           algorithm=alg, init=_init(T, TC_2D))
            s = sample(sim, cfg["therm"], _n_meas(cfg, T, TC_2D), cfg["every"]).summary()
            e_ex = energy_torus(T, L)
            row = {"backend": "hip", "variant": variant, "algorithm": alg, "L": L,
                   "T": float(T), **s, "e_exact_L": e_ex,
                   "m_exact_inf": float(magnetization_2d(T))}
            row["e_z"] = (s["e"] - e_ex) / s["e_err"] if s["e_err"] > 0 else float("nan")
            rows.append(row)
            log(f"  2D {variant:<5} {alg:<10} L={L} T={T:.4f} |m|={s['abs_m']:.5f} "
                f"(Yang {row['m_exact_inf']:.5f}) e={s['e']:.6f}+-{s['e_err']:.6f} "
                f"(exact L={L}: {e_ex:.6f}, z={row['e_z']:+.2f})")
    return rows


def binder_scan(dim: int, Ls: list[int], Ts: list[float], cfg: dict, log=print) -> dict:
    """Binder cumulants of every (L, T) from one batched scan (device moment sums, blocked
    jackknife errors) and the crossings of consecutive sizes (src/pyising/validate.py:128-152)."""
    from .batch import BatchScan

    sims, keys = [], []
    for L in Ls:
        for j, T in enumerate(Ts):
            sims.append(B.create("hip", "int8", shape=(L,) * dim, beta=1.0 / T,
                                 seed=1000 * L + j + 17 * dim, init="hot"))
            keys.append((L, j))
    t0 = time.time()
    scan = BatchScan(sims)
    scan.sweep(cfg["binder_therm"])
    res = scan_binder(scan, cfg["binder_blocks"], cfg["binder_block_sweeps"])
    scan.close()
    data: dict = {"dim": dim, "L": Ls, "T": [float(t) for t in Ts], "U": {}, "U_err": {}}
    for (L, j), (u, err) in zip(keys, res, strict=True):
        data["U"].setdefault(str(L), [None] * len(Ts))[j] = u
        data["U_err"].setdefault(str(L), [None] * len(Ts))[j] = err
    for L in Ls:
        log(f"  Binder {dim}D L={L}: " + " ".join(f"{u:.4f}" for u in data["U"][str(L)]))
    data["crossings"] = []
    for a, b in zip(Ls[:-1], Ls[1:], strict=True):
        tc = crossing(np.array(Ts), np.array(data["U"][str(a)]), np.array(data["U"][str(b)]))
        data["crossings"].append({"L1": a, "L2": b, "Tc": tc})
        log(f"  crossing {dim}D L={a}/{b}: Tc = {tc:.5f}")
    ref = TC_2D if dim == 2 else TC_3D
    data["Tc_reference"] = ref
    data["Tc_estimate"] = data["crossings"][-1]["Tc"]
    data["Tc_rel_error"] = abs(data["Tc_estimate"] - ref) / ref
    data["seconds"] = time.time() - t0
    return data


def scan_3d(cfg: dict, log=print) -> list[dict]:
    """3D Metropolis vs heat-bath energies at the same L and T (the reference compares both with
    Wolff, src/pyising/validate.py:155-174)."""
    rows = []
    L = cfg["L3"]
    for i, T in enumerate(cfg["T3"]):
        s = {}
        for k, alg in enumerate(("metropolis", "heatbath")):
            sim = B.create("hip", "int8", shape=(L, L, L), beta=1.0 / T, seed=11 + 1000 * k + i,
                           algorithm=alg, init=_init(T, TC_3D))
            s[alg] = sample(sim, cfg["therm"], _n_meas(cfg, T, TC_3D), cfg["every"]).summary()
        me, hb = s["metropolis"], s["heatbath"]
        err = float(np.hypot(me["e_err"], hb["e_err"]))
        rows.append({"backend": "hip", "variant": "int8", "L": L, "T": float(T),
                     "metropolis": me, "heatbath": hb, "e_z": (me["e"] - hb["e"]) / err})
        log(f"  3D L={L} T={T:.4f} metropolis e={me['e']:.5f}+-{me['e_err']:.5f} heatbath "
            f"{hb['e']:.5f}+-{hb['e_err']:.5f} (z={rows[-1]['e_z']:+.2f})")
    return rows


def summarize(res: dict) -> dict:
    """Headline numbers (src/pyising/validate.py:200-215) and the pass verdict."""
    s2 = res["scan2d"]
    away = [r for r in s2 if abs(r["T"] - TC_2D) > 0.2 and r["T"] < TC_2D]
    s = {
        "all_identical": all(r["identical_to_default"] for r in res["identity"]),
        "n_identity_checks": len(res["identity"]),
        "e_max_abs_err_vs_exact_L": max(abs(r["e"] - r["e_exact_L"]) for r in s2),
        "e_max_abs_z": max(abs(r["e_z"]) for r in s2),
        "m_max_abs_err_vs_yang_away_from_Tc": max(
            (abs(r["abs_m"] - r["m_exact_inf"]) for r in away), default=float("nan")),
        "Tc2d_estimate": res["binder2d"]["Tc_estimate"],
        "Tc2d_rel_error": res["binder2d"]["Tc_rel_error"],
        "Tc3d_estimate": res["binder3d"]["Tc_estimate"],
        "Tc3d_rel_error": res["binder3d"]["Tc_rel_error"],
        "e3d_max_abs_z_metropolis_vs_heatbath": max(abs(r["e_z"]) for r in res["scan3d"]),
    }
    m_err = s["m_max_abs_err_vs_yang_away_from_Tc"]
    s["all_passed"] = bool(s["all_identical"] and s["e_max_abs_z"] < E_Z_MAX
                           and (not np.isfinite(m_err) or m_err < M_ABS_MAX)
                           and s["Tc2d_rel_error"] < TC_REL_MAX
                           and s["Tc3d_rel_error"] < TC_REL_MAX
                           and s["e3d_max_abs_z_metropolis_vs_heatbath"] < E_Z_MAX)
    return s


def run(quick: bool = False, out: str | Path | None = "results/validation.json", log=print,
        cfg: dict | None = None) -> dict:
    """Run the whole suite; returns the result dict and writes it as JSON (``out=None``: no
    file). ``cfg`` overrides the QUICK / FULL configuration."""
    cfg = dict(cfg or (QUICK if quick else FULL))
    t0 = time.time()
    res: dict = {"quick": quick, "config": cfg}
    log("[1/5] identity across kernels")
    res["identity"] = check_identity(log)
    log("[2/5] 2D energy and magnetisation vs exact results")
    res["scan2d"] = energy_scan_2d(cfg, log)
    log("[3/5] 2D Binder crossing (batched scan)")
    res["binder2d"] = binder_scan(2, cfg["binder2_L"], cfg["binder2_T"], cfg, log)
    log("[4/5] 3D Binder crossing (batched scan)")
    res["binder3d"] = binder_scan(3, cfg["binder3_L"], cfg["binder3_T"], cfg, log)
    log("[5/5] 3D Metropolis vs heat bath")
    res["scan3d"] = scan_3d(cfg, log)
    res["summary"] = summarize(res)
    res["seconds"] = time.time() - t0
    if out is not None:
        Path(out).parent.mkdir(parents=True, exist_ok=True)
        Path(out).write_text(json.dumps(res, indent=1, default=float))
        log(f"wrote {out} ({res['seconds']:.0f} s)")
    return res
