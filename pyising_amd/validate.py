"""Physics validation suite on the GPU (src/pyising/validate.py:68-215, with this backend's own
samplers) -> ``results/validation.json``.

1. Identity across kernels: the default msc kernel, the lockstep kernel, the eager ablation and
   the batched scan leave the same lattice after the same sweeps; int8 alone and batched too.
2. 2D energy per spin vs the exact finite-torus value (Kaufman), |m| vs Onsager-Yang, for
   Metropolis int8, Metropolis msc and heat-bath int8.
3. 2D Binder-cumulant crossings of a batched temperature scan (every pair of sizes; the two
   largest give the estimate) vs Onsager's T_c.
4. 3D Binder crossings likewise vs the simple-cubic critical coupling (Ferrenberg, Xu &
   Landau 2018).
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

def _around(center: float, half_width: float, n: int, digits: int = 4) -> list[float]:
    """n temperatures evenly spaced over center +- half_width."""
    return [round(float(t), digits) for t in np.linspace(center - half_width,
                                                          center + half_width, n)]


def _plan(L2: int, t2: list, sweeps: tuple[int, int, int], binder_sizes: tuple[list, list],
          binder_points: tuple[int, int], binder_sweeps: tuple[int, int, int], L3: int,
          t3: list) -> dict:
    """A validation configuration. sweeps = (thermalisation, measurements, sweeps between
    measurements) of the single-lattice runs; binder_points = temperatures of the 2D and 3D
    Binder scans (over 2.27 +- 0.07 and 4.51 +- 0.05); binder_sweeps = (thermalisation,
    blocks, sweeps per block) of the batched Binder scans."""
    therm, meas, every = sweeps
    b_therm, b_blocks, b_sweeps = binder_sweeps
    return {"L2": L2, "T2": list(t2), "therm": therm, "meas": meas, "every": every,
            "binder2_L": binder_sizes[0], "binder2_T": _around(2.27, 0.07, binder_points[0]),
            "binder3_L": binder_sizes[1], "binder3_T": _around(4.51, 0.05, binder_points[1]),
            "binder_therm": b_therm, "binder_blocks": b_blocks, "binder_block_sweeps": b_sweeps,
            "L3": L3, "T3": list(t3)}


#: about 20 s on one MI355X
QUICK = _plan(64, [1.8, 2.0, TC_2D, 2.5, 3.0], (20000, 2000, 20), ([16, 32], [8, 16]), (6, 6),
              (20000, 32, 2000), 16, [4.0, TC_3D, 5.0])
#: the full suite: 2D temperatures over [1.6, 3.4] and T_c, three Binder sizes per dimension
FULL = _plan(128, sorted(_around(2.5, 0.9, 19, 3) + [TC_2D]), (50000, 5000, 40),
             ([16, 32, 64], [8, 16, 32]), (15, 13), (100000, 40, 10000), 32,
             [3.5, 4.2, TC_3D, 5.0, 6.0])
#: pass criteria: |z| of an energy, |m| error vs Onsager-Yang away from T_c, relative T_c error
E_Z_MAX, M_ABS_MAX, TC_REL_MAX = 5.0, 0.01, 0.02
#: relative distance from T_c within which autocorrelation times are long enough to need more
#: measurements, and how many more
NEAR_TC, NEAR_TC_FACTOR = 0.065, 4


def _n_meas(cfg: dict, T: float, tc: float) -> int:
    """Measurements at temperature T: more within NEAR_TC (relative) of the critical
    temperature, where local updates decorrelate slowly."""
    near = abs(T / tc - 1.0) < NEAR_TC
    return cfg["meas"] * (NEAR_TC_FACTOR if near else 1)


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
    # every pair of sizes crosses near T_c; the pair of the two largest sizes, whose
    # corrections to scaling are smallest, gives the estimate
    temps = np.array(Ts)
    pairs = [(a, b) for i, a in enumerate(Ls) for b in Ls[i + 1:]]
    data["crossings"] = [{"L1": a, "L2": b,
                          "Tc": crossing(temps, np.array(data["U"][str(a)]),
                                         np.array(data["U"][str(b)]))} for a, b in pairs]
    for c in data["crossings"]:
        log(f"  crossing {dim}D L={c['L1']}/{c['L2']}: Tc = {c['Tc']:.5f}")
    largest = max(data["crossings"], key=lambda c: (c["L2"], c["L1"]))
    data["Tc_reference"] = TC_2D if dim == 2 else TC_3D
    data["Tc_estimate"] = largest["Tc"]
    data["Tc_rel_error"] = abs(largest["Tc"] / data["Tc_reference"] - 1.0)
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
    """The headline numbers of a validation result (the key names of the reference's
    results/validation.json) and the pass verdict over all of them."""
    s2, s3 = res["scan2d"], res["scan3d"]
    ordered = [r for r in s2 if r["T"] < TC_2D - 0.2]  # |m| is compared where it is large
    b2, b3 = res["binder2d"], res["binder3d"]
    checks = [  # (summary key, value, limit or None)
        ("all_identical", all(r["identical_to_default"] for r in res["identity"]), None),
        ("n_identity_checks", len(res["identity"]), None),
        ("e_max_abs_err_vs_exact_L", max(abs(r["e"] - r["e_exact_L"]) for r in s2), None),
        ("e_max_abs_z", max(abs(r["e_z"]) for r in s2), E_Z_MAX),
        ("m_max_abs_err_vs_yang_away_from_Tc",
         max((abs(r["abs_m"] - r["m_exact_inf"]) for r in ordered), default=float("nan")),
         M_ABS_MAX),
        ("Tc2d_estimate", b2["Tc_estimate"], None),
        ("Tc2d_rel_error", b2["Tc_rel_error"], TC_REL_MAX),
        ("Tc3d_estimate", b3["Tc_estimate"], None),
        ("Tc3d_rel_error", b3["Tc_rel_error"], TC_REL_MAX),
        ("e3d_max_abs_z_metropolis_vs_heatbath", max(abs(r["e_z"]) for r in s3), E_Z_MAX),
    ]
    out = {key: value for key, value, _ in checks}
    # a NaN (no ordered temperature in the scan) is no failure
    within = all(not (v >= lim) for _, v, lim in checks if lim is not None)
    out["all_passed"] = bool(out["all_identical"] and within)
    return out


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
