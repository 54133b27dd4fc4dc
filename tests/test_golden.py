"""The committed golden fixtures (tests/golden/, written by tests/golden/make_golden.py).

CPU: the oracle reproduces the known answers the reference's own tests hold
(reference_known_answers.json: tests/test_philox.py:6-34, tests/test_lattice.py:33-38,
tests/test_backends.py:73-77, tests/test_exact.py:30-34 of the reference) and its own recorded
end states (oracle_end_states.json), so a later change of the oracle cannot go unnoticed.
GPU: the hip backend, called through the C ABI, reproduces every recorded end state (sha256 of
the full int8 lattice and exact M, E) without running the oracle.
"""

import hashlib
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
KA = json.loads((GOLDEN / "reference_known_answers.json").read_text())
STATES = json.loads((GOLDEN / "oracle_end_states.json").read_text())["cases"]
# recomputed on the CPU by the oracle in well under a second each; the larger ones are checked
# on the GPU only
CPU_CASES = [c for c in STATES if int(np.prod(c["shape"])) <= 2048 * 2048]


def digest(lat) -> str:
    return hashlib.sha256(np.ascontiguousarray(lat, dtype=np.int8).tobytes()).hexdigest()


def initial(c):
    """The case's initial lattice: "hot" / "cold", or the reference's random_lattice
    (numpy default_rng(init_seed) draws, mapped to +1 / -1)."""
    if c["init"] != "random":
        return c["init"]
    rng = np.random.default_rng(c["init_seed"])
    return rng.integers(0, 2, tuple(c["shape"]), dtype=np.int8) * 2 - 1


# ------------------------------------------------------------------------ CPU: oracle ----
@pytest.mark.parametrize("case", KA["philox4x32_10"]["cases"])
def test_philox_known_answers(case):
    from oracle import c_oracle as CO
    from oracle import ising_oracle as O

    ctr, key = case["counter"], case["key"]
    assert [int(x) for x in O.philox4x32(*ctr, *key)] == case["expected"]
    assert [int(x) for x in CO.philox(np.array(ctr), np.array(key))[0]] == case["expected"]


def test_seed_split_and_stream_layout():
    from oracle import ising_oracle as O

    for c in KA["split_seed"]["cases"]:
        assert list(O.split_seed(c["seed"])) == c["expected"]
    s = KA["site_uniforms_layout"]
    u = O.site_uniforms(s["n_sites"], s["sweep"], s["color"], s["seed"])
    k0, k1 = O.split_seed(s["seed"])
    r = O.philox4x32(s["quad"], s["sweep"], s["color"], 0, k0, k1)
    q = 4 * s["quad"]
    assert [int(x) for x in u[q:q + 4]] == [int(x) for x in r]


@pytest.mark.parametrize("case", KA["threshold_table"]["cases"])
def test_threshold_known_answers(case):
    from oracle import ising_oracle as O
    from pyising_amd import lattice as L

    for tab in (O.threshold_table(case["beta"], case["ndim"], case["algorithm"]),
                L.threshold_table(case["beta"], case["ndim"], case["algorithm"])):
        if "length" in case:
            assert len(tab) == case["length"]
        for i, v in case["entries"].items():
            assert int(tab[int(i)]) == v


def test_config_threshold_tables_recorded():
    from pyising_amd import lattice as L

    rec = KA["threshold_table"]["config_tables"]
    bc2, bc3 = 0.5 * math.log(1 + math.sqrt(2)), 0.221654626
    assert [int(x) for x in L.threshold_table(bc2, 2, "metropolis")] == rec["2d_metropolis_beta_c"]
    assert [int(x) for x in L.threshold_table(0.4406868, 2, "metropolis")] == rec["2d_metropolis_0.4406868"]
    assert [int(x) for x in L.threshold_table(bc3, 3, "metropolis")] == rec["3d_metropolis_beta_c"]
    assert [int(x) for x in L.threshold_table(bc3, 3, "heatbath")] == rec["3d_heatbath_beta_c"]


def test_cold_start_known_answers():
    from oracle import ising_oracle as O

    c = KA["cold_start_observables"]
    for shape, key in (((128, 128), "2d"), ((16, 16, 16), "3d")):
        sim = O.Int8Oracle(shape, 0.5, init="cold")
        assert list(sim.observables()) == pytest.approx(c[key])
    assert list(O.MscOracle((128, 128), 0.5, init="cold").observables()) == pytest.approx(c["2d"])


def test_exact_known_answers():
    from pyising_amd.exact import TC_2D, energy_2d, magnetization_2d

    c = KA["exact_2d"]
    assert TC_2D == pytest.approx(c["tc"], rel=1e-14)
    assert energy_2d(TC_2D) == pytest.approx(c["energy_at_tc"], rel=1e-9)
    assert magnetization_2d(2.0) == pytest.approx(c["magnetization_at_T2"])
    assert magnetization_2d(3.0) == c["magnetization_at_T3"]


@pytest.mark.parametrize("case", CPU_CASES, ids=[c["name"] for c in CPU_CASES])
def test_oracle_reproduces_end_state(case):
    from oracle import c_oracle as CO
    from oracle import ising_oracle as O

    if case["engine"] == "numpy":
        cls = O.Int8Oracle if case["variant"] == "int8" else O.MscOracle
    else:
        cls = CO.CInt8 if case["variant"] == "int8" else CO.CMsc
    sim = cls(tuple(case["shape"]), case["beta"], seed=case["seed"],
              algorithm=case["algorithm"], init=initial(case))
    sim.sweep(case["sweeps"])
    if case["then"]:
        sim.set_beta(case["then"][0])
        sim.sweep(case["then"][1])
    assert list(sim.observables_exact()) == [case["M"], case["E"]]
    assert digest(sim.get_lattice()) == case["sha256"]


def test_planted_threshold_swap_changes_the_end_state(monkeypatch):
    """Swapped dE = 4 and dE = 8 thresholds (planted in the oracle) must not reproduce the
    recorded state."""
    from oracle import ising_oracle as O

    case = next(c for c in STATES if c["name"] == "msc_128x256_beta_c_20")
    real = O.threshold_table

    def swapped(beta, ndim, algorithm):
        t = real(beta, ndim, algorithm).copy()
        t[3], t[4] = t[4], t[3]
        return t

    monkeypatch.setattr(O, "threshold_table", swapped)
    sim = O.MscOracle(tuple(case["shape"]), case["beta"], seed=case["seed"], init=initial(case))
    sim.sweep(case["sweeps"])
    assert digest(sim.get_lattice()) != case["sha256"]


def test_fixture_covers_the_configurations():
    names = {c["name"] for c in STATES}
    for need in ("int8_2d_L1024_hot_100", "msc_L16384_beta_c_hot_3",
                 "int8_3d_L256_Kc_metropolis_hot_10", "int8_3d_L256_Kc_heatbath_hot_10",
                 "int8_2d_L8192_T2.269_hot_10"):
        assert need in names


# ------------------------------------------------------------------- GPU: hip backend ----
@pytest.mark.gpu
@pytest.mark.parametrize("case", STATES, ids=[c["name"] for c in STATES])
def test_hip_reproduces_end_state(case):
    from pyising_amd.backends import hip_backend as H

    cls = H.HipInt8 if case["variant"] == "int8" else H.HipMultiSpin
    print(f"golden case {case['name']}", flush=True)
    sim = cls(tuple(case["shape"]), case["beta"], seed=case["seed"],
              algorithm=case["algorithm"], init=initial(case))
    sim.sweep(case["sweeps"])
    if case["then"]:
        sim.set_beta(case["then"][0])
        sim.sweep(case["then"][1])
    assert list(sim.observables_exact()) == [case["M"], case["E"]]
    assert digest(sim.get_lattice()) == case["sha256"]
