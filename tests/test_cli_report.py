"""Command line, benchmark report and validation driver: argument handling, the results layout
of bench/run.py:196-240 and the validation verdict on the CPU; on the GPU, each command runs on
small sizes and its physics checks pass."""

import json

import numpy as np
import pytest

from pyising_amd import cli, report, validate
from pyising_amd.lattice import TC_2D, TC_3D

#: keys of one run record and of the ceiling dict in the reference's results files
#: (bench/run.py:196-217, bench/roofline.py:107-117)
RUN_KEYS = {"backend", "variant", "algorithm", "size", "dim", "dtype", "metric", "value",
            "repeats", "median", "min", "max", "sweeps_per_repeat", "bytes_per_unit",
            "effective_GBs", "pct_copy_bw", "pct_spec_bw", "flops_per_unit", "pct_peak_flops",
            "memory_ceiling_flips_per_ns", "philox_calls_per_update", "rng_ceiling_flips_per_ns",
            "pct_roofline", "temperature", "final_m", "final_e", "notes"}
CEILING_KEYS = {"copy_bandwidth_GBs", "copy_bandwidth_measured_GBs", "copy_bandwidth_source",
                "spec_bandwidth_GBs", "fp32_tflops", "philox_calls_per_ns",
                "msc_philox_calls_per_word_at_Tc", "msc_philox_calls_per_word_warp_max_at_Tc",
                "copy_bandwidth_bytes"}
CEILING = {"copy_bandwidth_GBs": 5000.0, "copy_bandwidth_measured_GBs": 5000.0,
           "copy_bandwidth_source": "measured", "spec_bandwidth_GBs": 8000.0,
           "fp32_tflops": 157.3, "philox_calls_per_ns": 1000.0,
           "msc_philox_calls_per_word_at_Tc": 3.8, "msc_philox_calls_per_word_warp_max_at_Tc": 6.3,
           "msc_philox_slots_per_word_queue_at_Tc": 4.0, "copy_bandwidth_bytes": 1 << 30}


class _Int8:
    stream, variant, bytes_per_update = "int8", "int8", 3.0


class _Msc:
    stream, variant, bytes_per_update = "msc", "msc", 3.0 / 8.0


class _Eager(_Msc):
    variant = "msc-eager32"


def test_parser_accepts_every_command():
    p = cli.parser()
    a = p.parse_args(["run", "--L", "256", "--dim", "3", "--algorithm", "heatbath", "--T", "4.5"])
    assert (a.cmd, a.L, a.dim, a.algorithm, a.T, a.backend) == ("run", 256, 3, "heatbath", 4.5,
                                                                 "hip")
    assert p.parse_args(["bench", "--repeats", "2"]).repeats == 2
    assert p.parse_args(["validate", "--quick"]).quick
    assert p.parse_args(["list"]).cmd == "list"
    with pytest.raises(SystemExit):
        p.parse_args(["run", "--backend", "numpy"])


def test_list_runs_without_a_gpu(capsys, monkeypatch):
    from pyising_amd import backends as B

    monkeypatch.setattr(B, "load", lambda be, var: None)
    assert cli.main(["list"]) == 0
    out = capsys.readouterr().out
    assert out.count("unavailable") == len(B.REGISTRY)
    assert "msc" in out and "int8" in out


def test_report_main_needs_a_mode():
    with pytest.raises(SystemExit):
        report.main([])


def test_calls_per_update():
    assert report.calls_per_update(_Int8, CEILING) == 0.25
    assert report.calls_per_update(_Eager, CEILING) == 16 / 64
    assert report.calls_per_update(_Msc, CEILING) == 4.0 / 64


def test_run_record_layout_and_roofline():
    r = {"median": 6000.0, "min": 5900.0, "max": 6100.0, "sweeps_per_repeat": 100,
         "m": 0.01, "e": -1.41}
    rec = report.run_record(_Msc, "msc", "metropolis", (16384, 16384), r, CEILING, 5,
                            1 / TC_2D)
    assert set(rec) == RUN_KEYS
    assert rec["effective_GBs"] == pytest.approx(6000 * 3 / 8)
    assert rec["memory_ceiling_flips_per_ns"] == pytest.approx(5000 / (3 / 8), abs=0.1)
    # RNG ceiling 1000 / (4/64) = 16000 > memory ceiling 13333: the memory ceiling bounds it
    assert rec["rng_ceiling_flips_per_ns"] == pytest.approx(16000.0)
    assert rec["pct_roofline"] == pytest.approx(100 * 6000 / (5000 / (3 / 8)), abs=0.01)
    assert rec["pct_copy_bw"] == pytest.approx(100 * 2250 / 5000)
    rec8 = report.run_record(_Int8, "int8", "heatbath", (256,) * 3, r | {"median": 1000.0},
                             CEILING, 5, 1 / TC_3D)
    assert rec8["dtype"] == "int8" and rec8["dim"] == 3
    assert rec8["pct_roofline"] == pytest.approx(100 * 1000 / (5000 / 3), abs=0.01)


def _fake_report():
    r = {"median": 1.0, "min": 1.0, "max": 1.0, "sweeps_per_repeat": 1, "m": 0.0, "e": -1.0}
    runs = []
    for cls, var, med in ((_Int8, "int8", 1500.0), (_Int8, "int8", 1600.0),
                          (_Msc, "msc", 6500.0), (_Eager, "msc-eager32", 2000.0)):
        runs.append(report.run_record(cls, var, "metropolis", (8192, 8192), r | {"median": med},
                                      CEILING, 3, 1 / TC_2D))
    return {"machine": {"gpu": "test"}, "ceiling": CEILING, "provisional": True,
            "mode": "quick", "runs": runs, "seconds": 1.0}


def test_best_per_variant_markdown_and_figure(tmp_path):
    rep = _fake_report()
    best = report.best_per_variant(rep)
    assert best[("hip", "int8")]["median"] == 1600.0 and len(best) == 3
    md = report.markdown(rep)
    rows = [ln for ln in md.splitlines() if ln.startswith("|")]
    assert len(rows) == 1 + 1 + 4  # header, rule, four runs
    report.roofline_figure(rep, tmp_path / "roofline.png")
    assert (tmp_path / "roofline.png").stat().st_size > 1000
    report.scaling_figure(rep, tmp_path / "scaling.png")
    assert (tmp_path / "scaling.png").stat().st_size > 1000


def test_report_redraws_from_json(tmp_path):
    path = tmp_path / "r.json"
    path.write_text(json.dumps(_fake_report()))
    assert report.main(["--quick", "--from-json", str(path), "--markdown", str(tmp_path / "r.md"),
                        "--scaling-figure", str(tmp_path / "s.png")]) == 0
    assert "| msc | metropolis | 8192x8192 |" in (tmp_path / "r.md").read_text()
    assert (tmp_path / "s.png").exists()


def _fake_validation(e_z=1.0, m=0.8, tc2=TC_2D * 1.005, tc3=TC_3D, same=True):
    from pyising_amd.exact import magnetization_2d

    scan = [{"T": 1.8, "e": -1.8, "e_exact_L": -1.8, "e_z": e_z, "abs_m": m,
             "m_exact_inf": float(magnetization_2d(1.8))},
            {"T": TC_2D, "e": -1.4, "e_exact_L": -1.4, "e_z": -0.5, "abs_m": 0.6,
             "m_exact_inf": 0.0}]
    return {"identity": [{"identical_to_default": same}],
            "scan2d": scan,
            "binder2d": {"Tc_estimate": tc2, "Tc_rel_error": abs(tc2 - TC_2D) / TC_2D},
            "binder3d": {"Tc_estimate": tc3, "Tc_rel_error": abs(tc3 - TC_3D) / TC_3D},
            "scan3d": [{"e_z": 0.3}]}


def test_validation_verdict():
    from pyising_amd.exact import magnetization_2d

    m = float(magnetization_2d(1.8))
    assert validate.summarize(_fake_validation(m=m))["all_passed"]
    assert not validate.summarize(_fake_validation(m=m, e_z=6.0))["all_passed"]
    assert not validate.summarize(_fake_validation(m=m - 0.05))["all_passed"]
    assert not validate.summarize(_fake_validation(m=m, tc2=2.4))["all_passed"]
    assert not validate.summarize(_fake_validation(m=m, same=False))["all_passed"]
    s = validate.summarize(_fake_validation(m=m))
    assert s["m_max_abs_err_vs_yang_away_from_Tc"] == pytest.approx(0.0)
    assert s["n_identity_checks"] == 1


def test_more_measurements_near_tc():
    cfg = {"meas": 10}
    assert validate._n_meas(cfg, TC_2D, TC_2D) == 40
    assert validate._n_meas(cfg, 1.8, TC_2D) == 10
    assert validate._n_meas(cfg, 4.6, TC_3D) == 40


# ----------------------------------------------------------------------------------- GPU --

SMALL = {
    "L2": 32, "T2": [1.8, 3.0], "therm": 5000, "meas": 1000, "every": 5,
    "binder2_L": [8, 16], "binder2_T": list(np.linspace(2.1, 2.45, 8)),
    "binder3_L": [8, 16], "binder3_T": list(np.linspace(4.3, 4.7, 8)),
    "binder_therm": 5000, "binder_blocks": 16, "binder_block_sweeps": 2000,
    "L3": 8, "T3": [4.0, 5.0],
}


@pytest.mark.gpu
def test_validation_suite_small(tmp_path):
    res = validate.run(cfg=SMALL, out=tmp_path / "v.json", log=print)
    s = res["summary"]
    print(json.dumps(s))
    assert s["all_identical"] and s["n_identity_checks"] == 7
    assert s["e_max_abs_z"] < validate.E_Z_MAX
    assert s["m_max_abs_err_vs_yang_away_from_Tc"] < validate.M_ABS_MAX
    assert s["Tc2d_rel_error"] < 0.05 and s["Tc3d_rel_error"] < 0.05
    assert s["e3d_max_abs_z_metropolis_vs_heatbath"] < validate.E_Z_MAX
    assert json.loads((tmp_path / "v.json").read_text())["summary"] == json.loads(json.dumps(s))


@pytest.mark.gpu
def test_report_quick_layout(tmp_path, monkeypatch):
    monkeypatch.setattr(report, "PLAN_QUICK", [("int8", 2, [256], "metropolis"),
                                               ("msc", 2, [1024], "metropolis"),
                                               ("int8", 3, [64], "heatbath")])
    res = report.run("quick", tmp_path / "r.json", repeats=2, log=print)
    assert set(res) == {"machine", "ceiling", "provisional", "mode", "runs", "seconds"}
    assert CEILING_KEYS <= set(res["ceiling"])
    assert [r["variant"] for r in res["runs"]] == ["int8", "msc", "int8"]
    for r in res["runs"]:
        assert set(r) == RUN_KEYS
        assert r["median"] > 0 and 0 < r["pct_roofline"] < 150
        assert -1.0 <= r["final_m"] <= 1.0 and -3.0 <= r["final_e"] <= 0.0
    assert json.loads((tmp_path / "r.json").read_text())["runs"][1]["size"] == [1024, 1024]
    assert "gfx" in res["machine"]["gpu"]


@pytest.mark.gpu
def test_cli_run_and_bench(capsys):
    assert cli.main(["run", "--L", "256", "--sweeps", "20", "--T", "1.5", "--init", "cold"]) == 0
    out = capsys.readouterr().out
    assert "hip/int8 metropolis shape=(256, 256)" in out and "flips/ns" in out
    m = float(out.split("m = ")[1].split()[0])
    assert m > 0.95  # cold start far below T_c stays ordered
    assert cli.main(["bench", "--variant", "msc", "--L", "1024", "--sweeps", "50",
                     "--repeats", "2"]) == 0
    assert "hip/msc L=1024 dim=2: median" in capsys.readouterr().out


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with status 2 and says why, before any
    device work."""
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_committed_pmc_profile_matches_the_kernel_sources():
    """bench.py reports roofline.traffic from profiles/msc_L16384_pmc.json only when that
    profile was measured on the msc kernel sources and flags built now (their digest is recorded
    in it); a change to those sources must come with a new profile, or the line's traffic turns
    null. Checked here so that the mismatch shows before a GPU run."""
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    sys.path.insert(0, str(root))
    try:
        import bench
    finally:
        sys.path.remove(str(root))
    traffic, note = bench.traffic_from_profile()
    assert traffic is not None, note
    prof = json.loads((root / "profiles" / "msc_L16384_pmc.json").read_text())
    assert traffic == prof["hbm_bytes_per_launch"]
    assert 0 < traffic < prof["algorithmic_bytes_per_launch"]  # each XCD keeps its band in L2
