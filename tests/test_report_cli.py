"""Convergence report writer and command-line driver (CPU: no solver runs)."""
import csv
import json
import os
from types import SimpleNamespace

import pytest

import convergence_report as cr
import run_neohookean


def test_cli_defaults_are_the_reference_drivers():
    """run_neohookean_usd.py:15-75 defaults."""
    a = run_neohookean.build_parser().parse_args([])
    assert a.model == "assets/tetmesh/cow.node" and a.scale == 1.0 and a.arch == "gpu"
    assert (a.frames, a.fps, a.dt, a.substeps, a.iterations) == (300, 60, 0.01, 1, 200)
    assert (a.young, a.poisson, a.hydro_relax, a.dev_relax) == (5e8, 0.4999, 1.2, 1.2)
    assert a.block_neohookean and a.cheb and a.cheb_warmup == 5 and a.cheb_gamma == 0.666
    assert not a.monitor_convergence and not a.export_csv and a.plot_frame_ids is None
    b = run_neohookean.build_parser().parse_args(["--split-neohookean", "--no-cheb"])
    assert not b.block_neohookean and not b.cheb


def test_output_paths_follow_the_reference(tmp_path):
    """usd_export.py:10-34 naming, which the comparison tools expect."""
    assert run_neohookean.method_tag(True, True) == "block_neohookean_dynamic_cheby"
    assert run_neohookean.method_tag(False, False) == "split_neohookean_no_cheby"
    assert run_neohookean.young_tag(5e8) == "5p000e08"
    out, run = run_neohookean.build_output_paths("assets/tetmesh/cow.node", True, False, 1e5,
                                                 root=str(tmp_path))
    assert out == os.path.join(str(tmp_path), "output", "cow")
    assert run == os.path.join(out, "xpbd_block_neohookean_no_cheby_young_1p000e05")
    assert os.path.isdir(run)


def test_cpu_arch_is_refused():
    assert run_neohookean.main(["--arch", "cpu"]) == 2


def fake_solver(n_iter=4):
    recs = [dict(frame=f, substep=s, iter=i, iter_compute_time_sec=1e-4,
                 hydro_error=1.0 / (i + 1) + f, dev_error=2.0 / (i + 1) + s)
            for f in range(2) for s in range(2) for i in range(n_iter)]
    return SimpleNamespace(
        solve_compute_calls=4, solve_compute_iterations=16, solve_compute_time_total=0.016,
        constraints_iter=n_iter, hydrostatic_relaxation=1.2, deviatoric_relaxation=1.2,
        cheb_enable=True, cheb_gamma=0.666, cheb_warmup=5, monitor_convergence=True,
        convergence_history=recs, convergence_callback=None)


def test_substep_json_has_the_reference_schema(tmp_path):
    s = fake_solver()
    data = [r for r in s.convergence_history if r["frame"] == 1 and r["substep"] == 0]
    path = tmp_path / "f0001_s00.png"
    cr.save_substep_convergence(s, 1, 0, data, str(path))
    doc = json.loads((tmp_path / "f0001_s00.json").read_text())
    assert set(doc) == {"frame", "substep", "iterations", "iter_compute_times_sec",
                        "hydro_errors", "dev_errors", "records", "solver_type",
                        "compute_timing", "config"}
    assert doc["iterations"] == [0, 1, 2, 3] and doc["hydro_errors"][0] == 2.0
    assert set(doc["compute_timing"]) == {
        "solve_step_compute_count", "constraint_iteration_count", "total_compute_time_sec",
        "avg_compute_time_per_solve_step_sec", "avg_compute_time_per_iteration_sec"}
    assert doc["compute_timing"]["avg_compute_time_per_iteration_sec"] == pytest.approx(1e-3)
    assert set(doc["config"]) == {"constraints_iter", "hydrostatic_relaxation",
                                  "deviatoric_relaxation", "cheb_enable", "cheb_gamma",
                                  "cheb_warmup"}
    assert path.exists() and path.stat().st_size > 0


def test_stats_csv_and_callback(tmp_path, capsys):
    s = fake_solver()
    st = cr.get_convergence_stats(s)
    assert st["total_records"] == 16 and st["num_substeps"] == 4
    assert st["hydro"]["max_final"] == pytest.approx(1.25)
    cr.print_convergence_analysis(s, iterations=4, run_output_dir=str(tmp_path), export_csv=True)
    rows = list(csv.DictReader(open(tmp_path / "convergence_data.csv")))
    assert len(rows) == 16 and list(rows[0]) == cr.RECORD_FIELDS
    assert "CONSTRAINT CONVERGENCE ANALYSIS" in capsys.readouterr().out
    d = cr.setup_convergence_callback(s, str(tmp_path), frame_ids=[1], plot=False)
    s.convergence_callback(0, 0, s.convergence_history[:4])
    s.convergence_callback(1, 1, s.convergence_history[12:])
    assert sorted(os.listdir(d)) == ["f0001_s01.json"]
