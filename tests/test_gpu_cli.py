"""The command-line driver end to end on a small TetGen mesh."""
import csv
import json
import os

import pytest

import run_neohookean
from xpbd_gpu.mesh import box_grid, write_tetgen

pytestmark = pytest.mark.gpu


def test_cli_run_writes_the_reference_outputs(tmp_path, capsys):
    print("[case] cli run", flush=True)
    m = box_grid(3, origin=(-0.5, -1.0, -0.5))
    write_tetgen(str(tmp_path / "box"), m)
    model = str(tmp_path / "box.node")
    rc = run_neohookean.main(["--model", model, "--frames", "2", "--iterations", "10",
                              "--young", "1e6", "--poisson", "0.45", "--monitor-convergence",
                              "--export-csv", "--plot-frames", "--output-root", str(tmp_path)])
    assert rc == 0
    out = capsys.readouterr().out
    assert "PERFORMANCE" in out and "CONSTRAINT CONVERGENCE ANALYSIS" in out
    run = tmp_path / "output" / "box" / "xpbd_block_neohookean_dynamic_cheby_young_1p000e06"
    rows = list(csv.DictReader(open(run / "convergence_data.csv")))
    # dt 0.01 at 60 fps: two chunks of one substep per frame, 10 iterations each
    assert len(rows) == 2 * 2 * 10
    plots = sorted(os.listdir(run / "convergence_plots"))
    assert "f0000_s00.json" in plots and "f0000_s00.png" in plots
    doc = json.load(open(run / "convergence_plots" / "f0001_s00.json"))
    assert len(doc["hydro_errors"]) == 10 and doc["solver_type"] == "XPBDNeoHookeanSolver"
    # the animation: --use-binary (default) asks for .usdc, written as the .usda text stage
    usd = tmp_path / "output" / "box" / "xpbd_block_neohookean_dynamic_cheby_young_1p000e06.usda"
    text = usd.read_text()
    assert "points.timeSamples" in text and "\n            2: [" in text
    assert 'def Sphere "sdf_sphere_1"' in text and 'def Mesh "sdf_plane_0"' in text
