"""The experiment drivers (controllable_pbd_3d_amd/experiments.py) at tiny sizes: each writes
the summary list that the reference's plot helpers read (experiments/plot_utils.py:
plot_resolution_scaling needs solver_name, n_tets, avg_frame_time, fps;
plot_stiffness_scaling solver_name, mu, avg_iterations, avg_frame_time; plot_timing_bars
solver_name, avg_frame_time) and one SolverBenchmark document per run. As in the reference's
drivers, the benchmark's "frames" are the solver's solve() calls: one per substep, so one
animation frame of 5 substeps records 5 and fps is 1 / avg_frame_time."""

import glob
import json
import os

import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

KEYS = {"resolution": {"solver_name", "n_tets", "avg_frame_time", "fps"},
        "stiffness": {"solver_name", "mu", "avg_iterations", "avg_frame_time"},
        "ablation": {"solver_name", "avg_frame_time", "speedup"}}


@pytest.mark.parametrize("experiment", ["resolution", "stiffness", "ablation"])
def test_experiment_writes_plot_records(tmp_path, experiment):
    from controllable_pbd_3d_amd import experiments
    from controllable_pbd_3d_amd.benchmark import SolverBenchmark
    mesh_dir = os.path.join(os.path.dirname(__file__), "golden", "meshes")
    args = [experiment, "--out", str(tmp_path), "--frames", "1", "--sizes", "3",
            "--mus", "1e4", "1e6", "--mesh-dir", mesh_dir]
    print(f"[experiments] {experiment}", flush=True)
    experiments.main(args)
    recs = json.load(open(tmp_path / f"{experiment}.json"))
    n_var = len(experiments.VARIANTS)
    expect = {"resolution": n_var * 4, "stiffness": n_var * 2, "ablation": 5}[experiment]
    assert len(recs) == expect
    for r in recs:
        assert KEYS[experiment] <= set(r)
        assert r["avg_frame_time"] > 0 and r["n_frames"] == experiments.SUBSTEPS
    if experiment == "resolution":   # the reference's three Medit meshes ran too
        assert {r["mesh_name"] for r in recs} >= {"simple_sphere", "spot", "cube_tet"}
    runs = sorted(glob.glob(str(tmp_path / "runs" / "*.json")))
    assert len(runs) == expect
    b = SolverBenchmark.load(runs[0])
    assert b.summary()["n_frames"] == experiments.SUBSTEPS
