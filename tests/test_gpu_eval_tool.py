"""The evaluation tool end to end on a small problem: the v4 JSON keys the
reference's analysis scripts read, the estimates CSV columns, the slice mask
and points files, and sane numbers (NC and 2LMC near the FP64 solution)."""
import csv
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "n2wos-cuda_amd", "tools", "n2wos_eval_nc_wos.py")


def test_eval_tool_slice_outputs(tmp_path, torch_gpu):
    out = tmp_path / "eval.json"
    pre = str(tmp_path / "est")
    args = [sys.executable, TOOL, "--mesh", "mesh1", "--bc", "harmonic_x2_minus_y2", "--eval-mode", "slice",
            "--slice-width", "48", "--slice-height", "40", "--train-points", "4096", "--label-refreshes", "2",
            "--walks-per-label-refresh", "16", "--train-steps-per-refresh", "60", "--n-levels", "4",
            "--pure-walks-per-point", "64", "--hybrid-walks-per-point", "32", "--coarse-walks-per-point", "32",
            "--residual-walks-per-point", "16", "--depth-m", "1", "--epsilon", "1e-3", "--output", str(out),
            "--save-estimates-prefix", pre, "--save-train-points-prefix", pre]
    print("case: eval tool slice 48x40 mesh1", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    doc = json.loads(out.read_text())
    assert doc["schema"] == "n2wos_tcnn_nc_wos_eval_v4"
    for k in ("implementation_mode", "options", "slice_eval", "training_point_sampler", "mesh_stats", "bvh_stats",
              "training", "runs", "comparison", "estimate_outputs"):
        assert k in doc, k
    assert doc["implementation_mode"]["cache_backend"] == "hip_hashgrid_mlp_mfma_gfx950"
    runs = doc["runs"]
    for k in ("pure_wos", "nc_wos", "nc_2lmc", "nc_2lmc_m1"):
        assert k in runs, k
    for k in ("rmse", "mae", "max_abs_error", "mean_bias", "mean_sample_variance", "mean_steps", "elapsed_ms",
              "training_plus_elapsed_ms", "us_per_sample", "cache_query_fraction"):
        assert k in runs["pure_wos"] and k in runs["nc_wos"], k
    n = doc["slice_eval"]["inside_pixels"]
    assert 0 < n < 48 * 40 and runs["pure_wos"]["eval_points"] == n
    assert runs["pure_wos"]["rmse"] < 0.05 and runs["nc_2lmc"]["rmse"] < 0.1
    assert runs["nc_wos"]["cache_query_fraction"] > 0.9
    with open(pre + "_estimates.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["point_id", "x", "y", "z", "analytic_value", "pure_mean", "pure_sample_variance",
                       "nc_wos_mean", "nc_wos_sample_variance", "nc_2lmc_mean", "nc_2lmc_coarse_mean",
                       "nc_2lmc_residual_mean", "nc_2lmc_coarse_sample_variance", "nc_2lmc_residual_sample_variance"]
    assert len(rows) == n + 1
    a = np.array(rows[1:], np.float64)
    assert np.allclose(a[:, 4], a[:, 1] ** 2 - a[:, 2] ** 2, atol=1e-6)
    assert np.allclose(a[:, 9], a[:, 10] + a[:, 11], atol=1e-6)
    mask = (tmp_path / "eval_slice_mask.ppm").read_bytes()
    assert mask.startswith(b"P6\n48 40\n255\n") and len(mask) == len(b"P6\n48 40\n255\n") + 48 * 40 * 3
    with open(tmp_path / "eval_slice_points.csv") as f:
        assert len(f.read().strip().splitlines()) == n + 1
    with open(pre + "_train_points.csv") as f:
        assert len(f.read().strip().splitlines()) == 4096 + 1


def test_eval_tool_nano_preset_network(tmp_path, torch_gpu):
    """--cache-preset nano runs the reference preset's own network (6 levels,
    2^13 table, base 8, scale 1.5, 16 x 1 MLP) and the JSON says so; NC and
    2LMC estimates stay near the FP64 solution."""
    out = tmp_path / "nano.json"
    args = [sys.executable, TOOL, "--mesh", "mesh1", "--bc", "harmonic_x2_minus_y2", "--eval-mode", "ball",
            "--eval-points", "2048", "--train-points", "4096", "--label-refreshes", "2",
            "--walks-per-label-refresh", "16", "--train-steps-per-refresh", "100", "--cache-preset", "nano",
            "--pure-walks-per-point", "64", "--hybrid-walks-per-point", "32", "--coarse-walks-per-point", "32",
            "--residual-walks-per-point", "16", "--depth-m", "1", "--epsilon", "1e-3", "--output", str(out)]
    print("case: eval tool nano preset", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    o = json.loads(out.read_text())["options"]
    assert o["cache_preset"] == "nano"
    assert (o["n_levels"], o["log2_hashmap_size"], o["base_resolution"], o["n_neurons"], o["n_hidden_layers"]) == \
        (6, 13, 8, 16, 1)
    assert o["network"] == "hashgrid_relu_mlp_1x16_fp16_mfma"
    runs = json.loads(out.read_text())["runs"]
    assert runs["pure_wos"]["rmse"] < 0.05 and runs["nc_2lmc"]["rmse"] < 0.1


def test_slice_mask_equals_reference_run(eng, torch_gpu):
    """The slice sampler on the reference's own evaluation setting (bumpy
    sphere 128 x 256, amplitude 0.15, normalised, xy slice at z = 0, 512 x 512,
    mesh-bounds frame with 2% padding) gives the interior mask that the
    reference's tool wrote for that run (figures/zebra_k8_m2_slice_mask.ppm),
    bit for bit: frame, pixel positions and inside test all agree."""
    from n2wos_amd import eval_nc as EV, geometry as G
    g = np.load(os.path.join(ROOT, "tests", "golden", "slice_mask_zebra_k8_m2.npz"))
    w, h = int(g["width"]), int(g["height"])
    want = np.unpackbits(g["bits"])[: w * h].reshape(h, w).astype(bool)
    opt = EV.parse_args(["--mesh", "procedural_bumpy_sphere", "--eval-mode", "slice", "--slice-view", "xy",
                         "--slice-plane", "0.0", "--slice-width", str(w), "--slice-height", str(h)])
    v, t, _, _, _ = EV.make_mesh(opt)
    bvh = G.Bvh(v, t, leaf_size=opt.cubql_leaf_size, build_method=opt.cubql_build_method)
    print("case: reference slice mask 512x512", flush=True)
    pts, mask, _ = EV.slice_points(opt, bvh, v.min(0), v.max(0))
    got = mask.reshape(h, w)[::-1].astype(bool)  # image rows: v descending
    assert int(got.sum()) == int(g["inside_pixels"]) == len(pts)
    assert np.array_equal(got, want)


WTOOL = os.path.join(ROOT, "n2wos-cuda_amd", "tools", "n2wos_eval_wavefront_wos.py")


@pytest.mark.parametrize("method", ["pure_wos", "oracle_2lmc", "oracle_residual"])
def test_eval_wavefront_methods(tmp_path, torch_gpu, method):
    """The wavefront evaluation tool at x0 on mesh 1, boundary x^2 - y^2:
    the v1 JSON keys, every estimate within 4.5 standard errors of the exact
    value (the residual's of zero), and the analytic-oracle coarse level far
    less noisy than plain walks."""
    out = tmp_path / "wf.json"
    args = [sys.executable, WTOOL, "--mesh", "mesh1", "--method", method, "--samples", "16384",
            "--residual-samples", "4096", "--depth-m", "4", "--epsilon", "1e-3", "--x0", "0.2,-0.1,0.05",
            "--output", str(out)]
    print("case: eval wavefront %s" % method, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    doc = json.loads(out.read_text())
    assert doc["schema"] == "n2wos_wavefront_wos_eval_v1"
    for k in ("implementation_mode", "options", "mesh_stats", "bvh_stats", "target", "runs", "estimator"):
        assert k in doc, k
    assert doc["implementation_mode"]["cache_backend"] == "analytic_oracle_debug_only"
    est = doc["estimator"]
    exact = 0.2 ** 2 - 0.1 ** 2
    if method == "oracle_residual":
        run = doc["runs"]["oracle_residual"]
        assert abs(run["mean"]) < 4.5 * run["stderr"] + 1e-12
    else:
        assert abs(est["exact"] - exact) < 1e-12
        assert abs(est["error"]) < 4.5 * est["stderr"], est
    if method == "oracle_2lmc":
        coarse, resid = doc["runs"]["coarse"], doc["runs"]["residual"]
        assert coarse["samples"] == 16384 and resid["samples"] == 4096
        assert coarse["mean_steps"] <= 4.0
        assert abs(est["estimator_variance"] - coarse["estimator_variance"] - resid["estimator_variance"]) < 1e-15
