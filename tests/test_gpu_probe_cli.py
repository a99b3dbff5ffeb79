"""The probe CLI on the GPU: benchmark problem 1 (UV sphere 64 x 128,
r = 1 + 0.1 sin 6t sin 6p, normalised; 65,536 queries uniform in
[-1.2, 1.2]^3) exits 0, writes the n2wos_cuda_probe_bvh_v1 keys, and every
validated query equals the host brute force bit for bit."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

KEYS = {"schema", "patch", "generated_at_utc", "command_line", "cuda", "implementation_mode", "options", "mesh_stats",
        "bvh_stats", "timing", "validation"}
MODE_KEYS = {"gpu_resident_geometry_query", "geometry_query", "fcpw_gpu_public_api", "host_driven_walker_loop",
             "cache_inference", "rng", "accumulation", "timing_scope"}


def test_probe_cli_problem1(eng, torch_gpu, tmp_path):
    exe = os.path.join(os.path.dirname(eng.LIB_PATH), "..", "bin", "n2wos_probe_cuda_bvh")
    out = str(tmp_path / "probe.json")
    args = [exe, "--mesh", "sin6", "--bumpy-stacks", "64", "--bumpy-slices", "128", "--bumpy-amplitude", "0.1",
            "--queries", "65536", "--query-extent", "1.2", "--seed", "0", "--validate", "4096", "--repeat", "3",
            "--output", out]
    print("case: probe CLI problem 1", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    j = json.load(open(out))
    assert KEYS <= set(j) and MODE_KEYS <= set(j["implementation_mode"])
    assert j["schema"] == "n2wos_cuda_probe_bvh_v1"
    assert j["implementation_mode"]["geometry_query"] == "hip_bvh_gfx950"
    assert j["mesh_stats"]["triangles"] == 16128 and j["mesh_stats"]["vertices"] == 8066
    v = j["validation"]
    assert v["checked"] == 4096 and v["bad_distance_count_threshold_2e_4"] == 0
    assert v["bit_exact_matches"] == 4096 and v["triangle_id_mismatches"] == 0
    assert v["cuda_stack_overflow_count"] == 0
    assert len(j["timing"]["kernel_ms"]) == 3 and j["timing"]["median_mqueries_per_second"] > 0


@pytest.mark.parametrize("mode", ["wos_like_prefix", "near_boundary_shell", "uniform_box", "interior_slice",
                                  "surface_exact"])
def test_geometry_backend_sweep(eng, torch_gpu, tmp_path, mode):
    """The builder / leaf-size sweep probe: every backend (median and SAH
    builders x two leaf sizes, both query kernels) validates bit for bit
    against the host brute force, the v2 JSON carries the reference's keys and
    backend_policy labels, and the recommendation names the fastest passing
    backend."""
    exe = os.path.join(os.path.dirname(eng.LIB_PATH), "..", "bin", "n2wos_probe_geometry_backends")
    out = str(tmp_path / "sweep.json")
    args = [exe, "--mesh", "sin6", "--bumpy-stacks", "64", "--bumpy-slices", "128", "--bumpy-amplitude", "0.1",
            "--queries", "32768", "--validate", "1024", "--repeat", "2", "--query-mode", mode,
            "--cubql-build-methods", "spatial_median,sah", "--leaf-sizes", "2,8", "--query-kernels", "quad,lane",
            "--output", out]
    print("case: geometry backend sweep %s" % mode, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    j = json.load(open(out))
    assert j["schema"] == "n2wos_geometry_backend_probe_v2"
    for k in ("backend_policy", "options", "mesh_stats", "build_features", "recommendation", "backends"):
        assert k in j, k
    assert j["backend_policy"]["production_requires_gpu_resident_geometry_query"] is True
    b = {k: v for k, v in j["backends"].items() if v.get("built")}
    assert len(b) == 8 and not j["backends"]["cubql_cuda"]["enabled"]
    for name, v in b.items():
        val = v["validation"]
        assert v["validation_passed"] and v["overflow_count"] == 0, name
        assert val["checked"] == 1024 and val["bit_exact_matches"] == 1024, name
        assert val["max_backend_distance_delta"] == 0.0, name
    best = min(b, key=lambda k: b[k]["timing"]["median_us_per_query"])
    assert j["recommendation"]["selected_production_candidate"] == best


def test_cache_probe_reference_defaults(eng, torch_gpu, tmp_path):
    """The learned-cache probe at the reference probe's default workload
    (262,144 samples and batch, 200 Adam steps, 12 levels, 2^18 table, base 8,
    scale 1.5, 32 x 2 MLP, x^2 - y^2 on the unit cube): exits 0, writes the
    n2wos_tcnn_cache_probe_v1 keys, and the trained cache fits the target
    (RMSE < 0.05 over 20 repeats; mean of x^2 - y^2 over the cube is 0)."""
    exe = os.path.join(os.path.dirname(eng.LIB_PATH), "..", "bin", "n2wos_probe_tcnn_cache")
    out = str(tmp_path / "cache_probe.json")
    print("case: cache probe defaults", flush=True)
    r = subprocess.run([exe, "--output", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    j = json.load(open(out))
    assert j["schema"] == "n2wos_tcnn_cache_probe_v1"
    for k in ("patch", "generated_at_utc", "command_line", "cuda", "implementation_mode", "options", "training",
              "inference", "inference_plus_downstream_consumer", "validation"):
        assert k in j, k
    o = j["options"]
    assert (o["samples"], o["batch_size"], o["train_steps"], o["repeat"]) == (262144, 262144, 200, 20)
    assert (o["n_levels"], o["log2_hashmap_size"], o["n_neurons"], o["n_hidden_layers"]) == (12, 18, 32, 2)
    assert j["implementation_mode"]["cache_backend"] == "hip_hashgrid_mlp_mfma_gfx950"
    assert len(j["inference"]["kernel_ms"]) == 20 and j["inference"]["median_msamples_per_second"] > 0
    v = j["validation"]
    assert v["checked_repeated_samples"] == 262144 * 20
    assert abs(v["mean_exact"]) < 0.01
    assert abs(v["mean_prediction"] - v["mean_exact"]) < 0.01
    assert v["rmse"] < 0.05 and v["mae"] <= v["rmse"]
    print("cache probe: rmse %.4f, train %.3f ms/step, inference %.3f us/sample" %
          (v["rmse"], j["training"]["ms_per_step"], j["inference"]["median_us_per_sample"]), flush=True)


@pytest.mark.parametrize("shape", [("6", "13", "16", "1"), ("16", "19", "64", "3")])
def test_cache_probe_preset_shapes(eng, torch_gpu, tmp_path, shape):
    """The probe at the reference's nano (6 levels, 2^13, 16 x 1) and heavy
    (16 levels, 2^19, 64 x 3) shapes on a smaller workload: the JSON reports
    the shape that ran and the cache still fits x^2 - y^2."""
    exe = os.path.join(os.path.dirname(eng.LIB_PATH), "..", "bin", "n2wos_probe_tcnn_cache")
    out = str(tmp_path / "cache_probe.json")
    lv, t2, h, nl = shape
    args = [exe, "--samples", "65536", "--batch-size", "65536", "--train-steps", "150", "--repeat", "3",
            "--n-levels", lv, "--log2-hashmap-size", t2, "--n-neurons", h, "--n-hidden-layers", nl, "--output", out]
    print("case: cache probe shape %s" % (shape,), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    j = json.load(open(out))
    o = j["options"]
    assert (o["n_levels"], o["log2_hashmap_size"], o["n_neurons"], o["n_hidden_layers"]) == tuple(int(x) for x in shape)
    assert o["network_backend"] == "hashgrid_relu_mlp_%sx%s_fp16_mfma" % (nl, h)
    assert j["validation"]["rmse"] < 0.05
    bad = subprocess.run([exe, "--n-neurons", "48", "--output", out], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "n_neurons" in bad.stderr
