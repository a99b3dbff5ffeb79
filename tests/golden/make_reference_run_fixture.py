"""Writes the pure-WoS statistics of the reference's committed evaluation runs.

The reference tree holds the output of its evaluation tool for the bumpy
sphere and the Bunny with the zebra k = 8 boundary (figures/zebra_k8_m2.json,
figures/bunny_zebra_k8_m2.json; RTX 3070).  Their "pure_wos" run is plain
Walk-on-Spheres: eps 1e-4, max_steps 256, 64 walks per interior slice pixel.
This script copies the run parameters and the aggregate numbers that a
statistical comparison needs into reference_runs_zebra_k8.json beside it:
  eval_points, walks_per_point, samples, epsilon, max_steps,
  rmse, mean_bias, mean_sample_variance, mean_steps (JSON keys of runs.pure_wos
  and options).
Run where the reference tree is present: python tests/golden/make_reference_run_fixture.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/figures"


def main():
    out = {}
    for name in ("zebra_k8_m2", "bunny_zebra_k8_m2"):
        doc = json.load(open(os.path.join(REF, name + ".json")))
        pure = doc["runs"]["pure_wos"]
        opt = doc["options"]
        out[name] = {
            "source": f"figures/{name}.json",
            "command_line": doc["command_line"],
            "epsilon": opt["epsilon"],
            "max_steps": opt["max_steps"],
            "slice_plane": doc["slice_eval"]["plane"],
            **{k: pure[k] for k in ("eval_points", "walks_per_point", "samples", "rmse", "mean_bias",
                                     "mean_sample_variance", "mean_steps", "forced_max_steps")},
        }
    with open(os.path.join(HERE, "reference_runs_zebra_k8.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote reference_runs_zebra_k8.json")


if __name__ == "__main__":
    main()
