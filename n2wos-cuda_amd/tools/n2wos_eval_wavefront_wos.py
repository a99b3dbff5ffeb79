#!/usr/bin/env python3
"""Command-line entry of the wavefront evaluation tool (the counterpart of
the reference's n2wos_eval_wavefront_wos, src/tools/eval_wavefront_wos.cu):
plain WoS and the analytic-oracle two-level estimator at one point, with the
same flags and JSON.  See the package module eval_wavefront for the options."""
import importlib.util
import os
import sys

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _package():
    if "n2wos_amd" not in sys.modules:
        spec = importlib.util.spec_from_file_location("n2wos_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                      submodule_search_locations=[PKG_DIR])
        mod = importlib.util.module_from_spec(spec)
        sys.modules["n2wos_amd"] = mod
        spec.loader.exec_module(mod)
    return sys.modules["n2wos_amd"]


if __name__ == "__main__":
    _package()
    from n2wos_amd import eval_wavefront
    try:
        sys.exit(eval_wavefront.main(sys.argv[1:]))
    except SystemExit:
        raise
    except Exception as e:  # the reference prints "error: ..." and exits 1
        print(f"error: {e}", file=sys.stderr)
        sys.exit(1)
