"""CPU checks of host-side logic around the kernels: the evaluation tool's
cache presets (the reference's values, src/tools/eval_tcnn_nc_wos.cu:249-266,
and its preset / explicit-flag rule), the cache configuration's validation in
the C ABI (checked before any device call), the MLP weight count, and the
serving workload's seeded request trace."""
import ctypes as C

import numpy as np
import pytest


def test_eval_presets_follow_the_reference(eng):
    from n2wos_amd import eval_nc as EV
    shape = ("n_levels", "log2_hashmap_size", "base_resolution", "per_level_scale", "n_neurons", "n_hidden_layers")
    want = {"baseline": (12, 18, 8, 1.5, 32, 2), "paper_like": (12, 18, 8, 1.5, 32, 2),
            "light": (8, 15, 8, 1.5, 16, 1), "nano": (6, 13, 8, 1.5, 16, 1), "heavy": (16, 19, 8, 1.4, 64, 3)}
    for name, vals in want.items():
        o = EV.parse_args(["--cache-preset", name])
        assert tuple(getattr(o, k) for k in shape) == vals, name
        assert o.cache_preset == name
    o = EV.parse_args([])
    assert o.cache_preset == "baseline" and (o.n_neurons, o.n_hidden_layers) == (32, 2)
    # an explicit encoding or network flag overrides the preset and makes the run custom
    o = EV.parse_args(["--cache-preset", "heavy", "--n-levels", "8"])
    assert o.cache_preset == "custom" and o.n_levels == 8 and (o.n_neurons, o.n_hidden_layers) == (64, 3)
    o = EV.parse_args(["--n-neurons", "16"])
    assert o.cache_preset == "custom" and o.n_neurons == 16 and o.n_levels == 12
    for bad in (["--cache-preset", "xlarge"], ["--n-neurons", "48"], ["--n-hidden-layers", "5"],
                ["--cache-preset", "huge"], ["--n-levels", "17"]):
        with pytest.raises(SystemExit):
            EV.parse_args(bad)


def test_mlp_weight_count():
    from n2wos_amd import cache as K
    assert K.n_weights(64, 4) == K.N_WEIGHTS == 15360
    assert K.n_weights(16, 1) == 32 * 16 + 16 * 16
    assert K.n_weights(32, 2) == 32 * 32 + 32 * 32 + 16 * 32
    assert K.n_weights(64, 3) == 32 * 64 + 2 * 64 * 64 + 16 * 64


@pytest.mark.parametrize("field,value,word", [("n_neurons", 48, b"n_neurons"), ("n_hidden_layers", 5, b"n_hidden"),
                                              ("n_levels", 17, b"n_levels"), ("log2_table_size", 7, b"log2"),
                                              ("max_batch", (1 << 24) + 1, b"max_batch")])
def test_cache_config_rejected_before_any_device_call(eng, field, value, word):
    """n2wos_cache_create validates the configuration before it touches the
    device: these calls return N2WOS_ERR_INVALID on a machine without a GPU."""
    from n2wos_amd import cache as K
    cfg = K.CacheConfig(**{field: value}).to_c()
    h = C.c_void_p()
    rc = eng.lib().n2wos_cache_create(C.byref(cfg), C.byref(h))
    assert rc == eng._lib.ERR_INVALID
    assert word in eng.lib().n2wos_last_error()
    assert not h.value


def test_request_trace_is_seeded_and_bounded():
    from n2wos_amd import serve as S
    a1, s1, m1 = S.request_trace(500, 24.0, 3, seed=4)
    a2, s2, m2 = S.request_trace(500, 24.0, 3, seed=4)
    assert np.array_equal(a1, a2) and np.array_equal(s1, s2) and np.array_equal(m1, m2)
    assert a1[0] == 0.0 and np.all(np.diff(a1) >= 0)
    assert s1.min() >= 1000 and s1.max() <= 262144 and set(np.unique(m1)) <= {0, 1, 2}
    # mean gap of exponential arrivals at 24/s (500 samples: within 5 standard errors)
    gaps = np.diff(a1)
    assert abs(gaps.mean() - 1 / 24) < 5 * (1 / 24) / np.sqrt(len(gaps))
    # log-uniform sizes: the median of log size sits mid-range
    mid = 0.5 * (np.log(1000) + np.log(262144))
    assert abs(np.median(np.log(s1)) - mid) < 0.5
