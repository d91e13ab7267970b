"""bench.py's host-side assembly on the CPU: the roofline of each fused kernel from a scene's
statistics in both byte bases, and that every helper the JSON line needs exists (no GPU)."""
import importlib
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, REPO)
    return importlib.import_module("bench")


def test_helpers_exist(bench):
    for name in ("algorithmic_bytes", "reference_layout_bytes", "roofline", "scene_stats", "cpu_baseline",
                 "config_run", "learned_measurement", "run_batch", "run_learned", "load_traffic", "main"):
        assert callable(getattr(bench, name)), name


def test_roofline_both_bases(bench, built):
    from morphogenesis_amd import scenes

    sc = scenes.tissue(n_side=8, b=32, seed=1)
    st = bench.scene_stats(sc, neighbours=False)
    st["n_active"] = sc.n
    for phase in ("pressure", "force", "diffusion"):
        r = bench.roofline(phase, sc, st, 0.05, "tissue_8")
        ref, terms = bench.reference_layout_bytes(phase, sc, st)
        assert r["algorithmic_bytes_per_launch"] == ref > 0
        assert abs(r["achieved"] - ref / 0.05e-3 / 1e9) < 1e-9 * r["achieved"]
        assert r["frac"] == r["achieved"] / r["peak"]
        assert r["layout"]["bytes_per_launch"] == bench.algorithmic_bytes(phase, sc, st) > 0
        assert all(k in r["algorithmic_basis"] for k in terms)
    # the diffusion kernel also carries genes and the insert; the force kernel the advance
    assert set(bench.reference_layout_bytes("diffusion", sc, st)[1]) == {"diffusion", "insert", "genes"}
    assert set(bench.reference_layout_bytes("force", sc, st)[1]) == {"force", "advance"}
