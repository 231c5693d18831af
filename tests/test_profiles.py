"""The committed counter profiles that bench.py's roofline reads: they must
describe the traversal sources as they are (bench.py keys them by a digest of
those sources and drops a stale one, leaving the roofline empty), carry the
fields the roofline uses, and name only commands that run committed files."""
import glob
import importlib.util
import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traversal_profiles_match_the_sources(bench):
    digest = bench.traversal_digest()
    for rel, fields in ((bench.VALU_PROFILE, ("valu_insts_per_ray", "lane_utilisation", "valu_issue_per_simd_cycle")),
                        (bench.SHADOW_VALU_PROFILE, ("valu_insts_per_ray", "lane_utilisation",
                                                     "valu_issue_per_simd_cycle")),
                        (bench.TRAFFIC_PROFILE, ("hbm_bytes_per_launch",))):
        with open(os.path.join(ROOT, rel)) as f:
            p = json.load(f)
        assert p["source_digest"] == digest, f"{rel} was measured on other traversal sources: profile again"
        for k in fields:
            assert k in p, (rel, k)
        assert bench._profile(rel) is not None


def test_scatter_profile_matches_the_cache_sources(bench):
    with open(os.path.join(ROOT, bench.SCATTER_PROFILE)) as f:
        p = json.load(f)
    assert p["source_digest"] == bench.cache_digest(), "the scatter was counted on other cache sources: count again"
    assert p["rows"] == 1 << 18 and p["atomic_requests_per_step"] > 0
    assert bench._profile(bench.SCATTER_PROFILE, bench.cache_digest) is not None


def test_profile_issue_rates_are_below_peak(bench):
    # a fraction above 1 would mean the peak (one wave64 VALU instruction per
    # SIMD every 2 clocks) is not the bound
    for rel in (bench.VALU_PROFILE, bench.SHADOW_VALU_PROFILE):
        with open(os.path.join(ROOT, rel)) as f:
            p = json.load(f)
        assert 0.0 < p["valu_issue_per_simd_cycle"] / 0.5 <= 1.0
        assert 0.0 < p["lane_utilisation"] <= 1.0


def test_profiles_name_committed_commands_only():
    files = glob.glob(os.path.join(ROOT, "profiles", "*"))
    assert files
    for path in files:
        with open(path) as f:
            text = f.read()
        assert "_work" not in text, path
        assert "rocprofv3" in text, f"{path} does not give the command that made it"
