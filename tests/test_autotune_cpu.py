"""Autotuner host logic: the options parse, the cache file round-trips."""
import json

from conftest import CASES


def test_autotune_options_and_cache_file(pkg, tmp_path):
    from pbte_amd import autotune
    c = pkg.Case.from_yaml(CASES / "F1_1_1.yaml")
    assert c.performance.autotune is False and c.performance.autotune_cache == ""
    # the options change how the device sweeps, never the discrete problem
    h = c.config_hash
    c.performance.autotune = True
    assert c.config_hash == h
    path = str(tmp_path / "tiles.json")
    autotune._store(path, "k1", {"max_bundle": 2, "block": 0, "variant": 0, "ms": 1.5})
    autotune._store(path, "k2", {"max_bundle": 1, "block": 128, "variant": 2, "ms": 0.5})
    table = json.loads((tmp_path / "tiles.json").read_text())
    assert table["k1"]["max_bundle"] == 2 and table["k2"]["block"] == 128
    assert autotune._load(str(tmp_path / "missing.json")) == {}
    assert len(set(autotune.CANDIDATES)) == len(autotune.CANDIDATES) == 8
