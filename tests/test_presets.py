"""The reference's 14 new-file presets (tests/golden/presets, transcribed from
src/editor/presets/*.ts by tests/golden/make_preset_fixtures.py) through the scene loader."""
import json
import os

import numpy as np
import pytest

from fenix_amd.emitters import pack_emitter_sources
from fenix_amd.scene import graph_from_preset, scene_from_graph

HERE = os.path.dirname(os.path.abspath(__file__))
PRESETS = os.path.join(HERE, "golden", "presets")


def preset_ids():
    with open(os.path.join(PRESETS, "index.json")) as fh:
        return json.load(fh)["presets"]


def load_preset(pid):
    with open(os.path.join(PRESETS, pid + ".json")) as fh:
        return json.load(fh)


def test_fixture_set_matches_the_editor_menu():
    """newFilePresets.ts:20-35 lists 14 presets, blank first; every fixture file is listed."""
    ids = preset_ids()
    assert len(ids) == 14 and ids[0] == "blank" and len(set(ids)) == 14
    files = {f[:-5] for f in os.listdir(PRESETS) if f.endswith(".json") and f != "index.json"}
    assert files == set(ids)


@pytest.mark.parametrize("pid", preset_ids())
def test_preset_loads_into_sources_runtime_and_lights(pid):
    p = load_preset(pid)
    sc = scene_from_graph(graph_from_preset(p))
    kinds = [e["props"]["kind"] for e in p["emitters"]]
    assert set(kinds) <= {"scalar", "velocity", "burst", "igniter"}
    # every emitter reaches the combustion node unless the preset's graph edges leave it out
    if not p.get("graphEdges"):
        assert len(sc.sources) == len(p["emitters"])
    else:
        assert 0 < len(sc.sources) <= len(p["emitters"])
    assert sc.runtime_params["worldSize"] == p["runtimeParams"]["worldSize"]
    assert len(sc.lights) == len(p["lights"]) >= 1
    if sc.sources:
        packed = pack_emitter_sources(sc.sources)
        assert packed.shape == (32 * len(sc.sources),) and np.all(np.isfinite(packed))


def test_burst_presets_carry_burst_sources():
    """The eruptive (burst) emitters of the explosion presets reach the packed sources."""
    with_burst = [pid for pid in preset_ids() if any(e["props"]["kind"] == "burst" for e in load_preset(pid)["emitters"])]
    assert {"artillery-burst", "artillery-salvo", "fuel-tank-explosion", "nuclear-detonation",
            "volcanic-eruption", "demolition-dust"} <= set(with_burst)
