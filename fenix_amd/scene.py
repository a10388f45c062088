"""Scene loading: the reference's graph files (version 2 and 3) and new-file
presets, turned into what the native step consumes (emitter sources, runtime
parameters, quality settings, advection mode) and what the raymarch consumes
(lights, step count, scattering, display mode).

References:
* file format and validation: src/editor/serialization/graphSerializer.ts:18-45
  (SerializedGraph), :101-175 (parseGraphJson: version checks, defaults for
  older files, quality clamps);
* applying a file: src/ui/components/toolbar/GraphFilePopover.tsx:164-194
  (applySerializedGraph) and src/store/node-store/nodeStore.ts:119-222 (node <->
  runtime parameter mapping);
* graph -> sources: src/store/node-store/sourceGraph.ts:97-132
  (resolveEmitterSources) and nodeStore.ts:49-109 (emitterPropsToSource);
* default edges when a file has no graph, and preset edges:
  src/store/node-store/nodeGraphStore.ts:32-38, 81-118;
* lights reaching the renderer: src/features/viewport/useRenderOutputBridge.ts:8-37.

The reference's editor-only state (node positions, bloom) is read and kept but
not used: the bloom pass is outside this library.
"""
import copy
import json
import math
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

from .emitters import BurstEmitter, EmitterSource, IgniterEmitter, ScalarEmitter, VelocityEmitter
from .render import RenderLight

GRAPH_FILE_VERSION = 3

# SimulationQualitySettings defaults and clamps (combustion-volume-simulation/types.ts:16-44)
_QUALITY_DEFAULTS = {"pressureInterval": 1, "vorticityInterval": 1, "finePreIterations": 2, "finePostIterations": 3,
                     "midPreIterations": 2, "midPostIterations": 3, "coarseIterations": 10}
_QUALITY_LIMITS = {"pressureInterval": (1, 16), "vorticityInterval": (1, 16), "finePreIterations": (0, 32),
                   "finePostIterations": (0, 32), "midPreIterations": (0, 32), "midPostIterations": (0, 32),
                   "coarseIterations": (0, 64)}

# nodeGraphStore.ts:32-38
FIXED_EDGES = (
    {"id": "comb-adv", "source": "combustion", "target": "advection"},
    {"id": "wind-adv", "source": "wind", "target": "advection"},
    {"id": "vorticity-adv", "source": "vorticity", "target": "advection"},
    {"id": "gravity-adv", "source": "gravity", "target": "advection"},
    {"id": "adv-ro", "source": "advection", "target": "render-output"},
)


class GraphFileError(ValueError):
    """A graph file the reference's loader rejects (same messages)."""


def _js_round(v: float) -> int:
    return int(math.floor(v + 0.5))


def _js_number(v) -> float:
    """Number(v) for the JSON values a file can hold (null -> 0, booleans -> 0/1, numeric strings)."""
    if v is None:
        return 0.0
    if isinstance(v, bool):
        return float(v)
    if isinstance(v, (int, float)):
        return float(v)
    if isinstance(v, str):
        try:
            return float(v.strip()) if v.strip() else 0.0
        except ValueError:
            return math.nan
    return math.nan


def _js_str(v) -> str:
    """String(v) as a template literal shows a JSON value."""
    if v is _MISSING:
        return "undefined"
    if v is None:
        return "null"
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, float) and v.is_integer():
        return str(int(v))
    if isinstance(v, (int, float, str)):
        return str(v)
    if isinstance(v, list):
        return ",".join("" if x is None else _js_str(x) for x in v)
    return "[object Object]"


_MISSING = object()


def create_quality_settings(overrides: Optional[Dict] = None) -> Dict[str, int]:
    """createSimulationQualitySettings: defaults merged with overrides, rounded (Math.round)
    and clamped (Math.max(lo, Math.min(hi, .))); a non-numeric value clamps like NaN does
    there, to the lower bound."""
    merged = dict(_QUALITY_DEFAULTS)
    merged.update(overrides or {})
    out = {}
    for k, (lo, hi) in _QUALITY_LIMITS.items():
        x = _js_number(merged[k])
        out[k] = lo if math.isnan(x) else max(lo, min(hi, _js_round(x)))
    return out


def parse_graph_json(raw: str) -> Dict:
    """parseGraphJson: validates a graph file and fills the fields older files lack."""
    try:
        obj = json.loads(raw)
    except (TypeError, ValueError):
        raise GraphFileError("File is not valid JSON.")
    if not isinstance(obj, dict):
        raise GraphFileError("Invalid graph file.")
    version = obj.get("version", _MISSING)
    numeric = isinstance(version, (int, float)) and not isinstance(version, bool)
    if numeric and version == 1:
        raise GraphFileError("This file was saved with an older version of Fenix and is not compatible with the new "
                             "emitter model. Please recreate the effect using the new emitter types.")
    if not (numeric and version in (GRAPH_FILE_VERSION, 2)):
        raise GraphFileError(f"Unsupported graph version: {_js_str(version)}. Expected {GRAPH_FILE_VERSION}.")
    if not isinstance(obj.get("emitters"), list):
        raise GraphFileError("Missing emitters array.")
    if not isinstance(obj.get("lights"), list):
        raise GraphFileError("Missing lights array.")
    for key, message in (("runtimeParams", "Missing runtimeParams."), ("combustion", "Missing combustion node."),
                         ("advection", "Missing advection node."), ("renderOutput", "Missing renderOutput node.")):
        if not isinstance(obj.get(key), dict):
            raise GraphFileError(message)
    g = copy.deepcopy(obj)
    rp = g["runtimeParams"]
    if not isinstance(rp.get("worldSize"), (int, float)) or isinstance(rp.get("worldSize"), bool):
        rp["worldSize"] = 10
    if not isinstance(rp.get("gravity"), list):
        rp["gravity"] = [0, -1, 0]
    if not isinstance(rp.get("gravityStrength"), (int, float)) or isinstance(rp.get("gravityStrength"), bool):
        rp["gravityStrength"] = 0.45
    if not g.get("wind"):
        g["wind"] = {"directionX": rp["wind"][0], "directionY": rp["wind"][1], "directionZ": rp["wind"][2],
                     "strength": rp.get("windStrength")}
    if not g.get("gravity"):
        g["gravity"] = {"directionX": rp["gravity"][0], "directionY": rp["gravity"][1], "directionZ": rp["gravity"][2],
                        "strength": rp["gravityStrength"], "buoyancy": rp.get("buoyancy")}
    if not g.get("vorticity"):
        g["vorticity"] = {"strength": rp.get("vorticityStrength")}
    vort = g["vorticity"]
    for key, rp_key, default in (("constantMask", "vorticityConstantMask", 1), ("velocityMask", "vorticityVelocityMask", 0.4),
                                 ("heatMask", "vorticityHeatMask", 0.8), ("densityMask", "vorticityDensityMask", 0.35)):
        if vort.get(key) is None:
            vort[key] = rp.get(rp_key) if rp.get(rp_key) is not None else default
    g["simulationQuality"] = create_quality_settings(g.get("simulationQuality") or {})
    return g


def emitter_props_to_source(props: Dict) -> EmitterSource:
    """emitterPropsToSource (nodeStore.ts:49-109)."""
    pos = (props["positionX"], props["positionY"], props["positionZ"])
    kind = props.get("kind")
    if kind == "scalar":
        return ScalarEmitter(position=pos, radius=props["radius"], startTime=props["startTime"], duration=props["duration"],
                             densityRate=props["densityRate"], heatRate=props["heatRate"], fuelRate=props["fuelRate"],
                             noiseScale=props["noiseScale"], noiseMix=props["noiseMix"], seed=props["seed"])
    if kind == "velocity":
        t = props.get("tightness")
        return VelocityEmitter(position=pos, radius=props["radius"], startTime=props["startTime"],
                               duration=props["duration"], mode=props["mode"], speed=props["speed"],
                               direction=(props["directionX"], props["directionY"], props["directionZ"]),
                               falloff=props["falloff"], tightness=0 if t is None else t, seed=props["seed"])
    if kind == "burst":
        return BurstEmitter(position=pos, radius=props["radius"], startTime=props["startTime"], duration=props["duration"],
                            densityAmount=props["densityAmount"], heatAmount=props["heatAmount"],
                            fuelAmount=props["fuelAmount"], reactionAmount=props["reactionAmount"],
                            expansionSpeed=props["expansionSpeed"], liftSpeed=props["liftSpeed"],
                            turbulenceSpeed=props["turbulenceSpeed"], falloff=props["falloff"],
                            noiseScale=props["noiseScale"], noiseMix=props["noiseMix"], seed=props["seed"])
    return IgniterEmitter(position=pos, radius=props["radius"], startTime=props["startTime"], duration=props["duration"],
                          intensity=props["intensity"], seed=props["seed"])


def _is_material(emitter: Dict) -> bool:
    return emitter["props"].get("kind") in ("scalar", "burst")


def resolve_emitter_sources(emitters: Sequence[Dict], edges: Sequence[Dict]) -> List[EmitterSource]:
    """resolveEmitterSources (sourceGraph.ts:97-132): emitters with a directed path to
    the combustion node, in emitter order; a velocity or igniter emitter linked (in
    either direction, through emitter-emitter edges) to a material (scalar or burst)
    emitter takes the nearest one's position, and an igniter the larger radius."""
    outgoing: Dict[str, List[str]] = {}
    for e in edges:
        outgoing.setdefault(e["source"], []).append(e["target"])
    links: Dict[str, List[str]] = {}
    for e in edges:
        if e["source"].startswith("emitter-") and e["target"].startswith("emitter-"):
            links.setdefault(e["source"], []).append(e["target"])
            links.setdefault(e["target"], []).append(e["source"])
    by_id = {em["id"]: em for em in emitters}

    def reaches_combustion(start: str) -> bool:
        seen, stack = set(), [start]
        while stack:
            cur = stack.pop()
            if cur in seen:
                continue
            if cur == "combustion":
                return True
            seen.add(cur)
            stack.extend(outgoing.get(cur, []))
        return False

    def nearest_material(start: str) -> Optional[Dict]:
        first = by_id.get(start)
        if first is None or _is_material(first):
            return None
        seen, queue = set(), deque([start])
        while queue:
            cur = queue.popleft()
            if cur in seen:
                continue
            seen.add(cur)
            em = by_id.get(cur)
            if em is not None and _is_material(em):
                return em
            queue.extend(links.get(cur, []))
        return None

    out = []
    for em in emitters:
        if not reaches_combustion(em["id"]):
            continue
        src = emitter_props_to_source(em["props"])
        anchor = nearest_material(em["id"])
        if anchor is not None:
            a = emitter_props_to_source(anchor["props"])
            src.position = tuple(a.position)
            if src.kind == "igniter":
                src.radius = max(src.radius, a.radius)
        out.append(src)
    return out


def default_graph_edges(emitter_count: int, light_count: int, preset_edges: Optional[Sequence[Dict]] = None) -> List[Dict]:
    """createGraphState's edges (nodeGraphStore.ts:81-118): preset edges (emitter indices,
    target an index or 'combustion') or every emitter -> combustion; every light ->
    render-output; the fixed node edges."""
    if preset_edges is not None:
        em_edges = []
        for e in preset_edges:
            src = f"emitter-{e['source']}"
            tgt = "combustion" if e["target"] == "combustion" else f"emitter-{e['target']}"
            em_edges.append({"id": f"{src}->{tgt}", "source": src, "target": tgt})
    else:
        em_edges = [{"id": f"emitter-{i}->combustion", "source": f"emitter-{i}", "target": "combustion"}
                    for i in range(emitter_count)]
    light_edges = [{"id": f"light-{i}->render-output", "source": f"light-{i}", "target": "render-output"}
                   for i in range(light_count)]
    return em_edges + light_edges + [dict(e) for e in FIXED_EDGES]


def runtime_params_from_nodes(graph: Dict) -> Dict:
    """runtimeParamsFromNodeStore (nodeStore.ts:179-196) over the file's runtime params."""
    rp = dict(graph["runtimeParams"])
    w, gr, v = graph["wind"], graph["gravity"], graph["vorticity"]
    rp.update({
        "wind": [w["directionX"], w["directionY"], w["directionZ"]], "windStrength": w["strength"],
        "gravity": [gr["directionX"], gr["directionY"], gr["directionZ"]], "gravityStrength": gr["strength"],
        "buoyancy": gr["buoyancy"], "vorticityStrength": v["strength"],
        "vorticityConstantMask": 1 if v.get("constantMask") is None else v["constantMask"],
        "vorticityVelocityMask": 0.4 if v.get("velocityMask") is None else v["velocityMask"],
        "vorticityHeatMask": 0.8 if v.get("heatMask") is None else v["heatMask"],
        "vorticityDensityMask": 0.35 if v.get("densityMask") is None else v["densityMask"],
    })
    return rp


@dataclass
class Scene:
    """Everything a loaded graph file sets on the step and on the raymarch."""
    runtime_params: Dict
    quality: Dict[str, int]
    advection_mode: str
    sources: List[EmitterSource]
    lights: List[RenderLight]
    render_output: Dict
    edges: List[Dict] = field(default_factory=list)

    def bloom(self):
        """The file's bloom settings (RenderOutputNodeProps) as (enabled, threshold, strength, radius)."""
        ro = self.render_output
        return (bool(ro.get("bloomEnabled", True)), float(ro.get("bloomThreshold", 0.6)),
                float(ro.get("bloomStrength", 0.9)), float(ro.get("bloomRadius", 0.4)))

    def apply(self, sim, raymarch=None) -> None:
        """applySerializedGraph for the simulation (and the raymarch pass, when given):
        runtime params, quality, advection mode, sources; then reset."""
        sim.setRuntimeParams(self.runtime_params)
        sim.setQualitySettings(self.quality)
        sim.setScalarAdvectionMode(self.advection_mode)
        sim.updateSources(self.sources)
        sim.reset()
        if raymarch is not None:
            ro = self.render_output
            raymarch.setRenderParams(stepCount=ro.get("stepCount"), lights=self.lights,
                                     scatteringForward=ro.get("scatteringForward"),
                                     scatteringBack=ro.get("scatteringBack"))


def scene_from_graph(graph: Dict) -> Scene:
    """A parsed graph file -> Scene (file edges, or the default edges when it has none)."""
    edges = graph["graph"]["edges"] if graph.get("graph") else default_graph_edges(len(graph["emitters"]),
                                                                                   len(graph["lights"]))
    lights = []
    for e in edges:
        if e["target"] == "render-output" and e["source"].startswith("light-"):
            light = next((l for l in graph["lights"] if l["id"] == e["source"]), None)
            if light is None:
                continue
            p = light["props"]
            lights.append(RenderLight(type=p["lightType"], direction=(p["dirX"], p["dirY"], p["dirZ"]),
                                      position=(p["posX"], p["posY"], p["posZ"]), color=(p["colorR"], p["colorG"], p["colorB"]),
                                      intensity=p["intensity"]))
    return Scene(runtime_params=runtime_params_from_nodes(graph), quality=dict(graph["simulationQuality"]),
                 advection_mode=graph["advection"].get("mode", "maccormack"),
                 sources=resolve_emitter_sources(graph["emitters"], edges), lights=lights,
                 render_output=dict(graph["renderOutput"]), edges=edges)


def load_scene(raw: str) -> Scene:
    """Reads a graph file's text (what the reference's Open button accepts)."""
    return scene_from_graph(parse_graph_json(raw))


def graph_from_preset(preset: Dict, combustion: Optional[Dict] = None, render_output: Optional[Dict] = None) -> Dict:
    """A new-file preset (presets/types.ts:16-25: emitters, lights, optional graphEdges,
    runtimeParams, renderOutput overrides) as the graph the editor builds from it
    (nodeStore.ts:111-117, 146-182; nodeGraphStore.ts:97-118)."""
    emitters = [{"id": f"emitter-{i}", "label": e["label"], "props": dict(e["props"])}
                for i, e in enumerate(preset["emitters"])]
    lights = [{"id": f"light-{i}", "label": l["label"], "props": dict(l["props"])} for i, l in enumerate(preset["lights"])]
    rp = copy.deepcopy(preset["runtimeParams"])
    grav = rp.get("gravity") or [0, -1, 0]
    ro = {"displayMode": "temperature", "stepCount": 400, "scatteringForward": 0.32, "scatteringBack": -0.18,
          "bloomEnabled": True, "bloomThreshold": 0.6, "bloomStrength": 0.9, "bloomRadius": 0.4}
    ro.update(render_output or {})
    ro.update(preset.get("renderOutput") or {})
    graph = {
        "version": GRAPH_FILE_VERSION,
        "runtimeParams": rp,
        "emitters": emitters,
        "lights": lights,
        "combustion": dict(combustion or {}),
        "advection": {"mode": "maccormack"},
        "wind": {"directionX": rp["wind"][0], "directionY": rp["wind"][1], "directionZ": rp["wind"][2],
                 "strength": rp["windStrength"]},
        "gravity": {"directionX": grav[0], "directionY": grav[1], "directionZ": grav[2],
                    "strength": 0.45 if rp.get("gravityStrength") is None else rp["gravityStrength"],
                    "buoyancy": rp["buoyancy"]},
        "vorticity": {"strength": rp["vorticityStrength"],
                      "constantMask": 1 if rp.get("vorticityConstantMask") is None else rp["vorticityConstantMask"],
                      "velocityMask": 0.4 if rp.get("vorticityVelocityMask") is None else rp["vorticityVelocityMask"],
                      "heatMask": 0.8 if rp.get("vorticityHeatMask") is None else rp["vorticityHeatMask"],
                      "densityMask": 0.35 if rp.get("vorticityDensityMask") is None else rp["vorticityDensityMask"]},
        "renderOutput": ro,
        "simulationQuality": create_quality_settings({}),
        "graph": {"nodePositions": {}, "edges": default_graph_edges(len(emitters), len(lights), preset.get("graphEdges"))},
    }
    return graph
