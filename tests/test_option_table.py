"""Option validation pinned to the reference's error classes, case by case,
in both host mirrors (Python: GsRangeError is a ValueError, GsTypeError a
TypeError; Node: RangeError / TypeError).

Every row cites the reference check it restates. Three kinds of row differ
from the reference on purpose and say so: depthSortMode has no run-time
check there (its type system rejects other values), unknown options are
rejected here instead of ignored, and the reference's WebGPU dispatch
limits (65,535 workgroups per dimension: intersectionCapacity above
256 * 65,535, or more than 65,535 worst-case raster chunk tasks) do not
exist on this path, so those values are accepted.
"""
from __future__ import annotations

import shutil

import pytest

from test_node_facade import ADDON, run_node

R, T, OK = "RangeError", "TypeError", "ok"

TABLE = [
    ({}, OK, "defaults"),
    ({"antialiasMode": "classic"}, OK, "GaussianPass.ts:161-165"),
    ({"antialiasMode": "msaa"}, R, "GaussianPass.ts:161-165"),
    ({"redrawStrategy": "auto"}, OK, "GaussianPass.ts:166-174"),
    ({"redrawStrategy": "never"}, OK, "GaussianPass.ts:166-174"),
    ({"redrawStrategy": "sometimes"}, R, "GaussianPass.ts:166-174"),
    ({"radixBackend": "workgroup"}, OK, "radixBackend.ts:7-15"),
    ({"radixBackend": "gpu"}, R, "radixBackend.ts:7-15"),
    ({"dataMode": "external"}, OK, "GaussianPass.ts:211-213"),
    ({"dataMode": "shared"}, R, "GaussianPass.ts:211-213"),
    ({"intersectionCapacity": 0}, R, "GaussianPass.ts:179-185"),
    ({"intersectionCapacity": 1.5}, R, "GaussianPass.ts:179-185"),
    ({"intersectionCapacity": 20_000_000}, OK, "deliberate: no WebGPU dispatch limit"),
    ({"rasterChunkSize": 128}, R, "rasterChunks.ts:85-94"),
    ({"rasterChunkSize": 1000}, R, "rasterChunks.ts:85-94"),
    ({"rasterChunkSize": None}, OK, "rasterChunks.ts:84"),
    ({"rasterChunkSize": 256, "intersectionCapacity": 10_000_000}, OK,
     "deliberate: no 65,535 chunk-task dispatch limit"),
    ({"depthAlphaThreshold": 1.5}, R, "GaussianPass.ts:236-242"),
    ({"depthAlphaThreshold": float("nan")}, R, "GaussianPass.ts:236-242"),
    ({"rasterTransmittanceThreshold": 0.0}, R, "GaussianPass.ts:246-255"),
    ({"rasterTransmittanceThreshold": 1.0}, R, "GaussianPass.ts:246-255"),
    ({"rasterTransmittanceThreshold": float("nan")}, R, "GaussianPass.ts:246-255"),
    ({"fixedSize": {"width": 0, "height": 10}}, R, "GaussianPass.ts:386-395"),
    ({"fixedSize": {"width": 64, "height": 2.5}}, R, "GaussianPass.ts:386-395"),
    ({"fixedSize": {"width": 64, "height": 10}}, OK, "GaussianPass.ts:386-395"),
    ({"depthRange": [5.0, 1.0]}, R, "GaussianPass.ts:410-423"),
    ({"depthRange": [0.0, 1.0]}, R, "GaussianPass.ts:410-423"),
    ({"depthRange": [1e-50, 1.0]}, R, "GaussianPass.ts:410-423 (fround(near) = 0)"),
    ({"depthRange": [1.0, 1e39]}, R, "GaussianPass.ts:410-423 (far beyond float32)"),
    ({"depthRange": [1.0, 1.00000001]}, R, "GaussianPass.ts:410-423 (equal in float32)"),
    ({"depthRange": [0.5, 50.0]}, OK, "GaussianPass.ts:410-423"),
    ({"filterVariance": -1.0}, R, "GaussianPass.ts:436-441"),
    ({"filterVariance": 1e39}, R, "GaussianPass.ts:436-441"),
    ({"filterVariance": 0.0}, OK, "GaussianPass.ts:436-441"),
    ({"invalidDepthValue": float("nan")}, R, "GaussianPass.ts:449-455"),
    ({"invalidDepthValue": 1e39}, R, "GaussianPass.ts:449-455"),
    ({"invalidDepthValue": float("inf")}, OK, "GaussianPass.ts:449-455"),
    ({"boundaryPolicy": "wrap"}, R, "GaussianPass.ts:462-466"),
    ({"boundaryPolicy": "repeat-y"}, OK, "GaussianPass.ts:462-466"),
    ({"intersectionCapacity": 16, "initialIntersectionCapacity": 12}, R,
     "IntersectionCapacityPolicy.ts:58-65"),
    ({"maxIntersectionCapacity": 0}, R, "IntersectionCapacityPolicy.ts:49-56"),
    ({"maxIntersectionCapacity": (1 << 28) + 1}, R, "IntersectionCapacityPolicy.ts:66-70"),
    ({"initialIntersectionCapacity": 2000, "maxIntersectionCapacity": 1000}, R,
     "IntersectionCapacityPolicy.ts:75-78"),
    ({"intersectionGrowthFactor": 1.0}, R, "IntersectionCapacityPolicy.ts:82-83"),
    ({"intersectionHeadroom": 0.5}, R, "IntersectionCapacityPolicy.ts:84-85"),
    ({"intersectionReadbackInterval": -1}, R, "IntersectionCapacityPolicy.ts:86-87"),
    ({"intersectionReadbackInterval": 0}, OK, "IntersectionCapacityPolicy.ts:86-87"),
    ({"depthSortMode": "float16"}, T, "deliberate: a type error in the reference's types.ts:45"),
    ({"unknownOption": 1}, T, "deliberate: unknown options are rejected"),
]


def python_outcome(gs, opts) -> str:
    try:
        gs.IntersectionCapacityPolicy(gs.validate_options(dict(opts)))
    except TypeError:
        return T
    except ValueError:
        return R
    return OK


@pytest.mark.parametrize("opts,want,ref", TABLE, ids=[f"{i}:{r[2]}" for i, r in enumerate(TABLE)])
def test_python_mirror_matches_reference(gs, opts, want, ref):
    assert python_outcome(gs, opts) == want


@pytest.mark.skipif(shutil.which("node") is None or not ADDON.exists(),
                    reason="node or the built N-API addon is missing")
def test_node_facade_matches_reference():
    # JSON has no NaN / Infinity: send them as strings and map them back.
    def enc(v):
        if isinstance(v, float) and v != v:
            return "__nan__"
        if v == float("inf"):
            return "__inf__"
        if isinstance(v, dict):
            return {k: enc(x) for k, x in v.items()}
        if isinstance(v, (list, tuple)):
            return [enc(x) for x in v]
        return v
    got = run_node("""
      const dec = (v) => v === '__nan__' ? NaN : v === '__inf__' ? Infinity :
        Array.isArray(v) ? v.map(dec) : (v && typeof v === 'object') ?
        Object.fromEntries(Object.entries(v).map(([k, x]) => [k, dec(x)])) : v;
      const out = input.map((o) => {
        try { new facade.IntersectionCapacityPolicy(facade.validateOptions(dec(o))); return 'ok'; }
        catch (e) { return e.constructor.name; }
      });
      console.log(JSON.stringify(out));""", [enc(o) for o, _, _ in TABLE])
    bad = [(TABLE[i][0], g, TABLE[i][1]) for i, g in enumerate(got) if g != TABLE[i][1]]
    assert not bad, bad
