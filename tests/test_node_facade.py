"""The Node host facade (3dgs_tile_webgpu_amd/node): the N-API addon over the
C ABI and the JavaScript pass API above it.

CPU tests check that the addon loads and exports what the facade calls, that
option validation throws the reference's error classes for the same cases as
the Python mirror, that the capacity policy and the object-frame marshalling
agree with the Python side, and that creating a pass without a device throws.
The GPU test renders a frame from JavaScript and compares the image bits,
counts and sorted records with the oracle (bit-exact), and a two-view batch
with the single renders.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
FACADE = ROOT / "3dgs_tile_webgpu_amd" / "node"
ADDON = ROOT / "3dgs_tile_webgpu_amd" / "_lib" / "gs_napi.node"

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not ADDON.exists(),
                                reason="node or the built N-API addon is missing")


def run_node(script: str, payload=None, timeout: float = 120.0):
    """Runs `script` with `facade` bound and `input` parsed from stdin; the
    script prints one JSON line."""
    prelude = (f"const facade = require({json.dumps(str(FACADE))});\n"
               "const input = JSON.parse(require('fs').readFileSync(0, 'utf8'));\n")
    proc = subprocess.run(["node", "-e", prelude + script], input=json.dumps(payload),
                          capture_output=True, text=True, timeout=timeout)
    assert proc.returncode == 0, proc.stderr[-4000:]
    return json.loads(proc.stdout.strip().splitlines()[-1])


def test_addon_exports_what_the_facade_calls():
    names = run_node("console.log(JSON.stringify(Object.keys(facade.native)));")
    used = set()
    for line in (FACADE / "index.js").read_text().splitlines():
        used.update(w.split("(")[0] for w in line.split("native.")[1:])
    assert used and used <= set(names), used - set(names)
    assert {"create", "setData", "render", "renderBatch", "readState", "readImage",
            "readDebug", "resizeIntersections", "destroy"} <= set(names)


CASES = [
    {},
    {"depthSortMode": "float16"},
    {"antialiasMode": "msaa"},
    {"rasterTransmittanceThreshold": 0.0},
    {"rasterTransmittanceThreshold": 1.0},
    {"rasterChunkSize": 128},
    {"rasterChunkSize": 1000},
    {"rasterChunkSize": None},
    {"filterVariance": -1.0},
    {"boundaryPolicy": "wrap"},
    {"boundaryPolicy": "repeat-x"},
    {"background": [0, 0, 0]},
    {"fixedSize": {"width": 0, "height": 10}},
    {"fixedSize": {"width": 64, "height": 10}},
    {"depthRange": [5.0, 1.0]},
    {"unknownOption": 1},
    {"tileSize": 12},
    {"tileSize": 32},
    {"radixBits": 5},
    {"rasterBatch": 512},
    {"intersectionCapacity": 16, "initialIntersectionCapacity": 12},
    {"maxIntersectionCapacity": (1 << 28) + 1},
    {"initialIntersectionCapacity": 2000, "maxIntersectionCapacity": 1000},
    {"intersectionGrowthFactor": 1.0},
    {"intersectionHeadroom": 0.5},
    {"intersectionCapacity": 0},
    {"heavySplitUnits": 0},
    {"heavySplitUnits": -1},
    {"heavySplitUnits": 1.5},
    {"heavyFineSplitUnits": 0},
    {"heavyFineSplitUnits": 4096},
    {"heavyFineSplitUnits": -2},
    {"heavyFineSplitUnits": 0xFFFFFFFF},
    {"heavyPairUnits": 0},
    {"heavyPairUnits": 1408},
    {"heavyPairUnits": -1},
    {"heavyPairUnits": True},
    {"viewGroup": 1},
    {"viewGroup": 8},
    {"viewGroup": 0},
    {"viewGroup": 9},
    {"viewGroup": 2.5},
]


def python_outcome(gs, opts) -> str:
    try:
        gs.IntersectionCapacityPolicy(gs.validate_options(dict(opts)))
    except TypeError:
        return "TypeError"
    except ValueError:
        return "RangeError"
    return "ok"


def test_option_validation_matches_python_mirror(gs):
    got = run_node("""
      const out = input.map((o) => {
        try { new facade.IntersectionCapacityPolicy(facade.validateOptions(o)); return 'ok'; }
        catch (e) { return e.constructor.name; }
      });
      console.log(JSON.stringify(out));""", CASES)
    want = [python_outcome(gs, c) for c in CASES]
    assert got == want
    assert got.count("ok") >= 5 and "TypeError" in got and "RangeError" in got


def test_capacity_policy_matches_python(gs):
    rng = np.random.default_rng(5)
    pairs = [[int(rng.integers(1, 5_000_000)), int(rng.integers(1, 12_000_000))] for _ in range(200)]
    pairs += [[512, 900], [1000, 1001], [256, 700]]
    got = run_node("""
      const p = new facade.IntersectionCapacityPolicy({maxIntersectionCapacity: 10000000});
      const q = new facade.IntersectionCapacityPolicy({initialIntersectionCapacity: 256, maxIntersectionCapacity: 1000});
      console.log(JSON.stringify({big: input.map(([c, d]) => p.next(c, d)),
                                  small: [q.next(512, 900), q.next(1000, 1001)],
                                  initial: new facade.IntersectionCapacityPolicy({}).initial}));""", pairs)
    p = gs.IntersectionCapacityPolicy({"maxIntersectionCapacity": 10_000_000})
    assert got["big"] == [p.next(c, d) for c, d in pairs]
    assert got["small"] == [1000, None]  # test/IntersectionCapacity.test.ts:26-60
    assert got["initial"] == 262_144


def test_object_frame_marshalling_matches_python(gs):
    cam = gs.PerspectiveCamera(position=(3.0, -2.0, 11.0), target=(0.5, 0.2, -0.3))
    world = np.eye(4)
    world[:3, :3] = np.linalg.qr(np.random.default_rng(2).normal(size=(3, 3)))[0] * 1.3
    world[:3, 3] = [1.0, 2.0, -3.0]
    got = run_node("""
      const cam = {matrixWorld: {elements: input.cam}};
      console.log(JSON.stringify({a: Array.from(facade.objectFrames(cam, input.world)),
                                  b: Array.from(facade.objectFrames(cam, null, false))}));""",
                   {"cam": gs.elements(cam.matrix_world).tolist(), "world": gs.elements(world).tolist()})
    np.testing.assert_array_equal(np.float32(got["a"]), gs.object_frames_for(cam, world))
    np.testing.assert_array_equal(np.float32(got["b"]), gs.object_frames_for(cam, None, visible=False))


def test_create_without_device_throws():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be visible")
    got = run_node("""
      try { facade.gaussianPass({}, {}, {}); console.log(JSON.stringify(null)); }
      catch (e) { console.log(JSON.stringify([e.constructor.name, e.code])); }""")
    assert got == ["Error", -4]


def write_scene(tmp: Path, scene) -> dict:
    files = {}
    for name, arr in (("means", scene.means), ("scalesOpacity", scene.scales_opacity),
                      ("rotations", scene.rotations), ("shCoefficients", scene.sh)):
        path = tmp / f"{name}.bin"
        np.ascontiguousarray(arr).tofile(path)
        files[name] = str(path)
    return {"files": files, "count": scene.count, "shDegree": scene.sh_degree}


LOAD_SCENE = """
  const fs = require('fs');
  const load = (p, T) => { const b = new Uint8Array(fs.readFileSync(p)); return new T(b.buffer); };
  const data = {count: input.scene.count, shDegree: input.scene.shDegree,
                means: load(input.scene.files.means, Float32Array),
                scalesOpacity: load(input.scene.files.scalesOpacity, Uint32Array),
                rotations: load(input.scene.files.rotations, Uint32Array),
                shCoefficients: load(input.scene.files.shCoefficients, Uint32Array)};
  const camera = (c) => ({projectionMatrix: {elements: c.projection}, matrixWorld: {elements: c.world},
                          near: c.near, far: c.far});
  const dump = (p, a) => fs.writeFileSync(p, Buffer.from(a.buffer, a.byteOffset, a.byteLength));
"""


def camera_json(gs, cam) -> dict:
    return {"projection": gs.elements(cam.projection_matrix).tolist(),
            "world": gs.elements(cam.matrix_world).tolist(), "near": cam.near, "far": cam.far}


@pytest.mark.gpu
def test_node_pass_renders_bit_exact(gs, oracle, gpu_available, tmp_path):
    scene = gs.workload_scene("small", count=20_000)
    cam = gs.WORKLOADS["small"].camera()
    w, h = 320, 240
    payload = {"scene": write_scene(tmp_path, scene), "camera": camera_json(gs, cam), "w": w, "h": h,
               "out": str(tmp_path)}
    print("node pass: small 20K 320x240", flush=True)
    got = run_node(LOAD_SCENE + """
      const pass = facade.gaussianPass(camera(input.camera), data, {initialIntersectionCapacity: 4096});
      pass.render(input.w, input.h);
      dump(input.out + '/image.bin', pass.readImage());
      const res = pass.getResources();
      dump(input.out + '/sorted.bin', res.sorted_records);
      pass.readStats().then((stats) => {
        console.log(JSON.stringify({stats, status: pass.getOutputStatus(), state: pass.readState()}));
        pass.dispose();
      });""", payload)
    ref = oracle.run_frame(scene, gs.frame_for(cam, w, h), gs.object_frames_for(cam),
                           capacity=got["stats"]["intersectionCapacity"], threads=4)
    st = got["stats"]
    assert (st["visibleGaussianCount"], st["intersectionCount"]) == (ref.counts.visible, ref.counts.count)
    assert not st["overflow"] and got["state"]["errorFlags"] == 0
    # the first attempt overflowed the 4096-record capacity and the retry published
    assert got["status"]["growthCount"] >= 1 and got["status"]["state"] == "ready"
    sorted_records = np.fromfile(tmp_path / "sorted.bin", np.uint32).reshape(-1, 2)
    np.testing.assert_array_equal(sorted_records, ref.get("sorted_records"))
    np.testing.assert_array_equal(np.fromfile(tmp_path / "image.bin", np.uint16),
                                  ref.get("image_f16").reshape(-1))


@pytest.mark.gpu
def test_node_batch_matches_single_renders(gs, gpu_available, tmp_path):
    scene = gs.workload_scene("small", count=20_000)
    cams = [gs.PerspectiveCamera(fov=60, aspect=w / h, position=pos, target=(0, 0, 0))
            for pos, (w, h) in (((0.0, 0.0, 6.0), (256, 192)), ((2.0, 1.0, 5.0), (200, 150)))]
    sizes = [(256, 192), (200, 150)]
    payload = {"scene": write_scene(tmp_path, scene), "cams": [camera_json(gs, c) for c in cams],
               "sizes": sizes, "out": str(tmp_path)}
    print("node batch: small 20K two views", flush=True)
    got = run_node(LOAD_SCENE + """
      const cams = input.cams.map(camera);
      const single = cams.map((c, i) => {
        const pass = facade.gaussianPass(c, data, {initialIntersectionCapacity: 4096});
        pass.render(input.sizes[i][0], input.sizes[i][1]);
        const img = pass.readImage();
        pass.dispose();
        return img;
      });
      const batch = new facade.GaussianBatch(data, {initialIntersectionCapacity: 4096});
      batch.render(cams, input.sizes);
      const same = cams.map((c, i) => batch.readImage(i).every((v, k) => v === single[i][k]));
      const states = batch.readStates();
      const capacities = batch.capacities;
      const ctx = batch._ctx;
      batch.dispose();
      let threw = null;
      try { facade.native.readState(ctx); } catch (e) { threw = e.code; }
      console.log(JSON.stringify({same, states, capacities, threw}));""", payload)
    assert got["capacities"][0] > 4096  # both views overflowed once and grew
    assert got["same"] == [True, True]
    assert all(s["overflow"] == 0 and s["errorFlags"] == 0 and s["visible"] > 0 for s in got["states"])
    assert got["threw"] == -5  # a disposed handle is GS_E_STATE


def test_short_object_tables_are_rejected_before_the_library(gs):
    """Every camera's object table must have the rows the library reads
    (40 floats per object of setObjects); both host mirrors refuse a batch
    whose tables differ, before any native call reads them."""
    import importlib
    batch = importlib.import_module("3dgs_tile_webgpu_amd.batch")
    one, two = np.zeros((1, 40), np.float32), np.zeros((2, 40), np.float32)
    batch.check_object_tables([two, two])
    with pytest.raises(gs.GsRangeError):
        batch.check_object_tables([two, one])
    got = run_node("""
      const t = (n) => new Float32Array(40 * n);
      const out = [[t(2), t(2)], [t(2), t(1)], [t(1), new Float32Array(39)]].map((tables) => {
        try { facade.checkObjectTables(tables); return 'ok'; } catch (e) { return e.constructor.name; }
      });
      console.log(JSON.stringify(out));""")
    assert got == ["ok", "RangeError", "RangeError"]


@pytest.mark.gpu
def test_node_host_contract(gs, gpu_available, tmp_path):
    """From JavaScript: a short per-frame object table is a RangeError from
    the addon (not an over-read); redrawStrategy "auto" reuses a clean frame;
    updateActive shrinks the active list; the batch autotuner returns the 18
    shapes; the state carries the heavy-unit counter."""
    scene = gs.workload_scene("small", count=20_000)
    cam = gs.WORKLOADS["small"].camera()
    payload = {"scene": write_scene(tmp_path, scene), "camera": camera_json(gs, cam)}
    print("node host contract: small 20K", flush=True)
    got = run_node(LOAD_SCENE + """
      const c = camera(input.camera);
      const pass = facade.gaussianPass(c, data, {redrawStrategy: 'auto', rasterStats: true,
                                                 intersectionCapacity: 1 << 20});
      pass.render(320, 180);
      pass.render(320, 180);
      const counts = [pass.renderCount, pass.cacheHitCount];
      const st = pass.readState();
      const frame = {projection: Float32Array.from(c.projectionMatrix.elements), width: 320,
                     height: 180, near: c.near, far: c.far, objectFrames: new Float32Array(20)};
      let short = null;
      try { facade.native.render(pass._ctx, frame); } catch (e) { short = e.constructor.name; }
      const slots = new Uint32Array(5000).map((v, i) => i);
      pass.updateActive(0, slots, 5000);
      pass.render(320, 180);
      const visibleAfter = pass.readState().visible;
      pass.dispose();
      const batch = new facade.GaussianBatch(data, {initialIntersectionCapacity: 1 << 20});
      const tuned = batch.autotune([c, c], [[320, 180], [256, 144]], 1);
      batch.dispose();
      console.log(JSON.stringify({counts, st, short, visibleAfter, tuned}));""", payload)
    assert got["counts"] == [1, 1]
    assert got["short"] == "RangeError"
    st = got["st"]
    assert st["rasterChecked"] > st["rasterCheckedHeavy"] >= 0 and st["visible"] > 5000
    assert 0 < got["visibleAfter"] <= 5000
    assert len(got["tuned"]["table"]) == 18 and got["tuned"]["best"]["tileSize"] in (8, 16, 32)


@pytest.mark.gpu
def test_node_trains_calibrates_and_enables_the_cull(gs, gpu_available, tmp_path):
    """A Node script drives the learned cull end to end on the dense
    3M-Gaussian scene at 1920x1080: collects training frames, fits the
    normalisation, trains, calibrates the bias on 8 validation views (mean
    PSNR >= 46 dB) and enables it; on 8 held-out views the culled frames
    remove keys while keeping a mean PSNR >= 44 dB against the plain ones
    (8 views vary by several dB around the validation mean)."""
    scene = gs.workload_scene("3m_cull")
    views = gs.orbit_cameras(112, seed=41, radius=(12.0, 25.0), resolutions=[(1920, 1080)])
    payload = {"scene": write_scene(tmp_path, scene),
               "cams": [camera_json(gs, c) for c, _ in views], "sizes": [s for _, s in views]}
    print("node learned cull: 3M, 96 train + 8 valid + 8 held-out views at 1920x1080", flush=True)
    got = run_node(LOAD_SCENE + """
      const cams = input.cams.map(camera);
      const batch = new facade.GaussianBatch(data, {initialIntersectionCapacity: 1 << 24});
      const cull = new facade.LearnedCull(batch, 3);
      cull.reserve(20000000);
      for (let k = 0; k < 96; k++) cull.collect(cams[k], input.sizes[k][0], input.sizes[k][1], 0.05, k);
      const n = cull.datasetSize();
      cull.fitNormalization();
      const loss0 = cull.train(1, {underWeight: 300, seed: 1});
      cull.train(10000, {lr: 2e-3, underWeight: 300, seed: 2});
      cull.train(5000, {lr: 5e-4, underWeight: 300, seed: 3});
      const loss = cull.train(5000, {lr: 1e-4, underWeight: 300, seed: 4});
      const cal = cull.calibrate(cams.slice(96, 104), input.sizes.slice(96, 104), 46);
      const held = cams.slice(104), sizes = input.sizes.slice(104);
      cull.disable();
      batch.render(held, sizes);
      const plain = held.map((c, v) => batch.readImage(v));
      const keysPlain = batch.readStates().reduce((a, s) => a + s.intersections, 0);
      cull.enable(1e-4);
      batch.render(held, sizes);
      const keysCull = batch.readStates().reduce((a, s) => a + s.intersections, 0);
      const ps = held.map((c, v) => Math.min(999, facade.psnr(batch.readImage(v), plain[v])));
      batch.dispose();
      console.log(JSON.stringify({n, loss0, loss, shift: cal.shift, rows: cal.rows, keysPlain,
                                  keysCull, ps}));""", payload, timeout=600)
    print({k: v for k, v in got.items() if k != "rows"}, got["rows"][-3:])
    assert got["n"] > 1_000_000 and got["loss"] < 0.5 * got["loss0"]
    assert got["shift"] is not None
    assert got["keysCull"] < 0.7 * got["keysPlain"]
    assert np.mean(got["ps"]) >= 44.0


@pytest.mark.gpu
def test_node_cull_checkpoint_round_trip(gs, gpu_available, tmp_path):
    """The learned cull's training checkpoint from JavaScript: checkpoint()
    returns the library's blob, restore() resumes from it (threshold
    included) and a truncated blob throws RangeError with the state kept."""
    scene = gs.workload_scene("small", count=20_000)
    cam = gs.WORKLOADS["small"].camera()
    payload = {"scene": write_scene(tmp_path, scene), "camera": camera_json(gs, cam)}
    got = run_node(LOAD_SCENE + """
      const pass = facade.gaussianPass(camera(input.camera), data, {initialIntersectionCapacity: 1 << 20});
      const cull = new facade.LearnedCull(pass, 2);
      cull.reserve(1 << 16);
      cull.collect(camera(input.camera), 320, 240, 1.0, 1);
      cull.fitNormalization();
      cull.train(50, {batch: 4096, underWeight: 300, seed: 1});
      cull.threshold = 2e-4;
      cull.disable();
      const blob = cull.checkpoint();
      cull.train(20, {batch: 4096, underWeight: 300, seed: 2});
      const moved = cull.params();
      cull.threshold = 1e-4;
      cull.disable();
      cull.restore(blob);
      const again = cull.checkpoint();
      let same = again.length === blob.length;
      for (let i = 0; same && i < blob.length; i++) same = again[i] === blob[i];
      let error = null;
      try { cull.restore(blob.subarray(0, blob.length - 4)); } catch (e) { error = e.name; }
      const kept = cull.checkpoint();
      let keptSame = true;
      for (let i = 0; keptSame && i < blob.length; i++) keptSame = kept[i] === blob[i];
      const p = cull.params();
      let diff = 0;
      for (let i = 0; i < p.length; i++) diff = Math.max(diff, Math.abs(p[i] - moved[i]));
      pass.dispose();
      console.log(JSON.stringify({bytes: blob.length, same, error, keptSame,
                                  threshold: cull.threshold, diff}));""", payload)
    assert got["bytes"] == 32 + (3 * 5313 + 2 * 16) * 4
    assert got["same"] and got["keptSame"] and got["error"] == "RangeError"
    assert abs(got["threshold"] - 2e-4) < 1e-9
    assert got["diff"] > 0.0  # the restore took the parameters back before the 20 steps
