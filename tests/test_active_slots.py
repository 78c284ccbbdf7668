"""The store's active slot list and GaussianData's buffer checks, on the host.

ActiveSlots (Python and the Node facade) restates GaussianData.stageActivation
(src/renderer/GaussianData.ts:170-232); the CPU tests replay
test/GaussianData.test.ts:94-123 (staged batches, commit, listeners, the
capacity error) and check the committed ranges, including a randomized
comparison of the two implementations against a plain restatement of the
list semantics. check_data / checkData restate the constructor checks
(GaussianData.ts:76-166, 265-318) with the reference's error classes; the
reference's itemSize checks have no counterpart for flat arrays, so an array
of the wrong element type is the TypeError case here (the reference's
representation-mismatch RangeError of test/GaussianData.test.ts:60-92 is
covered by its typed-array TypeError). The GPU test commits changes through
a tracked list and renders bit for bit what the oracle renders for the same
active list.
"""
from __future__ import annotations

import json
import shutil

import numpy as np
import pytest

from test_node_facade import ADDON, run_node

node_missing = shutil.which("node") is None or not ADDON.exists()


def test_staged_activation_reference_sequence(gs):
    """test/GaussianData.test.ts:94-123."""
    a = gs.ActiveSlots(6, np.array([0, 5], np.uint32))
    first, second = [], []
    a.subscribe(first.append)
    a.subscribe(second.append)
    assert a.stage_activation([2], [0], False) is None
    assert a.count == 2 and a.slots.tolist() == [0, 5]
    assert first == []
    rng = a.stage_activation([3], [], True)
    assert a.count == 3 and a.slots.tolist() == [5, 2, 3]
    assert first == second == [rng] == [{"start": 0, "count": 3}]
    a.stage_activation([], [5, 3, 2], True)
    assert a.count == 0
    with pytest.raises(gs.GsRangeError, match="capacity"):
        a.stage_activation([6], [], True)
    with pytest.raises(gs.GsRangeError, match="Removed GPU slot exceeds capacity"):
        a.stage_activation([], [6], True)


def test_ranges_cover_every_changed_ordinal(gs):
    """A commit's range covers every ordinal whose slot changed, and only
    ordinals below the new count; an empty staging is not a commit."""
    rng = np.random.default_rng(5)
    a = gs.ActiveSlots(64, rng.choice(64, 20, replace=False).astype(np.uint32))
    assert a.stage_activation([], [], True) is None
    for _ in range(200):
        before = a.slots.copy()
        removed = rng.choice(64, rng.integers(0, 6), replace=False)
        added = rng.choice(64, rng.integers(0, 6), replace=False)
        r = a.stage_activation(added, removed, True)
        after = a.slots
        n = min(before.size, after.size)
        changed = [i for i in range(n) if before[i] != after[i]] + list(range(n, after.size))
        if r is None:
            assert removed.size == 0 and added.size == 0
            continue
        assert 0 <= r["start"] and r["start"] + r["count"] <= after.size
        assert all(r["start"] <= i < r["start"] + r["count"] for i in changed)
        assert len(set(after.tolist())) == after.size  # each slot once


def list_model(ops, capacity, initial):
    """The list semantics written out plainly: removals swap the last slot
    into the hole, additions append, both skip no-ops."""
    values = list(initial)
    for added, removed in ops:
        for s in removed:
            if s in values:
                i = values.index(s)
                last = values.pop()
                if i < len(values):
                    values[i] = last
        for s in added:
            if s not in values:
                values.append(s)
    return values


@pytest.mark.skipif(node_missing, reason="node or the built N-API addon is missing")
def test_python_and_node_lists_agree_with_the_model(gs):
    rng = np.random.default_rng(11)
    initial = rng.choice(40, 12, replace=False).tolist()
    ops = [(rng.choice(40, rng.integers(0, 5), replace=False).tolist(),
            rng.choice(40, rng.integers(0, 5), replace=False).tolist()) for _ in range(60)]
    a = gs.ActiveSlots(40, np.array(initial, np.uint32))
    py_ranges = [a.stage_activation(ad, rm, True) for ad, rm in ops]
    js = run_node("""
const a = new facade.ActiveSlots(40, Uint32Array.from(input.initial));
const ranges = input.ops.map(([ad, rm]) => a.stageActivation(Uint32Array.from(ad), Uint32Array.from(rm), true));
console.log(JSON.stringify({slots: Array.from(a.slots), ranges, version: a.version}));
""", {"initial": initial, "ops": ops})
    assert a.slots.tolist() == js["slots"] == list_model(ops, 40, initial)
    assert py_ranges == js["ranges"]
    assert js["version"] == a.version


@pytest.mark.skipif(node_missing, reason="node or the built N-API addon is missing")
def test_node_staged_activation_reference_sequence():
    out = run_node("""
const a = new facade.ActiveSlots(6, Uint32Array.from([0, 5]));
const first = [], second = [];
a.subscribe((r) => first.push(r)); a.subscribe((r) => second.push(r));
const r0 = a.stageActivation(Uint32Array.from([2]), Uint32Array.from([0]), false);
const s0 = Array.from(a.slots);
const r1 = a.stageActivation(Uint32Array.from([3]), new Uint32Array(0), true);
const s1 = Array.from(a.slots);
const f1 = first.slice(), g1 = second.slice();
a.stageActivation(new Uint32Array(0), Uint32Array.from([5, 3, 2]), true);
let err = null;
try { a.stageActivation(Uint32Array.from([6]), new Uint32Array(0), true); } catch (e) { err = [e.constructor.name, e.message]; }
console.log(JSON.stringify({r0, s0, r1, s1, n: a.count, first: f1, second: g1, err}));
""", {})
    assert out["r0"] is None and out["s0"] == [0, 5]
    assert out["s1"] == [5, 2, 3] and out["r1"] == {"start": 0, "count": 3}
    assert out["first"] == out["second"] == [out["r1"]]
    assert out["n"] == 0
    assert out["err"][0] == "RangeError" and "capacity" in out["err"][1]


def scene_of(gs, n=4, degree=1):
    means = np.zeros((n, 4), np.float32)
    return gs.PackedScene(means, np.zeros((n, 2), np.uint32), np.zeros(n, np.uint32),
                          np.zeros((n, (degree + 1) ** 2), np.uint32), degree)


DATA_CASES = [
    ("means_vec3", "RangeError", "itemSize 4"),
    ("means_f64", "TypeError", "Float32Array storage"),
    ("sh_short", "RangeError", "at least 4"),
    ("sh_float", "TypeError", "wrong typed array"),
    ("shape_short", "TypeError", "Invalid compact Gaussian shape attribute"),
    ("degree4", "RangeError", "shDegree must be 0, 1, 2, or 3"),
    ("slot_out", "RangeError", "Active GPU slot exceeds capacity"),
]


def bad_python(gs, case):
    s = scene_of(gs, 1, 1)
    if case == "means_vec3":
        s.means = np.zeros((1, 3), np.float32)
    elif case == "means_f64":
        s.means = np.zeros((1, 4), np.float64)
    elif case == "sh_short":
        s.sh = np.zeros((1, 1), np.uint32)
    elif case == "sh_float":
        s.sh = np.zeros((1, 4), np.float32)
    elif case == "shape_short":
        s.rotations = np.zeros(0, np.uint32)
    elif case == "degree4":
        s.sh_degree = 4
    elif case == "slot_out":
        s.active_slots = np.array([1], np.uint32)
    return s


@pytest.mark.parametrize("case,cls,message", DATA_CASES, ids=[c[0] for c in DATA_CASES])
def test_check_data_reference_errors(gs, case, cls, message):
    """GaussianData.ts:76-166, 265-318 (test/GaussianData.test.ts:33-92)."""
    gs.check_data(scene_of(gs))
    want = gs.GsRangeError if cls == "RangeError" else gs.GsTypeError
    with pytest.raises(want, match=message):
        gs.check_data(bad_python(gs, case))


@pytest.mark.skipif(node_missing, reason="node or the built N-API addon is missing")
def test_node_check_data_reference_errors():
    out = run_node("""
const ok = () => ({count: 1, shDegree: 1, means: new Float32Array(4), scalesOpacity: new Uint32Array(2),
                   rotations: new Uint32Array(1), shCoefficients: new Uint32Array(4)});
const cases = {
  means_vec3: (d) => { d.means = new Float32Array(3); },
  means_f64: (d) => { d.means = new Float64Array(4); },
  sh_short: (d) => { d.shCoefficients = new Uint32Array(1); },
  sh_float: (d) => { d.shCoefficients = new Float32Array(4); },
  shape_short: (d) => { d.rotations = new Uint32Array(0); },
  degree4: (d) => { d.shDegree = 4; },
  slot_out: (d) => { d.activeSlots = Uint32Array.from([1]); },
};
const out = {};
facade.checkData(ok());
for (const [k, f] of Object.entries(cases)) {
  const d = ok(); f(d);
  try { facade.checkData(d); out[k] = null; } catch (e) { out[k] = [e.constructor.name, e.message]; }
}
console.log(JSON.stringify(out));
""", {})
    for case, cls, message in DATA_CASES:
        got = out[case]
        assert got is not None, case
        # a short means array is a count error in the flat-array form
        want_msg = "at least 1" if case == "means_vec3" else message
        assert got[0] == cls and want_msg in got[1], (case, got)


@pytest.mark.gpu
def test_tracked_list_renders_like_the_oracle(gs, oracle, gpu_available):
    """Commits through a tracked ActiveSlots list reach the device: each
    rendered frame is bit-exact against the oracle's frame of a scene whose
    active list is the committed one; a staged, uncommitted change leaves
    the image as it was."""
    scene = gs.workload_scene("small")
    w = gs.WORKLOADS["small"]
    cam = w.camera()
    rng = np.random.default_rng(21)
    active = gs.ActiveSlots(scene.count, rng.permutation(scene.count)[: scene.count // 2].astype(np.uint32))
    data = gs.PackedScene(scene.means, scene.scales_opacity, scene.rotations, scene.sh,
                          scene.sh_degree, active.slots.copy())
    p = gs.gaussian_pass(cam, data, {"intersectionCapacity": 1 << 21})
    p.track_active(active)
    frame = gs.frame_for(cam, w.width, w.height)
    objects = gs.object_frames_for(cam)[None]

    def check():
        p.render(w.width, w.height)
        st = p.read_state()
        assert st.overflow == 0 and st.error_flags == 0
        ref_scene = gs.PackedScene(scene.means, scene.scales_opacity, scene.rotations, scene.sh,
                                   scene.sh_degree, active.slots.copy())
        ref = oracle.run_frame(ref_scene, frame, objects, capacity=1 << 21, threads=0)
        np.testing.assert_array_equal(p.read_image().view(np.uint16), ref.get("image_f16"))
        assert st.intersections == ref.counts.count
        return p.read_image().view(np.uint16).copy()

    img0 = check()
    inactive = np.setdiff1d(np.arange(scene.count), active.slots)
    active.stage_activation(inactive[:3000], active.slots[:2000].copy(), False)
    p.invalidate()
    p.render(w.width, w.height)
    np.testing.assert_array_equal(p.read_image().view(np.uint16), img0)  # nothing committed
    active.stage_activation(inactive[3000:5000], [], True)
    img1 = check()
    assert not np.array_equal(img0, img1)
    active.stage_activation([], active.slots[::3].copy(), True)
    check()
    p.dispose()
