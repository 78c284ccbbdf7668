'use strict';
// Node facade over the N-API addon (gs_napi.node, built from gs_napi.cpp).
//
// gaussianPass(camera, data, options) mirrors the reference's
// gaussianPass(renderer, camera, store, options) (createGaussianPass.ts:7-14,
// GaussianPass.ts:144-307) for the frame path: the same option names,
// defaults and validation errors (RangeError / TypeError), the capacity
// feedback of IntersectionCapacityPolicy.ts:13-101 (an overflowing frame
// publishes nothing; capacity grows and the next attempt retries), and the
// introspection calls readStats / getOutputStatus / getResources. `camera`
// is any object with three.js-style `projectionMatrix.elements`,
// `matrixWorld.elements`, `near` and `far`; `data` holds the GaussianData
// packed arrays (GaussianData.ts:10-33).
const path = require('path');

const addonPath = process.env.GS_NAPI_ADDON || path.join(__dirname, '..', '_lib', 'gs_napi.node');
const native = require(addonPath);

const DEFAULTS = Object.freeze({
  depthSortMode: 'float32',
  antialiasMode: 'compensated',
  filterVariance: 0.3,
  boundaryPolicy: 'clip',
  rasterTransmittanceThreshold: 1e-4,
  depthAlphaThreshold: 0.95,
  invalidDepthValue: 1.0,
  background: [0, 0, 0, 0],
  subpixelSampleCulling: true,
  rasterChunkSize: 8192,
  outputDepth: false,
  rasterStats: false,
  intersectionGrowthFactor: 2.0,
  intersectionHeadroom: 1.25,
  tileSize: 16,
  radixBits: 8,
  rasterBatch: 256,
});
// Options of the reference that take no part in this path's arithmetic,
// accepted (and validated where the reference validates them) for drop-in
// use: radixBackend picks a WebGPU sort kernel (the library's sort gives the
// same order), colorSpace tags the output texture, profileKernels asks for
// timestamp queries.
const ACCEPTED = ['redrawStrategy', 'radixBackend', 'dataMode', 'colorSpace', 'profileKernels',
  'intersectionReadbackInterval'];
const KNOWN = new Set([...Object.keys(DEFAULTS), ...ACCEPTED, 'fixedSize', 'depthRange',
  'intersectionCapacity', 'initialIntersectionCapacity', 'maxIntersectionCapacity', 'device',
  'heavySplitUnits', 'heavyFineSplitUnits', 'heavyPairUnits', 'viewGroup']);
const BOUNDARY = { clip: 0, 'repeat-x': 1, 'repeat-y': 2 };
const MAX_CAPACITY = 2 ** 28;
const RESOURCE_NAMES = ['projected_mean', 'projected_conic', 'projected_color', 'visible_offsets',
  'depth_records', 'tile_counts', 'ordered_counts', 'intersection_offsets', 'sorted_records',
  'tile_offsets'];

const isSafeInt = (v) => Number.isSafeInteger(v);
// Node 12 has no nullish-coalescing operator.
const dflt = (v, d) => (v === undefined || v === null ? d : v);

// IntersectionCapacityPolicy.ts:13-101 for the options this API takes.
class IntersectionCapacityPolicy {
  constructor(options, deviceLimit = MAX_CAPACITY) {
    for (const name of ['intersectionCapacity', 'initialIntersectionCapacity', 'maxIntersectionCapacity']) {
      const v = options[name];
      if (v !== undefined && (!isSafeInt(v) || v < 1)) throw new RangeError(`${name} must be a positive safe integer`);
    }
    const fixed = options.intersectionCapacity;
    if (fixed !== undefined && (options.initialIntersectionCapacity !== undefined ||
        options.maxIntersectionCapacity !== undefined)) {
      throw new RangeError('intersectionCapacity cannot be combined with dynamic capacity options');
    }
    this.maximum = fixed !== undefined ? fixed : dflt(options.maxIntersectionCapacity, deviceLimit);
    if (this.maximum > deviceLimit) throw new RangeError('Intersection capacity exceeds device buffer or dispatch limits');
    const initial = fixed !== undefined ? fixed : options.initialIntersectionCapacity;
    this.initial = initial !== undefined ? initial : Math.min(262144, this.maximum);
    if (this.initial > this.maximum) throw new RangeError('initialIntersectionCapacity exceeds maximum capacity');
    this.growth = dflt(options.intersectionGrowthFactor, 2.0);
    this.headroom = dflt(options.intersectionHeadroom, 1.25);
    if (!Number.isFinite(this.growth) || this.growth <= 1) throw new RangeError('intersectionGrowthFactor must exceed 1');
    if (!Number.isFinite(this.headroom) || this.headroom < 1) throw new RangeError('intersectionHeadroom must be at least 1');
    this.interval = dflt(options.intersectionReadbackInterval, 100);
    if (!Number.isFinite(this.interval) || this.interval < 0) throw new RangeError('intersectionReadbackInterval must be nonnegative');
  }

  // The capacity after an overflow with `demand` records, or null when the
  // demand exceeds the maximum (the reference's budget-exceeded state).
  next(current, demand) {
    return native.capacityNext(current, demand, this.maximum, this.growth, this.headroom);
  }
}

// GaussianPass.ts:157-260 and the render-domain setters (:386-466) for the
// supported option subset, with the reference's error classes (every value
// check there is a RangeError). depthSortMode has no run-time check in the
// reference (its types reject other values) and unknown options are
// rejected: both are TypeErrors here.
function validateOptions(options) {
  for (const k of Object.keys(options)) {
    if (!KNOWN.has(k)) throw new TypeError(`unsupported option: ${k}`);
  }
  const o = { ...DEFAULTS, ...options };
  const oneOf = (v, d, allowed) => allowed.includes(v === undefined ? d : v);
  if (o.depthSortMode !== 'float32' && o.depthSortMode !== 'packed16') throw new TypeError('depthSortMode must be float32 or packed16');
  if (o.antialiasMode !== 'compensated' && o.antialiasMode !== 'classic') {
    throw new RangeError('antialiasMode must be either "compensated" or "classic"');
  }
  if (!oneOf(o.redrawStrategy, 'always', ['always', 'auto', 'never'])) {
    throw new RangeError('redrawStrategy must be "always", "auto", or "never"');
  }
  if (!oneOf(o.radixBackend, 'auto', ['auto', 'subgroup', 'workgroup'])) {
    throw new RangeError('radixBackend must be "auto", "subgroup", or "workgroup"');
  }
  if (!oneOf(o.dataMode, 'managed', ['managed', 'external'])) throw new RangeError('dataMode must be "managed" or "external"');
  if (!(o.boundaryPolicy in BOUNDARY)) throw new RangeError('boundaryPolicy must be "clip", "repeat-x", or "repeat-y"');
  const t = o.rasterTransmittanceThreshold;
  if (!(Number.isFinite(t) && t > 0 && t < 1)) throw new RangeError('rasterTransmittanceThreshold must be finite and in (0, 1)');
  const fv = o.filterVariance;
  if (!(Number.isFinite(fv) && fv >= 0 && Number.isFinite(Math.fround(fv)))) {
    throw new RangeError('filterVariance must be finite and nonnegative');
  }
  const iv = o.invalidDepthValue;
  if (typeof iv !== 'number' || Number.isNaN(iv) || (Number.isFinite(iv) && !Number.isFinite(Math.fround(iv)))) {
    throw new RangeError('invalidDepthValue must be representable as float32 and cannot be NaN');
  }
  if (!Array.isArray(o.background) || o.background.length !== 4 || !o.background.every(Number.isFinite)) {
    throw new RangeError('background must be four finite numbers');
  }
  const cs = o.rasterChunkSize;
  if (cs !== null && (!Number.isInteger(cs) || cs < 256 || cs % 256 !== 0)) {
    throw new RangeError('rasterChunkSize must be a multiple of 256 and at least 256');
  }
  const da = o.depthAlphaThreshold;
  if (!(Number.isFinite(da) && da >= 0 && da <= 1)) throw new RangeError('depthAlphaThreshold must be finite and in [0, 1]');
  if (o.fixedSize !== undefined) {
    const { width, height } = o.fixedSize;
    if (!(isSafeInt(width) && isSafeInt(height) && width >= 1 && height >= 1)) {
      throw new RangeError('fixedSize width/height must be positive safe integers');
    }
  }
  if (o.depthRange !== undefined && o.depthRange !== null) {
    const [n, f] = o.depthRange;
    if (!(Number.isFinite(n) && Number.isFinite(f) && n > 0 && Math.fround(n) > 0 && f > n &&
          Number.isFinite(Math.fround(f)) && Math.fround(n) < Math.fround(f))) {
      throw new RangeError('depthRange must contain finite positive increasing float32 values');
    }
  }
  if (![8, 16, 32].includes(o.tileSize) || ![4, 6, 8].includes(o.radixBits) || ![128, 256].includes(o.rasterBatch)) {
    throw new RangeError('tileSize must be 8, 16 or 32, radixBits 4, 6 or 8, rasterBatch 128 or 256');
  }
  if (o.viewGroup !== undefined && !(Number.isInteger(o.viewGroup) && o.viewGroup >= 1 && o.viewGroup <= 8)) {
    throw new RangeError('viewGroup must be an integer in [1, 8]');
  }
  for (const name of ['heavySplitUnits', 'heavyFineSplitUnits', 'heavyPairUnits']) {
    const v = o[name];
    if (v !== undefined && !(Number.isInteger(v) && v >= 0 && v < 0xffffffff)) {
      throw new RangeError(name + ' must be a non-negative integer');
    }
  }
  return o;
}

function nativeConfig(o, policy) {
  return {
    device: dflt(o.device, 0),
    depthSortMode: o.depthSortMode === 'float32' ? 0 : 1,
    antialiasMode: o.antialiasMode === 'compensated' ? 0 : 1,
    subpixelSampleCulling: o.subpixelSampleCulling ? 1 : 0,
    rasterChunkSize: o.rasterChunkSize === null ? 0 : o.rasterChunkSize,
    intersectionCapacity: policy.initial,
    maxIntersectionCapacity: policy.maximum,
    outputDepth: o.outputDepth ? 1 : 0,
    transmittanceThreshold: o.rasterTransmittanceThreshold,
    depthAlphaThreshold: o.depthAlphaThreshold,
    invalidDepthValue: o.invalidDepthValue,
    background: o.background,
    rasterStats: o.rasterStats ? 1 : 0,
    tileSize: o.tileSize,
    radixBits: o.radixBits,
    rasterBatch: o.rasterBatch,
    heavySplitUnits: o.heavySplitUnits === undefined ? 0xffffffff : o.heavySplitUnits,
    heavyFineSplitUnits: o.heavyFineSplitUnits === undefined ? 0xffffffff : o.heavyFineSplitUnits,
    heavyPairUnits: o.heavyPairUnits === undefined ? 0xffffffff : o.heavyPairUnits,
    viewGroup: dflt(o.viewGroup, 1),
  };
}

// ObjectFrameState.writeCloud for one cloud (ObjectFrameState.ts:97-120).
function objectFrames(camera, world = null, visible = true) {
  const w = world ? Float64Array.from(world) : Float64Array.of(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1);
  return native.objectFrame(Float64Array.from(camera.matrixWorld.elements), w, visible);
}

// Every camera of a batch brings its own object table (40 floats per
// object); the library reads as many rows from each as setObjects set, so
// they must all have the first one's length.
function checkObjectTables(tables) {
  const n = tables.length ? tables[0].length : 0;
  for (const t of tables) {
    if (!(t instanceof Float32Array) || t.length % 40 !== 0 || t.length !== n) {
      throw new RangeError("every camera's object table must have the same rows");
    }
  }
}

const sameArray = (a, b) => a !== null && b !== null && a.length === b.length && a.every((v, i) => v === b[i]);

// The store's active slot list (GaussianData.ts:86-97, 170-232): a compact
// list of active slots with each slot's ordinal in it. Staged removals move
// the list's last slot into the removed slot's place, staged additions
// append; a commit swaps the front and back lists, bumps the version and
// hands every subscriber the changed ordinal range, which
// GaussianPass.trackActiveSlots uploads (ProjectionStage.updateActiveSlots,
// ProjectionStage.ts:298-329).
class ActiveSlots {
  constructor(capacity, slots) {
    if (!Number.isInteger(capacity) || capacity <= 0) {
      throw new RangeError('GaussianData count must be a positive integer');
    }
    this.capacity = capacity;
    this._values = new Uint32Array(capacity);
    this._ordinals = new Int32Array(capacity).fill(-1);
    this._stagedValues = new Uint32Array(capacity);
    this._stagedOrdinals = new Int32Array(capacity);
    this.count = 0;
    this.version = 0;
    this._stagedCount = 0;
    this._staging = false;
    this._first = Infinity;
    this._last = -1;
    this._listeners = new Set();
    if (slots) this.stageActivation(slots, new Uint32Array(0), true);
  }

  get slots() { return this._values.subarray(0, this.count); }

  subscribe(listener) { this._listeners.add(listener); }
  unsubscribe(listener) { this._listeners.delete(listener); }

  stageActivation(added, removed, commit) {
    if (!this._staging && added.length === 0 && removed.length === 0) return null;
    for (const slot of added) {
      if (slot >= this.capacity) throw new RangeError('Active GPU slot exceeds capacity');
    }
    for (const slot of removed) {
      if (slot >= this.capacity) throw new RangeError('Removed GPU slot exceeds capacity');
    }
    if (!this._staging) {
      this._stagedValues.set(this._values.subarray(0, this.count));
      this._stagedOrdinals.set(this._ordinals);
      this._stagedCount = this.count;
      this._first = Infinity;
      this._last = -1;
      this._staging = true;
    }
    // Removals come before additions across delta batches.
    for (const slot of removed) {
      const ordinal = this._stagedOrdinals[slot];
      if (ordinal < 0) continue;
      const last = this._stagedValues[--this._stagedCount];
      this._stagedValues[ordinal] = last;
      this._stagedOrdinals[last] = ordinal;
      this._stagedOrdinals[slot] = -1;
      this._first = Math.min(this._first, ordinal);
      this._last = Math.max(this._last, ordinal);
    }
    for (const slot of added) {
      if (this._stagedOrdinals[slot] >= 0) continue;
      const ordinal = this._stagedCount++;
      this._stagedValues[ordinal] = slot;
      this._stagedOrdinals[slot] = ordinal;
      this._first = Math.min(this._first, ordinal);
      this._last = Math.max(this._last, ordinal);
    }
    if (!commit) return null;
    const values = this._values;
    this._values = this._stagedValues;
    this._stagedValues = values;
    const ordinals = this._ordinals;
    this._ordinals = this._stagedOrdinals;
    this._stagedOrdinals = ordinals;
    this.count = this._stagedCount;
    this._staging = false;
    this.version++;
    const start = Math.min(this._first, this.count);
    const end = Math.min(this._last + 1, this.count);
    const range = { start, count: Math.max(0, end - start) };
    for (const listener of this._listeners) listener(range);
    return range;
  }
}

// GaussianData's buffer checks (GaussianData.ts:76-166, 265-318) for the
// compact + rgb8e8 contract this path renders, with its error classes and
// messages.
function checkData(data) {
  const count = data.count;
  if (!Number.isInteger(count) || count <= 0) {
    throw new RangeError('GaussianData count must be a positive integer');
  }
  const shDegree = dflt(data.shDegree, 3);
  if (!Number.isInteger(shDegree) || shDegree < 0 || shDegree > 3) {
    throw new RangeError('GaussianData shDegree must be 0, 1, 2, or 3');
  }
  const shFormat = dflt(data.shFormat, 'rgb8e8');
  if (shFormat !== 'float32' && shFormat !== 'rgb8e8') {
    throw new RangeError('GaussianData shFormat must be float32 or rgb8e8');
  }
  if (!(data.means instanceof Float32Array)) {
    throw new TypeError('GaussianData means must use Float32Array storage');
  }
  if (data.means.length < count * 4) {
    throw new RangeError(`GaussianData means has ${Math.floor(data.means.length / 4)} items; at least ${count} are required`);
  }
  if (!(data.scalesOpacity instanceof Uint32Array) || data.scalesOpacity.length < count * 2 ||
      !(data.rotations instanceof Uint32Array) || data.rotations.length < count) {
    throw new TypeError('Invalid compact Gaussian shape attribute');
  }
  if (!(data.shCoefficients instanceof Uint32Array) || shFormat !== 'rgb8e8') {
    throw new TypeError(`GaussianData ${shFormat} shCoefficients use the wrong typed array`);
  }
  const need = count * (shDegree + 1) * (shDegree + 1);
  if (data.shCoefficients.length < need) {
    throw new RangeError(`GaussianData shCoefficients has ${data.shCoefficients.length} items; at least ${need} are required`);
  }
  if (data.activeSlots) {
    const n = Math.min(dflt(data.activeCount, data.activeSlots.length), data.activeSlots.length);
    for (let i = 0; i < n; i++) {
      if (data.activeSlots[i] >= count) throw new RangeError('Active GPU slot exceeds capacity');
    }
  }
}

class GaussianPass {
  constructor(camera, data, options = {}) {
    this.options = validateOptions(options);
    this.policy = new IntersectionCapacityPolicy(this.options);
    this._ctx = native.create(nativeConfig(this.options, this.policy));
    this.camera = camera;
    this._capacity = this.policy.initial;
    this.growthCount = 0;
    this.latestAttempt = 0;
    // redrawStrategy (GaussianPass.ts:159-174, 684-693, 776-790, 1187+):
    // "always" encodes every frame; "auto" only when the frame's inputs
    // (size, depth range, projection, camera pose, data) changed or the pass
    // was invalidated; "never" only when invalidated.
    this.redrawStrategy = dflt(this.options.redrawStrategy, 'always');
    this.renderCount = 0;
    this.cacheHitCount = 0;
    this._dirty = true;
    this._snapshot = null;
    this.allocationError = null;  // set when growth failed (allocation-failed)
    this.width = 0;
    this.height = 0;
    this.updateData(data);
  }

  // Force the next render to encode a frame (GaussianPass.ts:604-606).
  invalidate() { this._dirty = true; }

  get intersectionCapacity() { return this._capacity; }
  get filterVariance() { return this.options.filterVariance; }
  get invalidDepthValue() { return this.options.invalidDepthValue; }
  get boundaryPolicy() { return this.options.boundaryPolicy; }
  get depthRange() { return dflt(this.options.depthRange, null); }

  // TiledGaussianPipeline.updateData (TiledGaussianPipeline.ts:203-218).
  updateData(data) {
    checkData(data);
    native.setData(this._ctx, data);
    this.data = data;
    this._dirty = true;
  }

  // Upload every committed change of an ActiveSlots list as it is committed
  // (GaussianData.subscribeActiveSlots, ProjectionStage.ts:298-329).
  trackActiveSlots(list) {
    if (list.capacity !== this.data.count) throw new RangeError('Active GPU slot exceeds capacity');
    const listener = (range) => {
      this.updateActive(range.start, list.slots.subarray(range.start, range.start + range.count), list.count);
    };
    list.subscribe(listener);
    if (!this._tracked) this._tracked = [];
    this._tracked.push([list, listener]);
  }

  // Active-slot update of [start, start + slots.length)
  // (GaussianData.ts:170-232; ProjectionStage.updateActiveSlots :298-329).
  updateActive(start, slots, activeCount) {
    native.updateActive(this._ctx, start, slots, activeCount);
    this._dirty = true;
  }

  setFixedSize(width, height) {
    if (!(Number.isInteger(width) && Number.isInteger(height) && width >= 1 && height >= 1)) {
      throw new RangeError('fixedSize must be positive integers');
    }
    this.options.fixedSize = { width, height };
    return this;
  }

  _frame(width, height) {
    const range = this.options.depthRange;
    return {
      projection: Float32Array.from(this.camera.projectionMatrix.elements),
      width,
      height,
      near: range ? range[0] : this.camera.near,
      far: range ? range[1] : this.camera.far,
      filterVariance: this.options.filterVariance,
      boundaryAxis: BOUNDARY[this.options.boundaryPolicy],
      attempt: ++this.latestAttempt,
    };
  }

  // prepareFrame + render (TiledGaussianPipeline.ts:286-341). On overflow the
  // capacity grows by the policy and the frame is retried once.
  render(width, height) {
    const fixed = this.options.fixedSize;
    if (width === undefined) {
      width = fixed ? fixed.width : this.width;
      height = fixed ? fixed.height : this.height;
    }
    if (!width || !height) throw new RangeError('render size is required');
    if (this.allocationError !== null) return;  // GaussianPass.ts:683
    const table = objectFrames(this.camera);
    if (width !== this.width || height !== this.height) this._dirty = true;
    let snapshot = null;
    if (this.redrawStrategy === 'auto') {
      const range = this.options.depthRange;
      snapshot = [width, height, range ? range[0] : this.camera.near, range ? range[1] : this.camera.far,
        ...Float32Array.from(this.camera.projectionMatrix.elements), ...table];
      if (!sameArray(snapshot, this._snapshot)) this._dirty = true;
    }
    if (this.redrawStrategy !== 'always' && !this._dirty) {
      this.cacheHitCount++;
      return;
    }
    this.width = width;
    this.height = height;
    native.setObjects(this._ctx, table);
    native.render(this._ctx, this._frame(width, height));
    this.renderCount++;
    this._dirty = false;
    this._snapshot = snapshot;
    const st = native.readState(this._ctx);
    if (st.overflow) {
      const next = this.policy.next(this._capacity, st.requested);
      if (next !== null && next > this._capacity) {
        try {
          this.resizeIntersections(next);
        } catch (e) {  // GaussianPass.ts:1010-1017 (failAllocation)
          this.allocationError = String(e);
          return;
        }
        native.render(this._ctx, this._frame(width, height));
      }
    }
  }

  resizeIntersections(capacity) {
    native.resizeIntersections(this._ctx, capacity);
    this._capacity = capacity;
    this.growthCount++;
  }

  warmup() {
    return Promise.resolve();
  }

  readState() { return native.readState(this._ctx); }

  // GaussianPassStats (types.ts:94-103).
  readStats() {
    const st = native.readState(this._ctx);
    return Promise.resolve({
      visibleGaussianCount: st.visible,
      intersectionCount: st.intersections,
      requestedIntersections: st.requested,
      intersectionCapacity: st.capacity,
      overflow: st.overflow !== 0,
      rasterWork: this.options.rasterStats ? { checked: st.rasterChecked, blended: st.rasterBlended } : null,
    });
  }

  // getOutputStatus (GaussianPass.ts:989-1008), synchronous form.
  getOutputStatus() {
    const st = native.readState(this._ctx);
    let state = 'ready';
    if (this.allocationError !== null) state = 'allocation-failed';
    else if (st.attempt === 0) state = 'awaiting-first-result';
    else if (st.overflow) state = this.policy.next(this._capacity, st.requested) === null ? 'budget-exceeded' : 'overflow-pending';
    return {
      state,
      valid: st.version !== 0 && !st.overflow && this.allocationError === null,
      version: st.version,
      confirmedAttempt: st.attempt,
      latestAttempt: this.latestAttempt,
      requestedIntersections: st.requested,
      intersectionCapacity: this._capacity,
      maxIntersectionCapacity: this.policy.maximum,
      growthCount: this.growthCount,
    };
  }

  // The rgba16float colour target as raw half-float bits, row-major.
  readImage() {
    const out = new Uint16Array(this.width * this.height * 4);
    native.readImage(this._ctx, out, null);
    return out;
  }

  readDepth() {
    const out = new Float32Array(this.width * this.height);
    native.readImage(this._ctx, null, out);
    return out;
  }

  // getResources (TiledGaussianPipeline.ts:391-406) as host copies.
  getResources() {
    const r = {};
    for (const name of RESOURCE_NAMES) {
      const buf = native.readDebug(this._ctx, name);
      r[name] = name.startsWith('projected_') ? new Float32Array(buf) : new Uint32Array(buf);
    }
    return r;
  }

  dispose() {
    for (const [list, listener] of this._tracked || []) list.unsubscribe(listener);
    this._tracked = [];
    if (this._ctx) native.destroy(this._ctx);
    this._ctx = null;
  }
}

// Several cameras over one store in one captured launch (the reference's
// several passes sharing a store through dataMode "external",
// docs/render-extensions.md:5-47). Each view has its own capacity and
// grows on overflow like a single pass; the batch then re-renders.
class GaussianBatch {
  constructor(data, options = {}) {
    this.options = validateOptions(options);
    this.policy = new IntersectionCapacityPolicy(this.options);
    this._ctx = native.create(nativeConfig(this.options, this.policy));
    this.capacities = [];
    this.sizes = [];
    this.attempt = 0;
    checkData(data);
    native.setData(this._ctx, data);
  }

  _frames(cameras, sizes) {
    const range = this.options.depthRange;
    this.attempt++;
    return cameras.map((cam, i) => ({
      projection: Float32Array.from(cam.projectionMatrix.elements),
      width: sizes[i][0],
      height: sizes[i][1],
      near: range ? range[0] : cam.near,
      far: range ? range[1] : cam.far,
      filterVariance: this.options.filterVariance,
      boundaryAxis: BOUNDARY[this.options.boundaryPolicy],
      attempt: this.attempt,
      objectFrames: objectFrames(cam),
    }));
  }

  _prepare(cameras, sizes) {
    if (cameras.length !== sizes.length || cameras.length === 0) {
      throw new RangeError('one size per camera is required');
    }
    const frames = this._frames(cameras, sizes);
    checkObjectTables(frames.map((f) => f.objectFrames));
    // The object table's size comes from the context; each view brings its own frames.
    native.setObjects(this._ctx, frames[0].objectFrames);
    return frames;
  }

  render(cameras, sizes) {
    const frames = this._prepare(cameras, sizes);
    this.sizes = sizes.map((s) => [s[0], s[1]]);
    while (this.capacities.length < frames.length) this.capacities.push(this.policy.initial);
    for (let round = 0; round < 8; round++) {
      native.renderBatch(this._ctx, frames);
      let grown = false;
      native.readBatchState(this._ctx, frames.length).forEach((st, v) => {
        if (!st.overflow) return;
        const next = this.policy.next(this.capacities[v], st.requested);
        if (next === null) throw new RangeError('intersection demand exceeds maximum');
        native.resizeBatchIntersections(this._ctx, v, next);
        this.capacities[v] = next;
        grown = true;
      });
      if (!grown) return;
    }
    throw new Error('batch capacity did not settle');
  }

  readStates() { return native.readBatchState(this._ctx, this.sizes.length); }

  readImage(view) {
    const [w, h] = this.sizes[view];
    const out = new Uint16Array(w * h * 4);
    native.readImage(this._ctx, out, null, view);
    return out;
  }

  setTuning(tuning) { native.setTuning(this._ctx, tuning); }

  getTuning() { return native.getTuning(this._ctx); }

  // Times every launch shape (tile 8/16/32 x digits 4/6/8 x raster batch
  // 128/256) on these views, keeps the fastest and returns {best, table}.
  autotune(cameras, sizes, reps = 3) {
    const frames = this._prepare(cameras, sizes);
    this.sizes = sizes.map((s) => [s[0], s[1]]);
    const result = native.autotune(this._ctx, frames, reps);
    this.capacities = native.readBatchState(this._ctx, frames.length).map((st) => st.capacity);
    return result;
  }

  dispose() {
    if (this._ctx) native.destroy(this._ctx);
    this._ctx = null;
  }
}

// PSNR of two rgba16float images (raw half bits) over their rgb channels.
function halfToFloat(h) {
  const s = h & 0x8000 ? -1 : 1;
  const e = (h >> 10) & 31;
  const m = h & 1023;
  if (e === 0) return s * m * 2 ** -24;
  if (e === 31) return m ? NaN : s * Infinity;
  return s * (1 + m / 1024) * 2 ** (e - 15);
}

function psnr(a, b) {
  let sum = 0;
  let n = 0;
  for (let i = 0; i < a.length; i++) {
    if (i % 4 === 3) continue;
    const d = halfToFloat(a[i]) - halfToFloat(b[i]);
    sum += d * d;
    n++;
  }
  const mse = sum / n;
  return mse === 0 ? Infinity : 10 * Math.log10(1 / mse);
}

const CULL_PARAMS = 5313;

// The learned cull of one pass or batch (no counterpart in the reference;
// its nearest mechanisms are the opacity and subpixel culls,
// ProjectionStage.ts:533, 572-587): a 16-64-64-1 MLP on the matrix cores,
// fused into projection, trained by the library's own kernels on targets
// the plain rasterizer records. The plain path is its fallback (disable())
// and the yardstick of its image loss (calibrate()).
class LearnedCull {
  constructor(owner, seed = 0) {
    this.owner = owner;
    this.threshold = 1e-4;
    this.enabled = false;
    native.cullInit(owner._ctx, seed);
  }

  reserve(capacity) { native.cullReserve(this.owner._ctx, capacity); }

  // Records one plain frame of `camera` and appends a hashed fraction `rate`
  // of its visible Gaussians; an overflowing frame grows the owner's
  // capacity and is recorded again.
  collect(camera, width, height, rate, seed) {
    const ctx = this.owner._ctx;
    const table = objectFrames(camera);
    native.setObjects(ctx, table);
    const range = this.owner.options.depthRange;
    for (let i = 0; i < 6; i++) {
      const st = native.cullCollect(ctx, {
        projection: Float32Array.from(camera.projectionMatrix.elements), width, height,
        near: range ? range[0] : camera.near, far: range ? range[1] : camera.far,
        filterVariance: this.owner.options.filterVariance,
        boundaryAxis: BOUNDARY[this.owner.options.boundaryPolicy], attempt: 1, objectFrames: table,
      }, rate, seed);
      if (!st.overflow) return st;
      const next = this.owner.policy.next(st.capacity, st.requested);
      if (next === null) throw new RangeError('intersection demand exceeds maximum');
      native.resizeIntersections(ctx, next);
      if (this.owner._capacity !== undefined) this.owner._capacity = next;
    }
    throw new Error('collect frame kept overflowing');
  }

  datasetSize() { return native.cullDatasetSize(this.owner._ctx); }
  readDataset(first = 0, n = undefined) {
    const count = n === undefined ? this.datasetSize() - first : n;
    return native.cullReadDataset(this.owner._ctx, first, count);
  }
  fitNormalization() { native.cullFitNormalization(this.owner._ctx); }
  train(steps, { batch = 65536, lr = 2e-3, underWeight = 4.0, seed = 0 } = {}) {
    return native.cullTrain(this.owner._ctx, steps, batch, lr, underWeight, seed);
  }
  params() { return native.cullGetParams(this.owner._ctx); }
  setParams(p) { native.cullSetParams(this.owner._ctx, p); }
  // Training checkpoint (parameters, Adam moments and steps, standardisation,
  // threshold) as a Uint8Array, and resuming from one.
  checkpoint() { return native.cullSave(this.owner._ctx); }
  restore(blob) {
    native.cullLoad(this.owner._ctx, blob);
    this.threshold = new DataView(blob.buffer, blob.byteOffset, blob.byteLength).getFloat32(24, true);
  }
  normalization() { return native.cullGetNormalization(this.owner._ctx); }
  setNormalization(mu, invSigma) { native.cullSetNormalization(this.owner._ctx, mu, invSigma); }
  predict(features) { return native.cullEval(this.owner._ctx, features); }
  gradient(features, targets, underWeight = 4.0) {
    return native.cullGradient(this.owner._ctx, features, targets, underWeight);
  }

  enable(threshold = undefined) {
    if (threshold !== undefined) this.threshold = threshold;
    native.cullEnable(this.owner._ctx, true, this.threshold);
    this.enabled = true;
    if (this.owner.invalidate) this.owner.invalidate();
  }

  disable() {
    native.cullEnable(this.owner._ctx, false, this.threshold);
    this.enabled = false;
    if (this.owner.invalidate) this.owner.invalidate();
  }

  // Moves the output bias by the lowest shift (most culling) whose culled
  // frames of these cameras (a GaussianBatch owner) keep a mean PSNR of at
  // least targetPsnr against their plain frames. Returns {shift, rows}; shift
  // is null (parameters unchanged, cull disabled) when no shift reaches it.
  calibrate(cameras, sizes, targetPsnr = 46, shifts = [-3, -2.5, -2, -1.875, -1.75, -1.625, -1.5,
    -1.375, -1.25, -1.125, -1, -0.75, -0.5, -0.25, 0, 0.5, 1, 1.5, 2, 3, 4, 6, 8]) {
    const b = this.owner;
    const was = this.enabled;
    this.disable();
    b.render(cameras, sizes);
    const plain = cameras.map((c, v) => b.readImage(v));
    const keys = b.readStates().reduce((a, st) => a + st.intersections, 0);
    const p0 = this.params();
    const rows = [];
    let chosen = null;
    for (const shift of shifts) {
      const p = Float32Array.from(p0);
      p[CULL_PARAMS - 1] += shift;
      this.setParams(p);
      this.enable();
      b.render(cameras, sizes);
      const kept = b.readStates().reduce((a, st) => a + st.intersections, 0) / Math.max(1, keys);
      const ps = cameras.map((c, v) => psnr(b.readImage(v), plain[v]));
      const mean = ps.reduce((a, x) => a + x, 0) / ps.length;
      rows.push({ shift, keysKept: kept, psnrMean: mean, psnrMin: Math.min(...ps) });
      if (mean >= targetPsnr) {
        chosen = shift;
        break;
      }
    }
    const p = Float32Array.from(p0);
    if (chosen !== null) p[CULL_PARAMS - 1] += chosen;
    this.setParams(p);
    if (!was || chosen === null) this.disable();
    return { shift: chosen, rows };
  }
}

function gaussianPass(camera, data, options = {}) {
  return new GaussianPass(camera, data, options);
}

module.exports = {
  gaussianPass, GaussianPass, GaussianBatch, IntersectionCapacityPolicy, LearnedCull, ActiveSlots,
  validateOptions, checkData, objectFrames, checkObjectTables, psnr, DEFAULTS, native,
};
