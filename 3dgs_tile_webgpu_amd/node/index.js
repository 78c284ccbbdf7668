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
const KNOWN = new Set([...Object.keys(DEFAULTS), 'fixedSize', 'depthRange', 'intersectionCapacity',
  'initialIntersectionCapacity', 'maxIntersectionCapacity', 'device', 'heavySplitUnits',
  'heavyFineSplitUnits']);
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
  }

  // The capacity after an overflow with `demand` records, or null when the
  // demand exceeds the maximum (the reference's budget-exceeded state).
  next(current, demand) {
    return native.capacityNext(current, demand, this.maximum, this.growth, this.headroom);
  }
}

// GaussianPass.ts:157-260 checks for the supported option subset.
function validateOptions(options) {
  for (const k of Object.keys(options)) {
    if (!KNOWN.has(k)) throw new TypeError(`unsupported option: ${k}`);
  }
  const o = { ...DEFAULTS, ...options };
  if (o.depthSortMode !== 'float32' && o.depthSortMode !== 'packed16') throw new TypeError('depthSortMode must be float32 or packed16');
  if (o.antialiasMode !== 'compensated' && o.antialiasMode !== 'classic') throw new TypeError('antialiasMode must be compensated or classic');
  if (!(o.boundaryPolicy in BOUNDARY)) throw new RangeError('boundaryPolicy must be clip, repeat-x or repeat-y');
  const t = o.rasterTransmittanceThreshold;
  if (!(Number.isFinite(t) && t > 0 && t < 1)) throw new RangeError('rasterTransmittanceThreshold must be finite and in (0, 1)');
  if (!(Number.isFinite(o.filterVariance) && o.filterVariance >= 0)) throw new RangeError('filterVariance must be finite and nonnegative');
  if (!Array.isArray(o.background) || o.background.length !== 4 || !o.background.every(Number.isFinite)) {
    throw new RangeError('background must be four finite numbers');
  }
  const cs = o.rasterChunkSize;
  if (cs !== null && (!Number.isInteger(cs) || cs < 256 || cs % 256 !== 0)) {
    throw new RangeError('rasterChunkSize must be a multiple of 256 and at least 256');
  }
  if (!(o.depthAlphaThreshold >= 0 && o.depthAlphaThreshold <= 1)) throw new RangeError('depthAlphaThreshold must be in [0, 1]');
  if (o.fixedSize !== undefined) {
    const { width, height } = o.fixedSize;
    if (!(Number.isInteger(width) && Number.isInteger(height) && width >= 1 && height >= 1)) {
      throw new RangeError('fixedSize must be positive integers');
    }
  }
  if (o.depthRange !== undefined) {
    const [n, f] = o.depthRange;
    if (!(n > 0 && n < f && Number.isFinite(f))) throw new RangeError('depthRange must be positive and increasing');
  }
  if (![8, 16, 32].includes(o.tileSize) || ![4, 6, 8].includes(o.radixBits) || ![128, 256].includes(o.rasterBatch)) {
    throw new RangeError('tileSize must be 8, 16 or 32, radixBits 4, 6 or 8, rasterBatch 128 or 256');
  }
  for (const name of ['heavySplitUnits', 'heavyFineSplitUnits']) {
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
  };
}

// ObjectFrameState.writeCloud for one cloud (ObjectFrameState.ts:97-120).
function objectFrames(camera, world = null, visible = true) {
  const w = world ? Float64Array.from(world) : Float64Array.of(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1);
  return native.objectFrame(Float64Array.from(camera.matrixWorld.elements), w, visible);
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
    this.renderCount = 0;
    this.width = 0;
    this.height = 0;
    this.updateData(data);
  }

  get intersectionCapacity() { return this._capacity; }
  get filterVariance() { return this.options.filterVariance; }
  get invalidDepthValue() { return this.options.invalidDepthValue; }
  get boundaryPolicy() { return this.options.boundaryPolicy; }
  get depthRange() { return dflt(this.options.depthRange, null); }

  // TiledGaussianPipeline.updateData (TiledGaussianPipeline.ts:203-218).
  updateData(data) {
    native.setData(this._ctx, data);
    this.data = data;
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
    this.width = width;
    this.height = height;
    native.setObjects(this._ctx, objectFrames(this.camera));
    native.render(this._ctx, this._frame(width, height));
    this.renderCount++;
    const st = native.readState(this._ctx);
    if (st.overflow) {
      const next = this.policy.next(this._capacity, st.requested);
      if (next !== null && next > this._capacity) {
        this.resizeIntersections(next);
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
    if (st.attempt === 0) state = 'awaiting-first-result';
    else if (st.overflow) state = this.policy.next(this._capacity, st.requested) === null ? 'budget-exceeded' : 'overflow-pending';
    return {
      state,
      valid: st.version !== 0 && !st.overflow,
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

  render(cameras, sizes) {
    if (cameras.length !== sizes.length || cameras.length === 0) {
      throw new RangeError('one size per camera is required');
    }
    const frames = this._frames(cameras, sizes);
    // The object table's size comes from the context; each view brings its own frames.
    native.setObjects(this._ctx, frames[0].objectFrames);
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

  dispose() {
    if (this._ctx) native.destroy(this._ctx);
    this._ctx = null;
  }
}

function gaussianPass(camera, data, options = {}) {
  return new GaussianPass(camera, data, options);
}

module.exports = {
  gaussianPass, GaussianPass, GaussianBatch, IntersectionCapacityPolicy, validateOptions, objectFrames,
  DEFAULTS, native,
};
