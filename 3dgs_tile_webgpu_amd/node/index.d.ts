// Type declarations of the Node facade (index.js) over the MI355X frame
// pipeline. The names follow the reference's pass API (types.ts,
// GaussianPass.ts) for the frame path this library implements.

export type DepthSortMode = 'float32' | 'packed16';
export type AntialiasMode = 'compensated' | 'classic';
export type GaussianBoundaryPolicy = 'clip' | 'repeat-x' | 'repeat-y';
export type GaussianPassRedrawStrategy = 'always' | 'auto' | 'never';

export interface GaussianPassOptions {
  depthSortMode?: DepthSortMode;
  antialiasMode?: AntialiasMode;
  filterVariance?: number;
  boundaryPolicy?: GaussianBoundaryPolicy;
  rasterTransmittanceThreshold?: number;
  depthAlphaThreshold?: number;
  invalidDepthValue?: number;
  background?: [number, number, number, number];
  subpixelSampleCulling?: boolean;
  rasterChunkSize?: number | null;
  outputDepth?: boolean;
  rasterStats?: boolean;
  fixedSize?: { width: number; height: number };
  depthRange?: [number, number];
  intersectionCapacity?: number;
  initialIntersectionCapacity?: number;
  maxIntersectionCapacity?: number;
  intersectionGrowthFactor?: number;
  intersectionHeadroom?: number;
  intersectionReadbackInterval?: number;
  /** When frames are encoded (GaussianPass.ts:159-174); default "always". */
  redrawStrategy?: GaussianPassRedrawStrategy;
  /** Accepted for drop-in use; no part in this path's arithmetic. */
  radixBackend?: 'auto' | 'subgroup' | 'workgroup';
  dataMode?: 'managed' | 'external';
  colorSpace?: string;
  profileKernels?: boolean;
  /** Launch shape; only 16 gives tile keys comparable with the reference. */
  tileSize?: 8 | 16 | 32;
  radixBits?: 4 | 6 | 8;
  rasterBatch?: 128 | 256;
  /** Heavy raster units above which 16-pixel tiles run with two lanes per
   * pixel instead of four, as whole tiles and then row slabs (default: 1.25
   * per compute unit); never changes any output. */
  heavySplitUnits?: number;
  /** Heavy raster units above which row slabs have 4 rows instead of 8
   * (default: 16 per compute unit); never changes any output. */
  heavyFineSplitUnits?: number;
  /** Heavy raster units above which the two-lane shape is row slabs rather
   * than whole tiles (above heavySplitUnits; default: 5.5 per compute unit);
   * never changes any output. */
  heavyPairUnits?: number;
  /** Views of a batch rendered by each launch, 1..8 (default 1): every
   * stage is one kernel over the group's views; never changes any output. */
  viewGroup?: number;
  device?: number;
}

/** GaussianData's packed arrays (compact geometry + rgb8e8 SH). */
export interface GaussianDataBuffers {
  count: number;
  activeCount?: number;
  shDegree?: 0 | 1 | 2 | 3;
  /** Only rgb8e8 words are rendered by this path (float32 SH is a TypeError). */
  shFormat?: 'rgb8e8' | 'float32';
  means: Float32Array;          // count * 4
  scalesOpacity: Uint32Array;   // count * 2
  rotations: Uint32Array;       // count
  shCoefficients: Uint32Array;  // count * (shDegree + 1)^2
  activeSlots?: Uint32Array;
}

export interface CameraLike {
  projectionMatrix: { elements: ArrayLike<number> };
  matrixWorld: { elements: ArrayLike<number> };
  near: number;
  far: number;
}

export interface GaussianPassStats {
  visibleGaussianCount: number;
  intersectionCount: number;
  requestedIntersections: number;
  intersectionCapacity: number;
  overflow: boolean;
  rasterWork: { checked: number; blended: number } | null;
}

export type GaussianPassOutputState =
  'awaiting-first-result' | 'ready' | 'overflow-pending' | 'budget-exceeded' | 'allocation-failed';

export interface GaussianPassOutputStatus {
  state: GaussianPassOutputState;
  valid: boolean;
  version: number;
  confirmedAttempt: number;
  latestAttempt: number;
  requestedIntersections: number;
  intersectionCapacity: number;
  maxIntersectionCapacity: number;
  growthCount: number;
}

export interface NativeState {
  attempt: number;
  version: number;
  requested: number;
  overflow: number;
  visible: number;
  intersections: number;
  capacity: number;
  tilesX: number;
  tilesY: number;
  rasterChecked: number;
  rasterBlended: number;
  /** The part of rasterChecked in heavy units (tiles above 1024 records, chunk tasks). */
  rasterCheckedHeavy: number;
  errorFlags: number;
}

export class IntersectionCapacityPolicy {
  constructor(options: GaussianPassOptions, deviceLimit?: number);
  readonly initial: number;
  readonly maximum: number;
  next(current: number, demand: number): number | null;
}

export class GaussianPass {
  constructor(camera: CameraLike, data: GaussianDataBuffers, options?: GaussianPassOptions);
  readonly options: GaussianPassOptions;
  readonly intersectionCapacity: number;
  readonly filterVariance: number;
  readonly invalidDepthValue: number;
  readonly boundaryPolicy: GaussianBoundaryPolicy;
  readonly depthRange: [number, number] | null;
  readonly redrawStrategy: GaussianPassRedrawStrategy;
  /** Frames encoded. */
  readonly renderCount: number;
  /** Frames that reused the last output (redrawStrategy "auto" / "never"). */
  readonly cacheHitCount: number;
  camera: CameraLike;
  updateData(data: GaussianDataBuffers): void;
  /** Active slots [start, start + slots.length) and the new active count. */
  updateActive(start: number, slots: Uint32Array, activeCount: number): void;
  /** Upload each committed change of the list (its capacity is data.count). */
  trackActiveSlots(list: ActiveSlots): void;
  /** Force the next render to encode a frame. */
  invalidate(): void;
  /** The error of a failed capacity growth; no frame is encoded after it. */
  readonly allocationError: string | null;
  setFixedSize(width: number, height: number): this;
  render(width?: number, height?: number): void;
  resizeIntersections(capacity: number): void;
  warmup(): Promise<void>;
  readState(): NativeState;
  readStats(): Promise<GaussianPassStats>;
  getOutputStatus(): GaussianPassOutputStatus;
  /** rgba16float bits, row-major, width * height * 4. */
  readImage(): Uint16Array;
  readDepth(): Float32Array;
  getResources(): Record<string, Float32Array | Uint32Array>;
  dispose(): void;
}

/** Several cameras over one store in one captured launch. */
export class GaussianBatch {
  constructor(data: GaussianDataBuffers, options?: GaussianPassOptions);
  readonly capacities: number[];
  readonly sizes: [number, number][];
  render(cameras: CameraLike[], sizes: [number, number][]): void;
  readStates(): NativeState[];
  readImage(view: number): Uint16Array;
  setTuning(tuning: LaunchShape): void;
  getTuning(): LaunchShape;
  /** Times the 18 launch shapes on these views and keeps the fastest. */
  autotune(cameras: CameraLike[], sizes: [number, number][], reps?: number):
    { best: LaunchShape; table: (LaunchShape & { ms: number })[] };
  dispose(): void;
}

export interface LaunchShape { tileSize: number; radixBits: number; rasterBatch: number }

export interface CalibrationRow { shift: number; keysKept: number; psnrMean: number; psnrMin: number }

/** The learned cull of a pass or batch (no counterpart in the reference). */
export class LearnedCull {
  constructor(owner: GaussianPass | GaussianBatch, seed?: number);
  threshold: number;
  readonly enabled: boolean;
  reserve(capacity: number): void;
  collect(camera: CameraLike, width: number, height: number, rate: number, seed: number): NativeState;
  datasetSize(): number;
  readDataset(first?: number, n?: number): { features: Float32Array; targets: Float32Array };
  fitNormalization(): void;
  train(steps: number, opts?: { batch?: number; lr?: number; underWeight?: number; seed?: number }): number;
  /** W1[64][16], b1[64], W2[64][64], b2[64], w3[64], b3 (5313 floats). */
  params(): Float32Array;
  /** Training state (parameters, Adam moments and step count,
   * standardisation, threshold) as the library's checkpoint blob. */
  checkpoint(): Uint8Array;
  /** Resume from a checkpoint; throws RangeError for a blob of another
   * size, version or network, or with non-finite values. */
  restore(blob: Uint8Array): void;
  setParams(params: Float32Array): void;
  normalization(): { mu: Float32Array; invSigma: Float32Array };
  setNormalization(mu: Float32Array, invSigma: Float32Array): void;
  /** Predicted log max alpha * T for raw feature rows (16 floats each). */
  predict(features: Float32Array): Float32Array;
  gradient(features: Float32Array, targets: Float32Array, underWeight?: number):
    { grad: Float32Array; loss: number };
  enable(threshold?: number): void;
  disable(): void;
  calibrate(cameras: CameraLike[], sizes: [number, number][], targetPsnr?: number,
            shifts?: number[]): { shift: number | null; rows: CalibrationRow[] };
}

export interface ActiveSlotRange { start: number; count: number }

/** GaussianData's active slot list (GaussianData.ts:170-232). */
export class ActiveSlots {
  constructor(capacity: number, slots?: Uint32Array);
  readonly capacity: number;
  readonly count: number;
  readonly version: number;
  /** The committed list, a view of its first `count` entries. */
  readonly slots: Uint32Array;
  subscribe(listener: (range: ActiveSlotRange) => void): void;
  unsubscribe(listener: (range: ActiveSlotRange) => void): void;
  /** Removals first, then additions; with commit, publish the changed range. */
  stageActivation(added: Uint32Array, removed: Uint32Array, commit: boolean): ActiveSlotRange | null;
}

/** GaussianData's buffer checks (RangeError / TypeError as the reference). */
export function checkData(data: GaussianDataBuffers): void;

/** PSNR of two rgba16float images (raw half bits) over rgb. */
export function psnr(a: Uint16Array, b: Uint16Array): number;
export function checkObjectTables(tables: Float32Array[]): void;

export function gaussianPass(camera: CameraLike, data: GaussianDataBuffers,
                             options?: GaussianPassOptions): GaussianPass;
export function validateOptions(options: GaussianPassOptions): GaussianPassOptions;
export function objectFrames(camera: CameraLike, world?: ArrayLike<number> | null,
                             visible?: boolean): Float32Array;
export const DEFAULTS: Readonly<GaussianPassOptions>;
export const native: Record<string, (...args: any[]) => any>;
