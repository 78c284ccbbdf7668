// Type declarations of the Node facade (index.js) over the MI355X frame
// pipeline. The names follow the reference's pass API (types.ts,
// GaussianPass.ts) for the frame path this library implements.

export type DepthSortMode = 'float32' | 'packed16';
export type AntialiasMode = 'compensated' | 'classic';
export type GaussianBoundaryPolicy = 'clip' | 'repeat-x' | 'repeat-y';

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
  /** Launch shape; only 16 gives tile keys comparable with the reference. */
  tileSize?: 8 | 16 | 32;
  radixBits?: 4 | 6 | 8;
  rasterBatch?: 128 | 256;
  /** Heavy raster units above which 16-pixel tiles run as row slabs
   * (default: 1.25 per compute unit); never changes any output. */
  heavySplitUnits?: number;
  /** Heavy raster units above which those slabs have 4 rows instead of 8
   * (default: 16 per compute unit); never changes any output. */
  heavyFineSplitUnits?: number;
  device?: number;
}

/** GaussianData's packed arrays (compact geometry + rgb8e8 SH). */
export interface GaussianDataBuffers {
  count: number;
  activeCount?: number;
  shDegree?: 0 | 1 | 2 | 3;
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
  'awaiting-first-result' | 'ready' | 'overflow-pending' | 'budget-exceeded';

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
  readonly renderCount: number;
  camera: CameraLike;
  updateData(data: GaussianDataBuffers): void;
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
  setTuning(tuning: { tileSize: number; radixBits: number; rasterBatch: number }): void;
  dispose(): void;
}

export function gaussianPass(camera: CameraLike, data: GaussianDataBuffers,
                             options?: GaussianPassOptions): GaussianPass;
export function validateOptions(options: GaussianPassOptions): GaussianPassOptions;
export function objectFrames(camera: CameraLike, world?: ArrayLike<number> | null,
                             visible?: boolean): Float32Array;
export const DEFAULTS: Readonly<GaussianPassOptions>;
export const native: Record<string, (...args: any[]) => any>;
