// Type declarations of index.js for the TypeScript host (the reference's
// src/renderers). Field names follow the reference types they stand in for:
// TrainingCameraData (src/camera/training-camera.ts:4-15), TrainingParams,
// LossWeights, TrainingStatus (src/renderers/gaussian-renderer.ts:99-122),
// PruneParams (src/utils/training-controller.ts:49-54).

export type Vec = ArrayLike<number>;

export interface CameraData {
  view: Vec;        // world -> camera, column-major 4x4
  view_inv?: Vec;
  proj: Vec;        // get_pinhole_projection_matrix layout
  proj_inv?: Vec;
  viewport: Vec;    // (w, h)
  focal: Vec;       // (fx, fy) pixels
}

export interface PointCloudArrays {
  pos: Float32Array;       // 3 per Gaussian
  logscale: Float32Array;  // 3
  quat: Float32Array;      // 4: (x, y, z, w)
  logit: Float32Array;     // 1
  sh: Float32Array;        // 48: coefficient-major, rgb inner
}

export interface TrainingImage { data: Float32Array; width: number; height: number; }
export interface TrainingParams {
  lr_means: number; lr_sh: number; lr_opacity: number; lr_scale: number; lr_rotation: number;
  sh_update_interval: number;
}
export interface LossWeights { w_l2: number; w_l1: number; w_ssim: number; }
export interface TrainingStatus {
  isTraining: boolean; currentIteration: number; targetIterations: number;
  iterationsPerSecond: number; gaussianCount: number;
}
export interface PruneParams { minOpacity: number; maxScale?: number; maxScreenRadius?: number; useScreenSizePruning?: boolean; }
export interface StepStats { nIsect: number; nViews: number; sortPasses: number; keyBits: number; l1Sum: number; }
export interface Grads extends PointCloudArrays { geom: Float32Array; }
export interface MultiViewScores { importance: Uint32Array; pruning: Float32Array; sampleCount: number; }
// TrainingScheduleConfig (src/utils/training-controller.ts:1-21)
export interface TrainingScheduleConfig {
  densifyFromIter: number; densifyUntilIter: number; densificationInterval: number;
  pruneFromIter: number; pruneUntilIter: number; pruneInterval: number;
  opacityResetInterval: number; opacityResetValue: number;
  minOpacity: number; maxScreenSize: number; maxScaleRatio: number; maxScaleAbsolute: number;
}
// DensificationConfig (src/utils/densification.ts:14-24)
export interface DensificationConfig {
  gradThreshold: number; gradAbsThreshold: number; denseSizeRatio: number; minVisibility: number;
  importanceThreshold: number; postDensifyOpacityCap: number;
}
export declare const DEFAULT_TRAINING_SCHEDULE: TrainingScheduleConfig;
export declare const DEFAULT_DENSIFICATION_CONFIG: DensificationConfig;
export declare class TrainingController {
  constructor(sceneExtent: number, config?: Partial<TrainingScheduleConfig>);
  setIteration(iteration: number): void;
  getIteration(): number;
  shouldPrune(): boolean;
  shouldDensify(): boolean;
  shouldResetOpacity(): boolean;
  getPruneParams(): Required<PruneParams>;
  getOpacityResetValue(): number;
}

export declare class Context {
  constructor(capacity: number, device?: number);
  readonly handle: unknown;
  setViews(cameras: CameraData[]): void;
  size(view: number): [number, number];
  close(): void;
}

export declare class RasterizeFunction {
  static forward(ctx: Context, views: number[]): { images: Float32Array[]; saved: { views: number[] } };
  static backward(ctx: Context, saved: { views: number[] }, dLdImages: Float32Array[]): Grads;
}

export interface MI355XGaussianRenderer {
  getGaussianCount(): number;
  setTrainingCameras(cameras: CameraData[]): void;
  setTrainingImages(images: TrainingImage[]): void;
  getTrainingParams(): TrainingParams;
  setTrainingParams(params: Partial<TrainingParams>): void;
  getLossWeights(): LossWeights;
  setLossWeights(weights: Partial<LossWeights>): void;
  buildTrainingTiles(): Promise<void>;
  trainStep(): Promise<StepStats>;
  train(numFrames: number): Promise<void>;
  stopTraining(): void;
  getTrainingStatus(): TrainingStatus;
  prune(customParams?: PruneParams): Promise<number>;
  resetOpacity(maxOpacity?: number): Promise<void>;
  initializeKnnAndSync(sceneExtent?: number): Promise<Float32Array>;
  renderCamera(camera: CameraData): { image: Float32Array; alpha: Float32Array; width: number; height: number };
  getMultiViewScores(views?: number[]): Promise<MultiViewScores>;
  densify(opts?: { rule?: 'gradient' | 'tile_l1'; scores?: MultiViewScores; views?: number[]; percentile?: number;
                   splitPx?: number; seed?: number; maxNew?: number }): Promise<{ cloned: number; split: number }>;
  densifyAndPrune(): Promise<{ densified: boolean; pruned: boolean; pruneAttempted: boolean }>;
  enableProfiling(enabled: boolean): void;
  printProfilingReport(): void;
  profiledTrainStep(): Promise<StepStats>;
  getTrainingController(): TrainingController | null;
  dumpTrainingDiagnostics(): Promise<void>;
  readImage(view: number): Float32Array;
  getGaussians(): PointCloudArrays;
  destroy(): void;
}

export declare function createGaussianRenderer(
  pointCloud: PointCloudArrays,
  options?: { capacity?: number; device?: number; viewOrder?: 'cycle' | 'batched'; config?: Record<string, number>;
             sceneExtent?: number; schedule?: Partial<TrainingScheduleConfig> | false;
             densification?: Partial<DensificationConfig> },
): MI355XGaussianRenderer;

export declare const native: Record<string, unknown>;
