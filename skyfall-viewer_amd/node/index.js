'use strict';
// JS host side of the MI355X training step: a drop-in for the training
// part of the reference's GaussianRenderer (src/renderers/gaussian-renderer.ts:30-58,
// returned at :4110-4153) and a Function-style differentiable rasterizer,
// both over the N-API addon (sv_napi.cpp) and libsv.so.
//
// Differences from the reference renderer, by design:
//  - images are Float32Array (h * w * 3) instead of GPU textures
//    (setTrainingImages takes {data, width, height});
//  - the camera per iteration is chosen by the caller's view order
//    (option viewOrder: 'cycle' | 'batched') instead of Math.random;
//  - prune uses the opacity rule (minOpacity) decided on the device, not
//    the weighted random draw of applyMultiViewPrune;
//  - the multi-view scores sample up to 10 views by a seeded shuffle of the
//    iteration instead of Math.random;
//  - errors of the device path throw; missing state warns and skips as the
//    reference does (console.warn + return).

const native = require('./skyfall_sv.node');

// Reference defaults (gaussian-renderer.ts:88-96, 506-522, 2421-2423).
const DEFAULT_PARAMS = {
  lr_means: 0.00016, lr_sh: 0.0025, lr_opacity: 0.025, lr_scale: 0.005, lr_rotation: 0.001,
  sh_update_interval: 16,
};
const DEFAULT_LOSS = { w_l2: 0.0, w_l1: 0.8, w_ssim: 0.2 };
// DEFAULT_TRAINING_SCHEDULE (training-controller.ts:24-44)
const DEFAULT_TRAINING_SCHEDULE = {
  densifyFromIter: 500, densifyUntilIter: 15000, densificationInterval: 100,
  pruneFromIter: 500, pruneUntilIter: 25000, pruneInterval: 100,
  opacityResetInterval: 3000, opacityResetValue: 0.01,
  minOpacity: 0.005, maxScreenSize: 20.0, maxScaleRatio: 0.0, maxScaleAbsolute: 1.0,
};
// DEFAULT_DENSIFICATION_CONFIG (densification.ts:26-33); performDensification
// gates on MULTI_VIEW_IMPORTANCE_THRESHOLD and samples MULTI_VIEW_SAMPLE_COUNT
// views (gaussian-renderer.ts:71-72)
const DEFAULT_DENSIFICATION_CONFIG = {
  gradThreshold: 0.003, gradAbsThreshold: 0.015, denseSizeRatio: 0.001, minVisibility: 25,
  importanceThreshold: 1.5, postDensifyOpacityCap: 0.8,
};
const MULTI_VIEW_IMPORTANCE_THRESHOLD = 5;
const MULTI_VIEW_SAMPLE_COUNT = 10;

// TrainingController (training-controller.ts:59-160): the iteration
// predicates of the prune / densify / opacity-reset schedule.
class TrainingController {
  constructor(sceneExtent, config = {}) {
    this.config = { ...DEFAULT_TRAINING_SCHEDULE, ...config };
    this.sceneExtent = sceneExtent;
    this.iteration = 0;
  }
  setIteration(i) { this.iteration = i; }
  getIteration() { return this.iteration; }
  shouldPrune() {
    const i = this.iteration, c = this.config;
    if (i < c.pruneFromIter || i > c.pruneUntilIter) return false;
    return i % c.pruneInterval === 0;
  }
  shouldDensify() {
    const i = this.iteration, c = this.config;
    if (i < c.densifyFromIter || i > c.densifyUntilIter) return false;
    return i % c.densificationInterval === 0;
  }
  shouldResetOpacity() {
    const i = this.iteration, c = this.config;
    if (i === 0) return false;
    if (i < c.densifyFromIter || i > c.densifyUntilIter) return false;
    return i % c.opacityResetInterval === 0;
  }
  getPruneParams() {
    const c = this.config, useScreenSizePruning = this.iteration > c.opacityResetInterval;
    let maxScale = 1e6;
    if (c.maxScaleAbsolute > 0) maxScale = c.maxScaleAbsolute;
    else if (c.maxScaleRatio > 0) maxScale = this.sceneExtent * c.maxScaleRatio;
    return { minOpacity: c.minOpacity, maxScale, maxScreenRadius: useScreenSizePruning ? c.maxScreenSize : 10000.0,
             useScreenSizePruning };
  }
  getOpacityResetValue() { return this.config.opacityResetValue; }
}

// Up to `count` distinct view indices of `n`, a seeded Fisher-Yates shuffle
// (mulberry32 of the seed) standing in for sampleCameraIndices' Math.random.
function sampleViews(n, count, seed) {
  const idx = Array.from({ length: n }, (_, i) => i);
  if (n <= count) return idx;
  let s = seed >>> 0;
  const rand = () => {
    s = (s + 0x6D2B79F5) >>> 0;
    let t = s;
    t = Math.imul(t ^ (t >>> 15), t | 1);
    t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
  for (let i = n - 1; i > 0; i--) {
    const j = Math.floor(rand() * (i + 1));
    [idx[i], idx[j]] = [idx[j], idx[i]];
  }
  return idx.slice(0, count);
}

function cameraToNative(cam) {
  const pick = (k) => Array.from(cam[k]);
  const out = { view: pick('view'), proj: pick('proj'), viewport: pick('viewport'), focal: pick('focal') };
  if (cam.view_inv) out.view_inv = pick('view_inv');
  if (cam.proj_inv) out.proj_inv = pick('proj_inv');
  return out;
}

// torch.autograd.Function shape without torch: forward renders the views
// and saves only what the backward needs (final alpha and last index per
// pixel stay on the device, 8 bytes a pixel); backward takes dL/dimage per
// view and returns dL/dparameters (transmittance is recomputed by reverse
// traversal, no per-pixel lists are stored).
class RasterizeFunction {
  static forward(ctx, views) {
    const images = [];
    for (const v of views) {
      native.forward(ctx.handle, v);
      const [w, h] = ctx.size(v);
      images.push(native.readImage(ctx.handle, v, w, h));
    }
    return { images, saved: { views: views.slice() } };
  }

  static backward(ctx, saved, dLdImages) {
    native.zeroGrad(ctx.handle);
    saved.views.forEach((v, i) => {
      native.forward(ctx.handle, v);
      native.setResidual(ctx.handle, v, dLdImages[i]);
      native.backward(ctx.handle, v);
    });
    return native.readGrads(ctx.handle);
  }
}

class Context {
  constructor(capacity, device = 0) {
    this.handle = native.create(capacity, device);
    this.sizes = [];
  }
  setViews(cameras) {
    native.setViews(this.handle, cameras.map(cameraToNative));
    this.sizes = cameras.map((c) => [Math.round(c.viewport[0]), Math.round(c.viewport[1])]);
  }
  size(v) { return this.sizes[v]; }
  close() { native.destroy(this.handle); }
}

// pointCloud: {pos, logscale, quat (x,y,z,w), logit, sh} Float32Arrays.
function createGaussianRenderer(pointCloud, options = {}) {
  const n = pointCloud.logit.length;
  const capacity = options.capacity || Math.max(1, 2 * n);
  const ctx = new Context(capacity, options.device || 0);
  native.setGaussians(ctx.handle, pointCloud);
  let params = { ...DEFAULT_PARAMS };
  let loss = { ...DEFAULT_LOSS };
  let cameras = [];
  let imagesSet = false;
  let training = false;
  let stopRequested = false;
  let iteration = 0;
  let target = 0;
  let ips = 0;
  let cursor = 0;
  let profiling = false;
  const viewOrder = options.viewOrder || 'cycle';
  const sceneExtent = options.sceneExtent || 1.0;
  // the schedule runs inside train(); options.schedule = false turns it off
  const controller = options.schedule === false ? null : new TrainingController(sceneExtent, options.schedule || {});
  const densification = { ...DEFAULT_DENSIFICATION_CONFIG, ...(options.densification || {}) };

  function applyConfig(extra = {}) {
    native.setConfig(ctx.handle, {
      lr_pos: params.lr_means, lr_sh: params.lr_sh, lr_opacity: params.lr_opacity,
      lr_scale: params.lr_scale, lr_rot: params.lr_rotation,
      sh_update_interval: params.sh_update_interval,
      w_l2: loss.w_l2, w_l1: loss.w_l1, w_ssim: loss.w_ssim, ...(options.config || {}), ...extra,
    });
  }
  applyConfig();

  function nextViews() {
    if (viewOrder === 'batched') return cameras.map((_, i) => i);
    const v = cursor % cameras.length;
    cursor += 1;
    return [v];
  }

  const renderer = {
    context: ctx,
    getGaussianCount: () => native.getCount(ctx.handle),
    setTrainingCameras(cams) {
      cameras = cams.slice();
      ctx.setViews(cameras);
      imagesSet = false;
    },
    setTrainingImages(images) {
      if (images.length !== cameras.length) {
        console.warn('[train] image count does not match camera count; skipping');
        return;
      }
      images.forEach((img, v) => native.setTarget(ctx.handle, v, img.data));
      imagesSet = true;
    },
    getTrainingParams: () => ({ ...params }),
    setTrainingParams(p) { params = { ...params, ...p }; applyConfig(); },
    getLossWeights: () => ({ ...loss }),
    setLossWeights(w) { loss = { ...loss, ...w }; applyConfig(); },
    async buildTrainingTiles() {
      if (!cameras.length) { console.warn('[tiles] no training cameras'); return; }
      native.buildTiles(ctx.handle, nextViews()[0]);
    },
    // One training iteration (trainStep, gaussian-renderer.ts:3638-3660).
    async trainStep() {
      const mode = viewOrder === 'batched' ? native.MODE_BATCHED : native.MODE_SEQUENTIAL;
      const st = await native.stepAsync(ctx.handle, nextViews(), mode, 1);
      iteration += 1;
      return st;
    },
    async train(numFrames) {
      if (training) { console.warn('[train] already training'); return; }
      if (!cameras.length || !imagesSet) { console.warn('[train] cameras and images are required'); return; }
      training = true;
      stopRequested = false;
      target = iteration + numFrames;
      const t0 = Date.now();
      const start = iteration;
      try {
        while (iteration < target && !stopRequested) {
          await renderer.trainStep();
          if (!controller) continue;
          // gaussian-renderer.ts:3796-3828
          controller.setIteration(iteration);
          let densifyHandledPrune = false;
          if (controller.shouldDensify()) {
            await renderer.densifyAndPrune();
            densifyHandledPrune = true;
          }
          if (!densifyHandledPrune && controller.shouldPrune()) await renderer.prune();
          if (controller.shouldResetOpacity()) await renderer.resetOpacity(controller.getOpacityResetValue());
        }
      } finally {
        training = false;
        const dt = (Date.now() - t0) / 1000;
        ips = dt > 0 ? (iteration - start) / dt : 0;
      }
    },
    stopTraining() { stopRequested = true; },
    getTrainingStatus: () => ({
      isTraining: training, currentIteration: iteration, targetIterations: target,
      iterationsPerSecond: ips, gaussianCount: native.getCount(ctx.handle),
    }),
    async prune(customParams) {
      const params = customParams || (controller ? controller.getPruneParams() : { minOpacity: 0.005 });
      const minOpacity = params.minOpacity !== undefined ? params.minOpacity : 0.005;
      native.pruneOpacity(ctx.handle, minOpacity);
      return native.getCount(ctx.handle);
    },
    // runDensifyAndPruneCombined (gaussian-renderer.ts:3468-3488): one set
    // of multi-view scores, densify, then prune
    async densifyAndPrune() {
      if (!native.getCount(ctx.handle)) return { densified: false, pruned: false, pruneAttempted: false };
      const scores = await renderer.getMultiViewScores();
      const d = await renderer.densify({ scores });
      const before = native.getCount(ctx.handle);
      await renderer.prune();
      return { densified: d.cloned + d.split > 0, pruned: native.getCount(ctx.handle) !== before, pruneAttempted: true };
    },
    async resetOpacity(maxOpacity = 0.1) { native.resetOpacity(ctx.handle, maxOpacity); },
    // A live view of training from any camera (the viewer's render, drawn
    // with the training tile pipeline): {image, alpha, width, height}.
    renderCamera(camera) { return native.renderCamera(ctx.handle, camera); },
    // initializeKnnAndSync (gaussian-renderer.ts:2159-2172): density-aware
    // initial log-scales from the 3 nearest neighbours; the trainer's
    // parameters are the device masters, so there is nothing to sync.
    async initializeKnnAndSync(sceneExtent = options.sceneExtent || 1.0) { return native.knnInit(ctx.handle, sceneExtent); },
    // generateMultiViewScores (gaussian-renderer.ts:2754-2781) over the given
    // views (the reference samples MULTI_VIEW_SAMPLE_COUNT = 10 at random;
    // here a seeded sample of the iteration)
    async getMultiViewScores(views = sampleViews(cameras.length, MULTI_VIEW_SAMPLE_COUNT, iteration)) {
      return native.multiViewScores(ctx.handle, views);
    },
    // performDensification (gaussian-renderer.ts:3211-3466): at most
    // max(1, floor(0.1 N)) new splats per event, highest score first
    // (:3297-3306). rule 'gradient' (the trainer's): mean screen-space
    // gradients since the last event and multi-view importance
    // (:3249-3290); rule 'tile_l1': candidates on tiles whose L1 error is
    // above the views' `percentile`, split above `splitPx` px.
    async densify({ rule = 'gradient', scores, views, percentile = 0.9, splitPx = 8, seed = iteration,
                    maxNew = Math.max(1, Math.floor(0.1 * native.getCount(ctx.handle))) } = {}) {
      let d;
      if (rule === 'tile_l1') {
        d = native.densifyDecide(ctx.handle, views || cameras.map((_, i) => i), percentile, splitPx, maxNew);
      } else if (rule === 'gradient') {
        const sc = scores || await renderer.getMultiViewScores(views);
        if (!sc || !sc.sampleCount) {
          console.warn('[densify] Unable to compute multi-view scores; skipping.');
          return { cloned: 0, split: 0 };
        }
        const width = cameras.length ? Math.round(cameras[0].viewport[0]) : 1;
        d = native.densifyDecideGrad(ctx.handle, {
          gradThreshold: densification.gradThreshold, gradAbsThreshold: densification.gradAbsThreshold,
          sizeThreshold: sceneExtent * densification.denseSizeRatio, minVisibility: densification.minVisibility,
          importanceThreshold: MULTI_VIEW_IMPORTANCE_THRESHOLD, pixelScale: Math.max(width, 1) / 2.0,
        }, sc.importance, sc.sampleCount, maxNew);
      } else {
        throw new Error(`unknown densify rule ${rule}`);
      }
      if (d.clones + d.splits > 0) native.densify(ctx.handle, d.actions, seed);
      return { cloned: d.clones, split: d.splits };
    },
    enableProfiling(enabled) { profiling = !!enabled; native.enableTiming(ctx.handle, profiling); },
    printProfilingReport() {
      for (const t of native.timings(ctx.handle)) {
        console.log(`${t.name.padEnd(16)} ${(t.totalMs / t.launches).toFixed(3)} ms x ${t.launches}`);
      }
    },
    async profiledTrainStep() {
      native.enableTiming(ctx.handle, true);
      const st = await renderer.trainStep();
      renderer.printProfilingReport();
      native.enableTiming(ctx.handle, profiling);
      return st;
    },
    getTrainingController: () => controller,
    async dumpTrainingDiagnostics() {
      const s = renderer.getTrainingStatus();
      console.log(`[diag] iteration ${s.currentIteration}, ${s.gaussianCount} Gaussians, ${cameras.length} views`);
    },
    readImage(v) { const [w, h] = ctx.size(v); return native.readImage(ctx.handle, v, w, h); },
    getGaussians: () => native.getGaussians(ctx.handle),
    destroy() { ctx.close(); },
  };
  return renderer;
}

module.exports = {
  native, Context, RasterizeFunction, createGaussianRenderer, cameraToNative, TrainingController,
  DEFAULT_TRAINING_SCHEDULE, DEFAULT_DENSIFICATION_CONFIG,
};
