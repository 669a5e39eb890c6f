'use strict';
// Drives the JS drop-in renderer and RasterizeFunction on one view for the
// GPU test: reads the target image (float32 h*w*3) from argv[2], writes the
// parameters after one training iteration and the gradients of
// RasterizeFunction.backward(dL/dimage = target - image) as raw float32 to
// argv[3] (concatenated in the order printed on stdout as JSON).
const fs = require('fs');
const path = require('path');
const sv = require(path.join(__dirname, '..', '..', 'skyfall-viewer_amd', 'node', 'index.js'));

async function main() {
  const [targetFile, outFile, n, w, h] = [process.argv[2], process.argv[3], +process.argv[4], +process.argv[5], +process.argv[6]];
  const native = sv.native;
  const g = native.sceneGaussians(0, n, 1.0);
  const cam = native.sceneCamera([0, 0, 3], [0, 0, 0], w, h, 60);
  const buf = fs.readFileSync(targetFile);
  const target = new Float32Array(buf.buffer, buf.byteOffset, buf.byteLength / 4);
  const r = sv.createGaussianRenderer(g, {
    capacity: 2 * n,
    config: { sh_deg: 3, beta1: 0.9, beta2: 0.999, eps: 1e-15, pos_lr_schedule: 0 },
  });
  r.setTrainingCameras([cam]);
  r.setTrainingImages([{ data: target, width: w, height: h }]);
  r.setTrainingParams({ lr_means: 1e-3, lr_sh: 1e-3, lr_opacity: 1e-3, lr_scale: 1e-3, lr_rotation: 1e-3,
                        sh_update_interval: 1 });
  r.setLossWeights({ w_l2: 0.8, w_l1: 0.2, w_ssim: 0 });
  await r.train(1);
  const status = r.getTrainingStatus();
  const after = r.getGaussians();
  // RasterizeFunction on a fresh context with the initial parameters
  const ctx = new sv.Context(2 * n);
  native.setGaussians(ctx.handle, g);
  ctx.setViews([cam]);
  const fw = sv.RasterizeFunction.forward(ctx, [0]);
  const img = fw.images[0];
  const dL = new Float32Array(img.length);
  for (let i = 0; i < img.length; i++) dL[i] = target[i] - img[i];
  const grads = sv.RasterizeFunction.backward(ctx, fw.saved, [dL]);
  const parts = [];
  const order = [];
  for (const k of ['pos', 'logscale', 'quat', 'logit', 'sh']) { parts.push(after[k]); order.push('after_' + k); }
  parts.push(img); order.push('image');
  for (const k of ['pos', 'logscale', 'quat', 'logit', 'sh', 'geom']) { parts.push(grads[k]); order.push('grad_' + k); }
  fs.writeFileSync(outFile, Buffer.concat(parts.map((a) => Buffer.from(a.buffer, a.byteOffset, a.byteLength))));
  console.log(JSON.stringify({ order, iteration: status.currentIteration, count: status.gaussianCount }));
  ctx.close();
  r.destroy();
}

main().catch((e) => { console.error(e); process.exit(1); });
