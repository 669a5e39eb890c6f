'use strict';
// Drives the JS drop-in renderer's train loop with its prune / densify /
// opacity-reset schedule (gaussian-renderer.ts:3796-3828) brought forward:
// densify and prune every 10 iterations from 10, opacity reset at 30, with
// the reference trainer's densify rule (lowered thresholds). Prints the
// splat count after each iteration as JSON.
const path = require('path');
const sv = require(path.join(__dirname, '..', '..', 'skyfall-viewer_amd', 'node', 'index.js'));

async function main() {
  const native = sv.native;
  const n = 8192, w = 96, h = 72;
  const g = native.sceneGaussians(0, n, 1.0);
  const cams = [0, 1, 2].map((k) => {
    const az = (2 * Math.PI * k) / 3;
    return native.sceneCamera([3 * Math.sin(az), 0, 3 * Math.cos(az)], [0, 0, 0], w, h, 60);
  });
  const r = sv.createGaussianRenderer(g, {
    capacity: n, config: { sh_deg: 3 }, sceneExtent: 2.0,
    schedule: { densifyFromIter: 10, densificationInterval: 10, pruneFromIter: 10, pruneInterval: 10,
                opacityResetInterval: 30 },
    densification: { gradThreshold: 0.0005, gradAbsThreshold: 0.002, minVisibility: 2 },
  });
  r.setTrainingCameras(cams);
  const images = cams.map(() => ({ data: new Float32Array(w * h * 3).fill(0.3), width: w, height: h }));
  r.setTrainingImages(images);
  const counts = [];
  for (let i = 0; i < 40; i++) {
    await r.train(1);
    counts.push(r.getGaussianCount());
  }
  const p = r.getGaussians();
  let finite = true;
  for (const k of ['pos', 'logscale', 'quat', 'logit', 'sh']) for (const v of p[k]) if (!Number.isFinite(v)) finite = false;
  const maxOpacity = Math.max(...Array.from(p.logit, (x) => 1 / (1 + Math.exp(-x))));
  console.log(JSON.stringify({ counts, finite, maxOpacity, iteration: r.getTrainingStatus().currentIteration }));
  r.destroy();
}

main().catch((e) => { console.error(e); process.exit(1); });
