'use strict';
// Drives the JS drop-in renderer's train loop for the GPU test: starts
// train(1000) on a small scene, asks for a second train() while the first
// runs (the reference warns and returns, gaussian-renderer.ts:3746-3749),
// then stopTraining() once two iterations have finished
// (gaussian-renderer.ts:3856-3861). Prints the status as JSON.
const path = require('path');
const sv = require(path.join(__dirname, '..', '..', 'skyfall-viewer_amd', 'node', 'index.js'));

const tick = () => new Promise((res) => setTimeout(res, 1));

async function main() {
  const native = sv.native;
  const n = 4096, w = 64, h = 64;
  const g = native.sceneGaussians(0, n, 1.0);
  const cam = native.sceneCamera([0, 0, 3], [0, 0, 0], w, h, 60);
  const target = new Float32Array(w * h * 3).fill(0.25);
  const r = sv.createGaussianRenderer(g, { capacity: n, config: { sh_deg: 3 } });
  r.setTrainingCameras([cam]);
  r.setTrainingImages([{ data: target, width: w, height: h }]);
  const warn = console.warn;
  let warned = 0;
  console.warn = () => { warned += 1; };
  const run = r.train(1000);
  while (r.getTrainingStatus().currentIteration < 2) await tick();
  await r.train(5);  // re-entrant: returns at once
  const during = r.getTrainingStatus();
  r.stopTraining();
  await run;
  console.warn = warn;
  const after = r.getTrainingStatus();
  console.log(JSON.stringify({ warned, during_training: during.isTraining, iterations: after.currentIteration,
                               training: after.isTraining, target: after.targetIterations }));
  r.destroy();
}

main().catch((e) => { console.error(e); process.exit(1); });
