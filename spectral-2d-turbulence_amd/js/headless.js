'use strict';
// Headless frame loop: the sequence of the reference's main.js loop (src/main.js:389-447) run
// against the MI355X solver, without a canvas, a renderer or a pointer.
//
// Per frame: configure(config) and setSplats(splats of the frame); a command encoder and a compute
// pass; encodePrime when the solver is not primed; `substeps` x (encodeStep, encodeDye,
// encodeReduce) unless paused; encodeDiagnostics every 6th frame; pass.end(); copyDiagnostics;
// queue.submit; readDiagnostics().then(...) without waiting for it; simTime += dt * substeps from
// the last diagnostics received (main.js:445). Frames are separated by a turn of the event loop,
// as requestAnimationFrame separates them in a browser.
//
// Usage: node headless.js [--n 256] [--dye-scale 2] [--frames 60] [--substeps S]
//                         [--scenario dipole] [--seed-amplitude 6] [--out state.f32]
// prints one JSON line per diagnostics record received, then a summary line.

const fs = require('fs');
const { Solver, SCENARIOS, DEFAULT_CONFIG, requestDevice } = require('./index.js');

const nextFrame = () => new Promise((resolve) => setImmediate(resolve));

// runFrames: the frame loop for `frames` frames; returns the diagnostics records received (each
// with the frame it was encoded in) and the solver (destroy it when done).
async function runFrames({ n = 256, dyeScale = 2, frames = 60, config = {}, scenario = null,
                           seed = {}, splats = () => [], onDiagnostics = null } = {}) {
  const device = await requestDevice();
  const cfg = { ...DEFAULT_CONFIG, substeps: 4, paused: false, ...config };
  if (scenario) Object.assign(cfg, SCENARIOS[scenario].config, config);
  const solver = await Solver.create(device, { n, dyeScale });
  solver.configure(cfg);
  if (scenario) solver.reset({ field: SCENARIOS[scenario].build(n, cfg) });
  else solver.reset(seed);

  const records = [];
  const pending = [];
  let diagnostics = null;
  for (let tick = 0; tick < frames; tick++) {
    solver.configure(cfg);
    solver.setSplats(cfg.paused ? [] : splats(tick));
    const wantDiagnostics = tick % 6 === 0;
    const encoder = device.createCommandEncoder();
    const pass = encoder.beginComputePass();
    if (!solver.primed) solver.encodePrime(pass);
    if (!cfg.paused) {
      for (let i = 0; i < cfg.substeps; i++) {
        solver.encodeStep(pass);
        solver.encodeDye(pass);
        solver.encodeReduce(pass);
      }
    }
    if (wantDiagnostics) solver.encodeDiagnostics(pass);
    pass.end();
    if (wantDiagnostics) solver.copyDiagnostics(encoder);
    device.queue.submit([encoder.finish()]);
    if (wantDiagnostics) {
      const frame = tick;
      pending.push(solver.readDiagnostics().then((result) => {
        if (!result) return;
        diagnostics = result;
        const rec = { frame, ...result };
        records.push(rec);
        if (onDiagnostics) onDiagnostics(rec);
      }));
    }
    if (diagnostics && !cfg.paused) solver.simTime += diagnostics.dt * cfg.substeps;
    await nextFrame();
  }
  await Promise.all(pending);
  return { solver, records, config: cfg };
}

function parseArgs(argv) {
  const out = {};
  for (let i = 0; i < argv.length; i += 2) {
    const key = argv[i].replace(/^--/, '').replace(/-([a-z])/g, (_, c) => c.toUpperCase());
    out[key] = argv[i + 1];
  }
  return out;
}

async function main() {
  const a = parseArgs(process.argv.slice(2));
  const n = Number(a.n || 256);
  const config = {};
  if (a.substeps) config.substeps = Number(a.substeps);
  if (a.seedAmplitude) config.seedAmplitude = Number(a.seedAmplitude);
  const { solver, records } = await runFrames({
    n, dyeScale: Number(a.dyeScale || 2), frames: Number(a.frames || 60), config,
    scenario: a.scenario || null,
    seed: { k0: 14, amplitude: config.seedAmplitude || 6, seed: 1 },
    onDiagnostics: (r) => {
      console.log(JSON.stringify({ frame: r.frame, stepIndex: r.stepIndex, time: r.time,
                                   dt: r.dt, maxSpeed: r.maxSpeed, energy: r.energy,
                                   enstrophy: r.enstrophy }));
    },
  });
  if (a.out) fs.writeFileSync(a.out, Buffer.from(solver.readState().buffer));
  const d = solver.diagnostics();
  console.log(JSON.stringify({ done: true, records: records.length, stepIndex: d.stepIndex,
                               simTime: solver.simTime, time: d.time, energy: d.energy }));
  solver.destroy();
}

if (require.main === module) {
  main().catch((e) => {
    console.error(e);
    process.exit(1);
  });
}

module.exports = { runFrames };
