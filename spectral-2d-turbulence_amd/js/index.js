'use strict';
// Node.js host interface of the MI355X solver: the reference's Solver / Fft2D / SCENARIOS surface
// (src/gpu/solver.js, src/gpu/fft.js, src/scenarios.js) over the N-API addon s2t_napi.node,
// which wraps the C ABI of include/s2t.h.
//
// The frame-loop surface is the reference's: encodePrime / encodeStep / encodeDye / encodeReduce /
// encodeDiagnostics record work into a compute pass of the device in device.js, copyDiagnostics
// into its command encoder, queue.submit runs it and readDiagnostics resolves with the copied
// record (null when nothing was copied). Differences, by design:
//  * the CFL reduction runs inside every step, from the step's initial field
//    (reference/spectral2d.py:66-71), so encodeReduce records nothing;
//  * encodeDye straight after encodeStep of the same solver in one pass fuses into that step (its
//    row kernel writes the dye velocity); on its own it advects with the last dye velocity;
//  * the solver's full-size buffers (buf.packA, packB, physZ, force) are device buffers for the
//    Fft2D.register route; buf.state[k] reads and writes the current state in the fixture layout;
//  * solver.step(n, {dye}) and the host-array hooks (readState / writeState, velocityPack,
//    assemble, rhs, Fft2D on host arrays) remain as direct calls.

const path = require('path');
const { Device, DeviceBuffer, GPUBufferUsage, GPUMapMode, requestDevice } = require('./device.js');

let native = null;
function addon() {
  if (!native) {
    // throws when the addon has not been built: there is no JavaScript fallback
    native = require(path.join(__dirname, 's2t_napi.node'));
  }
  return native;
}

const MAX_SPLATS = 8;

// radixSchedule(n) of src/gpu/fft.js:20-31 (same result object and error text)
function radixSchedule(n) {
  return addon().radixSchedule(n);
}

// config defaults of src/main.js:9-36 for the keys the solver reads
const DEFAULT_CONFIG = Object.freeze({
  beta: 0,
  drag: 0.01,
  viscosity: 0.67,
  hyperP: 4,
  forceAmp: 0,
  forceK: 22,
  forceWidth: 2.5,
  cfl: 0.5,
  dtMin: 1e-5,
  dtMax: 0.02,
  dyeFade: 0.999,
});

// buf.state[k] of the reference (the current spectrum, fixture layout) and buf.dt (a dt written
// there becomes the solver's fixed step, as the reference harness writes dtbuf directly,
// tests/index.html:188)
class StateView {
  constructor(solver) {
    this.solver = solver;
    this.size = solver.ncells * 8;
    this.label = 'state';
  }

  _write(offset, bytes) {
    if (offset !== 0 || bytes.byteLength !== this.size) throw new Error('write the whole state');
    const f = new Float32Array(bytes.buffer.slice(bytes.byteOffset, bytes.byteOffset + this.size));
    this.solver.writeState(f);
  }

  _copyTo(offset, dst, dstOffset, size) {
    const st = this.solver.readState();
    dst._write(dstOffset, new Uint8Array(st.buffer, offset, size));
  }
}

class DtView {
  constructor(solver) {
    this.solver = solver;
    this.size = 16;
    this.label = 'dt';
  }

  _write(offset, bytes) {
    const f = new Float32Array(bytes.buffer.slice(bytes.byteOffset, bytes.byteOffset + 4));
    this.solver.configure({ fixedDt: f[0] });
  }

  _copyTo(offset, dst, dstOffset, size) {
    const d = this.solver.diagnostics();
    const f = new Float32Array([d.dt, d.maxSpeed, 0, 0]);
    dst._write(dstOffset, new Uint8Array(f.buffer, offset, size));
  }
}

class Solver {
  constructor(device, handle, n, dyeScale) {
    this.device = device || new Device();
    this.handle = handle;
    this.n = n;
    this.ncells = n * n;
    this.m = n * dyeScale;
    this.nbins = n / 2 + 1;
    this.dyeScale = dyeScale;
    this.phase = 0;
    this.dyeParity = 0;
    this.stepIndex = 0;
    this.simTime = 0;
    this.readState_ = 'idle';
    this.primed = false;
    this.disposed = false;
    this.cfg = { ...DEFAULT_CONFIG };
    this.fft = new Fft2D(this.device, n, this);
    this._buf = null;
  }

  // Solver.create(device, {n, dyeScale}) (solver.js:71-75)
  static async create(device, { n = 256, dyeScale = 2 } = {}) {
    const s = new Solver(device, addon().create(n, dyeScale), n, dyeScale);
    s.configure(s.cfg);
    return s;
  }

  _h() {
    if (this.disposed) throw new Error('solver was destroyed');
    return this.handle;
  }

  _sync() {
    if (!this.disposed) addon().sync(this.handle);
  }

  // the reference's named buffers (solver.js:88-104): full-size fields registered with this
  // solver's Fft2D, created on first use
  get buf() {
    if (!this._buf) {
      const bytes = this.ncells * 8;
      const mk = (label) => this.device.createBuffer({ size: bytes, usage: GPUBufferUsage.STORAGE, label });
      const b = {};
      for (const name of ['packA', 'packB', 'physZ', 'force']) {
        let made = null;
        Object.defineProperty(b, name, {
          enumerable: true,
          get: () => {
            if (!made) {
              made = mk(name);
              this.fft.register(name, made);
            }
            return made;
          },
        });
      }
      const state = new StateView(this);
      b.state = [state, state, state];
      b.dt = new DtView(this);
      this._buf = b;
    }
    return this._buf;
  }

  // configure(cfg) (solver.js:306-324); cfg.fixedDt > 0 replaces the CFL step
  configure(cfg) {
    this.cfg = { ...this.cfg, ...cfg };
    addon().configure(this._h(), this.cfg);
  }

  // hyperviscosity coefficient in effect (paramsF[2] of the reference)
  get nuH() {
    return addon().nuH(this._h());
  }

  get paramsF() {
    return Float32Array.from([this.cfg.beta, this.cfg.drag, this.nuH, this.cfg.hyperP,
      this.cfg.forceAmp, this.cfg.forceK, this.cfg.forceWidth, 0, (2 * Math.PI) / this.n,
      this.cfg.cfl, this.cfg.dtMin, this.cfg.dtMax]);
  }

  // setSplats(list) (solver.js:326-345), clamped to MAX_SPLATS
  setSplats(list) {
    addon().setSplats(this._h(), list.slice(0, MAX_SPLATS));
  }

  // seed({k0, amplitude, seed}) (solver.js:347-359)
  seed({ k0 = 14, amplitude = 6, seed = 1 } = {}) {
    addon().seedSpectrum(this._h(), k0, amplitude, seed);
    this.phase = 0;
    this.primed = false;
  }

  // seedPhysical(values) (solver.js:361-382)
  seedPhysical(values) {
    const f = values instanceof Float32Array ? values : Float32Array.from(values);
    addon().seedPhysical(this._h(), f, this.n);
    this.phase = 0;
    this.primed = false;
  }

  // reset(opts) (solver.js:420-444): blank dye, then the pattern; clock at zero, dt = dtMax
  reset(opts = {}) {
    this.stepIndex = 0;
    this.simTime = 0;
    this.dyeParity = 0;
    if (opts.field) {
      this.seedPhysical(opts.field);
    } else {
      this.seed(opts);
    }
    addon().resetDye(this._h());
    this.paintDye(opts.dyePattern !== undefined ? opts.dyePattern : 'bands');
    this.primed = false;
  }

  // paintDye(pattern) (solver.js:446-473); 'none' leaves the dye as it is
  paintDye(pattern) {
    if (pattern === 'none') return;
    addon().paintDye(this._h(), 1);
  }

  // encodePrime(pass) (solver.js:475-486): the velocity of the current state
  encodePrime(pass) {
    pass._record(this, () => addon().prime(this._h()));
    this.primed = true;
  }

  // encodeStep(pass) (solver.js:488-521): one IF-RK3 step with its CFL reduction
  encodeStep(pass) {
    const op = pass._record(this, () => addon().step(this._h(), 1, op.tag === 'step+dye'), 'step');
    this.phase = (this.phase + 1) % 3;
    this.stepIndex += 1;
    this.primed = true;
  }

  // encodeDye(pass) (solver.js:523-536): straight after this solver's step it fuses into it
  encodeDye(pass) {
    const last = pass._last();
    if (last && last.owner === this && last.tag === 'step') {
      last.tag = 'step+dye';
    } else {
      pass._record(this, () => addon().dyeStep(this._h(), 0), 'dye');
    }
    this.dyeParity = 1 - this.dyeParity;
  }

  // encodeReduce(pass) (solver.js:538-545): the reduction is part of every step
  encodeReduce() {}

  // encodeDiagnostics(pass) (solver.js:547-558)
  encodeDiagnostics(pass) {
    pass._record(this, () => addon().encodeDiagnostics(this._h()), 'diagnostics');
  }

  // copyDiagnostics(encoder) (solver.js:560-568): skipped while a read is outstanding
  copyDiagnostics(encoder) {
    if (this.readState_ !== 'idle') return;
    encoder._record(this, () => addon().copyDiagnostics(this._h()), 'copy');
    this.readState_ = 'copied';
  }

  // readDiagnostics() (solver.js:570-608): the copied record in FP64, or null when nothing was
  // copied, the copy failed or the solver was destroyed
  async readDiagnostics() {
    if (this.readState_ !== 'copied' || this.disposed) return null;
    this.readState_ = 'mapping';
    try {
      const d = await addon().readCopiedDiagnosticsAsync(this.handle, this.n);
      return this.disposed ? null : d;
    } catch (e) {
      return null;
    } finally {
      this.readState_ = 'idle';
    }
  }

  // the diagnostics of the state after all calls made so far, on a worker thread
  async diagnosticsAsync() {
    if (this.disposed) return null;
    const d = await addon().diagnosticsAsync(this._h(), this.n);
    this.simTime = d.time;
    return d;
  }

  diagnostics() {
    const d = addon().diagnostics(this._h(), this.n);
    this.simTime = d.time;
    return d;
  }

  // `count` steps (+ dye): encodeStep (+ encodeDye) + encodeReduce, enqueued at once
  step(count = 1, { dye = false } = {}) {
    addon().step(this._h(), count, !!dye);
    this.stepIndex += count;
    this.phase = (this.phase + count) % 3;
    if (dye && count % 2) this.dyeParity = 1 - this.dyeParity;
    if (count > 0) this.primed = true;
  }

  // the same steps on a worker thread: the Promise resolves once the device has finished them
  // (calls on the solver meanwhile wait for them)
  stepAsync(count = 1, { dye = false } = {}) {
    this.stepIndex += count;
    this.phase = (this.phase + count) % 3;
    if (dye && count % 2) this.dyeParity = 1 - this.dyeParity;
    if (count > 0) this.primed = true;
    return addon().stepAsync(this._h(), count, !!dye);
  }

  // column-group width of the column kernel, tuned on a scratch solver (s2t_autotune):
  // {width, ms: {width: ms per step}}; results are the same at every width
  autotune(steps = 5) {
    return addon().autotune(this._h(), this.n, steps);
  }

  sync() {
    addon().sync(this._h());
  }

  // current dye texture as binary16 rgba (Solver.currentDye)
  get currentDye() {
    return addon().readDye(this._h());
  }

  readState() {
    return addon().readState(this._h(), this.n);
  }

  writeState(spectrum) {
    const f = spectrum instanceof Float32Array ? spectrum : Float32Array.from(spectrum);
    addon().writeState(this._h(), f, this.n);
    this.primed = false;
  }

  // simulated time and step index (resuming a checkpoint of readState() and the diagnostics'
  // time and stepIndex: the forcing phases follow the step index)
  setClock({ time = 0, stepIndex = 0 } = {}) {
    this.simTime = time;
    this.stepIndex = stepIndex;
    addon().setClock(this._h(), time, stepIndex);
  }

  // the pointwise stages on host arrays (reference harness, tests/index.html:161-182)
  velocityPack(spectrum) {
    const f = spectrum instanceof Float32Array ? spectrum : Float32Array.from(spectrum);
    return addon().velocityPack(this._h(), f, this.n);
  }

  assemble(z, fhat = null, { stepIndex = 0, dt = 0.01 } = {}) {
    const zf = z instanceof Float32Array ? z : Float32Array.from(z);
    const ff = fhat === null ? null : fhat instanceof Float32Array ? fhat : Float32Array.from(fhat);
    return addon().assemble(this._h(), zf, ff, this.n, stepIndex, dt);
  }

  rhs({ stepIndex = 0, dt = 0.01 } = {}) {
    return addon().rhs(this._h(), this.n, stepIndex, dt);
  }

  destroy() {
    this.disposed = true;
    if (this._buf) {
      for (const name of ['packA', 'packB', 'physZ', 'force']) {
        const d = Object.getOwnPropertyDescriptor(this._buf, name);
        if (d && d.get && this.fft.groups.has(name)) this.fft.groups.get(name).destroy();
      }
    }
    // the native handle is released by its finalizer when the last reference goes
    this.handle = null;
  }
}

// Fft2D (fft.js:33-101): register(name, buffer) binds a device buffer; forward(pass, name) /
// inverse(pass, name) record an in-place 2-D transform of it (row pass, then column pass; forward
// unscaled, inverse 1/N per pass). forward(data) / inverse(data) on a host array (Float32Array in
// the fixture layout) transform it at once and return the result.
class Fft2D {
  constructor(device, n, solver = null) {
    this.device = device;
    this.n = n;
    this.solver = solver;
    this.groups = new Map();
    this.handle = solver ? null : addon().create(n, 1);
  }

  static async create(device, n) {
    radixSchedule(n);
    return new Fft2D(device || new Device(), n);
  }

  _h() {
    return this.solver ? this.solver._h() : this.handle;
  }

  _sync() {
    if (this.solver) this.solver._sync();
    else addon().sync(this.handle);
  }

  register(name, buffer) {
    this.groups.set(name, buffer);
  }

  run(pass, name, inverse) {
    if (!(pass && typeof pass._record === 'function')) return this._host(pass, inverse);
    const buffer = this.groups.get(name);
    if (!buffer) throw new Error(`buffer ${name} was never registered with the FFT`);
    pass._record(this.solver || this, () => addon().fft2Device(this._h(), buffer._ext, this.n, !!inverse));
  }

  _host(data, inverse) {
    const f = data instanceof Float32Array ? data : Float32Array.from(data);
    return addon().fft2(this._h(), f, this.n, !!inverse);
  }

  forward(pass, name) {
    return this.run(pass, name, false);
  }

  inverse(pass, name) {
    return this.run(pass, name, true);
  }
}

// ---------------------------------------------------------------------------------------------
// scenarios: physical vorticity fields on x_j = 2 pi j / N (src/scenarios.js:57-165)
// ---------------------------------------------------------------------------------------------
const TWO_PI = 2 * Math.PI;

function onGrid(n, fn) {
  const out = new Float32Array(n * n);
  const h = TWO_PI / n;
  for (let r = 0; r < n; r++) {
    for (let c = 0; c < n; c++) out[r * n + c] = fn(c * h, r * h);
  }
  return out;
}

function periodicDelta(a, b) {
  const d = a - b;
  if (d > Math.PI) return d - TWO_PI;
  if (d < -Math.PI) return d + TWO_PI;
  return d;
}

// Gaussian vortices a^2-scaled, truncated at r^2 < 25 a^2
function gaussians(n, cores) {
  return onGrid(n, (x, y) => {
    let sum = 0;
    for (const v of cores) {
      const dx = periodicDelta(x, v.x);
      const dy = periodicDelta(y, v.y);
      const r2 = dx * dx + dy * dy;
      const a2 = v.a * v.a;
      if (r2 < 25 * a2) sum += v.amp * Math.exp(-r2 / a2);
    }
    return sum;
  });
}

const SCENARIOS = {
  'shear layer': {
    config: { substeps: 2, viscosity: 0.5, hyperP: 4, beta: 0, drag: 0, forceAmp: 0 },
    build(n) {
      const delta = Math.max(TWO_PI / 30, (6 * TWO_PI) / n);
      return onGrid(n, (x, y) => {
        const low = y <= Math.PI;
        const z = (low ? y - Math.PI / 2 : 1.5 * Math.PI - y) / delta;
        const c = Math.cosh(z);
        return 0.05 * Math.cos(x) - (low ? 1 : -1) / (c * c) / delta;
      });
    },
  },
  'taylor-green': {
    config: { substeps: 4, viscosity: 90, hyperP: 1, beta: 0, drag: 0, forceAmp: 0 },
    build(n) {
      return onGrid(n, (x, y) => 6 * Math.cos(2 * x) * Math.cos(2 * y));
    },
  },
  dipole: {
    config: { substeps: 4, viscosity: 0.3, hyperP: 4, beta: 0, drag: 0, forceAmp: 0 },
    build(n) {
      const a = 0.3;
      const amp = 12 / (Math.PI * a * a);
      const x = 0.6 * Math.PI;
      return gaussians(n, [
        { x, y: Math.PI - 0.45, a, amp },
        { x, y: Math.PI + 0.45, a, amp: -amp },
      ]);
    },
  },
  merger: {
    config: { substeps: 8, viscosity: 0.3, hyperP: 4, beta: 0, drag: 0, forceAmp: 0 },
    build(n, cfg = {}) {
      const d = cfg.mergerSep !== undefined ? cfg.mergerSep : 0.7;
      return gaussians(n, [
        { x: Math.PI - d / 2, y: Math.PI, a: 0.22, amp: 30 },
        { x: Math.PI + d / 2, y: Math.PI, a: 0.22, amp: 30 },
      ]);
    },
  },
  polygon: {
    config: { substeps: 8, viscosity: 0.22, hyperP: 4, beta: 0, drag: 0, forceAmp: 0 },
    build(n, cfg = {}) {
      const count = Math.round(cfg.polygonN !== undefined ? cfg.polygonN : 6);
      const mode = Math.max(2, Math.floor(count / 2));
      const cores = [];
      for (let j = 0; j < count; j++) {
        const th = (TWO_PI * j) / count;
        const r = 1.25 * (1 + 0.02 * Math.cos((TWO_PI * mode * j) / count));
        cores.push({ x: Math.PI + r * Math.cos(th), y: Math.PI + r * Math.sin(th), a: 0.19, amp: 34 });
      }
      return gaussians(n, cores);
    },
  },
};

const SCENARIO_NAMES = Object.keys(SCENARIOS);

// Ensemble (include/s2t.h s2t_batch_*): members of mixed grid sizes step together, members of
// one size in the same fused kernels, each bit-identical to the same member run as a Solver.
class Batch {
  constructor(sizes) {
    this.sizes = sizes.slice();
    this.handle = addon().batchCreate(this.sizes);
    this.cfg = this.sizes.map(() => ({ ...DEFAULT_CONFIG }));
    this.sizes.forEach((_, i) => this.configure(i, {}));
  }

  get count() {
    return this.sizes.length;
  }

  configure(i, cfg) {
    this.cfg[i] = { ...this.cfg[i], ...cfg };
    addon().batchConfigure(this.handle, i, this.cfg[i]);
  }

  seed(i, { k0 = 14, amplitude = 6, seed = 1 } = {}) {
    addon().batchSeedSpectrum(this.handle, i, k0, amplitude, seed);
  }

  seedPhysical(i, values) {
    const f = values instanceof Float32Array ? values : Float32Array.from(values);
    addon().batchSeedPhysical(this.handle, i, f);
  }

  // one hipGraph per RK3 step of all members (default) or eager launches
  useGraphs(on) {
    addon().batchUseGraphs(this.handle, !!on);
  }

  // grouped launches for the 256 / 512 / 1024 classes (s2t_batch_use_grouped)
  useGrouped(on) {
    addon().batchUseGrouped(this.handle, !!on);
  }

  step(count = 1) {
    addon().batchStep(this.handle, count);
  }

  // column-group width per size class (s2t_batch_autotune): {n: width}
  autotune(steps = 5) {
    addon().batchAutotune(this.handle, steps);
    const out = {};
    this.sizes.forEach((n, i) => { out[n] = addon().batchColumnWidth(this.handle, i); });
    return out;
  }

  sync() {
    addon().batchSync(this.handle);
  }

  readState(i) {
    return addon().batchReadState(this.handle, i);
  }

  writeState(i, spectrum) {
    const f = spectrum instanceof Float32Array ? spectrum : Float32Array.from(spectrum);
    addon().batchWriteState(this.handle, i, f);
  }

  diagnostics(i) {
    return addon().batchDiagnostics(this.handle, i);
  }
}

// request statuses and server-side scenarios of include/s2t.h
const REQ = Object.freeze({ QUEUED: 0, RUNNING: 1, DONE: 2, REJECTED: 3, CANCELLED: 4, FAILED: 5 });
const SERVER_SCENARIOS = Object.freeze({ turbulence: 0, 'taylor-green': 1, dipole: 2, merger: 3 });

// Serving loop (include/s2t.h s2t_server_*): requests (a scenario, a grid size and a step count)
// share pre-allocated solver slots per grid size and run on the library's worker thread. submit()
// returns the request id at once and a Promise that settles when the request reaches its final
// status: resolved with the FP64 diagnostics (and the final spectrum with keepState) when done,
// rejected with an Error whose `status` is REQ.REJECTED, REQ.CANCELLED or REQ.FAILED otherwise
// (FAILED: the server's worker stopped on an error, whose text is in the message).
class Server {
  // classes: [{n, slots}], budgetBytes: device memory limit for the slots (0 = none)
  constructor(classes, { budgetBytes = 0 } = {}) {
    this.pending = new Map();
    this.destroyed = false;
    this.handle = addon().serverCreate(
      classes.map((c) => c.n), classes.map((c) => c.slots), budgetBytes,
      (id, status, result) => this._final(id, status, result));
    addon().serverRef(this.handle, false);
  }

  // request: {n, steps, scenario ('turbulence' | 'taylor-green' | 'dipole' | 'merger'),
  //           k0, amplitude, seed, mergerSep, keepState}; config: solver keys (DEFAULT_CONFIG)
  submit(request, config = {}) {
    const scen = SERVER_SCENARIOS[request.scenario || 'turbulence'];
    if (scen === undefined) throw new Error(`unknown server scenario ${request.scenario}`);
    const req = {
      ...request,
      scenario: scen,
      scenarioParam: request.mergerSep !== undefined ? request.mergerSep : 0.7,
      keepState: request.keepState ? 1 : 0,
    };
    let settle;
    const result = new Promise((resolve, reject) => { settle = { resolve, reject }; });
    const first = this.pending.size === 0;
    if (first) addon().serverRef(this.handle, true);
    let id;
    try {
      // a rejection's callback is queued during submit and runs after this returns
      id = addon().serverSubmit(this.handle, req, { ...DEFAULT_CONFIG, ...config });
    } catch (e) {
      if (first) addon().serverRef(this.handle, false);
      throw e;
    }
    this.pending.set(id, settle);
    return { id, result };
  }

  poll(id) {
    return addon().serverPoll(this.handle, id);
  }

  cancel(id) {
    addon().serverCancel(this.handle, id);
  }

  stats() {
    return addon().serverStats(this.handle);
  }

  // make the worker fail at the admission after `afterAdmissions` admissions (tests of a caller's
  // handling of a failed server)
  injectFailure(afterAdmissions) {
    addon().serverInjectFailure(this.handle, afterAdmissions);
  }

  // result: diagnostics (and `state`) read by the library's thread when the request finished
  _final(id, status, result) {
    const s = this.pending.get(id);
    if (!s) return;
    this.pending.delete(id);
    if (this.pending.size === 0 && !this.destroyed) addon().serverRef(this.handle, false);
    if (status === REQ.DONE) {
      s.resolve(result);
    } else {
      const e = new Error(status === REQ.CANCELLED ? `request ${id} cancelled`
        : status === REQ.FAILED ? `request ${id} failed: ${result}` : `request ${id} rejected`);
      e.status = status;
      s.reject(e);
    }
  }

  // cancels what is still queued or running (their Promises are rejected with REQ.CANCELLED) and
  // frees the slots; call it when done, as the completion function holds this object alive
  destroy() {
    if (this.destroyed) return;
    this.destroyed = true;
    addon().serverDestroy(this.handle);
  }
}

module.exports = {
  Device,
  DeviceBuffer,
  GPUBufferUsage,
  GPUMapMode,
  requestDevice,
  Solver,
  Batch,
  Server,
  REQ,
  Fft2D,
  SCENARIOS,
  SCENARIO_NAMES,
  MAX_SPLATS,
  DEFAULT_CONFIG,
  radixSchedule,
};
