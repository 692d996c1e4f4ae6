'use strict';
// The part of the WebGPU device API that the reference's solver code and its callers use
// (src/main.js:389-447, src/gpu/fft.js:77-100, tests/index.html:99-112), over the MI355X addon:
// buffers, command encoders, compute passes and a queue. A pass records work instead of shader
// dispatches: Solver.encode* and Fft2D.forward / inverse append native calls to it, and
// queue.submit runs the recorded work in order. Nothing runs before submit, as in WebGPU.
//
// Ordering: each recorded call names the native stream it enqueues on (a solver's) or none (a
// host-blocking buffer copy). Consecutive calls on one stream stay asynchronous; before a call on
// another stream, or a buffer copy, the device waits for the stream used last, so work takes
// effect in queue order.

const path = require('path');

let native = null;
function addon() {
  if (!native) native = require(path.join(__dirname, 's2t_napi.node'));
  return native;
}

// GPUBufferUsage / GPUMapMode bit values of the WebGPU specification
const GPUBufferUsage = Object.freeze({
  MAP_READ: 0x0001, MAP_WRITE: 0x0002, COPY_SRC: 0x0004, COPY_DST: 0x0008, INDEX: 0x0010,
  VERTEX: 0x0020, UNIFORM: 0x0040, STORAGE: 0x0080, INDIRECT: 0x0100, QUERY_RESOLVE: 0x0200,
});
const GPUMapMode = Object.freeze({ READ: 0x0001, WRITE: 0x0002 });

function bytesOf(data, dataOffset = 0, size) {
  const view = ArrayBuffer.isView(data) ? data : new Float32Array(data);
  const elem = view.BYTES_PER_ELEMENT || 1;
  const count = size === undefined ? view.length - dataOffset : size;
  return new Uint8Array(view.buffer, view.byteOffset + dataOffset * elem, count * elem);
}

// Device memory (s2t_buffer_*), zero-filled at creation. Mapping reads the device copy when the
// map resolves; only MAP_READ is supported.
class DeviceBuffer {
  constructor(device, { size, usage = 0, label = '' }) {
    this.device = device;
    this.size = size;
    this.usage = usage;
    this.label = label;
    this.mapState = 'unmapped';
    this._ext = addon().bufferCreate(size);
    this._mapped = null;
  }

  _write(offset, bytes) {
    addon().bufferWrite(this._ext, offset, bytes);
  }

  _copyTo(offset, dst, dstOffset, size) {
    if (dst instanceof DeviceBuffer) {
      addon().bufferCopy(this._ext, offset, dst._ext, dstOffset, size);
    } else {
      dst._write(dstOffset, new Uint8Array(addon().bufferRead(this._ext, offset, size)));
    }
  }

  async mapAsync(mode, offset = 0, size = this.size - offset) {
    if (!(mode & GPUMapMode.READ)) throw new Error('only GPUMapMode.READ is supported');
    this.mapState = 'pending';
    this.device._drain();
    this._mapped = { offset, data: addon().bufferRead(this._ext, offset, size) };
    this.mapState = 'mapped';
  }

  getMappedRange(offset = this._mapped ? this._mapped.offset : 0, size) {
    if (!this._mapped) throw new Error(`buffer ${this.label} is not mapped`);
    const start = offset - this._mapped.offset;
    const end = size === undefined ? this._mapped.data.byteLength : start + size;
    return this._mapped.data.slice(start, end);
  }

  unmap() {
    this._mapped = null;
    this.mapState = 'unmapped';
  }

  destroy() {
    if (this._ext) addon().bufferDestroy(this._ext);
    this._ext = null;
  }
}

// A compute pass records calls; WGSL pipelines are not part of this build.
class ComputePass {
  constructor(encoder) {
    this.encoder = encoder;
  }

  _record(owner, run, tag = null) {
    return this.encoder._record(owner, run, tag);
  }

  _last() {
    const ops = this.encoder.ops;
    return ops.length ? ops[ops.length - 1] : null;
  }

  setPipeline() {
    throw new Error('WGSL pipelines are not part of the MI355X build: use the Solver / Fft2D methods');
  }

  setBindGroup() {
    throw new Error('bind groups are not part of the MI355X build');
  }

  dispatchWorkgroups() {
    throw new Error('WGSL dispatches are not part of the MI355X build');
  }

  end() {}
}

class CommandEncoder {
  constructor(device) {
    this.device = device;
    this.ops = [];
  }

  _record(owner, run, tag = null) {
    const op = { owner, run, tag };
    this.ops.push(op);
    return op;
  }

  beginComputePass() {
    return new ComputePass(this);
  }

  copyBufferToBuffer(src, srcOffset, dst, dstOffset, size) {
    this._record(null, () => src._copyTo(srcOffset, dst, dstOffset, size));
  }

  finish() {
    return { ops: this.ops };
  }
}

class Queue {
  constructor(device) {
    this.device = device;
  }

  writeBuffer(buffer, offset, data, dataOffset = 0, size) {
    this.device._drain();
    buffer._write(offset, bytesOf(data, dataOffset, size));
  }

  submit(commandBuffers) {
    for (const cb of commandBuffers) for (const op of cb.ops) this.device._run(op);
  }

  async onSubmittedWorkDone() {
    this.device._drain();
  }
}

class Device {
  constructor() {
    this.label = 'MI355X (HIP)';
    this.queue = new Queue(this);
    this.lost = new Promise(() => {});
    this._owner = null;  // the object whose stream was used last (has _sync())
  }

  createBuffer(desc) {
    return new DeviceBuffer(this, desc);
  }

  createCommandEncoder() {
    return new CommandEncoder(this);
  }

  _drain() {
    if (this._owner) this._owner._sync();
    this._owner = null;
  }

  _run(op) {
    if (op.owner !== this._owner) this._drain();
    this._owner = op.owner;
    op.run();
  }

  destroy() {
    this._drain();
  }
}

// navigator.gpu.requestAdapter().requestDevice() of the reference's initDevice (device.js:8-42)
async function requestDevice() {
  return new Device();
}

module.exports = { Device, DeviceBuffer, GPUBufferUsage, GPUMapMode, requestDevice };
