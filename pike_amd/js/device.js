'use strict';
// PikeDevice stands where the reference's GPUDevice stands
// (render/pike.js:92-97): passes receive it in their static create(), record
// work into a command encoder in generateCommands(), and
// device.queue.submit([encoder.finish()]) runs the recorded work in order on
// the HIP stream behind the context.  Kernels are asynchronous; reads
// (readImage, counters) synchronize.
const native = require('./native');

class CommandBuffer {
  constructor(commands) {
    this.commands = commands;
  }
}

class CommandEncoder {
  constructor(device) {
    this.device = device;
    this.commands = [];
    // bounce index of the next TracePass in this encoder; CameraPass resets it
    this.nextBounce = 0;
  }

  record(fn) {
    this.commands.push(fn);
  }

  finish() {
    const cb = new CommandBuffer(this.commands);
    this.commands = [];
    return cb;
  }
}

class Queue {
  constructor(device) {
    this.device = device;
  }

  submit(buffers) {
    for (const b of buffers) {
      for (const fn of b.commands) fn();
    }
  }

  onSubmittedWorkDone() {
    this.device.native.sync(this.device.ctx);
    return Promise.resolve();
  }
}

class PikeDevice {
  constructor(index) {
    this.native = native.load();
    this.ctx = this.native.ctxCreate(index || 0);
    this.queue = new Queue(this);
  }

  static async create(index) {
    return new PikeDevice(index);
  }

  createCommandEncoder() {
    return new CommandEncoder(this);
  }

  destroy() {
    if (this.ctx) this.native.ctxDestroy(this.ctx);
    this.ctx = null;
  }
}

module.exports = { PikeDevice, CommandEncoder, Queue };
