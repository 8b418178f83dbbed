'use strict';
// RenderState and CameraPass with the reference's surface
// (render/passes/camera_pass.js:3-221).  RenderState owns the frame on the
// device (color accumulator, queues, counters: one pk_frame); CameraPass
// records ray generation.  The counter clears of the reference
// (clearNumHitsAndMisses / clearNumRays / clearColorBuffer) are performed by
// the kernels themselves — each kernel resets the counters its successor
// uses — so those methods record nothing and exist for drop-in callers.

class RenderState {
  constructor(device, resolution, batchSize) {
    this.device = device;
    this.samples = 0;
    this.envTheta = 0;
    this.resolution = [Number(resolution[0]), Number(resolution[1])];
    this.batchSize = batchSize;
    this.frame = device.native.frameCreate(device.ctx, this.resolution[0], this.resolution[1]);
  }

  incrementSamples() {
    this.samples++;
  }

  resetSamples() {
    this.samples = 0;
  }

  getSamples() {
    return this.samples;
  }

  setEnvRotation(theta) {
    this.envTheta = theta;
  }

  getRenderStateBuffer() {
    return this.frame;
  }

  async printDebugInfo() {
    const c = this.device.native.counters(this.frame);
    console.log(c);
    return c;
  }

  // Latches {samples, envTheta} for the frame about to be recorded
  // (camera_pass.js:53-64); CameraPass.generateCommands writes them to the
  // device together with the camera state.
  generateCommands(commandEncoder) {
    commandEncoder.frameState = { samples: this.samples, envTheta: this.envTheta };
  }

  clearNumHitsAndMisses(commandEncoder) {}

  clearNumRays(commandEncoder) {}

  clearColorBuffer(commandEncoder) {}

  // Running mean and the current frame's color buffer (RGBA float32).
  readImage() {
    return this.device.native.readImage(this.frame, this.resolution[0], this.resolution[1], false);
  }

  readColorBuffer() {
    return this.device.native.readImage(this.frame, this.resolution[0], this.resolution[1], true);
  }

  destroy() {
    if (this.frame) this.device.native.frameDestroy(this.frame);
    this.frame = null;
  }
}

class CameraPass {
  constructor(device, resolution, batchSize) {
    this.device = device;
    this.resolution = [Number(resolution[0]), Number(resolution[1])];
    this.batchSize = batchSize || this.resolution[0] * this.resolution[1];
    if (this.batchSize !== this.resolution[0] * this.resolution[1]) {
      throw new Error('CameraPass: the batch must cover the frame (render/pike.js:18 uses W*H)');
    }
    this.renderState = new RenderState(device, this.resolution, this.batchSize);
    this.cameraState = null;
  }

  static async create(...args) {
    return new CameraPass(...args);
  }

  getRenderState() {
    return this.renderState;
  }

  // {pos, dir, fov, focalDepth, apertureSize, distortion, bokeh,
  //  invocationOffset} in the tracer's normalized scene space
  setCameraState(state) {
    if (state.invocationOffset) {
      throw new Error('CameraPass: invocationOffset batching is not supported (the batch is the whole frame)');
    }
    this.cameraState = Object.assign({}, state);
  }

  generateCommands(commandEncoder) {
    if (!this.cameraState) throw new Error('CameraPass: setCameraState before generateCommands');
    const n = this.device.native;
    const rs = this.renderState;
    const frame = rs.frame;
    const cam = this.cameraState;
    const fs = commandEncoder.frameState || { samples: rs.samples, envTheta: rs.envTheta };
    commandEncoder.nextBounce = 0;
    commandEncoder.record(() => {
      n.frameSetCamera(frame, cam);
      n.frameSetEnvTheta(frame, fs.envTheta);
      n.frameBegin(frame, fs.samples);
      n.raygen(frame);
    });
  }
}

module.exports = { RenderState, CameraPass };
