'use strict';
// ShadeHitPass (render/passes/shade_hit_pass.js), TraceShadowPass
// (trace_shadow_pass.js), ShadeMissPass (shade_miss_pass.js) and
// AccumulatePass (accumulate_pass.js).  Each keeps the reference's static
// create(...) signature and generateCommands(commandEncoder).

class ShadeHitPass {
  constructor(device, cameraPass, tracePass, scene, envGenerator) {
    this.device = device;
    this.cameraPass = cameraPass;
    this.tracePass = tracePass;
    this.scene = scene;
    this.envGenerator = envGenerator;
  }

  static async create(...args) {
    return new ShadeHitPass(...args);
  }

  // shades the hits of the most recent TracePass in this encoder
  generateCommands(commandEncoder) {
    const n = this.device.native;
    const frame = this.cameraPass.getRenderState().frame;
    const scene = this.tracePass.getSceneHandle();
    const bounce = commandEncoder.nextBounce - 1;
    if (bounce < 0) throw new Error('ShadeHitPass: record a TracePass first');
    commandEncoder.record(() => n.shade(frame, scene, bounce));
  }
}

class TraceShadowPass {
  constructor(device, cameraPass, tracePass) {
    this.device = device;
    this.cameraPass = cameraPass;
    this.tracePass = tracePass;
  }

  static async create(...args) {
    return new TraceShadowPass(...args);
  }

  // any-hit for the NEE shadow rays; unoccluded ones add env radiance
  generateCommands(commandEncoder) {
    const n = this.device.native;
    const frame = this.cameraPass.getRenderState().frame;
    const scene = this.tracePass.getSceneHandle();
    commandEncoder.record(() => n.traceShadow(frame, scene));
  }
}

// Miss shading runs inside the two trace kernels (each pixel's color is
// written only by the thread that owns its path), so this pass records no
// device work; it stays for callers that create and record it.
class ShadeMissPass {
  constructor(device, cameraPass, tracePass, envGenerator) {
    this.device = device;
    this.cameraPass = cameraPass;
    this.tracePass = tracePass;
    this.envGenerator = envGenerator;
  }

  static async create(...args) {
    return new ShadeMissPass(...args);
  }

  generateCommands(commandEncoder) {}
}

class AccumulatePass {
  constructor(device, resolution, renderState) {
    this.device = device;
    this.resolution = resolution;
    this.renderState = renderState;
  }

  static async create(device, resolution, renderState) {
    return new AccumulatePass(device, resolution, renderState);
  }

  getAccumulateTexture() {
    const rs = this.renderState;
    return { width: rs.resolution[0], height: rs.resolution[1], read: () => rs.readImage() };
  }

  generateCommands(commandEncoder) {
    const n = this.device.native;
    const frame = this.renderState.frame;
    commandEncoder.record(() => n.accumulate(frame));
  }
}

module.exports = { ShadeHitPass, TraceShadowPass, ShadeMissPass, AccumulatePass };
