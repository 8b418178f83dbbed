'use strict';
// PikeRenderer (render/pike.js:13-170) without the browser: the same pass
// construction order (initWebGpu, :92-118) and the same per-frame command
// recording (tick, :120-162), driven by a loop instead of
// requestAnimationFrame.  Camera state is given in the normalized scene
// space the tracer works in (TracePass.normalize maps scene-space points).
const { PikeDevice } = require('./device');
const { CameraPass } = require('./passes/camera_pass');
const { TracePass } = require('./passes/trace_pass');
const { ShadeHitPass, TraceShadowPass, ShadeMissPass, AccumulatePass } = require('./passes/shade_passes');
const { PostProcessPass } = require('./passes/postprocess_pass');
const { EnvironmentGenerator } = require('./env_sampler');
const { Raycaster } = require('./raycaster');

class PikeRenderer {
  constructor(scene, resolution, options) {
    const opts = options || {};
    this.scene = scene;
    this.resolution = [Number(resolution[0]), Number(resolution[1])];
    this.BATCH_SIZE = this.resolution[0] * this.resolution[1];
    this.numBounces = opts.numBounces === undefined ? 4 : opts.numBounces;
    this.camera = opts.camera || { dir: [0, 0, -1], origin: [0, 0, 2], fov: 1.0 };
    this.focalDepth = opts.focalDepth === undefined ? 0.5 : opts.focalDepth;
    this.apertureSize = opts.apertureSize || 0.0;
    this.distortion = opts.distortion || 0.0;
    this.bokeh = opts.bokeh === undefined ? 0.5 : opts.bokeh;
    this.deviceIndex = opts.device || 0;
  }

  async initDevice() {
    this.device = await PikeDevice.create(this.deviceIndex);
    const envGenerator = this.scene.env ? new EnvironmentGenerator(this.scene.env) : null;
    this.cameraPass = await CameraPass.create(this.device, this.resolution, this.BATCH_SIZE);
    this.renderState = this.cameraPass.getRenderState();
    this.tracePass = await TracePass.create(this.device, this.cameraPass, this.scene);
    this.shadePass = await ShadeHitPass.create(this.device, this.cameraPass, this.tracePass, this.scene, envGenerator);
    this.traceShadowPass = await TraceShadowPass.create(this.device, this.cameraPass, this.tracePass);
    this.shadeMissPass = await ShadeMissPass.create(this.device, this.cameraPass, this.tracePass, envGenerator);
    this.accumulatePass = await AccumulatePass.create(this.device, this.resolution, this.renderState);
    this.postProcessPass = await PostProcessPass.create(this.device, 'rgba8unorm', null, this.accumulatePass);
    this.raycaster = new Raycaster(this.tracePass.getBVH());   // pike.js:117
  }

  // autofocus on what the camera ray hits (pike.js:88-90)
  focusCamera() {
    this.focalDepth = 1 - 1 / this.raycaster.cast(this.camera);
    return this.focalDepth;
  }

  // The reference builds its CameraController over the canvas with an
  // onChange that refocuses and restarts accumulation (pike.js:65-76); a
  // headless host hands over the controller it drives.  The controller's ray
  // and field of view become the camera of the following ticks.
  useController(controller) {
    this.controller = controller;
    const chained = controller.onChange;
    const sync = () => {
      const r = controller.getCameraRay();
      this.camera = { origin: r.origin.slice(), dir: r.dir.slice(), fov: controller.getFov() };
    };
    sync();
    controller.onChange = (state) => {
      sync();
      if (this.raycaster) this.focusCamera();
      if (this.renderState) this.renderState.resetSamples();
      chained(state);
    };
    return controller;
  }

  tick() {
    const commandEncoder = this.device.createCommandEncoder();
    this.renderState.generateCommands(commandEncoder);
    this.renderState.incrementSamples();
    this.renderState.clearColorBuffer(commandEncoder);
    this.renderState.clearNumHitsAndMisses(commandEncoder);
    this.renderState.clearNumRays(commandEncoder);
    this.cameraPass.setCameraState({
      pos: this.camera.origin,
      dir: this.camera.dir,
      fov: this.camera.fov,
      focalDepth: this.focalDepth,
      apertureSize: this.apertureSize,
      distortion: this.distortion,
      bokeh: this.bokeh,
      invocationOffset: 0,
    });
    this.cameraPass.generateCommands(commandEncoder);
    for (let i = 0; i < this.numBounces; i++) {
      this.tracePass.generateCommands(commandEncoder);
      this.shadePass.generateCommands(commandEncoder);
      this.traceShadowPass.generateCommands(commandEncoder);
      this.shadeMissPass.generateCommands(commandEncoder);
    }
    this.accumulatePass.generateCommands(commandEncoder);
    this.postProcessPass.generateCommands(commandEncoder);
    this.device.queue.submit([commandEncoder.finish()]);
  }

  async start(frames) {
    await this.initDevice();
    for (let f = 0; f < frames; f++) this.tick();
    await this.device.queue.onSubmittedWorkDone();
  }

  readImage() {
    return this.renderState.readImage();
  }

  destroy() {
    this.renderState.destroy();
    this.tracePass.destroy();
    this.device.destroy();
  }
}

module.exports = { PikeRenderer };
