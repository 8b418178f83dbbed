'use strict';
// TracePass (render/passes/trace_pass.js:11-276): builds the BVH with
// bvh.js's split decisions (native builder), uploads geometry, materials,
// atlas and environment tables once (pk_scene_upload), and records the
// closest-hit traversal of each bounce.  Misses are shaded inside the same
// kernel (the reference's ShadeMissPass work for bounce rays).
const { BVH } = require('../bvh');
const { sceneDesc } = require('../scene_desc');

class TracePass {
  constructor(device, cameraPass, scene) {
    this.device = device;
    this.cameraPass = cameraPass;
    this.scene = scene;
    this.desc = sceneDesc(scene);
    this.bvh = new BVH(device.native, this.desc, scene.bvhOptions);
    this.sceneHandle = device.native.sceneUpload(device.ctx, this.desc, this.bvh.handle);
    // what Raycaster needs to cast rays against the uploaded scene
    this.bvh.tracer = { native: device.native, ctx: device.ctx, scene: this.sceneHandle };
  }

  static async create(device, cameraPass, scene) {
    return new TracePass(device, cameraPass, scene);
  }

  getBVH() {
    return this.bvh;
  }

  getSceneHandle() {
    return this.sceneHandle;
  }

  // The normalizer the reference applies to scene space (trace_pass.js:202-211):
  // camera states are given in this normalized space.
  normalize(v) {
    const m = this.bvh.normalizer();
    return [(v[0] - m[0]) * m[3], (v[1] - m[1]) * m[3], (v[2] - m[2]) * m[3]];
  }

  generateCommands(commandEncoder) {
    const n = this.device.native;
    const frame = this.cameraPass.getRenderState().frame;
    const scene = this.sceneHandle;
    const bounce = commandEncoder.nextBounce++;
    commandEncoder.record(() => n.traceClosest(frame, scene, bounce));
  }

  destroy() {
    if (this.sceneHandle) this.device.native.sceneDestroy(this.sceneHandle);
    this.sceneHandle = null;
    this.bvh.destroy();
  }
}

module.exports = { TracePass };
