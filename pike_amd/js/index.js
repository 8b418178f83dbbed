'use strict';
// pike_amd JavaScript host: the reference's render-pass API over the
// MI355X C ABI (N-API addon pike_amd/napi/pike_napi.node).
const { PikeDevice } = require('./device');
const { RenderState, CameraPass } = require('./passes/camera_pass');
const { TracePass } = require('./passes/trace_pass');
const { ShadeHitPass, TraceShadowPass, ShadeMissPass, AccumulatePass } = require('./passes/shade_passes');
const { PostProcessPass } = require('./passes/postprocess_pass');
const { BVH } = require('./bvh');
const { EnvironmentGenerator } = require('./env_sampler');
const { PikeRenderer } = require('./pike');
const { loadRawScene, saveRawScene } = require('./raw_scene');
const { sceneDesc } = require('./scene_desc');
const { Scene } = require('./scene');
const { CameraController } = require('./camera_controller');
const { Raycaster } = require('./raycaster');
const { GLTFLoader } = require('./gltf_loader');
const { TexturePacker } = require('./texture_packer');
const png = require('./png');
const native = require('./native');

module.exports = {
  PikeDevice, RenderState, CameraPass, TracePass, ShadeHitPass, TraceShadowPass, ShadeMissPass, AccumulatePass,
  PostProcessPass, BVH, EnvironmentGenerator, PikeRenderer, loadRawScene, saveRawScene, sceneDesc, Scene,
  GLTFLoader, TexturePacker, png, native, CameraController, Raycaster,
};
