'use strict';
// Command-line driver of the JavaScript host over a raw scene directory:
//   node cli.js bvh <scene_dir> <out_dir>      reference BVH buffers (host only)
//   node cli.js env <scene_dir> <out_dir>      environment tables (host only)
//   node cli.js render <scene_dir> <out.f32> <width> <height> <frames> <depth>
//        PikeRenderer tick loop; writes the running mean (RGBA float32)
//   node cli.js adjoint <scene_dir> <out_dir> <d_image.f32> <width> <height> <spp> <depth>
//        path-replay adjoint of the frame's image (pk_adjoint): writes the
//        atlas and environment gradients (float32, PK_PARAM_* layouts)
//   node cli.js load-scene <scene.json> <out_dir> [--instanced]
//        Scene.load (scene JSON + glTF props + environment PNG) baked into
//        the raw scene format above (--instanced: props sharing a file become
//        placements of one mesh)
//   node cli.js render-scene <scene.json> <out.png> <width> <height> <frames> <depth> [out.f32]
//        loads, renders (PikeRenderer), tonemaps (PostProcessPass) and writes
//        a PNG (and optionally the linear running mean).  The camera is the
//        scene JSON's "camera" {origin, target, fov} or a view of the whole
//        scene from +z.
const fs = require('fs');
const path = require('path');
const { loadRawScene, saveRawScene } = require('./raw_scene');
const { Scene } = require('./scene');
const { sceneDesc } = require('./scene_desc');
const png = require('./png');
const { BVH } = require('./bvh');
const { EnvironmentGenerator } = require('./env_sampler');
const { PikeRenderer } = require('./pike');
const native = require('./native');

function write(file, arr) {
  fs.writeFileSync(file, Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength));
}

// camera of a loaded scene JSON: its "camera" entry, else the bounding box
// seen from +z
function sceneCamera(scene) {
  const c = scene.json.camera;
  const lo = [Infinity, Infinity, Infinity], hi = [-Infinity, -Infinity, -Infinity];
  const add = (pos, T) => {
    for (let i = 0; i < pos.length; i += 3)
      for (let k = 0; k < 3; k++) {
        const x = T ? T[4 * k] * pos[i] + T[4 * k + 1] * pos[i + 1] + T[4 * k + 2] * pos[i + 2] + T[4 * k + 3]
          : pos[i + k];
        lo[k] = Math.min(lo[k], x);
        hi[k] = Math.max(hi[k], x);
      }
  };
  add(scene.desc.positions, null);
  for (const inst of scene.desc.instances || []) add(scene.desc.meshes[inst.mesh].positions, inst.transform);
  const center = [0, 1, 2].map((k) => (lo[k] + hi[k]) / 2);
  const extent = Math.max(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
  const origin = c && c.origin ? c.origin : [center[0], center[1], center[2] + 1.5 * extent];
  const target = c && c.target ? c.target : center;
  const d = [0, 1, 2].map((k) => target[k] - origin[k]);
  const l = Math.hypot(d[0], d[1], d[2]);
  return { origin, dir: d.map((x) => x / l), fov: c && c.fov ? c.fov : 1.0, focal_depth: 0.5, aperture: 0,
    distortion: 0, bokeh: 0.5 };
}

async function main() {
  // --instanced: props sharing a glTF file become placements of one mesh
  const instancing = process.argv.includes('--instanced');
  process.argv = process.argv.filter((a) => a !== '--instanced');
  const [cmd, dir, out] = process.argv.slice(2);
  if (cmd === 'load-scene' || cmd === 'render-scene') {
    const scene = await new Scene().load(dir, { instancing });
    sceneDesc(scene);
    const cam = sceneCamera(scene);
    if (cmd === 'load-scene') {
      saveRawScene(scene.desc, cam, out, path.basename(dir));
      return;
    }
    const [w, h, frames, depth] = process.argv.slice(5, 9).map(Number);
    const r = new PikeRenderer(scene, [w, h], { numBounces: depth });
    await r.initDevice();
    r.camera = { origin: r.tracePass.normalize(cam.origin), dir: cam.dir, fov: cam.fov };
    for (let f = 0; f < frames; f++) r.tick();
    await r.device.queue.onSubmittedWorkDone();
    const img = r.readImage();
    fs.writeFileSync(out, png.encode(w, h, r.postProcessPass.tonemap(img, w, h)));
    if (process.argv[9]) write(process.argv[9], img);
    r.destroy();
    return;
  }
  const scene = loadRawScene(dir);
  if (cmd === 'bvh') {
    const bvh = new BVH(native.load(), scene.desc);
    const r = bvh.exportReference();
    fs.mkdirSync(out, { recursive: true });
    write(path.join(out, 'nodes.i32'), r.nodes);
    write(path.join(out, 'tris.i32'), r.triangles);
    write(path.join(out, 'verts.f32'), r.vertices);
    write(path.join(out, 'normalizer.f64'), r.normalizer);
    fs.writeFileSync(path.join(out, 'info.json'), JSON.stringify(bvh.info));
  } else if (cmd === 'env') {
    const e = scene.desc.env;
    const gen = new EnvironmentGenerator({ width: e.width, height: e.height, data: e.data },
      { format: e.format, sampler: e.sampler });
    fs.mkdirSync(out, { recursive: true });
    write(path.join(out, 'bins.i32'), gen.createHistogramBuffer());
    write(path.join(out, 'coords.i32'), gen.createEnvCoordBuffer());
    write(path.join(out, 'pdf.f32'), gen.createPdfTexture());
  } else if (cmd === 'render') {
    const [w, h, frames, depth] = process.argv.slice(5).map(Number);
    const r = new PikeRenderer({ desc: scene.desc, env: null }, [w, h], { numBounces: depth });
    await r.initDevice();
    const cam = scene.camera;
    r.camera = { origin: r.tracePass.normalize(cam.origin), dir: cam.dir, fov: cam.fov };
    r.focalDepth = cam.focal_depth;
    r.apertureSize = cam.aperture;
    r.distortion = cam.distortion;
    r.bokeh = cam.bokeh;
    for (let f = 0; f < frames; f++) r.tick();
    await r.device.queue.onSubmittedWorkDone();
    write(out, r.readImage());
    r.destroy();
  } else if (cmd === 'adjoint') {
    const dImage = process.argv[5];
    const [w, h, spp, depth] = process.argv.slice(6).map(Number);
    const r = new PikeRenderer({ desc: scene.desc, env: null }, [w, h], { numBounces: depth });
    await r.initDevice();
    const n = native.load();
    const cam = scene.camera;
    const frame = r.cameraPass.getRenderState().frame;
    n.frameSetCamera(frame, { pos: r.tracePass.normalize(cam.origin), dir: cam.dir, fov: cam.fov,
      focalDepth: cam.focal_depth, apertureSize: cam.aperture, distortion: cam.distortion, bokeh: cam.bokeh });
    const buf = fs.readFileSync(dImage);
    const d = new Float32Array(buf.buffer, buf.byteOffset, buf.byteLength / 4);
    const g = n.adjoint(frame, r.tracePass.getSceneHandle(),
      { width: w, height: h, maxDepth: depth, numFrames: spp }, d);
    fs.mkdirSync(out, { recursive: true });
    write(path.join(out, 'atlas.f32'), g.atlas);
    write(path.join(out, 'env.f32'), g.env);
    r.destroy();
  } else {
    throw new Error(`unknown command ${cmd}`);
  }
}

main().catch((e) => {
  console.error(e.stack || e.message);
  process.exit(1);
});
