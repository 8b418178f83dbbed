'use strict';
// Command-line driver of the JavaScript host over a raw scene directory:
//   node cli.js bvh <scene_dir> <out_dir>      reference BVH buffers (host only)
//   node cli.js env <scene_dir> <out_dir>      environment tables (host only)
//   node cli.js render <scene_dir> <out.f32> <width> <height> <frames> <depth>
//        PikeRenderer tick loop; writes the running mean (RGBA float32)
//   node cli.js adjoint <scene_dir> <out_dir> <d_image.f32> <width> <height> <spp> <depth>
//        path-replay adjoint of the frame's image (pk_adjoint): writes the
//        atlas and environment gradients (float32, PK_PARAM_* layouts)
const fs = require('fs');
const path = require('path');
const { loadRawScene } = require('./raw_scene');
const { BVH } = require('./bvh');
const { EnvironmentGenerator } = require('./env_sampler');
const { PikeRenderer } = require('./pike');
const native = require('./native');

function write(file, arr) {
  fs.writeFileSync(file, Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength));
}

async function main() {
  const [cmd, dir, out] = process.argv.slice(2);
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
