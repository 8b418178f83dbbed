'use strict';
// Raycaster (render/raycaster.js:5-131): cast(ray) returns the distance to
// the closest hit of a ray {origin, dir} given in the tracer's normalized
// space, or MAX_T = 1e6 when it hits nothing; PikeRenderer.focusCamera uses
// it for autofocus (pike.js:88-90: focalDepth = 1 - 1 / t).  The reference
// walks the binary BVH on the CPU in float64 with a 1e-12 epsilon; this one
// asks the GPU's closest-hit traversal (pk_trace_rays) over the uploaded
// scene, so t is the tracer's own float32 canonical hit distance (hits
// nearer than 1e-7 are ignored, as in rendering).
const MAX_T = 1e6;

class Raycaster {
  // bvh: TracePass.getBVH() (it carries the uploaded scene), or any object
  // with a `tracer` = {native, ctx, scene}
  constructor(bvh) {
    this.bvh = bvh;
    if (!bvh || !bvh.tracer) throw new Error('Raycaster needs the BVH of an uploaded scene (TracePass.getBVH())');
    this.tracer = bvh.tracer;
  }

  // rays: array of {origin, dir}; returns a Float64Array of distances
  castMany(rays) {
    const buf = new Float32Array(8 * rays.length);
    rays.forEach((r, i) => {
      buf.set([r.origin[0], r.origin[1], r.origin[2], 0, r.dir[0], r.dir[1], r.dir[2], 0], 8 * i);
    });
    const { native, ctx, scene } = this.tracer;
    const res = native.traceRays(ctx, scene, buf, false);
    const out = new Float64Array(rays.length);
    for (let i = 0; i < rays.length; i++) out[i] = res.ids[i] >= 0 ? res.t[i] : MAX_T;
    return out;
  }

  cast(ray) {
    return this.castMany([ray])[0];
  }
}

module.exports = { Raycaster, MAX_T };
