'use strict';
// BVH (render/bvh.js:7-108): same SAH split decisions, built natively
// (pk_bvh_build) and collapsed into the 128-byte wide layout the kernels
// traverse.  exportReference() returns the reference's flattened buffers
// (trace_pass.js:220-263: BFS nodes of 5 words, leaf-ordered triangles,
// normalized float32 vertices).

// A desc with meshes and instances gets the two-level build
// (pk_bvh_build_instanced), which has no reference export.
// options: { spatialSplits: build with SBVH spatial splits (the branch
// bvh.js:128-135 leaves disabled), threads: builder threads (0 = all) }
class BVH {
  constructor(native, desc, options) {
    const opts = options || {};
    this.native = native;
    this.numVertices = desc.positions.length / 3;
    this.numTriangles = desc.triangles.length / 4;
    this.instanced = !!(desc.instances && desc.instances.length);
    this.handle = this.instanced
      ? native.bvhBuildInstanced(desc, desc.meshes, desc.instances, opts.threads || 0, !!opts.spatialSplits)
      : native.bvhBuild(desc.positions, desc.triangles, opts.threads || 0, !!opts.spatialSplits);
    const info = native.bvhInfo(this.handle);
    this.info = info;
    this.depth = info.binaryDepth;
    this._ref = null;
  }

  getNumLeafTris() {
    return this.info.leafTriangles;
  }

  exportReference() {
    if (!this._ref) this._ref = this.native.bvhExportReference(this.handle, this.numVertices, this.numTriangles);
    return this._ref;
  }

  normalizer() {
    return this.native.bvhNormalizer(this.handle);
  }

  destroy() {
    if (this.handle) this.native.bvhDestroy(this.handle);
    this.handle = null;
  }
}

module.exports = { BVH };
