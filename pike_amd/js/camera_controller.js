'use strict';
// CameraController (render/camera_controller.js:5-98): mouse-look, wheel
// field of view and WASD/RF flying for a camera ray {origin, dir} in the
// tracer's normalized space, with the reference's constants:
//   drag by (dx, dy) pixels: dir = rotateY(dir, dx / 180), then a rotation
//     by -dy / 180 about normalize(dir x up) (:21-32);
//   wheel: fovScale -= wheelDelta / 1200 * fovScale (:37-40);
//   per integration step of dt seconds, each held key moves the origin by dt
//     along dir (w/s), the strafe axis normalize(dir x up) (d/a) or
//     normalize(dir x strafe) (f/r) (:67-97).
// The reference listens on a canvas and the document and integrates on
// requestAnimationFrame; headless hosts call mouseDown/mouseMove/mouseUp/
// wheel/keyDown/keyUp and integrate(nowMs) themselves.  attach(canvas, doc)
// wires DOM-style event targets the way the reference does.
const { rotateArbitrary } = require('./scene');

const KEYS = new Set(['w', 'a', 's', 'd', 'r', 'f']);
const UP = [0, 1, 0];

function normalize(v) {
  const l = Math.hypot(v[0], v[1], v[2]);
  return [v[0] / l, v[1] / l, v[2] / l];
}

function cross(a, b) {
  return [a[1] * b[2] - a[2] * b[1], -(a[0] * b[2] - a[2] * b[0]), a[0] * b[1] - a[1] * b[0]];
}

// rotation about +y (vector.js:211-218)
function rotateY(v, a) {
  return [v[2] * Math.sin(a) + v[0] * Math.cos(a), v[1], v[2] * Math.cos(a) - v[0] * Math.sin(a)];
}

class CameraController {
  // ray: {origin, dir} (mutated in place, as the reference does);
  // onChange({ray, fov}) runs after a wheel step and after an integration
  // step with keys or the mouse active
  constructor(ray, onChange, nowMs) {
    this.ray = ray;
    this.onChange = onChange || (() => {});
    this.activeEvents = new Set();
    this.fovScale = 1;
    this.time = nowMs === undefined ? Date.now() : nowMs;
    this._drag = null;
  }

  mouseDown(x, y, button) {
    // the reference tracks the left button only (e.which === 1)
    this._drag = (button === undefined || button === 0) ? [x, y] : null;
  }

  mouseMove(x, y) {
    if (!this._drag) return;
    this.activeEvents.add('mouse');
    const rx = (this._drag[0] - x) / 180.0;
    const ry = (this._drag[1] - y) / 180.0;
    this.ray.dir = normalize(rotateY(this.ray.dir, rx));
    const axis = normalize(cross(this.ray.dir, UP));
    this.ray.dir = normalize(rotateArbitrary(this.ray.dir, axis, -ry));
    this._drag = [x, y];
  }

  mouseUp() {
    this._drag = null;
    this.activeEvents.delete('mouse');
  }

  wheel(wheelDelta) {
    this.fovScale -= wheelDelta / 1200 * this.fovScale;
    this.onChange({ ray: this.ray, fov: this.fovScale });
  }

  keyDown(key) {
    if (KEYS.has(key)) this.activeEvents.add(key);
  }

  keyUp(key) {
    if (KEYS.has(key)) this.activeEvents.delete(key);
  }

  getCameraRay() {
    return this.ray;
  }

  getFov() {
    return this.fovScale;
  }

  // one step of the reference's _integrate with dt = (nowMs - last) / 1000
  integrate(nowMs) {
    const now = nowMs === undefined ? Date.now() : nowMs;
    const delta = (now - this.time) / 1000;
    this.time = now;
    const strafe = normalize(cross(this.ray.dir, UP));
    const move = (v, s) => {
      this.ray.origin = [this.ray.origin[0] + v[0] * s, this.ray.origin[1] + v[1] * s,
        this.ray.origin[2] + v[2] * s];
    };
    for (const key of this.activeEvents) {
      switch (key) {
        case 'w': move(this.ray.dir, delta); break;
        case 'a': move(strafe, -delta); break;
        case 's': move(this.ray.dir, -delta); break;
        case 'd': move(strafe, delta); break;
        case 'r': move(normalize(cross(this.ray.dir, strafe)), -delta); break;
        case 'f': move(normalize(cross(this.ray.dir, strafe)), delta); break;
        default: break;
      }
    }
    if (this.activeEvents.size) this.onChange({ ray: this.ray, fov: this.fovScale });
  }

  // DOM-style wiring (camera_controller.js:16-51): canvas mouse events and
  // document key events
  attach(canvas, doc) {
    canvas.addEventListener('mousedown', (e) => this.mouseDown(e.layerX, e.layerY, e.which === 1 ? 0 : 1), false);
    canvas.addEventListener('mousemove', (e) => this.mouseMove(e.layerX, e.layerY), false);
    canvas.addEventListener('mouseup', () => this.mouseUp(), false);
    canvas.addEventListener('mousewheel', (e) => this.wheel(e.wheelDelta), false);
    doc.addEventListener('keypress', (e) => this.keyDown(e.key), false);
    doc.addEventListener('keyup', (e) => this.keyUp(e.key), false);
    return this;
  }
}

module.exports = { CameraController, rotateY };
