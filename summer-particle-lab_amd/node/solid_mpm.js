'use strict';
// Drop-in `SolidMPM` for Summer-Particle-Lab's 2D elastic MPM
// (src/solid-mpm.js), backed by the MI355X 2D engine (mpm2_* in
// include/mpm_engine.h) through the N-API addon.
//
// Same surface: resetParticle(i), removeParticle(i, last), reset(),
// hasParticles(s), stableStep(s), step(s, dt), create(s, opts), stats(s),
// and the public arrays f00..c11 (Float64Array). step(s, dt) reads the
// fields the reference step reads from `s` (count, x, y, vx, vy, mass,
// volume, material, pinned, settings.solidStiffness, settings.boundary,
// scene) and writes x, y, vx, vy and its own arrays back. advance(s, dt) is
// the driver's all-solid frame (src/simulation.js:240-300) with its substep
// loop on the device.
const native = require('./mpm_napi.node');

const FIELDS = ['f00', 'f01', 'f10', 'f11', 'c00', 'c01', 'c10', 'c11'];
const GRAVITY = {none: 0, uniform: 1, center: 2};
// Wall boxes (bound, floor, top) of src/solid-mpm.js:69-70.
const WALLS = {softbody: [170, -145, 145], basin: [245, -175, 245]};

// The reference's MPM solid is material 11 (src/materials.js, `mpm: true`).
const isMPMSolid = (mat) => mat === 11;

// 2D fixed-corotated Kirchhoff stress P F^T with the closed-form polar
// rotation (src/solid-mpm.js:10-18): [t00, t01, t10, t11].
function solidStress(a, b, c, d, mu, lambda) {
  const J = a * d - b * c, q = Math.hypot(a + d, c - b);
  const r = q > 1e-12 ? (a + d) / q : 1, t = q > 1e-12 ? (c - b) / q : 0;
  const vol = lambda * (J - 1) * J;
  return [2 * mu * ((a - r) * a + (b + t) * b) + vol, 2 * mu * ((a - r) * c + (b + t) * d),
    2 * mu * ((c - t) * a + (d - r) * b), 2 * mu * ((c - t) * c + (d - r) * d) + vol];
}

class SolidMPM {
  constructor(capacity, options = {}) {
    this.capacity = capacity;
    this.isMPMSolid = options.isMPMSolid || isMPMSolid;
    for (const k of FIELDS) this[k] = new Float64Array(capacity);
    this.engine = native.create2(capacity);
    this.reset();
  }

  resetParticle(i) { for (const k of FIELDS) this[k][i] = 0; this.f00[i] = this.f11[i] = 1; }
  removeParticle(i, last) { for (const k of FIELDS) this[k][i] = this[k][last]; }
  reset() {
    for (const k of FIELDS) this[k].fill(k === 'f00' || k === 'f11' ? 1 : 0);
    this.substeps = this.activeNodes = 0;
  }
  hasParticles(s) { for (let i = 0; i < s.count; i++) if (this.isMPMSolid(s.material[i])) return true; return false; }

  _settings(s) {
    const st = s.settings || {};
    const box = WALLS[s.scene] || [555, -555, 555];
    const stiffness = Number.isFinite(st.solidStiffness) ? st.solidStiffness : 1;
    return {stiffness, walls: st.boundary || s.scene === 'softbody' || s.scene === 'basin' ? 1 : 0,
      bound: box[0], floor: box[1], top: box[2]};
  }

  _upload(s) {
    const ids = [];
    for (let p = 0; p < s.count; p++) if (this.isMPMSolid(s.material[p])) ids.push(p);
    const n = ids.length, v = {count: n};
    for (const k of ['x', 'y', 'vx', 'vy', 'mass', 'volume']) v[k] = new Float64Array(n);
    v.F = new Float64Array(4 * n); v.C = new Float64Array(4 * n); v.pinned = new Uint8Array(n);
    for (let k = 0; k < n; k++) {
      const p = ids[k];
      v.x[k] = s.x[p]; v.y[k] = s.y[p]; v.vx[k] = s.vx[p]; v.vy[k] = s.vy[p];
      v.mass[k] = s.mass[p]; v.volume[k] = s.volume[p]; v.pinned[k] = s.pinned ? s.pinned[p] : 0;
      for (let j = 0; j < 4; j++) { v.F[4 * k + j] = this[FIELDS[j]][p]; v.C[4 * k + j] = this[FIELDS[4 + j]][p]; }
    }
    native.upload2(this.engine, v);
    this._ids = ids; this._view = v;
  }

  _download(s) {
    const ids = this._ids, v = this._view;
    native.download2(this.engine, v);
    for (let k = 0; k < ids.length; k++) {
      const p = ids[k];
      s.x[p] = v.x[k]; s.y[p] = v.y[k]; s.vx[p] = v.vx[k]; s.vy[p] = v.vy[k];
      for (let j = 0; j < 4; j++) { this[FIELDS[j]][p] = v.F[4 * k + j]; this[FIELDS[4 + j]][p] = v.C[4 * k + j]; }
    }
    this.activeNodes = native.stats2(this.engine).nodes;
  }

  stableStep(s) {
    this._upload(s);
    return native.stableDt2(this.engine, this._settings(s));
  }

  step(s, dt) {
    this._upload(s);
    native.step2(this.engine, this._settings(s), dt, 1);
    this._download(s);
  }

  // One frame of the 2D driver for a scene whose particles are all MPM solids
  // (src/simulation.js:240-300): gravity 'none', 'uniform' or 'center' from
  // s.settings.gravity and strength. Returns {substeps, advanced}.
  advance(s, dt = 1 / 60) {
    const st = s.settings || {};
    const mode = GRAVITY[st.gravity === undefined ? 'center' : st.gravity];
    if (mode === undefined) throw new Error(`gravity ${st.gravity} is not supported on the device`);
    this._upload(s);
    const r = native.advance2(this.engine, this._settings(s), dt, mode,
      Number.isFinite(st.strength) ? st.strength : 1, 96);
    this._download(s);
    this.substeps = r.substeps;
    s.time = (s.time || 0) + r.advanced;
    return r;
  }

  create(s, {x1, y1, x2, y2, material = 11, shape = 'panel', fixed = false}) {
    if (!this.isMPMSolid(material) || ![x1, y1, x2, y2].every(Number.isFinite)) return false;
    const cl = (x) => Math.max(-530, Math.min(530, x));
    x1 = cl(x1); x2 = cl(x2); y1 = cl(y1); y2 = cl(y2);
    const length = Math.hypot(x2 - x1, y2 - y1), panel = shape === 'panel';
    const width = panel ? Math.abs(x2 - x1) : length, height = panel ? Math.abs(y2 - y1) : 22;
    if (width < 7 || height < 7) return false;
    const nx = Math.floor(width / 3.6) + 1, ny = Math.floor(height / 3.6) + 1;
    if (nx * ny + s.count > s.x.length) return false;
    const ux = (x2 - x1) / length, uy = (y2 - y1) / length;
    for (let row = 0; row < ny; row++) for (let col = 0; col < nx; col++) {
      const x = panel ? x1 + Math.sign(x2 - x1) * col * 3.6 : x1 + ux * col * 3.6 - uy * (row * 3.6 - height / 2);
      const y = panel ? y1 + Math.sign(y2 - y1) * row * 3.6 : y1 + uy * col * 3.6 + ux * (row * 3.6 - height / 2);
      s.add(x, y, material);
      if (fixed && col < 2) s.structures.pin(s.count - 1, true);
    }
    return true;
  }

  stats(s) {
    let count = 0, min = Infinity, max = 0, strain = 0;
    for (let p = 0; p < s.count; p++) if (this.isMPMSolid(s.material[p])) {
      count++;
      const a = this.f00[p], b = this.f01[p], c = this.f10[p], d = this.f11[p], J = a * d - b * c;
      min = Math.min(min, J); max = Math.max(max, J);
      strain = Math.max(strain, .5 * Math.hypot(a * a + c * c - 1, b * b + d * d - 1, Math.SQRT2 * (a * b + c * d)));
    }
    return {count, minJ: count ? min : 1, maxJ: count ? max : 1, strain, nodes: this.activeNodes,
      substeps: this.substeps};
  }

  destroy() { if (this.engine) native.destroy2(this.engine); this.engine = null; }
}

module.exports = {SolidMPM, solidStress, isMPMSolid};
