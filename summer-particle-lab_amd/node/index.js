'use strict';
// Drop-in `MPM3` for Summer-Particle-Lab's 3D MPM step API
// (src/three/mpm.js:7-136), backed by the MI355X engine through the N-API
// addon (mpm_napi.node) and the C ABI (include/mpm_engine.h).
//
// Same method surface: configure(domain), reset(), resetParticle(p),
// remove(p, last), stableStep(s), step(s, dt), refreshDiagnostics(s), minJ(s),
// and the same public arrays F, C, Jp, dilation, stress, plastic, cachedTau,
// cacheValid (Float64Array/Uint8Array, 9 values per particle for tensors).
// `step(s, dt)` reads the fields the reference step reads from `s` (count,
// x..vz, mass, volume, model, pinned, isSolid, settings.stiffness, lab.part,
// temperature, plasticity[], thermalMaterials[]) and writes x..vz and its own
// arrays back, so callers and tests that mutate F/C between steps keep
// working. Reference units are kept: E = m/(64 V) 16000 stiffness,
// mu = lambda = 0.4 E, particle volume 64 V (src/three/mpm.js:34,37,68),
// separating walls (src/three/mpm.js:93), no gravity inside the step.
//
// Option `resident: true` keeps particle state on the device between steps
// (upload once, call `download(s)` to read it back); the default copies state
// in and out on every step like the reference's in-place arrays.
//
// Option `guard: true` turns on the device state guard with the reference's
// stable-step constants (cap h/8/240, Courant number 1.4/8): `checkState()`
// then returns the GPU service's status code, 7 (CFL exceeded) or 8 (invalid
// state), or 0 (gpu/mixed_solver.py:381-396). Off by default, because the
// reference's MPM3.step never throws.
//
// Scene fields honoured besides the reference step's inputs: `s.press` (the
// platen on the grid, src/three/mpm.js:97-101, through a split step),
// `s.profile` (device kernel times added to its `transfer` and `grid`
// buckets; the constitutive update is fused into the transfer kernel, so
// `constitutive` gains nothing). `s.energyObserver` and
// `s.mechanicalDiagnostics` need the host-side grid of the reference step
// (unforced momenta, per-node wall work) and are refused with an error.
const native = require('./mpm_napi.node');

const DEFAULT_DOMAIN = Object.freeze({lower: [-64, -48, -64], upper: [64, 72, 64], cellSize: 8, particlesPerAxis: 2});
const THERMAL_KEYS = ['Tref', 'Thot', 'hotRatio', 'A', 'H', 'rate', 'Q', 'n', 'stress'];
const PLASTIC_KEYS = ['A', 'B', 'n', 'C', 'rate0', 'Tref', 'Tm', 'm'];
const PRESSURE_KEYS = ['cohesion', 'slope', 'dilatancy', 'tension', 'alpha', 'plateau', 'gamma', 'densify', 'beta',
  'epsD'];
const SUPPORTED = new Set([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11]);

function identity() { return [1, 0, 0, 0, 1, 0, 0, 0, 1]; }
function det(a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

class MPM3 {
  constructor(capacity, options = {}) {
    this.capacity = capacity;
    this.options = Object.assign({sortEvery: 8, resident: false, guard: false}, options);
    this.F = new Float64Array(capacity * 9); this.C = new Float64Array(capacity * 9);
    this.Jp = new Float64Array(capacity); this.dilation = new Float64Array(capacity);
    this.stress = new Float64Array(capacity); this.plastic = new Float64Array(capacity);
    this.tau = new Float64Array(capacity * 9); this.cachedTau = new Float64Array(capacity * 9);
    this.cacheValid = new Uint8Array(capacity);
    this.engine = null;
    this.configure(this.options.domain || DEFAULT_DOMAIN);
    this.reset();
  }

  configure(domain) {
    this.domain = domain; this.H = domain.cellSize;
    // The grid's origin and edge (src/three/discretization.js:13): two guard
    // cells below, nodes = ceil((upper - origin) / h) + 3 per axis.
    this.O = domain.lower.map((v) => Math.floor(v / domain.cellSize) * domain.cellSize - 2 * domain.cellSize);
    this.N = Math.max(...domain.upper.map((v, a) => Math.ceil((v - this.O[a]) / domain.cellSize) + 3));
    this.mass = this.v = this.fixed = null; this.base = new Int32Array(this.capacity * 3);
    if (this.engine) native.destroy(this.engine);
    this.engine = native.create({lower: domain.lower, upper: domain.upper, cellSize: domain.cellSize,
      capacity: this.capacity, wallMode: [0, 0, 0, 0, 0, 0], gravity: [0, 0, 0], gravityOnParticles: 0,
      massEps: 1e-12, sortEvery: this.options.sortEvery, strictDomain: 0,
      guardOff: this.options.guard ? 0 : 1, cflNumber: 1.4 / 8, cflCap: domain.cellSize / 8 / 240});
    this.onDevice = 0;
  }

  // The state guard's finding since the last call: 0, 7 (CFL exceeded) or 8
  // (invalid state), as the GPU service reports them.
  checkState() { return native.checkState(this.engine); }

  reset() { this.C.fill(0); this.F.fill(0); for (let p = 0; p < this.F.length / 9; p++) this.resetParticle(p); }

  resetParticle(p) {
    this.F.set(identity(), p * 9); this.C.fill(0, p * 9, p * 9 + 9); this.Jp[p] = 1; this.dilation[p] = 0;
    this.stress[p] = 0; this.plastic[p] = 0; this.cacheValid[p] = 0; this.onDevice = 0;
  }

  remove(p, last) {
    for (const key of ['F', 'C', 'cachedTau']) this[key].copyWithin(p * 9, last * 9, last * 9 + 9);
    for (const key of ['Jp', 'dilation', 'stress', 'plastic', 'cacheValid']) this[key][p] = this[key][last];
    this.onDevice = 0;
  }

  // Particles the step acts on and the material table of (model, part, E).
  _gather(s) {
    const ids = [];
    const keys = new Map(), materials = [];
    const material = new Uint8Array(s.count);
    for (let p = 0; p < s.count; p++) {
      if (s.isSolid && !s.isSolid(p)) continue;
      if (this._part !== undefined && (s.lab && s.lab.part ? s.lab.part[p] : 0) !== this._part) continue;
      const model = s.model ? s.model[p] : 0;
      if (!SUPPORTED.has(model)) throw new Error(`model ${model} is not supported by the MI355X engine`);
      const part = s.lab && s.lab.part ? s.lab.part[p] : 0;
      const E = s.mass[p] / (64 * s.volume[p]) * 16000 * ((s.settings && s.settings.stiffness) || 1);
      // Snow's yield strain and sand's friction angle are per particle in the
      // reference (src/three/simulation.js:40) and constant within a part.
      const ys = s.yieldStrain ? s.yieldStrain[p] : 0.025, fr = s.friction ? s.friction[p] : 35;
      // Frame code of the particle's material: 11 (MPM solid) shows as 11 + model.
      const code = s.material ? s.material[p] : 11;
      const key = `${model}|${part}|${E}|${ys}|${fr}|${code}`;
      let m = keys.get(key);
      if (m === undefined) {
        m = materials.length;
        if (m >= 64) throw new Error('more than 64 distinct materials');
        keys.set(key, m);
        const th = (s.thermalMaterials && s.thermalMaterials[part]) || {};
        const pl = (s.plasticity && s.plasticity[part]) || {};
        const pr = (s.pressureMaterials && s.pressureMaterials[part]) || {};
        const fm = (s.fractureMaterials && s.fractureMaterials[part]) || {};
        const thermal = {}, plastic = {}, pressure = {};
        for (const k of THERMAL_KEYS) if (Number.isFinite(th[k])) thermal[k] = th[k];
        for (const k of PLASTIC_KEYS) if (Number.isFinite(pl[k])) plastic[k] = pl[k];
        for (const k of PRESSURE_KEYS) if (Number.isFinite(pr[k])) pressure[k] = pr[k];
        // Damage is the scene's (s.damage, src/three/fracture.js:11-18).
        const fracture = {enabled: fm.enabled === true && !!s.damage};
        if (Number.isFinite(fm.onset)) fracture.onset = fm.onset;
        if (Number.isFinite(fm.failure)) fracture.failure = fm.failure;
        materials.push({model, E, mu: 0.4 * E, lambda: 0.4 * E, thermal, plastic, pressure, fracture,
          yieldStrain: ys, friction: fr, code, part});
      }
      material[ids.length] = m;
      ids.push(p);
    }
    return {ids, materials, material: material.subarray(0, ids.length)};
  }

  _view(n) {
    const v = {count: n};
    for (const k of ['x', 'y', 'z', 'vx', 'vy', 'vz', 'mass', 'volume', 'temperature', 'plastic', 'jp', 'dilation', 'damage']) v[k] = new Float64Array(n);
    v.F = new Float64Array(n * 9); v.C = new Float64Array(n * 9); v.cachedTau = new Float64Array(n * 9);
    v.pinned = new Uint8Array(n); v.cacheValid = new Uint8Array(n); v.material = new Uint8Array(n);
    return v;
  }

  // Copy the step's inputs from `s` and this object to the device.
  upload(s) {
    const {ids, materials, material} = this._gather(s);
    const n = ids.length, v = this._view(n);
    this._scale = new Float64Array(n);
    this._snow = new Float64Array(n);
    for (let k = 0; k < n; k++) {
      const p = ids[k];
      v.x[k] = s.x[p]; v.y[k] = s.y[p]; v.z[k] = s.z[p]; v.vx[k] = s.vx[p]; v.vy[k] = s.vy[p]; v.vz[k] = s.vz[p];
      v.mass[k] = s.mass[p]; v.volume[k] = 64 * s.volume[p];
      v.temperature[k] = s.temperature ? s.temperature[p] : 20; v.plastic[k] = this.plastic[p];
      v.jp[k] = this.Jp[p]; v.dilation[k] = this.dilation[p]; v.damage[k] = s.damage ? s.damage[p] : 0;
      v.pinned[k] = s.pinned ? s.pinned[p] : 0; v.material[k] = material[k];
      v.cacheValid[k] = this.cacheValid[p];
      // Cached return stress: unit-modulus in the reference, absolute here
      // (and for snow times its hardening, src/three/mpm.js:35-37).
      const mat = materials[material[k]];
      this._snow[k] = mat.model === 2 ? Math.min(4, Math.max(.25, Math.exp(10 * (1 - this.Jp[p])))) : 0;
      this._scale[k] = 0.4 * mat.E * (this._snow[k] || 1);
      for (let j = 0; j < 9; j++) {
        v.F[k * 9 + j] = this.F[p * 9 + j]; v.C[k * 9 + j] = this.C[p * 9 + j];
        v.cachedTau[k * 9 + j] = this.cachedTau[p * 9 + j] * this._scale[k];
      }
    }
    native.setMaterials(this.engine, materials.length ? materials : [{model: 0, E: 1, mu: .4, lambda: .4}]);
    native.upload(this.engine, v);
    this._ids = ids; this._view_ = v; this._materials = materials; this.onDevice = 1;
  }

  // Copy device state back into `s` and this object.
  download(s) {
    const ids = this._ids, v = this._view_;
    if (!ids) return;
    native.download(this.engine, v);
    for (let k = 0; k < ids.length; k++) {
      const p = ids[k];
      s.x[p] = v.x[k]; s.y[p] = v.y[k]; s.z[p] = v.z[k]; s.vx[p] = v.vx[k]; s.vy[p] = v.vy[k]; s.vz[p] = v.vz[k];
      this.plastic[p] = v.plastic[k]; this.cacheValid[p] = v.cacheValid[k];
      this.Jp[p] = v.jp[k]; this.dilation[p] = v.dilation[k];
      // A failed particle steps on as sand (src/three/fracture.js:17).
      const failed = v.damage[k] >= 1;
      if (s.damage) s.damage[p] = v.damage[k];
      if (failed && s.model) s.model[p] = 3;
      // The device's cached snow stress carries the hardening of the new Jp.
      const scale = this._snow && this._snow[k] && !failed ? this._scale[k] / this._snow[k] *
        Math.min(4, Math.max(.25, Math.exp(10 * (1 - v.jp[k])))) : this._scale[k] / (this._snow[k] || 1);
      for (let j = 0; j < 9; j++) {
        this.F[p * 9 + j] = v.F[k * 9 + j]; this.C[p * 9 + j] = v.C[k * 9 + j];
        if (v.cacheValid[k]) this.cachedTau[p * 9 + j] = v.cachedTau[k * 9 + j] / scale;
      }
    }
  }

  stableStep(s) {
    if (!this.options.resident || !this.onDevice) this.upload(s);
    // Defaults of the C ABI are the reference constants h/8/240 and 1.4/8.
    return native.stableDt(this.engine, 0, 0);
  }

  step(s, dt) {
    if (s.energyObserver) {
      throw new Error('MPM3 on the MI355X engine: s.energyObserver needs the host grid of the reference step ' +
        '(unforced momenta); it is not supported');
    }
    if (s.mechanicalDiagnostics) {
      throw new Error('MPM3 on the MI355X engine: s.mechanicalDiagnostics needs the per-node wall work of the ' +
        'reference grid loop; it is not supported');
    }
    const profiled = !!s.profile;
    if (profiled) native.profile(this.engine, 1, 1);
    if (s.separatePartGrids) {
      this._stepParts(s, dt);
    } else {
      if (!this.options.resident || !this.onDevice) this.upload(s);
      // The stress this step applies (src/three/mpm.js:37), as the reference
      // leaves it in this.tau; resident mode keeps it on the device.
      if (!this.options.resident) this._readTau();
      if (s.externalGridContact || s.contactGridDeltas || (s.press && s.press.enabled)) this._splitStep(s, dt);
      else native.step(this.engine, dt, 1);
      if (!this.options.resident) this.download(s);
    }
    if (profiled) this._addProfile(s.profile);
  }

  // this.tau of the particles on the device (9 values each, reference units).
  _readTau() {
    const ids = this._ids;
    if (!ids || !ids.length) return;
    const t = new Float64Array(ids.length * 9);
    native.stressTensors(this.engine, t);
    for (let k = 0; k < ids.length; k++) this.tau.set(t.subarray(k * 9, k * 9 + 9), ids[k] * 9);
  }

  // Device kernel times into the reference's profile buckets (milliseconds,
  // src/three/mpm.js:62,86,105,124,133).
  _addProfile(profile) {
    const engines = this._children ? [...this._children.values()].map((c) => c.engine) : [this.engine];
    for (const e of engines) {
      const t = native.kernelTimes(e);
      native.profile(e, 0, 1);
      for (const [name, ms] of Object.entries(t)) {
        const bucket = name === 'grid' || name === 'pieces' ? 'grid' : 'transfer';
        profile[bucket] = (profile[bucket] || 0) + ms;
      }
    }
  }

  // The platen of the reference's grid loop (src/three/mpm.js:97-101) on a
  // downloaded grid {mass, v, fixed}: nodes at or above press.y - H/2 that no
  // pinned particle holds move down at most at press.velocity, and the
  // velocity taken away, times the node mass, adds to press.impulse.
  _applyPress(s, f) {
    const press = s.press, N = this.N, H = this.H, O1 = this.O[1];
    const y0 = press.y - H * .5;
    for (let n = 0; n < f.mass.length; n++) {
      if (!(f.mass[n] > 1e-12) || f.fixed[n]) continue;
      if (O1 + (Math.floor(n / N) % N) * H < y0) continue;
      const before = f.v[n * 3 + 1], after = Math.min(before, press.velocity);
      f.v[n * 3 + 1] = after;
      press.impulse += (before - after) * f.mass[n];
    }
  }

  // One engine per part for s.separatePartGrids (src/three/mpm.js:58,69):
  // parts never meet on a grid, so each part's particles step on their own
  // engine and grid; the children share this object's per-particle arrays.
  _child(part) {
    this._children = this._children || new Map();
    let c = this._children.get(part);
    if (!c) {
      c = new MPM3(this.capacity, Object.assign({}, this.options, {domain: this.domain, resident: false}));
      for (const k of ['F', 'C', 'Jp', 'dilation', 'stress', 'plastic', 'tau', 'cachedTau', 'cacheValid']) c[k] = this[k];
      c._part = part;
      this._children.set(part, c);
    }
    return c;
  }

  // The hooks see this object with `fields[part]` = {mass, v, fixed}, as the
  // reference's part-separated grids (src/three/mpm.js:58).
  _stepParts(s, dt) {
    if (this.options.resident) throw new Error('separatePartGrids needs the default (non-resident) mode');
    const parts = new Set();
    for (let p = 0; p < s.count; p++) if (!s.isSolid || s.isSolid(p)) parts.add(s.lab && s.lab.part ? s.lab.part[p] : 0);
    const kids = [...parts].sort((a, b) => a - b).map((k) => this._child(k));
    const n3 = this.N * this.N * this.N;
    for (const c of kids) { c.upload(s); c._readTau(); native.stepToGrid(c.engine, dt); }
    const pressed = !!(s.press && s.press.enabled);
    if (pressed && !s.externalGridContact) {
      const f = {mass: new Float64Array(n3), v: new Float64Array(3 * n3), fixed: new Uint8Array(n3)};
      for (const c of kids) {
        native.gridDownload(c.engine, f.mass, f.v, f.fixed);
        this._applyPress(s, f);
        native.gridUploadVelocity(c.engine, f.v);
      }
    }
    if (s.externalGridContact) {
      const count = Math.max((s.projectPartCount || 0) + 1, ...kids.map((c) => c._part + 1));
      if (!this.fields || this.fields.length !== count || !this.fields[0] || this.fields[0].mass.length !== n3)
        this.fields = Array.from({length: count}, () => ({mass: new Float64Array(n3), v: new Float64Array(3 * n3),
          fixed: new Uint8Array(n3)}));
      if (!this.mass || this.mass.length !== n3) {
        this.mass = new Float64Array(n3); this.v = new Float64Array(3 * n3); this.fixed = new Uint8Array(n3);
      }
      for (const f of this.fields) { f.mass.fill(0); f.v.fill(0); f.fixed.fill(0); }
      for (const c of kids) {
        const f = this.fields[c._part];
        native.gridDownload(c.engine, f.mass, f.v, f.fixed);
        if (pressed) this._applyPress(s, f);
      }
      for (let p = 0; p < s.count; p++) {
        const x = [s.x[p], s.y[p], s.z[p]];
        for (let a = 0; a < 3; a++) {
          this.base[p * 3 + a] = Math.max(0, Math.min(this.N - 3, Math.floor((x[a] - this.O[a]) / this.H - .5)));
        }
      }
      s.externalGridContact(this, dt);
      for (const c of kids) native.gridUploadVelocity(c.engine, this.fields[c._part].v);
    }
    for (const c of kids) {
      const d = s.contactGridDeltas ? s.contactGridDeltas[c._part] : null;
      const slots = new Int32Array(64).fill(d ? 0 : -1);
      native.stepFromGrid(c.engine, dt, slots, d ? Float64Array.from(d.subarray(0, 3 * n3)) : null, d ? 1 : 0);
      c.download(s);
    }
  }

  // A step with the reference's grid hooks (src/three/mpm.js:104,114): P2G
  // and the grid update on the device, then s.externalGridContact(this, dt)
  // on this object's mass, v, fixed, base, N, O, H (the reference's own
  // fields), the changed velocities back to the device, and G2P with each
  // part's s.contactGridDeltas added.
  _splitStep(s, dt) {
    native.stepToGrid(this.engine, dt);
    const n3 = this.N * this.N * this.N;
    const pressed = !!(s.press && s.press.enabled);
    if (s.externalGridContact || pressed) {
      if (!this.mass || this.mass.length !== n3) {
        this.mass = new Float64Array(n3); this.v = new Float64Array(3 * n3); this.fixed = new Uint8Array(n3);
      }
      native.gridDownload(this.engine, this.mass, this.v, this.fixed);
      if (pressed) this._applyPress(s, this);
      if (s.externalGridContact) {
        for (let p = 0; p < s.count; p++) {  // the step's stencil bases (src/three/mpm.js:25-28)
          const x = [s.x[p], s.y[p], s.z[p]];
          for (let a = 0; a < 3; a++) {
            this.base[p * 3 + a] = Math.max(0, Math.min(this.N - 3, Math.floor((x[a] - this.O[a]) / this.H - .5)));
          }
        }
        s.externalGridContact(this, dt);
      }
      native.gridUploadVelocity(this.engine, this.v);
    }
    const slots = new Int32Array(64).fill(-1), arrays = [];
    if (s.contactGridDeltas) {
      this._materials.forEach((m, k) => {
        const d = s.contactGridDeltas[m.part];
        if (!d) return;
        if (d.length < 3 * n3) throw new Error(`contactGridDeltas[${m.part}] needs ${3 * n3} values`);
        let slot = arrays.indexOf(d);
        if (slot < 0) { slot = arrays.length; arrays.push(d); }
        slots[k] = slot;
      });
    }
    let deltas = null;
    if (arrays.length) {
      deltas = new Float64Array(arrays.length * 3 * n3);
      arrays.forEach((d, k) => deltas.set(d.subarray(0, 3 * n3), k * 3 * n3));
    }
    native.stepFromGrid(this.engine, dt, slots, deltas, arrays.length);
  }

  // One frame of Simulation3.step for an all-solid scene, run on the device
  // (src/three/simulation.js:305-388 with regularizeSolids,
  // src/three/solid-sampling.js:10-42): adaptive substeps, particle gravity,
  // bounds, velocity reconstruction and anti-clump. Reads s.settings,
  // s.boundsLo/boundsHi and s.scene like the reference; returns
  // {substeps, advanced} and advances s.time.
  advanceFrame(s, dt = 1 / 60) {
    if (!this.options.resident || !this.onDevice) this.upload(s);
    const st = s.settings || {};
    const DIRS = {down: [0, -1, 0], up: [0, 1, 0], left: [-1, 0, 0], right: [1, 0, 0], front: [0, 0, 1], back: [0, 0, -1]};
    let mode = st.gravityMode === 'scene' || st.gravityMode === undefined
      ? (s.scene === 'million' ? 'none' : 'uniform') : st.gravityMode;
    if (st.gravity === false || st.gravityStrength === 0) mode = 'none';
    const code = {none: 0, uniform: 1, attract: 2, repel: 3}[mode];
    if (code === undefined) throw new Error(`gravity mode ${mode} is not supported on the device`);
    const num = (v, d) => (Number.isFinite(v) ? v : d);
    const settings = {timeStepScale: num(st.timeStepScale, 1), antiClump: st.solidAntiClump === false ? 0 : 1,
      gravityMode: code, gravityStrength: num(st.gravityStrength, 1),
      gravityDirection: DIRS[st.gravityDirection || 'down'], gravityCenter: st.gravityCenter || [0, 0, 0],
      gravitySoftening: num(st.gravitySoftening, 8)};
    if (s.boundsLo) settings.boundsLo = Array.from(s.boundsLo);
    if (s.boundsHi) settings.boundsHi = Array.from(s.boundsHi);
    const r = native.advanceFrame(this.engine, dt, settings);
    if (!this.options.resident) this.download(s);
    s.time = (s.time || 0) + r.advanced; s.substeps = r.substeps;
    return r;
  }

  // The service's frame of the device state (gpu/solver.py:986-1032):
  // {state, data} with the header totals and one preview row of 8 floats
  // (x, y, z, material code, T, pinned, stress, plastic) per block of
  // sampleStride particles. Rows follow the solid particles' order.
  snapshot(s, previewLimit = 100000) {
    if (!this.options.resident || !this.onDevice) this.upload(s);
    const codes = this._materials.map((m) => m.code);
    const r = native.snapshot(this.engine, previewLimit, codes);
    if (r.nonfinite) throw new Error('Non-finite GPU state; simulation stopped');
    const data = r.data;
    delete r.data;
    const state = Object.assign({type: 'frame', capacity: this.capacity, energy: 0, time: s.time || 0,
      substeps: s.substeps || 0}, r);
    delete state.nonfinite;
    return {state, data};
  }

  // One binary frame packet (gpu/server.py:17-20) as an ArrayBuffer.
  framePacket(state, data) { return native.framePacket(JSON.stringify(state), data); }

  refreshDiagnostics(s) {
    if (!this.options.resident || !this.onDevice) this.upload(s);
    const out = new Float64Array(this._ids.length);
    native.refreshDiagnostics(this.engine, out);
    this.stress.fill(0, 0, s.count);
    for (let k = 0; k < this._ids.length; k++) this.stress[this._ids[k]] = out[k];
    this._readTau();  // updateStress also leaves this.tau (src/three/mpm.js:37)
  }

  minJ(s) {
    let value = 1;
    for (let i = 0; i < s.count; i++) if (!s.isSolid || s.isSolid(i)) value = Math.min(value, det(this.F.subarray(i * 9, i * 9 + 9)));
    return value;
  }

  conservation() { return native.conservation(this.engine); }

  destroy() {
    if (this.engine) native.destroy(this.engine);
    this.engine = null;
    if (this._children) for (const c of this._children.values()) c.destroy();
    this._children = null;
  }
}

const {SolidMPM, solidStress, isMPMSolid} = require('./solid_mpm.js');

module.exports = {MPM3, DEFAULT_DOMAIN, native, SolidMPM, solidStress, isMPMSolid};
