'use strict';
// Drives the drop-in MPM3 (summer-particle-lab_amd/node) on a scene given as
// JSON {points, velocity, models, F, dt, steps, thermal?} in reference units
// and prints the resulting state as JSON. The scene object carries exactly
// the fields MPM3.step reads in the reference (src/three/mpm.js:56-134).
const path = require('path');
const fs = require('fs');
const {MPM3} = require(path.join(__dirname, '..', '..', 'summer-particle-lab_amd', 'node', 'index.js'));

const input = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const n = input.points.length;
const s = {count: n, settings: {stiffness: 1}, lab: {part: new Uint8Array(n)},
  thermalMaterials: [input.thermal || {}], plasticity: [input.plastic || {}], isSolid: () => true};
for (const k of ['x', 'y', 'z', 'vx', 'vy', 'vz', 'mass', 'volume', 'temperature']) s[k] = new Float64Array(n);
s.model = new Uint8Array(n); s.pinned = new Uint8Array(n);
for (let i = 0; i < n; i++) {
  [s.x[i], s.y[i], s.z[i]] = input.points[i];
  [s.vx[i], s.vy[i], s.vz[i]] = Array.isArray(input.velocity[0]) ? input.velocity[i] : input.velocity;
  s.mass[i] = 1.1; s.volume[i] = 1; s.temperature[i] = input.temperature || 20; s.model[i] = input.models[i];
}
if (input.pinned) for (const i of input.pinned) s.pinned[i] = 1;
if (input.part) for (let i = 0; i < n; i++) s.lab.part[i] = input.part[i];
if (input.settings) Object.assign(s.settings, input.settings);
if (input.boundsLo) { s.boundsLo = input.boundsLo; s.boundsHi = input.boundsHi; }
const mpm = new MPM3(n, {sortEvery: 1, guard: !!input.guard});
// The platen (src/three/mpm.js:97-101) and the profile buckets (:62-133).
if (input.press) s.press = Object.assign({enabled: true, impulse: 0}, input.press);
if (input.profile) s.profile = {constitutive: 0, transfer: 0, grid: 0};
// Grid hooks (src/three/mpm.js:104,114): a rigid plate below plateY that
// stops nodes moving into it, and a constant velocity delta for one part.
if (input.plateY !== undefined) {
  s.externalGridContact = (grid, dt) => {
    const N = grid.N;
    for (let node = 0; node < N * N * N; node++) {
      const y = grid.O[1] + (Math.floor(node / N) % N) * grid.H;
      if (grid.mass[node] > 1e-10 && y < input.plateY && grid.v[node * 3 + 1] < 0) grid.v[node * 3 + 1] = 0;
    }
  };
}
if (input.separate) { s.separatePartGrids = true; s.projectPartCount = Math.max(...input.part); }
if (input.deltaPart !== undefined) {
  const d = new Float64Array(3 * mpm.N ** 3);
  for (let k = 0; k < d.length; k += 3) d.set(input.delta, k);
  s.contactGridDeltas = [];
  s.contactGridDeltas[input.deltaPart] = d;
}
if (input.F) for (let i = 0; i < n; i++) mpm.F.set(input.F, i * 9);
const dtStable = mpm.stableStep(s);
const frames = [];
if (input.frames) for (let k = 0; k < input.frames; k++) frames.push(mpm.advanceFrame(s, input.dt));
else for (let k = 0; k < input.steps; k++) mpm.step(s, input.dt);
// this.tau as the last step left it: the stress of its start state.
const tauStep = Array.from(mpm.tau.subarray(0, n * 9));
const status = input.guard ? mpm.checkState() : null;
mpm.refreshDiagnostics(s);
let snapshot = null;
if (input.snapshot) {
  const {state, data} = mpm.snapshot(s, input.snapshot);
  snapshot = {state, data: Array.from(data), packet: Buffer.from(mpm.framePacket(state, data)).toString('base64')};
}
const out = {dtStable, x: Array.from(s.x), y: Array.from(s.y), z: Array.from(s.z), vx: Array.from(s.vx),
  vy: Array.from(s.vy), vz: Array.from(s.vz), F: Array.from(mpm.F.subarray(0, n * 9)),
  plastic: Array.from(mpm.plastic.subarray(0, n)), stress: Array.from(mpm.stress.subarray(0, n)),
  jp: Array.from(mpm.Jp.subarray(0, n)), dilation: Array.from(mpm.dilation.subarray(0, n)),
  minJ: mpm.minJ(s), conservation: mpm.conservation(), frames, snapshot, tauStep,
  tau: Array.from(mpm.tau.subarray(0, n * 9)), press: s.press || null, profile: s.profile || null, status};
mpm.destroy();
console.log(JSON.stringify(out));
