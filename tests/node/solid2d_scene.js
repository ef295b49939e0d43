'use strict';
// Drives the drop-in SolidMPM (summer-particle-lab_amd/node) on a 2D scene
// given as JSON {panels: [{x1, y1, x2, y2, fixed}], scene, frames, gravity,
// boundary, mode} and prints the resulting state as JSON. The scene object
// carries the fields the reference's SolidMPM reads (src/solid-mpm.js:38-92)
// in the reference's storage types (float32 particle arrays,
// src/simulation.js:18-19). mode 'frames' runs advance(s, 1/60); 'steps'
// runs the reference driver's loop of stableStep and step on the host
// (src/simulation.js:246-254) without gravity.
const path = require('path');
const fs = require('fs');
const {SolidMPM} = require(path.join(__dirname, '..', '..', 'summer-particle-lab_amd', 'node', 'index.js'));

const input = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const cap = 4096;
const s = {count: 0, scene: input.scene || 'empty', time: 0,
  settings: {solidStiffness: 1, gravity: input.gravity || 'none', strength: 1, boundary: !!input.boundary}};
for (const k of ['x', 'y', 'vx', 'vy', 'mass', 'volume']) s[k] = new Float32Array(cap);
s.material = new Uint8Array(cap); s.pinned = new Uint8Array(cap);
const mpm = new SolidMPM(cap);
s.add = (x, y, material) => {
  const i = s.count++;
  mpm.resetParticle(i);
  s.x[i] = x; s.y[i] = y; s.vx[i] = 0; s.vy[i] = 0; s.material[i] = material; s.mass[i] = 1.1; s.volume[i] = 1;
  s.pinned[i] = 0;
  return true;
};
s.structures = {pin: (i, on) => { s.pinned[i] = on ? 1 : 0; }};
const created = input.panels.map((p) => mpm.create(s, p));
const frames = [];
for (let f = 0; f < input.frames; f++) {
  if (input.mode === 'steps') {
    let remaining = 1 / 60, steps = 0;
    while (remaining > 1e-9 && steps < 96) { const h = Math.min(remaining, mpm.stableStep(s)); mpm.step(s, h); remaining -= h; steps++; }
    frames.push(steps);
  } else {
    frames.push(mpm.advance(s, 1 / 60).substeps);
  }
}
const n = s.count;
console.log(JSON.stringify({created, frames, x: Array.from(s.x.subarray(0, n)), y: Array.from(s.y.subarray(0, n)),
  pinned: Array.from(s.pinned.subarray(0, n)), f00: Array.from(mpm.f00.subarray(0, n)), stats: mpm.stats(s)}));
mpm.destroy();
