'use strict';
// Builds a frame packet with the addon's framePacket and reads it back with
// the browser worker's checks (src/three/gpu-worker.js:36-42): header length
// within the packet, integer count and renderCount with
// 0 <= renderCount <= count <= capacity, and exactly renderCount rows of 8
// float32 after the 4-byte-aligned header. Input JSON {header, rows}; prints
// {ok, state, data, base64}.
const path = require('path');
const fs = require('fs');
const native = require(path.join(__dirname, '..', '..', 'summer-particle-lab_amd', 'node', 'mpm_napi.node'));

function read(packet) {
  const length = new DataView(packet).getUint32(0, true), offset = 4 + Math.ceil(length / 4) * 4;
  if (length > packet.byteLength - 4) throw Error('header length past the end');
  const state = JSON.parse(new TextDecoder().decode(new Uint8Array(packet, 4, length)));
  const rendered = state.renderCount === undefined ? state.count : state.renderCount;
  if (!Number.isInteger(state.count) || !Number.isInteger(rendered) || rendered < 0 || rendered > state.count ||
      state.count < 0 || state.count > state.capacity || offset + rendered * 32 !== packet.byteLength)
    throw Error('invalid particle frame');
  return {state, data: new Float32Array(packet, offset, rendered * 8)};
}

const input = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const rows = Float32Array.from([].concat(...input.rows));
const packet = native.framePacket(JSON.stringify(input.header), rows);
let out;
try {
  const r = read(packet);
  out = {ok: true, state: r.state, data: Array.from(r.data)};
} catch (e) {
  out = {ok: false, error: e.message};
}
out.base64 = Buffer.from(packet).toString('base64');
console.log(JSON.stringify(out));
