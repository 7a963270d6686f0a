// Reads a scene written by tests/test_node_addon.py (JSON of the Python
// mirror's arrays) into the reference's ClothScene shape (scene.ts:5-14):
// typed arrays, hinges as objects, pins as a Map, contact or null.
"use strict";
const fs = require("fs");

function readScene(file) {
  const j = JSON.parse(fs.readFileSync(file, "utf8"));
  const m = j.mesh;
  const hinges = [];
  for (let k = 0; k < m.hinges.length; k++) {
    const h = m.hinges[k];
    hinges.push({ v0: h[0], v1: h[1], v2: h[2], v3: h[3], restAngle: h[4], edgeLen: h[5], areaSum: h[6] });
  }
  const mesh = {
    count: m.count,
    triCount: m.triCount,
    positions: Float32Array.from(m.positions),
    restPositions: Float32Array.from(m.restPositions),
    uv: Float32Array.from(m.uv),
    indices: Uint32Array.from(m.indices),
    invDm: Float32Array.from(m.invDm),
    areas: Float32Array.from(m.areas),
    masses: Float32Array.from(m.masses),
    hinges: hinges,
  };
  const pinned = new Map();
  for (const p of j.pinned) pinned.set(p[0], [p[1], p[2], p[3]]);
  let contact = null;
  if (j.contact) {
    contact = {
      params: j.contact.params,
      floorY: j.contact.floorY,
      staticPos: j.contact.staticPos ? Float32Array.from(j.contact.staticPos) : null,
      staticIdx: j.contact.staticIdx ? Uint32Array.from(j.contact.staticIdx) : null,
    };
  }
  return {
    mesh: mesh,
    material: j.material,
    gravity: j.gravity,
    pinned: pinned,
    positions: Float64Array.from(j.positions),
    velocities: Float64Array.from(j.velocities),
    contact: contact,
    meta: j.meta || {},
  };
}

module.exports = { readScene };
