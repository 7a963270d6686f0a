// Node-API addon on the CPU (no device): argument checks, the ClothSolver
// contract before initialize, and the host-only entry points against the
// values the Python binding computed.  Usage: node addon_cpu.js scene.json
// expected.json ; exits 0 when every check holds.
"use strict";
const assert = require("assert");
const fs = require("fs");
const path = require("path");
const { readScene } = require("./scene_io");
const { HipSolver, packScene, addon } = require(path.join(__dirname, "..", "..", "closim_amd", "node", "hip-solver.js"));

const scene = readScene(process.argv[2]);
const expected = JSON.parse(fs.readFileSync(process.argv[3], "utf8"));

assert.strictEqual(addon.abiVersion(), 2);

// the contract before initialize (cpu-solver.ts:15-43)
const s = new HipSolver();
for (const f of [() => s.step(1 / 60), () => s.getPositions(), () => s.getVelocities(),
                 () => s.setMaterial(scene.material), () => s.pinVertex(0)])
  assert.throws(f, /HipSolver not initialized/);
s.unpinVertex(0);  // optional chaining in the reference: no throw
assert.strictEqual(s.lastStats, null);

// malformed scenes are TypeErrors raised before the library is reached
function withField(k, v) {
  const d = packScene(scene);
  d[k] = v;
  return d;
}
const bad = [
  [withField("positions", Array.from(scene.positions)), /positions: expected a Float64Array/],
  [withField("positions", scene.positions.subarray(3)), /positions: length/],
  [withField("indices", new Int32Array(scene.mesh.indices)), /indices: expected a Uint32Array/],
  [withField("masses", new Float32Array(1)), /masses: length/],
  [withField("material", { thickness: 1 }), /arealDensityKgM2: missing/],
  [withField("nVerts", -1), /nVerts: expected a non-negative integer/],
  [withField("pinPos", new Float64Array(1)), /pinPos: length/],
];
for (const [d, re] of bad) assert.throws(() => addon.create(d, {}), (e) => e instanceof TypeError && re.test(e.message));
assert.throws(() => addon.create(packScene(scene), { precision: "fp16" }), /precision: expected/);
assert.throws(() => addon.create(packScene(scene), { preconditioner: "ilu" }), /preconditioner: expected/);
assert.throws(() => addon.step({}, 1 / 60, scene.positions, scene.velocities), /not initialized or already destroyed/);

// a well-formed scene reaches the library, which reports the missing device
assert.throws(() => new HipSolver().initialize(scene), (e) => e instanceof Error && !(e instanceof TypeError));

// host-only entry points: buildSchwarzDomains and the HVP chunk layout
const part = addon.schwarzPartition(packScene(scene), 16, 8, 32);
assert.strictEqual(part.count, expected.schwarz.count);
assert.deepStrictEqual(Array.from(part.domainOf), expected.schwarz.domainOf);
assert.deepStrictEqual(Array.from(part.localOf), expected.schwarz.localOf);
const tl = addon.hvpTiles(packScene(scene));
assert.strictEqual(tl.tiles, expected.tiles.tiles);
for (const k of ["hdr", "loc", "lidx", "off", "ent"]) assert.deepStrictEqual(Array.from(tl[k]), expected.tiles[k]);
console.log("addon cpu checks OK");
