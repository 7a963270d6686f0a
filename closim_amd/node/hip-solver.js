// hip-solver.js — HipSolver: the ClothSolver backend contract
// (src/backend/solver.ts:5-13) on the GPU, a drop-in for CpuSolver
// (src/backend/cpu-solver.ts:7-49), over the Node-API addon
// hip_solver_addon.node (the C ABI of include/cloth_hip.h).  Plain
// CommonJS that Node 12 parses (no `??` / `?.`).
//
// Semantics kept from CpuSolver: initialize() keeps the caller's scene by
// reference and step() writes the new state into scene.positions /
// scene.velocities in place, so getPositions() returns the live array;
// setMaterial() copies; pinVertex() records the pin in scene.pinned (the
// current position when none is given) and enforces it at once; unpinVertex()
// on an uninitialized solver does nothing; lastStats is the StepStats record.
"use strict";

const path = require("path");

function loadAddon() {
  const candidates = [
    process.env.CLOSIM_HIP_ADDON,
    path.join(__dirname, "..", "_build", "hip_solver_addon.node"),
  ].filter(Boolean);
  let last = null;
  for (const p of candidates) {
    try {
      return require(p);
    } catch (e) {
      last = e;
    }
  }
  throw new Error("HIP cloth addon not built (" + candidates.join(", ") + "): " + (last ? last.message : ""));
}

const addon = loadAddon();

function f64(a) {
  return a instanceof Float64Array ? a : Float64Array.from(a);
}

// The scene in the flat form the addon reads: the ClothMeshData arrays
// (mesh.ts:3-14), the hinges as parallel arrays, the pins in Map insertion
// order and the contact configuration of scene.contact
// (contact-assembly.ts:46-73).
function packScene(scene) {
  const mesh = scene.mesh;
  const h = mesh.hinges.length;
  const hingeVerts = new Uint32Array(4 * h);
  const hingeRestAngle = new Float64Array(h);
  const hingeEdgeLen = new Float64Array(h);
  const hingeAreaSum = new Float64Array(h);
  for (let k = 0; k < h; k++) {
    const g = mesh.hinges[k];
    hingeVerts[4 * k] = g.v0;
    hingeVerts[4 * k + 1] = g.v1;
    hingeVerts[4 * k + 2] = g.v2;
    hingeVerts[4 * k + 3] = g.v3;
    hingeRestAngle[k] = g.restAngle;
    hingeEdgeLen[k] = g.edgeLen;
    hingeAreaSum[k] = g.areaSum;
  }
  const ids = [];
  const pos = [];
  for (const [id, p] of scene.pinned) {
    ids.push(id);
    pos.push(p[0], p[1], p[2]);
  }
  let contact = null;
  if (scene.contact) {
    const c = scene.contact;
    contact = {
      params: c.params,
      floorY: c.floorY === undefined ? null : c.floorY,
      staticPos: c.staticPos || null,
      staticIdx: c.staticIdx || null,
    };
  }
  return {
    nVerts: mesh.count,
    nTris: mesh.triCount,
    nHinges: h,
    positions: scene.positions,
    velocities: scene.velocities,
    indices: mesh.indices,
    invDm: mesh.invDm,
    areas: mesh.areas,
    masses: mesh.masses,
    restPositions: mesh.restPositions || null,
    hingeVerts: hingeVerts,
    hingeRestAngle: hingeRestAngle,
    hingeEdgeLen: hingeEdgeLen,
    hingeAreaSum: hingeAreaSum,
    material: scene.material,
    gravity: scene.gravity,
    pinIds: Uint32Array.from(ids),
    pinPos: Float64Array.from(pos),
    contact: contact,
  };
}

class HipSolver {
  // options: { precision: "fp64" | "mixed", preconditioner: "jacobi" |
  // "block-jacobi" | "schwarz", schwarz: [target, min, max], hvp: "fd" |
  // "analytic", newtonIters, pcgMaxIters, lineSearchIters, gradTol, flags,
  // device }; the defaults are CpuSolver's (FP64, FD HVP, Jacobi, 10 Newton,
  // 60 PCG, 10 halvings, gradTol 1e-5).
  constructor(options) {
    this.options = Object.assign({}, options || {});
    this.scene = null;
    this.handle = null;
    this.lastStats = null;
  }

  need() {
    if (!this.scene || !this.handle) throw new Error("HipSolver not initialized");
  }

  initialize(scene) {
    this.dispose();
    if (!(scene.positions instanceof Float64Array) || !(scene.velocities instanceof Float64Array))
      throw new TypeError("scene.positions and scene.velocities must be Float64Arrays");
    this.handle = addon.create(packScene(scene), this.options);
    this.scene = scene;
    this.lastStats = null;
    this.staticSeen = this.staticState();
  }

  staticState() {
    const c = this.scene.contact;
    if (!c || !c.staticPos) return null;
    return { ref: c.staticPos, copy: Float32Array.from(c.staticPos) };
  }

  // the caller moves a collider with scene.contact.setStaticMesh(pos, idx)
  // (avatar-collider.ts:205-230); the device copy follows before each step
  syncStatic() {
    const c = this.scene.contact;
    if (!c || !c.staticPos) return;
    const seen = this.staticSeen;
    if (seen && seen.ref === c.staticPos && seen.copy.length === c.staticPos.length &&
        seen.copy.every((v, i) => v === c.staticPos[i]))
      return;
    addon.setStaticMesh(this.handle, c.staticPos, c.staticIdx);
    this.staticSeen = this.staticState();
  }

  step(dt) {
    this.need();
    this.syncStatic();
    addon.step(this.handle, dt, this.scene.positions, this.scene.velocities);
    this.lastStats = addon.lastStats(this.handle);
  }

  setMaterial(material) {
    this.need();
    addon.setMaterial(this.handle, material);
    this.scene.material = Object.assign({}, material);
  }

  pinVertex(vertexId, position) {
    this.need();
    const x = this.scene.positions;
    const p = position !== undefined && position !== null
      ? [position[0], position[1], position[2]]
      : [x[vertexId * 3], x[vertexId * 3 + 1], x[vertexId * 3 + 2]];
    addon.pinVertex(this.handle, vertexId, p, x);
    this.scene.pinned.set(vertexId, p);
  }

  unpinVertex(vertexId) {
    if (!this.scene || !this.handle) return;
    addon.unpinVertex(this.handle, vertexId);
    this.scene.pinned.delete(vertexId);
  }

  getPositions() {
    this.need();
    return this.scene.positions;
  }

  getVelocities() {
    this.need();
    return this.scene.velocities;
  }

  // checkpoint restore (no CpuSolver counterpart): device state and the live arrays
  setState(positions, velocities) {
    this.need();
    const x = f64(positions), v = f64(velocities);
    addon.setState(this.handle, x, v);
    this.scene.positions.set(x);
    this.scene.velocities.set(v);
  }

  dispose() {
    if (this.handle) addon.destroy(this.handle);
    this.handle = null;
  }
}

module.exports = { HipSolver, packScene, addon };
