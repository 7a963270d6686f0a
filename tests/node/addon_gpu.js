// Node-API addon on the GPU: the JS HipSolver steps a scene and writes the
// trajectory for tests/test_gpu_node_addon.py to compare.  Usage:
// node addon_gpu.js scene.json options.json steps pinStep pinVertex out.json
// (pinVertex < 0: no pin; the pin is taken at the current position, as
// CpuSolver.pinVertex does without a position, before step pinStep).
"use strict";
const fs = require("fs");
const path = require("path");
const { readScene } = require("./scene_io");
const { HipSolver } = require(path.join(__dirname, "..", "..", "closim_amd", "node", "hip-solver.js"));

const [sceneFile, optFile, stepsS, pinStepS, pinVertexS, outFile] = process.argv.slice(2);
const scene = readScene(sceneFile);
const opts = JSON.parse(fs.readFileSync(optFile, "utf8"));
const steps = parseInt(stepsS, 10), pinStep = parseInt(pinStepS, 10), pinV = parseInt(pinVertexS, 10);
const s = new HipSolver(opts);
s.initialize(scene);
const live = s.getPositions();
const stats = [];
for (let k = 0; k < steps; k++) {
  if (pinV >= 0 && k === pinStep) s.pinVertex(pinV);
  s.step(1 / 60);
  stats.push(s.lastStats);
  console.log("step " + k + " newton " + s.lastStats.newtonUsed + " pcg " + s.lastStats.pcgIters);
}
if (s.getPositions() !== live || live !== scene.positions) throw new Error("getPositions() is not the live array");
fs.writeFileSync(outFile, JSON.stringify({
  positions: Array.from(scene.positions),
  velocities: Array.from(scene.velocities),
  stats: stats,
  pinned: Array.from(scene.pinned.entries()),
}));
s.dispose();
console.log("addon gpu run OK");
