// The reference's callers of a ClothSolver, driven over the JS HipSolver:
// (1) a session that steps on a fixed dt from real frame times (the frame
// time clamped to 0.25 s, at most 8 steps per frame, the backlog dropped
// after a stall), publishes the live positions after every step and keeps
// the last finite state (sim-session.ts:116-152); (2) a fitting run:
// relaxation then simulation steps, each checked for non-finite positions or
// velocities and for a displacement above a bound, a thrown step reported
// as a failure (garment/assembly.ts:554-640).  Writes what it saw to JSON.
// Usage: node callers_gpu.js scene.json options.json out.json
"use strict";
const fs = require("fs");
const path = require("path");
const { readScene } = require("./scene_io");
const { HipSolver } = require(path.join(__dirname, "..", "..", "closim_amd", "node", "hip-solver.js"));

const [sceneFile, optFile, outFile] = process.argv.slice(2);
const opts = JSON.parse(fs.readFileSync(optFile, "utf8"));
const DT = 1 / 60;

function finite(a) {
  for (let i = 0; i < a.length; i++) if (!Number.isFinite(a[i])) return false;
  return true;
}

// (1) frame-driven session
function session(scene) {
  const solver = new HipSolver(opts);
  solver.initialize(scene);
  const published = [];
  let acc = 0, steps = 0, stable = scene.positions.slice();
  const frame = (real) => {
    acc += Math.max(0, Math.min(real, 0.25));
    let k = 0;
    while (acc >= DT && k < 8) {
      solver.step(DT);
      steps++;
      const x = solver.getPositions();
      if (x !== scene.positions) throw new Error("getPositions() is not the live scene array");
      if (finite(x)) stable = x.slice();
      published.push(x[1]);
      acc -= DT;
      k++;
    }
    if (k >= 8) acc = 0;
    return k;
  };
  // frames: regular, a stall (clamped, capped at 8 steps), sub-step frames
  const perFrame = [1 / 60, 1 / 30, 2.0, 0.005, 0.005, 0.005, 0.005, 1 / 60].map(frame);
  const stats = solver.lastStats;
  solver.dispose();
  return { perFrame, steps, published, stableY: stable[1], stats };
}

// (2) fitting run
function fitting(scene, relax, sim, maxStep) {
  const solver = new HipSolver(opts);
  solver.initialize(scene);
  let taken = 0, failure = null;
  for (let s = 0; s < relax + sim; s++) {
    const before = Float64Array.from(solver.getPositions());
    try {
      solver.step(DT);
    } catch (e) {
      failure = { code: "nan-inf-state", message: "solver threw at step " + s + ": " + e.message };
      break;
    }
    taken++;
    const now = solver.getPositions();
    if (!finite(now) || !finite(solver.getVelocities())) {
      failure = { code: "nan-inf-state", message: "non-finite state at step " + s };
      break;
    }
    let m = 0;
    for (let i = 0; i < now.length; i += 3)
      m = Math.max(m, Math.hypot(now[i] - before[i], now[i + 1] - before[i + 1], now[i + 2] - before[i + 2]));
    if (m > maxStep) {
      failure = { code: "exploded-initialization", message: "unstable displacement at step " + s };
      break;
    }
  }
  const out = { taken, failure, positions: Array.from(scene.positions) };
  solver.dispose();
  return out;
}

const a = session(readScene(sceneFile));
const b = fitting(readScene(sceneFile), 5, 10, 1.0);
// a blown-up state: the step throws the reference's error and the run reports it
const bad = readScene(sceneFile);
bad.contact = null;  // the non-finite state on the contact-free path
bad.positions[4] = NaN;
const c = fitting(bad, 1, 1, 1.0);
fs.writeFileSync(outFile, JSON.stringify({ session: a, fitting: b, blowup: c }));
console.log("callers run OK");
