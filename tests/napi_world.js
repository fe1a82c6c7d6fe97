'use strict';
// Drives one GI update through the Node-API addon the way the reference's host drives its passes
// (createVoxWorld -> createRc -> update -> encode -> readCounters / read SH) and prints the results as JSON.
const path = require('path');
const r = require(path.join(__dirname, '..', 'real-time-gi-lab_amd', 'napi', 'rtgi.js'));
const args = process.argv.slice(2);
const [kind, seed, v0, levels, G, cascades, probes0, s0] = args.slice(0, 8).map((x, i) => (i ? Number(x) : x));
// 'full': the reference's default extras (voxel relocation, per-voxel sun visibility, multi-bounce), two updates, and
// the gather / probe / bounce-volume readbacks.
// 'dynamic': the scene's moving part (triangles from firstDynamicTri on) goes to the dynamic coverage layer.
// 'motion': the scene's objects posed on the device at frames 3 and 11 (the runner's pose(t)); the poses sent are
// printed so that the other side can send the same.
// 'textured': a surface on the scene: UVs by planar projection along each normal's dominant axis (4 m tiles), one
// 64 x 64 layer (rgb = 4 x, 4 y, 128; alpha 0 on its left half, 255 on its right), every material masked, the odd
// ones toon.
const full = args[8] === 'full', dynamic = args[8] === 'dynamic', motion = args[8] === 'motion';
const tex = args[8] === 'textured';
const scene = r.generateScene(kind, seed, 0);
if (tex) {
  const nv = scene.pos.length / 3, uvs = new Float32Array(nv * 2), texels = new Uint8Array(64 * 64 * 4);
  for (let i = 0; i < nv; i++) {
    const n = [0, 1, 2].map((a) => Math.abs(scene.nrm[i * 3 + a]));
    const ax = n[0] >= n[1] && n[0] >= n[2] ? 0 : (n[1] >= n[2] ? 1 : 2);
    const o = [0, 1, 2].filter((a) => a !== ax);
    uvs[i * 2] = scene.pos[i * 3 + o[0]] / 4;
    uvs[i * 2 + 1] = scene.pos[i * 3 + o[1]] / 4;
  }
  for (let y = 0; y < 64; y++) for (let x = 0; x < 64; x++) texels.set([x * 4, y * 4, 128, x < 32 ? 0 : 255], (y * 64 + x) * 4);
  const materials = [];
  for (let m = 0; m < scene.materials.length / 4; m++) {
    materials.push({ baseColorLayer: 1, flags: 1 | (m % 2 ? 4 : 0), alphaCutoff: 0.5 });
  }
  scene.surface = { uvs, materials, texels, size: 64, layers: 1 };
}
const vw = r.createVoxWorld(scene, { v0, levels, G, maxBricks: levels * G * G * G, maxVoxels: 1 << 20 },
  { eye: scene.eye, sunDir: scene.sunDir }, dynamic ? scene.firstDynamicTri : undefined);
const rc = r.createRc(vw, full ? { cascades, probes0, s0, relocate: 1, sunTrace: 1, bounce: 1 } : { cascades, probes0, s0 });
vw.update(scene.eye);
const sent = [];
if (motion) {
  const spec = r.dynamicSpec(scene);
  vw.setDynamicSpec(scene, spec);
  for (const f of [3, 11]) {
    sent.push(vw.pose(spec.at(f)));
    rc.encode();
  }
} else {
  rc.encode();
}
if (full) rc.encode();
const sum = (a) => { let s = 0; for (let i = 0; i < a.length; i++) s += a[i]; return s; };
// Gather points: an 8 x 8 grid 1 m below the eye, facing up.
const pos = new Float32Array(64 * 3), nrm = new Float32Array(64 * 3);
for (let i = 0; i < 64; i++) {
  pos.set([scene.eye[0] + (i % 8) - 3.5, scene.eye[1] - 1, scene.eye[2] + Math.floor(i / 8) - 3.5], i * 3);
  nrm.set([0, 1, 0], i * 3);
}
const extra = full ? [rc.gather(pos, nrm, true), rc.readProbes(0), rc.readBounceVolume()]
  : dynamic ? [vw.dyn.readCounters()] : [];
Promise.all([vw.readCounters(), rc.readSh(), rc.readCascade(cascades - 1), ...extra]).then(([c, sh, top, irr, pr, bv]) => {
  const out = { counters: c, shSum: sum(sh), sh0: Array.from(sh.slice(0, 27)), topLen: top.length };
  if (dynamic) out.dyn = irr;
  if (motion) out.poses = sent.map((ps) => ps.map((p) => [p.c, p.s, ...p.offset]));
  if (full) {
    out.irr = Array.from(irr);
    out.probeSum = sum(pr.pos) + sum(pr.nrm);
    out.parentsSum = sum(pr.parents);
    out.texelSum = sum(bv.texels);
    out.origin = Array.from(bv.origin);
  }
  process.stdout.write(JSON.stringify(out) + '\n');
  vw.destroy();
});
