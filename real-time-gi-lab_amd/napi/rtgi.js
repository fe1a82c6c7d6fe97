'use strict';
// JS face of the Node-API addon: the reference host's pass interface over librtgi.
//   createVoxWorld(scene, params, frame, firstDynamicTri)
//                                  <- createVoxWorld (src/core/voxel.ts:746): update(eye), encode(), readCounters(),
//                                     (scene.surface: UVs, base-colour texels and material words of a textured,
//                                     alpha-tested scene),
//                                     usedBytes(c), destroy(); with firstDynamicTri below the triangle count, the
//                                     triangles from it on go to the dynamic coverage layer (dynBricks bricks;
//                                     src/core/dynlayer.ts) and dyn.readCounters() reads its counters
//   createRc(voxWorld, params)     <- createRc (src/core/rc.ts:1019): encode(), readSh(), readCascade(c), destroy(),
//                                     gather(pos, nrm, occlusion) (gatherWorld), readProbes(c), readBounceVolume()
//   voxWorld.setDynamicSpec(scene, spec), voxWorld.pose(dynamicPose)
//                                  <- posedVertices + the positions / normals writeBuffer of the runner's pose(t)
//                                     (src/core/scenes.ts:80-102, src/core/rcrun.ts:206-217): the rest mesh stays on
//                                     the device and each update poses the objects there; cos / sin of rotY are
//                                     taken here with Math.cos / Math.sin, as posedVertices takes them
//   dynamicSpec(scene)             <- the generated scene's DynamicSpec: objects and at(frame)
// One library world backs both: the voxel world owns it, and createRc re-creates it with the cascade parameters
// (its scene upload is repeated). encode() on either runs the whole update (rebuild, trace, merge, SH): the reference
// records both passes into one command buffer per frame (src/core/rcrun.ts:283-298).
const path = require('path');
const addon = require(path.join(__dirname, 'rtgi_napi.node'));

const DEFAULT_VOX = { v0: 0.2, levels: 5, G: 32, maxBricks: 4608, maxVoxels: 524288, dynBricks: 512 };
const DEFAULT_PARAMS = { cascades: 5, probes0: 64, s0: 0.2, kInterval: 1, levelMap: 1, voxScale: 0.25, voxOffset: 1,
  voxSkip: 0.5, resCap: 8, recordHits: 0, relocate: 0, lift: 0.25, sunTrace: 0, bounce: 0, objects: 0, fp16: 0,
  bf16: 0, history: 1 };
const SUN_E = [9, 8.5, 7.8];
const SKY = [0.45, 0.6, 0.9];

function checked(defaults, o, what) {
  for (const k of Object.keys(o)) if (!(k in defaults)) throw new Error(`unknown ${what} parameter ${k}`);
  return Object.assign({}, defaults, o);
}
function voxParams(o) { return checked(DEFAULT_VOX, o, 'voxel'); }
function rcParams(o) { return checked(DEFAULT_PARAMS, o, 'cascade'); }

function createVoxWorld(scene, params, frame, firstDynamicTri) {
  const ntris = scene.tri.length / 4;
  const first = firstDynamicTri === undefined ? ntris : firstDynamicTri;
  const vw = {
    params: voxParams(params || {}),
    rcParams: { cascades: 1, probes0: 1, s0: 1 },
    frame: Object.assign({ sunE: SUN_E, sky: SKY, sunVis: 1 }, frame || {}, { eye: (frame && frame.eye) || [0, 0, 0] }),
    handle: null,
  };
  vw.rebuild = () => {
    if (vw.handle) addon.destroy(vw.handle);
    vw.handle = addon.createWorld(vw.params, vw.rcParams);
    addon.uploadScene(vw.handle, scene.pos, scene.nrm, scene.tri, scene.materials);
    // scene.surface: {uvs, materials: [{baseColorLayer, flags, alphaCutoff}], texels, size, layers}, what surfaceAt
    // samples (uploadScene's uvs, baseColor array and material words, src/core/scene_gpu.ts:56-106)
    const sf = scene.surface;
    if (sf) addon.setSurface(vw.handle, sf.uvs, sf.materials, sf.texels, sf.size, sf.layers);
    if (first < ntris) addon.setDynamic(vw.handle, first, vw.params.dynBricks);
    addon.setFrame(vw.handle, vw.frame);
  };
  vw.dyn = first < ntris ? { readCounters: () => Promise.resolve().then(() => addon.readDynCounters(vw.handle)) } : null;
  vw.update = (eye) => { vw.frame = Object.assign({}, vw.frame, { eye }); addon.setFrame(vw.handle, vw.frame); };
  vw.encode = () => addon.encode(vw.handle);
  vw.readCounters = () => Promise.resolve().then(() => addon.readCounters(vw.handle));
  vw.usedBytes = (c) => {
    const { levels: NL, G, maxVoxels } = vw.params;
    return { persistent: (64 + NL * G * G * G) * 4 + c.bricks * 72 + Math.min(c.voxels, maxVoxels) * 12,
      buildTemp: c.bricks * 512 * 4 };
  };
  vw.destroy = () => { if (vw.handle) addon.destroy(vw.handle); vw.handle = null; };
  // Moving objects survive a rebuild (createRc re-creates the world): the spec and the last pose are replayed.
  let objects = null, lastPose = null;
  // sc: the scene whose positions / normals are the rest mesh (SceneData), or a generated scene (its restPos /
  // restNrm are the objects' span at rest).
  vw.setDynamicSpec = (sc, spec) => {
    const o = spec.objects, f0 = o.length ? o[0].firstVertex : 0;
    const end = o.length ? o[o.length - 1].firstVertex + o[o.length - 1].vertexCount : 0;
    const rest = sc.restPos ? [sc.restPos, sc.restNrm] : [sc.pos.slice(f0 * 3, end * 3), sc.nrm.slice(f0 * 3, end * 3)];
    objects = { list: o, rest };
    addon.setObjects(vw.handle, o, rest[0], rest[1]);
  };
  vw.pose = (dp) => {
    lastPose = dp.poses.map((p) => ({ c: Math.cos(p.rotY), s: Math.sin(p.rotY), offset: p.offset }));
    addon.setPose(vw.handle, lastPose);
    return lastPose;
  };
  const rebuild = vw.rebuild;
  vw.rebuild = () => {
    rebuild();
    if (objects) addon.setObjects(vw.handle, objects.list, objects.rest[0], objects.rest[1]);
    if (objects && lastPose) addon.setPose(vw.handle, lastPose);
  };
  vw.rebuild();
  return vw;
}

// DynamicSpec of a generated scene (src/core/scenes.ts:48-53): its objects, and at(frame) with rotY = rot0 + w f and
// offset = amp sin(wt f + ph); the scene's vertices are its rest mesh posed at the frame it was generated for, so a
// world that poses them takes the rest mesh (restPos / restNrm) as the objects' span.
function dynamicSpec(scene) {
  return {
    objects: scene.objects,
    at: (f) => ({ poses: scene.motion.map((m) => ({ rotY: m.rot0 + m.w * f,
      offset: m.amp.map((a) => a * Math.sin(m.wt * f + m.ph)) })) }),
  };
}

function createRc(voxWorld, params) {
  voxWorld.rcParams = rcParams(params || {});
  voxWorld.rebuild();
  return {
    params: voxWorld.rcParams,
    encode: () => addon.encode(voxWorld.handle),
    readSh: () => Promise.resolve().then(() => addon.readSh(voxWorld.handle)),
    readCascade: (c) => Promise.resolve().then(() => addon.readCascade(voxWorld.handle, c)),
    gather: (pos, nrm, occlusion) => Promise.resolve().then(() => addon.gather(voxWorld.handle, pos, nrm, !!occlusion)),
    readProbes: (c) => Promise.resolve().then(() => addon.readProbes(voxWorld.handle, c)),
    readBounceVolume: () => Promise.resolve().then(() => addon.readBounceVolume(voxWorld.handle)),
    destroy: () => {},
  };
}

module.exports = { addon, DEFAULT_VOX, DEFAULT_PARAMS, SUN_E, SKY, voxParams, rcParams, createVoxWorld, createRc,
  dynamicSpec, generateScene: addon.generateScene, deviceCount: addon.deviceCount, version: addon.version };
