'use strict';
// tracer_native.js — the Renderer / Denoiser host methods of tracer.js over the
// Node-API addon (pt_addon.node -> libpt.so).
//
// Drop-in use from the reference's host code:
//   const { NativeRenderer, NativeDenoiser } = require('./tracer_native.js');
//   const r = new NativeRenderer(0);
//   r.setScene(renderer.packSceneData(true), skybox, skybox.buildHDRCDF(), scene.bounces);
//   r.resize(w, h);  r.render(reps);  const rgba = r.readOutput();
// `sceneData` is exactly what Renderer.packSceneData returns (tracer.js:1819-1917);
// the skybox is an HDRTexture-like {data, width, height, exposure} whose data
// holds rgba16float bit patterns (a Uint16Array, or a Float16Array's buffer).
const fs = require('fs');
const path = require('path');

const addon = require(path.join(__dirname, 'build', 'pt_addon.node'));

// pt_set_option keys and values (pt_abi.h)
const OPT = {
  BACKEND: 1, COUNT_STATS: 2, WAVEFRONT_POOL: 3, SCHEDULE: 4, TRAVERSAL: 5, LDS_STACK: 6, WAVEFRONT_REFILL: 7,
  BLAS_BUILD: 8, TLAS_BUILD: 9,
};
const BACKEND = { MEGAKERNEL: 0, WAVEFRONT: 1 };
const BLAS_BUILD = { REFERENCE: 0, DEVICE: 1 };
const TLAS_BUILD = { REFERENCE: 0, DEVICE: 1 };

// override constants shader.wgsl:2-12 <- packSceneData flags (tracer.js:1901-1913)
const FLAG_BITS = {
  hasSpheres: 1, hasCubes: 2, hasPlanes: 4, hasCylinders: 8, hasTori: 16, hasMeshes: 32,
  hasListMeshes: 64, hasHeightMaps: 128, hasLights: 256, hasSkybox: 512, misSkybox: 1024,
};

function flagsWord(flags) {
  let w = 0;
  for (const k of Object.keys(FLAG_BITS)) if (flags[k]) w |= FLAG_BITS[k];
  return w;
}

function bytesOf(x) {
  if (!x) return null;
  if (x.data !== undefined && !(x instanceof ArrayBuffer) && !ArrayBuffer.isView(x)) x = x.data;
  return x;
}

// Renderer.updateUniforms (tracer.js:2207-2240): the 112-byte SceneParams.
function packParams(f) {
  const buf = new ArrayBuffer(112);
  const fl = new Float32Array(buf);
  const dv = new DataView(buf);
  fl.set(f.eye, 0); dv.setUint32(12, f.sampleNumber >>> 0, true);
  fl.set(f.ray00, 4); dv.setUint32(28, f.width, true);
  fl.set(f.ray10, 8); dv.setUint32(44, f.height, true);
  fl.set(f.ray01, 12); dv.setFloat32(60, f.exposure === undefined ? 1 : f.exposure, true);
  fl.set(f.ray11, 16); dv.setUint32(76, f.seed >>> 0, true);
  dv.setUint32(80, f.skyWidth, true);
  dv.setUint32(84, f.skyHeight, true);
  dv.setFloat32(88, f.skyTotalLum, true);
  dv.setFloat32(92, f.totalLightPower || 0, true);
  dv.setUint32(96, (f.section && f.section.x) || 0, true);
  dv.setUint32(100, (f.section && f.section.y) || 0, true);
  dv.setUint32(104, f.skyCdfIndex, true);
  return buf;
}

class NativeRenderer {
  constructor(device = 0) {
    this.ctx = addon.create(device);
    this.frame = 0;
    this.width = 0;
    this.height = 0;
    this.isgbuf = false;
    this.sky = { width: 1, height: 1, totalLum: 0, cdfIndex: 2 };
  }

  _desc(sceneData, sky, cdf, bounces) {
    const d = {
      materials: bytesOf(sceneData.mat), objects: bytesOf(sceneData.object), tlas: bytesOf(sceneData.tlas),
      meshes: bytesOf(sceneData.mesh), bvh: bytesOf(sceneData.bvh), triangles: bytesOf(sceneData.triangle),
      planes: bytesOf(sceneData.plane), lights: bytesOf(sceneData.light),
      flags: typeof sceneData.flags === 'number' ? sceneData.flags : flagsWord(sceneData.flags),
      bounceLimit: bounces === undefined ? 8 : bounces,
      sky: sky.data, skyWidth: sky.width, skyHeight: sky.height,
      skyExposure: sky.exposure === undefined ? 1 : sky.exposure,
      skyCDF: cdf ? cdf.skyCDF : null, skyTotalLum: cdf ? cdf.totalWeight : 0,
      // the texture array setScene builds (tracer.js:1919-1953): 512x512 rgba8 layers
      textures: sceneData.textures ? bytesOf(sceneData.textures) : null,
    };
    return d;
  }

  // Renderer.setScene (tracer.js:1955-2102).  cdf = skybox.buildHDRCDF(), or
  // null to build the CDF on the device.
  setScene(sceneData, sky, cdf, bounces, isgbuf = false) {
    this.isgbuf = isgbuf;
    addon.uploadScene(this.ctx, this._desc(sceneData, sky, cdf, bounces));
    // without a host CDF the device built one: its total weight goes into the uniforms
    const total = cdf ? cdf.totalWeight
      : addon.readSkyCdf(this.ctx, new Float32Array(sky.height + 1 + sky.height * (sky.width + 1)));
    this.sky = { width: sky.width, height: sky.height, totalLum: total, cdfIndex: sky.height + 1 };
    this.lastScene = [sceneData, sky, cdf, bounces];
    if (this.width) this.resize(this.width, this.height);
  }

  // Renderer.updateObjects (tracer.js:2132-2164): fast path, else full setScene.
  updateObjects(sceneData, bounces) {
    const [, sky, cdf, b] = this.lastScene;
    const rc = addon.updateObjects(this.ctx, this._desc(sceneData, sky, cdf, bounces === undefined ? b : bounces));
    if (rc !== 0) this.setScene(sceneData, sky, cdf, bounces === undefined ? b : bounces, this.isgbuf);
  }

  resize(width, height) {
    addon.resize(this.ctx, width, height, this.isgbuf);
    this.width = width;
    this.height = height;
    this.reset();
  }

  reset() { this.frame = 0; }

  // camera: {eye, ray00, ray10, ray01, ray11, exposure} as Camera.updateRays2 leaves them
  params(camera, seed, section) {
    return packParams(Object.assign({}, camera, {
      sampleNumber: this.frame, width: this.width, height: this.height, seed: seed >>> 0,
      skyWidth: this.sky.width, skyHeight: this.sky.height, skyTotalLum: this.sky.totalLum,
      skyCdfIndex: this.sky.cdfIndex, section: section || { x: 0, y: 0 },
    }));
  }

  // Renderer.renderSection: `passes` progressive samples over one section,
  // each with its own seed (seeds: Uint32Array, one per pass).
  renderSection(camera, section, seeds) {
    const p = this.params(camera, seeds[0], section);
    const st = addon.trace(this.ctx, p, section ? section.width : 0, section ? section.height : 0, seeds.length,
                           seeds);
    this.frame += seeds.length;
    return st;
  }

  // Renderer.render(tsize, reps): the whole frame in one dispatch, `reps`
  // samples (the reference's per-tile submit/await loop is not needed).
  render(camera, reps = 1, seedBase = 0) {
    const seeds = new Uint32Array(reps);
    for (let i = 0; i < reps; i++) seeds[i] = (Math.imul(this.frame + i, 0x85EBCA77) + seedBase) >>> 0;
    return this.renderSection(camera, null, seeds);
  }

  // The same, returning a Promise that resolves with the stats once the GPU is
  // done, like the reference's `await onSubmittedWorkDone()` (tracer.js:2280-2281,
  // 2294); the event loop keeps running meanwhile.  Other calls on this
  // renderer throw until it settles.
  renderSectionAsync(camera, section, seeds) {
    const p = this.params(camera, seeds[0], section);
    const pr = addon.traceAsync(this.ctx, p, section ? section.width : 0, section ? section.height : 0,
                                seeds.length, seeds);
    this.frame += seeds.length;
    return pr;
  }

  renderAsync(camera, reps = 1, seedBase = 0) {
    const seeds = new Uint32Array(reps);
    for (let i = 0; i < reps; i++) seeds[i] = (Math.imul(this.frame + i, 0x85EBCA77) + seedBase) >>> 0;
    return this.renderSectionAsync(camera, null, seeds);
  }

  gbuffer(camera) { addon.gbuffer(this.ctx, this.params(camera, 0)); }

  setOption(key, value) { addon.setOption(this.ctx, key, value); }

  readAccum() { return addon.readAccum(this.ctx, new Float32Array(this.width * this.height * 4)); }
  readOutput() { return addon.readOutput(this.ctx, new Uint8Array(this.width * this.height * 4)); }
  readGBuffer() { return addon.readGBuffer(this.ctx, new Float32Array(this.width * this.height * 12)); }
  readDenoised() { return addon.readDenoised(this.ctx, new Float32Array(this.width * this.height * 4)); }
}

// Denoiser (tracer.js:2418-2637) with the reference defaults.
class NativeDenoiser {
  constructor(renderer) {
    this.r = renderer;
    this.passes = 4;
    this.sigmaColor = 0.30;
    this.sigmaNormal = 128.0;
    this.sigmaDepth = 0.2;
    this.sigmaAlbedo = 0.07;
    this.sigmaMaterial = 0.07;
    this.glossyProtect = 0.75;
    this.exposure = 1.0;
    this.useMedian = true;
    this.useDemodulate = true;
  }

  uniforms() {
    const buf = new ArrayBuffer(48);
    const f32 = new Float32Array(buf);
    const u32 = new Uint32Array(buf);
    u32[0] = this.r.width; u32[1] = this.r.height; u32[2] = 1; u32[3] = this.passes;
    f32[4] = this.sigmaColor; f32[5] = this.sigmaNormal; f32[6] = this.sigmaDepth;
    f32[7] = this.sigmaAlbedo; f32[8] = this.sigmaMaterial; f32[9] = this.glossyProtect;
    f32[10] = this.exposure;
    u32[11] = (this.useMedian ? 1 : 0) | (this.useDemodulate ? 2 : 0);
    return buf;
  }

  // strideMode 0: steps 1,2,4..; 1: the reference's constant 1<<(passes-1).
  execute(strideMode = 0) {
    addon.denoise(this.r.ctx, this.uniforms(), strideMode);
    return this.r.readOutput();
  }

  // Promise of the rgba8 output
  executeAsync(strideMode = 0) {
    return addon.denoiseAsync(this.r.ctx, this.uniforms(), strideMode).then(() => this.r.readOutput());
  }
}

// A little-endian float32 .npy file (the package's trained weights) as a Float32Array.
function loadNpyF32(file) {
  const buf = fs.readFileSync(file);
  if (buf.toString('latin1', 1, 6) !== 'NUMPY') throw new Error(file + ': not a .npy file');
  const major = buf[6];
  const hlen = major >= 2 ? buf.readUInt32LE(8) : buf.readUInt16LE(8);
  const start = (major >= 2 ? 12 : 10) + hlen;
  const header = buf.toString('latin1', major >= 2 ? 12 : 10, start);
  if (!/'descr':\s*'<f4'/.test(header) || /'fortran_order':\s*True/.test(header))
    throw new Error(file + ': expected a C-order little-endian float32 array');
  const n = (buf.length - start) / 4;
  const out = new Float32Array(n);
  for (let i = 0; i < n; i++) out[i] = buf.readFloatLE(start + 4 * i);
  return out;
}

const PRETRAINED = path.join(__dirname, '..', 'weights', 'unet_small_scenes.npy');

// The learned alternative to NativeDenoiser for the same frame (pt_unet.h):
// the renderer needs the G-buffer (isgbuf) and a gbuffer() pass, as the
// a-trous filter does.
class NativeLearnedDenoiser {
  constructor(renderer, weights = PRETRAINED, seed = 0) {
    this.r = renderer;
    this.exposure = 1.0;
    addon.unetInit(this.r.ctx, seed);
    if (weights) addon.unetSetParams(this.r.ctx, weights instanceof Float32Array ? weights : loadNpyF32(weights));
  }

  execute() {
    addon.unetDenoise(this.r.ctx, this.exposure);
    return this.r.readOutput();
  }

  // Promise of the rgba8 output
  executeAsync() {
    return addon.unetDenoiseAsync(this.r.ctx, this.exposure).then(() => this.r.readOutput());
  }

  readResult() { return this.r.readDenoised(); }

  // fp8 inference (true: calibrated on the renderer's current frame) or bf16
  setPrecision(fp8) { addon.unetSetPrecision(this.r.ctx, fp8 ? 1 : 0); }
}

// Several NativeRenderers rendered one progressive frame at a time as one
// graph launch (pt_batch.h); each job keeps its own scene, frame and options.
class NativeBatch {
  constructor(device = 0) {
    this.h = addon.batchCreate(device);
    this.jobs = [];
  }

  add(renderer) {
    this.jobs.push(renderer);
    return addon.batchAdd(this.h, renderer.ctx);
  }

  // cameras[i], passes[i] for job i (0 skips it); seeds as NativeRenderer.render derives them
  _args(cameras, passes, seedBase) {
    const params = [], seeds = [];
    this.jobs.forEach((r, i) => {
      const s = new Uint32Array(passes[i]);
      for (let k = 0; k < passes[i]; k++) s[k] = (Math.imul(r.frame + k, 0x85EBCA77) + seedBase) >>> 0;
      params.push(r.params(cameras[i], passes[i] ? s[0] : 0));
      seeds.push(passes[i] ? s : null);
    });
    return [params, seeds];
  }

  frame(cameras, passes, seedBase = 0) {
    const [params, seeds] = this._args(cameras, passes, seedBase);
    const st = addon.batchFrame(this.h, params, passes, seeds);
    this.jobs.forEach((r, i) => { r.frame += passes[i]; });
    return st;
  }

  // Promise of the per-job stats; the batch and its renderers are busy until it settles
  frameAsync(cameras, passes, seedBase = 0) {
    const [params, seeds] = this._args(cameras, passes, seedBase);
    const pr = addon.batchFrameAsync(this.h, params, passes, seeds);
    this.jobs.forEach((r, i) => { r.frame += passes[i]; });
    return pr;
  }
}

module.exports = {
  NativeRenderer, NativeDenoiser, NativeLearnedDenoiser, NativeBatch, packParams, flagsWord, loadNpyF32,
  FLAG_BITS, OPT, BACKEND, BLAS_BUILD, TLAS_BUILD, PRETRAINED, addon,
};
