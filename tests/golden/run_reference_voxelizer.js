// Fixture generator (runs in the development container only, never on the
// GPU test machines): evaluates the reference's voxelizar() from
// /root/reference/src/geom/voxel.js as an ES module and runs it on the meshes
// listed in a manifest written by make_voxel_golden.py.  Prints one JSON
// object per case with hashes of the outputs; nothing of the reference's text
// is stored.
//
// usage: node --experimental-vm-modules run_reference_voxelizer.js manifest.json
'use strict';
const fs = require('fs');
const vm = require('vm');
const crypto = require('crypto');
const { performance } = require('perf_hooks');

const REF = '/root/reference/src/geom/voxel.js';

async function loadVoxel() {
  const text = fs.readFileSync(REF, 'utf8');
  const ctx = vm.createContext({
    performance, Math, Number, Infinity, NaN, Error,
    Float32Array, Float64Array, Uint8Array, Uint32Array, Int32Array, Array, Object,
  });
  const mod = new vm.SourceTextModule(text, { context: ctx, identifier: 'voxel.js' });
  await mod.link(() => { throw new Error('voxel.js is expected to have no imports'); });
  await mod.evaluate();
  return mod.namespace;
}

function sha(buf) {
  return crypto.createHash('sha256').update(Buffer.from(buf.buffer, buf.byteOffset, buf.byteLength)).digest('hex');
}

function rle(u8) {
  const out = [];
  let v = u8[0], n = 0;
  for (let i = 0; i < u8.length; i++) {
    if (u8[i] === v) { n++; continue; }
    out.push([v, n]); v = u8[i]; n = 1;
  }
  out.push([v, n]);
  return out;
}

(async () => {
  const manifest = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const { voxelizar } = await loadVoxel();
  for (const c of manifest.cases) {
    const pb = fs.readFileSync(c.pos);
    const ib = fs.readFileSync(c.idx);
    const pos = new Float32Array(pb.buffer, pb.byteOffset, pb.byteLength / 4);
    const idx = new Uint32Array(ib.buffer, ib.byteOffset, ib.byteLength / 4);
    const r = voxelizar(pos, idx, c.grid, { banda: c.banda, selarFrestas: c.selar });
    const s = r.stats;
    const out = {
      name: c.name,
      tipo_sha256: sha(r.tipo),
      sdf_sha256: sha(r.sdf),
      nCasca: s.nCasca, nDentro: s.nDentro, nSolido: s.nSolido,
      trianguloFora: s.trianguloFora, ocoSuspeito: s.ocoSuspeito,
      ms_reference: s.msTotal,
    };
    if (c.with_rle) out.tipo_rle = rle(r.tipo);
    process.stdout.write(JSON.stringify(out) + '\n');
  }
})().catch((e) => { console.error(e); process.exit(1); });
