// Fixture generator (runs in the development container only, never on the
// GPU test machines): evaluates the reference's ler() (src/geom/parse.js) and
// preparar() (src/geom/prepare.js) as ES modules from /root/reference and runs
// them on the meshes listed in a manifest written by make_prepare_golden.py.
// Prints one JSON object per case: hashes of the read and of the placed
// positions, and the placement's scalars; nothing of the reference's text is
// stored.  The image's Node 12 evaluates `??` and `?.` behind V8 flags.
//
// usage: node --experimental-vm-modules --harmony-nullish
//             --harmony-optional-chaining run_reference_prepare.js manifest.json
'use strict';
const fs = require('fs');
const vm = require('vm');
const crypto = require('crypto');

const GEOM = '/root/reference/src/geom/';

// the browser's base64 decoder, which parse.js calls for glTF data URIs and
// which Node 12 does not define
if (typeof globalThis.atob !== 'function') {
  globalThis.atob = (s) => Buffer.from(s, 'base64').toString('binary');
}

async function loadModules() {
  const mods = {};
  const make = (name) => {
    if (!mods[name]) {
      mods[name] = new vm.SourceTextModule(fs.readFileSync(GEOM + name, 'utf8'),
                                           { identifier: name });
    }
    return mods[name];
  };
  const prep = make('prepare.js');
  await prep.link((spec) => {
    if (spec !== './parse.js') throw new Error('unexpected import ' + spec);
    return make('parse.js');
  });
  await prep.evaluate();
  return { parse: mods['parse.js'].namespace, prepare: prep.namespace };
}

function sha(a) {
  return crypto.createHash('sha256')
    .update(Buffer.from(a.buffer, a.byteOffset, a.byteLength)).digest('hex');
}

(async () => {
  const manifest = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const { parse, prepare } = await loadModules();
  for (const c of manifest.cases) {
    const b = fs.readFileSync(c.file);
    const buf = b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength);
    const m = parse.ler(buf, c.name);
    const p = prepare.preparar(m, c.grade, c.opcoes);
    console.log(JSON.stringify({
      case: c.id,
      formato: m.formato, nTriangulos: m.nTriangulos,
      read_positions_sha256: sha(m.positions), read_indices_sha256: sha(m.indices),
      positions_sha256: sha(p.positions), n_positions: p.positions.length,
      ordem: p.eixos.ordem, inverter: p.sentido.aplicado,
      alturaFrente: p.sentido.alturaFrente, alturaTras: p.sentido.alturaTras,
      unidade: p.unidade.nome, fator: p.unidade.fator, escala: p.escala,
      comprimentoM: p.comprimentoM, bloqueioCaixa: p.bloqueioCaixa,
      min: p.extentos.min, max: p.extentos.max, n_avisos: p.avisos.length,
    }));
  }
})().catch((e) => { console.error(e); process.exit(1); });
