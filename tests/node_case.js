// Drives SolverMI355X from Node.js the way the reference's test pages drive
// SolverWebGPU, for tests/test_node.py.
// usage: node tests/node_case.js <tipo.u32> nx ny nz <out prefix>
'use strict';
const fs = require('fs');
const path = require('path');
const { SolverMI355X } = require(path.join(__dirname, '..', 'cfd__amd', 'node', 'solver.js'));

(async () => {
  const [tipoFile, nx, ny, nz, out] = process.argv.slice(2);
  const grade = { nx: +nx, ny: +ny, nz: +nz };
  const b = fs.readFileSync(tipoFile);
  const tipo = new Uint32Array(b.buffer, b.byteOffset, b.byteLength / 4);
  const ctx = await SolverMI355X.criarDevice();
  const s = await SolverMI355X.criar(ctx, grade);
  s.definirTipos(tipo);
  const esponja = { inicio: grade.nx - 1 - 8, comprimento: 8 };
  s.configurar({ omegaPlus: 1.6, lesCs: 0.1, inletU: [0.05, 0, 0], beltU: [0.05, 0, 0], esponja });
  const off = await s.calibrarRepouso();
  s.rampa(40, (f) => s.configurar({ omegaPlus: 1.6, lesCs: 0.1, inletU: [0.05 * f, 0, 0],
    beltU: [0.05 * f, 0, 0], esponja }));
  s.inicializar();
  for (let k = 0; k < 6; k++) s.passo(10);
  const f = await s.medirForcas();
  const m = await s.lerMacros();
  fs.writeFileSync(out + '.macros', Buffer.from(m.buffer));
  fs.writeFileSync(out + '.json', JSON.stringify({ off, f, frente: s.frente, passos: s.passos,
    bytes: s.bytesTotais, massa: await s.massa() }));
  s.destruir();
  let threw = false;
  try { s.passo(1); } catch (e) { threw = /destroyed/.test(e.message); }
  if (!threw) throw new Error('use after destruir did not throw');
})().catch((e) => { console.error(e); process.exit(1); });
