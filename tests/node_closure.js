// The learned closure driven from Node.js, for tests/test_node.py: the
// shipped model read from its JSON export, the wall distance computed on the
// GPU, then steps from a given state.
// usage: node tests/node_closure.js <tipo.u32> <pops.f32> <model.json> nx ny nz <out prefix>
'use strict';
const fs = require('fs');
const path = require('path');
const { SolverMI355X, carregarFechamento } =
  require(path.join(__dirname, '..', 'cfd__amd', 'node', 'solver.js'));

(async () => {
  const [tipoFile, popFile, modelo, nx, ny, nz, out] = process.argv.slice(2);
  const grade = { nx: +nx, ny: +ny, nz: +nz };
  const ler = (f, T) => {
    const b = fs.readFileSync(f);
    return new T(b.buffer, b.byteOffset, b.byteLength / T.BYTES_PER_ELEMENT);
  };
  const s = await SolverMI355X.criar(await SolverMI355X.criarDevice(), grade, 'fp16sr');
  s.definirTipos(ler(tipoFile, Uint32Array));
  s.configurar({ omegaPlus: 1.9, lesCs: 0.1, inletU: [0.08, 0, 0], beltU: [0, 0, 0] });
  s.frente = 'A';
  s.definirPopulacoes(ler(popFile, Float32Array));
  s.calcularDistanciaParede();
  s.definirFechamento(carregarFechamento(modelo));
  s.passo(30);
  const p = await s.lerPopulacoes();
  fs.writeFileSync(out + '.pops', Buffer.from(p.buffer));
  fs.writeFileSync(out + '.json', JSON.stringify({ fallbacks: s.fallbacksFechamento(true),
    depois: s.fallbacksFechamento(true) }));
  s.definirFechamento(null);
  s.destruir();
})().catch((e) => { console.error(e); process.exit(1); });
