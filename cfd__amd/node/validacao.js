/* cfd__amd/node/validacao.js - the reference's validation page
 * (tests/validacao.html:169-272) as a headless Node script over SolverMI355X.
 *
 * For each case: the analytic body (tunel.js:144-168) in the validation
 * tunnel with a mirror floor (validacao.html:198-202), omega fixed at 1.90
 * and u_lb derived from Re and D (validacao.js:126-129), an optional inlet yaw
 * to seed shedding, the transient, then the force every 100 steps into the
 * Cd window and the lift series into the Strouhal count (forces.js:39-177);
 * the result is classified against the correlations (validacao.js:26-101).
 *
 * usage: node cfd__amd/node/validacao.js [case id ...]   (default: all fast cases)
 * Prints one JSON object per case.  Node 12 compatible.
 */
'use strict';

const path = require('path');
const { SolverMI355X } = require(path.join(__dirname, 'solver.js'));

const CELULA = { FLUIDO: 0, SOLIDO: 1, ENTRADA: 2, SAIDA: 3, SOLIDO_MOVEL: 4,
  ESPELHO_Y: 5, ESPELHO_Z: 6, PAREDE: 7 };

/* ---- tunnel assembly (tunel.js:39-168) ---- */
function corpoAnalitico(g, f) {
  const { nx, ny, nz } = g;
  const c = new Uint8Array(nx * ny * nz);
  const r2 = f.raio * f.raio;
  for (let z = 0; z < nz; z++) {
    for (let y = 0; y < ny; y++) {
      for (let x = 0; x < nx; x++) {
        const dx = x + 0.5 - f.cx, dy = y + 0.5 - f.cy, dz = z + 0.5 - f.cz;
        const d = f.tipo === 'esfera' ? dx * dx + dy * dy + dz * dz : dx * dx + dz * dz;
        if (d <= r2) c[(z * ny + y) * nx + x] = 2;
      }
    }
  }
  return c;
}

function montarTipos(g, corpo, esteira) {
  const { nx, ny, nz } = g;
  const t = new Uint32Array(nx * ny * nz);
  for (let i = 0; i < t.length; i++) if (corpo[i] !== 0) t[i] = CELULA.SOLIDO;
  const piso = esteira ? CELULA.SOLIDO_MOVEL : CELULA.PAREDE;
  const at = (x, y, z) => (z * ny + y) * nx + x;
  for (let z = 0; z < nz; z++) {
    for (let y = 0; y < ny; y++) {
      t[at(0, y, z)] = CELULA.ENTRADA;
      t[at(nx - 1, y, z)] = CELULA.SAIDA;
    }
  }
  for (let z = 0; z < nz; z++) {
    for (let x = 0; x < nx; x++) {
      t[at(x, 0, z)] = CELULA.ESPELHO_Y;
      t[at(x, ny - 1, z)] = CELULA.ESPELHO_Y;
    }
  }
  for (let y = 0; y < ny; y++) {
    for (let x = 0; x < nx; x++) {
      t[at(x, y, 0)] = piso;
      t[at(x, y, nz - 1)] = CELULA.ESPELHO_Z;
    }
  }
  for (let z = 0; z < nz; z++) {        // inlet / outlet win the corners
    for (let y = 0; y < ny; y++) {
      t[at(0, y, z)] = CELULA.ENTRADA;
      t[at(nx - 1, y, z)] = CELULA.SAIDA;
    }
  }
  const comprimento = Math.max(4, Math.round(nx * 0.18));
  return { tipo: t, esponja: { inicio: nx - 1 - comprimento, comprimento } };
}

function areaFrontalCorpo(corpo, g) {
  const { nx, ny, nz } = g;
  let a = 0;
  for (let z = 1; z < nz - 1; z++) {
    for (let y = 1; y < ny - 1; y++) {
      for (let x = 1; x < nx - 1; x++) {
        if (corpo[(z * ny + y) * nx + x] !== 0) { a++; break; }
      }
    }
  }
  return a;
}

/* ---- forces (forces.js:39-177) ---- */
function maskell(cd, bloqueio, theta) {
  if (!(bloqueio > 0) || !Number.isFinite(cd) || cd <= 0) return cd;
  return cd / (1 + theta * cd * bloqueio);
}

function coeficientes(f, uLb, area, g, theta) {
  const q = 0.5 * uLb * uLb;
  const denom = Math.max(q * area, 1e-30);
  const cdBruto = f[0] / denom;
  const bloqueio = area / (g.ny * g.nz);
  return { cd: maskell(cdBruto, bloqueio, theta), cdBruto, cl: f[2] / denom, cs: f[1] / denom,
    bloqueio };
}

class JanelaCd {
  constructor(capacidade) { this.capacidade = capacidade; this.buf = []; this.i = 0; }
  adicionar(v) {
    if (!Number.isFinite(v)) return;
    if (this.buf.length < this.capacidade) this.buf.push(v);
    else { this.buf[this.i] = v; this.i = (this.i + 1) % this.capacidade; }
  }
  get n() { return this.buf.length; }
  get media() {
    if (!this.buf.length) return NaN;
    let s = 0;
    for (const v of this.buf) s += v;
    return s / this.buf.length;
  }
  get desvio() {
    const n = this.buf.length;
    if (n < 2) return NaN;
    const m = this.media;
    let s = 0;
    for (const v of this.buf) s += (v - m) ** 2;
    return Math.sqrt(s / (n - 1));
  }
}

function strouhal(serie, passosPorAmostra, d, uLb) {
  const n = serie.length;
  if (n < 32) return { st: NaN, motivo: 'series too short' };
  let m = 0;
  for (const v of serie) m += v;
  m /= n;
  const osc = serie.map((v) => v - m);
  let acc = 0;
  for (const b of osc) acc += b * b;
  const amp = Math.sqrt(acc / n);
  if (amp < 1e-6) return { st: 0, motivo: 'no oscillation: steady wake' };
  const limiar = 0.25 * amp;
  let cruz = 0, ultimo = 0;
  for (const v of osc) {
    const s = v > limiar ? 1 : (v < -limiar ? -1 : 0);
    if (s !== 0 && ultimo !== 0 && s !== ultimo) cruz++;
    if (s !== 0) ultimo = s;
  }
  if (cruz < 4) return { st: NaN, motivo: `only ${cruz} crossings; simulate longer` };
  const periodos = cruz / 2;
  const freqLb = periodos / (n * passosPorAmostra);
  return { st: freqLb * d / uLb, periodos, freqLb, amplitude: amp };
}

/* ---- correlations and cases (validacao.js:26-205) ---- */
function esferaCd(re) {
  if (!(re > 0)) return Infinity;
  if (re < 0.1) return 24 / re;
  return (24 / re) * (1 + 0.15 * re ** 0.687) + 0.42 / (1 + 4.25e4 * re ** -1.16);
}
function cilindroCd(re) {
  if (!(re > 0)) return Infinity;
  if (re < 1) return 8 * Math.PI / (re * (2.002 - Math.log(re)));
  if (re < 2e5) return 1.0 + 10 * re ** (-2 / 3);
  return 0.35;
}
function cilindroStrouhal(re) {
  if (re < 47) return 0;
  if (re < 300) return 0.2684 - 1.0356 / Math.sqrt(re);
  if (re < 2e5) return 0.198 * (1 - 19.7 / re);
  return 0.21;
}
function classificar(medido, ref) {
  if (!Number.isFinite(medido) || !Number.isFinite(ref) || ref === 0) {
    return { erro: NaN, pct: NaN, nivel: 'indefinido', ok: false };
  }
  const erro = (medido - ref) / ref;
  const a = Math.abs(erro) * 100;
  if (a <= 5) return { erro, pct: a, nivel: 'excelente', ok: true };
  if (a <= 15) return { erro, pct: a, nivel: 'aceitavel', ok: true };
  if (a <= 30) return { erro, pct: a, nivel: 'qualitativo', ok: false };
  return { erro, pct: a, nivel: 'discordante', ok: false };
}
function uLbParaOmega(re, d, omegaAlvo = 1.90, uLbMax = 0.1) {
  const nu = (1 / 3) * (1 / omegaAlvo - 0.5);
  return Math.min(nu * re / d, uLbMax);
}

const CASOS = [
  { id: 'esfera-re50', forma: 'esfera', re: 50, diametroCelulas: 24,
    grade: { nx: 240, ny: 120, nz: 96 }, referencia: () => esferaCd(50),
    travessiasTransiente: 2.0, travessiasMedida: 1.0 },
  { id: 'esfera-re100', forma: 'esfera', re: 100, diametroCelulas: 24,
    grade: { nx: 240, ny: 120, nz: 96 }, referencia: () => esferaCd(100),
    travessiasTransiente: 2.0, travessiasMedida: 1.5 },
  { id: 'cilindro-re100', forma: 'cilindro', re: 100, diametroCelulas: 16,
    grade: { nx: 480, ny: 8, nz: 320 }, referencia: () => cilindroCd(100),
    referenciaSt: () => cilindroStrouhal(100),
    perturbacao: { grausGuinada: 2.0, passos: 600 },
    travessiasTransiente: 2.5, travessiasMedida: 4.0 },
  { id: 'esfera-re200', forma: 'esfera', re: 200, diametroCelulas: 40,
    grade: { nx: 480, ny: 240, nz: 192 }, referencia: () => esferaCd(200),
    travessiasTransiente: 2.0, travessiasMedida: 1.5, profundo: true },
];

const AMOSTRA = 100;

async function rodarCaso(caso, precisao = 'fp32', janela = 4000) {
  const t0 = Date.now();
  const g = caso.grade;
  const D = caso.diametroCelulas;
  const uLb = uLbParaOmega(caso.re, D, 1.90);
  const nuLb = uLb * D / caso.re;
  const omega = 1 / (nuLb / (1 / 3) + 0.5);
  const corpo = corpoAnalitico(g, { tipo: caso.forma, cx: g.nx * 0.30, cy: g.ny * 0.5,
    cz: g.nz * 0.5, raio: D / 2 });
  const { tipo, esponja } = montarTipos(g, corpo, false);
  for (let y = 0; y < g.ny; y++) {       // mirror floor: the body in free air
    for (let x = 1; x < g.nx - 1; x++) tipo[y * g.nx + x] = CELULA.ESPELHO_Z;
  }
  const ctx = await SolverMI355X.criarDevice();
  const s = await SolverMI355X.criar(ctx, g, precisao);
  try {
    s.definirTipos(tipo);
    s.configurar({ omegaPlus: omega, lesCs: 0, inletU: [uLb, 0, 0], esponja });
    s.inicializar();
    const af = areaFrontalCorpo(corpo, g);
    const travessia = Math.ceil(g.nx / uLb);
    const transiente = Math.floor(caso.travessiasTransiente * travessia + 0.5);
    const medida = Math.floor(caso.travessiasMedida * travessia + 0.5);
    if (caso.perturbacao) {
      const t = caso.perturbacao.grausGuinada * Math.PI / 180;
      s.configurar({ omegaPlus: omega, lesCs: 0, esponja,
        inletU: [uLb * Math.cos(t), 0, uLb * Math.sin(t)] });
      s.passo(caso.perturbacao.passos);
      s.configurar({ omegaPlus: omega, lesCs: 0, inletU: [uLb, 0, 0], esponja });
    }
    for (let k = 0; k < transiente; k += 2000) s.passo(Math.min(2000, transiente - k));
    const jan = new JanelaCd(janela);
    const serieCl = [];
    let divergiu = null;
    for (let k = 0; k < medida; k += AMOSTRA) {
      s.passo(AMOSTRA);
      const f = await s.medirForcas();
      const co = coeficientes(f, uLb, af, g, caso.forma === 'cilindro' ? 1.0 : 2.5);
      if (!Number.isFinite(co.cd)) { divergiu = s.passos; break; }
      jan.adicionar(co.cd);
      serieCl.push(co.cl);
    }
    const massa = await s.massa();
    const ref = caso.referencia();
    const out = { id: caso.id, cd: jan.media, cdDesvio: jan.desvio, referencia: ref,
      classificacao: classificar(jan.media, ref), omega, uLb, areaFrontal: af,
      passos: transiente + medida, divergiuEm: divergiu, naoFinitos: massa[1],
      segundos: (Date.now() - t0) / 1000 };
    if (caso.referenciaSt) {
      const st = strouhal(serieCl, AMOSTRA, D, uLb);
      out.st = st.st;
      out.stReferencia = caso.referenciaSt();
      out.stClassificacao = classificar(st.st, out.stReferencia);
    }
    return out;
  } finally {
    s.destruir();
  }
}

module.exports = { CASOS, rodarCaso, corpoAnalitico, montarTipos, areaFrontalCorpo, coeficientes,
  maskell, JanelaCd, strouhal, esferaCd, cilindroCd, cilindroStrouhal, classificar, uLbParaOmega };

if (require.main === module) {
  (async () => {
    const ids = process.argv.slice(2);
    const casos = ids.length ? CASOS.filter((c) => ids.includes(c.id))
      : CASOS.filter((c) => !c.profundo);
    if (ids.length && casos.length !== ids.length) throw new Error(`unknown case in ${ids}`);
    for (const c of casos) console.log(JSON.stringify(await rodarCaso(c)));
  })().catch((e) => { console.error(e); process.exit(1); });
}
