/* cfd__amd/node/solver.js - SolverMI355X for Node.js.
 *
 * Same public API as the reference's SolverWebGPU
 * (src/backend/webgpu/solver.js) so the app and its test pages can swap it
 * in: criarDevice / criar / configurar / definirTipos / rampa / fatorRampa /
 * inicializar / definirCampo / passo / lerMacros / atualizarMacros /
 * calibrarRepouso / medirForcas / destruir, and the fields nx, ny, nz, N,
 * passos, frente, bytesTotais, cfg.  The device work runs in libcfdmi.so
 * (hand-written HIP for MI355X) through the N-API addon cfdmi.node.
 *
 * Node 12 compatible (no optional chaining, no nullish coalescing).
 */
'use strict';

const path = require('path');
const addon = require(path.join(__dirname, 'cfdmi.node'));

const Q = 19;
const MAGIC_WALL = 3 / 16;

class SolverMI355X {
  constructor(handle, grade, precisao) {
    this._h = handle;
    this.nx = grade.nx; this.ny = grade.ny; this.nz = grade.nz;
    this.N = this.nx * this.ny * this.nz;
    this.precisao = precisao;
    this.passos = 0;
    this.cfg = null;
    this.offsetForca = null;
    this._rampa = null;
  }

  /** The GPU context is process-wide; kept for API parity (solver.js:61-82). */
  static async criarDevice() { return { tipo: 'mi355x', versao: addon.version() }; }

  /** Allocate a solver on the GPU (solver.js:93-124).  precisao 'fp32' |
   *  'fp32exact' (every operation rounded as the CPU restatement rounds it) |
   *  'fp16' | 'fp16sr'. */
  static async criar(ctx, grade, precisao) {
    const p = precisao || 'fp32';
    const flags = { fp32: 0, fp32exact: 4, fp16: 1, fp16sr: 3 }[p];
    if (flags === undefined) {
      throw new Error(`precisao deve ser fp32, fp32exact, fp16 ou fp16sr: ${p}`);
    }
    const h = addon.create(grade.nx, grade.ny, grade.nz, flags);
    return new SolverMI355X(h, grade, p);
  }

  get bytesTotais() { return addon.bytes(this._h); }

  get frente() { return addon.getFront(this._h) ? 'B' : 'A'; }
  set frente(v) {
    if (v !== 'A' && v !== 'B') throw new Error(`frente deve ser 'A' ou 'B': ${v}`);
    addon.setFront(this._h, v === 'B' ? 1 : 0);
  }

  /** solver.js:277-293 */
  configurar({
    omegaPlus, magic = MAGIC_WALL, lesCs = 0.1, inletU,
    beltU = [0, 0, 0], esponja = null, espessuraCL = 0,
  }) {
    const esp = esponja !== null && esponja !== undefined ? esponja : {
      inicio: Math.round(this.nx * 0.82),
      comprimento: Math.max(4, Math.round(this.nx * 0.18)),
    };
    addon.configure(this._h, {
      omegaPlus, magic, lesCs, beltU, inletU: inletU || [0, 0, 0],
      spongeStart: esp.inicio, spongeLen: esp.comprimento, blThickness: espessuraCL,
    });
    this.cfg = { omegaPlus, magic, lesCs, inletU, beltU, esponja: esp, espessuraCL };
  }

  /** Uint32Array of N cell types (solver.js:296-301). */
  definirTipos(tipos) {
    if (tipos.length !== this.N) {
      throw new Error(`campo de tipos com ${tipos.length}, esperado ${this.N}`);
    }
    const t = tipos instanceof Uint32Array ? tipos : Uint32Array.from(tipos);
    addon.setTypes(this._h, t);
  }

  /** solver.js:323-335 */
  rampa(passos, aplicarVelocidade) {
    this._rampa = { passos, aplicar: aplicarVelocidade, atual: 0 };
  }

  get fatorRampa() {
    const r = this._rampa;
    if (!r || r.atual >= r.passos) return 1;
    const t = r.atual / r.passos;
    return t * t * (3 - 2 * t);
  }

  /** Free-stream equilibrium into B; front = B (solver.js:338-349). */
  inicializar() {
    addon.initEquilibrium(this._h);
    this.passos = 0;
    if (this._rampa) this._rampa.atual = 0;
  }

  /** Field from fn(x, y, z) -> {delta, u} through the caller's float64
   *  equilibrium, stored as float32 into the current front (solver.js:363-389). */
  definirCampo(fn, equilibrium) {
    const { nx, ny, nz, N } = this;
    const dados = new Float32Array(Q * N);
    const g = new Float64Array(Q);
    for (let z = 0; z < nz; z++) {
      for (let y = 0; y < ny; y++) {
        for (let x = 0; x < nx; x++) {
          const cell = (z * ny + y) * nx + x;
          const r = fn(x, y, z) || {};
          const delta = r.delta === undefined ? 0 : r.delta;
          const u = r.u === undefined ? [0, 0, 0] : r.u;
          equilibrium(delta, u, g);
          for (let i = 0; i < Q; i++) dados[i * N + cell] = g[i];
        }
      }
    }
    addon.setPopulations(this._h, dados);
    this.passos = 0;
  }

  _avancarRampa(n) {
    const r = this._rampa;
    if (r && r.atual < r.passos) {
      r.atual += n;
      if (r.aplicar) r.aplicar(this.fatorRampa);
    }
  }

  /** Enqueue n steps; the ramp callback runs first (solver.js:395-415). */
  passo(n = 1) {
    this._avancarRampa(n);
    addon.step(this._h, n);
    this.passos += n;
  }

  /** passo(n), then wait; returns the device time of the n launches (ms). */
  passoCronometrado(n) {
    this._avancarRampa(n);
    const ms = addon.stepTimed(this._h, n);
    this.passos += n;
    return ms;
  }

  /** Launch tuning of the step kernel (no reference counterpart): block,
   *  remap, layout, slicePad, sweep, rotate.  Setting merges the given keys
   *  into the current tuning; results are bit-identical under any tuning. */
  get ajuste() { return addon.getTuning(this._h); }
  set ajuste(t) { addon.setTuning(this._h, t); }

  /** Launch-order autotuner (no reference counterpart): for each workgroup
   *  size, times `passos` live steps in ascending and in alternating cell
   *  order and with the step start rotated by each fraction of the launch
   *  grid, keeps the fastest and returns the tuning.  Steps are taken in
   *  calls of `bloco` steps (one settling call, then ceil(passos / bloco)
   *  timed ones per candidate), as a caller's loop of passo(bloco) takes
   *  them: the inflow ramp fires at the same steps and every order gives the
   *  same bits, so the run equals one that calls passo(bloco) throughout. */
  autoajustarPasso(passos = 200, fracoes = [0.02, 0.1, 0.36], blocos = [256, 512], bloco = 100) {
    const remap = this.ajuste.remap;
    const span = remap > 1 ? 8 * remap : 1;
    let cand = [];
    for (const block of blocos) {
      const grade = Math.ceil(Math.ceil(this.N / block) / span) * span;
      cand = cand.concat([{ block, sweep: 0, rotate: 0 }, { block, sweep: 1, rotate: 0 }],
        fracoes.map((f) => ({ block, sweep: 0, rotate: Math.max(1, Math.floor(f * grade)) })));
    }
    const vezes = Math.max(1, Math.ceil(passos / bloco));
    let melhor = null;
    let melhorMs = Infinity;
    for (const c of cand) {
      this.ajuste = c;
      this.passo(bloco);
      let ms = 0;
      for (let k = 0; k < vezes; ++k) ms += this.passoCronometrado(bloco);
      if (ms < melhorMs) { melhor = c; melhorMs = ms; }
    }
    this.ajuste = melhor;
    return this.ajuste;
  }

  /** Float32Array(N*4) of (ux, uy, uz, delta) (solver.js:427-442). */
  async lerMacros() { return addon.readMacros(this._h, 4 * this.N); }

  /** solver.js:511-520 */
  atualizarMacros() { addon.updateMacros(this._h); }

  /** solver.js:480-484 */
  async calibrarRepouso() {
    this.offsetForca = [0, 0, 0];
    this.offsetForca = await this.medirForcas();
    return this.offsetForca;
  }

  /** Momentum-exchange force minus the rest offset (solver.js:486-508). */
  async medirForcas() {
    const v = addon.forces(this._h);
    const o = this.offsetForca || [0, 0, 0];
    return [v[0] - o[0], v[1] - o[1], v[2] - o[2]];
  }

  /** Learned sub-grid closure in every following step (no reference
   *  counterpart): modelo = {params (1377), ativacao ('silu' or
   *  'hard_silu', default 'silu'), shift, scale, lo, hi (8 each),
   *  outScale}, as carregarFechamento returns it; null returns to
   *  Smagorinsky.  Needs the wall distance (calcularDistanciaParede). */
  definirFechamento(modelo) {
    if (modelo === null || modelo === undefined) {
      addon.setClosure(this._h, null);
      return;
    }
    const nomeAtivacao = modelo.ativacao === undefined ? 'silu' : modelo.ativacao;
    if (!Object.prototype.hasOwnProperty.call(ATIVACOES, nomeAtivacao))
      throw new Error(`fechamento: ativacao desconhecida ${nomeAtivacao}`);
    const codigo = ATIVACOES[nomeAtivacao];
    const c = new Float32Array(1411);
    const partes = [['params', 1377], ['shift', 8], ['scale', 8], ['lo', 8], ['hi', 8]];
    let k = 0;
    partes.forEach(([nome, n]) => {
      const v = modelo[nome];
      if (!v || v.length !== n) throw new Error(`fechamento: ${nome} deve ter ${n} valores`);
      c.set(v, k);
      k += n;
      if (nome === 'params') c[k++] = codigo;
    });
    c[k] = modelo.outScale;
    addon.setClosure(this._h, c);
  }

  /** Wall distance of the closure, computed on the GPU from the types set. */
  calcularDistanciaParede() { addon.computeWallDistance(this._h); }

  /** Wall distance from the caller (Float32Array of N distances in cells). */
  definirDistanciaParede(d) {
    addon.setWallDistance(this._h, d instanceof Float32Array ? d : Float32Array.from(d));
  }

  /** Fluid cell-steps that kept the Smagorinsky tau since the last reset. */
  fallbacksFechamento(zerar = true) { return addon.closureFallbacks(this._h, !!zerar); }

  /** Float64 sum of delta over non-wall cells and the non-finite cell count. */
  async massa() { return addon.mass(this._h); }

  /** Current front as Float32Array(19 N), layout pop[i*N + cell]. */
  async lerPopulacoes() { return addon.getPopulations(this._h, Q * this.N); }

  /** Writes Float32Array(19 N) (same layout) into the current front. */
  definirPopulacoes(pops) {
    if (pops.length !== Q * this.N) {
      throw new Error(`populacoes com ${pops.length} valores, esperado ${Q * this.N}`);
    }
    addon.setPopulations(this._h, pops instanceof Float32Array ? pops : Float32Array.from(pops));
    this.passos = 0;
  }

  sincronizar() { addon.synchronize(this._h); }

  /** Frees the GPU buffers (solver.js:523-527). */
  destruir() { addon.destroy(this._h); }
}

/** voxelizar() of src/geom/voxel.js:376 on the GPU, bit-identical. */
function voxelizar(positions, indices, g, opts) {
  const o = opts || {};
  const banda = o.banda === undefined ? 1.5 : o.banda;
  const selar = o.selarFrestas === undefined ? true : !!o.selarFrestas;
  const pos = positions instanceof Float32Array ? positions : Float32Array.from(positions);
  const idx = indices instanceof Uint32Array ? indices : Uint32Array.from(indices);
  return addon.voxelize(pos, idx, g.nx, g.ny, g.nz, banda, selar, true);
}

const VOXEL = { FLUIDO: 0, CASCA: 1, DENTRO: 2 };
// hidden activation codes of cfd_mlp.activation (CFD_ACT_SILU, CFD_ACT_HARD_SILU)
const ATIVACOES = { silu: 0, hard_silu: 1 };

/** A closure model exported by cfd__amd.fechamento.Fechamento.salvarJson. */
function carregarFechamento(caminho) {
  const d = JSON.parse(require('fs').readFileSync(caminho, 'utf8'));
  return {
    params: Float32Array.from(d.params), shift: Float32Array.from(d.shift),
    scale: Float32Array.from(d.scale), lo: Float32Array.from(d.lo), hi: Float32Array.from(d.hi),
    outScale: d.out_scale, ativacao: d.ativacao === undefined ? 'silu' : d.ativacao,
  };
}

module.exports = { SolverMI355X, voxelizar, carregarFechamento, VOXEL, addon };
