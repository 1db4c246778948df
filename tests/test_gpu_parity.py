"""GPU kernels against the CPU oracle on identical inputs (float32, 1e-5).

The oracle (oracle/lbm_oracle.c) restates the reference's emitted shader
operation by operation; these tests run the HIP path through the C ABI on the
same seeded populations and cell types and compare:

  * populations after 1, 10 and 100 steps, relative to max |g| (non-wall
    cells; wall cells are left stale by both, as in the reference);
  * the macro field of the last step, relative to max |u| (and max |delta|);
  * init, atualizarMacros and the momentum-exchange force.

Tolerance: 1e-5 relative for the default float32 step, whose per-cell
arithmetic uses 1-ulp reciprocals and square roots and fused multiply-adds
(inside WGSL's own accuracy bounds); with CFD_ROUND_EXACT the step is the
oracle's float32 arithmetic and must agree bit for bit.
"""
import zlib

import numpy as np
import pytest

from cfd__amd import lattice, tunel

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _solver():
    from cfd__amd.solver import SolverMI355X
    return SolverMI355X


def _random_state(grid, seed, amp=0.02, noise=2e-4):
    """Equilibrium of a smooth random velocity / density field plus noise."""
    nx, ny, nz = grid
    rng = np.random.default_rng(seed)
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    u = []
    for a in range(3):
        k = rng.integers(1, 3, size=3)
        ph = rng.random(3) * 2 * np.pi
        u.append(amp * np.sin(2 * np.pi * k[0] * X / nx + ph[0]) *
                 np.cos(2 * np.pi * k[1] * Y / ny + ph[1]) *
                 np.sin(2 * np.pi * k[2] * Z / nz + ph[2]) + (0.05 if a == 0 else 0.0))
    delta = 1e-3 * np.sin(2 * np.pi * X / nx) * np.cos(2 * np.pi * Z / nz)
    g = lattice.equilibrium(delta, u).reshape(19, -1)
    g = g + noise * rng.standard_normal(g.shape) * lattice.W[:, None]
    return np.ascontiguousarray(g, dtype=np.float32)


def _tunnel_types(grid, esteira=True, seed=0):
    nx, ny, nz = grid
    g = {"nx": nx, "ny": ny, "nz": nz}
    corpo = tunel.corpoAnalitico(g, {"tipo": "esfera", "cx": nx * 0.35, "cy": ny * 0.5,
                                     "cz": nz * 0.45, "raio": min(ny, nz) * 0.2})
    tipo, esponja, _ = tunel.montarTipos(g, corpo, esteira=esteira)
    return tipo, esponja


def _nonwall(tipo):
    t = np.asarray(tipo)
    return ~((t == 1) | (t == 4) | (t == 7))


def _rel(a, b, mask=None):
    if mask is not None:
        a, b = a[..., mask], b[..., mask]
    scale = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) / scale


CASES = [
    # name, grid, tipo-builder, configurar kwargs
    ("tunnel_rolling_road_les_bl", (40, 24, 20), "tunnel",
     dict(omegaPlus=1.7, lesCs=0.1, inletU=[0.05, 0.0, 0.0], beltU=[0.05, 0.0, 0.0],
          espessuraCL=5.0)),
    ("tunnel_fixed_floor_yaw", (37, 13, 11), "tunnel_fixed",
     dict(omegaPlus=1.2, lesCs=0.16, inletU=[0.06, 0.004, 0.002], beltU=[0.0, 0.0, 0.0])),
    ("periodic_fluid", (64, 16, 12), "fluid",
     dict(omegaPlus=1.9, lesCs=0.1, inletU=[0.05, 0.0, 0.0])),
    ("all_types_random", (33, 17, 9), "random",
     dict(omegaPlus=1.5, lesCs=0.1, inletU=[0.04, 0.01, -0.01], beltU=[0.03, 0.01, 0.0],
          espessuraCL=3.0, esponja={"inicio": 20, "comprimento": 9})),
]


def _types(kind, grid, seed=3):
    nx, ny, nz = grid
    N = nx * ny * nz
    if kind == "tunnel":
        return _tunnel_types(grid, True)[0]
    if kind == "tunnel_fixed":
        return _tunnel_types(grid, False)[0]
    if kind == "fluid":
        return np.zeros(N, dtype=np.uint32)
    rng = np.random.default_rng(seed)
    # every code 0..7 plus out-of-range values, mostly fluid
    t = rng.choice(np.array([0, 0, 0, 0, 0, 0, 1, 2, 3, 4, 5, 6, 7, 9, 1000], dtype=np.uint32),
                   size=N)
    return t.astype(np.uint32)


@pytest.mark.parametrize("precisao", ["fp32", "fp32exact"])
@pytest.mark.parametrize("name,grid,kind,cfg", CASES, ids=[c[0] for c in CASES])
def test_step_parity(orc, name, grid, kind, cfg, precisao):
    S = _solver()
    nx, ny, nz = grid
    tipo = _types(kind, grid)
    g0 = _random_state(grid, seed=zlib.crc32(name.encode()) % 1000)
    par = orc.params_array(nx=nx, **cfg)
    s = S.criar({"nx": nx, "ny": ny, "nz": nz}, precisao=precisao)
    try:
        s.definirTipos(tipo)
        s.configurar(**cfg)
        s.frente = "A"
        s.definirPopulacoes(g0)
        ref = g0.copy()
        mask = _nonwall(tipo)
        done = 0
        for target in (1, 10, 100):
            n = target - done
            s.passo(n)
            ref, mref = orc.run(grid, ref, tipo, par, n)
            done = target
            got = s.lerPopulacoes()
            e = _rel(got, ref.reshape(19, -1), mask)
            assert e <= TOL, f"{name}: populations after {target} steps rel err {e:.3e}"
            m = s.lerMacros().reshape(-1, 4)
            mr = mref.reshape(-1, 4)
            eu = _rel(m[:, :3].T, mr[:, :3].T)
            ed = _rel(m[:, 3], mr[:, 3])
            assert eu <= TOL, f"{name}: macros u after {target} steps rel err {eu:.3e}"
            assert ed <= 1e-4, f"{name}: macros delta after {target} steps rel err {ed:.3e}"
            # forces from the current front
            fg = s.forcaBruta()
            fo = orc.forces(grid, ref, tipo)
            scale = max(np.abs(fo).max(), 1e-12)
            assert np.abs(fg - fo).max() <= 1e-5 * scale + 1e-9, (name, target, fg, fo)
    finally:
        s.destruir()


@pytest.mark.parametrize("name,grid,kind,cfg", CASES, ids=[c[0] for c in CASES])
def test_exact_rounding_is_the_oracle_bit_for_bit(orc, name, grid, kind, cfg):
    """CFD_ROUND_EXACT: every operation of the step rounds as the oracle's
    float32 restatement of the shader text (correctly rounded division and
    square root, no contraction), so populations and macros after 1, 10 and
    40 steps are the oracle's bits in every non-wall cell, and the force (a
    float64 sum in another order) agrees to 1e-12."""
    S = _solver()
    nx, ny, nz = grid
    tipo = _types(kind, grid)
    g0 = _random_state(grid, seed=zlib.crc32(name.encode()) % 1000 + 1)
    par = orc.params_array(nx=nx, **cfg)
    s = S.criar({"nx": nx, "ny": ny, "nz": nz}, precisao="fp32exact")
    try:
        s.definirTipos(tipo)
        s.configurar(**cfg)
        s.frente = "A"
        s.definirPopulacoes(g0)
        ref = g0.copy()
        mask = _nonwall(tipo)
        done = 0
        for target in (1, 10, 40):
            s.passo(target - done)
            ref, mref = orc.run(grid, ref, tipo, par, target - done)
            done = target
            got = s.lerPopulacoes()
            r = ref.reshape(19, -1)
            same = got[:, mask].view(np.uint32) == r[:, mask].view(np.uint32)
            assert same.all(), (name, target, int((~same).sum()))
            m = s.lerMacros().reshape(-1, 4)
            mr = mref.reshape(-1, 4)
            assert np.array_equal(m[mask].view(np.uint32), mr[mask].view(np.uint32)), (
                name, target)
            fg, fo = s.forcaBruta(), orc.forces(grid, ref, tipo)
            assert np.abs(fg - fo).max() <= 1e-12 * max(np.abs(fo).max(), 1.0), (fg, fo)
    finally:
        s.destruir()


def test_init_and_macros_parity(orc):
    S = _solver()
    grid = (40, 24, 20)
    tipo, esponja = _tunnel_types(grid, True)
    cfg = dict(omegaPlus=1.6, lesCs=0.0, inletU=[0.05, 0.01, 0.0], beltU=[0.05, 0.0, 0.0],
               esponja=esponja)
    s = S.criar(grid)
    try:
        s.definirTipos(tipo)
        s.configurar(**cfg)
        s.inicializar()
        assert s.frente == "B"
        got = s.lerPopulacoes()
        ref = orc.init(grid, tipo, orc.params_array(nx=grid[0], **cfg)).reshape(19, -1)
        assert _rel(got, ref) <= 1e-6
        s.atualizarMacros()
        m = s.lerMacros()
        mr = orc.macros(grid, ref)
        assert _rel(m, mr) <= 1e-6
    finally:
        s.destruir()


def test_populations_roundtrip_exact():
    S = _solver()
    grid = (19, 7, 5)
    g = np.random.default_rng(1).standard_normal((19, 19 * 7 * 5)).astype(np.float32)
    s = S.criar(grid)
    try:
        for f in ("A", "B"):
            s.frente = f
            s.definirPopulacoes(g)
            assert np.array_equal(s.lerPopulacoes(), g)
    finally:
        s.destruir()


def test_rest_offset_closed_form():
    """On zeroed buffers the force is sum over body links of 2 w_i c_i
    (solver.js:454-479): zero for an immersed closed body, non-zero and
    equal to the link count formula for a body standing on the floor."""
    S = _solver()
    nx, ny, nz = 30, 20, 16
    g = {"nx": nx, "ny": ny, "nz": nz}
    # immersed sphere: symmetric, closed -> offset 0
    corpo = tunel.corpoAnalitico(g, {"tipo": "esfera", "cx": 12, "cy": 10, "cz": 8, "raio": 4})
    tipo, esp, _ = tunel.montarTipos(g, corpo, esteira=False)
    s = S.criar(g)
    try:
        s.definirTipos(tipo)
        off = s.calibrarRepouso()
        assert np.abs(off).max() < 1e-12
        # block resting on the floor: only links from fluid cells count
        t = tipo.reshape(nz, ny, nx).copy()
        t[1:5, 8:12, 10:14] = 1
        t = t.reshape(-1)
        s.definirTipos(t)
        off = np.array(s.calibrarRepouso())
        # brute force: sum over FLUIDO cells and SOLIDO neighbours of 2 w_i c_i
        tt = t.reshape(nz, ny, nx)
        exp = np.zeros(3)
        for i in range(1, 19):
            c = lattice.C[i]
            nb = np.roll(tt, shift=(-c[2], -c[1], -c[0]), axis=(0, 1, 2))
            n = np.count_nonzero((tt == 0) & (nb == 1))
            exp += 2 * lattice.W[i] * c * n
        assert np.allclose(off, exp, rtol=0, atol=1e-12), (off, exp)
        assert abs(off[2]) > 0.1  # the floor contact leaves a downward offset
    finally:
        s.destruir()


def test_force_link_list_follows_the_types(orc):
    """The body links are listed when the types are set: re-setting the types
    (a larger body, then a smaller one, then none) must change the force like
    the oracle's scan over the type field does.  Periodic fluid around the
    bodies; the grid (31x21x17 = 11 067 cells) is not a multiple of the 16
    cells a lane scans, and the second field puts a solid bar in the last row,
    so link cells fall in the final partial word."""
    S = _solver()
    grid = (31, 21, 17)
    nx, ny, nz = grid
    g = {"nx": nx, "ny": ny, "nz": nz}
    pops = _random_state(grid, 5)

    def field(body, tail=False):
        t = np.zeros((nz, ny, nx), dtype=np.uint32)
        if body:
            t.reshape(-1)[tunel.corpoAnalitico(g, body).reshape(-1) != 0] = 1
        if tail:
            t[nz - 1, ny - 1, 24:28] = 1
        return t.reshape(-1)

    fields = [
        field({"tipo": "esfera", "cx": 10, "cy": 10, "cz": 8, "raio": 3}),
        field({"tipo": "esfera", "cx": 20, "cy": 11, "cz": 9, "raio": 6}, tail=True),
        field({"tipo": "esfera", "cx": 8, "cy": 6, "cz": 5, "raio": 2}),
        field(None),
    ]
    s = S.criar(g)
    try:
        for k, tipo in enumerate(fields):
            s.definirTipos(tipo)
            s.definirPopulacoes(pops)
            fg = s.forcaBruta()
            fo = orc.forces(grid, pops, tipo)
            scale = max(np.abs(fo).max(), 1e-12)
            assert np.abs(fg - fo).max() <= 1e-9 * scale + 1e-12, (k, fg, fo)
        assert np.all(fg == 0.0)
    finally:
        s.destruir()


def test_layouts_and_launch_shapes_are_bit_identical():
    """The population layout (SoA or 512-cell blocks) and the launch shape
    (workgroup size, XCD chunking, ascending, alternating or rotated cell
    order) change addressing only: the populations after 25 steps agree bit
    for bit."""
    grid = (72, 40, 36)                 # 103 680 cells: not a multiple of 512
    tipo, esponja = _tunnel_types(grid, esteira=True)
    g0 = _random_state(grid, 11)
    cfg = dict(omegaPlus=1.85, lesCs=0.12, inletU=[0.06, 0.0, 0.0], beltU=[0.06, 0.0, 0.0],
               esponja=esponja, espessuraCL=5.0)
    out = {}
    for t in ({"block": 256, "remap": 8, "layout": 1, "sweep": 0},
              {"block": 512, "remap": 0, "layout": 1, "sweep": 0},
              {"block": 256, "remap": 0, "layout": 1, "sweep": 1},
              {"block": 512, "remap": 64, "layout": 1, "sweep": 1},
              {"block": 256, "remap": 8, "layout": 1, "sweep": 1},
              {"block": 256, "remap": 2, "layout": 0, "sweep": 0},
              {"block": 512, "remap": 0, "layout": 0, "sweep": 1},
              {"block": 256, "remap": 8, "layout": 1, "sweep": 0, "rotate": 37},
              {"block": 512, "remap": 0, "layout": 0, "sweep": 0, "rotate": 150},
              {"block": 256, "remap": 0, "layout": 1, "sweep": 0, "rotate": 1000}):
        s = _solver().criar({"nx": grid[0], "ny": grid[1], "nz": grid[2]})
        try:
            s.ajuste = t
            s.definirTipos(tipo)
            s.configurar(**cfg)
            s.frente = "A"
            s.definirPopulacoes(g0)
            s.passo(25)
            s.atualizarMacros()
            out[str(t)] = (s.lerPopulacoes(), s.lerMacros())
        finally:
            s.destruir()
    ref_p, ref_m = next(iter(out.values()))
    wall = np.isin(tipo, [1, 4, 7])
    for k, (p, m) in out.items():
        assert np.array_equal(p[:, ~wall], ref_p[:, ~wall]), k
        assert np.array_equal(m, ref_m), k


@pytest.mark.parametrize("grid", [(72, 40, 36), (37, 23, 19)])
@pytest.mark.parametrize("precisao,fechamento", [("fp16", None), ("fp16sr", None),
                                                  ("fp16sr", "silu"), ("fp32", "silu"),
                                                  ("fp16sr", "hard_silu")])
def test_float16_and_closure_layouts_are_bit_identical(precisao, fechamento, grid):
    """The float16 steps (stored units, paired conversions, stochastic rounding
    drawn from the cell, direction and step only) and the closure step (SiLU
    and hard-SiLU networks) give the same bits in the SoA layout (the default
    for float16 sets up to 128 MB) and in 512-cell blocks, in several launch
    shapes of the plain step, also on a grid whose last block and last wave
    are partly empty (37x23x19)."""
    from cfd__amd.fechamento import Fechamento, distancia_parede, iniciar_mlp
    nx, ny, nz = grid
    g = {"nx": nx, "ny": ny, "nz": nz}
    tipo, esponja = _tunnel_types(grid, esteira=True)
    g0 = _random_state(grid, 13)
    cfg = dict(omegaPlus=1.85, lesCs=0.12, inletU=[0.06, 0.0, 0.0], beltU=[0.06, 0.0, 0.0],
               esponja=esponja, espessuraCL=5.0)
    fech = None
    if fechamento:
        rng = np.random.default_rng(4)
        prm = iniciar_mlp(4, bias_saida=0.3)
        shift = rng.normal(0, 1e-3, 8).astype(np.float32)
        scale = (1.0 / rng.uniform(5e-4, 2e-3, 8)).astype(np.float32)
        shift[7], scale[7] = 5.0, 0.1
        fech = Fechamento(prm, shift, scale, np.full(8, -60, np.float32),
                          np.full(8, 60, np.float32), 2e-3, fechamento)
        wd = distancia_parede(tipo, g)
    shapes = ({"layout": 0, "block": 256, "sweep": 0}, {"layout": 1, "block": 256, "sweep": 0},
              {"layout": 1, "block": 512, "sweep": 1}, {"layout": 0, "block": 512, "sweep": 0,
                                                         "rotate": 77})
    out = {}
    for t in shapes:
        s = _solver().criar(g, precisao=precisao)
        try:
            s.ajuste = t
            s.definirTipos(tipo)
            s.configurar(**cfg)
            s.frente = "A"
            s.definirPopulacoes(g0)
            if fech is not None:
                s.definirDistanciaParede(wd)
                s.definirFechamento(fech)
            s.passo(12)
            s.atualizarMacros()
            out[str(t)] = (s.lerPopulacoes(), s.lerMacros())
        finally:
            s.destruir()
    ref_p, ref_m = next(iter(out.values()))
    wall = np.isin(tipo, [1, 4, 7])
    for k, (pp, m) in out.items():
        assert np.array_equal(pp[:, ~wall], ref_p[:, ~wall]), k
        assert np.array_equal(m, ref_m), k


def test_default_sweep_and_tuning_errors():
    """Alternating order is the default for sets above 128 MB in float16
    and for float32 sets of 128-256 MB (one set fits the last-level cache,
    two do not); float16 sets up to 128 MB are SoA, every other set below
    4 GiB blocked; bad tuning values raise."""
    from cfd__amd._lib import CFDError
    S = _solver()
    for g, prec, want in (({"nx": 72, "ny": 40, "nz": 36}, "fp32", 0),
                          ({"nx": 240, "ny": 120, "nz": 96}, "fp32", 1),
                          ({"nx": 320, "ny": 160, "nz": 128}, "fp32", 0),
                          ({"nx": 320, "ny": 160, "nz": 128}, "fp16", 1),
                          ({"nx": 160, "ny": 80, "nz": 64}, "fp16", 0),
                          ({"nx": 240, "ny": 120, "nz": 96}, "fp16", 0)):
        s = S.criar(g, precisao=prec)
        try:
            assert s.ajuste["sweep"] == want, (g, prec)
            # float32 sets above the 256 MB cache: 512 threads, start rotated
            big32 = prec == "fp32" and g["nx"] * g["ny"] * g["nz"] * 76 > 256 << 20
            assert s.ajuste["block"] == (512 if big32 else 256), (g, prec)
            assert (s.ajuste["rotate"] > 0) == big32, (g, prec)
            small16 = prec == "fp16" and g["nx"] * g["ny"] * g["nz"] * 38 <= 128 << 20
            assert s.ajuste["layout"] == (0 if small16 else 1), (g, prec)
            for bad in ({"sweep": 2}, {"block": 128}, {"remap": 3}, {"layout": 2},
                        {"sweep": 0, "rotate": -1}, {"sweep": 1, "rotate": 5}):
                with pytest.raises(CFDError):
                    s.ajuste = bad
            assert s.ajuste["sweep"] == want
        finally:
            s.destruir()


def test_step_parity_soa_beyond_4gib(orc):
    """A float32 population set above 4 GiB (480x360x330, 57 M cells, 4.33 GB)
    falls back to the SoA layout with 64-bit slice offsets; two steps (A -> B
    -> A) of a sphere tunnel agree with the oracle, and so do the macros and
    the force."""
    from cfd__amd._lib import CFDError
    grid = (480, 360, 330)
    nx, ny, nz = grid
    N = nx * ny * nz
    assert N * 19 * 4 > 1 << 32
    tipo, esponja = _tunnel_types(grid, esteira=True)
    cfg = dict(omegaPlus=1.8, lesCs=0.1, inletU=[0.05, 0.0, 0.0], beltU=[0.05, 0.0, 0.0],
               esponja=esponja, espessuraCL=6.0)
    rng = np.random.default_rng(404)
    g0 = rng.standard_normal((19, N), dtype=np.float32)
    g0 *= (2e-3 * lattice.W[:, None]).astype(np.float32)
    s = _solver().criar({"nx": nx, "ny": ny, "nz": nz})
    try:
        assert s.ajuste["layout"] == 0
        with pytest.raises(CFDError):
            s.ajuste = {"layout": 1}
        s.definirTipos(tipo)
        s.configurar(**cfg)
        s.frente = "A"
        s.definirPopulacoes(g0)
        s.passo(2)
        got = s.lerPopulacoes()
        m = s.lerMacros().reshape(-1, 4)
        fg = s.forcaBruta()
    finally:
        s.destruir()
    ref, mref = orc.run(grid, g0, tipo, orc.params_array(nx=nx, **cfg), 2)
    del g0
    ref = ref.reshape(19, N)
    mask = _nonwall(tipo)
    scale = float(np.abs(ref[:, mask]).max())
    for i in range(19):
        e = float(np.abs(got[i, mask] - ref[i, mask]).max()) / scale
        assert e <= TOL, (i, e)
    mr = mref.reshape(-1, 4)
    assert _rel(m[:, :3].T, mr[:, :3].T) <= TOL
    fo = orc.forces(grid, ref, tipo)
    assert np.abs(fg - fo).max() <= 1e-5 * max(np.abs(fo).max(), 1e-12) + 1e-9, (fg, fo)


def test_launch_order_autotuner_keeps_the_trajectory():
    """autoajustarPasso times the candidate launch orders on live steps, in
    calls of the caller's chunk size, and keeps the fastest; since every
    order is bit-identical and the inflow ramp fires at the same steps, the
    state after the tuning steps plus 20 more equals a loop of passo(chunk)
    for as many steps, with the ramp running through both."""
    grid = (72, 40, 36)
    tipo, esponja = _tunnel_types(grid, esteira=True)
    g0 = _random_state(grid, 5)
    cfg = dict(omegaPlus=1.8, lesCs=0.1, beltU=[0.05, 0.0, 0.0], esponja=esponja,
               espessuraCL=4.0)
    chunk = 6
    out = []
    for tune in (True, False):
        s = _solver().criar({"nx": grid[0], "ny": grid[1], "nz": grid[2]})
        try:
            s.definirTipos(tipo)

            def aplicar(f, s=s):
                s.configurar(inletU=[0.05 * f, 0.0, 0.0], **cfg)

            aplicar(1.0)
            s.frente = "A"
            s.definirPopulacoes(g0)
            s.rampa(400, aplicar)
            total = s.passosAutoajuste(passos=12, bloco=chunk)
            assert total == 10 * 3 * chunk
            if tune:
                t = s.autoajustarPasso(passos=12, bloco=chunk)
                assert t["sweep"] in (0, 1) and t["rotate"] >= 0
                assert t["sweep"] == 0 or t["rotate"] == 0
                assert s.passos == total
                assert t["block"] in (256, 512)
            else:
                for _ in range(total // chunk):
                    s.passo(chunk)
            assert 0.0 < s.fatorRampa < 1.0
            s.passo(20)
            out.append(s.lerPopulacoes())
        finally:
            s.destruir()
    wall = np.isin(tipo, [1, 4, 7])
    assert np.array_equal(out[0][:, ~wall], out[1][:, ~wall])
