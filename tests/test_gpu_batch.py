"""Batched tunnels (one cooperative launch) against the single-tunnel path."""
import numpy as np
import pytest

from cfd__amd import tunel

pytestmark = pytest.mark.gpu


def _tunnel(nx, ny, nz, forma, D, u, omega):
    g = {"nx": nx, "ny": ny, "nz": nz}
    corpo = tunel.corpoAnalitico(g, {"tipo": forma, "cx": nx * 0.3, "cy": ny * 0.5,
                                     "cz": nz * 0.5, "raio": D / 2})
    tipo, esp, _ = tunel.montarTipos(g, corpo, esteira=False)
    t3 = tipo.reshape(nz, ny, nx)
    t3[0, :, 1:nx - 1] = 6
    return {"grade": g, "tipo": tipo,
            "cfg": {"omegaPlus": omega, "lesCs": 0.1, "inletU": [u, 0, 0], "esponja": esp}}


CASES = [_tunnel(40, 20, 20, "esfera", 6, 0.05, 1.7), _tunnel(64, 32, 24, "cilindro", 8, 0.06, 1.8),
         _tunnel(33, 17, 19, "esfera", 5, 0.04, 1.5), _tunnel(96, 40, 40, "esfera", 10, 0.05, 1.85),
         _tunnel(48, 24, 24, "cilindro", 6, 0.05, 1.6)]


def test_batch_matches_single_tunnel_solver():
    from cfd__amd.batch import LoteMI355X
    from cfd__amd.solver import SolverMI355X
    steps, amostra = 120, 30
    lote = LoteMI355X(CASES)
    try:
        # a cap of 110 tiles of 512 cells admits the three smallest tunnels
        # together, then the two larger ones one after the other
        lote.rodar(steps, amostra, {"block": 256, "tile_cells": 512, "wg_per_cu": 0,
                                    "max_active_tiles": 110})
        res = lote.resultados()
        pops = [lote.populacoes(k) for k in range(len(CASES))]
    finally:
        lote.destruir()
    assert res["forcas"].shape == (len(CASES), steps // amostra, 3)
    assert res["rodadas"] == 3 * steps
    fin = res["fim_ms"]
    assert (fin > 0).all()
    for k, c in enumerate(CASES):
        s = SolverMI355X.criar(c["grade"])
        try:
            s.definirTipos(c["tipo"])
            s.configurar(**c["cfg"])
            s.inicializar()
            forces = []
            for n in range(1, steps + 1):
                if n % amostra == 0:     # the batch samples the state after n-1 steps
                    forces.append(s.forcaBruta())
                s.passo(1)
            ref = s.lerPopulacoes()
        finally:
            s.destruir()
        diff = pops[k] != ref
        if diff.any():
            cells = np.nonzero(diff.any(axis=0))[0]
            kinds = np.bincount(c["tipo"][cells], minlength=8)
            rel = np.abs(pops[k] - ref).max() / np.abs(ref).max()
            pytest.fail(f"tunnel {k}: {cells.size} cells differ (types {kinds.tolist()}), "
                        f"max rel {rel:.3g}")
        # same links, float64 sums in another order
        assert np.abs(res["forcas"][k] - np.array(forces)).max() <= 1e-10, k


def _init_pops(c):
    from cfd__amd.solver import SolverMI355X
    s = SolverMI355X.criar(c["grade"])
    try:
        s.definirTipos(c["tipo"])
        s.configurar(**c["cfg"])
        s.inicializar()
        return s.lerPopulacoes()
    finally:
        s.destruir()


def test_batch_admission_orders_completion_and_autotuner_restores_state():
    from cfd__amd.batch import LoteMI355X
    lote = LoteMI355X(CASES)
    try:
        tun = lote.autoajustar(passos=8, candidatos=[
            {"block": 256, "tile_cells": 512, "wg_per_cu": 0, "max_active_tiles": 20},
            {"block": 512, "tile_cells": 1024, "wg_per_cu": 0, "max_active_tiles": 20}])
        assert tun["melhor"] in [{k: v for k, v in m.items() if k != "ms"}
                                 for m in tun["medidas"]]
        after_tune = lote.populacoes(0)
        lote.rodar(60, 0, {"block": 256, "tile_cells": 512, "wg_per_cu": 0,
                           "max_active_tiles": 12})
        fin = lote.resultados()["fim_ms"]
    finally:
        lote.destruir()
    # the first tunnel is back at its initial state after tuning
    init = _init_pops(CASES[0])
    assert np.array_equal(after_tune, init)
    cells = [c["grade"]["nx"] * c["grade"]["ny"] * c["grade"]["nz"] for c in CASES]
    order = np.argsort(cells, kind="stable")
    # smallest first: completion times are non-decreasing in admission order
    assert all(fin[order[i]] <= fin[order[i + 1]] + 1e-9 for i in range(len(order) - 1))


def test_batch_runs_continue_and_accept_a_state():
    """Two launches of 3 and 4 steps equal one tunnel stepped 7 times from the
    same uploaded (perturbed) state, whatever the parity between launches."""
    from cfd__amd.batch import LoteMI355X
    from cfd__amd.solver import SolverMI355X
    c = CASES[1]
    g0 = _init_pops(c)
    rng = np.random.default_rng(5)
    g0 = (g0 * (1 + 0.01 * rng.standard_normal(g0.shape))).astype(np.float32)
    lote = LoteMI355X([CASES[0], c])
    try:
        lote.definirPopulacoes(1, g0)
        assert np.array_equal(lote.populacoes(1), g0)
        lote.rodar(3, 0, {"block": 512, "tile_cells": 2048})
        lote.rodar(4, 0, {"block": 256, "tile_cells": 1024})
        got = lote.populacoes(1)
    finally:
        lote.destruir()
    s = SolverMI355X.criar(c["grade"])
    try:
        s.definirTipos(c["tipo"])
        s.configurar(**c["cfg"])
        s.inicializar()
        s.definirPopulacoes(g0)
        s.passo(7)
        ref = s.lerPopulacoes()
    finally:
        s.destruir()
    fluid = ~np.isin(c["tipo"], [1, 4, 7])   # wall cells are never written
    assert np.array_equal(got[:, fluid], ref[:, fluid])
