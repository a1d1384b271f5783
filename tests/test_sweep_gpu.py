"""Batched sweeps: persistent launches over uneven instances, hipGraph replay,
autotuned launch shapes -- against single propagations and the CPU checker."""

import numpy as np
import pytest

from helpers import flat, group_err
from oracle import bind as OB
from pypo_amd import sweep as SW

pytestmark = pytest.mark.gpu


def small_sweep(**kw):
    return SW.frequency_offset_sweep(n_freq=3, f_lo=90.0, f_hi=110.0, offsets=(0.0, 3.0),
                                     diameter=120.0, focal=48.0, n_lo=31, n_hi=47, n_feed=9,
                                     n_ff=21, ff_deg=6.0, **kw)


@pytest.fixture(scope="module")
def small(hip):
    feeds, dishes, ks, curs, ff, meta = small_sweep()
    s = SW.Sweep(feeds, dishes, ks, curs, ff, use_graph=False, tpl=2, chunk_tiles=2)
    yield s, (feeds, dishes, ks, curs, ff, meta)
    s.close()


def test_sweep_matches_cpu_checker(small, port):
    s, (feeds, dishes, ks, curs, ff, meta) = small
    s.run()
    for i in range(s.n_inst):
        jm_ref, _, _ = OB.propagate(port, "JM", feeds[i], dishes[i], curs[i], ks[i])
        e_jm = group_err(flat(s.read_currents(i)), flat(jm_ref))
        ff_ref, _, _ = OB.propagate(port, "FF", dishes[i], ff, jm_ref, ks[i])
        e_ff = group_err(flat(s.read(i)), flat(ff_ref))
        print(f"instance {i} {meta[i]}: JM {e_jm:.2e} FF {e_ff:.2e}")
        assert e_jm <= 1e-5 and e_ff <= 1e-5


def test_sweep_matches_single_propagations(small):
    from pypo_amd import device as D
    s, (feeds, dishes, ks, curs, ff, meta) = small
    s.run()
    g_ff = D.generate_grid(ff)
    for i in (0, s.n_inst - 1):
        g_feed, g_dish = D.generate_grid(feeds[i]), D.generate_grid(dishes[i])
        jm = D.propagate("JM", g_feed, g_dish, D.currents_tensor(curs[i]), feeds[i], ks[i])
        e = D.propagate("FF", g_dish, g_ff, jm, dishes[i], ks[i])
        assert group_err(flat(s.read_currents(i)), D.tensor_to_complex6(jm)) < 1e-9
        assert group_err(flat(s.read(i)), D.tensor_to_complex6(e)) < 1e-6


@pytest.mark.parametrize("tpl,chunk,wgs", [(1, 1, 1), (2, 1, 3), (1, 4096, 2), (4, 2, 1), (4, 4096, 1),
                                          (2, 4096, 64), (2, 1, 4096), (4, 3, 64)])
def test_launch_shapes_agree(small, tpl, chunk, wgs):
    s, _ = small
    s.plan(tpl=2, chunk_tiles=2, wgs_per_cu=2, use_graph=False)
    s.run()
    ref = [flat(s.read(i)) for i in range(s.n_inst)]
    s.plan(tpl=tpl, chunk_tiles=chunk, wgs_per_cu=wgs, use_graph=False)
    s.run()
    for i in range(s.n_inst):
        assert group_err(flat(s.read(i)), ref[i]) < 1e-6
    s.plan(tpl=2, chunk_tiles=2, wgs_per_cu=2, use_graph=False)


@pytest.mark.parametrize("tpl", [1, 2])
def test_vector_kernel_sweep_matches_checker(small, port, tpl):
    """The far-field stage on the vector ALUs (k_sweep_ff, selected with
    pypo_hip_set_ff_kernel) against the CPU checker and the matrix-core stage."""
    from pypo_amd import bind as B
    s, (feeds, dishes, ks, curs, ff, meta) = small
    s.plan(tpl=2, chunk_tiles=2, wgs_per_cu=2, use_graph=False)
    s.run()
    matrix = [flat(s.read(i)) for i in range(s.n_inst)]
    B.set_ff_kernel("vector")
    try:
        s.plan(tpl=tpl, chunk_tiles=2, wgs_per_cu=3, use_graph=False)
        s.run()
        for i in range(s.n_inst):
            jm_ref, _, _ = OB.propagate(port, "JM", feeds[i], dishes[i], curs[i], ks[i])
            ff_ref, _, _ = OB.propagate(port, "FF", dishes[i], ff, jm_ref, ks[i])
            got = flat(s.read(i))
            e_ref, e_mat = group_err(got, flat(ff_ref)), group_err(got, matrix[i])
            print(f"vector sweep tpl {tpl} instance {i}: vs checker {e_ref:.2e}, vs matrix {e_mat:.2e}")
            assert e_ref <= 1e-5 and e_mat <= 1e-5
    finally:
        B.set_ff_kernel("matrix")
        s.plan(tpl=2, chunk_tiles=2, wgs_per_cu=2, use_graph=False)


def test_graph_replays_match_eager(small):
    import torch
    s, _ = small
    s.plan(tpl=2, chunk_tiles=2, wgs_per_cu=2, use_graph=False)
    s.run()
    eager = [flat(s.read(i)) for i in range(s.n_inst)]
    s.plan(tpl=2, chunk_tiles=2, wgs_per_cu=2, use_graph=True)
    for rep in range(3):
        # unrelated eager work and synchronisation between replays
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()
        s.run()
        torch.cuda.synchronize()
        for i in range(s.n_inst):
            np.testing.assert_array_equal(flat(s.read(i)), eager[i])
    s.plan(tpl=2, chunk_tiles=2, wgs_per_cu=2, use_graph=False)


def test_autotune_picks_and_caches(small, tmp_path, monkeypatch):
    from pypo_amd import autotune as AT
    s, _ = small
    monkeypatch.setenv("PYPO_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    cands = [{"tpl": 1, "chunk_tiles": 1, "wgs_per_cu": 1}, {"tpl": 2, "chunk_tiles": 4096, "wgs_per_cu": 2}]
    best, secs, table = AT.tune(s, candidates=cands, use_graph=True)
    assert best in cands and len(table) == 2 and secs > 0
    assert (tmp_path / "tune.json").exists()
    best2, _, _ = AT.tune(s, candidates=cands, use_graph=False)
    assert best2 == best
    s.plan(tpl=2, chunk_tiles=2, wgs_per_cu=2, use_graph=False)


def test_bad_sweep_arguments_raise(hip):
    from pypo_amd import bind as B
    feeds, dishes, ks, curs, ff, _ = small_sweep()
    with pytest.raises(B.PyPOHIPError):
        SW.Sweep(feeds[:2], dishes[:2], [1.0, -1.0], curs[:2], ff)
    with pytest.raises(B.PyPOHIPError):
        SW.Sweep(feeds[:1], [dict(dishes[0], rms=0.1)], ks[:1], curs[:1], ff)


def test_full_size_sweep(hip, port):
    """The sweep at full size: 64 frequencies 80-120 GHz x feed offsets
    (0, +2, -2, +4) mm = 256 instances, dish uv grids 201^2 .. 401^2, 25^2
    feeds, one 181^2 AoE far field over +-1 deg, one hipGraph replay.  Three
    instances across the size range against single device propagations (JM
    to 1e-9, FF to 1e-6 of peak), and the largest instance (120 GHz, +4 mm,
    401^2) against the CPU checker within the mixed bar: JM on every dish
    point, the far field on every 97th target plus the beam centre."""
    from pypo_amd import device as D
    feeds, dishes, ks, curs, ff, meta = SW.frequency_offset_sweep()
    s = SW.Sweep(feeds, dishes, ks, curs, ff, tpl=2, chunk_tiles=4096, wgs_per_cu=4, use_graph=True)
    try:
        assert s.n_inst == 256 and meta[-1]["n_dish"] == 401 and meta[0]["n_dish"] == 201
        s.run()
        g_ff = D.generate_grid(ff)
        for i in (0, 130, 255):
            g_feed, g_dish = D.generate_grid(feeds[i]), D.generate_grid(dishes[i])
            jm = D.propagate("JM", g_feed, g_dish, D.currents_tensor(curs[i]), feeds[i], ks[i])
            e = D.propagate("FF", g_dish, g_ff, jm, dishes[i], ks[i])
            e_jm = group_err(flat(s.read_currents(i)), D.tensor_to_complex6(jm))
            e_ff = group_err(flat(s.read(i)), D.tensor_to_complex6(e))
            print(f"instance {i} {meta[i]}: vs single JM {e_jm:.2e} FF {e_ff:.2e}")
            assert e_jm < 1e-9 and e_ff < 1e-6
        i = 255
        jm_ref, _, _ = OB.propagate(port, "JM", feeds[i], dishes[i], curs[i], ks[i])
        e_jm = group_err(flat(s.read_currents(i)), flat(jm_ref))
        n_ff = 181 * 181
        idx = np.union1d(np.arange(0, n_ff, 97), [90 * 181 + 90])
        ff_ref, _ = OB.farfield_subset(port, dishes[i], ff, jm_ref, ks[i], idx)
        e_ff = group_err(flat(s.read(i))[:, idx], ff_ref)
        print(f"instance {i} vs CPU checker: JM {e_jm:.2e}, FF ({idx.size} targets) {e_ff:.2e}")
        assert e_jm <= 1e-5 and e_ff <= 1e-5
    finally:
        s.close()


@pytest.mark.parametrize("seed", range(8))
def test_random_sweeps_against_checker(hip, port, seed):
    """Random sweeps: 1-9 instances, feeds of 3-19 points a side at the focus
    of uv or xy paraboloids of 8-69 points a side (some under a scaling
    transform), random wavelengths, a shared far field of odd grid sizes
    (so the main beam is sampled) over a random cone, random launch shapes,
    graph or eager.  Every instance's JM and FF within 1e-5 of peak of the
    checker."""
    from pypo_amd import geometry as G
    from pypo_amd import types as T
    rng = np.random.default_rng(7100 + seed)
    feeds, dishes, ks, curs = [], [], [], []
    for _ in range(int(rng.integers(1, 10))):
        nf = [int(rng.integers(3, 20)), int(rng.integers(3, 20))]
        feed = G.plane({"gmode": "xy", "lims_x": [-4, 4], "lims_y": [-4, 4], "gridsize": nf, "flip": True})
        f = float(rng.uniform(30, 60))
        feed["transf"] = G.mat_translate([rng.uniform(-3, 3), rng.uniform(-3, 3), f])
        nd = [int(rng.integers(8, 70)), int(rng.integers(8, 70))]
        lims = {"gmode": "uv", "lims_u": [0, 40], "lims_v": [0, 360]} if rng.random() < 0.5 else \
            {"gmode": "xy", "lims_x": [-30, 30], "lims_y": [-30, 30]}
        dish = G.paraboloid(dict(lims, pmode="focus", vertex=[0, 0, 0], focus_1=[0, 0, f], gridsize=nd))
        if rng.random() < 0.3:
            m = np.eye(4)
            m[:3, :3] *= rng.uniform(0.7, 1.5)
            dish["transf"] = m @ dish["transf"]
        x, y = np.linspace(-4, 4, nf[0])[:, None], np.linspace(-4, 4, nf[1])[None, :]
        amp = np.exp(-(x ** 2 + y ** 2) / rng.uniform(2, 8)) * np.exp(1j * rng.uniform(0, 0.3) * x)
        z = np.zeros(nf, complex)
        feeds.append(feed)
        dishes.append(dish)
        ks.append(2 * np.pi / float(rng.uniform(0.8, 4.0)))
        curs.append(T.currents(-amp + 0.1 * rng.normal(size=nf), z, z, z, amp, z))
    deg = float(rng.uniform(0.5, 30))
    ff = G.plane({"gmode": "AoE", "lims_Az": [-deg, deg], "lims_El": [-deg, deg],
                  "gridsize": [2 * int(rng.integers(1, 20)) + 1, 2 * int(rng.integers(1, 20)) + 1]})
    shape = dict(tpl=int(rng.choice([1, 2, 4])), chunk_tiles=int(rng.choice([1, 2, 5, 4096])),
                 wgs_per_cu=int(rng.choice([1, 3, 64, 4096])), use_graph=bool(rng.random() < 0.5))
    s = SW.Sweep(feeds, dishes, ks, curs, ff, **shape)
    try:
        s.run()
        for i in range(len(feeds)):
            jm, _, _ = OB.propagate(port, "JM", feeds[i], dishes[i], curs[i], ks[i])
            e, _, _ = OB.propagate(port, "FF", dishes[i], ff, jm, ks[i])
            ej, ef = group_err(flat(s.read_currents(i)), flat(jm)), group_err(flat(s.read(i)), flat(e))
            print(f"seed {seed} instance {i} {shape}: JM {ej:.2e} FF {ef:.2e}")
            assert ej <= 1e-5 and ef <= 1e-5
    finally:
        s.close()
