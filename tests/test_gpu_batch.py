"""Ensembles (hv_batch_*): instance i of a batch equals a solo run of instance i,
bit for bit, in the same precision mode (solo runs with the per-sweep SOR
form, the one a batch uses)."""
import numpy as np
import pytest

import hvamd
from hvamd import configs
from hvamd.batch import Batch

pytestmark = pytest.mark.gpu


def _instances():
    """Three uneven instances: different grids, particle
    counts, Skyrme sets, boosts and boost conventions."""
    ws = configs.c3(3, seed=7)
    ws[0].grid, ws[0].dx = 20, 0.8
    ws[1].grid, ws[1].dx = 24, 0.7
    ws[2].grid, ws[2].dx = 16, 1.0
    for w in ws:
        w.ppn = 20 + (w.ppn % 17)
    params = [w.params(8) for w in ws]
    params[1].boost_mode = hvamd.HV_BOOST_REF_WEIGHTED
    params[2].t_exc = 1.5  # boost at step 3
    return ws, params


def _solo(w, prm, prec, chunks, sort_every, resident=False, owned=0):
    sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, prec)
    sim.set_option(hvamd.HV_OPT_SOR_RESIDENT_F32, 1 if resident else 0)
    sim.set_option(hvamd.HV_OPT_SOR_F32_OWNED, owned)
    sim.set_option(hvamd.HV_OPT_SORT_EVERY, sort_every)
    sim.upload_particles(*w.initial_state())
    it0 = sim.run_init(w.sky(), prm)
    stats, its, first = [], [], 0
    for c in chunks:
        s, it = sim.run_steps(w.sky(), prm, first, c)
        stats.append(s)
        its.append(it)
        first += c
    if resident:
        assert sim.get_option(hvamd.HV_OPT_SOR_RESIDENT_F32) == 1
        assert sim.get_option(hvamd.HV_OPT_SOR_F32_OWNED) == owned
    out = (it0, np.concatenate(stats), np.concatenate(its), sim.download_particles(),
           sim.total_energy(w.sky(), w.ppn))
    sim.close()
    return out


def _instances_res():
    """Grids with a compile-time tile body in the ensemble kernel (64^3) and
    without one (32^3 and 20^3 run the run-time-shape body)."""
    ws, params = _instances()
    ws[0].grid, ws[0].dx = 64, 0.35
    ws[2].grid, ws[2].dx = 32, 0.55
    return ws, params


@pytest.mark.parametrize("form", ["f64", "mixed_sweeps", "mixed_resident"])
def test_batch_equals_solo_runs(form):
    """F64; mixed with per-sweep FP32 corrections; mixed with the ensemble
    LDS-resident kernel (deep-halo lane-owned body) against solo LDS-resident
    solves (the body that derives node indexes per sweep)."""
    print(f"batch vs solo form={form}", flush=True)
    prec = hvamd.HV_PREC_F64 if form == "f64" else hvamd.HV_PREC_MIXED
    resident = form == "mixed_resident"
    ws, params = _instances_res() if resident else _instances()
    chunks = [5, 3]
    b = Batch([w.world() for w in ws], [w.n_protons for w in ws], [w.n_neutrons for w in ws], prec)
    b.set_option(hvamd.HV_OPT_SORT_EVERY, 2)
    b.set_option(hvamd.HV_OPT_SOR_RESIDENT_F32, 1 if resident else 0)
    b.profile_enable(True)
    for i, w in enumerate(ws):
        b.instance(i).upload_particles(*w.initial_state())
    sks = [w.sky() for w in ws]
    it0 = b.run_init(sks, params)
    stats, its, first = [], [], 0
    for c in chunks:
        s, it = b.run_steps(sks, params, first, c)
        stats.append(s)
        its.append(it)
        first += c
    stats = np.concatenate(stats, axis=1)
    its = np.concatenate(its, axis=1)
    parts = [b.instance(i).download_particles() for i in range(len(ws))]
    energies = [b.instance(i).total_energy(ws[i].sky(), ws[i].ppn) for i in range(len(ws))]
    res_launches = b.profile_read()["sor_resident_f32"][1]
    owned = b.get_option(hvamd.HV_OPT_SOR_F32_OWNED)
    b.close()
    assert res_launches == (sum(chunks) if resident else 0)
    if resident:
        assert owned == 2
    for i, w in enumerate(ws):
        s_it0, s_stats, s_its, s_part, s_e = _solo(w, params[i], prec, chunks, 2, resident)
        print(i, w.name, "iters", list(its[i]), "solo", list(s_its))
        assert it0[i] == s_it0
        assert np.array_equal(its[i], s_its)
        assert np.array_equal(stats[i], s_stats), i
        for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz"):
            assert np.array_equal(parts[i][k], s_part[k]), (i, k)
        assert np.array_equal(energies[i], s_e)


def test_batch_options_do_not_change_bits():
    """Sort cadence, smearing tile width (the autotuned knobs), the smearing
    form (fused; two-pass with the vector tile and staged halo coordinates;
    two-pass with narrow tiles) and the SOR instance grouping (1 KiB: every
    instance a group of its own) leave every bit of a batched run
    unchanged."""
    print("batch options", flush=True)
    ws, params = _instances()
    outs = []
    for sort_every, tile, group_kib, fused in ((0, 64, 0, 1), (3, 16, 1, 0), (4, 64, 0, 0)):
        b = Batch([w.world() for w in ws], [w.n_protons for w in ws], [w.n_neutrons for w in ws],
                  hvamd.HV_PREC_MIXED)
        b.set_option(hvamd.HV_OPT_SORT_EVERY, sort_every)
        b.set_option(hvamd.HV_OPT_SMEAR_TILE, tile)
        b.set_option(hvamd.HV_OPT_BATCH_GROUP_KIB, group_kib)
        b.set_option(hvamd.HV_OPT_SMEAR_FUSED, fused)
        for i, w in enumerate(ws):
            b.instance(i).upload_particles(*w.initial_state())
        sks = [w.sky() for w in ws]
        b.run_init(sks, params)
        s, it = b.run_steps(sks, params, 0, 6)
        outs.append((s, it, [b.instance(i).download_particles()["kz"] for i in range(len(ws))]))
        b.close()
    for o in outs[1:]:
        assert np.array_equal(outs[0][0], o[0])
        assert np.array_equal(outs[0][1], o[1])
        for a, c in zip(outs[0][2], o[2]):
            assert np.array_equal(a, c)


@pytest.mark.parametrize("prec", [hvamd.HV_PREC_F64, hvamd.HV_PREC_MIXED])
def test_batch_step_graph_replays_equal_eager_steps(prec):
    """Steps replayed from the captured step graphs (the batched sort inside
    the graph of every other step, a boost between replays, and in FP64 the
    host-steered SOR) give the bits of eagerly launched steps, step by
    step."""
    print(f"batch graph vs eager prec={prec}", flush=True)
    ws, params = _instances_res()
    outs = []
    for graph in (0, 1):
        b = Batch([w.world() for w in ws], [w.n_protons for w in ws], [w.n_neutrons for w in ws], prec)
        b.set_option(hvamd.HV_OPT_SORT_EVERY, 2)
        b.set_option(hvamd.HV_OPT_BATCH_GRAPH, graph)
        for i, w in enumerate(ws):
            b.instance(i).upload_particles(*w.initial_state())
        sks = [w.sky() for w in ws]
        b.run_init(sks, params)
        s1, i1 = b.run_steps(sks, params, 0, 6)
        s2, i2 = b.run_steps(sks, params, 6, 3)
        gi = b.graph_info()
        parts = [b.instance(i).download_particles() for i in range(len(ws))]
        b.close()
        print("graph", graph, gi, flush=True)
        assert gi["graph_steps"] == (5 + 2 if graph else 0)
        # per call one graph of the step that sorts first and one of the step
        # that does not (sorting every 2 steps)
        assert gi["captures"] == (4 if graph else 0)
        outs.append((s1, i1, s2, i2, parts))
    e, g = outs
    for k in range(4):
        assert np.array_equal(e[k], g[k]), k
    for pe, pg in zip(e[4], g[4]):
        for key in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz"):
            assert np.array_equal(pe[key], pg[key]), key


@pytest.mark.parametrize("owned", [0, 1, 2])
@pytest.mark.parametrize("big", [96, 32])
def test_batch_resident_shape_bodies_equal_solo(big, owned):
    """With the per-row body the ensemble kernel runs its compile-time tile
    body on the grid shape with the most nodes (16 x 12 rows x 48 slots on
    96^3, 32 x 16 x 16 on 32^3) and the run-time-shape body on the others;
    with the lane-owned (1) or deep-halo (2) body, that body on every tile:
    each against solo LDS-resident solves with another body, bit for bit."""
    print(f"batch resident shapes big={big} owned={owned}", flush=True)
    ws, params = _instances()
    ws = ws[:2]
    params = params[:2]
    ws[0].grid, ws[0].dx = big, 16.0 / big
    ws[1].grid, ws[1].dx = 20, 0.8
    b = Batch([w.world() for w in ws], [w.n_protons for w in ws], [w.n_neutrons for w in ws], hvamd.HV_PREC_MIXED)
    b.set_option(hvamd.HV_OPT_SORT_EVERY, 2)
    b.set_option(hvamd.HV_OPT_SOR_F32_OWNED, owned)
    for i, w in enumerate(ws):
        b.instance(i).upload_particles(*w.initial_state())
    sks = [w.sky() for w in ws]
    it0 = b.run_init(sks, params)
    s1, i1 = b.run_steps(sks, params, 0, 4)
    parts = [b.instance(i).download_particles() for i in range(len(ws))]
    assert b.get_option(hvamd.HV_OPT_SOR_F32_OWNED) == owned
    b.close()
    for i, w in enumerate(ws):
        s_it0, s_stats, s_its, s_part, _ = _solo(w, params[i], hvamd.HV_PREC_MIXED, [4], 2, resident=True,
                                                 owned=2 if owned == 0 else 0)
        print(i, w.grid, "iters", list(i1[i]), "solo", list(s_its), flush=True)
        assert it0[i] == s_it0
        assert np.array_equal(i1[i], s_its)
        assert np.array_equal(s1[i], s_stats)
        for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz"):
            assert np.array_equal(parts[i][k], s_part[k]), (i, k)


def test_batch_device_steered_step_equals_host_steered():
    """The mixed step with the FP64 finishing sweeps, the stop rule and the
    extrapolation history on the device (one cooperative launch, the whole
    step in the graph, read-backs every 16 steps) gives the bits of the step
    the host steers, across a re-plan and across two calls (the history
    handed back to the host and up again), with sorting and a boost."""
    print("batch device-steered vs host-steered", flush=True)
    ws, params = _instances_res()
    outs = []
    for steer in (0, 1):
        b = Batch([w.world() for w in ws], [w.n_protons for w in ws], [w.n_neutrons for w in ws],
                  hvamd.HV_PREC_MIXED)
        b.set_option(hvamd.HV_OPT_SORT_EVERY, 3)
        b.set_option(hvamd.HV_OPT_BATCH_DEVICE_STEER, steer)
        for i, w in enumerate(ws):
            b.instance(i).upload_particles(*w.initial_state())
        sks = [w.sky() for w in ws]
        b.run_init(sks, params)
        s1, i1 = b.run_steps(sks, params, 0, 19)
        s2, i2 = b.run_steps(sks, params, 19, 4)
        used = b.get_option(hvamd.HV_OPT_BATCH_DEVICE_STEER)
        counts = b.sor_counts()
        parts = [b.instance(i).download_particles() for i in range(len(ws))]
        b.close()
        print("steer", steer, "used", used, "iters", i1[:, -3:].tolist(), i2.tolist(), flush=True)
        assert used == steer
        outs.append((s1, i1, s2, i2, parts, counts["node_iters_f32"], counts["node_iters_f64"]))
    h, d = outs
    for k in range(4):
        assert np.array_equal(h[k], d[k]), k
    for ph, pd in zip(h[4], d[4]):
        for key in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz"):
            assert np.array_equal(ph[key], pd[key]), key
    assert h[5] == d[5] and h[6] == d[6]
