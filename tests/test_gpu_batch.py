"""GPU tests of fused modelling batches (ddx_model_batch): shots of contexts
with uneven grids run in one launch sequence, through the C ABI.

Bar: bit-identical seismograms to each context's own ddx_model_shot (the
fused launches apply the same operations to the same cells; only the block
to XCD assignment differs), which the single-shot parity tests tie to the
oracle at relative L2 <= 1e-5.
"""
import numpy as np
import pytest

import ddx
import pyoracle as po
from cases import case, oracle_for, oracle_rows
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def engine_for(c, divide=(1, 1, 1), storage="fp32", ring_storage="fp32"):
    d = ddx.make_desc(c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["dt"], c["fm"], c["ntime"], divide=divide,
                      rec_width=c["rec_width"], cfl_policy=1, storage=storage, ring_storage=ring_storage)
    e = ddx.Engine(d)
    e.set_model(c["vp"], c["vs"], c["rho"])
    e.set_wavelet(ddx.ricker(c["fm"], c["dt"], c["ntime"]))
    e.set_geometry(c["shots"], c["recv"])
    return e


def uneven_cases(ntime=60):
    """Three shots: different grids, CPML widths, wavelets and models."""
    return [(case(nx=30, ny=26, nz=22, pml=6, fm=25.0, ntime=ntime, seed=11), (1, 1, 1)),
            (case(nx=41, ny=20, nz=24, pml=8, fm=20.0, ntime=ntime, seed=12), (2, 1, 1)),
            (case(nx=24, ny=33, nz=18, pml=7, fm=30.0, ntime=ntime, seed=13), (1, 1, 1))]


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_batch_matches_single_shots(storage):
    engines, ref = [], []
    for c, divide in uneven_cases():
        e = engine_for(c, divide, storage)
        e.model_shot(0)
        ref.append(e.seismograms(0))
        engines.append(e)
    ddx.model_batch(engines)
    for q, e in enumerate(engines):
        assert np.array_equal(e.seismograms(0), ref[q]), q
    # a second batch in another order gives the same seismograms again
    ddx.model_batch(engines[::-1])
    for q, e in enumerate(engines):
        assert np.array_equal(e.seismograms(0), ref[q]), q
    for e in engines:
        e.close()


def test_batch_member_against_oracle():
    """One member of a batch against the FP32 oracle directly."""
    cases = uneven_cases()
    engines = [engine_for(c, divide) for c, divide in cases]
    ddx.model_batch(engines)
    c, divide = cases[1]
    rows, _ = oracle_rows(c, divide)
    o = oracle_for(c)
    ref = o.model_shot(o.node(*c["shots"][0]), po.ricker(c["fm"], c["dt"], c["ntime"]), [n for _, _, n in rows])
    assert rel_l2(engines[1].seismograms(0), ref) <= 1e-5
    for e in engines:
        e.close()


def test_batch_counts_cell_updates_and_rejects_mismatches():
    cases = uneven_cases(ntime=20)
    engines = [engine_for(c, divide) for c, divide in cases]
    before = [e.cell_updates() for e in engines]
    ddx.model_batch(engines)
    for (c, _), e, b in zip(cases, engines, before):
        ext = (c["nx"] + 2 * c["pml"]) * (c["ny"] + 2 * c["pml"]) * (c["nz"] + 2 * c["pml"])
        assert e.cell_updates() - b == ext * c["ntime"]
    with pytest.raises(ddx.DdxError, match="twice"):
        ddx.model_batch([engines[0], engines[0]])
    other = case(nx=20, ny=20, nz=20, ntime=30)  # a different ntime
    f = engine_for(other)
    with pytest.raises(ddx.DdxError, match="equal dt"):
        ddx.model_batch([engines[0], f])
    g = engine_for(case(nx=20, ny=20, nz=20, ntime=20), storage="bf16")
    with pytest.raises(ddx.DdxError, match="storage"):
        ddx.model_batch([engines[0], g])
    with pytest.raises(ValueError):
        ddx.model_batch(engines, ishots=[0])
    for e in engines + [f, g]:
        e.close()


def test_continuous_batch_members_join_and_leave():
    """Members of different length and grid join at different steps
    (one after 25 steps of the others) and leave when done: each one's
    seismograms equal its own model_shot bit for bit."""
    cs = [(case(nx=30, ny=26, nz=22, pml=6, fm=25.0, ntime=60, seed=11), (1, 1, 1)),
          (case(nx=41, ny=20, nz=24, pml=8, fm=20.0, ntime=40, seed=12), (2, 1, 1)),
          (case(nx=24, ny=33, nz=18, pml=7, fm=30.0, ntime=50, seed=13), (1, 1, 1))]
    engines, ref = [], []
    for c, divide in cs:
        e = engine_for(c, divide)
        e.model_shot(0)
        ref.append(e.seismograms(0))
        engines.append(e)
    b = ddx.Batch()
    b.add(engines[0])
    b.add(engines[1])
    assert b.run(25) == [] and len(b) == 2
    b.add(engines[2])
    finished = []
    steps = 0
    while len(b):
        out = b.run(7)
        finished += out
        steps += 1
        assert steps < 100
    # shot 1 (40 steps) ends first, at step 40; shot 0 at 60; shot 2 (joined at 25) at 75
    assert finished == [engines[1], engines[0], engines[2]]
    for q, e in enumerate(engines):
        assert np.array_equal(e.seismograms(0), ref[q]), q
    # the batch is reusable once empty
    b.add(engines[1])
    while len(b):
        b.run(100)
    assert np.array_equal(engines[1].seismograms(0), ref[1])
    b.close()
    for e in engines:
        e.close()


def test_continuous_batch_rejects_mismatches():
    b = ddx.Batch()
    e = engine_for(case(nx=20, ny=20, nz=20, ntime=20))
    b.add(e)
    with pytest.raises(ddx.DdxError, match="already"):
        b.add(e)
    f = engine_for(case(nx=20, ny=20, nz=20, ntime=20), storage="bf16")
    with pytest.raises(ddx.DdxError, match="storage"):
        b.add(f)
    assert b.run(100) == [e]
    b.close()
    e.close()
    f.close()


def test_failed_first_add_does_not_key_the_batch():
    """A context that fails to join (not ready) sets nothing: a context of
    another storage format can still be the first member."""
    c = case(nx=20, ny=20, nz=20, ntime=20)
    unready = ddx.Engine(ddx.make_desc(c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["dt"], c["fm"], c["ntime"],
                                       cfl_policy=1))
    unready.set_model(c["vp"], c["vs"], c["rho"])  # no wavelet, no geometry
    b = ddx.Batch()
    with pytest.raises(ddx.DdxError, match="geometry"):
        b.add(unready)
    assert len(b) == 0
    e = engine_for(c, storage="bf16")
    e.model_shot(0)
    ref = e.seismograms(0)
    b.add(e)
    assert b.run(100) == [e]
    assert np.array_equal(e.seismograms(0), ref)
    b.close()
    e.close()
    unready.close()


def test_batch_with_lds_tile_bricks():
    """The fused launches take the first member's brick shapes; with the LDS
    y-tile bricks (form 4: 512-thread blocks, an LDS tile per block) every
    member's seismograms still equal its default-brick single shot."""
    engines, ref = [], []
    for c, divide in uneven_cases():
        e = engine_for(c, divide)
        e.model_shot(0)
        ref.append(e.seismograms(0))
        engines.append(e)
    for k in (0, 1):
        engines[0].set_brick(k, po.PURE_FORWARD, 64, 8, 1, 2, px=4)
    ddx.model_batch(engines)
    for q, e in enumerate(engines):
        assert np.array_equal(e.seismograms(0), ref[q]), q
    b = ddx.Batch()
    for e in engines:
        b.add(e)
    while len(b):
        b.run(25)
    for q, e in enumerate(engines):
        assert np.array_equal(e.seismograms(0), ref[q]), q
    b.close()
    for e in engines:
        e.close()


def test_graph_replayed_batches_equal_single_shots():
    """ddx_set_graphs on the first member: the fused batch and a continuous
    batch (a member joining late and one removed early) replay one captured
    fused step per time step; every finished member's seismograms equal its
    own model_shot bit for bit."""
    engines, ref = [], []
    for c, divide in uneven_cases():
        e = engine_for(c, divide)
        e.model_shot(0)
        ref.append(e.seismograms(0))
        engines.append(e)
    engines[0].set_graphs(True)
    ddx.model_batch(engines)
    for q, e in enumerate(engines):
        assert np.array_equal(e.seismograms(0), ref[q]), q
    b = ddx.Batch()
    b.add(engines[0])
    b.add(engines[2])
    extra = engine_for(case(nx=26, ny=22, nz=20, ntime=60, seed=14))
    b.add(extra)
    assert b.run(11) == []
    b.remove(extra)  # leaves before its shot ends; the others run on
    b.add(engines[1])  # joins at step 11
    finished = []
    while len(b):
        finished += b.run(9)
    assert sorted(map(id, finished)) == sorted(map(id, engines))
    for q, e in enumerate(engines):
        assert np.array_equal(e.seismograms(0), ref[q]), q
    b.close()
    for e in engines + [extra]:
        e.close()


@pytest.mark.parametrize("storage,ring", [("fp32", "fp32"), ("bf16", "fp32"), ("fp32", "bfp16")])
def test_rtm_batch_matches_single_shots(storage, ring):
    """ddx_rtm_batch over three RTM shots of uneven grids, ring widths and
    decompositions (one member with a smoothed migration model): each
    member's nine images equal its own ddx_rtm_shot's bit for bit, and one
    member's against the FP32 oracle (FP32 storage; with the block-scaled
    store, the oracle's mirror of it)."""
    cs = [(case(nx=30, ny=26, nz=22, pml=6, fm=25.0, ntime=50, rec_width=1, seed=21), (1, 1, 1)),
          (case(nx=41, ny=20, nz=24, pml=8, fm=20.0, ntime=50, rec_width=2, seed=22), (2, 1, 1)),
          (case(nx=24, ny=33, nz=18, pml=7, fm=30.0, ntime=50, rec_width=4, seed=23), (1, 2, 1))]

    def member(c, divide):
        e = engine_for(c, divide, storage, ring)
        if c["rec_width"] == 4:
            e.set_migration_model(*[ddx.smooth_model(a, (2, 2, 2)) for a in (c["vp"], c["vs"], c["rho"])])
        e.model_shot(0)  # observed data from the modelling shot
        return e

    ref, single = [], []
    for c, divide in cs:
        e = member(c, divide)
        e.rtm_shot(0)
        e.finalize_images()
        ref.append([[e.image(q, r) for q in range(9)] for r in range(e.nsub)])
        single.append(e)
    batch = [member(c, divide) for c, divide in cs]
    ddx.rtm_batch(batch)
    for e in batch:
        e.finalize_images()
    for q, e in enumerate(batch):
        for r in range(e.nsub):
            for w in range(9):
                assert np.array_equal(e.image(w, r), ref[q][r][w]), (q, r, w)
    if storage == "fp32":
        c, divide = cs[0]
        rows, subs = oracle_rows(c, divide)
        o = oracle_for(c)
        if ring == "bfp16":
            o.set_ring_bfp(subs)
        obs = batch[0].seismograms(0)
        o.rtm_shot(o.node(*c["shots"][0]), po.ricker(c["fm"], c["dt"], c["ntime"]), [n for _, _, n in rows], obs)
        img, _ = o.finalize_box(0, 0, 0, c["nx"], c["ny"], c["nz"])
        assert rel_l2(batch[0].image(ddx.IPP, 0), img[0]) <= 1e-5
    for e in single + batch:
        e.close()


@pytest.mark.parametrize("ntimes", [(40, 40), (55, 40), (40, 55)])
def test_continuous_batch_rtm_members_join_at_any_step(ntimes):
    """RTM shots in a continuous batch (ddx_batch_add_rtm): two RTM members
    that join 7 steps apart, beside a modelling member, each run their
    forward pass and reverse loop at their own time index; every RTM
    member's nine images equal its own ddx_rtm_shot's and the modelling
    member's seismograms its ddx_model_shot's, bit for bit.  The members
    may differ in ntime: each saves its boundary store for its own time
    slices 0 .. ntime - 2, whichever member the fused tables were built
    from (the first, here the longer or the shorter one)."""
    cs = [(case(nx=30, ny=26, nz=22, pml=6, fm=25.0, ntime=ntimes[0], rec_width=1, seed=31), (1, 1, 1)),
          (case(nx=41, ny=20, nz=24, pml=8, fm=20.0, ntime=ntimes[1], rec_width=2, seed=32), (2, 1, 1))]
    mc = case(nx=24, ny=33, nz=18, pml=7, fm=30.0, ntime=40, seed=33)

    def member(c, divide):
        e = engine_for(c, divide)
        e.model_shot(0)  # observed data
        return e

    ref = []
    for c, divide in cs:
        e = member(c, divide)
        e.rtm_shot(0)
        e.finalize_images()
        ref.append([[e.image(q, r) for q in range(9)] for r in range(e.nsub)])
        e.close()
    m = engine_for(mc, (1, 2, 1))
    m.model_shot(0)
    mref = m.seismograms(0)
    rtm = [member(c, divide) for c, divide in cs]
    b = ddx.Batch()
    b.add(rtm[0], 0, rtm=True)
    b.add(m, 0)
    done = b.run(7)
    assert done == [] and len(b) == 2
    b.add(rtm[1], 0, rtm=True)
    finished = []
    while len(b):
        finished += b.run(1000)
    assert {id(e) for e in finished} == {id(e) for e in rtm + [m]}
    assert np.array_equal(m.seismograms(0), mref)
    for q, e in enumerate(rtm):
        e.finalize_images()
        for r in range(e.nsub):
            for w in range(9):
                assert np.array_equal(e.image(w, r), ref[q][r][w]), (q, r, w)
    b.close()
    for e in rtm + [m]:
        e.close()
