"""A zeroed source (st_field_zero_source) is written lazily: the next sorted owner-flush
pass starts every brick's sums from zero and stores every cell it owns, and every other
use of the source writes the zeros first.  Each case runs on a field whose source holds
the sums of an earlier dense step (old values in nearly every cell), then zeroes it and
steps a sparse cloud (most bricks empty, most cells never deposited into), and checks
the source against the oracle started from zero (FP64: to 1e-12 of the largest |source|).
"""
import numpy as np
import pytest

import scaletrack as stk
from scaletrack import inputs
from oracle import st_oracle

pytestmark = pytest.mark.gpu

NCELL = 16


def dense_case():
    return inputs.taylor_green_box(np_parcels=20000, ncell=NCELL, boundary=0, seed=5)


def sparse_case():
    return inputs.taylor_green_box(np_parcels=300, ncell=NCELL, boundary=0, seed=9)


def make_cloud(ctx, case):
    c = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, stk.FP64)
    c.set_state(*case.state())
    c.upload()
    return c


def dirty_field(ctx):
    """A field whose source holds a dense step's sums (and whose stray sums are merged)."""
    d = dense_case()
    field = stk.TwoWayEulerian(ctx, stk.Mesh(d.n, d.origin, d.ending), stk.FP64)
    field.upload_U(d.U)
    dense = make_cloud(ctx, d)
    stk.evolve_device(ctx, dense, field, d.dt_fluid, d.nsub, stk.params(boundary=0, brick=(4, 4, 4)))
    old = field.download_source()
    assert np.count_nonzero(np.any(old != 0, axis=1)) > 0.9 * NCELL ** 3
    dense.close()
    return field


def oracle_source(case, state=None):
    src, _, st = st_oracle.evolve_case(case, state=state, precision=stk.FP64)
    return src, st


def assert_source(got, want):
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want).max() / scale
    assert err <= 1e-12, err


@pytest.mark.parametrize("p", [
    dict(brick=(2, 2, 2)),                          # most bricks empty: the taken zero
    dict(brick=(4, 4, 4)),
    dict(brick=(2, 2, 2), flush=stk.FLUSH_ATOMIC),  # zeros written before the launch
    dict(method=stk.PLAIN),
    dict(brick=(4, 4, 4), passes_per_launch=10, substeps_per_pass=1),
], ids=["sorted-2", "sorted-4", "atomic", "plain", "multi-pass"])
def test_zeroed_source_then_sparse_step(ctx, p):
    s = sparse_case()
    field = dirty_field(ctx)
    cloud = make_cloud(ctx, s)
    field.zero_source()
    stk.evolve_device(ctx, cloud, field, s.dt_fluid, s.nsub, stk.params(boundary=0, block=128, **p))
    want, _ = oracle_source(s)
    assert_source(field.download_source(), want)
    # a second zero and step: the cloud is sorted now, and the source holds this
    # cloud's own sums
    field.zero_source()
    stk.evolve_device(ctx, cloud, field, s.dt_fluid, s.nsub, stk.params(boundary=0, block=128, **p))
    want2, _ = oracle_source(s, state=oracle_source(s)[1])
    assert_source(field.download_source(), want2)
    cloud.close()
    field.close()


def test_zeroed_source_reads_as_zero(ctx):
    field = dirty_field(ctx)
    field.zero_source()
    assert not field.download_source().any()
    field.merge_source()
    assert not field.download_source().any()
    field.close()


def test_two_clouds_after_a_zero(ctx):
    """The first cloud's pass takes the zero; the second adds to the first's sums."""
    s, d = sparse_case(), dense_case()
    field = dirty_field(ctx)
    a, b = make_cloud(ctx, s), make_cloud(ctx, d)
    field.zero_source()
    p = stk.params(boundary=0, brick=(2, 2, 2), block=128)
    stk.evolve_device(ctx, a, field, s.dt_fluid, s.nsub, p)
    stk.evolve_device(ctx, b, field, d.dt_fluid, d.nsub, p)
    assert_source(field.download_source(), oracle_source(s)[0] + oracle_source(d)[0])
    a.close()
    b.close()
    field.close()


@pytest.mark.parametrize("zero_in_graph", [False, True])
def test_graph_after_a_zero(ctx, zero_in_graph):
    """A zero pending when a graph is captured or replayed is written before the replay:
    the replays add to zero (or zero the source themselves), and an eager step after a
    replay does not take a zero that the replay already consumed."""
    s = sparse_case()
    field = dirty_field(ctx)
    cloud = make_cloud(ctx, s)
    p = stk.params(boundary=0, brick=(2, 2, 2), block=128)
    field.zero_source()
    g = stk.Graph(ctx, cloud, field, s.dt_fluid, s.nsub, p, zero_source=zero_in_graph)
    st = None
    for _ in range(2):
        field.zero_source()
        g.launch()
        want, st = oracle_source(s, state=st)
        assert_source(field.download_source(), want)
    # replay, then an eager step on top of it without a zero: the sums add up
    field.zero_source()
    g.launch()
    want_g, st = oracle_source(s, state=st)
    stk.evolve_device(ctx, cloud, field, s.dt_fluid, s.nsub, p)
    want_e, st = oracle_source(s, state=st)
    assert_source(field.download_source(), want_g + want_e)
    g.close()
    cloud.close()
    field.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_after_a_zero(ctx, graph):
    """A batch of two sparse clouds on two fields that hold old sums: zeroed, the batch's
    first pass takes the zeros (eagerly, or inside a graph captured with zero_source),
    and each field's source equals the oracle's for its own cloud."""
    cases = [sparse_case(), inputs.taylor_green_box(np_parcels=500, ncell=NCELL, boundary=0, seed=13)]
    fields = [dirty_field(ctx) for _ in cases]
    clouds = [make_cloud(ctx, c) for c in cases]
    p = stk.params(boundary=0, substeps_per_pass=1)
    batch = stk.Batch(ctx, clouds, fields, p, target_parcels_per_brick=64)
    g = batch.graph(cases[0].dt_fluid, cases[0].nsub, zero_source=True) if graph else None
    states = [None, None]
    for _ in range(2):
        if graph:
            g.launch()  # the replay zeroes the sources itself: they still hold old sums here
        else:
            for f in fields:
                f.zero_source()
            batch.evolve(cases[0].dt_fluid, cases[0].nsub)
        for i, (c, f) in enumerate(zip(cases, fields)):
            want, states[i] = st_oracle.evolve_case(c, state=states[i], precision=stk.FP64, substeps_per_pass=1)[0::2]
            assert_source(f.download_source(), want)
    if g is not None:
        g.close()
    batch.close()
    for o in clouds + fields:
        o.close()
