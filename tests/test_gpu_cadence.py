"""axs_rk4_advance_windings: RK4 steps with the windings / tagger pass at a
step cadence, the segments replayed from one captured graph.  Every
segment's counts, the final tags and the final state equal the same steps
and passes launched directly (axs_rk4_advance + axs_count_plaquettes /
axs_tag_cells with R at that time), for every stage fusion mode, odd and
even segment lengths, with eager work between calls, and on a cached graph
reused by a second call; the FP64 counts equal the oracle's."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _level(capi, ctx, n, precision, fusion=None):
    lvl = capi.Level(ctx, n, 0.5, precision)
    lvl.fourier_ic(0, n / 8.0, 0.5, 42, 0)
    lvl.fourier_ic(1, n / 8.0, 0.5, 42, 1)
    lvl.zero_components(2, 2)
    if fusion is not None:
        lvl.set_fusion(fusion)
    return lvl


def _direct(capi, lvl, bg, tau, dt, nseg, every, thr):
    """The reference sequence, one launch at a time."""
    counts, tags = [], None
    t = tau
    for _ in range(nseg):
        lvl.advance(bg, t, dt, every, use_graph=False)
        for _ in range(every):
            t += dt
        counts.append(lvl.count_plaquettes())
        if thr is not None:
            tags, c2 = lvl.tag(grad_thr=thr[0], amp_thr=thr[1],
                               R=capi.background_eval(bg, 0, t))
            assert c2 == counts[-1]
    return np.array(counts, np.int64), tags


@pytest.mark.parametrize("precision,fusion", [(0, 1), (0, 0), (1, 3), (1, 2),
                                              (1, 0)])
@pytest.mark.parametrize("every", [1, 2, 3])
def test_cadence_graph_equals_direct(ctx, precision, fusion, every):
    from axion_strings_adaptive_amd import capi
    n, dt, nseg = 64, 0.1, 4
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.0,
                              has_switch=True)
    thr = (0.02, 0.5)  # Laplacian criterion and the amplitude extension
    ref = _level(capi, ctx, n, precision, fusion)
    want, want_tags = _direct(capi, ref, bg, 1.0, dt, nseg, every, thr)
    assert want[:, 0].min() > 0
    for use_graph in (True, False):
        lvl = _level(capi, ctx, n, precision, fusion)
        tags = capi.DeviceBuffer(n ** 3)
        got = lvl.advance_windings(bg, 1.0, dt, nseg * every, every, tags,
                                   grad_thr=thr[0], amp_thr=thr[1],
                                   use_graph=use_graph)
        assert np.array_equal(got, want), (use_graph, got, want)
        t = tags.to_numpy(np.uint8).reshape(n, n, n)
        assert np.array_equal(t, want_tags)
        assert np.array_equal(lvl.download(), ref.download())
        lvl.close()
    ref.close()


def test_cadence_graph_reused_across_calls(ctx, oracle):
    """The second call replays the graph the first captured; eager work
    (an energy pass, a count) between calls; counts only (no tags);
    FP64 counts against the oracle's windings."""
    from axion_strings_adaptive_amd import capi
    n, dt, every = 32, 0.1, 5
    bg = capi.make_background(2.0)
    lvl = _level(capi, ctx, n, capi.AXS_FP64)
    s = lvl.download()
    a = lvl.advance_windings(bg, 1.0, dt, 3 * every, every, use_graph=True)
    e1 = lvl.energy_sums(1.0, 1.0, 0.5)
    lvl.count_plaquettes()
    t = 1.0
    for _ in range(3 * every):  # the tau sequence of the steps: += dt each
        t += dt
    b = lvl.advance_windings(bg, t, dt, 2 * every, every, use_graph=True)
    got = np.concatenate([a, b])
    tau = 1.0
    for j in range(5):
        tau = oracle.advance(s, 0.5, oracle.bg(2.0), tau, dt, every)
        assert tuple(got[j]) == oracle.count_plaquettes(s), j
    assert np.array_equal(lvl.download(), s)
    assert np.all(np.isfinite(e1))
    with pytest.raises(capi.AxsError, match="multiple of every"):
        lvl.advance_windings(bg, 1.0, dt, 7, every, use_graph=True)
    lvl.close()
