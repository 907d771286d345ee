"""Batched aberration search (phase-image-variance objective) through the C
ABI: against the FP64 oracle, independent of batch composition, graph
replay bit-identical to direct launches (also after new trials are set),
several datasets of uneven scan size in one launch equal to each alone."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import ssb_oracle as O
from tests import _cases as C
from tests._gpu import engine_for_data

pytestmark = pytest.mark.gpu


def _prepared_engine(cuda, data, radius=3.0, **kw):
    eng = engine_for_data(cuda, data, semiangle_mrad=kw.get("semi", 21.4), det_sampling=kw.get("ds", 1.0),
                          step=kw.get("step", 0.5))
    d = cuda.from_numpy(data).cuda()
    eng.mean_dp(d)
    eng.select_bf(0.0, radius)
    eng.prepare_g(d)
    return eng


def _trials(n, seed=0):
    rng = np.random.default_rng(seed)
    mags, angs = np.zeros((n, 14)), np.zeros((n, 14))
    mags[:, 0] = rng.uniform(-300.0, 300.0, n)
    mags[:, 1] = rng.uniform(0.0, 60.0, n)
    angs[:, 1] = rng.uniform(-1.5, 1.5, n)
    mags[:, 4] = rng.uniform(-2.0e5, 2.0e5, n)
    mags[1, 1] = 0.0                                   # a radial (C10 + C30) trial
    rot = np.radians(rng.uniform(-5.0, 5.0, n))
    return mags, angs, rot


def test_phase_variance_matches_fp64_oracle(cuda):
    from quantem.gpu_amd.ssb.compute.hip.trials import TrialBatch
    data = C.poisson_data(124)
    eng = _prepared_engine(cuda, data)
    mags, angs, rot = _trials(11)
    tb = TrialBatch([eng], 16)
    tb.set(mags, angs, rot)
    got = tb.phase_variance(use_graph=False).cpu().numpy()[0]
    p = O.prepare(data, bf_radius=3, **C.POISSON_KW)
    for t in range(11):
        geom = O.detector_geometry(p.bf.rows, p.bf.cols, p.bf.center, (16, 16), (1.0, 1.0), p.geom.wavelength,
                                   21.4, rot[t], 128, 0.5)
        ref = O.phase_variance(p.G, p.dc, geom, mags[t], angs[t])
        assert got[t] == pytest.approx(ref, rel=1e-4), t
    tb.close()


def test_batch_composition_graph_replay_and_new_trials(cuda):
    from quantem.gpu_amd.ssb.compute.hip.trials import TrialBatch
    torch = cuda
    eng = _prepared_engine(cuda, C.poisson_data(125))
    mags, angs, rot = _trials(21, seed=1)
    tb = TrialBatch([eng], 24)
    tb.set(mags, angs, rot)
    direct = tb.phase_variance(use_graph=False).clone()
    graph = [tb.phase_variance(use_graph=True).clone() for _ in range(3)]
    torch.cuda.synchronize()
    for g in graph:
        assert torch.equal(g, direct)
    # every trial alone gives its own bits (blocks of trials share G loads only)
    for t in (0, 7, 20):
        tb.set(mags[t:t + 1], angs[t:t + 1], rot[t:t + 1])
        one = tb.phase_variance(use_graph=False)
        assert torch.equal(one[0, 0], direct[0, t])
    # new trials of the same count replay the captured graph; eager work in between
    tb.set(mags, angs, rot)
    tb.phase_variance(use_graph=True)
    m2, a2, r2 = _trials(21, seed=2)
    tb.set(m2, a2, r2)
    again = tb.phase_variance(use_graph=True).clone()
    _ = torch.ones(1 << 20, device="cuda").sum()
    eager = tb.phase_variance(use_graph=False).clone()
    torch.cuda.synchronize()
    assert torch.equal(again, eager)
    assert not torch.equal(again, direct)
    tb.close()


def test_uneven_scans_in_one_launch(cuda):
    """Scans 128 and 192 (padded to 256) at detector 16: one launch over both
    equals each dataset's own batch."""
    from quantem.gpu_amd.ssb.compute.hip.engine import normalize_scan
    from quantem.gpu_amd.ssb.compute.hip.trials import TrialBatch
    torch = cuda
    a = C.poisson_data(123)
    b = normalize_scan(torch.from_numpy(C.poisson_data(126, scan=(192, 192))).cuda()).cpu().numpy()
    ea, eb = _prepared_engine(cuda, a), _prepared_engine(cuda, b)
    mags, angs, rot = _trials(9, seed=3)
    both = TrialBatch([ea, eb], 9)
    both.set(mags, angs, rot)
    res = both.phase_variance(use_graph=True).clone()
    for i, e in enumerate((ea, eb)):
        one = TrialBatch([e], 9)
        one.set(mags, angs, rot)
        assert torch.equal(one.phase_variance(use_graph=False)[0], res[i])
        one.close()
    both.close()


def test_search_finds_the_generating_defocus(cuda):
    """The objective peaks at the forward model's aberrations: a 1-D design
    over C10 on a weak-phase dataset (C10 = 200 A) picks the design point
    nearest 200 A."""
    from quantem.gpu_amd.ssb.compute.hip.trials import TrialBatch
    from quantem.gpu_amd.ssb.synthetic import SyntheticConfig, generate
    cfg = SyntheticConfig(scan=128, det=64, dose=1e6, seed=5, sigma_px=3.0, aberrations={"C10": 200.0})
    data = generate(cfg, device="cuda").cpu().numpy()
    eng = _prepared_engine(cuda, data, radius=None, semi=20.0, ds=1.25, step=cfg.step_A)
    grid = np.arange(0.0, 401.0, 25.0)
    mags = np.zeros((grid.size, 14))
    mags[:, 0] = grid
    tb = TrialBatch([eng], grid.size)
    tb.set(mags, np.zeros_like(mags))
    var = tb.phase_variance().cpu().numpy()[0]
    assert grid[int(np.argmax(var))] == 200.0
    tb.close()


def test_odd_aberrations_and_missing_stack_are_rejected(cuda):
    from quantem.gpu_amd.ssb.compute.hip.trials import TrialBatch
    from tests._gpu import make_engine
    eng = _prepared_engine(cuda, C.poisson_data(123))
    tb = TrialBatch([eng], 4)
    mags = np.zeros((1, 14))
    mags[0, 2] = 10.0                                   # C21 (odd)
    with pytest.raises(NotImplementedError, match="even aberrations"):
        tb.set(mags, np.zeros((1, 14)))
    tb.close()
    bare = make_engine(cuda, 128, (16, 16), 8, O.wavelength_angstrom(300), 21.4, (1.0, 1.0), 0.5)
    with pytest.raises(ValueError, match="Fourier stack"):
        TrialBatch([bare], 4)


def test_exact_batch_graph_replay_equals_per_trial_batch(cuda):
    """The reference's exact objective over a batch of trials replayed from a
    captured hipGraph: bit-identical to ``ssb_batch_phase_loss`` (trials
    enqueued one by one), on the first capture, on replays with new trials of
    the same pattern, and after a pattern change (recapture)."""
    torch = cuda
    eng = _prepared_engine(cuda, C.poisson_data(125))
    eng.cache_rotation(math.radians(1.0))
    out = torch.empty(7, dtype=torch.float32, device="cuda")     # one output: the graph is replayed
    mags, angs, rot = _trials(7, seed=4)
    mags[2, 4] = 0.0                                   # a quadratic-chi trial among C_s trials
    for seed in (4, 5):
        if seed == 5:
            m2, a2, r2 = _trials(7, seed=5)
            m2[2, 4] = 0.0
            mags, angs, rot = m2, a2, r2
        eager = eng.batch_phase_loss(mags, angs, rot).clone()
        graph = eng.batch_phase_loss(mags, angs, rot, out=out, graph=True).clone()
        again = eng.batch_phase_loss(mags, angs, rot, out=out, graph=True).clone()
        torch.cuda.synchronize()
        assert torch.equal(graph, eager) and torch.equal(again, eager)
    mags[3, 4] = 0.0                                   # new pattern: recaptured
    eager = eng.batch_phase_loss(mags, angs, rot).clone()
    graph = eng.batch_phase_loss(mags, angs, rot, out=out, graph=True).clone()
    torch.cuda.synchronize()
    assert torch.equal(graph, eager)
    # the cached rotation is restored after the batch
    single = eng.reconstruct_with_loss(-120.0, 0.0, 0.0)[1]
    eng.cache_rotation(math.radians(1.0), force=True)
    assert eng.reconstruct_with_loss(-120.0, 0.0, 0.0)[1] == single
