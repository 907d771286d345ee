"""The bright-field superset of the reconstruction path: the stream pass
copies a sampled box of every uint16 frame compactly and the gather reads
it when the exact selection lies inside.  Phases equal the frame gather's
bit for bit; a selection outside the box (an explicit radius wider than the
disk) falls back to the frames with the same bits."""
from __future__ import annotations

import numpy as np
import pytest

from tests import _cases as C
from tests._gpu import engine_for_data, synthetic

pytestmark = pytest.mark.gpu

SUPERSET_KNOB = 3          # QSSB_TUNE_SUPERSET: 1 compact copy, 2 frames (default)


@pytest.mark.parametrize("case", ["synthetic-128x64", "poisson-radius3", "wide-radius24", "synthetic-256x128"])
def test_superset_copy_gives_the_frame_gather_bits(cuda, case):
    torch = cuda
    radius = None
    if case == "synthetic-128x64":
        cfg, data = synthetic(128, 64, {"C10": 150.0}, rotation_deg=2.0)
        kw = dict(semiangle_mrad=20.0, det_sampling=1.25, step=cfg.step_A)
    elif case == "synthetic-256x128":
        cfg, data = synthetic(256, 128, {"C10": 200.0})
        kw = dict(semiangle_mrad=20.0, det_sampling=1.25, step=cfg.step_A)
    elif case == "wide-radius24":
        cfg, data = synthetic(128, 128, {"C10": 150.0})
        kw = dict(semiangle_mrad=20.0, det_sampling=1.25, step=cfg.step_A)
        radius = 24.0
    else:
        data = C.poisson_data(123)
        kw = dict(semiangle_mrad=21.4, det_sampling=1.0, step=0.5)
        radius = 3.0
    n = data.shape[0]
    eng = engine_for_data(torch, data, **kw)
    d = torch.from_numpy(data).cuda()
    out = {}
    for mode in (1, 2):
        eng.tune_set(SUPERSET_KNOB, mode)
        ph = torch.empty((n, n), dtype=torch.float32, device="cuda")
        obj = torch.empty((n, n), dtype=torch.complex64, device="cuda")
        eng.recon_u16(d, 150.0, 0.0, 0.0, 0.0, 0.0, radius, obj_out=obj, phase_out=ph)
        out[mode] = (ph, obj)
    torch.cuda.synchronize()
    assert torch.equal(out[1][0], out[2][0]) and torch.equal(out[1][1], out[2][1])
    assert eng.tune_get(SUPERSET_KNOB) == 2


def test_superset_buffers_survive_a_growing_graph_batch(cuda):
    # the hipGraph batch reallocates only its own trial table when a larger
    # batch arrives; the superset buffers of the plan stay allocated and the
    # next reconstruction gives the same bits
    torch = cuda
    cfg, data = synthetic(128, 64, {"C10": 150.0}, rotation_deg=2.0)
    eng = engine_for_data(torch, data, semiangle_mrad=20.0, det_sampling=1.25, step=cfg.step_A)
    eng.tune_set(SUPERSET_KNOB, 1)
    d = torch.from_numpy(data).cuda()
    n = data.shape[0]

    def recon():
        ph = torch.empty((n, n), dtype=torch.float32, device="cuda")
        eng.recon_u16(d, 150.0, 0.0, 0.0, 0.0, 0.0, None, phase_out=ph)
        return ph

    first = recon()
    eng.prepare_g(d)
    for t in (2, 5):
        mags = np.zeros((t, 14))
        mags[:, 0] = np.linspace(100.0, 200.0, t)
        losses = eng.batch_phase_loss(mags, np.zeros((t, 14)), graph=True)
        assert torch.isfinite(losses).all()
    second = recon()
    torch.cuda.synchronize()
    assert torch.equal(first, second)
