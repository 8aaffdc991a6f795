"""Near-offset mute and weighted L2 misfit of seidart_hip.fwi (CPU only)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _shot(src, recs):
    return dict(src=src, receivers=np.asarray(recs, dtype=np.int32))


def test_near_offset_weights_mark_receivers_at_the_source():
    from seidart_hip import fwi
    recs = [[x, 1, 10] for x in (1, 7, 13, 19, 25)] + [[13, 1, 13], [13, 1, 14]]
    w = fwi.near_offset_weights([_shot((13, 10), recs), _shot((20, 10), recs)], radius=2)
    # shot 0: x = 13 at depth 10 (offset 0) and at depth 13 (offset 3) -> only the first is muted
    assert w[0].tolist() == [1, 1, 0, 1, 1, 1, 1]
    # shot 1: source x = 20 -> receiver x = 19 (offset 1) muted, x = 25 (offset 5) kept
    assert w[1].tolist() == [1, 1, 1, 0, 1, 1, 1]
    assert fwi.near_offset_weights([_shot((13, 10), recs)], radius=0)[0].tolist() == [1, 1, 0, 1, 1, 1, 1]


def test_weighted_misfit_is_the_sum_over_kept_receivers():
    from seidart_hip import fwi
    rng = np.random.default_rng(0)
    pred = torch.tensor(rng.standard_normal((3, 50, 6, 2)))
    obs = torch.tensor(rng.standard_normal((3, 50, 6, 2)))
    w = np.ones((3, 6))
    w[1, 2] = 0.0
    w[2, 5] = 0.0
    want = sum(0.5 * float(torch.sum((pred[q][:, r] - obs[q][:, r]) ** 2))
               for q in range(3) for r in range(6) if w[q, r])
    assert abs(float(fwi.l2_misfit(pred, obs, w)) - want) <= 1e-12 * want
    assert abs(float(fwi.l2_misfit(pred[1], obs[1], w[1])) -
               sum(0.5 * float(torch.sum((pred[1][:, r] - obs[1][:, r]) ** 2)) for r in range(6) if w[1, r])) < 1e-9
    assert float(fwi.l2_misfit(pred, obs)) == pytest.approx(0.5 * float(torch.sum((pred - obs) ** 2)), rel=1e-14)
