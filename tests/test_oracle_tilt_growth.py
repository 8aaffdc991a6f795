"""Growth of the reference 2D elastic scheme for tilted symmetry axes, on the
CPU oracle.

The reference's sigma_xx / sigma_zz update adds c15 * (dvz_dx + dvx_dz) and
c35 * (dvz_dx + dvx_dz) with both shear derivatives taken at the normal-stress
point, without the averaging a staggered grid needs there
(cpmlfdtd.f08:325-345, the same form at 349-357 for sigma_zz).  A tilted
symmetry axis makes c15 and c35 non-zero, and the run then grows
exponentially over a few thousand steps.  This file shows that property of
the reference scheme, and shows that c15/c35 are what drives it:

* vertical axes (tilt 0) and small tilts stay bounded over 3000 steps;
* a 30 degree tilt grows by more than 1e6 over its early peak, with the
  reference's 3-term velocity differences and with the 4-term ones alike
  (so the truncated velocity stencils, cpmlfdtd.f08:405-446, are not the
  cause);
* the same 30 degree tensor with c15 = c35 = 0 (every other coefficient
  kept) stays bounded.

The uneven-batch and FWI workloads run for 1500-3000 steps; these numbers
are the evidence for running them with vertical symmetry axes.
"""
import numpy as np
import pytest

from oracle.oracle_ctypes import run_oracle

STEPS = 3000
EARLY = 600      # the source pulse has left the source region by then


def _run(tilt, stencil="reference", zero=()):
    from seidart_hip import models
    cfg = models.elastic2(nx=100, nz=100, steps=STEPS, seed=5, src=(50, 50), rec_z=50,
                          tilt=(tilt, tilt))
    cfg["stencil"] = stencil
    if tilt:
        assert np.abs(cfg["arrays"]["c15"]).max() > 0 and np.abs(cfg["arrays"]["c35"]).max() > 0
    for k in zero:
        cfg["arrays"][k] = np.zeros_like(cfg["arrays"][k])
    _, (vn, _), _, unstable = run_oracle(cfg, nthreads=4)
    assert unstable == 0          # far below the 1e25 blow-up threshold
    assert np.isfinite(vn).all()
    growth = vn[-1] / vn[:EARLY].max()
    print(f"tilt {tilt} {stencil} zero={zero}: early peak {vn[:EARLY].max():.2e}, "
          f"final {vn[-1]:.2e}, growth {growth:.2e}", flush=True)
    return growth


@pytest.mark.parametrize("tilt", [0.0, 5.0])
def test_vertical_and_small_tilt_stay_bounded(tilt):
    assert _run(tilt) < 1.0


@pytest.mark.parametrize("stencil", ["reference", "corrected"])
def test_thirty_degree_tilt_grows(stencil):
    assert _run(30.0, stencil) > 1e6


def test_growth_comes_from_c15_c35():
    assert _run(30.0, zero=("c15", "c35")) < 1.0


def test_growth_increases_with_tilt():
    g15, g20, g30 = _run(15.0), _run(20.0), _run(30.0)
    assert 1.0 < g15 < g20 < g30
