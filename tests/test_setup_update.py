"""Profile updates (mt_setup_update_cn2, setup_recon_tomo_update): only the
regulariser depends on the layer weights and r0, so the FDPCG blocks' WFS
sums are kept from the setup and the update redoes the diagonal, the drop
tolerances and the block pseudo inverses.  The result must be the same bits
as a setup built from scratch with the new profile."""
import numpy as np
import pytest

from maos_amd import configs, tomo


@pytest.mark.parametrize("name", ["small8", "large30"])
def test_update_equals_fresh_setup(name):
    cfg = getattr(configs, name)()
    r = tomo.Recon(cfg)
    n = r.npsr
    for k, r0 in enumerate((0.15, 0.22)):
        wt = list(np.linspace(1.0, 0.2 + 0.3 * k, n) / np.linspace(1.0, 0.2 + 0.3 * k, n).sum())
        r.update_cn2(wt, r0)
        fresh = tomo.Recon(dict(cfg, wt=wt, r0z=r0))
        a, b = r.fdpcg(), fresh.fdpcg()
        for key in ("Mhat", "Minv", "perm", "permhf"):
            assert np.array_equal(a[key], b[key]), (name, r0, key)
        assert np.array_equal(r.l2c(), fresh.l2c())
