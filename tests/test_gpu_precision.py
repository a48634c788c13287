"""The precision study on the synthetic turbulent field (random Fourier
modes, flows.RandomFourierField), scaled to a size the FP64 oracle runs in
seconds: 24^3 periodic box, wavenumbers |k| <= 4 (the study's 42 at 256^3
scaled with the grid), dt = 0.005, forward and backward FTLE at T = 0.2 with
h = 0.1, RK3 + cubic B-spline.  Every storage mode of the library runs on
the same snapshots and is compared with the FP64 oracle at every output
(both directions):

  mode                                  flowmap rel.   FTLE abs. (max |FTLE| ~ 14)
  FP64 library                          1e-10          1e-9
  FP64 compute, FP32 velocity+flowmap   1e-6           1e-5
  FP32                                  1e-5           1e-4
  FP32 + COMPENSATED                    1e-5           1e-4
  FP16 stage fields + COMP.             2e-4           2e-3
  BF16 stage fields + COMP.             2e-3           2e-2

and the errors must order as the roundoffs do (FP64 < FP32 storage under
FP64 compute < FP32 < FP16 < BF16)."""
import numpy as np
import pytest

import flows as F
from scenarios import rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N, KMAX, DT, T, H, NSTEPS = 24, 4, 0.005, 0.2, 0.1, 44


@pytest.fixture(scope="module")
def study():
    """Snapshots and the oracle's outputs, computed once."""
    from oracle_api import Oracle
    field = F.RandomFourierField(N, kmax=KMAX, seed=2, device="cuda:0")
    X, Y, Z, fl = F.abc_grid(N)
    vel = []
    for s in range(NSTEPS):
        u = field.at(s * DT)
        vel.append(tuple(np.ascontiguousarray(a.cpu().numpy().transpose(2, 1, 0)) for a in u))
    o = Oracle(X, Y, Z, fl)
    hds = [o.diagnostic_init(d, 0, T, H, "t") for d in (F.FTLE_FWD, F.FTLE_BKWD)]
    o.set_option("INTEGRATOR", F.RK3)
    o.set_option("INTERPOLATOR", F.CUBIC)
    outs = {}
    for s in range(NSTEPS):
        o.update(*vel[s], s * DT)
        assert o.error == 0
        if o.noutput(hds[0]) > len(outs):
            outs[s] = [(o.composed(hd), o.ftle(hd)) for hd in hds]
    assert len(outs) == 2
    o.close()
    return (X, Y, Z, fl), vel, outs


def run_mode(study, variant, opts):
    import libcfd2lcs_amd as L
    (X, Y, Z, fl), vel, outs = study
    lib = L.Cfd2lcs(variant)
    lib.finalize()
    lib.init(X, Y, Z, fl)
    # storage options take effect at a reset: set them before the first
    for k, v in opts:
        lib.set_option(k, v)
    hds = [lib.diagnostic_init(d, 0, T, H, "t") for d in (F.FTLE_FWD, F.FTLE_BKWD)]
    lib.set_option("INTEGRATOR", F.RK3)
    lib.set_option("INTERPOLATOR", F.CUBIC)
    efm = eft = 0.0
    scale = 0.0
    for s in range(NSTEPS):
        lib.update(*vel[s], s * DT)
        assert lib.error == 0
        if s in outs:
            for hd, (comp, ftle) in zip(hds, outs[s]):
                assert lib.noutput(hd) == len([a for a in outs if a <= s])
                g = lib.composed(hd)
                efm = max(efm, max(rel_err(g[c], comp[c]) for c in range(3)))
                eft = max(eft, float(np.abs(lib.ftle(hd) - ftle).max()))
                scale = max(scale, float(np.abs(ftle).max()))
    lib.finalize()
    return efm, eft, scale


def test_turbulence_precision_study(study):
    import libcfd2lcs_amd as L
    modes = [
        ("FP64", "dp", (), 1e-10, 1e-9),
        ("FP64/FP32-storage", "dp", (("VELOCITY_STORAGE", L.LCS_FP32), ("FLOWMAP_STORAGE", L.LCS_FP32)), 1e-6, 1e-5),
        ("FP32", "sp", (), 1e-5, 1e-4),
        ("FP32+COMPENSATED", "sp", (("COMPENSATED", L.LCS_TRUE),), 1e-5, 1e-4),
        ("FP16+COMPENSATED", "sp", (("VELOCITY_STORAGE", L.LCS_FP16), ("COMPENSATED", L.LCS_TRUE)), 2e-4, 2e-3),
        ("BF16+COMPENSATED", "sp", (("VELOCITY_STORAGE", L.LCS_BF16), ("COMPENSATED", L.LCS_TRUE)), 2e-3, 2e-2),
    ]
    res = {}
    for name, variant, opts, tfm, tft in modes:
        efm, eft, scale = run_mode(study, variant, opts)
        res[name] = eft
        print(f"{name:18s} flowmap rel {efm:.3e}  FTLE abs {eft:.3e}  (max |FTLE| {scale:.3f})")
        assert efm <= tfm and eft <= tft, name
    assert res["FP64"] < res["FP64/FP32-storage"] < res["FP32"] < res["FP16+COMPENSATED"] < res["BF16+COMPENSATED"]
