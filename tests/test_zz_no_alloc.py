"""dlpx_two_body_forces_noalloc (the allocation-free variant of
dlpx_two_body_forces) against dlpx_two_body_forces on the same context, on two
calls in a row, the second with fewer atoms, and against the oracle: FP64
forces to 1e-10 RMS relative, mixed to 1e-5.  Needs an MI355X."""
import numpy as np
import pytest

import oracle
from dlpx import lib, systems

pytestmark = pytest.mark.gpu


def rms_rel(f, ref):
    return float(np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum()))


@pytest.mark.parametrize("precision,tol,same", [(lib.FP64, 1e-10, 1e-13), (lib.MIXED, 1e-5, 1e-6)])
def test_noalloc_matches_two_body_forces_twice_smaller_second(precision, tol, same):
    big = systems.nacl(ncells=6, kvec=(32, 32, 32), seed=11)    # 1728 ions
    small = systems.nacl(ncells=5, kvec=(32, 32, 32), seed=12)  # 1000 ions
    ctx = lib.Context(big, precision=precision, natms_max=big.natms)
    for s in (big, small):
        print("case noalloc precision", precision, "natms", s.natms, flush=True)
        ctx.set_system(s)
        ctx.link_cell_pairs(s.pos)
        ref = ctx.two_body_forces(s.pos)
        got = ctx.two_body_forces(s.pos, noalloc=True)
        # the same kernels and launch shapes: equal up to the order of the
        # spreading atomics
        assert rms_rel(got["frc"], ref["frc"]) <= same
        for k in ("engcpe", "vircpe", "engsrp", "virsrp"):
            assert abs(got[k] - ref[k]) <= same * max(abs(ref[k]), abs(ref["buffer"][9]), 1.0), k
        assert np.abs(got["stress"] - ref["stress"]).max() <= same * np.abs(ref["stress"]).max()
        o = oracle.two_body(s)
        assert rms_rel(got["frc"], o["frc"]) <= tol
        for k in ("engcpe", "vircpe", "engsrp", "virsrp"):
            assert abs(got[k] - o[k]) <= 10 * tol * max(abs(o[k]), abs(o["buffer"][9]), 1.0), k
    ctx.close()


def test_noalloc_rejects_per_particle_output():
    s = systems.nacl(ncells=5, kvec=(32, 32, 32), seed=12)
    ctx = lib.Context(s)
    ctx.link_cell_pairs(s.pos)
    with pytest.raises(lib.DlpxError) as e:
        ctx.two_body_forces(s.pos, per_particle=True, noalloc=True)
    assert e.value.code == 1002
    ctx.close()
