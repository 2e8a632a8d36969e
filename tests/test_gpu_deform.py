"""Forces after a compressive remap that keeps the lists, against the oracle.

cac_deform rescales the box, atoms and nodes; while the rebuild check allows
it, the lists of the last build stay in use.  A lattice with 0.02 A of noise
(which gives the atoms nonzero forces, as a perfect lattice would not) is
built with a 5.95 A list cutoff and compressed by 0.88 or 0.95 without a
rebuild: 4.95 / 0.88 < 5.95, so the old lists still hold every pair inside
the 4.95 A force cutoff, and the density pass must fall back from its inner
mask to the whole list (the remap invalidates it).  At 0.88 the first
neighbor shell moves to 2.25 A, below the force tables' LDS window
(r >= 2.415 A), so every first-neighbor pair is read from the global tables;
at 0.95 (2.43 A) they stay in the window.  Forces must equal the oracle's at
the compressed coordinates: FP64 to 1e-10, mixed precision within its FP32
error.
"""
import pytest

from cac_code_amd import Engine, PairEAM, _lib
from cac_code_amd.lattice import build_cac_system

from gpu_common import CUTNEIGH_EAM, jitter, oracle_state, rel_err, rms_rel

pytestmark = pytest.mark.gpu


def eam_pair(cu_funcfl):
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    return p


@pytest.mark.parametrize("scale", [0.88, 0.95])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_deform_without_rebuild_matches_oracle(orc, cu_funcfl, cu_tables, scale, precision):
    s = jitter(build_cac_system(10, 5, 4.0, temperature=0.0), 0.02, 95)
    prec = _lib.PREC_FP64 if precision == "fp64" else _lib.PREC_MIXED
    print(f"[deform] atoms={s.n_atoms} elements={s.n_elements} scale={scale} {precision}", flush=True)
    eng = Engine(0)
    try:
        eng.load_system(s)
        eng.set_pair(eam_pair(cu_funcfl))
        eng.set_neighbor(CUTNEIGH_EAM, skin=1.0, every=1, delay=0, check=True)
        eng.set_coupling(True)
        eng.build()
        eng.force(precision=prec)            # a full density pass: the inner mask is set
        eng.deform([scale, scale, scale])
        assert not eng.decide()              # the lists of the build stay
        eng.force(precision=prec)
        x, _, f = eng.atoms()
        nodex, _, nf = eng.elements()
    finally:
        eng.close()
    r = s.__class__(**{**s.__dict__, "x": x, "nodex": nodex, "box_hi": s.box_lo + scale * (s.box_hi - s.box_lo)})
    ref = oracle_state(orc, r, CUTNEIGH_EAM, tables=cu_tables, coupled=True)
    fa, fn = ref["atoms"]["f"], ref["elems"]["nodef"]
    if precision == "fp64":
        assert rel_err(f, fa) < 1e-10
        assert rel_err(nf, fn) < 1e-10
    else:
        assert rms_rel(f, fa) < 2e-5 and rel_err(f, fa) < 1e-4
        assert rms_rel(nf, fn) < 2e-4 and rel_err(nf, fn) < 1e-3
