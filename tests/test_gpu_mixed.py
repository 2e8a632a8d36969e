"""Mixed precision (FP32 spline and pair arithmetic, FP64 positions and
accumulation) against the FP64 path on the same configurations.  Bars:
forces within 1e-4 (atoms) and 1e-3 (nodes) relative RMS, energy within 1e-6
relative, virial within 1e-4 of its largest component; over a 2,000-step NVE
run of a crystal, total energy within 1e-6 eV per atom."""
import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM, _lib
from cac_code_amd.lattice import build_cac_system

from gpu_common import CUTNEIGH_EAM, rms_rel, small_system

pytestmark = pytest.mark.gpu

SYSTEMS = {
    "small": lambda: small_system(seed=7, L=10, nae=5, core=4.0),
    "elements_only": lambda: small_system(seed=11, L=10, nae=5, core=0.0),
    "medium": lambda: small_system(seed=3, L=20, nae=5, core=8.0, sigma=0.08),
    "finest": lambda: small_system(seed=5, L=8, nae=2, core=3.0, sigma=0.03),
}


@pytest.mark.parametrize("coupled", [False, True], ids=["plain", "coupled"])
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_mixed_against_fp64(name, coupled, cu_funcfl):
    s = SYSTEMS[name]()
    print(f"[mixed] {name} coupled={coupled}: atoms={s.n_atoms} elements={s.n_elements}", flush=True)
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    with Engine(0) as eng:
        eng.load_system(s)
        eng.set_pair(p)
        eng.set_neighbor(CUTNEIGH_EAM, every=1, delay=0, check=True)
        eng.set_coupling(coupled)
        eng.build()
        out = {}
        for prec in (_lib.PREC_FP64, _lib.PREC_MIXED):
            eng.force(eflag=True, vflag=True, precision=prec)
            e, w = eng.energy_virial()
            _, _, f = eng.atoms()
            _, _, nf = eng.elements()
            out[prec] = (f, nf, float(np.sum(e)), np.asarray(w))
    f64, n64, e64, w64 = out[_lib.PREC_FP64]
    fmx, nmx, emx, wmx = out[_lib.PREC_MIXED]
    if s.n_atoms:
        assert rms_rel(fmx, f64) < 1e-4
    if s.n_elements:
        assert rms_rel(nmx, n64) < 1e-3
    assert abs(emx - e64) <= 1e-6 * abs(e64)
    assert np.max(np.abs(wmx - w64)) <= 1e-4 * np.max(np.abs(w64))


def _nve_energy(s, p, precision, steps, chunk):
    """Total energy per particle every `chunk` steps of a device NVE run."""
    from cac_code_amd.lattice import kinetic_energy
    with Engine(0) as eng:
        eng.load_system(s)
        eng.set_pair(p)
        eng.set_neighbor(CUTNEIGH_EAM, every=1, delay=10, check=True)
        eng.build()
        series = []
        for k in range(0, steps + 1, chunk):
            if k:
                eng.run(chunk, precision=precision)
            eng.force(eflag=True, precision=precision)
            _, v, _ = eng.atoms()
            _, nv, _ = eng.elements()
            series.append((float(np.sum(eng.energy_virial()[0])) + kinetic_energy(s, v, nv)) / s.n_integrated)
    return np.asarray(series)


def test_mixed_nve_drift_atoms(cu_funcfl):
    """Velocity Verlet on a periodic 300 K crystal (atoms only, the finest
    mesh, where the forces are the gradient of the energy): FP64 conserves
    the total energy once the integrator's first offset is in (no trend over
    2,000 steps), and the mixed trajectory's energy stays within 1e-6 eV per
    atom of FP64's."""
    s = build_cac_system(10, 5, 1.0e6, temperature=300.0, seed=21)
    assert s.n_elements == 0 and s.n_atoms == 4000
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    e64 = _nve_energy(s, p, _lib.PREC_FP64, 2000, 500)
    emx = _nve_energy(s, p, _lib.PREC_MIXED, 2000, 500)
    print(f"[mixed] atoms-only NVE dE/atom fp64 {e64 - e64[0]} mixed {emx - emx[0]}", flush=True)
    assert np.max(np.abs(e64[1:] - e64[1])) < 1e-6
    assert abs(e64[-1] - e64[0]) < 1e-4
    assert np.max(np.abs((emx - emx[0]) - (e64 - e64[0]))) < 1e-6
