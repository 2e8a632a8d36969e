"""cac_eam_atoms_noalloc (the EAM atom passes over caller-owned device arrays,
allocating nothing) against cac_force's atom forces and densities on the
same positions and lists: two calls in a row on one device, the second with
fewer atoms (the fixed scratch keeps the first call's values beyond the
second's atoms, which must not leak in), and the size limit."""
import ctypes as C

import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM, _lib
from cac_code_amd.lattice import build_cac_system

from gpu_common import jitter, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


class Hip:
    """Device buffers for the test, the caller's memory (libamdhip64)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0      # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def eam_pair(cu_funcfl):
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    return p


def atoms_only(L, seed):
    """A thermal all-atom box (the core covers the whole box)."""
    return jitter(build_cac_system(L, 2, 100.0, temperature=300.0, seed=seed), 0.06, seed + 1)


def reference_and_noalloc(eng, hip, s):
    eng.load_system(s)
    eng.build()
    eng.force()
    x, _, f_ref = eng.atoms()
    rho_ref, _ = eng.rho()
    num, nb = eng.atom_neighbors()
    lists = np.ascontiguousarray((nb & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    n, M = lists.shape
    dx, dnum, dnb = hip.upload(x), hip.upload(num.astype(np.int32)), hip.upload(lists)
    df, drho = hip.upload(np.zeros((n, 3))), hip.upload(np.zeros(n))
    rc = eng.L.cac_eam_atoms_noalloc(eng.ctx, n, dx, M, dnum, dnb, df, drho)
    assert rc == _lib.CAC_OK, eng.L.cac_last_error(eng.ctx)
    eng.synchronize()
    return f_ref, rho_ref, hip.download(df, f_ref), hip.download(drho, rho_ref)


def test_noalloc_matches_force(cu_funcfl):
    hip = Hip()
    eng = Engine(0)
    try:
        eng.set_pair(eam_pair(cu_funcfl))
        eng.set_neighbor(5.95, skin=1.0)
        for L, seed in ((10, 3), (8, 5)):            # the second call has fewer atoms
            s = atoms_only(L, seed)
            print(f"[noalloc] atoms={s.n_atoms}", flush=True)
            f_ref, rho_ref, f, rho = reference_and_noalloc(eng, hip, s)
            assert np.abs(f_ref).max() > 0.05         # thermal forces, not a perfect lattice
            assert rel_err(f, f_ref) < TOL
            assert rel_err(rho, rho_ref) < TOL
    finally:
        eng.close()
        hip.free()


def test_noalloc_size_limit(cu_funcfl):
    eng = Engine(0)
    try:
        eng.set_pair(eam_pair(cu_funcfl))
        rc = eng.L.cac_eam_atoms_noalloc(eng.ctx, (1 << 20) + 1, None, 0, None, None, None, None)
        assert rc == _lib.CAC_ERR_ARG
        assert b"exceed the limit" in eng.L.cac_last_error(eng.ctx)
    finally:
        eng.close()
