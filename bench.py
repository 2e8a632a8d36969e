#!/usr/bin/env python3
"""Throughput of the CAC force-and-neighbor path on one MI355X.

One step = FixNVE initial integrate, rebuild check (and rebuild when an atom
or node moved more than half the skin), EAM force on atoms and element
integration points with the node scatter, FixNVE final integrate
(verlet.cpp:149-219), all device-resident through libcac_hip.so.

Metric: Matom-steps/s = (N_atoms + 8*N_elements) * steps / s / 1e6.

Default workload: EAM Cu CAC, periodic box of 198^3 fcc cells (715.8 A),
elements of 11 sites per edge with 8 corner integration points, atomistic
core of dropped blocks (500,456 atoms, 22,952 elements), Johnson-type
analytic EAM tabulated with nrho = nr = 10000, rc = 4.95 A, neighbor cutoff
5.95 A, 300 K seeded velocities, dt = 1 fs.  Synthetic data, generated here.

The JSON line carries the roofline of the dominant kernel (algorithmic
bytes / measured average launch time, HIP events on the context stream), the
CPU oracle timed on this host on a bounded sample of the same workload, and
the parity of the device's first force call with the oracle's on the full
workload (neighbor sets compared sorted, FP64 forces norm-relative).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
K_RC = 42  # fcc Cu neighbours inside rc = 4.95 A (12 + 6 + 24)

# Algorithmic bytes per unit for each kernel family: compulsory FP64 state
# (positions 24 B, F' 8 B, forces 24 B per atom; 8 nodes x 24 B, erho 8 x 8 B
# per element) plus 4-byte atom-list entries, or 8-byte integration-point
# entries for 8 points, for the K_RC neighbors inside the force cutoff.
ALG_BYTES = {
    "atom_rho": ("atoms", 24 + 8 + 4 * K_RC),
    "atom_force": ("atoms", 24 + 8 + 24 + 4 * K_RC),
    "ip_erho": ("elements", 192 + 64 + 8 * 8 * K_RC),
    "ip_force": ("elements", 192 + 64 + 192 + 8 * 8 * K_RC),
}
B_ATOM = 616
B_ELEM = 7640
# HBM bytes per launch of the FP64 kernels on the eam_c1 workload, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE runs scaled by a plain 1 GiB copy's
# counts (fetch x2.000, write x1.000).  A profile-derived constant, not
# measured by this script: TRAFFIC_SOURCE names the summary it comes from.
TRAFFIC_SOURCE = "profiles/eam_c1_kernels.md (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)"
TRAFFIC = {"atom_force": 189.9e6 + 28.3e6, "atom_rho": 184.8e6 + 27.2e6, "ip_force": 135.2e6 + 18.8e6,
           "ip_erho": 120.2e6 + 12.2e6}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(name, seed):
    from cac_code_amd.lattice import build_cac_system, eam_c1_system, eam_c3_system, lj_c0_system
    if name == "eam_c1":
        return eam_c1_system(seed=seed)
    if name == "eam_c3":
        return eam_c3_system(seed=seed)
    if name == "lj_c0":
        return lj_c0_system(seed=seed)
    if name == "small":
        return build_cac_system(20, 5, 8.0, temperature=300.0, seed=seed)
    raise ValueError(name)


def make_pair(workload):
    from cac_code_amd import PairEAM, PairLJCut, johnson_cu_funcfl
    if workload == "lj_c0":
        p = PairLJCut()
        p.settings(["5.0"])
        p.coeff(["1", "1", "0.4093", "2.338"])
        return p, 6.0, 1.0, None
    ff = johnson_cu_funcfl()
    p = PairEAM()
    p.settings([])
    p.coeff(["*", "*", "Cu_johnson.eam"], funcfl=ff)
    return p, 5.95, 1.0, ff


COUPLED = True  # atom-element coupling (set from --no-coupling in main)


KERNELS = ("atom_rho", "atom_force", "site_pos", "ip_erho", "site_fp", "ip_force", "site_react", "scatter", "surr_feat",
           "surr_mlp",
           "surr_fit", "atom_neigh", "ip_neigh", "elem_neigh", "atom_sites")


def kernel_profile(eng, steps, precision):
    """Per-kernel average durations (HIP events on the context stream) over
    `steps` further steps; kept out of the timed region, whose clock must not
    include the event bookkeeping."""
    eng.synchronize()
    eng.profile(True)
    eng.run(steps, precision=precision)
    eng.synchronize()
    prof = {}
    for name in KERNELS:
        ms, cnt = eng.profile_avg_ms(name)
        if cnt:
            prof[name] = (ms, cnt)
    eng.profile(False)
    return prof


STRAIN_RATE_PS = 0.0  # uniaxial strain rate along x per ps (set from --strain-rate in main)


def gpu_run(s, pair, cutneigh, skin, steps, warmup, precision, every, delay, prof_steps=100):
    eng = new_engine(s, pair, cutneigh, skin, every, delay)
    if STRAIN_RATE_PS:
        eng.set_deform_rate([STRAIN_RATE_PS, 0.0, 0.0])
    nb_warm = eng.run(warmup, precision=precision) if warmup else 0
    eng.synchronize()
    t0 = time.perf_counter()
    nb = eng.run(steps, precision=precision)
    eng.synchronize()
    t1 = time.perf_counter()
    prof = kernel_profile(eng, min(steps, prof_steps), precision)
    return eng, t1 - t0, nb, nb_warm, prof


def centro_leg(eng, lo=1.0, hi=10.0, nnn=12):
    """Centro-symmetry of every atom after the timed run, the reference's
    defect output (Output::centro_atom, output.cpp:352-441; the xyz dump keeps
    lo < c < hi, dump.cpp:355-365), over the current lists; outside the timed
    region.  Its kernel time comes from the HIP-event timer."""
    import numpy as np
    eng.centro(nnn)                 # first launch loads the kernel's code object
    eng.synchronize()
    eng.profile(True)
    c = eng.centro(nnn)
    ms, cnt = eng.profile_avg_ms("centro")
    eng.profile(False)
    return {"nnn": nnn, "kernel_ms": round(ms, 3) if cnt else None, "window": [lo, hi],
            "defect_atoms": int(np.count_nonzero((c > lo) & (c < hi))), "atoms": int(c.size)}


def cpu_baseline(s, pair, ff, cutneigh, threads, warm, timed, rebuild_rate):
    """The CPU FP64 oracle on this host with `threads` OpenMP threads: one
    neighbor build, `warm` untimed and `timed` timed full steps (force on
    atoms and elements, both NVE halves); the build is charged at the GPU
    run's rebuild rate."""
    import numpy as np
    from oracle import oracle as O
    from cac_code_amd.lattice import FTM2V_METAL
    O.build()
    O.set_num_threads(threads)
    x = s.x.copy(); v = s.v.copy(); nodex = s.nodex.copy(); nodev = s.nodev.copy()
    es = O.ElementSet(nodex, s.nae, s.intp_ijk, s.intp_w)
    t0 = time.perf_counter()
    if COUPLED:
        anum, anb, inum, inb = O.coupled_neighbors(x, es, s.box_lo, s.box_hi, cutneigh, 2, 128, 128)
    else:
        anum, anb = O.atom_neighbors(x, s.box_lo, s.box_hi, cutneigh, 128)
        inum, inb = O.ip_neighbors(es, s.box_lo, s.box_hi, cutneigh, 2, 128)
    t_build = time.perf_counter() - t0
    lj = pair.init_one(1, 1) if ff is None else None
    T = O.eam_tables(ff) if ff is not None else None
    if COUPLED:
        last = {}

        def fa():
            fn = O.eam_coupled if ff is not None else O.lj_coupled
            r = fn(x, es, s.box_lo, s.box_hi, anum, anb, inum, inb, T if ff is not None else lj)
            last["nodef"] = r["nodef"]
            return r["f"]

        fe = lambda: last["nodef"]
    elif ff is not None:
        fa = lambda: O.eam_atoms(x, s.box_lo, s.box_hi, anum, anb, T)["f"]
        fe = lambda: O.eam_elements(es, s.box_lo, s.box_hi, inum, inb, T)["nodef"]
    else:
        fa = lambda: O.lj_atoms(x, s.box_lo, s.box_hi, anum, anb, lj)["f"]
        fe = lambda: O.lj_elements(es, s.box_lo, s.box_hi, inum, inb, lj)["nodef"]
    dt, dtf = 0.001, 0.5 * 0.001 * FTM2V_METAL
    m, me = s.mass, s.mass_e
    f = fa(); nodef = fe()
    xc = np.zeros((s.n_elements, 3))
    state = None
    for i in range(warm + timed):
        if i == warm:
            state = dict(x=x.copy(), nodex=es.nodex.copy())     # the configuration for the parity check
            t0 = time.perf_counter()
        O.nve_atoms_initial(x, v, f, dt, dtf, m)
        O.nve_nodes_initial(es.nodex, nodev, nodef, xc, dt, dtf, me)
        f = fa(); nodef = fe()
        O.nve_atoms_final(v, f, dtf, m)
        O.nve_nodes_final(nodev, nodef, dtf, me)
    t_step = (time.perf_counter() - t0) / timed
    per_step = t_step + rebuild_rate * t_build
    return dict(per_step=per_step, steps=timed, warm=warm, t_build=t_build, t_step=t_step,
                cores=O.num_threads(), state=state)


def full_size_parity(s, pair, ff, cutneigh, skin, every, delay, state):
    """The device's lists and FP64 forces against the oracle's on the full
    workload, at the configuration the CPU baseline reached after its warm-up
    steps (thermal forces, not the vanishing forces of the perfect start),
    wrapped into the box (orc_pbc_*) first; outside the timed region.  Sets
    are compared sorted, row by row; forces as max |gpu - cpu| / max |cpu|
    over atoms and over nodes."""
    import dataclasses
    import numpy as np
    from oracle import oracle as O
    x = state["x"].copy()
    nodex = state["nodex"].copy()
    O.pbc_atoms(x, s.box_lo, s.box_hi)
    O.pbc_elements(O.element_centers(nodex), nodex, s.box_lo, s.box_hi)
    es = O.ElementSet(nodex, s.nae, s.intp_ijk, s.intp_w)
    if COUPLED:
        rnum_a, rnb_a, rnum_i, rnb_i = O.coupled_neighbors(x, es, s.box_lo, s.box_hi, cutneigh, 2, 128, 128)
        r = (O.eam_coupled(x, es, s.box_lo, s.box_hi, rnum_a, rnb_a, rnum_i, rnb_i, O.eam_tables(ff)) if ff is not None
             else O.lj_coupled(x, es, s.box_lo, s.box_hi, rnum_a, rnb_a, rnum_i, rnb_i, pair.init_one(1, 1)))
        rf, rnf = r["f"], r["nodef"]
    else:
        rnum_a, anb32 = O.atom_neighbors(x, s.box_lo, s.box_hi, cutneigh, 128)
        rnb_a = anb32.astype(np.uint64) | (np.uint64(O.ATOM_SITE) << np.uint64(32))
        rnum_i, rnb_i = O.ip_neighbors(es, s.box_lo, s.box_hi, cutneigh, 2, 128)
        if ff is not None:
            T = O.eam_tables(ff)
            rf = O.eam_atoms(x, s.box_lo, s.box_hi, rnum_a, anb32, T)["f"]
            rnf = O.eam_elements(es, s.box_lo, s.box_hi, rnum_i, rnb_i, T)["nodef"]
        else:
            lj = pair.init_one(1, 1)
            rf = O.lj_atoms(x, s.box_lo, s.box_hi, rnum_a, anb32, lj)["f"]
            rnf = O.lj_elements(es, s.box_lo, s.box_hi, rnum_i, rnb_i, lj)["nodef"]
    eng = new_engine(dataclasses.replace(s, x=x, nodex=nodex), pair, cutneigh, skin, every, delay)
    try:
        _, _, f = eng.atoms()
        _, _, nodef = eng.elements()
        sets = {}
        for name, (num, nb), rnum, rnb in (("atoms", eng.atom_neighbors(), rnum_a, rnb_a),
                                           ("integration_points", eng.ip_neighbors(), rnum_i, rnb_i)):
            same_counts = bool(np.array_equal(num, rnum))
            w = int(max(num.max(initial=0), rnum.max(initial=0)))
            pad = np.uint64(0xFFFFFFFFFFFFFFFF)
            a = np.where(np.arange(w)[None, :] < num[:, None], nb[:, :w], pad)
            b = np.where(np.arange(w)[None, :] < rnum[:, None], rnb[:, :w], pad)
            a.sort(axis=1)
            b.sort(axis=1)
            sets[name] = {"rows": int(num.size), "entries": int(num.sum()), "counts_equal": same_counts,
                          "sets_bit_exact": bool(same_counts and np.array_equal(a, b))}
    finally:
        eng.close()

    def rel(a, b):
        sc = float(np.max(np.abs(b))) if b.size else 0.0
        return float(np.max(np.abs(a - b))) / sc if sc > 0 else 0.0

    return {"neighbor_sets": sets, "atom_force_rel_max": rel(f, rf), "node_force_rel_max": rel(nodef, rnf),
            "atom_force_max_eV_A": float(np.max(np.abs(rf))) if rf.size else 0.0,
            "node_force_max_eV_A": float(np.max(np.abs(rnf))) if rnf.size else 0.0,
            "tolerance": 1e-10, "configuration": "after the CPU baseline's warm-up steps, FP64"}


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def new_engine(s, pair, cutneigh, skin, every, delay):
    from cac_code_amd import Engine
    eng = Engine(0)
    eng.load_system(s)
    eng.set_pair(pair)
    eng.set_neighbor(cutneigh, skin=skin, every=every, delay=delay, check=True)
    eng.set_coupling(COUPLED)
    eng.build()
    eng.force()
    return eng


def _err(test, ref):
    """Force difference: absolute RMS (eV/A) of |test - ref| per particle,
    RMS |ref|, and relative RMS / max (divided by RMS |ref|), the relative
    ones only where the reference forces do not vanish."""
    import numpy as np
    d = (test - ref).reshape(-1, 3)
    r = ref.reshape(-1, 3)
    rms = float(np.sqrt(np.mean(np.sum(r * r, axis=1))))
    drms = float(np.sqrt(np.mean(np.sum(d * d, axis=1))))
    dmax = float(np.sqrt(np.max(np.sum(d * d, axis=1))))
    out = {"abs_rms_eV_A": drms, "ref_rms_eV_A": rms}
    if rms > 1e-6:          # relative figures only against forces that do not vanish
        out.update(rel_rms=drms / rms, rel_max=dmax / rms)
    return out


def accuracy(s, pair, cutneigh, skin, every, delay, at_steps, models=None):
    """Along an FP64 trajectory: at each step in at_steps, the mixed (and,
    for each named model, the mixed + surrogate) forces against the FP64
    forces on the same configuration."""
    import numpy as np
    from cac_code_amd import _lib
    from cac_code_amd import surrogate as S
    eng = new_engine(s, pair, cutneigh, skin, every, delay)
    models = models or {}
    out = []
    done = 0
    for step in at_steps:
        if step > done:
            eng.run(step - done)
            done = step
        eng.force(precision=_lib.PREC_FP64)
        _, _, fa = eng.atoms()
        _, _, fn = eng.elements()
        eng.force(eflag=True, precision=_lib.PREC_FP64)
        e64 = float(eng.energy_virial()[0].sum())
        eng.force(eflag=True, precision=_lib.PREC_MIXED)
        emx = float(eng.energy_virial()[0].sum())
        eng.force(precision=_lib.PREC_MIXED)
        _, _, fam = eng.atoms()
        _, _, fnm = eng.elements()
        rec = {"step": step, "mixed": {"atoms": _err(fam, fa), "nodes": _err(fnm, fn),
                                       "energy_rel": abs(emx - e64) / abs(e64), "energy_fp64_eV": e64}}
        for name, model in models.items():
            S.attach(eng, model, s)
            eng.force(precision=_lib.PREC_MIXED)
            _, _, fas = eng.atoms()
            _, _, fns = eng.elements()
            _, cls = eng.surrogate_features()
            info = eng.surrogate_info()
            eng.surrogate_enable(False)
            sm = cls == 0
            rec[name] = {"atoms": _err(fas, fa), "nodes": _err(fns, fn),
                         "surrogate_nodes": _err(fns[sm], fn[sm]) if sm.any() else None,
                         "fallback_fraction": float(1.0 - sm.mean()), **info}
        eng.force(precision=_lib.PREC_FP64)      # the trajectory continues on FP64 forces
        out.append(rec)
        log(f"[bench] accuracy at step {step}: {json.dumps(rec)}")
    eng.close()
    return out


def drift(s, pair, cutneigh, skin, every, delay, nsteps, precision, chunk=1000, limit=0.5):
    """Total energy (build-defined PE + KE) every `chunk` steps of an NVE run,
    stopped once |E - E0| exceeds `limit` eV per integrated particle (the
    force-based CAC forces are not the gradient of an energy, so the
    element region and the atom-element interface can heat up)."""
    from cac_code_amd.lattice import kinetic_energy
    eng = new_engine(s, pair, cutneigh, skin, every, delay)
    n_int = s.n_atoms + 8 * s.n_elements

    def etot():
        eng.force(eflag=True, precision=precision)
        pe = float(eng.energy_virial()[0].sum())
        _, v, _ = eng.atoms()
        _, nv, _ = eng.elements()
        return pe + kinetic_energy(s, v, nv)

    e0 = etot()
    series = [0.0]
    done = 0
    diverged = None
    t0 = time.perf_counter()
    while done < nsteps and diverged is None:
        k = min(chunk, nsteps - done)
        nb = eng.run(k, precision=precision)
        done += k
        series.append((etot() - e0) / n_int)
        log(f"[bench] drift prec={precision} step {done}: {series[-1]:+.3e} eV/particle, {nb} rebuilds, "
            f"{time.perf_counter() - t0:.1f}s")
        if abs(series[-1]) > limit:
            diverged = done
    eng.close()
    return {"steps": done, "steps_asked": nsteps, "diverged_at": diverged, "limit_eV_per_particle": limit,
            "e0_eV": e0, "drift_eV_per_particle": series[-1],
            "max_abs_eV_per_particle": max(abs(x) for x in series),
            "every": chunk, "series_eV_per_particle": series, "seconds": round(time.perf_counter() - t0, 1)}


def drift_regions(pair, cutneigh, skin, args):
    """The NVE drift of each half of the method on its own, FP64 and mixed:
    a periodic 300 K crystal of 500,000 atoms (the finest mesh, where the
    forces are the gradient of the energy) and the workload's whole box
    meshed into nae = 11 elements (23,328), whose force-based element forces
    (pair_eam.cpp:1013-1134) are not.  The mixed-minus-FP64 energy difference
    is what the precision mode adds."""
    from cac_code_amd import _lib
    from cac_code_amd.lattice import build_cac_system
    regions = {"atoms_only": build_cac_system(50, 5, 1.0e6, temperature=300.0, seed=args.seed),
               "elements_only": build_cac_system(198, 11, 0.0, temperature=300.0, seed=args.seed)}
    out = {}
    for name, s in regions.items():
        rec = {"atoms": s.n_atoms, "elements": s.n_elements}
        for tag, prec in (("fp64", _lib.PREC_FP64), ("mixed", _lib.PREC_MIXED)):
            d = drift(s, pair, cutneigh, skin, args.every, args.delay, args.drift_steps, prec)
            rec[tag] = {k: d[k] for k in ("steps", "diverged_at", "drift_eV_per_particle", "max_abs_eV_per_particle",
                                          "series_eV_per_particle")}
        a, b = rec["mixed"]["series_eV_per_particle"], rec["fp64"]["series_eV_per_particle"]
        rec["mixed_minus_fp64_max_eV_per_particle"] = max(abs(x - y) for x, y in zip(a, b))
        log(f"[bench] drift {name}: fp64 {rec['fp64']['drift_eV_per_particle']:+.3e}, mixed "
            f"{rec['mixed']['drift_eV_per_particle']:+.3e} eV/particle after {rec['fp64']['steps']} steps; "
            f"|mixed - fp64| <= {rec['mixed_minus_fp64_max_eV_per_particle']:.2e}")
        out[name] = rec
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--workload", choices=["eam_c1", "eam_c3", "lj_c0", "small"], default="eam_c1")
    ap.add_argument("--strain-rate", type=float, default=None,
                    help="uniaxial engineering strain rate along x, per second (eam_c3 default 1e8)")
    ap.add_argument("--seed", type=int, default=4928459)
    ap.add_argument("--every", type=int, default=1)
    ap.add_argument("--delay", type=int, default=10)
    ap.add_argument("--cpu-warmup", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-steps-1t", type=int, default=3, help="timed CPU steps of the single-thread sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-coupling", action="store_true",
                    help="atoms and elements do not interact (the reference's behaviour)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the mixed-precision and surrogate lines, accuracy and drift")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="only train the surrogate and time mixed precision + surrogate (profiling runs)")
    ap.add_argument("--surrogate-samples", type=int, default=200000)
    ap.add_argument("--surrogate-epochs", type=int, default=400)
    ap.add_argument("--surrogate-data", choices=["trajectory", "box"], default="box",
                    help="training data: snapshots of the workload's own plain-path trajectory under random "
                         "strains (velocity seed 1234), or periodic all-element boxes with per-node noise")
    ap.add_argument("--error-steps", default="0,100,1000")
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()

    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copy_probe_bytes_written": b.numel() * 8}))
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the library has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    from cac_code_amd import _lib
    global COUPLED, STRAIN_RATE_PS
    COUPLED = not args.no_coupling
    rate = args.strain_rate if args.strain_rate is not None else (1e8 if args.workload == "eam_c3" else 0.0)
    STRAIN_RATE_PS = rate * 1e-12                       # metal units: time in ps
    t0 = time.perf_counter()
    s = make_workload(args.workload, args.seed)
    pair, cutneigh, skin, ff = make_pair(args.workload)
    log(f"[bench] workload {args.workload}: atoms={s.n_atoms} elements={s.n_elements} "
        f"(built in {time.perf_counter() - t0:.1f}s)")
    prec = _lib.PREC_FP64 if args.precision == "fp64" else _lib.PREC_MIXED
    if args.surrogate_only:
        n_int = s.n_atoms + 8 * s.n_elements
        _, rec = surrogate_variant(s, pair, cutneigh, skin, args, n_int, float("nan"), surrogate_data(s, pair, args))
        print(json.dumps({"surrogate": rec}), flush=True)
        return
    eng, secs, nbuilds, nbw, prof = gpu_run(s, pair, cutneigh, skin, args.steps, args.warmup, prec, args.every,
                                            args.delay)
    centro = centro_leg(eng) if s.n_atoms else None
    eng.close()
    n_int = s.n_atoms + 8 * s.n_elements
    n_rep = s.n_atoms + s.nae[0] * s.nae[1] * s.nae[2] * s.n_elements
    value = n_int * args.steps / secs / 1e6
    log(f"[bench] {args.precision}: {secs * 1e3 / args.steps:.3f} ms/step, {value:.1f} Matom-steps/s, "
        f"{nbuilds} rebuilds in {args.steps} steps")
    for k, (ms, c) in prof.items():
        log(f"[bench]   {k:12s} {ms * 1e3:9.1f} us x {c}")
    if centro:
        log(f"[bench] centro-symmetry: {centro['kernel_ms']} ms for {centro['atoms']} atoms, "
            f"{centro['defect_atoms']} in the dump window {centro['window']}")

    # dominant kernel by total time among those with an algorithmic byte count
    dom = max((k for k in prof if k in ALG_BYTES), key=lambda k: prof[k][0] * prof[k][1], default=None)
    roof = None
    if dom:
        unit, per = ALG_BYTES[dom]
        units = s.n_atoms if unit == "atoms" else s.n_elements
        ms = prof[dom][0]
        ach = units * per / (ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": TRAFFIC.get(dom) if (args.workload == "eam_c1" and args.precision == "fp64") else None,
                "traffic_source": TRAFFIC_SOURCE if (args.workload == "eam_c1" and args.precision == "fp64") else None,
                "alg_bytes_per_launch": units * per, "avg_launch_us": round(ms * 1e3, 2)}
    b_step = B_ATOM * s.n_atoms + B_ELEM * s.n_elements
    step_roof = {"alg_bytes_per_step": b_step,
                 "achieved_GBs": round(b_step * args.steps / secs / 1e9, 1),
                 "frac": round(b_step * args.steps / secs / 1e9 / HBM_PEAK_GBS, 4)}

    variants, acc = {}, {}
    if not args.no_variants and args.precision == "fp64":
        eng2, secs2, nb2, _, prof2 = gpu_run(s, pair, cutneigh, skin, args.steps, args.warmup, _lib.PREC_MIXED,
                                             args.every, args.delay)
        eng2.close()
        variants["mixed"] = {"value": round(n_int * args.steps / secs2 / 1e6, 2),
                             "ms_per_step": round(secs2 * 1e3 / args.steps, 4), "rebuilds": nb2,
                             "kernels_us": {k: round(v[0] * 1e3, 1) for k, v in prof2.items()}}
        log(f"[bench] mixed: {secs2 * 1e3 / args.steps:.3f} ms/step")
        models = {}
        if args.workload == "eam_c1" and not args.no_surrogate:
            data = surrogate_data(s, pair, args)
            models["surrogate"], variants["surrogate"] = surrogate_variant(s, pair, cutneigh, skin, args, n_int,
                                                                           secs, data)
            models["surrogate_config"], variants["surrogate_config"] = surrogate_variant(
                s, pair, cutneigh, skin, args, n_int, secs, data, config=True)
        steps = [int(x) for x in args.error_steps.split(",") if x.strip()]
        acc["forces"] = accuracy(s, pair, cutneigh, skin, args.every, args.delay, steps, models)
        if args.drift_steps:
            acc["drift_fp64"] = drift(s, pair, cutneigh, skin, args.every, args.delay, args.drift_steps,
                                      _lib.PREC_FP64)
            acc["drift_mixed"] = drift(s, pair, cutneigh, skin, args.every, args.delay, args.drift_steps,
                                       _lib.PREC_MIXED)
            for k in ("drift_fp64", "drift_mixed"):
                d = acc[k]
                end = (f"diverged (|dE| > {d['limit_eV_per_particle']} eV/particle) at step {d['diverged_at']}"
                       if d["diverged_at"] else f"after {d['steps']} steps")
                log(f"[bench] {k}: {d['drift_eV_per_particle']:+.3e} eV/particle, {end}")
            if args.workload == "eam_c1":
                acc["drift_regions"] = drift_regions(pair, cutneigh, skin, args)

    cpu = None
    if not args.no_cpu:
        rate = nbuilds / max(args.steps, 1)
        host = host_info()
        nthr = host["affinity"] or 1
        if host["omp_num_threads"]:
            nthr = min(nthr, int(host["omp_num_threads"]))
        log(f"[bench] cpu baseline (oracle): {args.cpu_warmup}+{args.cpu_steps} steps on {nthr} threads ...")
        cn = cpu_baseline(s, pair, ff, cutneigh, nthr, args.cpu_warmup, args.cpu_steps, rate)
        log(f"[bench] cpu: {cn['per_step']:.3f} s/step on {cn['cores']} threads; 1+{args.cpu_steps_1t} steps on 1 "
            f"thread ...")
        log("[bench] full-size parity of the device's first force call against the oracle's ...")
        parity = full_size_parity(s, pair, ff, cutneigh, skin, args.every, args.delay, cn.pop("state"))
        log(f"[bench] parity: {json.dumps(parity)}")
        c1 = cpu_baseline(s, pair, ff, cutneigh, 1, 1, args.cpu_steps_1t, rate)
        c1.pop("state")
        cpu = {"value": round(n_int / cn["per_step"] / 1e6, 4), "unit": "Matom-steps/s", "cores": cn["cores"],
               "kind": "port",
               "sample": f"{args.cpu_warmup} warm-up + {cn['steps']} timed full FP64 steps of the same system on "
                         f"the oracle ({cn['t_step']:.3f} s/step) plus one neighbor build ({cn['t_build']:.2f} s) "
                         f"charged at the GPU rebuild rate {rate:.4f}/step",
               "single_thread": {"value": round(n_int / c1["per_step"] / 1e6, 4), "t_step": round(c1["t_step"], 3),
                                 "t_build": round(c1["t_build"], 2),
                                 "sample": f"1 warm-up + {c1['steps']} timed steps, one neighbor build"},
               "coupling": COUPLED,
               "parity_full_size": parity,
               "host": host}
        log(f"[bench] cpu: {cn['per_step']:.3f} s/step on {cn['cores']} threads, {c1['per_step']:.3f} s/step on 1")

    out = {
        "metric": "Matom-steps/s on 1 MI355X (integrated particles: atoms + 8 x elements)",
        "value": round(value, 2),
        "unit": "Matom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-pair/f64-accum",
        "data": "synthetic",
        "config": {"workload": args.workload, "atoms": s.n_atoms, "elements": s.n_elements,
                   "nae": list(s.nae), "integration_points": int(s.intp_w.size), "box_cells": s.cells,
                   "represented_atoms": n_rep, "rebuilds": nbuilds, "neighbor": [cutneigh, skin],
                   "neigh_modify": {"every": args.every, "delay": args.delay, "check": "yes"},
                   "coupling": COUPLED, "strain_rate_per_s": rate,
                   "seed": args.seed, "parallelism": "single GPU"},
        "represented_Matom_steps_per_s": round(n_rep * args.steps / secs / 1e6, 1),
        "roofline": roof,
        "step_roofline": step_roof,
        "centro": centro,
        "variants": variants,
        "accuracy": acc,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def surrogate_data(s, pair, args):
    """Training set (dataset seed 1234) and held-out set (seed 77) of the
    element surrogate, from the plain integration-point path."""
    from cac_code_amd import surrogate as S
    t0 = time.perf_counter()
    if args.surrogate_data == "trajectory":
        system_fn = lambda sd: make_workload(args.workload, sd)
        ds = S.generate_dataset_trajectory(pair, system_fn, args.surrogate_samples, seed=1234, coupled=COUPLED,
                                           log=log)
        held = S.generate_dataset_trajectory(pair, system_fn, 20000, seed=77, coupled=COUPLED)
    else:
        ds = S.generate_dataset(pair, args.surrogate_samples, seed=1234, log=log)
        held = S.generate_dataset(pair, 20000, seed=77)
    return ds, held, time.perf_counter() - t0


def surrogate_variant(s, pair, cutneigh, skin, args, n_int, secs_fp64, data, config=False):
    """Train the element surrogate (init seed 0, Adam lr 1e-3) on the
    dataset, the structured 279-input one or (config) the configured
    102-input network, and time mixed precision + surrogate on the same
    workload."""
    import numpy as np
    from cac_code_amd import _lib
    from cac_code_amd import surrogate as S
    ds, held, t_data = data
    t0 = time.perf_counter()
    if config:
        model = S.train_config(ds, epochs=args.surrogate_epochs, batch=1024, lr=1e-3, seed=0, log=log)
        held_fit = None
    else:
        model = S.train(ds, epochs=args.surrogate_epochs, batch=1024, lr=1e-3, seed=0, log=log)
        held_fit = S.rel_rms(S.predict(model, held.x, structured_only=True), held.t)
    held_err = S.rel_rms(S.predict(model, held.x), held.t)
    t_train = time.perf_counter() - t0 + t_data
    name = "configured 102-input network" if config else "structured 279-input surrogate"
    log(f"[bench] {name} trained in {t_train:.0f}s: held-out rel RMS {held_err:.5f}"
        + (f" (least-squares part alone {held_fit:.5f})" if held_fit is not None else ""))
    eng = new_engine(s, pair, cutneigh, skin, args.every, args.delay)
    info = S.attach(eng, model, s)
    eng.force(precision=_lib.PREC_MIXED)
    nb_w = eng.run(args.warmup, precision=_lib.PREC_MIXED) if args.warmup else 0
    eng.synchronize()
    t1 = time.perf_counter()
    nb = eng.run(args.steps, precision=_lib.PREC_MIXED)
    eng.synchronize()
    secs = time.perf_counter() - t1
    prof = {k: round(v[0] * 1e3, 1) for k, v in kernel_profile(eng, min(args.steps, 100), _lib.PREC_MIXED).items()}
    info = eng.surrogate_info()
    eng.close()
    ne = s.n_elements
    f_fb = 1.0 - info["surrogate"] / ne
    b_step = B_ATOM * s.n_atoms + ne * ((1.0 - f_fb) * 2568 + f_fb * B_ELEM)
    mlp_flop = 2.0 * (model.nin * S.NHID + (S.NLAYERS - 1) * S.NHID * S.NHID + S.NHID * S.NOUT) * info["surrogate"]
    mlp_roof = None
    if "surr_mlp" in prof:
        tf = mlp_flop / (prof["surr_mlp"] * 1e-6) / 1e12
        mlp_roof = {"bound": "mfma", "kernel": "surr_mlp (3 hidden + output GEMMs, BF16 MFMA)",
                    "achieved": round(tf, 1), "peak": 2500.0, "unit": "TFLOP/s", "frac": round(tf / 2500.0, 4),
                    "flop_per_launch": mlp_flop}
    log(f"[bench] mixed+{'configured network' if config else 'surrogate'}: {secs * 1e3 / args.steps:.3f} ms/step, "
        f"fallback fraction {f_fb:.3f}, {prof}")
    return model, {"value": round(n_int * args.steps / secs / 1e6, 2),
                   "ms_per_step": round(secs * 1e3 / args.steps, 4),
                   "speedup_vs_fp64": round(secs_fp64 / secs, 3), "rebuilds": nb,
                   "fallback_fraction": round(f_fb, 4), **info,
                   "alg_bytes_per_step": b_step, "kernels_us": prof, "mlp_roofline": mlp_roof,
                   "training": {"samples": args.surrogate_samples, "epochs": args.surrogate_epochs,
                                "data": args.surrogate_data, "dataset": {k: v for k, v in ds.meta.items()
                                                                         if not isinstance(v, np.ndarray)},
                                "dataset_seed": 1234, "init_seed": 0, "lr": 1e-3, "seconds": round(t_train, 1),
                                "heldout_rel_rms": round(held_err, 6),
                                "heldout_least_squares_only": None if held_fit is None else round(held_fit, 6)},
                   "network": (f"MLP {model.nin}->256->256->256->24 SiLU BF16"
                               + ("" if config else " + FP32 linear and corner-local quadratic least-squares part"))}


if __name__ == "__main__":
    main()
