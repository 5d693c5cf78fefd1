"""Benchmark of the Rys ERI + J/K (Fock) build on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-grad]

Workload (one "step" = one J+K build over the screened canonical quartet list
with the density resident in HBM): the seeded 16-water cluster in cc-pVDZ
(400 Cartesian functions; O-H 1.81 bohr, HOH 104.5 deg, random orientations,
14 bohr cube, O-O >= 5 bohr), D = C C^T with C 400x80 ~ N(0, 1/384), Schwarz
threshold 1e-12, FP64.  The build is captured once in a hipGraph and replayed.
Also reported: the same build as one persistent launch for the register-shape
classes; FP64 and mixed-precision builds of the linear alkane C40H82 in
def2-SVP (FP32 recursion for quartets with Q_ij Q_kl max|D| < 1e-7) with the
errors of J, K and E against the FP64 build; the FP64 build of 64 seeded
cc-pVTZ molecules in one BatchEngine (and comparisons on its 23 smallest);
the nuclear-gradient time for the 32-water def2-SVP cluster; density-fitted
J/K; the per-class roofline of the dominant kernel and a CPU baseline timed
on a bounded sample with the reference CPU sources (oracle/_ref) or, where
those are not built, their C restatement.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "libaccint-public_amd"))

FP64_PEAK_TFLOPS = 78.6  # MI355X vector / matrix FP64
METRIC = "J+K build s & ERIs/s, FP64 vs mixed at |dE|<=1e-6 Eh; gradient s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_of(cls):
    """HBM bytes per launch of the class kernel from the committed FETCH_SIZE /
    WRITE_SIZE counter runs (profiles/), or None when that class was not measured."""
    path = os.path.join(REPO, "profiles", "eri_jk_16water_ccpvdz_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    import re
    for line in open(path):
        m = re.match(r"traffic_bytes_per_launch (\(\w\w\|\w\w\)) = ([0-9.eE+-]+)", line.strip())
        if m and m.group(1) == cls:
            return float(m.group(2))
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(threads):
    """The CPU baseline, run in a child process so that OMP_PROC_BIND and
    OMP_NUM_THREADS reach the OpenMP runtime before it starts (the parent has
    torch's runtime loaded).  Times, with the compiled reference CPU sources
    (oracle/_ref) where they are built, else their C restatement:
      * the canonical screened loop (thr 1e-12) on every 12th bra shell pair
        of the 16-water cc-pVDZ workload;
      * the reference's default path, the unscreened N^4 loop of
        CpuEngine::compute_and_consume_parallel, on four first-shell slices of
        the same workload (projected to all 192 slices), and in full on
        H2O/STO-3G.
    Prints one JSON object."""
    import oracle
    from libaccint import systems
    use_ref = oracle.ref_lib() is not None
    atoms = systems.water_cluster(16, 14.0, seed=2)
    rb = oracle.RawBasis("cc-pvdz", atoms)
    o = oracle.Oracle(rb, use_ref=use_ref)
    D = systems.density(rb.nbf, 80, 5, 1.0 / math.sqrt(384.0))
    Q = o.schwarz(n_threads=threads)
    stride, offset = 12, 7
    t = time.perf_counter()
    _, _, nq, ne = o.jk_canonical_sample(D, Q, 1e-12, stride, offset, n_threads=threads)
    dt = time.perf_counter() - t
    # N^4 default path: a seeded sample of first-shell pairs (i, j), each
    # with all n_shells^2 (k, l)
    ns = rb.n_shells
    ij = np.random.default_rng(17).integers(0, ns, size=(24, 2)).astype(np.int32)
    dt4 = None
    if use_ref:
        t = time.perf_counter()
        o.jk_full_pairs(D, ij, n_threads=threads)
        dt4 = time.perf_counter() - t
    h2o = oracle.Oracle(oracle.RawBasis("sto-3g", systems.h2o_sto3g_geometry()), use_ref=use_ref)
    rng = np.random.default_rng(1)
    C = rng.standard_normal((7, 5)) / np.sqrt(7)
    reps = 200
    t = time.perf_counter()
    for _ in range(reps):
        h2o.jk_full(C @ C.T, n_threads=threads)
    dt1 = (time.perf_counter() - t) / reps
    out = {"kind": "reference" if use_ref else "port", "threads": threads, "cpu_model": cpu_model(),
           "canonical_sample": {"stride": stride, "offset": offset, "quartets": nq, "eris": ne, "seconds": dt},
           "n4_sample": {"first_shell_pairs": len(ij), "shell_quartets": len(ij) * ns ** 2, "seconds": dt4,
                         "projected_full_seconds": None if dt4 is None else dt4 * ns * ns / len(ij)},
           "h2o_sto3g_n4_seconds": dt1}
    print(json.dumps(out), flush=True)


def class_name(c):
    names = "spdf"
    pb = int((math.isqrt(8 * c + 1) - 1) // 2)
    pk = c - pb * (pb + 1) // 2

    def pc(p):
        la = int((math.isqrt(8 * p + 1) - 1) // 2)
        return la, p - la * (la + 1) // 2
    a, b = pc(pb)
    cc, d = pc(pk)
    return "(%s%s|%s%s)" % (names[a], names[b], names[cc], names[d])


def mixed_probe(lai, _cabi, systems, torch, stream, L):
    """C40H82 def2-SVP: per-class kernel times of the FP64 build and of the same
    quartets all through the float instantiation (tau above every bound, no
    primitive cut), one class after another (lai_dev_jk_profile)."""
    import ctypes
    atoms = systems.alkane(40)
    basis = lai.basis_set("def2-svp", atoms)
    n = basis.n_basis_functions()
    D = systems.density(n, 161, 11, 1.0 / math.sqrt(n))
    eng = lai.Engine(basis)
    h = ctypes.c_void_p(eng.device_basis_handle())
    tD = torch.from_numpy(D).cuda()
    tJ = torch.zeros_like(tD)
    tK = torch.zeros_like(tD)
    dp = ctypes.POINTER(ctypes.c_double)
    out = {}
    for tag, o, flags in (("fp64", lai.JKOptions(schwarz_threshold=1e-12), 0),
                          ("all_fp32", lai.JKOptions(schwarz_threshold=1e-12, mixed_precision=True,
                                                     fp32_threshold=1e30, fp32_prim_cut=0.0), _cabi.LAI_JK_MIXED)):
        eng.prepare_worklist(o, D if o.mixed_precision else None)
        _cabi.check(L.lai_dev_set_mixed_prim_cut(h, ctypes.c_double(o.fp32_prim_cut)))
        a64, a32 = np.zeros(55), np.zeros(55)
        for _ in range(2):
            _cabi.check(L.lai_dev_jk_profile(h, ctypes.c_void_p(tD.data_ptr()), ctypes.c_void_p(tJ.data_ptr()),
                                             ctypes.c_void_p(tK.data_ptr()), flags, ctypes.c_void_p(stream.cuda_stream),
                                             a64.ctypes.data_as(dp), a32.ctypes.data_as(dp)))
        out[tag] = a64 + a32
    print(json.dumps({"workload": "C40H82 def2-SVP (1010 Cartesian bf), Schwarz 1e-12: per-class J/K kernel ms",
                      "class_ms_fp64": float(out["fp64"].sum()), "class_ms_all_fp32": float(out["all_fp32"].sum()),
                      "classes": {class_name(c): [round(float(out["fp64"][c]), 3), round(float(out["all_fp32"][c]), 3)]
                                  for c in np.argsort(-out["fp64"]) if out["fp64"][c] > 0.05}}), flush=True)


def main():
    if len(sys.argv) > 2 and sys.argv[1] == "--cpu-leg":
        cpu_leg(int(sys.argv[2]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-df", action="store_true")
    ap.add_argument("--no-scf", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="no autotuner run in the setup (the default choices, or the table named by "
                         "LIBACCINT_DISPATCH_TABLE)")
    ap.add_argument("--save-table", default="",
                    help="write the 16-water build's dispatch table (JSON) to this path")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="end with a device fill and copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--mixed-probe", action="store_true",
                    help="only the per-class FP64 and all-FP32 profiles of the C40H82 def2-SVP build (one JSON "
                         "line; for counter runs)")
    ap.add_argument("--no-batch-full", action="store_true",
                    help="skip the FP64 build of all 64 molecules of the batched workload (about 100 s)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this library has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    import libaccint as lai
    from libaccint import _cabi, systems
    import ctypes

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    L = _cabi.lib()
    if args.mixed_probe:
        mixed_probe(lai, _cabi, systems, torch, stream, L)
        return

    # ---------------------------------------------------------------- setup
    atoms = systems.water_cluster(16, 14.0, seed=2)
    basis = lai.basis_set("cc-pvdz", atoms)
    n = basis.n_basis_functions()
    D = systems.density(n, 80, 5, 1.0 / math.sqrt(384.0))
    engine = lai.Engine(basis)
    t0 = time.perf_counter()
    opts = lai.JKOptions(schwarz_threshold=1e-12)
    info = engine.prepare_worklist(opts)
    log("setup (Schwarz + worklist): %.2f s, %d quartets, %.3e ERIs" %
        (time.perf_counter() - t0, info["n_quartets"], info["n_eri"]))
    hb = ctypes.c_void_p(engine.device_basis_handle())
    tD = torch.from_numpy(D).to(dev)
    tJ = torch.zeros((n, n), dtype=torch.float64, device=dev)
    tK = torch.zeros_like(tJ)
    torch.cuda.synchronize()

    def timed_graph(o, steps, warmup):
        engine.capture_jk_graph(tD.data_ptr(), tJ.data_ptr(), tK.data_ptr(), o, stream.cuda_stream)
        for _ in range(warmup):
            engine.launch_jk_graph(stream.cuda_stream)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            engine.launch_jk_graph(stream.cuda_stream)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        return ev0.elapsed_time(ev1) / steps, wall

    def jk_once(o):
        tJ.zero_()
        tK.zero_()
        torch.cuda.synchronize()
        engine.build_jk_device(tD.data_ptr(), tJ.data_ptr(), tK.data_ptr(), o, stream.cuda_stream, D)
        torch.cuda.synchronize()
        return tJ.cpu().numpy().copy(), tK.cpu().numpy().copy()

    # -------------------------------------------------------------- FP64 J+K
    J64, K64 = jk_once(opts)
    E64 = 0.5 * np.sum(D * J64) - 0.25 * np.sum(D * K64)
    tune = {}
    if not args.no_tune:
        # the default kernel choices, then the autotuner (setup, not timed)
        ms_untuned, _ = timed_graph(opts, max(2, args.steps // 2), 1)
        t0 = time.perf_counter()
        table = engine.autotune(tD.data_ptr(), opts, stream.cuda_stream, 3)
        t_tune = time.perf_counter() - t0
        Jt_, Kt_ = jk_once(opts)
        dev_tuned = max(float(np.max(np.abs(Jt_ - J64))) / float(np.max(np.abs(J64))),
                        float(np.max(np.abs(Kt_ - K64))) / float(np.max(np.abs(K64))))
        changed = [e_ for e_ in table.entries if e_.ms < e_.ms_default]
        tune = {"autotune_seconds": t_tune, "jk_seconds_fp64_untuned": ms_untuned * 1e-3,
                "autotune_classes_changed": len(changed), "autotune_max_rel_dev_jk": dev_tuned}
        log("autotune: %.2f s, %d of %d classes changed, untuned %.3f ms, tuned J/K rel. deviation %.1e" %
            (t_tune, len(changed), len(table.entries), ms_untuned, dev_tuned))
    if args.save_table:
        engine.dispatch_table().to_json(args.save_table)
    ms64, wall64 = timed_graph(opts, args.steps, args.warmup)
    log("FP64 J+K: %.3f ms/step (events), wall %.3f s for %d steps" % (ms64, wall64, args.steps))
    # the same build with the register-shape classes as one persistent launch
    # over a device queue of their chunks
    ms64p, _ = timed_graph(lai.JKOptions(schwarz_threshold=1e-12, persistent=True), args.steps, args.warmup)
    log("FP64 J+K, persistent register-shape launch: %.3f ms/step" % ms64p)

    # per-class profile of the same build (HIP events around each class kernel)
    f64c = np.zeros(55)
    f32c = np.zeros(55)
    _cabi.check(L.lai_dev_worklist_class_flops(hb, f64c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               f32c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    ms_cls = np.zeros(55)
    reps = 3
    for _ in range(reps):
        m64 = np.zeros(55)
        m32 = np.zeros(55)
        _cabi.check(L.lai_dev_jk_profile(hb, ctypes.c_void_p(tD.data_ptr()), ctypes.c_void_p(tJ.data_ptr()),
                                         ctypes.c_void_p(tK.data_ptr()), 0, ctypes.c_void_p(stream.cuda_stream),
                                         m64.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         m32.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        ms_cls += (m64 + m32) / reps
    dom = int(np.argmax(ms_cls))

    # executed work (primitive pairs above the screening bound) of the
    # dominant class kernel over its event-timed duration
    achieved = (f64c[dom] + f32c[dom]) / (ms_cls[dom] * 1e-3) / 1e12
    exec_total = float(np.sum(f64c) + np.sum(f32c))
    build_tflops = exec_total / (ms64 * 1e-3) / 1e12
    # the same with every non-zero primitive counted (the work model of the
    # reference's algorithm)
    flops_total = info["flops_fp64"] + info["flops_fp32"]
    build_tflops_model = flops_total / (ms64 * 1e-3) / 1e12
    per_class = sorted(((class_name(c), round(ms_cls[c], 3), round((f64c[c] + f32c[c]) / max(ms_cls[c], 1e-9) / 1e9, 2))
                        for c in range(55) if ms_cls[c] > 0), key=lambda x: -x[1])
    log("dominant class %s: %.3f ms, %.2f TFLOP/s; whole build %.2f TFLOP/s" %
        (class_name(dom), ms_cls[dom], achieved, build_tflops))
    log("per-class (name, ms, TFLOP/s x1e3): %s" % per_class[:12])

    # ---------------------------------------------------------------- mixed
    # FP64 vs mixed precision on the linear alkane C40H82 / def2-SVP
    mixed = {}
    if not args.no_mixed:
        matoms = systems.alkane(40)
        mbasis = lai.basis_set("def2-svp", matoms)
        mn = mbasis.n_basis_functions()
        mD = systems.density(mn, 161, 11, 1.0 / math.sqrt(mn))
        meng = lai.Engine(mbasis)
        t0 = time.perf_counter()
        o64 = lai.JKOptions(schwarz_threshold=1e-12)
        mi64 = meng.prepare_worklist(o64)
        log("C40H82 setup: %.2f s, %d quartets" % (time.perf_counter() - t0, mi64["n_quartets"]))
        tmD = torch.from_numpy(mD).to(dev)
        tmJ = torch.zeros((mn, mn), dtype=torch.float64, device=dev)
        tmK = torch.zeros_like(tmJ)

        def m_once(o):
            tmJ.zero_()
            tmK.zero_()
            torch.cuda.synchronize()
            meng.build_jk_device(tmD.data_ptr(), tmJ.data_ptr(), tmK.data_ptr(), o, stream.cuda_stream, mD)
            torch.cuda.synchronize()
            return tmJ.cpu().numpy().copy(), tmK.cpu().numpy().copy()

        def m_timed(o, steps, warmup):
            meng.capture_jk_graph(tmD.data_ptr(), tmJ.data_ptr(), tmK.data_ptr(), o, stream.cuda_stream)
            for _ in range(warmup):
                meng.launch_jk_graph(stream.cuda_stream)
            torch.cuda.synchronize()
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for _ in range(steps):
                meng.launch_jk_graph(stream.cuda_stream)
            ev1.record(stream)
            torch.cuda.synchronize()
            return ev0.elapsed_time(ev1) / steps

        msteps = max(2, args.steps // 2)
        Jd, Kd = m_once(o64)
        Ed = 0.5 * np.sum(mD * Jd) - 0.25 * np.sum(mD * Kd)
        if not args.no_tune:
            meng.autotune(tmD.data_ptr(), o64, stream.cuda_stream, 2)
        ms_d = m_timed(o64, msteps, 1)
        om = lai.JKOptions(schwarz_threshold=1e-12, mixed_precision=True, fp32_threshold=1e-7)
        meng.prepare_worklist(om, mD)
        mim = meng.worklist_info()
        if not args.no_tune:
            meng.autotune(tmD.data_ptr(), om, stream.cuda_stream, 2, mD)
        Jm, Km = m_once(om)
        Em = 0.5 * np.sum(mD * Jm) - 0.25 * np.sum(mD * Km)
        ms_m = m_timed(om, msteps, 1)
        # the mixed build's FP64 and FP32 class launches timed one after
        # another, with the executed flops of each precision
        hm = ctypes.c_void_p(meng.device_basis_handle())
        dp = ctypes.POINTER(ctypes.c_double)
        fm64, fm32, pm64, pm32 = np.zeros(55), np.zeros(55), np.zeros(55), np.zeros(55)
        _cabi.check(L.lai_dev_worklist_class_flops(hm, fm64.ctypes.data_as(dp), fm32.ctypes.data_as(dp)))
        _cabi.check(L.lai_dev_jk_profile(hm, ctypes.c_void_p(tmD.data_ptr()), ctypes.c_void_p(tmJ.data_ptr()),
                                         ctypes.c_void_p(tmK.data_ptr()), _cabi.LAI_JK_MIXED,
                                         ctypes.c_void_p(stream.cuda_stream), pm64.ctypes.data_as(dp),
                                         pm32.ctypes.data_as(dp)))
        # the FP32-marked quartets run in float only in the wavefront and
        # workgroup shapes (the lane-per-quartet classes run them in FP64)
        tab = (_cabi.ClassTune * 55)()
        ntab = ctypes.c_int()
        _cabi.check(L.lai_dev_get_tuning(hm, tab, 55, ctypes.byref(ntab)))
        in_float = np.zeros(55, dtype=bool)
        for k in range(ntab.value):
            in_float[tab[k].cls] = tab[k].strategy != 0
        f_float = float(fm32[in_float].sum())
        f_all = float((fm64 + fm32).sum())
        split = {"class_ms_fp64_launches": float(pm64.sum()), "class_ms_fp32_launches": float(pm32.sum()),
                 "tflops_fp64_launches": (f_all - f_float) / max(float(pm64.sum()), 1e-9) / 1e9,
                 "tflops_fp32_launches": f_float / max(float(pm32.sum()), 1e-9) / 1e9,
                 "flop_fraction_in_float": f_float / max(f_all, 1.0)}
        log("C40H82 mixed: FP64 launches %.3f ms at %.2f TFLOP/s, FP32 launches %.3f ms at %.2f TFLOP/s, %.3f of "
            "the flops in float" % (split["class_ms_fp64_launches"], split["tflops_fp64_launches"],
                                    split["class_ms_fp32_launches"], split["tflops_fp32_launches"],
                                    split["flop_fraction_in_float"]))
        # the same FP32 split without the FP32 primitive cut
        on = lai.JKOptions(schwarz_threshold=1e-12, mixed_precision=True, fp32_threshold=1e-7, fp32_prim_cut=0.0)
        Jn, Kn = m_once(on)
        En = 0.5 * np.sum(mD * Jn) - 0.25 * np.sum(mD * Kn)
        ms_n = m_timed(on, msteps, 1)
        # every quartet through the float instantiation (tau above every bound,
        # no primitive cut) against the FP64 build, class by class: what the
        # FP32 recursion alone buys
        def m_profile(o, flags):
            meng.prepare_worklist(o, mD if o.mixed_precision else None)
            _cabi.check(L.lai_dev_set_mixed_prim_cut(hm, ctypes.c_double(o.fp32_prim_cut)))
            a64, a32 = np.zeros(55), np.zeros(55)
            for _ in range(2):
                _cabi.check(L.lai_dev_jk_profile(hm, ctypes.c_void_p(tmD.data_ptr()), ctypes.c_void_p(tmJ.data_ptr()),
                                                 ctypes.c_void_p(tmK.data_ptr()), flags,
                                                 ctypes.c_void_p(stream.cuda_stream), a64.ctypes.data_as(dp),
                                                 a32.ctypes.data_as(dp)))
            return a64 + a32
        p64 = m_profile(o64, 0)
        p32 = m_profile(lai.JKOptions(schwarz_threshold=1e-12, mixed_precision=True, fp32_threshold=1e30,
                                      fp32_prim_cut=0.0), _cabi.LAI_JK_MIXED)
        split["all_fp32_vs_fp64"] = {
            "class_ms_fp64": float(p64.sum()), "class_ms_all_fp32": float(p32.sum()),
            "top_classes_ms_fp64_fp32": [(class_name(c), round(float(p64[c]), 3), round(float(p32[c]), 3))
                                         for c in np.argsort(-p64)[:8]]}
        log("C40H82 all quartets FP32: class kernels %.3f ms against %.3f ms FP64" % (p32.sum(), p64.sum()))
        mixed = {"mixed_workload": "C40H82 def2-SVP (1010 Cartesian bf), Schwarz 1e-12, %d quartets; FP32 Rys "
                                   "recursion where Q_ij Q_kl max|D| < 1e-7" % mi64["n_quartets"],
                 "jk_seconds_fp64_c40h82": ms_d * 1e-3, "jk_seconds_mixed": ms_m * 1e-3,
                 "mixed_speedup": ms_d / ms_m,
                 "mixed_fp32_quartet_fraction": mim["n_fp32"] / max(mim["n_quartets"], 1),
                 "dE_mixed_Eh": float(Em - Ed), "dJ_mixed_max": float(np.max(np.abs(Jm - Jd))),
                 "dK_mixed_max": float(np.max(np.abs(Km - Kd))), "fp32_threshold": 1e-7,
                 "fp32_prim_cut": 1e-7, "jk_seconds_mixed_no_prim_cut": ms_n * 1e-3,
                 "dE_mixed_no_prim_cut_Eh": float(En - Ed), "mixed_class_split": split}
        log("C40H82: FP64 %.3f ms, mixed %.3f ms, speedup %.2fx, dE %.2e Eh, fp32 fraction %.3f" %
            (ms_d, ms_m, ms_d / ms_m, Em - Ed, mixed["mixed_fp32_quartet_fraction"]))
        del meng, tmD, tmJ, tmK

    # -------------------------------------------------------------- batched
    # 64 seeded molecules (3-60 atoms of H/C/N/O, random-walk bonding) in
    # cc-pVTZ through one BatchEngine: the FP64 build of all 64 (one replay of
    # the captured build), and on the 23 molecules with at most 20 atoms the
    # comparisons -- persistent launch, mixed precision, one engine per
    # molecule.  No autotuner run: the default launch choices.
    batch = {}
    if not args.no_batch:
        rng = np.random.default_rng(4)
        mols = [systems.random_molecule(rng) for _ in range(64)]

        def batch_setup(sel):
            bases = [lai.basis_set("cc-pvtz", mols[k]) for k in sel]
            be = lai.BatchEngine(bases)
            bnbf = be.n_basis_functions()
            bD = np.concatenate([systems.density(nb_, max(2, nb_ // 5), 200 + k, 1.0 / math.sqrt(nb_)).ravel()
                                 for k, nb_ in zip(sel, bnbf)])
            t0 = time.perf_counter()
            binfo = be.prepare_worklist(lai.JKOptions(schwarz_threshold=1e-12))
            bsetup = time.perf_counter() - t0
            log("batch (%d molecules): setup %.2f s, %d quartets, %.3e executed flops" %
                (len(sel), bsetup, binfo["n_quartets"], binfo["flops_fp64"] + binfo["flops_fp32"]))
            tbD = torch.from_numpy(bD).to(dev)
            return bases, be, bnbf, bD, tbD, torch.zeros_like(tbD), torch.zeros_like(tbD), binfo, bsetup

        def b_timed(be, tbD, tbJ, tbK, o, steps, warm, what):
            be.capture_jk_graph(tbD.data_ptr(), tbJ.data_ptr(), tbK.data_ptr(), o, stream.cuda_stream)
            for _ in range(warm):
                be.launch_jk_graph(stream.cuda_stream)
            torch.cuda.synchronize()
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for _ in range(steps):
                be.launch_jk_graph(stream.cuda_stream)
            ev1.record(stream)
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / steps
            log("batch %s build: %.3f ms" % (what, ms))
            return ms

        bo = lai.JKOptions(schwarz_threshold=1e-12)
        sel = [k for k, m in enumerate(mols) if len(m) <= 20]
        bases, be, bnbf, bD, tbD, tbJ, tbK, binfo, bsetup = batch_setup(sel)
        ms_b = b_timed(be, tbD, tbJ, tbK, bo, 3, 1, "FP64")
        ms_bp = b_timed(be, tbD, tbJ, tbK, lai.JKOptions(schwarz_threshold=1e-12, persistent=True), 3, 1,
                        "FP64 persistent")
        bm = lai.JKOptions(schwarz_threshold=1e-12, mixed_precision=True, fp32_threshold=1e-7)
        be.prepare_worklist(bm, bD)
        bminfo = be.worklist_info()
        ms_bm = b_timed(be, tbD, tbJ, tbK, bm, 3, 1, "mixed")
        # the same molecules one engine at a time (no batching), FP64
        engines = [lai.Engine(bb) for bb in bases]
        offs = be.matrix_offsets()
        for e_ in engines:
            e_.prepare_worklist(bo)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for m_, e_ in enumerate(engines):
            a_, z_ = offs[m_], offs[m_ + 1]
            e_.build_jk_device(tbD[a_:z_].data_ptr(), tbJ[a_:z_].data_ptr(), tbK[a_:z_].data_ptr(), bo,
                               stream.cuda_stream, None)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms_seq = ev0.elapsed_time(ev1)
        batch = {"batch_subset_workload": "23 of the 64 seeded H/C/N/O molecules in cc-pVTZ (those with <= 20 "
                                          "atoms), nbf %d..%d, Schwarz 1e-12, %d quartets, %.3e ERIs" %
                                          (min(bnbf), max(bnbf), binfo["n_quartets"], binfo["n_eri"]),
                 "batch_subset_setup_seconds": bsetup,
                 "batch_subset_jk_seconds_fp64": ms_b * 1e-3,
                 "batch_subset_eris_per_s_fp64": binfo["n_eri"] / (ms_b * 1e-3),
                 "batch_subset_jk_seconds_fp64_persistent_launch": ms_bp * 1e-3,
                 "batch_subset_jk_seconds_mixed": ms_bm * 1e-3,
                 "batch_subset_mixed_fp32_quartet_fraction": bminfo["n_fp32"] / max(bminfo["n_quartets"], 1),
                 "batch_subset_one_engine_per_molecule_seconds_fp64": ms_seq * 1e-3}
        log("batch (23 molecules): FP64 %.3f ms, persistent %.3f ms, mixed %.3f ms, one engine per molecule "
            "%.3f ms" % (ms_b, ms_bp, ms_bm, ms_seq))
        del be, engines, tbD, tbJ, tbK
        if not args.no_batch_full:
            _, be, bnbf, bD, tbD, tbJ, tbK, binfo, bsetup = batch_setup(list(range(64)))
            ms_f = b_timed(be, tbD, tbJ, tbK, bo, 1, 0, "FP64, all 64 molecules,")
            flops_f = binfo["flops_fp64"] + binfo["flops_fp32"]
            batch.update({
                "batch_workload": "all 64 seeded H/C/N/O molecules (3-60 atoms) in cc-pVTZ, nbf %d..%d, Schwarz "
                                  "1e-12, %d quartets, %.3e ERIs, FP64, one BatchEngine" %
                                  (min(bnbf), max(bnbf), binfo["n_quartets"], binfo["n_eri"]),
                "batch_setup_seconds": bsetup,
                "batch_jk_seconds_fp64": ms_f * 1e-3,
                "batch_eris_per_s_fp64": binfo["n_eri"] / (ms_f * 1e-3),
                "batch_tflops_executed": flops_f / (ms_f * 1e-3) / 1e12})
            del be, tbD, tbJ, tbK

    # ------------------------------------------------------------------ SCF
    # RHF on the 16-water cc-pVDZ cluster: eager iterations (eigh, host DIIS)
    # against iterations replayed from one captured hipGraph
    scf_out = {}
    if not args.no_scf:
        from libaccint import scf as scf_mod
        res = {}
        for mode in (False, True):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = scf_mod.rhf(atoms, "cc-pvdz", 160, graph=mode, options=lai.JKOptions(schwarz_threshold=1e-12))
            res[mode] = (time.perf_counter() - t0, r)
        (te, re_), (tg, rg) = res[False], res[True]
        scf_out = {"scf": {"workload": "RHF, 16-water cluster cc-pVDZ (400 bf), J/K with Schwarz 1e-12, DIIS(6), "
                                       "converged to "
                                       "|dE| < 1e-12, max |dD| < 1e-10; wall time including set-up",
                           "eager_seconds": te, "eager_iterations": re_.iterations, "eager_energy": re_.energy,
                           "graph_seconds": tg, "graph_iterations": rg.iterations, "graph_energy": rg.energy}}
        log("RHF 16 waters: eager %.3f s (%d it), graph %.3f s (%d it), dE %.1e Eh" %
            (te, re_.iterations, tg, rg.iterations, rg.energy - re_.energy))

    # ------------------------------------------------------------- gradient
    grad = {}
    if not args.no_grad:
        gatoms = systems.water_cluster(32, 18.0, seed=4)
        gbasis = lai.basis_set("def2-svp", gatoms)
        gn = gbasis.n_basis_functions()
        gD = systems.density(gn, 160, 6, 1.0 / math.sqrt(768.0))
        geng = lai.Engine(gbasis)
        ginfo = geng.prepare_worklist(lai.JKOptions(schwarz_threshold=1e-12))
        g = geng.compute_jk_gradient(gD, 1.0)  # warmup
        reps = 2
        t = time.perf_counter()
        for _ in range(reps):
            g = geng.compute_jk_gradient(gD, 1.0)
        gs = (time.perf_counter() - t) / reps
        g0 = np.zeros(55)
        g1 = np.zeros(55)
        _cabi.check(L.lai_dev_worklist_grad_flops(ctypes.c_void_p(geng.device_basis_handle()),
                                                  g0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                  g1.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        F0, F1 = float(g0.sum()), float(g1.sum())
        grad = {"gradient_seconds": gs, "gradient_workload": "32-water def2-SVP (800 Cartesian bf), Schwarz 1e-12, "
                "%d quartets, dE/dR for E = 1/2 tr(DJ) - 1/4 tr(DK)" % ginfo["n_quartets"],
                "gradient_translational_residual": float(np.max(np.abs(g.sum(axis=0)))),
                "gradient_roofline": {
                    "flops_derivative_classes": F0, "tflops_derivative_classes": F0 / gs / 1e12,
                    "frac_derivative_classes": F0 / gs / 1e12 / FP64_PEAK_TFLOPS,
                    "flops_raised_tables": F1, "tflops_raised_tables": F1 / gs / 1e12,
                    "frac_raised_tables": F1 / gs / 1e12 / FP64_PEAK_TFLOPS, "peak": FP64_PEAK_TFLOPS,
                    "units": "executed primitive pairs; derivative classes: per centre (3 of 4) F_prim of the "
                             "class with that l raised and with it lowered, + 9 N_cart per root; raised tables: "
                             "30 + n_r (50 + 3 F_dir(La+1, Lb+1, Lc+1, Ld) + 12 N_cart)"}}
        log("gradient (32-water def2-SVP): %.3f s, %.2f TFLOP/s (derivative classes) / %.2f (raised tables)" %
            (gs, F0 / gs / 1e12, F1 / gs / 1e12))

    # ------------------------------------------------------ density fitting
    df = {}
    if not args.no_df:
        t = time.perf_counter()
        fb = lai.make_df_fock_builder(basis, [z for z, _ in atoms], [list(p) for _, p in atoms], "cc-pVDZ-RI")
        fb.set_density(D)
        fb.initialize()
        init_s = time.perf_counter() - t
        hdf = ctypes.c_void_p(fb.device_handle())
        # J, K and, D = C C^T being symmetric, K from its upper part (LAI_DF_SYMMETRIC)
        DF_FLAGS = 1 | 2 | 4
        naux = fb.n_aux()
        dJ = torch.zeros_like(tJ)
        dK = torch.zeros_like(tK)
        ms = np.zeros(5)
        P = ctypes.POINTER(ctypes.c_double)
        rows = []
        for it in range(1 + max(3, args.steps // 4)):
            _cabi.check(L.lai_dev_df_jk_profile(hdf, ctypes.c_void_p(tD.data_ptr()), ctypes.c_void_p(dJ.data_ptr()),
                                                ctypes.c_void_p(dK.data_ptr()), DF_FLAGS,
                                                ctypes.c_void_p(stream.cuda_stream), ms.ctypes.data_as(P)))
            if it:
                rows.append(ms.copy())
        km = np.median(np.array(rows), axis=0)
        dsteps = max(3, args.steps // 2)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            e0.record(stream)
            for _ in range(dsteps):
                _cabi.check(L.lai_dev_df_jk(hdf, ctypes.c_void_p(tD.data_ptr()), ctypes.c_void_p(dJ.data_ptr()),
                                            ctypes.c_void_p(dK.data_ptr()), DF_FLAGS,
                                            ctypes.c_void_p(stream.cuda_stream)))
            e1.record(stream)
        stream.synchronize()
        df_ms = e0.elapsed_time(e1) / dsteps
        Jd = dJ.cpu().numpy()
        Jx, _ = engine.build_jk(D, opts)  # FP64 four-centre J
        ej_df, ej = 0.5 * float(np.sum(D * Jd)), 0.5 * float(np.sum(D * Jx))
        b_bytes = float(n) * naux * n * 8.0
        gemm_flops = 2.0 * n * naux * n * n
        # share of the exchange product's 64 x 80 tiles of K computed with a
        # symmetric D (tiles entirely below the diagonal are skipped)
        tm, tn = (128 if n >= 1024 else 64), 80
        tiles = [(mi, ni) for mi in range((n + tm - 1) // tm) for ni in range((n + tn - 1) // tn)]
        k_share = sum(1 for mi, ni in tiles if not mi * tm > ni * tn + tn - 1) / len(tiles)
        df = {"df": {
            "workload": "16-water cc-pVDZ (400 bf) / cc-pVDZ-RI (%d fitting functions), FP64 J+K from HBM-resident B"
                        % naux,
            "init_seconds": init_s, "jk_ms": df_ms,
            # J and K together: gamma_P = sum_a (B_P D)_aa from the diagonal of X
            "kernels_ms": {"gamma_from_X_diagonal": km[0], "coulomb_pass": km[1], "gemm_X_eq_BD": km[2],
                           "gemm_K_eq_XBt_slices": km[3], "slice_sum": km[4]},
            "density_symmetric": True,
            "roofline": {"gemm_X_tflops": gemm_flops / km[2] / 1e9,
                         "gemm_K_tflops_algorithmic": gemm_flops / km[3] / 1e9,
                         "gemm_K_tflops_executed": k_share * gemm_flops / km[3] / 1e9,
                         "gemm_K_tile_share": k_share,
                         "gemm_peak_tflops": FP64_PEAK_TFLOPS,
                         "coulomb_pass_GBs": b_bytes / km[1] / 1e6,
                         "hbm_peak_GBs": 8000.0,
                         "units": "gemm: 2 n^2 n_aux n flops each; Coulomb pass: one read of B (n^2 n_aux doubles)"},
            "E_J_df_minus_exact": ej_df - ej}}
        log("DF J+K (16 waters, cc-pVDZ-RI): %.3f ms, init %.2f s; X %.1f TF/s, K %.1f TF/s executed" %
            (df_ms, init_s, gemm_flops / km[2] / 1e9, k_share * gemm_flops / km[3] / 1e9))
        del fb

    # ----------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        import subprocess
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", str(threads)], env=env,
                           stdout=subprocess.PIPE, text=True, check=True)
        c = json.loads(r.stdout.strip().splitlines()[-1])
        cs = c["canonical_sample"]
        rate = cs["eris"] / cs["seconds"]
        cpu = {"value": rate, "unit": "ERIs/s", "cores": threads, "kind": c["kind"], "cpu_model": c["cpu_model"],
               "cpus_visible": os.cpu_count(), "omp": "OMP_PROC_BIND=close OMP_PLACES=cores",
               "sample": "canonical screened J/K (thr 1e-12) over every %dth bra shell pair (offset %d) of the same "
                         "16-water cc-pVDZ workload: %d quartets, %d ERIs in %.1f s" %
                         (cs["stride"], cs["offset"], cs["quartets"], cs["eris"], cs["seconds"]),
               "projected_full_jk_seconds": info["n_eri"] / rate,
               "default_n4_path": {"what": "unscreened N^4 loop of CpuEngine::compute_and_consume_parallel "
                                           "(ThreadLocal) over %d seeded first-shell pairs (i, j) of 192^2, each "
                                           "with all (k, l): %d shell quartets" %
                                           (c["n4_sample"]["first_shell_pairs"], c["n4_sample"]["shell_quartets"]),
                                   "seconds_for_sample": c["n4_sample"]["seconds"],
                                   "projected_full_jk_seconds": c["n4_sample"]["projected_full_seconds"]},
               "h2o_sto3g_default_n4_jk_seconds": c["h2o_sto3g_n4_seconds"]}
        log("CPU baseline (%s, %d threads): %.3e ERIs/s canonical screened; N^4 default path projected %s s; "
            "H2O/STO-3G N^4 %.3f ms" % (c["cpu_model"], threads, rate, c["n4_sample"]["projected_full_seconds"],
                                        c["h2o_sto3g_n4_seconds"] * 1e3))

    eris_per_s = info["n_eri"] / (ms64 * 1e-3)
    out = {
        "metric": METRIC,
        "value": eris_per_s,
        "unit": "ERIs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded geometry and density)",
        "config": {"workload": "16-water cluster cc-pVDZ (400 Cartesian bf), Schwarz 1e-12, FP64 J+K build",
                   "n_quartets": info["n_quartets"], "n_eri": info["n_eri"], "parallelism": "single GPU"},
        "jk_seconds_fp64": ms64 * 1e-3,
        "jk_seconds_fp64_persistent_launch": ms64p * 1e-3,
        "quartets_per_s": info["n_quartets"] / (ms64 * 1e-3),
        "jk_build_tflops_executed": build_tflops,
        "jk_build_frac_fp64_peak": build_tflops / FP64_PEAK_TFLOPS,
        "jk_build_tflops_all_nonzero_primitives": build_tflops_model,
        "flops_executed": exec_total,
        "flops_all_nonzero_primitives": flops_total,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic_of(class_name(dom)),
                     "kernel": "class %s ERI/J-K kernel" % class_name(dom),
                     "kernel_ms": float(ms_cls[dom]),
                     "pipe": "FP64 vector ALU (the Rys recursion and digestion are FMA chains, not matrix "
                             "tiles; MI355X's FP64 vector and matrix peaks are both 78.6 TFLOP/s)"},
        "cpu_baseline": cpu,
    }
    out.update(tune)
    out.update(mixed)
    out.update(batch)
    out.update(grad)
    out.update(df)
    out.update(scf_out)
    print(json.dumps(out), flush=True)
    if args.copy_probe_mb > 0:  # fill writes N bytes; copy reads N and writes N
        x = torch.empty(args.copy_probe_mb * (1 << 20) // 8, dtype=torch.float64, device=dev)
        x.fill_(1.0)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
