"""The dispatch table on the GPU: every compiled strategy, chunk cap and grid
cap of a class gives the same J/K (and gradient) as the default choice, the
autotuner keeps results unchanged, and tables travel through JSON and
LIBACCINT_DISPATCH_TABLE.  Tolerances: 1e-12 relative for FP64 J/K (only the
order of the FP64 atomic sums differs), 1e-12 absolute for the gradient;
1e-10 relative for the mixed-precision build, whose FP32 quartets change
their float operation order with the kernel shape."""
import ctypes

import numpy as np
import pytest

from libaccint import _cabi, systems
import libaccint as lai

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def strategies(c):
    s, a = ctypes.c_int(), ctypes.c_int()
    _cabi.check(_cabi.lib().lai_dev_class_strategies(c, ctypes.byref(s), ctypes.byref(a)))
    return s.value, a.value


def forced_table(use_alt):
    """Every class on its alternative strategy (or the default one) with
    chunk and grid caps far from the defaults."""
    S = lai.GpuExecutionStrategy
    small = {0: 300, 1: 5, 2: 3}  # 300: not a multiple of the 256-lane sort run
    ents = []
    for c in range(55):
        s, a = strategies(c)
        st = a if (use_alt and a >= 0) else s
        ents.append(lai.DispatchEntry(c, S(st), small[st], 1))
    return lai.DispatchTable(ents)


def test_every_strategy_matches_default_cc_pvtz():
    atoms = systems.water_cluster(2, 8.0, seed=5)
    b = lai.basis_set("cc-pvtz", atoms)  # s..f: every class family
    n = b.n_basis_functions()
    D = systems.density(n, 10, 17, 1.0 / np.sqrt(n))
    e = lai.Engine(b)
    opts = lai.JKOptions(schwarz_threshold=1e-12)
    mixed = lai.JKOptions(schwarz_threshold=1e-12, mixed_precision=True, fp32_threshold=1e-7)
    J0, K0 = e.build_jk(D, opts)
    Jm0, Km0 = e.build_jk(D, mixed)
    for use_alt in (True, False):
        t = forced_table(use_alt)
        e.set_dispatch_table(t)
        got = {x.cls: (x.strategy, x.chunk, x.blocks_per_cu) for x in e.dispatch_table().entries}
        assert got == {x.cls: (x.strategy, x.chunk, x.blocks_per_cu) for x in t.entries}
        J1, K1 = e.build_jk(D, opts)
        assert rel(J1, J0) < 1e-12 and rel(K1, K0) < 1e-12, use_alt
        Jm1, Km1 = e.build_jk(D, mixed)
        assert rel(Jm1, Jm0) < 1e-10 and rel(Km1, Km0) < 1e-10, use_alt
    e.set_dispatch_table(lai.DispatchTable())  # back to the defaults
    for x in e.dispatch_table().entries:
        assert (int(x.strategy), x.chunk, x.blocks_per_cu) == (strategies(x.cls)[0], 0, 0)


def test_gradient_unchanged_by_tuned_chunks():
    atoms = systems.water_cluster(2, 7.0, seed=21)
    b = lai.basis_set("def2-svp", atoms)
    n = b.n_basis_functions()
    D = systems.density(n, 10, 9, 1.0 / np.sqrt(n))
    e = lai.Engine(b)
    g0 = e.compute_jk_gradient(D, 1.0)
    e.set_dispatch_table(forced_table(True))
    g1 = e.compute_jk_gradient(D, 1.0)
    assert np.max(np.abs(g1 - g0)) < 1e-12


def test_autotune_json_env_and_profile_once(tmp_path, monkeypatch):
    torch = pytest.importorskip("torch")
    atoms = systems.water_cluster(3, 10.0, seed=8)
    b = lai.basis_set("cc-pvdz", atoms)
    n = b.n_basis_functions()
    D = systems.density(n, 15, 4, 1.0 / np.sqrt(n))
    opts = lai.JKOptions(schwarz_threshold=1e-12)
    e = lai.Engine(b)
    J0, K0 = e.build_jk(D, opts)
    counts = e.worklist_info()["class_counts"]
    tD = torch.from_numpy(D).cuda()
    tab = e.autotune(tD.data_ptr(), opts, reps=1)
    assert sorted(x.cls for x in tab.entries) == [c for c in range(55) if counts[c] > 0]
    for x in tab.entries:
        s, a = strategies(x.cls)
        assert int(x.strategy) in (s, a) and x.ms > 0 and x.ms_default > 0
        assert x.ms <= x.ms_default
    J1, K1 = e.build_jk(D, opts)
    assert rel(J1, J0) < 1e-12 and rel(K1, K0) < 1e-12
    chosen = {x.cls: (x.strategy, x.chunk, x.blocks_per_cu) for x in e.dispatch_table().entries}
    # JSON file -> LIBACCINT_DISPATCH_TABLE -> a fresh engine takes the same choices
    path = str(tmp_path / "table.json")
    e.dispatch_table().to_json(path)
    monkeypatch.setenv("LIBACCINT_DISPATCH_TABLE", path)
    e2 = lai.Engine(b)
    assert {x.cls: (x.strategy, x.chunk, x.blocks_per_cu) for x in e2.dispatch_table().entries} == chosen
    J2, K2 = e2.build_jk(D, opts)
    assert rel(J2, J0) < 1e-12 and rel(K2, K0) < 1e-12
    monkeypatch.delenv("LIBACCINT_DISPATCH_TABLE")
    # ProfileOnce: the first build of an engine with auto tuning tunes itself
    cfg = lai.DispatchConfig()
    cfg.enable_auto_tuning = True
    cfg.auto_tune_min_batch = 1
    e3 = lai.Engine(b, cfg)
    J3, K3 = e3.build_jk(D, opts)
    assert rel(J3, J0) < 1e-12 and rel(K3, K0) < 1e-12
    assert any(x.ms > 0 for x in e3.autotune(tD.data_ptr(), opts, reps=1).entries)


def test_batch_engine_autotune_keeps_results():
    torch = pytest.importorskip("torch")
    mols = [("cc-pvdz", systems.water_cluster(2, 8.0, seed=3)),
            ("def2-svp", systems.random_molecule(np.random.default_rng(7), 4, 6))]
    be = lai.BatchEngine([lai.basis_set(nm, at) for nm, at in mols])
    Ds = [systems.density(n, max(2, n // 5), 11 + m, 1.0 / np.sqrt(n)) for m, n in enumerate(be.n_basis_functions())]
    Db = np.concatenate([D.ravel() for D in Ds])
    opts = lai.JKOptions(schwarz_threshold=1e-12)
    J0, K0 = be.build_jk(Db, opts)
    tab = be.autotune(torch.from_numpy(Db).cuda().data_ptr(), opts, reps=1)
    assert len(tab) > 0
    J1, K1 = be.build_jk(Db, opts)
    assert rel(J1, J0) < 1e-12 and rel(K1, K0) < 1e-12


def test_concurrent_classes_equal_serial_build():
    """lai_dev_jk spreads the class kernels over four streams (fork/join);
    lai_dev_jk_profile runs them one after another on one stream.  Same J/K
    to the order of the FP64 atomic sums, and the four-stream build is
    repeatable on the caller's stream."""
    torch = pytest.importorskip("torch")
    atoms = systems.water_cluster(3, 9.0, seed=8)
    b = lai.basis_set("cc-pvdz", atoms)
    n = b.n_basis_functions()
    D = systems.density(n, 15, 21, 1.0 / np.sqrt(n))
    e = lai.Engine(b)
    e.prepare_worklist(lai.JKOptions(schwarz_threshold=1e-12))
    L = _cabi.lib()
    h = ctypes.c_void_p(e.device_basis_handle())
    s = torch.cuda.Stream()
    tD = torch.from_numpy(D).cuda()
    P = ctypes.POINTER(ctypes.c_double)
    out = []
    for prof in (False, True, False):
        tJ = torch.zeros_like(tD)
        tK = torch.zeros_like(tD)
        if prof:
            m64, m32 = np.zeros(55), np.zeros(55)
            _cabi.check(L.lai_dev_jk_profile(h, ctypes.c_void_p(tD.data_ptr()), ctypes.c_void_p(tJ.data_ptr()),
                                             ctypes.c_void_p(tK.data_ptr()), 0, ctypes.c_void_p(s.cuda_stream),
                                             m64.ctypes.data_as(P), m32.ctypes.data_as(P)))
        else:
            _cabi.check(L.lai_dev_jk(h, ctypes.c_void_p(tD.data_ptr()), ctypes.c_void_p(tJ.data_ptr()),
                                     ctypes.c_void_p(tK.data_ptr()), 0, ctypes.c_void_p(s.cuda_stream)))
        s.synchronize()
        out.append((tJ.cpu().numpy(), tK.cpu().numpy()))
    (J0, K0), (J1, K1), (J2, K2) = out
    assert rel(J0, J1) < 1e-12 and rel(K0, K1) < 1e-12
    assert rel(J0, J2) < 1e-12 and rel(K0, K2) < 1e-12


def test_persistent_launch_equals_per_class_launches():
    """LAI_JK_PERSISTENT: the register-shape classes drain one device queue in
    a single launch.  Same J/K as the per-class launches, for one molecule
    (cc-pVDZ, every covered class with s-d shells) and a batch of several
    molecules in cc-pVTZ (f shells, covered and uncovered classes side by
    side), also through a captured graph replayed twice."""
    waters = systems.water_cluster(4, 9.0, seed=3)
    b = lai.basis_set("cc-pvdz", waters)
    n = b.n_basis_functions()
    D = systems.density(n, 20, 8, 1.0 / np.sqrt(n))
    e = lai.Engine(b)
    J0, K0 = e.build_jk(D, lai.JKOptions(schwarz_threshold=1e-12))
    J1, K1 = e.build_jk(D, lai.JKOptions(schwarz_threshold=1e-12, persistent=True))
    assert rel(J1, J0) < 1e-12 and rel(K1, K0) < 1e-12
    torch = pytest.importorskip("torch")
    s = torch.cuda.Stream()
    # a graph captured before any persistent build of the engine
    e2 = lai.Engine(b)
    op = lai.JKOptions(schwarz_threshold=1e-12, persistent=True)
    e2.prepare_worklist(op)
    tD = torch.from_numpy(D).cuda()
    tJ = torch.zeros_like(tD)
    tK = torch.zeros_like(tD)
    e2.capture_jk_graph(tD.data_ptr(), tJ.data_ptr(), tK.data_ptr(), op, s.cuda_stream)
    e2.launch_jk_graph(s.cuda_stream)
    s.synchronize()
    assert rel(tJ.cpu().numpy(), J0) < 1e-12 and rel(tK.cpu().numpy(), K0) < 1e-12
    rng = np.random.default_rng(4)
    mols = [systems.random_molecule(rng) for _ in range(20)]
    mols = [m for m in mols if len(m) <= 13][:3]  # 13, 7 and 10 atoms
    assert len(mols) == 3
    be = lai.BatchEngine([lai.basis_set("cc-pvtz", m) for m in mols])
    nbf = be.n_basis_functions()
    Db = np.concatenate([systems.density(k, max(2, k // 5), 40 + i, 1.0 / np.sqrt(k)).ravel()
                         for i, k in enumerate(nbf)])
    Jb0, Kb0 = be.build_jk(Db, lai.JKOptions(schwarz_threshold=1e-12))
    op = lai.JKOptions(schwarz_threshold=1e-12, persistent=True)
    Jb1, Kb1 = be.build_jk(Db, op)
    assert rel(Jb1, Jb0) < 1e-12 and rel(Kb1, Kb0) < 1e-12
    tD = torch.from_numpy(Db).cuda()
    tJ = torch.zeros_like(tD)
    tK = torch.zeros_like(tD)
    be.capture_jk_graph(tD.data_ptr(), tJ.data_ptr(), tK.data_ptr(), op, s.cuda_stream)
    for rep in range(2):
        tJ.zero_()
        tK.zero_()
        torch.cuda.synchronize()
        be.launch_jk_graph(s.cuda_stream)
        s.synchronize()
        assert rel(tJ.cpu().numpy(), Jb0) < 1e-12 and rel(tK.cpu().numpy(), Kb0) < 1e-12, rep
