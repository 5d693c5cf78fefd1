"""DispatchPolicy::select_backend against the expectations of the reference's
own unit tests (tests/unit/engine/test_dispatch_policy.cpp:15-338; HIP where
they expect CUDA).  CPU only: the policy is host logic."""
import pytest

import libaccint as lai

W, H, B = lai.WorkUnitType, lai.BackendHint, lai.BackendType

# (work type, batch, total AM, primitives, hint, GPU available, expected) -- test_dispatch_policy.cpp:70-338
CASES = [
    (W.FullBasis, 1000, 10, 10000, H.ForceCPU, True, B.CPU),              # ForceCPU_AlwaysReturnsCPU
    (W.SingleShellPair, 1, 0, 1, H.ForceGPU, True, B.HIP),                # ForceGPU_ReturnsGPU_WhenAvailable
    (W.FullBasis, 1000, 10, 10000, H.ForceGPU, False, B.CPU),             # ForceGPU_FallsBackToCPU_WhenNotAvailable
    (W.FullBasis, 1000, 10, 10000, H.Auto, False, B.CPU),                 # NoGPU_AlwaysReturnsCPU
    (W.SingleShellPair, 1, 2, 9, H.Auto, True, B.CPU),                    # SingleShellPair_PrefersCPU
    (W.SingleShellQuartet, 1, 4, 81, H.Auto, True, B.CPU),               # SingleShellQuartet_PrefersCPU
    (W.ShellSetPair, 100, 4, 10000, H.Auto, True, B.HIP),                 # ShellSetPair_LargeBatch_PrefersGPU
    (W.ShellSetPair, 4, 2, 100, H.Auto, True, B.CPU),                     # ShellSetPair_SmallBatch_PrefersCPU
    (W.ShellSetQuartet, 10, 4, 1000, H.Auto, True, B.HIP),                # ShellSetQuartet_EvenModestBatch_CanPreferGPU
    (W.FullBasis, 50, 4, 10000, H.Auto, True, B.HIP),                     # FullBasis_LargeSystem_PrefersGPU
    (W.FullBasis, 5, 2, 100, H.Auto, True, B.CPU),                        # FullBasis_SmallSystem_PrefersCPU
    (W.ShellSetPair, 20, 3, 2000, H.PreferCPU, True, B.CPU),              # PreferCPU_RaisesBarForGPU
    (W.ShellSetPair, 8, 3, 500, H.PreferGPU, True, B.HIP),                # PreferGPU_LowersBarForGPU
    (W.ShellSetQuartet, 8, 16, 500, H.Auto, True, B.HIP),                 # HighAM_PrefersGPU
]


@pytest.mark.parametrize("wt,batch,am,prims,hint,gpu,expected", CASES)
def test_select_backend_matches_reference_tests(wt, batch, am, prims, hint, gpu, expected):
    assert lai.DispatchPolicy().select_backend(wt, batch, am, prims, hint, gpu) == expected


def test_single_shell_very_high_am():
    # SingleShellPair_VeryHighAM_MayPreferGPU: either answer is allowed there;
    # the rule gives GPU at total AM >= high_am_threshold + 2
    p = lai.DispatchPolicy()
    assert p.select_backend(W.SingleShellPair, 1, 8, 100, H.Auto, True) == B.HIP
    assert p.select_backend(W.SingleShellPair, 1, 5, 100, H.Auto, True) == B.CPU
    # PreferCPU never sends a single shell pair to the GPU
    assert p.select_backend(W.SingleShellPair, 1, 8, 100, H.PreferCPU, True) == B.CPU


def test_config_round_trip_and_thresholds():
    c = lai.DispatchConfig()
    assert (c.min_gpu_batch_size, c.min_gpu_primitives, c.min_gpu_shells) == (16, 1000, 10)
    c.min_gpu_shells = 20
    p = lai.DispatchPolicy(c)
    assert p.config().min_gpu_shells == 20
    assert p.select_backend(W.FullBasis, 15, 2, 100, H.Auto, True) == B.CPU
    assert p.select_backend(W.FullBasis, 20, 2, 100, H.Auto, True) == B.HIP
    assert not p.auto_tuning_enabled()
    p.set_auto_tuning(True)
    assert p.auto_tuning_enabled()
