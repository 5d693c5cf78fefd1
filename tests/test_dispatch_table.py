"""Dispatch table of the J/K kernels on the host side: its JSON form, the
strategy names of the reference's OptimalDispatchTable
(include/libaccint/kernels/optimal_dispatch_table.hpp:29-56), the compiled
strategies of every class and the C-ABI argument checks.  CPU only."""
import ctypes
import json

import pytest

from libaccint import _cabi
import libaccint as lai

NAMES = "spdf"


def class_am(c):
    import math
    pb = (math.isqrt(8 * c + 1) - 1) // 2
    pk = c - pb * (pb + 1) // 2

    def pam(p):
        la = (math.isqrt(8 * p + 1) - 1) // 2
        return la, p - la * (la + 1) // 2
    return pam(pb) + pam(pk)


def strategies(c):
    s, a = ctypes.c_int(), ctypes.c_int()
    assert _cabi.lib().lai_dev_class_strategies(c, ctypes.byref(s), ctypes.byref(a)) == 0
    return s.value, a.value


def test_strategy_names_match_reference_vocabulary():
    S = lai.GpuExecutionStrategy
    for st, name in ((S.ThreadPerQuartet, "thread_per_quartet"), (S.WarpPerQuartet, "warp_per_quartet"),
                     (S.BlockPerQuartet, "block_per_quartet")):
        assert lai.strategy_name(st) == name
        assert lai.gpu_execution_strategy_from_string(name) == st
    with pytest.raises(ValueError):
        lai.gpu_execution_strategy_from_string("lane_per_quartet")


def test_json_round_trip_and_class_by_am():
    S = lai.GpuExecutionStrategy
    t = lai.DispatchTable([lai.DispatchEntry(2, S.ThreadPerQuartet, 512, 4),
                           lai.DispatchEntry(14, S.BlockPerQuartet, 8, 0)])
    text = t.dump_json()
    doc = json.loads(text)  # valid JSON
    assert doc["version"] == 1 and len(doc["entries"]) == 2
    assert doc["entries"][0]["am"] == [1, 0, 1, 0]  # class 2 is (ps|ps)
    assert doc["entries"][1]["am"] == [2, 1, 2, 1]  # class 14 is (dp|dp)
    u = lai.DispatchTable.parse_json(text)
    assert [(e.cls, e.strategy, e.chunk, e.blocks_per_cu) for e in u.entries] == \
        [(2, S.ThreadPerQuartet, 512, 4), (14, S.BlockPerQuartet, 8, 0)]
    # an entry may name its class by angular momenta only
    v = lai.DispatchTable.parse_json(json.dumps(
        {"entries": [{"am": [2, 2, 1, 0], "strategy": "warp_per_quartet", "chunk": 16}]}))
    assert v.entries[0].cls == 16 and v.entries[0].chunk == 16 and v.entries[0].blocks_per_cu == 0


def test_json_file_round_trip(tmp_path):
    S = lai.GpuExecutionStrategy
    t = lai.DispatchTable([lai.DispatchEntry(c, S.ThreadPerQuartet, 256 + c, 0) for c in range(3)])
    p = str(tmp_path / "table.json")
    t.to_json(p)
    u = lai.DispatchTable.from_json(p)
    assert [(e.cls, e.chunk) for e in u.entries] == [(0, 256), (1, 257), (2, 258)]
    with pytest.raises(RuntimeError):
        lai.DispatchTable.from_json(str(tmp_path / "missing.json"))


@pytest.mark.parametrize("doc,msg", [
    ({"entries": [{"class": 3, "am": [1, 0, 1, 0], "strategy": "thread_per_quartet"}]}, "disagree"),
    ({"entries": [{"am": [0, 1, 0, 0], "strategy": "thread_per_quartet"}]}, "canonical"),
    ({"entries": [{"class": 55, "strategy": "thread_per_quartet"}]}, "valid class"),
    ({"entries": [{"class": 1}]}, "strategy"),
    ({"version": 2, "entries": []}, "version"),
    ({"items": []}, "entries"),
])
def test_json_rejects_malformed_tables(doc, msg):
    with pytest.raises(ValueError, match=msg):
        lai.DispatchTable.parse_json(json.dumps(doc))


def test_every_class_has_a_compiled_strategy():
    seen_alt = 0
    for c in range(55):
        s, a = strategies(c)
        assert s in (0, 1, 2) and a in (-1, 0, 1, 2) and a != s
        seen_alt += a >= 0
    # lane per quartet for the s/p classes, workgroup per quartet for (ff|ff)
    assert strategies(0) == (0, -1) and strategies(1) == (0, -1)
    assert strategies(54)[0] == 2
    assert class_am(14) == (2, 1, 2, 1) and strategies(14) == (1, 2)
    assert seen_alt >= 30
    s, a = ctypes.c_int(), ctypes.c_int()
    assert _cabi.lib().lai_dev_class_strategies(55, ctypes.byref(s), ctypes.byref(a)) == -1


def test_cabi_tuning_calls_reject_null_handles():
    L = _cabi.lib()
    tab = (_cabi.ClassTune * 55)()
    n = ctypes.c_int(-1)
    assert L.lai_dev_autotune(None, None, 0, 3, tab, 55, ctypes.byref(n), None) == -1
    assert L.lai_dev_get_tuning(None, tab, 55, ctypes.byref(n)) == -1
    assert L.lai_dev_set_tuning(None, tab, 1) == -1
    assert b"invalid arguments" in L.lai_dev_last_error()


def test_auto_tuning_config_fields():
    c = lai.DispatchConfig()
    assert c.enable_auto_tuning is False and c.auto_tune_min_batch == 100
    c.enable_auto_tuning = True
    c.auto_tune_min_batch = 1
    assert c.enable_auto_tuning and c.auto_tune_min_batch == 1


def test_class_index_helpers_agree_with_decode():
    for c in range(55):
        am = class_am(c)
        assert tuple(lai.quartet_class_am(c)) == am
        assert lai.quartet_class_index(list(am)) == c
    assert lai.quartet_class_index([0, 1, 0, 0]) == -1  # la < lb
    assert lai.quartet_class_index([0, 0, 1, 0]) == -1  # bra pair class below the ket's
    assert lai.quartet_class_index([4, 0, 0, 0]) == -1  # beyond f
    with pytest.raises(ValueError):
        lai.quartet_class_am(55)
