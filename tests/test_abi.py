"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public header declares, and rejects malformed arguments on the host before
any HIP call."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

import custingeralg_amd as T
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "custinger_tc.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tc_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_exported():
    names = declared_symbols()
    assert len(names) >= 30
    lib = T.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(T.EXPORTS) == names


def test_symbols_are_plain_c():
    # extern "C": no C++ mangling on the exported names
    out = os.popen(f"nm -D --defined-only {T.LIB_PATH}").read()
    for n in declared_symbols():
        assert re.search(rf"\bT {n}$", out, re.M), n


def test_status_strings_and_version():
    lib = T.lib()
    assert lib.tc_status_string(T.TC_OK) == b"ok"
    assert lib.tc_status_string(T.TC_ERR_INVALID) == b"invalid argument"
    assert lib.tc_status_string(99) == b"unknown status"
    assert T.lib().tc_version() >= 10000


def test_schedule_default_without_gpu():
    s = T.default_schedule()
    assert s.method == T.METHOD_AUTO and s.policy == T.POLICY_FIXED
    assert 1 <= s.small_max <= s.mid_max and s.job_items > 0
    # no merges by the fixed heuristic, built-in hub degree
    assert s.ratio_x16 == 0 and s.hub_degree == 0 and s.build == 0


def test_host_validation_rejects_bad_csr():
    lib = T.lib()
    h = C.c_void_p()
    off = np.array([0, 3, 2], np.int64)  # ends match nnz but not monotone
    adj = np.array([1, 0], np.int32)
    st = lib.tc_graph_create_csr(2, 2, off.ctypes.data, adj.ctypes.data, T.TC_HOST, 0, 0,
                                 C.byref(h))
    assert st == T.TC_ERR_INVALID and not h.value
    assert b"monotone" in lib.tc_last_error()
    off = np.array([0, 1, 3], np.int64)  # off[nv] != nnz
    st = lib.tc_graph_create_csr(2, 2, off.ctypes.data, adj.ctypes.data, T.TC_HOST, 0, 0,
                                 C.byref(h))
    assert st == T.TC_ERR_INVALID
    st = lib.tc_graph_create_csr(-1, 0, None, None, T.TC_HOST, 0, 0, C.byref(h))
    assert st == T.TC_ERR_INVALID
    st = lib.tc_graph_create_csr(2, 2, off.ctypes.data, adj.ctypes.data, 7, 0, 0, C.byref(h))
    assert st == T.TC_ERR_INVALID


def test_null_handles_are_rejected():
    lib = T.lib()
    assert lib.tc_static_count(None, None, None, None, None) == T.TC_ERR_INVALID
    assert lib.tc_batch_apply(None, None, None, None, 0, None, None, None, None) == T.TC_ERR_INVALID
    assert lib.tc_graph_get_info(None, None) == T.TC_ERR_INVALID
    assert lib.tc_graph_destroy(None) == T.TC_OK


def test_bad_schedule_rejected():
    lib = T.lib()
    s = T.default_schedule()
    s.policy = 42
    # a non-null fake handle is never dereferenced: the schedule check comes first
    assert lib.tc_static_count(C.c_void_p(8), None, None, C.byref(s), None) == T.TC_ERR_INVALID


def test_host_to_host_copy():
    a = np.arange(10, dtype=np.int32)
    b = np.zeros(10, np.int32)
    assert T.lib().tc_memcpy(b.ctypes.data, a.ctypes.data, a.nbytes, 3) == T.TC_OK
    assert np.array_equal(a, b)
    assert T.lib().tc_memcpy(b.ctypes.data, a.ctypes.data, a.nbytes, 9) == T.TC_ERR_INVALID


def test_host_rng_and_batched_spec_match_oracle():
    """The package's host copy of the counter RNG and the batched workload sizes
    agree with the oracle's C generator."""
    for seed, ctr in [(0, 0), (1, 2), (1000, 1), (2**63, 2**40 + 7)]:
        assert T.rng(seed, ctr) == O.rng(seed, ctr)
    for a, b in zip(T.batched_spec(G=300), O.batched_spec(G=300)):
        assert np.array_equal(a, b)
