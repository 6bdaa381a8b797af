"""The C-ABI library loads and exports every entry point of include/ttx.h; host-only helpers work
without a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ttx.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ttx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = set(re.findall(r" T (ttx_[a-z0-9_]+)$", out, flags=re.M))
    assert set(names) <= exported


def test_binding_covers_header():
    assert set(declared_symbols()) == set(L.SIGNATURES)


def test_params_init_and_core_size():
    lib = L.load()
    p = L.Params()
    dims = [32] * 6
    ranks = [1, 32, 32, 32, 32, 32, 1]
    lib.ttx_params_init(ctypes.byref(p), 6, L.i64_array(dims), L.i32_array(ranks))
    assert p.oversampling == 10 and p.power_iters == 2 and p.omega_seed == 1234
    assert p.precision == L.PREC_BF16 and p.rank_mode == L.RANK_FIXED
    assert lib.ttx_core_size(ctypes.byref(p), 0, 0) == 32 * 32
    assert lib.ttx_core_size(ctypes.byref(p), 0, 2) == 32 * 32 * 32
    assert lib.ttx_core_size(ctypes.byref(p), 0, 5) == 32 * 32
    # TR layouts: N1 x (r0 r1) and r_{d-1} x (N_d r0)
    tr = [4, 16, 16, 16, 16, 16, 4]
    lib.ttx_params_init(ctypes.byref(p), 6, L.i64_array(dims), L.i32_array(tr))
    assert lib.ttx_core_size(ctypes.byref(p), 1, 0) == 32 * 64
    assert lib.ttx_core_size(ctypes.byref(p), 1, 5) == 16 * 32 * 4
    assert lib.ttx_core_size(ctypes.byref(p), 1, 6) == 0


def test_status_strings():
    lib = L.load()
    assert lib.ttx_status_string(0) == b"ok"
    assert lib.ttx_status_string(1) == b"bad argument"
    assert lib.ttx_status_string(L.ERR_NO_CONVERGE) == b"no convergence"
    assert lib.ttx_status_string(99) == b"unknown status"
    assert lib.ttx_version() >= 100


def test_null_handle_rejected_without_gpu():
    lib = L.load()
    p = L.Params()
    assert lib.ttx_tt_svd(None, None, ctypes.byref(p), None, None, None) == L.ERR_BAD_ARG
    assert lib.ttx_permute_231(None, None, None, 1, 1, 1) == L.ERR_BAD_ARG
