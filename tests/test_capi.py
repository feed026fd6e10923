"""The C ABI library loads without a device and exports every entry point include/oic.h declares."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from canu_amd import oic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "oic.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oic_[a-z_]+)\s*\(", text)))


def test_header_and_python_mirror_agree():
    assert declared() == sorted(oic.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = oic.lib()
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", oic.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (oic_\w+)", out))
    assert set(declared()) <= exported


def test_params_default_matches_reference_defaults():
    p = oic.Params()
    oic.lib().oic_params_default(ctypes.byref(p))
    # oicParameters::initialize (overlapInCore.H:374-411)
    assert (p.kmer, p.hash_bits, p.unique, p.hopeless, p.partial) == (22, 22, 1, 1, 0)
    assert abs(p.max_erate - 0.06) < 1e-12 and abs(p.hash_load - 0.6) < 1e-12
    assert p.hash_data_len == 100000000 and p.frag_olap_limit == 2 ** 64 - 1


def test_dp_tables_rejects_bad_rate():
    me = ctypes.c_uint32()
    assert oic.lib().oic_dp_tables(0.0, 1.0, None, None, ctypes.byref(me)) == oic.OIC_EINVAL
    assert oic.lib().oic_dp_tables(0.045, 1.0, None, None, ctypes.byref(me)) == 0
    assert me.value == 94373


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(oic, "LIB_PATH", str(tmp_path / "liboic_hip.so"))
    monkeypatch.setattr(oic, "_lib", None)
    with pytest.raises(RuntimeError):
        oic.Engine()
