"""The C ABI boundary without a GPU: struct layout, exported symbols,
default configuration, and loud failure when no device is visible."""
import ctypes
import os
import re

import numpy as np
import pytest

import _oracle as O
from imau_vel import abi, config, solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_python_structs_match_c_layout():
    out = np.zeros(8, np.int64)
    O.lib().orc_abi_layout(O._ptr(out))
    assert out[0] == ctypes.sizeof(abi.ImauGrid)
    assert out[1] == ctypes.sizeof(abi.ImauCfg)
    assert out[2] == ctypes.sizeof(abi.ImauIceIO)
    assert out[3] == abi.ImauCfg.zeta.offset
    assert out[4] == abi.ImauCfg.check_interval.offset
    assert out[5] == abi.ImauIceIO.u_vav_cx.offset
    assert out[6] == abi.ImauIceIO.DIVA_SOR_nit.offset
    assert out[7] == abi.ImauIceIO.stat_seconds.offset
    out = np.zeros(11, np.int64)
    O.lib().orc_abi_layout2(O._ptr(out))
    sizes = [ctypes.sizeof(c) for c in (abi.ImauThicknessCfg, abi.ImauThicknessIO, abi.ImauProducerCfg,
                                        abi.ImauProducerIO, abi.ImauTimestepCfg, abi.ImauTimestepState,
                                        abi.ImauRefgeoIO)]
    assert list(out[:7]) == sizes
    assert out[7] == abi.ImauTimestepCfg.continental_shelf_min_height.offset
    assert out[8] == abi.ImauTimestepState.nsteps.offset
    assert out[9] == abi.ImauRefgeoIO.Hb_GIAeq.offset
    assert out[10] == abi.ImauTimestepCfg.choice_calving_law.offset


def test_library_exports_every_declared_entry_point():
    L = solver.load_library()
    txt = open(abi.HEADER).read()
    declared = set(re.findall(r"\b(imau_vel_\w+)\s*\(", txt))
    assert declared, "no entry points parsed"
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared == set(solver.EXPORTED)


def test_library_exports_every_thickness_entry_point():
    L = solver.load_library()
    txt = open(abi.HEADER).read()
    declared = set(re.findall(r"\b(imau_(?:thickness|critical)_\w+)\s*\(", txt))
    assert declared == set(solver.THICKNESS_EXPORTED)
    for name in sorted(declared):
        assert hasattr(L, name), name


def test_library_exports_every_producer_entry_point():
    L = solver.load_library()
    txt = open(abi.HEADER).read()
    declared = set(re.findall(r"\b(imau_(?:producer|update|calc)_\w+)\s*\(", txt))
    assert declared == set(solver.PRODUCER_EXPORTED)
    for name in sorted(declared):
        assert hasattr(L, name), name


def test_library_exports_the_time_loop():
    L = solver.load_library()
    txt = open(abi.HEADER).read()
    declared = set(re.findall(r"\b(imau_run_\w+)\s*\(", txt))
    assert declared == set(solver.TIMELOOP_EXPORTED)
    for name in sorted(declared):
        assert hasattr(L, name), name


def test_producer_defaults_match_the_python_mirror():
    from imau_vel import config
    L = solver.load_library()
    c = abi.ImauProducerCfg()
    L.imau_producer_default_config(ctypes.byref(c))
    ref = config.producer_config()
    for name, t in abi.ImauProducerCfg._fields_:
        if hasattr(t, "_length_"):
            assert list(getattr(c, name)) == list(getattr(ref, name)), name
        else:
            assert getattr(c, name) == getattr(ref, name), name


def test_thickness_defaults_match_the_python_mirror():
    from imau_vel import config
    L = solver.load_library()
    c = abi.ImauThicknessCfg()
    L.imau_thickness_default_config(ctypes.byref(c))
    ref = config.thickness_config()
    for name, _ in abi.ImauThicknessCfg._fields_:
        assert getattr(c, name) == getattr(ref, name), name


def test_library_exports_every_network_entry_point():
    from imau_vel import learned
    L = solver.load_library()
    txt = open(os.path.join(ROOT, "include", "imau_net.h")).read()
    declared = set(re.findall(r"\b(imau_net_\w+)\s*\(", txt))
    assert declared == set(learned.NET_EXPORTED)
    for name in sorted(declared):
        assert hasattr(L, name), name


def test_c_default_config_equals_python_defaults():
    L = solver.load_library()
    c = abi.ImauCfg()
    L.imau_vel_default_config(ctypes.byref(c))
    p = config.default_config()
    for name, typ in abi.ImauCfg._fields_:
        a, b = getattr(c, name), getattr(p, name)
        if name == "zeta":
            assert list(a) == list(b)
        else:
            assert a == b, name


def test_context_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(solver.ImauError) as e:
        solver.Context()
    assert e.value.code == abi.IMAU_ERR_NO_DEVICE


def test_fortran_bindings_are_current():
    """imau_vel_bindings.f90 is the generator's output for the current header."""
    import subprocess
    import sys
    gen = os.path.join(ROOT, "imau-ice_amd", "fortran", "gen_bindings.py")
    path = os.path.join(ROOT, "imau-ice_amd", "fortran", "imau_vel_bindings.f90")
    before = open(path).read()
    out = subprocess.run([sys.executable, gen, "--stdout"], capture_output=True, text=True, check=True).stdout
    assert out == before
