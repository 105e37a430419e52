"""ctypes mirror of include/imau_vel.h.

The struct layouts are read from the header itself so that the Python side
cannot drift from the C ABI.  The header mirrors IMAU-ICE's type_grid /
type_ice_model / config struct (src/data_types_module.f90:13-325,
src/configuration_module.f90:271-377).
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
REPO_ROOT = os.path.abspath(os.path.join(_HERE, "..", ".."))
HEADER = os.path.join(REPO_ROOT, "include", "imau_vel.h")

_CTYPE = {
    "double": ctypes.c_double,
    "int32_t": ctypes.c_int32,
    "int64_t": ctypes.c_int64,
}


def _header_text() -> str:
    with open(HEADER, "r") as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return txt


def _defines(txt: str) -> dict:
    out = {}
    for name, val in re.findall(r"#define\s+(\w+)\s+(-?\d+)", txt):
        out[name] = int(val)
    return out


def _enums(txt: str) -> dict:
    out = {}
    for body in re.findall(r"enum\s*\{(.*?)\}", txt, flags=re.S):
        for name, val in re.findall(r"(\w+)\s*=\s*(-?\d+)", body):
            out[name] = int(val)
    return out


def _struct_fields(txt: str, name: str, defines: dict):
    m = re.search(r"typedef\s+struct\s+%s\s*\{(.*?)\}\s*%s\s*;" % (name, name), txt, flags=re.S)
    if not m:
        raise RuntimeError("struct %s not found in %s" % (name, HEADER))
    fields = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        decl = decl.replace("const ", "")
        tm = re.match(r"(double|int32_t|int64_t)\s+(.*)$", decl, flags=re.S)
        if not tm:
            raise RuntimeError("cannot parse declaration %r" % decl)
        base = _CTYPE[tm.group(1)]
        for item in tm.group(2).split(","):
            item = item.strip()
            ptr = item.startswith("*")
            item = item.lstrip("*").strip()
            am = re.match(r"(\w+)\s*\[\s*(\w+)\s*\]", item)
            if am:
                n = am.group(2)
                n = defines[n] if n in defines else int(n)
                fields.append((am.group(1), base * n))
            elif ptr:
                fields.append((item, ctypes.POINTER(base)))
            else:
                fields.append((item, base))
    return fields


_TXT = _header_text()
DEFINES = _defines(_TXT)
ENUMS = _enums(_TXT)
NZ_MAX = DEFINES["IMAU_NZ_MAX"]


class ImauGrid(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_grid", DEFINES)


class ImauCfg(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_cfg", DEFINES)


class ImauIceIO(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_ice_io", DEFINES)


class ImauThicknessCfg(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_thickness_cfg", DEFINES)


class ImauThicknessIO(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_thickness_io", DEFINES)


class ImauProducerCfg(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_producer_cfg", DEFINES)


class ImauProducerIO(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_producer_io", DEFINES)


class ImauTimestepCfg(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_timestep_cfg", DEFINES)


class ImauTimestepState(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_timestep_state", DEFINES)


class ImauPcState(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_pc_state", DEFINES)


class ImauRefgeoIO(ctypes.Structure):
    _fields_ = _struct_fields(_TXT, "imau_refgeo_io", DEFINES)


IO_POINTER_FIELDS = [n for n, t in ImauIceIO._fields_ if hasattr(t, "contents")]

# export the enum constants at module level (IMAU_OK, IMAU_DYN_DIVA, ...)
globals().update(ENUMS)
