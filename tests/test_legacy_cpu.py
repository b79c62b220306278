"""CPU checks of the single-beam (legacy) manager library: it loads next to
libmdsm_amd, exports the reference's C++ names and its own C entry points,
and the XML reader gives the legacy manager's fields.  No GPU compute."""
import ctypes as C
import os
import re

import pytest

from mdsm_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LEGACY = os.path.join(ROOT, "mdsm_amd", "libmdsm_legacy.so")


class XmlParams(C.Structure):
    _fields_ = [("fch1", C.c_float), ("foff", C.c_float), ("tsamp", C.c_float), ("lowdm", C.c_float),
                ("dmstep", C.c_float), ("detection_threshold", C.c_float), ("nchans", C.c_uint32),
                ("tdms", C.c_uint32), ("nsamp", C.c_uint32), ("nbits", C.c_uint32),
                ("use_brute_force", C.c_int), ("fileprefix", C.c_char * 120), ("basedir", C.c_char * 120),
                ("secs_per_file", C.c_uint32), ("use_pc_time", C.c_char), ("single_file_mode", C.c_char),
                ("dialect", C.c_int)]


def _parse(path):
    lib = abi.lib()
    lib.mdsm_parse_observation_xml.restype = C.c_int
    lib.mdsm_parse_observation_xml.argtypes = [C.c_char_p, C.POINTER(XmlParams)]
    x = XmlParams()
    rc = lib.mdsm_parse_observation_xml(path.encode(), C.byref(x))
    return rc, x


def test_legacy_library_exports():
    lib = C.CDLL(LEGACY)
    txt = open(os.path.join(ROOT, "include", "mdsm_legacy.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    c_names = sorted(set(re.findall(r"\b(mdsm_legacy_[a-z_]+)\s*\(", txt)))
    assert len(c_names) == 8
    for n in c_names:
        assert hasattr(lib, n), n
    # the reference's C++ entry points (src/mdsm/dedispersion_manager.h:7-12)
    for mangled in ["_Z14initialiseMDSMP6SURVEY", "_Z10next_chunkjRjddPfS0_", "_Z16start_processingj",
                    "_Z12tearDownMDSMv"]:
        assert hasattr(lib, mangled), mangled


def test_observation_xml_legacy_fields():
    rc, x = _parse(os.path.join(ROOT, "tests", "golden", "survey_legacy_single.xml"))
    assert rc == 0
    assert (x.nchans, x.tdms, x.nsamp) == (64, 96, 4096)
    assert (x.fch1, x.foff, x.lowdm, x.dmstep) == (420.0, -0.25, 0.0, 0.25)
    assert abs(x.tsamp - 0.000256) < 1e-10 and x.detection_threshold == 3.5
    assert x.use_brute_force == 1 and x.fileprefix == b"legacy_run" and x.basedir == b"."
    assert x.single_file_mode == b"\x01" and x.secs_per_file == 600 and x.dialect & 2


def test_observation_xml_subband_flag(tmp_path):
    p = tmp_path / "sub.xml"
    p.write_text('<observation><frequencies top="400" offset="-1"/>'
                 '<dm lowDM="0" numDMs="8" dmStep="1" useBruteForce="0"/><channels number="8"/>'
                 '<timing tsamp="0.001"/></observation>')
    rc, x = _parse(str(p))
    assert rc == 0 and x.use_brute_force == 0
    p.write_text('<observation><dm lowDM="0" numDMs="8" dmStep="1"/></observation>')
    rc, x = _parse(str(p))
    assert rc == 0 and x.use_brute_force == 1
    rc, _ = _parse(str(tmp_path / "missing.xml"))
    assert rc != 0


def test_legacy_open_refuses_subband(tmp_path):
    # refused before any device call: subband dedispersion is not provided
    p = tmp_path / "sub.xml"
    p.write_text('<observation><frequencies top="400" offset="-1"/>'
                 '<dm lowDM="0" numDMs="8" dmStep="1" useBruteForce="0"/><channels number="8"/>'
                 '<timing tsamp="0.001"/><samples number="1024"/></observation>')
    lib = C.CDLL(LEGACY)
    lib.mdsm_legacy_open.argtypes = [C.c_char_p, C.c_uint, C.c_uint]
    lib.mdsm_legacy_last_error.restype = C.c_char_p
    assert lib.mdsm_legacy_open(str(p).encode(), 0, 0) != 0
    assert b"subband" in lib.mdsm_legacy_last_error()
    assert lib.mdsm_legacy_open(str(tmp_path / "missing.xml").encode(), 0, 0) != 0
    lib.mdsm_legacy_close()
