"""CPU-only checks of the product library: it loads, exports every symbol the
C header declares, and its host-side logic (delay tables, survey validation,
clustering) matches the oracle.  No GPU compute is called here."""
import os
import re

import numpy as np
import pytest

import oracle as O
from mdsm_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mdsm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    names = _declared("mdsm_abi.h")
    assert len(names) > 40
    lib = abi.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("mode", [abi.DELAY_FP64_ROUND, abi.DELAY_REFERENCE, abi.DELAY_LEGACY])
@pytest.mark.parametrize("cfg", [
    (418.0, -0.078125, 256, 51.2e-6, 0.0, 0.1, 256),
    (1500.0, -0.29296875, 1024, 64e-6, 0.0, 0.2, 4096),
    (1400.0, -0.5, 512, 100e-6, 3.0, 0.37, 999),
])
def test_delay_table_matches_oracle(cfg, mode):
    fch1, foff, C, ts, lo, st, M = cfg
    t, ms = abi.delay_table(abi.Survey.make(fch1, foff, C, ts, lo, st, M, 1024, mode))
    ref = {abi.DELAY_FP64_ROUND: O.delay_table, abi.DELAY_REFERENCE: O.delay_table_ref,
           abi.DELAY_LEGACY: O.delay_table_legacy}[mode](*cfg)
    assert np.array_equal(t, ref)
    assert ms == ref.max()


def test_survey_validation_errors():
    bad = [
        abi.Survey.make(1500.0, +0.3, 64, 64e-6, 0.0, 1.0, 8, 1024),   # ascending band
        abi.Survey.make(1500.0, -0.3, 0, 64e-6, 0.0, 1.0, 8, 1024),    # no channels
        abi.Survey.make(1500.0, -0.3, 64, 0.0, 0.0, 1.0, 8, 1024),     # tsamp 0
        abi.Survey.make(100.0, -1.0, 128, 64e-6, 0.0, 1.0, 8, 1024),   # negative frequency
        abi.Survey.make(1500.0, -0.3, 64, 64e-6, -1.0, 1.0, 8, 1024),  # negative DM
    ]
    for s in bad:
        with pytest.raises(abi.MdsmError):
            abi.delay_table(s)


def _cands(d, t, snr=None):
    c = np.zeros(len(d), abi.CAND_DTYPE)
    c["dm_idx"] = d
    c["sample"] = t
    c["snr"] = np.arange(len(d), dtype=np.float32) if snr is None else snr
    return c


def test_dbscan2d_matches_oracle_and_sklearn():
    from sklearn.cluster import DBSCAN
    rng = np.random.default_rng(7)
    pts = np.concatenate([rng.integers(0, 300, (400, 2)), rng.integers(0, 8, (80, 2)) + 100,
                          np.stack([np.arange(40) * 3, np.arange(40)], 1)]).astype(np.int32)
    pts = np.unique(pts, axis=0)
    c = _cands(pts[:, 0], pts[:, 1], rng.random(len(pts)).astype(np.float32))
    abi.sort_candidates(c)
    lab, n = abi.dbscan2d(c, 3, 4)
    olab, on = O.dbscan2d(c["dm_idx"], c["sample"], 3, 4)
    assert n == on and np.array_equal(lab, olab)
    sk = DBSCAN(eps=3, min_samples=4, metric="chebyshev").fit(
        np.stack([c["dm_idx"], c["sample"]], 1)).labels_
    assert np.array_equal(lab, sk)
    peaks, sizes = abi.cluster_peaks(c, lab, n)
    for k in range(n):
        m = lab == k
        assert sizes[k] == m.sum()
        assert peaks[k]["snr"] == c["snr"][m].max()


def test_dbscan2d_rejects_unsorted():
    c = _cands([5, 1], [0, 0])
    with pytest.raises(abi.MdsmError):
        abi.dbscan2d(c)


def test_dbscan_ref_matches_oracle():
    rng = np.random.default_rng(8)
    n = 300
    tm = np.round(rng.random(n) * 2.0, 3)
    dm = np.round(rng.random(n) * 10.0, 1)
    snr = 6 + rng.random(n) * 3
    for mp in (1, 2, 3, 5):
        lab, k = abi.dbscan_ref(tm, dm, snr, 0.05, 1.0, 1.5, mp)
        olab, ok = O.dbscan_ref(tm, dm, snr, 0.05, 1.0, 1.5, mp)
        assert k == ok and np.array_equal(lab, olab)


# ---------------------------------------------------------------- XML surveys
import ctypes as _C
_lib = abi.lib()
_lib.mdsm_compat_parse_xml.restype = _C.c_int
_lib.mdsm_compat_parse_xml.argtypes = [_C.c_char_p, _C.POINTER(abi.Survey), _C.c_uint32,
                                       _C.POINTER(_C.c_uint32), _C.POINTER(_C.c_float),
                                       _C.POINTER(_C.c_uint32), _C.POINTER(_C.c_int)]


def _parse(name):
    beams = (abi.Survey * 8)()
    nb, thr, mp, dia = _C.c_uint32(), _C.c_float(), _C.c_uint32(), _C.c_int()
    path = os.path.join(ROOT, "tests", "golden", name).encode()
    rc = _lib.mdsm_compat_parse_xml(path, beams, 8, _C.byref(nb), _C.byref(thr), _C.byref(mp), _C.byref(dia))
    assert rc == 0
    return list(beams[:nb.value]), thr.value, mp.value, dia.value


def test_xml_multibeam_dialect():
    beams, thr, mp, dia = _parse("survey_multibeam.xml")
    assert dia == 1 and len(beams) == 2
    b = beams[1]
    assert (b.fch1, b.foff, b.nchans, b.ndms, b.nsamp) == (1400.0, -0.25, 128, 300, 8192)
    assert abs(b.tsamp - 6.4e-5) < 1e-12 and abs(b.lowdm - 2.5) < 1e-7 and abs(b.dmstep - 0.25) < 1e-7
    assert abs(thr - 6.5) < 1e-6 and mp == 3


def test_xml_legacy_attribute_names():
    # the multibeam reader of the reference would read nsamp=0 and threshold=0
    # here; both names are accepted and the dialect is reported
    beams, thr, mp, dia = _parse("survey_legacy_names.xml")
    assert dia & 2 and len(beams) == 1
    b = beams[0]
    assert (b.nchans, b.ndms, b.nsamp) == (256, 512, 16384)
    assert abs(b.fch1 - 399.951172) < 1e-4 and abs(thr - 4.0) < 1e-6 and mp == 15


def test_xml_errors(tmp_path):
    beams = (abi.Survey * 1)()
    nb = _C.c_uint32()
    assert _lib.mdsm_compat_parse_xml(b"/nonexistent.xml", beams, 1, _C.byref(nb), None, None, None) != 0
    bad = str(tmp_path / "bad.xml")
    with open(bad, "w") as f:
        f.write("<notobservation><dm lowDM='0'/></notobservation>")
    assert _lib.mdsm_compat_parse_xml(bad.encode(), beams, 1, _C.byref(nb), None, None, None) != 0
    with open(bad, "w") as f:
        f.write("<observation><dm lowDM='0' </observation>")
    assert _lib.mdsm_compat_parse_xml(bad.encode(), beams, 1, _C.byref(nb), None, None, None) != 0


_lib.mdsm_compat_open.argtypes = [_C.c_char_p, _C.c_uint]
_lib.mdsm_compat_open.restype = _C.c_int
_lib.mdsm_compat_last_error.restype = _C.c_char_p

_XML_DIR = """<observation>
  <dm lowDM="0" numDMs="16" dmStep="1.0" />
  <channels nchans="32" npols="1" />
  <timing tsamp="0.000064" />
  <samples nsamp="4096" nbits="32" />
  {storage}
  <writer filePrefix="x" baseDirectory="{dir}" singleFileMode="1" />
  <beams><beam beamId="0" topFrequency="1500" frequencyOffset="-1.0" /></beams>
</observation>"""


def test_compat_unwritable_output_directory_fails_initialisation(tmp_path):
    """The reference exits when a candidate file cannot be created
    (multibeam_dedispersion_output.cpp:41-45); here initialiseMDSM fails
    (before any device call) and reports the directory."""
    xml = tmp_path / "obs.xml"
    xml.write_text(_XML_DIR.format(dir=str(tmp_path / "missing" / "dir"), storage=""))
    assert _lib.mdsm_compat_open(str(xml).encode(), 0) == 2
    assert b"Invalid output directory" in _lib.mdsm_compat_last_error()


def test_compat_storage_element(tmp_path):
    xml = tmp_path / "obs.xml"
    xml.write_text(_XML_DIR.format(dir=str(tmp_path), storage='<storage type="bogus" />'))
    assert _lib.mdsm_compat_open(str(xml).encode(), 0) == 1
    assert b"unknown storage type" in _lib.mdsm_compat_last_error()
