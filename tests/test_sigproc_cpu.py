"""SIGPROC ingest on the CPU: the oracle's unpack and corner turn against
the reference's own read_block (compiled from file_handler.cpp) and the
values the fixture writer packed; the product header parser against the
reference read_header; argument and format errors of the C ABI."""
import importlib.util
import os

import numpy as np
import pytest

import oracle
from mdsm_amd import abi

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
_spec = importlib.util.spec_from_file_location("make_filterbank", os.path.join(GOLDEN, "make_filterbank.py"))
mkfil = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(mkfil)

FIXTURES = mkfil.fixture_arrays()
NBITS = {name: int(name.split("_")[-1].split("bit")[0]) for name in FIXTURES}
needs_ref = pytest.mark.skipif(not oracle.ref_file_available(),
                               reason="oracle/_ref/libref_file.so not built (MDSM sources absent)")


def _payload(name):
    path = os.path.join(GOLDEN, name)
    h = abi.sigproc_read_header(path)
    with open(path, "rb") as f:
        return h, f.read()[h.data_offset:]


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_oracle_unpack_known_values(name):
    data = FIXTURES[name]
    h, payload = _payload(name)
    got = oracle.sigproc_unpack(payload, NBITS[name], data.size)
    np.testing.assert_array_equal(got, data.reshape(-1).astype(np.float32))


@needs_ref
@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_oracle_unpack_matches_reference_read_block(name):
    data = FIXTURES[name]
    h, payload = _payload(name)
    ref = oracle.ref_read_block(os.path.join(GOLDEN, name), h.data_offset, NBITS[name], data.size)
    assert ref.size == data.size
    np.testing.assert_array_equal(oracle.sigproc_unpack(payload, NBITS[name], data.size), ref)


def test_oracle_unpack_bit_order():
    # LSB-first fields: 0b11100100 = fields (0,1,2,3) at 2 bits, (4, 14) at 4 bits
    assert oracle.sigproc_unpack(b"\xe4", 2, 4).tolist() == [0, 1, 2, 3]
    assert oracle.sigproc_unpack(b"\xe4", 4, 2).tolist() == [4, 14]
    assert oracle.sigproc_unpack(b"\xe4", 1, 8).tolist() == [0, 0, 1, 0, 0, 1, 1, 1]
    assert oracle.sigproc_unpack(b"\x01\x02", 16, 1).tolist() == [513]


def test_oracle_corner_turn_is_a_transpose():
    rng = np.random.default_rng(5)
    nspec, nchans, nbeams, ld, col0 = 37, 5, 3, 50, 9
    spectra = rng.standard_normal((nspec, nchans * nbeams)).astype(np.float32)
    out = oracle.corner_turn(spectra, nchans, nbeams, ld=ld, col0=col0).reshape(nbeams, nchans, ld)
    want = spectra.reshape(nspec, nbeams, nchans).transpose(1, 2, 0)
    np.testing.assert_array_equal(out[:, :, col0:col0 + nspec], want)
    assert not out[:, :, :col0].any() and not out[:, :, col0 + nspec:].any()


@needs_ref
@pytest.mark.parametrize("name", sorted(n for n in FIXTURES if not n.endswith(".dat")))
def test_header_matches_reference(name):
    path = os.path.join(GOLDEN, name)
    h = abi.sigproc_read_header(path)
    table = name.startswith("fil_table")
    ref = oracle.ref_header_values(path, ntable=16 if table else 0)
    fields = ["nchans", "nbits", "nifs", "machine_id", "telescope_id", "data_type", "tstart", "tsamp",
              "fch1", "foff", "total_bytes"]
    if table:
        fields += ["nbeams", "ibeam", "barycentric", "pulsarcentric", "nbins", "nsamples", "refdm",
                   "az_start", "za_start", "src_raj", "src_dej", "period", "npuls"]
        assert h.rawdatafile.decode() == ref["rawdatafile"]
        assert h.nfchannel == 16
        np.testing.assert_array_equal(np.array(h.frequency_table[:16]), ref["frequency_table"])
    for k in fields:
        assert float(getattr(h, k)) == ref[k], k
    assert h.source_name.decode() == ref["source_name"]
    assert h.data_offset == ref["file_position"]
    assert h.has_header == 1


def test_header_telescope_not_counted():
    # the reference's total_bytes omits the Telescope value string (12 bytes)
    h = abi.sigproc_read_header(os.path.join(GOLDEN, "fil_table_8bit.fil"))
    assert h.data_offset - h.total_bytes == 4 + len("Medicina")
    assert h.fch1 == 0.0 and h.foff == 0.0  # a frequency table is in use


def test_raw_file_without_header():
    h = abi.sigproc_read_header(os.path.join(GOLDEN, "fil_raw_8bit.dat"))
    assert (h.has_header, h.nbits, h.data_offset) == (0, 8, 0)
    if oracle.ref_file_available():
        assert oracle.ref_header_values(os.path.join(GOLDEN, "fil_raw_8bit.dat")) is None


def _status(fn, *args):
    rc = fn(*args)
    return rc, abi.last_error()


def test_header_format_errors(tmp_path):
    lib = abi.lib()
    good = mkfil.header(8, 8)
    h = abi.SigprocHeader()
    cases = {
        "unknown": good.replace(b"nifs", b"nofs"),
        "truncated": good[:-20],
        "badlen": mkfil._s(b"", "HEADER_START") + b"\x00\x01\x00\x00xx",
    }
    for name, blob in cases.items():
        p = tmp_path / f"{name}.fil"
        p.write_bytes(blob)
        rc, msg = _status(lib.mdsm_sigproc_read_header, str(p).encode(), abi.C.byref(h))
        assert rc == abi.MDSM_ERR_FORMAT, (name, msg)
    p = tmp_path / "empty.fil"
    p.write_bytes(b"")
    assert lib.mdsm_sigproc_read_header(str(p).encode(), abi.C.byref(h)) == abi.MDSM_ERR_FORMAT
    missing = str(tmp_path / "missing.fil").encode()
    assert lib.mdsm_sigproc_read_header(missing, abi.C.byref(h)) == abi.MDSM_ERR_IO


def test_unpack_argument_errors():
    lib = abi.lib()
    fake = 0x1000  # never dereferenced: every case fails validation first
    assert lib.mdsm_unpack_corner_turn(fake, 3, 10, 8, 1, fake, 10, 80, 0, None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_unpack_corner_turn(fake, 1, 10, 3, 1, fake, 10, 30, 0, None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_unpack_corner_turn(fake, 8, 10, 8, 1, fake, 9, 72, 0, None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_unpack_corner_turn(fake, 8, 10, 8, 2, fake, 10, 70, 0, None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_unpack_corner_turn(fake + 1, 16, 10, 8, 1, fake, 10, 80, 0, None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_unpack_corner_turn(None, 8, 10, 8, 1, fake, 10, 80, 0, None) == abi.MDSM_ERR_ARG


def test_reader_open_errors():
    lib = abi.lib()
    r = abi.C.c_void_p()
    raw = os.path.join(GOLDEN, "fil_raw_8bit.dat").encode()
    # a raw file has no channel count: the caller must give it
    assert lib.mdsm_sigproc_open(raw, 0, 1, 16, abi.C.byref(r), None) == abi.MDSM_ERR_ARG
    fil = os.path.join(GOLDEN, "fil_8bit.fil").encode()
    # 64 channels in the header cannot be 3 beams x 20
    assert lib.mdsm_sigproc_open(fil, 20, 3, 16, abi.C.byref(r), None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_sigproc_open(fil, 0, 3, 16, abi.C.byref(r), None) == abi.MDSM_ERR_ARG


def test_dm_writer_round_trip(tmp_path):
    rng = np.random.default_rng(8)
    series = rng.standard_normal((130, 300)).astype(np.float32)
    prefix = str(tmp_path / "MDSM")
    w = abi.DmWriter(prefix, [0.0, 12.3, 3.5, 99.0], fch1=150.0, foff=-0.1953125, tsamp=2.56e-4)
    w.write(series[:, :200], lowdm=0.0, dmstep=0.1)
    w.write(series[:, 200:], lowdm=0.0, dmstep=0.1)
    w.close()
    import struct
    for dm, row in [(0.0, 0), (12.3, 123), (3.5, 35)]:
        path = f"{prefix}_{dm:g}.fil"
        h = abi.sigproc_read_header(path)
        assert (h.nchans, h.nbits, h.nifs, h.data_type, h.machine_id) == (1, 32, 1, 1, 0)
        assert (h.fch1, h.foff, h.tsamp, h.tstart) == (150.0, -0.1953125, 2.56e-4, 0.0)
        data = np.fromfile(path, np.float32, offset=h.data_offset)
        np.testing.assert_array_equal(data, series[row])
        # the header bytes, keyword by keyword, in the reference's order
        blob = open(path, "rb").read()[:h.data_offset]
        want = mkfil._s(b"", "HEADER_START")
        for k, fmt, v in [("machine_id", "<i", 0), ("telescope_id", "<i", 0), ("data_type", "<i", 1),
                          ("fch1", "<d", 150.0), ("foff", "<d", -0.1953125), ("nchans", "<i", 1),
                          ("tsamp", "<d", 2.56e-4), ("nbits", "<i", 32), ("tstart", "<d", 0.0),
                          ("nifs", "<i", 1)]:
            want = mkfil._s(want, k) + struct.pack(fmt, v)
        assert blob == mkfil._s(want, "HEADER_END")
    # a DM with no matching trial gets a header and no samples
    h = abi.sigproc_read_header(f"{prefix}_99.fil")
    assert os.path.getsize(f"{prefix}_99.fil") == h.data_offset
