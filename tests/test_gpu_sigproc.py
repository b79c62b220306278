"""SIGPROC unpack + corner turn on the GPU against the oracle (bit exact):
every fixture at every bit width, beams split out of the spectrum, row
padding and a column offset, larger random streams, and the streaming
reader over a file in uneven chunks."""
import os

import numpy as np
import pytest

import oracle
from mdsm_amd import abi

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = ["fil_1bit.fil", "fil_2bit.fil", "fil_4bit.fil", "fil_8bit.fil", "fil_16bit.fil",
            "fil_32bit.fil", "fil_table_8bit.fil", "fil_raw_8bit.dat"]


def _gpu_unpack(packed, nbits, nspec, nchans, nbeams, ld, col0, beam_stride=None):
    beam_stride = nchans * ld if beam_stride is None else beam_stride
    d_in = abi.DeviceBuffer.from_numpy(np.frombuffer(packed, np.uint8).copy())
    d_out = abi.DeviceBuffer(4 * nbeams * beam_stride)
    d_out.upload(np.full(nbeams * beam_stride, -7.0, np.float32))
    abi.unpack_corner_turn(d_in.ptr, nbits, nspec, nchans, nbeams, d_out.ptr, ld, beam_stride, col0)
    abi.synchronize()
    return d_out.download(np.float32, (nbeams * beam_stride,))


def _expected(packed, nbits, nspec, nchans, nbeams, ld, col0, beam_stride=None):
    beam_stride = nchans * ld if beam_stride is None else beam_stride
    vals = oracle.sigproc_unpack(packed, nbits, nspec * nchans * nbeams)
    out = np.full(nbeams * beam_stride, -7.0, np.float32)
    return oracle.corner_turn(vals, nchans, nbeams, ld=ld, beam_stride=beam_stride, col0=col0, out=out)


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("nbeams", [1, 2])
def test_unpack_fixtures(name, nbeams):
    path = os.path.join(GOLDEN, name)
    h = abi.sigproc_read_header(path)
    ctot = h.nchans if h.has_header else 64
    with open(path, "rb") as f:
        payload = f.read()[h.data_offset:]
    nspec = len(payload) * 8 // (ctot * h.nbits)
    nchans = ctot // nbeams
    ld, col0 = nspec + 13, 5
    print(f"unpack {name} nbits={h.nbits} nspec={nspec} beams={nbeams}", flush=True)
    got = _gpu_unpack(payload, h.nbits, nspec, nchans, nbeams, ld, col0)
    np.testing.assert_array_equal(got, _expected(payload, h.nbits, nspec, nchans, nbeams, ld, col0))


@pytest.mark.parametrize("nbits", [1, 2, 4, 8, 16, 32])
def test_unpack_random_streams(nbits):
    rng = np.random.default_rng(nbits)
    # ragged sizes: channels and spectra not multiples of the 64x64 tile
    for nspec, nchans, nbeams in [(1, 8, 1), (3001, 200, 5), (130, 1000, 1)]:
        nbytes = nspec * nchans * nbeams * nbits // 8
        if nbits == 32:
            packed = rng.standard_normal(nbytes // 4).astype(np.float32).tobytes()
        else:
            packed = rng.integers(0, 256, nbytes, dtype=np.uint8).tobytes()
        ld, col0 = nspec + 64, 17
        beam_stride = nchans * ld + 3
        print(f"unpack random nbits={nbits} nspec={nspec} nchans={nchans} beams={nbeams}", flush=True)
        got = _gpu_unpack(packed, nbits, nspec, nchans, nbeams, ld, col0, beam_stride)
        want = _expected(packed, nbits, nspec, nchans, nbeams, ld, col0, beam_stride)
        np.testing.assert_array_equal(got, want)


def test_reader_streams_file_in_chunks(tmp_path):
    import importlib.util
    spec = importlib.util.spec_from_file_location("mkfil", os.path.join(GOLDEN, "make_filterbank.py"))
    mkfil = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mkfil)
    rng = np.random.default_rng(77)
    nchans, nbeams, nspec, nbits = 96, 2, 5000, 4
    data = rng.integers(0, 16, (nspec, nchans * nbeams))
    path = tmp_path / "stream.fil"
    path.write_bytes(mkfil.header(nchans * nbeams, nbits) + mkfil.pack(data, nbits))
    ld = nspec + 1500 + 32  # every request must fit its row, even past the end of the file
    d_out = abi.DeviceBuffer(4 * nbeams * nchans * ld)
    d_out.zero()
    reader = abi.SigprocReader(str(path), nchans=0, nbeams=nbeams, max_spectra=1500)
    assert reader.nchans == nchans and reader.header.nbits == nbits
    col, counts = 0, []
    for want in [1, 1499, 700, 1500, 1500, 1500]:
        got = reader.read(want, d_out.ptr, ld, col0=col)
        counts.append(got)
        col += got
    reader.close()
    abi.synchronize()
    assert counts == [1, 1499, 700, 1500, 1300, 0]
    out = d_out.download(np.float32, (nbeams * nchans * ld,))
    want = oracle.corner_turn(data.astype(np.float32), nchans, nbeams, ld=ld)
    np.testing.assert_array_equal(out, want)


def test_reader_raw_file():
    path = os.path.join(GOLDEN, "fil_raw_8bit.dat")
    raw = np.fromfile(path, np.uint8).astype(np.float32).reshape(50, 64)
    ld = 64
    d_out = abi.DeviceBuffer(4 * 64 * ld)
    d_out.zero()
    reader = abi.SigprocReader(path, nchans=32, nbeams=2, max_spectra=64)
    got = reader.read(64, d_out.ptr, ld)
    reader.close()
    abi.synchronize()
    assert got == 50
    out = d_out.download(np.float32, (64 * ld,))
    np.testing.assert_array_equal(out, oracle.corner_turn(raw, 32, 2, ld=ld))
