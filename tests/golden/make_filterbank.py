"""Writes the small SIGPROC filterbank fixtures used by the ingest tests
(header keywords as length-prefixed strings, binary values, then
time-major samples packed per nbits: 1/2/4 bits least-significant first,
8/16 bits unsigned, 32 bits float).  Run: python tests/golden/make_filterbank.py"""
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _s(b, text):
    b += struct.pack("<i", len(text)) + text.encode()
    return b


def header(nchans, nbits, fch1=1500.0, foff=-0.5, tsamp=6.4e-5, tstart=60000.5, name="FAKE_PSR"):
    b = b""
    b = _s(b, "HEADER_START")
    b = _s(b, "source_name")
    b = _s(b, name)
    for key, fmt, val in [("machine_id", "<i", 10), ("telescope_id", "<i", 4), ("data_type", "<i", 1),
                          ("fch1", "<d", fch1), ("foff", "<d", foff), ("nchans", "<i", nchans),
                          ("nbits", "<i", nbits), ("tstart", "<d", tstart), ("tsamp", "<d", tsamp),
                          ("nifs", "<i", 1)]:
        b = _s(b, key)
        b += struct.pack(fmt, val)
    b = _s(b, "HEADER_END")
    return b


def rich_header(nchans, nbits):
    """Every keyword the reference parser knows, a Telescope string (read
    but not counted in total_bytes) and a FREQUENCY_START..END table."""
    b = b""
    b = _s(b, "HEADER_START")
    b = _s(b, "rawdatafile")
    b = _s(b, "raw_2026.dat")
    b = _s(b, "source_name")
    b = _s(b, "B0531+21")
    b = _s(b, "Telescope")
    b = _s(b, "Medicina")
    ints = [("machine_id", 7), ("telescope_id", 3), ("data_type", 1), ("nchans", nchans), ("nbits", nbits),
            ("nifs", 1), ("nbeams", 4), ("ibeam", 2), ("barycentric", 0), ("pulsarcentric", 1),
            ("nbins", 128), ("nsamples", 4096)]
    dbls = [("tstart", 59000.25), ("tsamp", 1.31072e-4), ("refdm", 56.77), ("az_start", 123.5),
            ("za_start", 12.25), ("src_raj", 53431.97), ("src_dej", 220052.1), ("period", 0.0334)]
    for key, val in ints:
        b = _s(b, key)
        b += struct.pack("<i", val)
    for key, val in dbls:
        b = _s(b, key)
        b += struct.pack("<d", val)
    b = _s(b, "npuls")
    b += struct.pack("<q", 123456789012)
    b = _s(b, "FREQUENCY_START")
    for c in range(nchans):
        b = _s(b, "fchannel")
        b += struct.pack("<d", 1420.0 - 0.25 * c)
    b = _s(b, "FREQUENCY_END")
    b = _s(b, "HEADER_END")
    return b


def pack(samples, nbits):
    s = samples.reshape(-1)
    if nbits in (1, 2, 4):
        per = 8 // nbits
        s = s.astype(np.uint8).reshape(-1, per)
        out = np.zeros(s.shape[0], np.uint8)
        for k in range(per):
            out |= (s[:, k] << (k * nbits)).astype(np.uint8)
        return out.tobytes()
    if nbits == 8:
        return s.astype(np.uint8).tobytes()
    if nbits == 16:
        return s.astype("<u2").tobytes()
    return s.astype("<f4").tobytes()


def fixture_arrays():
    """The sample values of every fixture, [t][chan], as written."""
    rng = np.random.default_rng(123)
    nchans, nsamp = 64, 200
    out = {}
    for nbits in (1, 2, 4, 8, 16, 32):
        if nbits == 32:
            out[f"fil_{nbits}bit.fil"] = rng.standard_normal((nsamp, nchans)).astype(np.float32) * 10 + 3
        else:
            out[f"fil_{nbits}bit.fil"] = rng.integers(0, 1 << nbits, (nsamp, nchans))
    out["fil_table_8bit.fil"] = rng.integers(0, 256, (40, 16))
    out["fil_raw_8bit.dat"] = rng.integers(0, 256, (50, 64))
    return out


def main():
    for name, data in fixture_arrays().items():
        nbits = int(name.split("_")[-1].split("bit")[0])
        with open(os.path.join(HERE, name), "wb") as f:
            if name.startswith("fil_table"):
                f.write(rich_header(data.shape[1], nbits))
            elif not name.endswith(".dat"):  # .dat: raw 8-bit samples, no header
                f.write(header(data.shape[1], nbits))
            f.write(pack(data, nbits))


if __name__ == "__main__":
    main()
