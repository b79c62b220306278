// sigproc.cpp -- SIGPROC filterbank header parser and streaming reader.
//
// Header grammar (read_header, src/mdsm-multibeam/file_handler.cpp:56-190):
// strings are a 32-bit length (1..80) followed by that many bytes; the
// header opens with "HEADER_START" and closes with "HEADER_END"; between
// them each keyword string is followed by a binary value whose type the
// keyword fixes (int32, double, int64 for npuls, or a string for
// source_name, rawdatafile and Telescope).  FREQUENCY_START/FREQUENCY_END
// bracket fchannel entries, which fill a frequency table and set fch1 and
// foff to 0.  A file whose first string is not HEADER_START is raw 8-bit
// data read from byte 0 (src/MDSM.cpp:53-56).
//
// Differences from the reference, all on malformed input: an unknown
// keyword, a bad string length or a truncated header return
// MDSM_ERR_FORMAT (the reference exits the process); a value string is
// always read as the value of the keyword before it (the reference first
// tries it as a keyword); unset fields are zero.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

using namespace mdsm;

namespace {

enum class Kind { I32, F64, I64, Str };

struct Key {
  const char *name;
  Kind kind;
  size_t offset;  // into mdsm_sigproc_header_t
};

#define HK(field, kind) {#field, kind, offsetof(mdsm_sigproc_header_t, field)}
const Key kKeys[] = {
    HK(machine_id, Kind::I32),   HK(telescope_id, Kind::I32), HK(data_type, Kind::I32),
    HK(nchans, Kind::I32),       HK(nbits, Kind::I32),        HK(nifs, Kind::I32),
    HK(nbeams, Kind::I32),       HK(ibeam, Kind::I32),        HK(barycentric, Kind::I32),
    HK(pulsarcentric, Kind::I32), HK(nbins, Kind::I32),       HK(nsamples, Kind::I32),
    HK(tstart, Kind::F64),       HK(tsamp, Kind::F64),        HK(fch1, Kind::F64),
    HK(foff, Kind::F64),         HK(refdm, Kind::F64),        HK(az_start, Kind::F64),
    HK(za_start, Kind::F64),     HK(src_raj, Kind::F64),      HK(src_dej, Kind::F64),
    HK(period, Kind::F64),       HK(npuls, Kind::I64),        HK(source_name, Kind::Str),
    HK(rawdatafile, Kind::Str),
};
#undef HK

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

// one length-prefixed string; false on EOF or a length outside 1..80
bool read_string(FILE *f, std::string &s, int64_t &pos) {
  int32_t n = 0;
  if (fread(&n, 4, 1, f) != 1 || n < 1 || n > 80) return false;
  char buf[81];
  if (fread(buf, 1, (size_t)n, f) != (size_t)n) return false;
  s.assign(buf, (size_t)n);
  pos += 4 + n;
  return true;
}

template <typename T>
bool read_value(FILE *f, void *dst, int64_t &pos) {
  T v;
  if (fread(&v, sizeof(T), 1, f) != 1) return false;
  std::memcpy(dst, &v, sizeof(T));
  pos += sizeof(T);
  return true;
}

// parse from the current position of f (byte 0); leaves f at the data
int parse_header(FILE *f, const char *path, mdsm_sigproc_header_t *h) {
  std::memset(h, 0, sizeof(*h));
  std::string key;
  int64_t pos = 0;
  if (!read_string(f, key, pos) || key != "HEADER_START") {
    // raw 8-bit data (an empty file is an error, as in the reference)
    MDSM_REQUIRE(fseek(f, 0, SEEK_END) == 0 && ftell(f) > 0, MDSM_ERR_FORMAT, "%s: empty file",
                 path);
    fseek(f, 0, SEEK_SET);
    h->nbits = 8;
    return MDSM_OK;
  }
  h->has_header = 1;
  int64_t skipped = 0;  // Telescope value bytes, which the reference does not count
  for (;;) {
    MDSM_REQUIRE(read_string(f, key, pos), MDSM_ERR_FORMAT,
                 "%s: truncated header or bad string at byte %lld", path, (long long)pos);
    if (key == "HEADER_END") break;
    if (key == "FREQUENCY_START") {
      h->nfchannel = 0;
      continue;
    }
    if (key == "FREQUENCY_END") continue;
    if (key == "fchannel") {
      MDSM_REQUIRE(h->nfchannel < 4096, MDSM_ERR_FORMAT, "%s: more than 4096 fchannel entries", path);
      MDSM_REQUIRE(read_value<double>(f, &h->frequency_table[h->nfchannel], pos), MDSM_ERR_FORMAT,
                   "%s: truncated fchannel", path);
      h->nfchannel++;
      h->fch1 = h->foff = 0.0;
      continue;
    }
    if (key == "Telescope") {
      std::string v;
      const int64_t before = pos;
      MDSM_REQUIRE(read_string(f, v, pos), MDSM_ERR_FORMAT, "%s: bad Telescope value", path);
      skipped += pos - before;
      continue;
    }
    const Key *k = nullptr;
    for (const Key &c : kKeys)
      if (key == c.name) k = &c;
    MDSM_REQUIRE(k, MDSM_ERR_FORMAT, "%s: unknown header parameter '%s'", path, key.c_str());
    char *dst = reinterpret_cast<char *>(h) + k->offset;
    bool ok = false;
    switch (k->kind) {
      case Kind::I32: ok = read_value<int32_t>(f, dst, pos); break;
      case Kind::F64: ok = read_value<double>(f, dst, pos); break;
      case Kind::I64: ok = read_value<int64_t>(f, dst, pos); break;
      case Kind::Str: {
        std::string v;
        ok = read_string(f, v, pos);
        if (ok) std::memcpy(dst, v.c_str(), v.size() + 1);
        break;
      }
    }
    MDSM_REQUIRE(ok, MDSM_ERR_FORMAT, "%s: truncated value of '%s'", path, key.c_str());
  }
  h->data_offset = pos;
  h->total_bytes = (int32_t)(pos - skipped);
  return MDSM_OK;
}

}  // namespace

extern "C" int mdsm_sigproc_read_header(const char *path, mdsm_sigproc_header_t *h) {
  MDSM_REQUIRE(path && h, MDSM_ERR_ARG, "null argument");
  File fh;
  fh.f = fopen(path, "rb");
  MDSM_REQUIRE(fh.f, MDSM_ERR_IO, "cannot open %s", path);
  return parse_header(fh.f, path, h);
}

static int check_unpack(int nbits, uint32_t nchans, uint32_t nbeams, uint64_t ld,
                        uint64_t beam_stride, uint64_t col0, uint32_t nspec) {
  MDSM_REQUIRE(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8 || nbits == 16 || nbits == 32,
               MDSM_ERR_ARG, "nbits must be 1, 2, 4, 8, 16 or 32 (got %d)", nbits);
  MDSM_REQUIRE(nchans >= 1 && nbeams >= 1 && (uint64_t)nchans * nbeams <= 65535u * 64u,
               MDSM_ERR_ARG, "nchans*nbeams must be in 1..4194240");
  MDSM_REQUIRE(((uint64_t)nchans * nbeams * nbits) % 8 == 0, MDSM_ERR_ARG,
               "a spectrum of %u %d-bit samples is not a whole number of bytes", nchans * nbeams,
               nbits);
  MDSM_REQUIRE(ld >= col0 + nspec, MDSM_ERR_ARG, "rows of %llu samples cannot hold col0 + nspectra",
               (unsigned long long)ld);
  MDSM_REQUIRE(nbeams == 1 || beam_stride >= (uint64_t)nchans * ld, MDSM_ERR_ARG,
               "beam_stride smaller than nchans*ld overlaps beams");
  return MDSM_OK;
}

extern "C" int mdsm_unpack_corner_turn(const void *d_packed, int nbits, uint32_t nspectra,
                                       uint32_t nchans, uint32_t nbeams, float *d_out, uint64_t ld,
                                       uint64_t beam_stride, uint64_t col0, void *stream) {
  MDSM_REQUIRE(d_packed && d_out, MDSM_ERR_ARG, "null buffer");
  int rc = check_unpack(nbits, nchans, nbeams, ld, beam_stride, col0, nspectra);
  if (rc != MDSM_OK) return rc;
  MDSM_REQUIRE(nbits < 16 || (reinterpret_cast<uintptr_t>(d_packed) % (nbits / 8)) == 0, MDSM_ERR_ARG,
               "packed %d-bit input must be %d-byte aligned", nbits, nbits / 8);
  return launch_unpack_corner_turn(d_packed, nbits, nspectra, nchans, nbeams, d_out, ld,
                                   beam_stride, col0, as_stream(stream));
}

// ---------------------------------------------------------------- reader
// Two staging slots (pinned host + device bytes) alternate between calls:
// the file read of one call overlaps the copy and unpack of the previous
// one; an event per slot guards its reuse.
struct mdsm_sigproc_reader_s {
  FILE *f = nullptr;
  int nbits = 8;
  uint32_t nchans = 0, nbeams = 1, max_spectra = 0;
  uint64_t spec_bytes = 0;
  void *host[2] = {nullptr, nullptr};
  void *dev[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  bool pending[2] = {false, false};
  int slot = 0;
};

static void free_reader(mdsm_sigproc_reader_s *r) {
  for (int i = 0; i < 2; ++i) {
    if (r->done[i]) {
      if (r->pending[i]) (void)hipEventSynchronize(r->done[i]);
      (void)hipEventDestroy(r->done[i]);
    }
    if (r->host[i]) (void)hipHostFree(r->host[i]);
    if (r->dev[i]) (void)hipFree(r->dev[i]);
  }
  if (r->f) fclose(r->f);
  delete r;
}

extern "C" int mdsm_sigproc_open(const char *path, uint32_t nchans, uint32_t nbeams,
                                 uint32_t max_spectra, mdsm_sigproc_reader *out,
                                 mdsm_sigproc_header_t *header) {
  MDSM_REQUIRE(path && out && nbeams >= 1 && max_spectra >= 1, MDSM_ERR_ARG, "bad arguments");
  *out = nullptr;
  mdsm_sigproc_header_t local;
  mdsm_sigproc_header_t *h = header ? header : &local;
  File fh;
  fh.f = fopen(path, "rb");
  MDSM_REQUIRE(fh.f, MDSM_ERR_IO, "cannot open %s", path);
  int rc = parse_header(fh.f, path, h);
  if (rc != MDSM_OK) return rc;
  if (nchans == 0) {
    MDSM_REQUIRE(h->has_header && h->nchans > 0 && (uint32_t)h->nchans % nbeams == 0, MDSM_ERR_ARG,
                 "%s: channel count not given and not in the header as a multiple of %u beams",
                 path, nbeams);
    nchans = (uint32_t)h->nchans / nbeams;
  } else if (h->has_header && h->nchans > 0) {
    MDSM_REQUIRE((uint64_t)nchans * nbeams == (uint64_t)h->nchans, MDSM_ERR_ARG,
                 "%s: header has %d channels, expected %u beams x %u", path, h->nchans, nbeams,
                 nchans);
  }
  rc = check_unpack(h->nbits, nchans, nbeams, max_spectra, (uint64_t)nchans * max_spectra, 0,
                    max_spectra);
  if (rc != MDSM_OK) return rc;
  auto *r = new mdsm_sigproc_reader_s();
  r->nbits = h->nbits;
  r->nchans = nchans;
  r->nbeams = nbeams;
  r->max_spectra = max_spectra;
  r->spec_bytes = (uint64_t)nchans * nbeams * h->nbits / 8;
  const size_t bytes = r->spec_bytes * max_spectra;
  bool ok = true;
  for (int i = 0; ok && i < 2; ++i)
    ok = hipHostMalloc(&r->host[i], bytes, hipHostMallocDefault) == hipSuccess &&
         hipMalloc(&r->dev[i], bytes) == hipSuccess &&
         hipEventCreateWithFlags(&r->done[i], hipEventDisableTiming) == hipSuccess;
  r->f = fh.f;
  fh.f = nullptr;
  if (!ok) {
    free_reader(r);
    set_error("reader staging allocation of 2 x %zu bytes failed", bytes);
    return MDSM_ERR_NOMEM;
  }
  *out = r;
  return MDSM_OK;
}

extern "C" int mdsm_sigproc_close(mdsm_sigproc_reader r) {
  if (r) free_reader(r);
  return MDSM_OK;
}

extern "C" int mdsm_sigproc_read(mdsm_sigproc_reader r, uint32_t nspectra, float *d_out,
                                 uint64_t ld, uint64_t beam_stride, uint64_t col0, uint32_t *got,
                                 void *stream) {
  MDSM_REQUIRE(r && d_out && got, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(nspectra <= r->max_spectra, MDSM_ERR_ARG, "%u spectra exceed the reader's %u",
               nspectra, r->max_spectra);
  *got = 0;
  int rc = check_unpack(r->nbits, r->nchans, r->nbeams, ld, beam_stride, col0, nspectra);
  if (rc != MDSM_OK) return rc;
  const int k = r->slot;
  if (r->pending[k]) {  // the slot's previous copy must have been consumed
    MDSM_HIP(hipEventSynchronize(r->done[k]));
    r->pending[k] = false;
  }
  const size_t want = r->spec_bytes * nspectra;
  const size_t have = fread(r->host[k], 1, want, r->f);
  MDSM_REQUIRE(have == want || !ferror(r->f), MDSM_ERR_IO, "read error");
  const uint32_t n = (uint32_t)(have / r->spec_bytes);
  // a partial spectrum at the end of the file is dropped, as readBinaryData does
  if (n) {
    hipStream_t s = as_stream(stream);
    MDSM_HIP(hipMemcpyAsync(r->dev[k], r->host[k], r->spec_bytes * n, hipMemcpyHostToDevice, s));
    rc = launch_unpack_corner_turn(r->dev[k], r->nbits, n, r->nchans, r->nbeams, d_out, ld,
                                   beam_stride, col0, s);
    if (rc != MDSM_OK) return rc;
    MDSM_HIP(hipEventRecord(r->done[k], s));
    r->pending[k] = true;
    r->slot ^= 1;
  }
  *got = n;
  return MDSM_OK;
}

// ---------------------------------------------------------------- writer
// Per-DM SIGPROC output (DedispersedDataWriter,
// src/pelican-mdsm/modules/DedispersedDataWriter.cpp:17-135): one file
// <prefix>_<dm>.fil per requested DM (the DM printed as %g), a header of
// machine_id 0, telescope_id 0, data_type 1, fch1, foff, nchans 1, tsamp,
// nbits 32, tstart, nifs 1 in that order, then the FP32 samples of the
// trial whose DM is within 1e-5 of the requested one, appended per chunk.
struct mdsm_dm_writer_s {
  std::vector<FILE *> files;
  std::vector<float> dms;
};

namespace {
void put_string(FILE *f, const char *s) {
  const int32_t n = (int32_t)std::strlen(s);
  fwrite(&n, 4, 1, f);
  fwrite(s, 1, (size_t)n, f);
}
void put_int(FILE *f, const char *k, int32_t v) {
  put_string(f, k);
  fwrite(&v, 4, 1, f);
}
void put_double(FILE *f, const char *k, double v) {
  put_string(f, k);
  fwrite(&v, 8, 1, f);
}
}  // namespace

extern "C" int mdsm_dm_writer_open(const char *prefix, const float *dms, uint32_t ndms, double fch1,
                                   double foff, double tsamp, double tstart, mdsm_dm_writer *out) {
  MDSM_REQUIRE(prefix && dms && ndms && out, MDSM_ERR_ARG, "null argument");
  *out = nullptr;
  auto *w = new mdsm_dm_writer_s();
  for (uint32_t i = 0; i < ndms; ++i) {
    char path[4096];
    snprintf(path, sizeof(path), "%s_%g.fil", prefix, (double)dms[i]);
    FILE *f = fopen(path, "wb");
    if (!f) {
      for (FILE *g : w->files) fclose(g);
      delete w;
      set_error("cannot create %s", path);
      return MDSM_ERR_IO;
    }
    put_string(f, "HEADER_START");
    put_int(f, "machine_id", 0);
    put_int(f, "telescope_id", 0);
    put_int(f, "data_type", 1);
    put_double(f, "fch1", fch1);
    put_double(f, "foff", foff);
    put_int(f, "nchans", 1);
    put_double(f, "tsamp", tsamp);
    put_int(f, "nbits", 32);
    put_double(f, "tstart", tstart);
    put_int(f, "nifs", 1);
    put_string(f, "HEADER_END");
    fflush(f);
    w->files.push_back(f);
    w->dms.push_back(dms[i]);
  }
  *out = w;
  return MDSM_OK;
}

extern "C" int mdsm_dm_writer_write(mdsm_dm_writer w, const float *series, uint32_t nrows,
                                    uint32_t nsamp, float lowdm, float dmstep) {
  MDSM_REQUIRE(w && (series || !nrows), MDSM_ERR_ARG, "null argument");
  for (size_t i = 0; i < w->files.size(); ++i)
    for (uint32_t j = 0; j < nrows; ++j) {
      const float dm = lowdm + dmstep * (float)j;
      if (std::fabs(dm - w->dms[i]) < 0.00001) {
        MDSM_REQUIRE(fwrite(series + (uint64_t)j * nsamp, 4, nsamp, w->files[i]) == nsamp, MDSM_ERR_IO,
                     "write failed");
        fflush(w->files[i]);
        break;
      }
    }
  return MDSM_OK;
}

extern "C" int mdsm_dm_writer_close(mdsm_dm_writer w) {
  if (!w) return MDSM_OK;
  for (FILE *f : w->files) fclose(f);
  delete w;
  return MDSM_OK;
}
