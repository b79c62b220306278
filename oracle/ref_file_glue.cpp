// C entry points around the reference's own SIGPROC reader
// (src/mdsm-multibeam/file_handler.cpp: read_header :56, read_block :193),
// compiled from the reference sources by oracle/Makefile (target ref) into
// oracle/_ref/libref_file.so.  TEST INFRASTRUCTURE ONLY.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "file_handler.h"

extern "C" int ref_read_header(const char *path, double *fch1, double *foff, double *tsamp,
                               double *tstart, int *nchans, int *nbits, int *nifs,
                               int *total_bytes, char *source_name) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  FILE_HEADER *h = read_header(f);
  if (!h) {
    fclose(f);
    return 0;  // no header: raw data file
  }
  *fch1 = h->fch1;
  *foff = h->foff;
  *tsamp = h->tsamp;
  *tstart = h->tstart;
  *nchans = h->nchans;
  *nbits = h->nbits;
  *nifs = h->nifs;
  *total_bytes = h->total_bytes;
  strcpy(source_name, h->source_name);
  free(h);
  fclose(f);
  return 1;
}

// every numeric field in a fixed order (as doubles), the two strings and the
// frequency table; fields the file does not set hold whatever the
// reference's uninitialised allocation held.
extern "C" int ref_header_values(const char *path, double *v, char *source_name, char *rawdatafile,
                                 double *ftable, int ntable) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  FILE_HEADER *h = read_header(f);
  if (!h) {
    fclose(f);
    return 0;
  }
  const double vals[] = {(double)h->machine_id, (double)h->telescope_id, (double)h->data_type,
                         (double)h->nchans, (double)h->nbits, (double)h->nifs, (double)h->nbeams,
                         (double)h->ibeam, (double)h->barycentric, (double)h->pulsarcentric,
                         (double)h->nbins, (double)h->itmp, h->tstart, h->tsamp, h->fch1, h->foff,
                         h->refdm, h->az_start, h->za_start, h->src_raj, h->src_dej, h->period,
                         (double)h->npuls, (double)h->total_bytes, (double)ftell(f)};
  memcpy(v, vals, sizeof(vals));
  strcpy(source_name, h->source_name);
  strcpy(rawdatafile, h->rawdatafile);
  for (int i = 0; i < ntable && i < 4096; ++i) ftable[i] = h->frequency_table[i];
  free(h);
  fclose(f);
  return 1;
}

// read nread samples of nbits each, starting byte_offset bytes into the file
extern "C" long ref_read_block(const char *path, long byte_offset, int nbits, float *block,
                               unsigned long nread) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  fseek(f, byte_offset, SEEK_SET);
  const unsigned long n = read_block(f, nbits, block, nread);
  fclose(f);
  return (long)n;
}
