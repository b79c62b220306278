// survey_xml.cpp -- <observation> XML reader for the survey parameters,
// without Qt.  Restates processSurveyParameters
// (src/mdsm-multibeam/multibeam_dedispersion_manager.cpp:69-321, defaults
// :114-147) and accepts the older attribute names of the legacy parser
// (src/mdsm/dedispersion_manager.cpp:104-185) that the shipped configs use
// (config/gbncc.xml, src/mdsm-multibeam/TestFile.xml): channels@number,
// samples@number/bits, detection@threshold, gpus@ids, output@..., and the
// clustering typo minPuleWidth.  Where the reference read a missing
// multibeam attribute as 0 (detectionThreshold, nsamp), the value given
// under the legacy name, else the default, is used.
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "mdsm_abi.h"
#include "mdsm_compat.h"

namespace mdsm_xml {

struct Element {
  std::string name;
  std::map<std::string, std::string> attrs;  // lower-cased keys
  std::vector<std::unique_ptr<Element>> children;
};

static std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  return s;
}

// Minimal XML reader: elements and attributes (text content ignored),
// comments, processing instructions and doctype skipped.
class Parser {
 public:
  explicit Parser(const std::string &t) : s_(t) {}
  std::unique_ptr<Element> parse(std::string &err) {
    std::unique_ptr<Element> root;
    while (skip_to_tag()) {
      if (starts("<?")) { skip_past("?>"); continue; }
      if (starts("<!--")) { skip_past("-->"); continue; }
      if (starts("<!")) { skip_past(">"); continue; }
      root = element(err);
      break;
    }
    if (!root && err.empty()) err = "no root element";
    return root;
  }

 private:
  const std::string &s_;
  size_t p_ = 0;
  bool starts(const char *x) const { return s_.compare(p_, strlen(x), x) == 0; }
  void skip_ws() { while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_; }
  bool skip_to_tag() {
    while (p_ < s_.size() && s_[p_] != '<') ++p_;
    return p_ < s_.size();
  }
  void skip_past(const char *x) {
    size_t q = s_.find(x, p_);
    p_ = q == std::string::npos ? s_.size() : q + strlen(x);
  }
  std::string name() {
    size_t b = p_;
    while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || strchr("_-:.", s_[p_]))) ++p_;
    return s_.substr(b, p_ - b);
  }
  std::unique_ptr<Element> element(std::string &err) {
    auto e = std::make_unique<Element>();
    ++p_;  // '<'
    e->name = lower(name());
    if (e->name.empty()) { err = "malformed tag at offset " + std::to_string(p_); return nullptr; }
    for (;;) {
      skip_ws();
      if (p_ >= s_.size()) { err = "unterminated tag <" + e->name + ">"; return nullptr; }
      if (starts("/>")) { p_ += 2; return e; }
      if (s_[p_] == '>') { ++p_; break; }
      std::string k = lower(name());
      if (k.empty()) { err = "malformed attribute in <" + e->name + ">"; return nullptr; }
      skip_ws();
      if (p_ >= s_.size() || s_[p_] != '=') { err = "attribute without value in <" + e->name + ">"; return nullptr; }
      ++p_;
      skip_ws();
      if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\'')) { err = "unquoted attribute in <" + e->name + ">"; return nullptr; }
      char q = s_[p_++];
      size_t b = p_;
      while (p_ < s_.size() && s_[p_] != q) ++p_;
      if (p_ >= s_.size()) { err = "unterminated attribute value"; return nullptr; }
      e->attrs[k] = s_.substr(b, p_ - b);
      ++p_;
    }
    for (;;) {  // children until </name>
      if (!skip_to_tag()) { err = "missing </" + e->name + ">"; return nullptr; }
      if (starts("<!--")) { skip_past("-->"); continue; }
      if (starts("</")) {
        p_ += 2;
        std::string n = lower(name());
        skip_past(">");
        if (n != e->name) { err = "mismatched </" + n + "> for <" + e->name + ">"; return nullptr; }
        return e;
      }
      auto c = element(err);
      if (!c) return nullptr;
      e->children.push_back(std::move(c));
    }
  }
};

}  // namespace mdsm_xml

using mdsm_xml::Element;

static thread_local std::string g_xml_err;
static int g_dialect = 0;
static int g_brute_force = 1;  // dm@useBruteForce of the last file read (legacy dialect)

extern "C" const char *mdsm_compat_xml_error(void) { return g_xml_err.c_str(); }
extern "C" int mdsm_compat_xml_dialect(void) { return g_dialect; }

namespace {

struct Attrs {
  const Element *e;
  int *dialect;
  bool second_is_multibeam = false;  // clustering/writer: both spellings are multibeam names
  // first present of (multibeam name, legacy name); marks the dialect used
  const std::string *get(const char *mb, const char *legacy = nullptr) const {
    auto it = e->attrs.find(mdsm_xml::lower(mb));
    if (it != e->attrs.end() && !it->second.empty()) { *dialect |= MDSM_XML_MULTIBEAM; return &it->second; }
    if (legacy) {
      it = e->attrs.find(mdsm_xml::lower(legacy));
      if (it != e->attrs.end() && !it->second.empty()) {
        *dialect |= second_is_multibeam ? MDSM_XML_MULTIBEAM : MDSM_XML_LEGACY;
        return &it->second;
      }
    }
    return nullptr;
  }
  void f(float &dst, const char *mb, const char *legacy = nullptr) const {
    if (auto v = get(mb, legacy)) dst = (float)std::strtod(v->c_str(), nullptr);
  }
  void u(unsigned &dst, const char *mb, const char *legacy = nullptr) const {
    if (auto v = get(mb, legacy)) dst = (unsigned)std::strtoul(v->c_str(), nullptr, 10);
  }
  void b(bool &dst, const char *mb, const char *legacy = nullptr) const {
    if (auto v = get(mb, legacy)) dst = std::strtol(v->c_str(), nullptr, 10) != 0;
  }
  void c(char &dst, const char *mb, const char *legacy = nullptr) const {
    if (auto v = get(mb, legacy)) dst = (char)std::strtol(v->c_str(), nullptr, 10);
  }
  void s(char *dst, size_t n, const char *mb, const char *legacy = nullptr) const {
    if (auto v = get(mb, legacy)) { strncpy(dst, v->c_str(), n - 1); dst[n - 1] = 0; }
  }
};

std::vector<unsigned> split_uints(const std::string &v) {
  std::vector<unsigned> out;
  std::stringstream ss(v);
  std::string tok;
  while (std::getline(ss, tok, ','))
    if (!tok.empty()) out.push_back((unsigned)std::strtoul(tok.c_str(), nullptr, 10));
  return out;
}

}  // namespace

SURVEY *processSurveyParameters(const std::string &filepath) {
  g_xml_err.clear();
  std::ifstream in(filepath);
  if (!in) { g_xml_err = "Cannot open observation file '" + filepath + "'"; return nullptr; }
  std::stringstream buf;
  buf << in.rdbuf();
  const std::string text = buf.str();
  mdsm_xml::Parser parser(text);
  std::string err;
  auto root = parser.parse(err);
  if (!root) { g_xml_err = "parse error: " + err; return nullptr; }
  if (root->name != "observation") {
    g_xml_err = "Invalid root element, should be 'observation'";
    return nullptr;
  }
  auto *s = (SURVEY *)calloc(1, sizeof(SURVEY));
  // defaults (multibeam_dedispersion_manager.cpp:114-147)
  s->nantennas = 32;
  s->detection_threshold = 5.0f;
  s->ncoeffs = 8;
  s->spectrum_thresh = 6;
  s->channel_thresh = 10;
  s->channel_block = 1024;
  strcpy(s->fileprefix, "output");
  strcpy(s->basedir, ".");
  s->secs_per_file = 600;
  s->use_pc_time = 1;
  s->dbscan_min_points = 1;
  s->dbscan_time_range = 1;
  s->dbscan_dm_range = 1;
  s->dbscan_snr_range = 1;
  s->min_pulse_width = 1;
  s->output_bits = 32;
  s->output_compression = 1;
  s->npols = 1;
  int dialect = 0;
  g_brute_force = 1;
  std::vector<BEAM> beams;
  float top = 0, offset = 0;
  bool have_freq = false;
  for (const auto &ch : root->children) {
    const Element *e = ch.get();
    Attrs a{e, &dialect};
    const std::string &n = e->name;
    if (n == "dm") {
      auto bf = e->attrs.find("usebruteforce");
      if (bf != e->attrs.end() && !bf->second.empty())
        g_brute_force = std::strtol(bf->second.c_str(), nullptr, 10) != 0;
      a.f(s->lowdm, "lowDM");
      a.u(s->tdms, "numDMs");
      a.f(s->dmstep, "dmStep");
    } else if (n == "channels") {
      a.u(s->nchans, "nchans", "number");
      a.u(s->npols, "npols");
      a.u(s->ncoeffs, "ncoeffs");
    } else if (n == "antennas") {
      a.u(s->nantennas, "number");
    } else if (n == "timing") {
      a.f(s->tsamp, "tsamp");
    } else if (n == "samples") {
      a.u(s->nsamp, "nsamp", "number");
      a.u(s->nbits, "nbits", "bits");
      a.b(s->voltage, "voltage");
    } else if (n == "frequencies") {
      a.f(top, "top");
      a.f(offset, "offset");
      have_freq = true;
      dialect |= MDSM_XML_LEGACY;
    } else if (n == "rfi") {
      a.b(s->apply_rfi_clipper, "applyRFIClipper");
      a.f(s->spectrum_thresh, "spectrumThreshold");
      a.f(s->channel_thresh, "channelThreshold");
      a.u(s->channel_block, "channelBlock");
      if (auto v = a.get("channelMask")) {
        std::vector<RANGE> masks;
        std::stringstream ss(*v);
        std::string tok;
        while (std::getline(ss, tok, ',')) {
          if (tok.empty()) continue;
          RANGE r;
          size_t dash = tok.find('-');
          r.from = (unsigned)std::strtoul(tok.substr(0, dash).c_str(), nullptr, 10);
          r.to = dash == std::string::npos ? r.from : (unsigned)std::strtoul(tok.substr(dash + 1).c_str(), nullptr, 10);
          masks.push_back(r);
        }
        s->num_masks = (unsigned)masks.size();
        s->channel_mask = (RANGE *)malloc(sizeof(RANGE) * std::max<size_t>(1, masks.size()));
        std::copy(masks.begin(), masks.end(), s->channel_mask);
      }
    } else if (n == "detection") {
      a.f(s->detection_threshold, "detectionThreshold", "threshold");
      a.b(s->apply_median_filter, "applyMedianFilter");
      a.b(s->apply_detrending, "applyDetrending");
      a.b(s->tbb_enabled, "enableTBB");
    } else if (n == "clustering") {
      a.second_is_multibeam = true;
      a.b(s->apply_clustering, "applyClustering");
      a.b(s->apply_classification, "applyClassification");
      a.u(s->dbscan_min_points, "minPoints", "clusteringMinPoints");
      a.f(s->dbscan_time_range, "timeRange", "clusteringTimeRange");
      a.f(s->dbscan_dm_range, "dmRange", "clusteringDmRange");
      a.f(s->dbscan_snr_range, "snrRange", "clusteringSnrRange");
      if (a.get("minPulseWidth", "clusteringMinPulseWidth")) a.f(s->min_pulse_width, "minPulseWidth", "clusteringMinPulseWidth");
      else a.f(s->min_pulse_width, "minPuleWidth");  // spelling used by config/gbncc.xml
    } else if (n == "writer" || n == "output") {
      a.second_is_multibeam = n == "writer";
      if (n == "output") dialect |= MDSM_XML_LEGACY;
      a.b(s->dump_to_disk, "writeToFile");
      a.u(s->output_bits, "outputBits");
      a.u(s->output_compression, "compression");
      a.s(s->fileprefix, sizeof(s->fileprefix), "filePrefix", "outputFilePrefix");
      a.s(s->basedir, sizeof(s->basedir), "baseDirectory", "outputBaseDirectory");
      a.u(s->secs_per_file, "secondsPerFile", "outputSecondsPerFile");
      a.c(s->use_pc_time, "usePCTime", "outputUsePCTime");
      a.c(s->single_file_mode, "singleFileMode", "outputSingleFileMode");
    } else if (n == "storage") {
      // extension: dedisperser input storage (no reference counterpart, so
      // it marks neither dialect)
      auto it = e->attrs.find("type");
      if (it != e->attrs.end() && !it->second.empty()) {
        const std::string *v = &it->second;
        std::string t = *v;
        for (auto &ch : t) ch = (char)std::tolower((unsigned char)ch);
        if (t == "fp32" || t == "f32" || t == "float32") s->input_storage = MDSM_F32;
        else if (t == "fp16" || t == "f16" || t == "float16" || t == "half") s->input_storage = MDSM_F16;
        else if (t == "u8" || t == "uint8") s->input_storage = MDSM_U8;
        else if (t == "fp16pairs" || t == "f16pairs" || t == "fp16_pairs") s->input_storage = MDSM_F16_PAIRS;
        else {
          g_xml_err = "unknown storage type: " + *v;
          free(s);
          return nullptr;
        }
      }
    } else if (n == "series") {
      // extension: whether next_chunk_multibeam returns the dedispersed
      // series (no reference counterpart, so it marks neither dialect)
      auto it = e->attrs.find("download");
      if (it != e->attrs.end()) {
        const std::string &v = it->second;
        if (v == "0" || v == "false") s->series_on_device = 1;
        else if (v == "1" || v == "true") s->series_on_device = 0;
        else {
          g_xml_err = "series download must be 0 or 1: " + v;
          free(s);
          return nullptr;
        }
      }
    } else if (n == "gpus") {
      if (auto v = a.get("gpuIDs", "ids")) {
        auto ids = split_uints(*v);
        s->num_gpus = (unsigned)ids.size();
        s->gpu_ids = (unsigned *)malloc(sizeof(unsigned) * std::max<size_t>(1, ids.size()));
        std::copy(ids.begin(), ids.end(), s->gpu_ids);
      }
    } else if (n == "beams") {
      a.b(s->apply_beamforming, "applyBeamforming");
      for (const auto &bc : e->children) {
        if (bc->name != "beam") continue;
        Attrs ba{bc.get(), &dialect};
        BEAM b{};
        ba.u(b.beam_id, "beamId");
        ba.f(b.fch1, "topFrequency");
        ba.f(b.foff, "frequencyOffset");
        beams.push_back(b);
      }
    }
  }
  if (beams.empty() && have_freq) {  // legacy single-band configuration
    BEAM b{};
    b.fch1 = top;
    b.foff = offset;
    beams.push_back(b);
  }
  if (!s->num_gpus) {
    s->num_gpus = 1;
    s->gpu_ids = (unsigned *)calloc(1, sizeof(unsigned));
  }
  s->nbeams = (unsigned)beams.size();
  s->beams = (BEAM *)calloc(std::max<size_t>(1, beams.size()), sizeof(BEAM));
  std::copy(beams.begin(), beams.end(), s->beams);
  if (s->dump_to_disk && s->tbb_enabled) s->dump_to_disk = 0;
  g_dialect = dialect;
  return s;
}

// C entry point (no GPU): per-beam survey parameters of an XML file in the
// reference's float/truncation delay convention, plus the detection settings.
extern "C" int mdsm_compat_parse_xml(const char *path, mdsm_survey_t *beams, uint32_t max_beams,
                                     uint32_t *nbeams, float *threshold, uint32_t *min_points,
                                     int *dialect) {
  SURVEY *s = processSurveyParameters(path ? std::string(path) : std::string());
  if (!s) return MDSM_ERR_ARG;
  if (nbeams) *nbeams = s->nbeams;
  for (uint32_t b = 0; b < s->nbeams && b < max_beams && beams; ++b) {
    mdsm_survey_t &o = beams[b];
    o.fch1 = s->beams[b].fch1;
    o.foff = s->beams[b].foff;
    o.tsamp = s->tsamp;
    o.lowdm = s->lowdm;
    o.dmstep = s->dmstep;
    o.nchans = s->nchans;
    o.ndms = s->tdms;
    o.nsamp = s->nsamp;
    o.delay_mode = MDSM_DELAY_REFERENCE;
  }
  if (threshold) *threshold = s->detection_threshold;
  if (min_points) *min_points = s->dbscan_min_points;
  if (dialect) *dialect = g_dialect;
  free(s->beams);
  free(s->gpu_ids);
  free(s->channel_mask);
  free(s);
  return MDSM_OK;
}

extern "C" int mdsm_parse_observation_xml(const char *path, mdsm_xml_params_t *out) {
  if (!out) return MDSM_ERR_ARG;
  SURVEY *s = processSurveyParameters(path ? std::string(path) : std::string());
  if (!s) return MDSM_ERR_ARG;
  memset(out, 0, sizeof(*out));
  if (s->nbeams) {
    out->fch1 = s->beams[0].fch1;
    out->foff = s->beams[0].foff;
  }
  out->tsamp = s->tsamp;
  out->lowdm = s->lowdm;
  out->dmstep = s->dmstep;
  out->detection_threshold = s->detection_threshold;
  out->nchans = s->nchans;
  out->tdms = s->tdms;
  out->nsamp = s->nsamp;
  out->nbits = s->nbits;
  out->use_brute_force = g_brute_force;
  strncpy(out->fileprefix, s->fileprefix, sizeof(out->fileprefix) - 1);
  strncpy(out->basedir, s->basedir, sizeof(out->basedir) - 1);
  out->secs_per_file = s->secs_per_file;
  out->use_pc_time = s->use_pc_time;
  out->single_file_mode = s->single_file_mode;
  out->dialect = g_dialect;
  free(s->beams);
  free(s->gpu_ids);
  free(s->channel_mask);
  free(s);
  return MDSM_OK;
}
