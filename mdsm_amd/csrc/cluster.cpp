// cluster.cpp -- host-side candidate clustering.
//
// mdsm_dbscan2d: the 2-D DBSCAN on (DM index, sample) with a
// Chebyshev neighbourhood, O(n (2 eps + 1) log n) through a row index over
// candidates sorted by (dm_idx, sample).
// mdsm_dbscan_ref: the reference's 3-D box DBSCAN
// (src/mdsm-multibeam/dbscan.cpp:194-276, getNeighbours :390-415), kept with
// its semantics (strict bounds, self / assigned / visited points excluded
// from the neighbour count, float comparisons, ids from 1, noise -1) so that
// the compatibility path clusters exactly as MDSM did.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

using namespace mdsm;

extern "C" int mdsm_sort_candidates(mdsm_cand_t *cands, uint32_t n) {
  MDSM_REQUIRE(cands || n == 0, MDSM_ERR_ARG, "null candidates");
  std::sort(cands, cands + n, [](const mdsm_cand_t &a, const mdsm_cand_t &b) {
    return a.dm_idx != b.dm_idx ? a.dm_idx < b.dm_idx : a.sample < b.sample;
  });
  return MDSM_OK;
}

extern "C" int mdsm_dbscan2d(const mdsm_cand_t *c, uint32_t n, int eps, int minpts,
                             int32_t *label, uint32_t *nclusters) {
  MDSM_REQUIRE((c && label) || n == 0, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(eps >= 0 && minpts >= 1, MDSM_ERR_ARG, "bad eps/minpts");
  for (uint32_t i = 1; i < n; ++i)
    MDSM_REQUIRE(c[i - 1].dm_idx < c[i].dm_idx ||
                     (c[i - 1].dm_idx == c[i].dm_idx && c[i - 1].sample <= c[i].sample),
                 MDSM_ERR_ARG, "candidates must be sorted by (dm_idx, sample)");
  // row index: for each DM value present, the [begin, end) range
  std::vector<int32_t> rows;
  std::vector<uint32_t> rbeg;
  for (uint32_t i = 0; i < n; ++i)
    if (i == 0 || c[i].dm_idx != c[i - 1].dm_idx) {
      rows.push_back(c[i].dm_idx);
      rbeg.push_back(i);
    }
  rbeg.push_back(n);
  auto for_neighbours = [&](uint32_t p, auto &&fn) {
    const int32_t d = c[p].dm_idx, t = c[p].sample;
    auto it = std::lower_bound(rows.begin(), rows.end(), d - eps);
    for (; it != rows.end() && *it <= d + eps; ++it) {
      const size_t r = (size_t)(it - rows.begin());
      const mdsm_cand_t *b = c + rbeg[r], *e = c + rbeg[r + 1];
      const mdsm_cand_t *q = std::lower_bound(
          b, e, t - eps, [](const mdsm_cand_t &x, int32_t v) { return x.sample < v; });
      for (; q != e && q->sample <= t + eps; ++q) fn((uint32_t)(q - c));
    }
  };
  std::vector<char> core(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    int cnt = 0;
    for_neighbours(i, [&](uint32_t) { ++cnt; });
    core[i] = cnt >= minpts;
    label[i] = -1;
  }
  std::vector<uint32_t> queue;
  queue.reserve(n);
  int32_t nc = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!core[i] || label[i] != -1) continue;
    const int32_t id = nc++;
    label[i] = id;
    queue.clear();
    queue.push_back(i);
    for (size_t h = 0; h < queue.size(); ++h) {
      for_neighbours(queue[h], [&](uint32_t j) {
        if (label[j] != -1) return;
        label[j] = id;
        if (core[j]) queue.push_back(j);
      });
    }
  }
  if (nclusters) *nclusters = (uint32_t)nc;
  return MDSM_OK;
}

extern "C" int mdsm_cluster_peaks(const mdsm_cand_t *c, const int32_t *label, uint32_t n,
                                  uint32_t nclusters, mdsm_cand_t *peaks, uint32_t *sizes) {
  MDSM_REQUIRE((c && label) || n == 0, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(peaks || nclusters == 0, MDSM_ERR_ARG, "null peaks");
  std::vector<int64_t> best(nclusters, -1);
  if (sizes) std::fill(sizes, sizes + nclusters, 0u);
  for (uint32_t i = 0; i < n; ++i) {
    const int32_t l = label[i];
    if (l < 0) continue;
    MDSM_REQUIRE((uint32_t)l < nclusters, MDSM_ERR_ARG, "label out of range");
    if (sizes) sizes[l]++;
    if (best[l] < 0 || c[i].snr > c[best[l]].snr) best[l] = i;
  }
  for (uint32_t k = 0; k < nclusters; ++k) peaks[k] = best[k] >= 0 ? c[best[k]] : mdsm_cand_t{-1, -1, -1, 0.0f};
  return MDSM_OK;
}

static unsigned ref_neighbours(const double *tm, const double *dm, const double *snr,
                               const int32_t *cluster, unsigned n, unsigned index, char *nb,
                               const char *visited, float min_time, float min_dm, float min_snr) {
  unsigned count = 0;
  for (unsigned i = 0; i < n; ++i) {
    if (index == i || cluster[i] > 0) continue;
    const float x = (float)std::fabs(tm[index] - tm[i]);
    const float y = (float)std::fabs(dm[index] - dm[i]);
    const float z = (float)std::fabs(snr[index] - snr[i]);
    if (x < min_time && y < min_dm && z < min_snr && nb[i] == 0 && visited[i] == 0) {
      nb[i] = 1;
      ++count;
    }
  }
  return count;
}

extern "C" int mdsm_dbscan_ref(const double *tm, const double *dm, const double *snr, uint32_t n,
                               float min_time, float min_dm, float min_snr, uint32_t min_points,
                               int32_t *cluster, uint32_t *nclusters) {
  MDSM_REQUIRE((tm && dm && snr && cluster) || n == 0, MDSM_ERR_ARG, "null argument");
  std::vector<char> nb(n, 0), visited(n, 0);
  std::fill(cluster, cluster + n, 0);
  int32_t nc = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (visited[i]) continue;
    visited[i] = 1;
    unsigned found = ref_neighbours(tm, dm, snr, cluster, n, i, nb.data(), visited.data(),
                                    min_time, min_dm, min_snr);
    if (found < min_points) {
      cluster[i] = -1;
      std::fill(nb.begin(), nb.end(), 0);
      continue;
    }
    const int32_t id = ++nc;
    cluster[i] = id;
    // the reference rescans from index 0 for the next flagged neighbour
    uint32_t j = 0;
    while (found != 0) {
      j = 0;
      while (nb[j] == 0) ++j;
      if (!visited[j]) {
        visited[j] = 1;
        found += ref_neighbours(tm, dm, snr, cluster, n, j, nb.data(), visited.data(), min_time,
                                min_dm, min_snr);
      }
      if (cluster[j] == 0) cluster[j] = id;
      nb[j] = 0;
      --found;
    }
  }
  if (nclusters) *nclusters = (uint32_t)nc;
  return MDSM_OK;
}

// Cluster classification (Cluster::computeTransientProbability,
// src/mdsm-multibeam/dbscan.cpp:25-171) with the reference's float/double
// arithmetic: a DM-S/N histogram, its (quirky) 3-tap smoothing, the median
// DM of the bins within 0.9 of the peak, rejection rules (DM <= 2, invalid
// width) and the mean squared distance to the dispersion-smearing model
// sqrt(pi)/2 * erf(y)/y.  The reference keeps clusters scoring < 0.1.
extern "C" int mdsm_classify_cluster_ref(const double *tm, const double *dm, const double *snr,
                                         uint32_t n, float min_dm, float dm_step, uint32_t num_dms,
                                         float min_pulse_width, float fch1, float foff,
                                         uint32_t nchans, float *score, double *info) {
  MDSM_REQUIRE(tm && dm && snr && score && n > 0 && num_dms > 2 && dm_step > 0, MDSM_ERR_ARG,
               "bad arguments");
  std::vector<float> hist(num_dms, 0.f), smooth(num_dms, 0.f), model(num_dms, 0.f);
  for (uint32_t i = 0; i < n; ++i) {
    const int b = (int)std::round((dm[i] - min_dm) / dm_step);
    MDSM_REQUIRE(b >= 0 && (uint32_t)b < num_dms, MDSM_ERR_ARG, "point %u outside the DM range", i);
    hist[b] = (float)(hist[b] + snr[i]);
  }
  for (uint32_t i = 1; i + 1 < num_dms; ++i) smooth[i] = (float)((hist[i] + hist[i] + hist[i + 1]) / 3.0);
  for (uint32_t i = 1; i + 1 < num_dms; ++i) hist[i] = smooth[i];
  float peak = 0.f;
  for (float v : hist) peak = std::max(peak, v);
  const float inv_peak = (float)(1.0 / peak);
  std::vector<int> high;
  float level = 0.f;
  for (uint32_t i = 0; i < num_dms; ++i)
    if (hist[i] * inv_peak > 0.9) {
      high.push_back((int)i);
      level += hist[i];
    }
  const size_t h = high.size();
  level /= (float)h;
  const float max_dm = (h % 2 == 0) ? (float)((high[h / 2 - 1] + high[h / 2]) * 0.5 * dm_step + min_dm)
                                    : (float)high[h / 2] * dm_step + min_dm;
  float max_snr = 0.f, width = 0.f;
  double position = 0.0;
  float result = 1.f;
  do {
    if (max_dm <= 2.0) break;  // zero-DM RFI
    for (uint32_t i = 0; i < n; ++i)
      if (std::fabs(dm[i] - max_dm) < dm_step * 0.5 && snr[i] > max_snr) max_snr = (float)snr[i];
    double t_lo = FLT_MAX, t_hi = FLT_MIN;
    for (uint32_t i = 0; i < n; ++i)
      if (std::fabs(dm[i] - max_dm) < dm_step) {
        t_lo = std::min(t_lo, tm[i]);
        t_hi = std::max(t_hi, tm[i]);
      }
    width = (float)((t_hi - t_lo) * 1e3 * 0.4);
    position = t_lo + (t_hi - t_lo) / 2;
    if (!(width <= FLT_MAX && width >= FLT_MIN && width > min_pulse_width)) break;
    const float band = std::fabs(foff * (float)nchans);
    const float freq3 = (float)std::pow(1e-3 * fch1 - (band * 1e-3) * 0.5, 3);
    float fit_max = 0.f;
    for (uint32_t i = 0; i < num_dms; ++i) {
      const float x = (float)i * dm_step + min_dm - max_dm;
      const float y = (float)(6.91e-3 * x * (band / (width * freq3)));
      model[i] = (float)(std::sqrt(M_PI) * 0.5 * (1.0 / y) * std::erf(y));
      fit_max = (fit_max > model[i]) ? fit_max : model[i];
    }
    model[(int)std::round((max_dm - min_dm) / dm_step)] = 1.f;
    if ((fit_max - 1.0) > 0.1) break;
    float mse = 0.f;
    for (uint32_t i = 0; i < num_dms; ++i) {
      const float d = hist[i] / level - model[i];
      mse += d * d;
    }
    result = mse / (float)num_dms;
  } while (false);
  *score = result;
  if (info) {
    info[0] = max_dm;
    info[1] = max_snr;
    info[2] = width;
    info[3] = position;
  }
  return MDSM_OK;
}
