// cluster.cpp -- host-side candidate clustering.
//
// mdsm_dbscan2d: the north-star 2-D DBSCAN on (DM index, sample) with a
// Chebyshev neighbourhood, O(n (2 eps + 1) log n) through a row index over
// candidates sorted by (dm_idx, sample).
// mdsm_dbscan_ref: the reference's 3-D box DBSCAN
// (src/mdsm-multibeam/dbscan.cpp:194-276, getNeighbours :390-415), kept with
// its semantics (strict bounds, self / assigned / visited points excluded
// from the neighbour count, float comparisons, ids from 1, noise -1) so that
// the compatibility path clusters exactly as MDSM did.
#include <algorithm>
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
