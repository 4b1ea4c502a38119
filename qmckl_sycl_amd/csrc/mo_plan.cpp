// Column order of the MO contraction.
//
// The contraction mo[(p,c)][i] = sum_k ao[(p,c)][k] C[i][k]
// (src/qmckl_mo_sycl.cpp:23-50) is dense in the reference.  MOs that are
// localised have zero coefficients on most AOs, but the zeros of MOs that
// sit next to each other in the caller's order rarely line up into the
// 16-MO x 16-AO blocks an MFMA tile covers.  This file picks an order of
// the MOs for the prepared coefficient copies in which they do: groups of
// 16 MOs whose nonzero coefficients cover few 16-AO blocks, the groups
// chained so that neighbouring groups (which the GEMM kernels hand to the
// four wavefronts of a workgroup side by side) cover similar blocks.  The
// kernels skip the blocks that are zero and store each column back at its
// MO's place, so the order changes no result bit.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <numeric>
#include <vector>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

constexpr int kGroup = 16;       // columns per MFMA column block
constexpr int64_t kWindow = 2048;  // candidates scanned per pick, in seed order

int64_t popcount_andnot(const uint64_t *a, const uint64_t *u, int64_t nw) {
  int64_t c = 0;
  for (int64_t w = 0; w < nw; ++w) c += __builtin_popcountll(a[w] & ~u[w]);
  return c;
}

}  // namespace

int64_t plan_mo_columns(const uint64_t *words, int64_t M, int64_t nw, int64_t nkb16,
                        std::vector<int32_t> &col_mo, std::vector<uint64_t> &blk) {
  const int64_t nblk = (M + kGroup - 1) / kGroup;
  col_mo.assign((size_t)M, 0);
  blk.assign((size_t)(nblk * nw), 0);
  if (M <= 0 || nw <= 0 || nkb16 <= 0) return 0;
  auto bits = [&](int64_t m) { return words + m * nw; };

  // Seed order: MOs by the mean index of their nonzero AO blocks.  AOs are
  // stored nucleus by nucleus, so this keeps MOs on nearby nuclei close.
  std::vector<double> centre((size_t)M, -1.0);
  for (int64_t m = 0; m < M; ++m) {
    double sum = 0.0, n = 0.0;
    for (int64_t w = 0; w < nw; ++w)
      for (uint64_t x = bits(m)[w]; x != 0; x &= x - 1) {
        sum += (double)(w * 64 + __builtin_ctzll(x));
        n += 1.0;
      }
    if (n > 0.0) centre[(size_t)m] = sum / n;
  }
  std::vector<int32_t> seed((size_t)M);
  std::iota(seed.begin(), seed.end(), 0);
  std::stable_sort(seed.begin(), seed.end(),
                   [&](int32_t a, int32_t b) { return centre[(size_t)a] < centre[(size_t)b]; });

  // Greedy groups: each starts from the first MO not yet taken (seed
  // order) and adds the MO that adds the fewest AO blocks to the group's
  // union (the first such in seed order).
  std::vector<char> taken((size_t)M, 0);
  std::vector<int32_t> grouped;
  grouped.reserve((size_t)M);
  std::vector<uint64_t> unions((size_t)(nblk * nw), 0);
  int64_t first = 0;
  for (int64_t g = 0; g < nblk; ++g) {
    while (taken[(size_t)seed[(size_t)first]]) ++first;
    uint64_t *U = unions.data() + g * nw;
    const int32_t m0 = seed[(size_t)first];
    taken[(size_t)m0] = 1;
    grouped.push_back(m0);
    std::copy(bits(m0), bits(m0) + nw, U);
    for (int n = 1; n < kGroup && (int64_t)grouped.size() < M; ++n) {
      int32_t best = -1;
      int64_t best_cost = INT64_MAX, scanned = 0;
      for (int64_t pos = first; pos < M && scanned < kWindow; ++pos) {
        const int32_t m = seed[(size_t)pos];
        if (taken[(size_t)m]) continue;
        ++scanned;
        const int64_t cost = popcount_andnot(bits(m), U, nw);
        if (cost < best_cost) {
          best_cost = cost;
          best = m;
          if (cost == 0) break;
        }
      }
      taken[(size_t)best] = 1;
      grouped.push_back(best);
      for (int64_t w = 0; w < nw; ++w) U[w] |= bits(best)[w];
    }
  }

  // Chain the full groups: from group 0, repeatedly the unchained group with
  // the largest |U_a & U_b| / |U_a | U_b| against the last one.  A partial
  // group (M not a multiple of 16) stays last, so columns >= M are padding.
  const int64_t nfull = M / kGroup;
  std::vector<int64_t> chain;
  chain.reserve((size_t)nblk);
  std::vector<char> used((size_t)nblk, 0);
  if (nfull > 0) {
    chain.push_back(0);
    used[0] = 1;
    for (int64_t step = 1; step < nfull; ++step) {
      const uint64_t *a = unions.data() + chain.back() * nw;
      int64_t best = -1;
      double best_score = -1.0;
      for (int64_t g = 0; g < nfull; ++g) {
        if (used[(size_t)g]) continue;
        const uint64_t *b = unions.data() + g * nw;
        int64_t in = 0, un = 0;
        for (int64_t w = 0; w < nw; ++w) {
          in += __builtin_popcountll(a[w] & b[w]);
          un += __builtin_popcountll(a[w] | b[w]);
        }
        const double score = un > 0 ? (double)in / (double)un : 1.0;
        if (score > best_score) {
          best_score = score;
          best = g;
        }
      }
      chain.push_back(best);
      used[(size_t)best] = 1;
    }
  }
  if (nblk > nfull) chain.push_back(nfull);

  // Balance the wavefronts: the FP64 contraction hands column block p of a
  // workgroup to wavefront p % 4 and waits at every 16-AO stage for the
  // wavefront with the most live blocks.  Swap full blocks between
  // wavefronts while that sum of per-stage maxima falls.
  {
    const int64_t nmov = nfull;  // the partial block, if any, stays last
    std::vector<int32_t> load((size_t)(4 * nkb16), 0);
    auto bit = [&](int64_t pos, int64_t st) {
      return (int32_t)((unions[(size_t)(chain[(size_t)pos] * nw + st / 64)] >> (st % 64)) & 1ull);
    };
    for (int64_t pos = 0; pos < nblk; ++pos)
      for (int64_t st = 0; st < nkb16; ++st) load[(size_t)((pos % 4) * nkb16 + st)] += bit(pos, st);
    auto stage_max = [&](int64_t st) {
      int32_t m = 0;
      for (int w = 0; w < 4; ++w) m = std::max(m, load[(size_t)(w * nkb16 + st)]);
      return m;
    };
    for (int sweep = 0; sweep < 8; ++sweep) {
      bool improved = false;
      for (int64_t p1 = 0; p1 < nmov; ++p1)
        for (int64_t p2 = p1 + 1; p2 < nmov; ++p2) {
          const int64_t w1 = p1 % 4, w2 = p2 % 4;
          if (w1 == w2) continue;
          int64_t delta = 0;
          for (int64_t st = 0; st < nkb16; ++st) {
            const int32_t d = bit(p2, st) - bit(p1, st);  // w1 gains d, w2 loses d
            if (d == 0) continue;
            const int32_t before = stage_max(st);
            load[(size_t)(w1 * nkb16 + st)] += d;
            load[(size_t)(w2 * nkb16 + st)] -= d;
            delta += stage_max(st) - before;
            load[(size_t)(w1 * nkb16 + st)] -= d;
            load[(size_t)(w2 * nkb16 + st)] += d;
          }
          if (delta < 0) {
            for (int64_t st = 0; st < nkb16; ++st) {
              const int32_t d = bit(p2, st) - bit(p1, st);
              load[(size_t)(w1 * nkb16 + st)] += d;
              load[(size_t)(w2 * nkb16 + st)] -= d;
            }
            std::swap(chain[(size_t)p1], chain[(size_t)p2]);
            improved = true;
          }
        }
      if (!improved) break;
    }
  }

  int64_t live = 0;
  for (int64_t k = 0; k < nblk; ++k) {
    const int64_t g = chain[(size_t)k];
    const int64_t n = std::min<int64_t>(kGroup, M - g * kGroup);
    for (int64_t j = 0; j < n; ++j) col_mo[(size_t)(k * kGroup + j)] = grouped[(size_t)(g * kGroup + j)];
    for (int64_t w = 0; w < nw; ++w) {
      blk[(size_t)(k * nw + w)] = unions[(size_t)(g * nw + w)];
      live += __builtin_popcountll(unions[(size_t)(g * nw + w)]);
    }
  }
  return live;
}

}  // namespace qmckl_amd
