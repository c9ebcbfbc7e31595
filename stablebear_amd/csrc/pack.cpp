// Host-side packing of PCF sets into the device record layout, and tile
// planning. Replaces the reference's CSR flattening (init_pcf_data,
// include/sbear/cuda/cuda_pcf_data_manager.cuh:13-23; OffsetDataManager::init,
// offset_data_manager.hpp:42-68) and its block scheduler
// (cuda_block_scheduler.hpp:29-129).
#include "sbamd_internal.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <memory>
#include <sstream>
#include <thread>
#include <unordered_map>

namespace sbamd
{
  namespace
  {
    template <typename T>
    double tval(const sbamd_pcf_set& s, int64_t k)
    {
      return static_cast<double>(static_cast<const T*>(s.t)[k]);
    }

    double time_at(const sbamd_pcf_set& s, int64_t k)
    {
      return s.dtype == SBAMD_F64 ? tval<double>(s, k) : tval<float>(s, k);
    }

    double value_at(const sbamd_pcf_set& s, int64_t k)
    {
      return s.dtype == SBAMD_F64 ? static_cast<const double*>(s.v)[k]
                                  : static_cast<double>(static_cast<const float*>(s.v)[k]);
    }

    void check_header(const sbamd_pcf_set& s)
    {
      check_set_offsets(s.n, s.offsets, s.dtype);
      if (s.n > 0 && s.offsets[s.n] > 0 && (!s.t || !s.v))
        fail(SBAMD_ERR_INVALID, "PCF set time/value arrays are NULL");
    }
  }

  void check_set_offsets(int64_t n, const int64_t* offsets, int32_t dtype)
  {
    if (dtype != SBAMD_F32 && dtype != SBAMD_F64)
      fail(SBAMD_ERR_INVALID, "PCF set dtype must be SBAMD_F32 or SBAMD_F64");
    if (n < 0)
      fail(SBAMD_ERR_INVALID, "PCF set size must be non-negative");
    if (n > std::numeric_limits<int32_t>::max() / 2)
      fail(SBAMD_ERR_INVALID, "PCF set too large (more than 2^30 PCFs)");
    if (n > 0 && !offsets)
      fail(SBAMD_ERR_INVALID, "PCF set offsets are NULL");
    if (n == 0)
      return;
    if (offsets[0] != 0)
      fail(SBAMD_ERR_INVALID, "PCF set offsets must start at 0");
    for (int64_t i = 0; i < n; ++i)
      if (offsets[i + 1] < offsets[i])
        fail(SBAMD_ERR_INVALID, "PCF set offsets must be non-decreasing");
  }

  namespace
  {
    // First PCF of [b, e) that breaks a rule of validate_set, or -1.
    template <typename T>
    int64_t first_bad(const sbamd_pcf_set& s, int64_t b, int64_t e)
    {
      const T* t = static_cast<const T*>(s.t);
      for (int64_t i = b; i < e; ++i)
      {
        const int64_t k0 = s.offsets[i], k1 = s.offsets[i + 1];
        if (k0 == k1)
          continue;
        if (t[k0] != T(0))
          return i;
        T prev = t[k0];
        for (int64_t k = k0 + 1; k < k1; ++k)
        {
          const T tk = t[k];
          // rejects NaN and +-inf as well: only finite, non-decreasing times pass
          if (!(tk >= prev) || !(tk <= std::numeric_limits<T>::max()))
            return i;
          prev = tk;
        }
      }
      return -1;
    }
  }

  // Breakpoint rules of the reference's Pcf constructor
  // (src/python/pypcf_support.hpp:44-87): an empty PCF is the zero PCF;
  // times non-decreasing; first time exactly 0. Times must also be finite
  // (the walk's end-of-PCF sentinel is +inf). Large sets are scanned by up
  // to 16 threads over ranges of about equal points; the message is that of
  // the first offending PCF, as in a serial scan.
  void validate_set(const sbamd_pcf_set& s)
  {
    check_header(s);
    const int64_t M = s.n > 0 ? s.offsets[s.n] : 0;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nth = M >= (int64_t(1) << 20) ? static_cast<int>(std::min(16u, hw)) : 1;
    int64_t from = 0, to = s.n;
    if (nth > 1)
    {
      std::vector<int64_t> cut(static_cast<size_t>(nth) + 1, s.n), bad(static_cast<size_t>(nth), -1);
      cut[0] = 0;
      for (int k = 1; k < nth; ++k)
        cut[k] = std::lower_bound(s.offsets, s.offsets + s.n, M * k / nth) - s.offsets;
      auto scan = [&](int k) {
        bad[k] = s.dtype == SBAMD_F64 ? first_bad<double>(s, cut[k], cut[k + 1])
                                      : first_bad<float>(s, cut[k], cut[k + 1]);
      };
      std::vector<std::thread> th;
      for (int k = 1; k < nth; ++k)
        th.emplace_back(scan, k);
      scan(0);
      for (auto& x : th)
        x.join();
      from = to = s.n;
      for (int k = 0; k < nth; ++k)
        if (bad[k] >= 0)
        {
          from = bad[k];
          to = bad[k] + 1;
          break;
        }
    }
    for (int64_t i = from; i < to; ++i)
    {
      const int64_t b = s.offsets[i], e = s.offsets[i + 1];
      if (e == b)
        continue;
      double prev = time_at(s, b);
      if (!std::isfinite(prev))
        fail(SBAMD_ERR_INVALID, "Breakpoint times must be finite.");
      for (int64_t k = b + 1; k < e; ++k)
      {
        const double tk = time_at(s, k);
        if (!std::isfinite(tk))
          fail(SBAMD_ERR_INVALID, "Breakpoint times must be finite.");
        if (tk < prev)
          fail(SBAMD_ERR_INVALID, "Breakpoints must be given in non-decreasing time order.");
        prev = tk;
      }
      const double t0 = time_at(s, b);
      if (t0 != 0.0)
      {
        std::ostringstream os;
        os << "The first breakpoint must have time t=0 (got t=" << t0 << ").";
        fail(SBAMD_ERR_INVALID, os.str());
      }
    }
  }

  // Records per PCF after dropping zero-length pieces (a point is kept when
  // the next point has a later time, or it is the last point). Keeping the
  // last of equal times also realises the reference's start rule "last point
  // with t <= 0" (iterate_rectangles.hpp:23-35, 44-45).
  std::vector<int64_t> record_counts(const sbamd_pcf_set& s)
  {
    std::vector<int64_t> m(static_cast<size_t>(s.n));
    for (int64_t i = 0; i < s.n; ++i)
    {
      const int64_t b = s.offsets[i], e = s.offsets[i + 1];
      if (e == b)
      {
        m[i] = 1;
        continue;
      }
      int64_t c = 1;
      for (int64_t k = b; k + 1 < e; ++k)
        if (time_at(s, k) < time_at(s, k + 1))
          ++c;
      m[i] = c;
    }
    return m;
  }

  void sort_by_records(const std::vector<int64_t>& m, std::vector<int64_t>& perm)
  {
    perm.resize(m.size());
    std::iota(perm.begin(), perm.end(), int64_t(0));
    std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return m[a] < m[b]; });
  }

  void pack_set(const sbamd_pcf_set& s, PackedSet& out)
  {
    validate_set(s);
    out.n = s.n;
    out.points = s.n > 0 ? s.offsets[s.n] : 0;
    out.in_dtype = s.dtype;
    std::vector<int64_t> m = record_counts(s);
    sort_by_records(m, out.perm);
    out.roff.assign(static_cast<size_t>(s.n) + 1, 0);
    for (int64_t q = 0; q < s.n; ++q)
      out.roff[q + 1] = out.roff[q] + m[out.perm[q]];
    const int64_t R = out.roff[s.n];
    out.src.resize(static_cast<size_t>(R));
    out.tend64.resize(static_cast<size_t>(R));
    out.val64.resize(static_cast<size_t>(R));
    const double inf = std::numeric_limits<double>::infinity();
    for (int64_t q = 0; q < s.n; ++q)
    {
      const int64_t i = out.perm[q];
      const int64_t b = s.offsets[i], e = s.offsets[i + 1];
      int64_t r = out.roff[q];
      if (e == b)
      {
        out.src[r] = -1;
        out.tend64[r] = inf;
        out.val64[r] = 0.0;
        continue;
      }
      for (int64_t k = b; k < e; ++k)
      {
        const bool last = k + 1 == e;
        if (!last && !(time_at(s, k) < time_at(s, k + 1)))
          continue;
        out.src[r] = k;
        out.tend64[r] = last ? inf : time_at(s, k + 1) + 0.0; // + 0.0 maps -0.0 to +0.0
        out.val64[r] = value_at(s, k);
        ++r;
      }
    }
  }

  // Lower-triangle (symmetric) or full (cdist) tile grid in sorted index
  // space, longest tiles first: rows and columns are sorted by record count,
  // so descending block indices put the longest walks at the head of the
  // launch and the short ones fill the tail.
  namespace
  {
    // Tiles of R x C pairs over rows [r_lo, r_hi) x cols [c_lo, c_hi) of
    // sorted index space, longest first (the rows and columns are sorted by
    // record count, so the last blocks hold the longest PCFs). With `tri`
    // (r_lo == c_lo: a block of a symmetric plan) only pairs (a, b) with
    // a > b, or a >= b with include_diag, are wanted: a tile is kept when it
    // holds one, and flagged TILE_DIAG when it also holds an unwanted pair
    // (the kernel then checks each pair). R and C need not be equal.
    void tile_grid(const std::vector<int64_t>& m_rows, const std::vector<int64_t>& m_cols, int64_t r_lo,
                   int64_t r_hi, int64_t c_lo, int64_t c_hi, int R, int C, bool tri, bool include_diag,
                   uint16_t prob, int64_t fold_threshold, std::vector<Tile>& tiles)
    {
      if (r_hi <= r_lo || c_hi <= c_lo)
        return;
      const int64_t nbr = (r_hi - r_lo + R - 1) / R;
      const int64_t nbc = (c_hi - c_lo + C - 1) / C;
      // longest PCF of every row and column block, once
      auto block_maxima = [](const std::vector<int64_t>& m, int64_t lo, int64_t hi, int B, int64_t nb) {
        std::vector<int64_t> mx(static_cast<size_t>(nb), 0);
        for (int64_t k = lo; k < hi; ++k)
        {
          const size_t blk = static_cast<size_t>((k - lo) / B);
          mx[blk] = std::max(mx[blk], m[static_cast<size_t>(k)]);
        }
        return mx;
      };
      const std::vector<int64_t> rmx = block_maxima(m_rows, r_lo, r_hi, R, nbr);
      const std::vector<int64_t> cmx = block_maxima(m_cols, c_lo, c_hi, C, nbc);
      const int64_t below = include_diag ? 0 : 1; // a wanted pair has col <= row - below
      for (int64_t bi = nbr - 1; bi >= 0; --bi)
      {
        const int64_t r0 = r_lo + bi * R;
        const int64_t rn = std::min<int64_t>(R, r_hi - r0);
        int64_t bj_end = nbc - 1;
        if (tri)
        {
          // the last column block holding a wanted pair of this row block
          const int64_t cmax = r0 + rn - 1 - below;
          if (cmax < c_lo)
            continue;
          bj_end = std::min<int64_t>(bj_end, (cmax - c_lo) / C);
        }
        for (int64_t bj = bj_end; bj >= 0; --bj)
        {
          const int64_t c0 = c_lo + bj * C;
          const int64_t cn = std::min<int64_t>(C, c_hi - c0);
          Tile t;
          t.r0 = static_cast<int32_t>(r0);
          t.c0 = static_cast<int32_t>(c0);
          t.nr = static_cast<int16_t>(rn);
          t.nc = static_cast<int16_t>(cn);
          t.prob = prob;
          // an unwanted pair: the first row against the last column
          t.flags = tri && c0 + cn - 1 > r0 - below ? TILE_DIAG : 0;
          if (rmx[static_cast<size_t>(bi)] + cmx[static_cast<size_t>(bj)] - 1 >= fold_threshold)
            t.flags |= TILE_FOLD;
          tiles.push_back(t);
        }
      }
    }
  }

  void build_tiles(const std::vector<int64_t>& m_rows, const std::vector<int64_t>* m_cols, int R,
                   int C, bool include_diag, uint16_t prob, int64_t fold_threshold,
                   std::vector<Tile>& tiles)
  {
    const bool sym = m_cols == nullptr;
    const std::vector<int64_t>& mc = sym ? m_rows : *m_cols;
    tile_grid(m_rows, mc, 0, static_cast<int64_t>(m_rows.size()), 0, static_cast<int64_t>(mc.size()), R, C, sym,
              include_diag, prob, fold_threshold, tiles);
  }

  // Tiles over rows [r_lo, r_hi) x cols [c_lo, c_hi) of sorted index space
  // (absolute indices). With `tri` the region is the square r_lo == c_lo,
  // r_hi == c_hi of a symmetric plan and only its lower triangle is tiled.
  void build_tiles_region(const std::vector<int64_t>& m_rows, const std::vector<int64_t>& m_cols,
                          int64_t r_lo, int64_t r_hi, int64_t c_lo, int64_t c_hi, int R, int C, bool tri,
                          bool include_diag, uint16_t prob, int64_t fold_threshold, std::vector<Tile>& tiles)
  {
    if (tri && (r_lo != c_lo || r_hi != c_hi))
      fail(SBAMD_ERR_INVALID, "a triangular tile region must be square");
    tile_grid(m_rows, m_cols, r_lo, r_hi, c_lo, c_hi, R, C, tri, include_diag, prob, fold_threshold, tiles);
  }

  // Reorders tiles [b0, b1) of a launch with one workgroup per tile so that
  // each XCD works through supertiles, blocks of tiles that share their rows
  // and columns, and re-reads those PCFs from its own L2. A supertile of size
  // G is 2G tiles tall and G/2 wide (G x G for G = 4): on the 20,000-PCF L1
  // pdist with 16 x 32 tiles, 512 rows x 256 columns of PCFs took 41.5 ms
  // against 42.5 for 256 x 512 (16 x 16 tiles) and 41.7 for 256 x 256.
  // The dispatcher hands workgroup w to XCD w mod nxcd, so XCD x runs the
  // tiles at positions congruent to x: the supertiles, taken in the order of
  // their first tile (the longest first, as the list was), are dealt to the
  // XCDs in turn, and the XCDs' sequences are interleaved. Consecutive
  // supertiles of one XCD mostly lie in the same block row, so the rows stay
  // in its L2 as well. G is the largest of 16, 8 and 4 that still gives every
  // XCD four supertiles (with fewer, one XCD would hold a large share of the
  // longest tiles); a range too small for G = 4 keeps its order. Returns G
  // (0: order kept).
  int xcd_supertile_order(std::vector<Tile>& tiles, size_t b0, size_t b1, int R, int C, int nxcd)
  {
    if (nxcd <= 1 || b1 <= b0 + 1)
      return 0;
    const size_t n = b1 - b0;
    auto key = [&](const Tile& t, int G) {
      const int sr = (G > 4 ? 2 * G : G) * R, sc = (G > 4 ? G / 2 : G) * C;
      // powers of two for the tile sides in use: shifts
      const uint64_t r = (sr & (sr - 1)) == 0 ? static_cast<uint64_t>(t.r0) >> __builtin_ctz(sr)
                                              : static_cast<uint64_t>(t.r0 / sr);
      const uint64_t c = (sc & (sc - 1)) == 0 ? static_cast<uint64_t>(t.c0) >> __builtin_ctz(sc)
                                              : static_cast<uint64_t>(t.c0 / sc);
      return (static_cast<uint64_t>(t.prob) << 48) | (r << 24) | c;
    };
    std::vector<uint32_t> gid(n);
    for (int G : {16, 8, 4})
    {
      // supertile of every tile, numbered by first appearance (consecutive
      // tiles mostly share one, so the map is consulted once per run)
      std::unordered_map<uint64_t, uint32_t> id;
      uint64_t last = ~0ull;
      uint32_t last_id = 0;
      for (size_t k = 0; k < n; ++k)
      {
        const uint64_t kk = key(tiles[b0 + k], G);
        if (kk != last)
        {
          auto it = id.emplace(kk, static_cast<uint32_t>(id.size())).first;
          last = kk;
          last_id = it->second;
        }
        gid[k] = last_id;
      }
      const size_t ng = id.size();
      if (ng < 4 * static_cast<size_t>(nxcd))
        continue;
      // supertiles in first-appearance order, tiles in list order within
      std::vector<size_t> start(ng + 1, 0);
      for (size_t k = 0; k < n; ++k)
        ++start[gid[k] + 1];
      for (size_t g = 0; g < ng; ++g)
        start[g + 1] += start[g];
      // XCD x takes supertiles x, x + nxcd, ...: the rank of a tile in its
      // XCD's sequence is the tiles of that XCD's earlier supertiles plus its
      // rank in its own supertile
      const size_t nx = static_cast<size_t>(nxcd);
      std::vector<size_t> seq_before(ng), seq_len(nx, 0);
      for (size_t g = 0; g < ng; ++g)
      {
        seq_before[g] = seq_len[g % nx];
        seq_len[g % nx] += start[g + 1] - start[g];
      }
      // the launch order interleaves the XCDs' sequences: rank k of XCD x
      // goes after every XCD's first k tiles (as far as they are that long)
      // and the tiles of rank k of XCDs below x
      size_t maxlen = 0;
      for (size_t x = 0; x < nx; ++x)
        maxlen = std::max(maxlen, seq_len[x]);
      std::vector<size_t> slot_base(maxlen + 1, 0); // output position of rank k's first tile
      {
        std::vector<size_t> lens(seq_len);
        std::sort(lens.begin(), lens.end());
        size_t alive = nx, li = 0;
        for (size_t k = 0; k < maxlen; ++k)
        {
          while (li < nx && lens[li] <= k)
          {
            ++li;
            --alive;
          }
          slot_base[k + 1] = slot_base[k] + alive;
        }
      }
      const size_t minlen = *std::min_element(seq_len.begin(), seq_len.end());
      std::unique_ptr<Tile[]> out(new Tile[n]);
      std::vector<size_t> fill(start.begin(), start.end() - 1);
      for (size_t k = 0; k < n; ++k)
      {
        const size_t g = gid[k], x = g % nx;
        const size_t r = seq_before[g] + (fill[g]++ - start[g]);
        size_t below = x; // XCDs below x with a tile of rank r: all of them
        if (r >= minlen)  // unless some sequences have ended
        {
          below = 0;
          for (size_t y = 0; y < x; ++y)
            below += seq_len[y] > r;
        }
        out[slot_base[r] + below] = tiles[b0 + k];
      }
      std::copy(out.get(), out.get() + n, tiles.begin() + static_cast<std::ptrdiff_t>(b0));
      return G;
    }
    return 0;
  }

  void xcd_queue_order(std::vector<Tile>& tiles, size_t b0, size_t b1, int R, int C, int nxcd,
                       const std::vector<int64_t>& work, int64_t* qoff)
  {
    const size_t n = b1 - b0;
    const size_t nx = static_cast<size_t>(nxcd);
    for (size_t x = 0; x <= nx; ++x)
      qoff[x] = x == 0 ? 0 : static_cast<int64_t>(n);
    if (nxcd <= 1 || n == 0)
      return;
    // units: G x G tiles of one problem (they share G R rows and G C
    // columns), G the largest that leaves at least 8 units per queue
    std::vector<uint32_t> uid(n);
    size_t nu = 0;
    for (int G : {16, 8, 4, 2, 1})
    {
      const int64_t sr = static_cast<int64_t>(G) * R, sc = static_cast<int64_t>(G) * C;
      std::unordered_map<uint64_t, uint32_t> id;
      for (size_t k = 0; k < n; ++k)
      {
        const Tile& t = tiles[b0 + k];
        const uint64_t key = (static_cast<uint64_t>(t.prob) << 48) |
                             (static_cast<uint64_t>(t.r0 / sr) << 24) | static_cast<uint64_t>(t.c0 / sc);
        uid[k] = id.emplace(key, static_cast<uint32_t>(id.size())).first->second;
      }
      nu = id.size();
      if (nu >= 8 * nx || G == 1)
        break;
    }
    // unit work and its first tile in the (longest-first) list
    std::vector<int64_t> uwork(nu, 0);
    std::vector<size_t> ufirst(nu, n);
    for (size_t k = 0; k < n; ++k)
    {
      uwork[uid[k]] += work[k];
      ufirst[uid[k]] = std::min(ufirst[uid[k]], k);
    }
    // units by decreasing work, each to the queue with the least work so far
    std::vector<uint32_t> order(nu);
    for (size_t u = 0; u < nu; ++u)
      order[u] = static_cast<uint32_t>(u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      return uwork[a] != uwork[b] ? uwork[a] > uwork[b] : ufirst[a] < ufirst[b];
    });
    std::vector<int64_t> load(nx, 0);
    std::vector<uint32_t> uq(nu), urank(nu);
    std::vector<std::vector<uint32_t>> qunits(nx);
    for (uint32_t u : order)
    {
      const size_t x = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
      load[x] += uwork[u];
      uq[u] = static_cast<uint32_t>(x);
      urank[u] = static_cast<uint32_t>(qunits[x].size());
      qunits[x].push_back(u);
    }
    // layout: queue 0's units, then queue 1's, ...; a unit's tiles stay in list order
    std::vector<size_t> ucount(nu, 0), ustart(nu, 0);
    for (size_t k = 0; k < n; ++k)
      ++ucount[uid[k]];
    size_t at = 0;
    for (size_t x = 0; x < nx; ++x)
    {
      qoff[x] = static_cast<int64_t>(at);
      for (uint32_t u : qunits[x])
      {
        ustart[u] = at;
        at += ucount[u];
      }
    }
    qoff[nx] = static_cast<int64_t>(at);
    std::unique_ptr<Tile[]> out(new Tile[n]);
    for (size_t k = 0; k < n; ++k)
      out[ustart[uid[k]]++] = tiles[b0 + k];
    std::copy(out.get(), out.get() + n, tiles.begin() + static_cast<std::ptrdiff_t>(b0));
  }

}
