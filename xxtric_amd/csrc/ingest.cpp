// ingest.cpp -- host-side graph ingest of libxtric (no device work).
//
//  * RGG of the reference (graph.hpp:745-1374, LCG mode, one rank or the
//    graph a run on P ranks builds): same LCG streams (utils.hpp:245-430),
//    same radius (graph.hpp:789-791), same
//    predicate sqrt(dx*dx + dy*dy) <= r evaluated with the same operand order
//    (graph.hpp:922-926).  The reference tests all n^2/2 pairs; here points
//    are binned into square cells of side slightly above r and only the 3x3
//    neighbouring cells are tested, which yields the identical edge set.
//    Options outside the reference (xt_rgg_generate_opts): the strict test
//    distance < r, and the connectivity radius sqrt(2 ln n / (pi n)).
//    This file is compiled with -ffp-contract=off so dx*dx + dy*dy rounds
//    exactly like the reference's host build (x86-64 baseline, no FMA).
//  * vite binary edge lists (graph.hpp:294-395).
//  * whitespace edge lists (cugraph/example-inputs/*.tsv) -> symmetric CSR.
//  * parameters of the seeded batch mix (log-uniform sizes, RGG / ER).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../../include/xtric.h"
#include "xt_guard.hpp"

namespace {

constexpr double kPi = 3.14159;          // deftypes.hpp:7
constexpr int64_t kMlcg = 2147483647;    // deftypes.hpp:14
constexpr int64_t kAlcg = 16807;         // deftypes.hpp:15

uint32_t seed_seq_first(uint32_t seed) {
  std::seed_seq seq({seed});
  std::vector<std::uint32_t> out(1);
  seq.generate(out.begin(), out.end());
  return out[0];
}

// 2x2 integer matrices (row-major) in wrapping 64-bit arithmetic.
struct M2 {
  uint64_t a[4];
};
M2 mul_wrap(const M2& x, const M2& y) {
  M2 c;
  c.a[0] = x.a[0] * y.a[0] + x.a[1] * y.a[2];
  c.a[1] = x.a[0] * y.a[1] + x.a[1] * y.a[3];
  c.a[2] = x.a[2] * y.a[0] + x.a[3] * y.a[2];
  c.a[3] = x.a[2] * y.a[1] + x.a[3] * y.a[3];
  return c;
}
M2 pow_wrap(M2 m, uint64_t k) {
  M2 r = {{1, 0, 0, 1}};
  for (; k; k >>= 1) {
    if (k & 1) r = mul_wrap(r, m);
    m = mul_wrap(m, m);
  }
  return r;
}

// First LCG number of rank r of a P-rank run.  The reference gives each rank
// the step matrix M = [[a, 0], [b, 1]] raised to its 2n/P numbers, multiplied
// without reduction mod 2^31-1 (utils.hpp:271-303), and combines the ranks'
// matrices by a parallel prefix (utils.hpp:308-340), so rank r starts from
// x0 * (M^(2n/P))^r in wrapping int64 arithmetic, reduced by C++'s signed
// remainder (utils.hpp:343-347).  The ranks' matrices are all powers of M,
// which commute, so the power is taken directly.
int64_t rank_first(int64_t x0, int64_t len, int r) {
  if (r == 0) return x0;
  const M2 step = {{(uint64_t)kAlcg, 0, 0, 1}};
  const M2 p = pow_wrap(pow_wrap(step, (uint64_t)len), (uint64_t)r);
  const int64_t t = (int64_t)((uint64_t)x0 * p.a[0] + p.a[2]);
  return t % kMlcg;
}

// LCG coordinates of a P-rank run, rank r's points at [r*n/P, (r+1)*n/P):
// X from the first n/P numbers of its stream, Y from the next n/P rescaled to
// [r/P, (r+1)/P) (utils.hpp:352-420, graph.hpp:836-872).
void lcg_coords(int64_t n, int nprocs, uint32_t seed, std::vector<double>& X, std::vector<double>& Y) {
  X.resize(n);
  Y.resize(n);
  const int64_t nl = n / nprocs, len = 2 * nl;
  const int64_t x0 = (int64_t)seed_seq_first(seed);
  const double mult = 1.0 / (double)(1.0 + (double)(kMlcg - 1));
  const double rec_np = (double)(1.0 / (double)nprocs);
  const double range = (1.0 / (double)nprocs);
  for (int p = 0; p < nprocs; p++) {
    int64_t r = rank_first(x0, len, p);
    const double lo = p * rec_np;
    for (int64_t i = 0; i < len; i++) {
      if (i > 0) r = (r * kAlcg) % kMlcg;
      const double d = (double)((double)std::fabs((double)r) * mult);
      if (i < nl)
        X[p * nl + i] = d;
      else
        Y[p * nl + i - nl] = lo + (double)(range * d);
    }
  }
}

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int hw_threads() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 1;
  return (int)std::min(16u, h);
}

template <typename Fn>
void parallel_for(int64_t n, Fn fn) {
  const int nt = (int)std::min<int64_t>(hw_threads(), std::max<int64_t>(1, n / 4096));
  if (nt <= 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) {
    const int64_t a = n * t / nt, b = n * (t + 1) / nt;
    th.emplace_back([=] { fn(a, b); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

double xt_rgg_radius(int64_t nv) {
  const double rc = std::sqrt((double)std::log(nv) / (double)(kPi * nv));
  const double rt = std::sqrt((double)2.0736 / (double)nv);
  return (rc + rt) / (double)2.0;
}

double xt_rgg_radius_connect(int64_t nv) {
  return std::sqrt(2.0 * std::log((double)nv) / (M_PI * (double)nv));
}

int xt_rgg_generate_opts(int64_t nv, int nprocs, uint32_t seed, double radius, uint32_t flags,
                         int64_t* offsets, int64_t* tails, int64_t cap, int64_t* ne_out) {
  return xt_guarded([&]() -> int {
    if (nv <= 0 || !offsets || !ne_out) return XT_ERR_ARG;
    if (flags & ~(XT_RGG_STRICT | XT_RGG_CONNECT)) return XT_ERR_ARG;
    // graph.hpp:765-793: a power-of-two rank count dividing nv, strips wider than r
    if (nprocs < 1 || (nprocs & (nprocs - 1)) != 0 || nv % nprocs != 0) return XT_ERR_ARG;
    const bool strict = (flags & XT_RGG_STRICT) != 0;
    const double r = radius > 0                  ? radius
                     : (flags & XT_RGG_CONNECT) ? xt_rgg_radius_connect(nv)
                                                : xt_rgg_radius(nv);
    if (!((1.0 / (double)nprocs) > r)) return XT_ERR_ARG;
    const int64_t nl = nv / nprocs;
    std::vector<double> X, Y;
    lcg_coords(nv, nprocs, seed, X, Y);
    double minx = X[0], maxx = X[0], miny = Y[0], maxy = Y[0];
    for (int64_t i = 1; i < nv; i++) {
      minx = std::min(minx, X[i]);
      maxx = std::max(maxx, X[i]);
      miny = std::min(miny, Y[i]);
      maxy = std::max(maxy, Y[i]);
    }
    // Cells of side cs > r: any pair within distance r lies in neighbouring cells.
    const double cs = r * 1.0000001 + 1e-300;
    const int64_t cap_cells = 1 << 24;
    int64_t gx = std::max<int64_t>(1, (int64_t)std::floor((maxx - minx) / cs) + 1);
    int64_t gy = std::max<int64_t>(1, (int64_t)std::floor((maxy - miny) / cs) + 1);
    double csx = cs, csy = cs;
    while (gx * gy > cap_cells) {  // very small radii: coarser cells, still correct
      csx *= 2;
      csy *= 2;
      gx = std::max<int64_t>(1, (int64_t)std::floor((maxx - minx) / csx) + 1);
      gy = std::max<int64_t>(1, (int64_t)std::floor((maxy - miny) / csy) + 1);
    }
    std::vector<int64_t> cell(nv), cstart(gx * gy + 1, 0), order(nv);
    for (int64_t i = 0; i < nv; i++) {
      int64_t cx = std::min(gx - 1, (int64_t)std::floor((X[i] - minx) / csx));
      int64_t cy = std::min(gy - 1, (int64_t)std::floor((Y[i] - miny) / csy));
      cell[i] = cy * gx + cx;
      cstart[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < gx * gy; c++) cstart[c + 1] += cstart[c];
    {
      std::vector<int64_t> cur(cstart.begin(), cstart.end() - 1);
      for (int64_t i = 0; i < nv; i++) order[cur[cell[i]]++] = i;  // ascending ids within a cell
    }
    std::vector<std::vector<int64_t>> rows(nv);
    parallel_for(nv, [&](int64_t a, int64_t b) {
      for (int64_t i = a; i < b; i++) {
        const int64_t cx = cell[i] % gx, cy = cell[i] / gx;
        auto& row = rows[i];
        for (int64_t yy = std::max<int64_t>(0, cy - 1); yy <= std::min(gy - 1, cy + 1); yy++)
          for (int64_t xx = std::max<int64_t>(0, cx - 1); xx <= std::min(gx - 1, cx + 1); xx++) {
            const int64_t c = yy * gx + xx;
            for (int64_t k = cstart[c]; k < cstart[c + 1]; k++) {
              const int64_t j = order[k];
              if (j == i) continue;
              // Between ranks the reference tests only pairs whose local
              // indices differ (its cross loops run j > i over the
              // neighbour's points and mirror each hit, graph.hpp:983-1066).
              if (nprocs > 1 && i / nl != j / nl && i % nl == j % nl) continue;
              const int64_t p = std::min(i, j), q = std::max(i, j);
              const double dx = X[p] - X[q];
              const double dy = Y[p] - Y[q];
              const double ed = std::sqrt(dx * dx + dy * dy);
              if (strict ? ed < r : ed <= r) row.push_back(j);
            }
          }
        std::sort(row.begin(), row.end());
      }
  });
  offsets[0] = 0;
  for (int64_t i = 0; i < nv; i++) offsets[i + 1] = offsets[i] + (int64_t)rows[i].size();
  *ne_out = offsets[nv];
  if (tails && cap >= offsets[nv])
    for (int64_t i = 0; i < nv; i++) std::copy(rows[i].begin(), rows[i].end(), tails + offsets[i]);
  return XT_OK;
  });
}

int xt_rgg_generate_ranks(int64_t nv, int nprocs, uint32_t seed, double radius, int64_t* offsets,
                          int64_t* tails, int64_t cap, int64_t* ne_out) {
  return xt_rgg_generate_opts(nv, nprocs, seed, radius, 0, offsets, tails, cap, ne_out);
}

int xt_rgg_generate(int64_t nv, uint32_t seed, double radius, int64_t* offsets, int64_t* tails,
                    int64_t cap, int64_t* ne_out) {
  return xt_rgg_generate_ranks(nv, 1, seed, radius, offsets, tails, cap, ne_out);
}

}  // extern "C"

// ---- vite ----------------------------------------------------------------
template <typename E, typename W>
static int vite_read_t(const char* path, int64_t* nv, int64_t* ne, int64_t* off, int64_t* tails) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return XT_ERR_ARG;
  E hdr[2];
  if (std::fread(hdr, sizeof(E), 2, f) != 2) {
    std::fclose(f);
    return XT_ERR_INPUT;
  }
  *nv = (int64_t)hdr[0];
  *ne = (int64_t)hdr[1];
  if (*nv < 0 || *ne < 0) {
    std::fclose(f);
    return XT_ERR_INPUT;
  }
  if (!off) {
    std::fclose(f);
    return XT_OK;
  }
  std::vector<E> o((size_t)*nv + 1);
  if (std::fread(o.data(), sizeof(E), o.size(), f) != o.size()) {
    std::fclose(f);
    return XT_ERR_INPUT;
  }
  // graph.hpp:386-388 rebases the offsets to 0.
  for (int64_t i = 0; i <= *nv; i++) off[i] = (int64_t)o[i] - (int64_t)o[0];
  if (!tails) {  // offsets only
    std::fclose(f);
    return XT_OK;
  }
  struct Rec {
    E tail;
    W w;
  };
  std::vector<Rec> buf(1 << 16);
  for (int64_t e = 0; e < *ne;) {
    const int64_t k = std::min<int64_t>(*ne - e, (int64_t)buf.size());
    if (std::fread(buf.data(), sizeof(Rec), (size_t)k, f) != (size_t)k) {
      std::fclose(f);
      return XT_ERR_INPUT;
    }
    if (tails)
      for (int64_t t = 0; t < k; t++) tails[e + t] = (int64_t)buf[t].tail;
    e += k;
  }
  std::fclose(f);
  return XT_OK;
}

template <typename E, typename W>
static int vite_write_t(const char* path, int64_t nv, const int64_t* off, const int64_t* tails) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return XT_ERR_ARG;
  const E hdr[2] = {(E)nv, (E)off[nv]};
  bool ok = std::fwrite(hdr, sizeof(E), 2, f) == 2;
  std::vector<E> o((size_t)nv + 1);
  for (int64_t i = 0; i <= nv; i++) o[i] = (E)off[i];
  ok = ok && std::fwrite(o.data(), sizeof(E), o.size(), f) == o.size();
  struct Rec {
    E tail;
    W w;
  };
  std::vector<Rec> buf(1 << 16);
  const int64_t ne = off[nv];
  for (int64_t e = 0; ok && e < ne;) {
    const int64_t k = std::min<int64_t>(ne - e, (int64_t)buf.size());
    for (int64_t t = 0; t < k; t++) {
      buf[t].tail = (E)tails[e + t];
      buf[t].w = (W)1.0;
    }
    ok = std::fwrite(buf.data(), sizeof(Rec), (size_t)k, f) == (size_t)k;
    e += k;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? XT_OK : XT_ERR_INPUT;
}

extern "C" {

int xt_vite_read(const char* path, int layout32, int64_t* nv, int64_t* ne, int64_t* offsets,
                 int64_t* tails) {
  return xt_guarded([&]() -> int {
    if (!path || !nv || !ne) return XT_ERR_ARG;
    return layout32 ? vite_read_t<int32_t, float>(path, nv, ne, offsets, tails)
                    : vite_read_t<int64_t, double>(path, nv, ne, offsets, tails);
  });
}

int xt_vite_write(const char* path, int layout32, int64_t nv, const int64_t* offsets,
                  const int64_t* tails) {
  return xt_guarded([&]() -> int {
    if (!path || nv < 0 || !offsets || (offsets[nv] > 0 && !tails)) return XT_ERR_ARG;
    return layout32 ? vite_write_t<int32_t, float>(path, nv, offsets, tails)
                    : vite_write_t<int64_t, double>(path, nv, offsets, tails);
  });
}

// ---- whitespace edge list -> symmetric CSR ----------------------------------
int xt_edgelist_to_csr(const char* path, int64_t* nv_out, int64_t* ne_out, int64_t* offsets,
                       int64_t* tails) {
  return xt_guarded([&]() -> int {
    if (!path || !nv_out || !ne_out) return XT_ERR_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return XT_ERR_ARG;
    std::vector<uint64_t> keys;
    std::vector<std::pair<int64_t, int64_t>> ed;
    char line[512];
    int64_t maxid = -1;
    while (std::fgets(line, sizeof(line), f)) {
      if (line[0] == '#' || line[0] == '%') continue;
      long long a, b;
      if (std::sscanf(line, "%lld %lld", &a, &b) != 2) continue;
      if (a < 0 || b < 0) {
        std::fclose(f);
        return XT_ERR_INPUT;
      }
      ed.emplace_back(a, b);
      maxid = std::max<int64_t>(maxid, std::max<int64_t>(a, b));
    }
    std::fclose(f);
    const int64_t nv = maxid + 1;
    std::vector<std::pair<int64_t, int64_t>> all;
    all.reserve(ed.size() * 2);
    for (auto& p : ed) {
      all.push_back(p);
      all.emplace_back(p.second, p.first);
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    *nv_out = nv;
    *ne_out = (int64_t)all.size();
    if (offsets) {
      std::fill(offsets, offsets + nv + 1, 0);
      for (auto& p : all) offsets[p.first + 1]++;
      for (int64_t i = 0; i < nv; i++) offsets[i + 1] += offsets[i];
    }
    if (tails)
      for (size_t e = 0; e < all.size(); e++) tails[e] = all[e].second;
    return XT_OK;
  });
}

// ---- batch mix -----------------------------------------------------------
int xt_batch_params(uint64_t base_seed, int64_t g, int64_t nmin, int64_t nmax, int64_t* n, int* kind,
                    uint64_t* gseed, uint64_t* er_thr64) {
  return xt_guarded([&]() -> int {
    if (nmin < 2 || nmax < nmin || g < 0) return XT_ERR_ARG;
    const uint64_t gk = mix64(mix64(base_seed) + (uint64_t)g);
    const double u1 = (double)(mix64(gk + 1) >> 11) * 0x1.0p-53;
    const double ln0 = std::log((double)nmin), ln1 = std::log((double)nmax);
    int64_t nn = (int64_t)std::llround(std::exp(ln0 + u1 * (ln1 - ln0)));
    nn = std::min(std::max(nn, nmin), nmax);
    const double u2 = (double)(mix64(gk + 2) >> 11) * 0x1.0p-53;
    const double deg = 4.0 + 60.0 * u2;
    const double p = deg / (double)(nn - 1);
    if (n) *n = nn;
    if (kind) *kind = (int)(g & 1);
    if (gseed) *gseed = gk;
    if (er_thr64) *er_thr64 = p >= 1.0 ? (1ULL << 32) : (uint64_t)std::floor(p * 4294967296.0);
    return XT_OK;
  });
}

int xt_batch_graph_csr(uint64_t base_seed, int64_t g, int64_t nmin, int64_t nmax, int64_t* nv,
                       int64_t* ne, int64_t* offsets, int32_t* tails) {
  return xt_guarded([&]() -> int {
    int64_t n;
    int kind;
    uint64_t gseed, thr;
    int rc = xt_batch_params(base_seed, g, nmin, nmax, &n, &kind, &gseed, &thr);
    if (rc != XT_OK) return rc;
    if (nv) *nv = n;
    if (kind == 0) {
      std::vector<int64_t> off(n + 1);
      int64_t m = 0;
      const uint32_t s = (uint32_t)(gseed ^ (gseed >> 32));
      rc = xt_rgg_generate(n, s, 0.0, off.data(), nullptr, 0, &m);
      if (rc != XT_OK) return rc;
      if (ne) *ne = m;
      if (offsets) std::copy(off.begin(), off.end(), offsets);
      if (tails) {
        std::vector<int64_t> t(m > 0 ? m : 1);
        rc = xt_rgg_generate(n, s, 0.0, off.data(), t.data(), m, &m);
        for (int64_t e = 0; e < m; e++) tails[e] = (int32_t)t[e];
      }
      return rc;
    }
    // Erdos-Renyi: edge {i<j} iff (mix64(mix64(gseed+3) + i*n + j) >> 32) < thr.
    const uint64_t key = mix64(gseed + 3);
    std::vector<std::vector<int32_t>> rows(n);
    parallel_for(n, [&](int64_t a, int64_t b) {
      for (int64_t i = a; i < b; i++)
        for (int64_t j = 0; j < n; j++) {
          if (j == i) continue;
          const uint64_t p = (uint64_t)std::min(i, j), q = (uint64_t)std::max(i, j);
          if ((mix64(key + p * (uint64_t)n + q) >> 32) < thr) rows[i].push_back((int32_t)j);
        }
  });
  int64_t m = 0;
  for (auto& r : rows) m += (int64_t)r.size();
  if (ne) *ne = m;
  if (offsets) {
    offsets[0] = 0;
    for (int64_t i = 0; i < n; i++) offsets[i + 1] = offsets[i] + (int64_t)rows[i].size();
  }
  if (tails && offsets)
    for (int64_t i = 0; i < n; i++) std::copy(rows[i].begin(), rows[i].end(), tails + offsets[i]);
  return XT_OK;
  });
}

}  // extern "C"
