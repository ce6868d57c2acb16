// Cell-block sparsity patterns and probing colourings for compressed Jacobian assembly, and the
// device-side probe seeds and response scatter into block-ELL storage.
//
// Reference: block_stencil_colouring, _saturation_colouring and materialize_block_jacobian
// (aquaflux/solve/sparse_jacobian.py:338-534, 722-848).  The Jacobian couples cell i to cell j
// only when they lie within `reach` steps on the interior-face cell graph; two column cells may
// share a probe colour only if no row couples to both, i.e. when their graph distance exceeds
// 2 reach.  Colours are chosen greedily by saturation degree (distinct neighbour colours), ties by
// conflict degree.  Any collision-free colouring recovers the same matrix; the colour count is
// the number of directional derivatives per field that an assembly costs.
//
// The pattern and colouring are host work done once per mesh; the conflict neighbourhood (the
// ball of radius 2 reach) is enumerated by breadth-first search rather than stored.
#include "afx_common.h"
#include "afx_profile.h"
#include "aquaflux_amd.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

using namespace afx;

namespace {

struct CellGraph {
  int64_t n = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> adj;
};

// Undirected cell graph from interior-face edges (self-loops dropped, duplicate edges merged).
CellGraph build_graph(int64_t n, int64_t n_edges, const int32_t* a, const int32_t* b) {
  CellGraph g;
  g.n = n;
  g.ptr.assign(n + 1, 0);
  for (int64_t e = 0; e < n_edges; ++e) {
    if (a[e] == b[e]) continue;
    g.ptr[a[e] + 1]++;
    g.ptr[b[e] + 1]++;
  }
  for (int64_t v = 0; v < n; ++v) g.ptr[v + 1] += g.ptr[v];
  g.adj.resize(g.ptr[n]);
  std::vector<int64_t> fill(g.ptr.begin(), g.ptr.end() - 1);
  for (int64_t e = 0; e < n_edges; ++e) {
    if (a[e] == b[e]) continue;
    g.adj[fill[a[e]]++] = b[e];
    g.adj[fill[b[e]]++] = a[e];
  }
  // sort and merge duplicates per row, compacting in place
  int64_t out = 0;
  std::vector<int64_t> nptr(n + 1, 0);
  for (int64_t v = 0; v < n; ++v) {
    const int64_t lo = g.ptr[v], hi = g.ptr[v + 1];
    std::sort(g.adj.begin() + lo, g.adj.begin() + hi);
    nptr[v] = out;
    for (int64_t i = lo; i < hi; ++i)
      if (i == lo || g.adj[i] != g.adj[i - 1]) g.adj[out++] = g.adj[i];
  }
  nptr[n] = out;
  g.adj.resize(out);
  g.ptr.swap(nptr);
  return g;
}

// Breadth-first ball of radius r around v (v first, then by increasing distance); `level` (when
// given) receives each member's distance.
struct BallSearch {
  std::vector<int32_t> stamp;
  int32_t epoch = 0;
  explicit BallSearch(int64_t n) : stamp(n, 0) {}
  void ball(const CellGraph& g, int32_t v, int r, std::vector<int32_t>& out, std::vector<int8_t>* level = nullptr) {
    if (epoch == INT_MAX) {
      std::fill(stamp.begin(), stamp.end(), 0);
      epoch = 0;
    }
    ++epoch;
    out.clear();
    out.push_back(v);
    if (level) {
      level->clear();
      level->push_back(0);
    }
    stamp[v] = epoch;
    size_t lo = 0;
    for (int d = 0; d < r && lo < out.size(); ++d) {
      const size_t hi = out.size();
      for (size_t i = lo; i < hi; ++i) {
        const int32_t u = out[i];
        for (int64_t k = g.ptr[u]; k < g.ptr[u + 1]; ++k) {
          const int32_t w = g.adj[k];
          if (stamp[w] != epoch) {
            stamp[w] = epoch;
            out.push_back(w);
            if (level) level->push_back((int8_t)(d + 1));
          }
        }
      }
      lo = hi;
    }
  }
};

// Saturation-degree greedy colouring of the distance-(2 reach) conflict graph.  A bucket per rank
// (saturation * (max degree + 1) + degree) holds the uncoloured vertices at that rank as a doubly
// linked list; the highest occupied bucket is popped each round (last-in first), so a round costs
// one ball search around the coloured vertex.  Vertices enter their buckets in index order and
// neighbours are re-bucketed in ascending index order, as the reference's _saturation_colouring
// does (aquaflux/solve/sparse_jacobian.py:458-534), so the colouring is the reference's, vertex for
// vertex.
void saturation_colouring(const CellGraph& g, int reach, int32_t* colour, int32_t* n_colours) {
  const int64_t n = g.n;
  const int r2 = 2 * reach;
  BallSearch bs(n);
  std::vector<int32_t> ball;
  std::vector<int32_t> degree(n);
  int32_t max_degree = 0;
  for (int64_t v = 0; v < n; ++v) {
    bs.ball(g, (int32_t)v, r2, ball);
    degree[v] = (int32_t)ball.size() - 1;
    max_degree = std::max(max_degree, degree[v]);
  }
  const int64_t tie = (int64_t)max_degree + 1;
  const int64_t n_rank = tie * max_degree + max_degree + 1;
  const int words = max_degree / 64 + 2;
  std::vector<uint64_t> taken((size_t)n * words, 0);
  std::vector<int32_t> sat(n, 0);
  std::vector<int64_t> head(n_rank, -1);
  std::vector<int64_t> next(n, -1), prev(n, -1);
  for (int64_t v = 0; v < n; ++v) colour[v] = -1;
  auto push = [&](int64_t v, int64_t rank) {
    const int64_t h = head[rank];
    next[v] = h;
    prev[v] = -1;
    if (h >= 0) prev[h] = v;
    head[rank] = v;
  };
  auto unlink = [&](int64_t v, int64_t rank) {
    if (prev[v] >= 0)
      next[prev[v]] = next[v];
    else
      head[rank] = next[v];
    if (next[v] >= 0) prev[next[v]] = prev[v];
  };
  int64_t top = 0;
  for (int64_t v = 0; v < n; ++v) {
    push(v, degree[v]);
    top = std::max<int64_t>(top, degree[v]);
  }
  int32_t used = 0;
  for (int64_t round = 0; round < n; ++round) {
    while (head[top] < 0) --top;
    const int64_t v = head[top];
    unlink(v, top);
    int32_t chosen = 0;
    for (int w = 0; w < words; ++w) {
      const uint64_t free_bits = ~taken[(size_t)v * words + w];
      if (free_bits) {
        chosen = w * 64 + __builtin_ctzll(free_bits);
        break;
      }
    }
    colour[v] = chosen;
    used = std::max(used, chosen + 1);
    const int cw = chosen / 64;
    const uint64_t bit = 1ull << (chosen % 64);
    bs.ball(g, (int32_t)v, r2, ball);
    // neighbours move to their new buckets in ascending index order (the conflict row's order in
    // the reference), which fixes the last-in-first-out tie order and so the colouring itself
    std::sort(ball.begin() + 1, ball.end());
    for (size_t i = 1; i < ball.size(); ++i) {
      const int64_t u = ball[i];
      if (colour[u] >= 0) continue;
      uint64_t& t = taken[(size_t)u * words + cw];
      if (t & bit) continue;
      t |= bit;
      const int64_t r_old = (int64_t)sat[u] * tie + degree[u];
      sat[u]++;
      const int64_t r_new = r_old + tie;
      unlink(u, r_old);
      push(u, r_new);
      if (r_new > top) top = r_new;
    }
  }
  *n_colours = used;
}

// y[i] = (field == b && colour[cell] == k) ? 1 : 0 over a field-major vector of `block` fields.
__global__ void __launch_bounds__(256) k_probe_seed(int64_t C, int32_t B, const int32_t* __restrict__ colour,
                                                    int32_t k, int32_t b, double* __restrict__ seed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C * B) return;
  const int64_t v = i / C, j = i - v * C;
  seed[i] = (v == b && colour[j] == k) ? 1.0 : 0.0;
}

// Column `b` of every block whose column cell has colour k and lies within max_depth steps, read
// from the probe response.
template <class T>
__global__ void __launch_bounds__(256) k_probe_scatter(int32_t C, int32_t B, int32_t nslot,
                                                       const int32_t* __restrict__ cols,
                                                       const int8_t* __restrict__ depth, int32_t max_depth,
                                                       const int32_t* __restrict__ colour, int32_t k, int32_t b,
                                                       const double* __restrict__ resp, T* __restrict__ vals) {
  const int32_t c = xcd_thread_index();
  if (c >= C) return;
  for (int32_t s = 0; s < nslot; ++s) {
    const int32_t j = cols[(int64_t)s * C + c];
    if (s > 0 && j == c) continue;  // padding slot
    if (colour[j] != k) continue;
    if (depth && depth[(int64_t)s * C + c] > max_depth) continue;  // beyond this column's reach
    for (int32_t a = 0; a < B; ++a) vals[(((int64_t)s * B + a) * B + b) * C + c] = (T)resp[(int64_t)a * C + c];
  }
}

}  // namespace

extern "C" int afx_block_colouring(int64_t n_cells, int64_t n_edges, const int32_t* owner, const int32_t* neighbour,
                                   int32_t reach, int64_t* pattern_ptr, int32_t* pattern_cols, int8_t* pattern_depth,
                                   int32_t* colour, int32_t* n_colours) {
  AFX_REQUIRE(n_cells > 0 && n_cells < INT32_MAX && n_edges >= 0, "afx_block_colouring: bad sizes");
  AFX_REQUIRE(reach >= 1 && reach <= 64, "afx_block_colouring: reach must be 1 to 64");
  AFX_REQUIRE(pattern_ptr && (n_edges == 0 || (owner && neighbour)), "afx_block_colouring: null argument");
  AFX_REQUIRE(!colour || n_colours, "afx_block_colouring: colour needs n_colours");
  for (int64_t e = 0; e < n_edges; ++e)
    AFX_REQUIRE(owner[e] >= 0 && owner[e] < n_cells && neighbour[e] >= 0 && neighbour[e] < n_cells,
                "afx_block_colouring: edge endpoint out of range");
  const CellGraph g = build_graph(n_cells, n_edges, owner, neighbour);
  BallSearch bs(n_cells);
  std::vector<int32_t> ball;
  std::vector<int8_t> level;
  std::vector<int64_t> order;
  pattern_ptr[0] = 0;
  for (int64_t v = 0; v < n_cells; ++v) {
    bs.ball(g, (int32_t)v, reach, ball, &level);
    const int64_t at = pattern_ptr[v];
    pattern_ptr[v + 1] = at + (int64_t)ball.size();
    if (pattern_cols) {  // the cell itself first, then ascending
      order.resize(ball.size());
      for (size_t i = 0; i < ball.size(); ++i) order[i] = (int64_t)i;
      std::sort(order.begin() + 1, order.end(), [&](int64_t x, int64_t y) { return ball[x] < ball[y]; });
      for (size_t i = 0; i < ball.size(); ++i) {
        pattern_cols[at + i] = ball[order[i]];
        if (pattern_depth) pattern_depth[at + i] = level[order[i]];
      }
    }
  }
  if (colour) saturation_colouring(g, reach, colour, n_colours);
  return 0;
}

extern "C" int afx_probe_seed(int64_t n_cells, int32_t block, const int32_t* colour, int32_t k, int32_t field,
                              double* seed, hipStream_t s) {
  AFX_REQUIRE(n_cells > 0 && block > 0 && colour && seed && field >= 0 && field < block,
              "afx_probe_seed: bad argument");
  k_probe_seed<<<grid_for(n_cells * block), dim3(kBlock), 0, s>>>(n_cells, block, colour, k, field, seed);
  AFX_CHECK_LAUNCH();
  return 0;
}

extern "C" int afx_probe_scatter(int64_t n_cells, int32_t block, int32_t nslot, const int32_t* cols,
                                 const int8_t* depth, int32_t max_depth, const int32_t* colour, int32_t k,
                                 int32_t field, const double* response, int precision, void* vals, hipStream_t s) {
  AFX_REQUIRE(n_cells > 0 && n_cells < INT32_MAX && block > 0 && nslot > 0 && cols && colour && response && vals &&
                  field >= 0 && field < block,
              "afx_probe_scatter: bad argument");
  AFX_REQUIRE(precision == 0 || precision == 1, "afx_probe_scatter: precision must be 0 (fp32) or 1 (fp64)");
  if (precision == 0)
    AFX_LAUNCH(kKProbeScatter, (k_probe_scatter<float>), grid_for(n_cells), dim3(kBlock), 0, s, (int32_t)n_cells,
               block, nslot, cols, depth, max_depth, colour, k, field, response, (float*)vals);
  else
    AFX_LAUNCH(kKProbeScatter, (k_probe_scatter<double>), grid_for(n_cells), dim3(kBlock), 0, s, (int32_t)n_cells,
               block, nslot, cols, depth, max_depth, colour, k, field, response, (double*)vals);
  AFX_CHECK_LAUNCH();
  return 0;
}
