// Host-side (CPU) setup of the MGPBD system: CSR pattern of A, face adjacency, tet
// aggregation, Galerkin coarse patterns and refill maps, vertex incidence.
//
// These run once per solver init. They are O(nnz) rewrites of the reference's Python
// set/dict code and reproduce its results exactly, including the tie-breaks that CPython's
// set iteration order decides (see PySet8 below). All entry points are extern "C" and take
// plain host pointers; see include/mgpbd.h.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "mgpbd_internal.h"

namespace {

// A CPython set of non-negative small ints that never holds more than 4 keys: such a set
// keeps its initial 8-slot table (it resizes only when fill*5 >= mask*3, i.e. at the 5th key)
// and, since mask + 9 > 7, every probe is the perturbed open-addressing step
// i = (5*i + 1 + (perturb >>= 5)) & 7 starting at i = hash & 7 with hash(k) = k
// (CPython Objects/setobject.c, set_add_entry / set_next). Iteration is slot order.
struct PySet8 {
  int64_t slot[8];
  int n = 0;
  PySet8() { for (int i = 0; i < 8; ++i) slot[i] = -1; }
  void add(int64_t key) {
    uint64_t perturb = (uint64_t)key;
    uint64_t i = (uint64_t)key & 7u;
    while (true) {
      if (slot[i] < 0) { slot[i] = key; ++n; return; }
      if (slot[i] == key) return;
      perturb >>= 5;
      i = (i * 5u + 1u + perturb) & 7u;
    }
  }
  // iteration order
  int items(int32_t* out) const {
    int m = 0;
    for (int i = 0; i < 8; ++i) if (slot[i] >= 0) out[m++] = (int32_t)slot[i];
    return m;
  }
};

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------------------
// CSR pattern of A (reference: cons/mgpbd_gpu.py:153-194). Row t lists every other tet that
// shares a vertex with t in ascending order, then t itself (diagonal last).
// ---------------------------------------------------------------------------------------
struct mgpbd_pattern {
  std::vector<int32_t> indptr, indices;
};

static void vertex_tet_lists(int n_tets, int n_verts, const int32_t* t_i,
                             std::vector<int32_t>& vptr, std::vector<int32_t>& vtet) {
  vptr.assign((size_t)n_verts + 1, 0);
  for (int64_t k = 0; k < 4LL * n_tets; ++k) vptr[(size_t)t_i[k] + 1]++;
  for (int v = 0; v < n_verts; ++v) vptr[v + 1] += vptr[v];
  vtet.resize(4LL * n_tets);
  // A tet naming one vertex twice appears twice in that vertex's list; the row builder
  // sorts and de-duplicates, which gives the reference's set semantics.
  std::vector<int32_t> fill(vptr.begin(), vptr.end() - 1);
  for (int t = 0; t < n_tets; ++t)
    for (int l = 0; l < 4; ++l) vtet[fill[t_i[4LL * t + l]]++] = t;
}

mgpbd_pattern* mgpbd_pattern_build(int n_tets, int n_verts, const int32_t* t_i) {
  if (n_tets < 0 || n_verts < 0 || (n_tets > 0 && !t_i)) {
    mgpbd_set_error("mgpbd_pattern_build: invalid arguments");
    return nullptr;
  }
  for (int64_t k = 0; k < 4LL * n_tets; ++k)
    if (t_i[k] < 0 || t_i[k] >= n_verts) {
      mgpbd_set_error("mgpbd_pattern_build: vertex id out of range");
      return nullptr;
    }
  auto* P = new mgpbd_pattern;
  std::vector<int32_t> vptr, vtet;
  vertex_tet_lists(n_tets, n_verts, t_i, vptr, vtet);
  P->indptr.assign((size_t)n_tets + 1, 0);
  // pass 1: row lengths
  std::vector<int32_t> rowlen(n_tets);
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int t = 0; t < n_tets; ++t) {
      buf.clear();
      for (int l = 0; l < 4; ++l) {
        int v = t_i[4LL * t + l];
        buf.insert(buf.end(), vtet.begin() + vptr[v], vtet.begin() + vptr[v + 1]);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      rowlen[t] = (int32_t)buf.size();  // includes t itself
    }
  }
  int64_t nnz = 0;
  for (int t = 0; t < n_tets; ++t) {
    nnz += rowlen[t];
    if (nnz > INT32_MAX) {
      delete P;
      mgpbd_set_error("mgpbd_pattern_build: nnz exceeds int32");
      return nullptr;
    }
    P->indptr[t + 1] = (int32_t)nnz;
  }
  P->indices.resize((size_t)nnz);
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int t = 0; t < n_tets; ++t) {
      buf.clear();
      for (int l = 0; l < 4; ++l) {
        int v = t_i[4LL * t + l];
        buf.insert(buf.end(), vtet.begin() + vptr[v], vtet.begin() + vptr[v + 1]);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      int32_t* out = P->indices.data() + P->indptr[t];
      int m = 0;
      for (int32_t j : buf)
        if (j != t) out[m++] = j;
      out[m] = t;
    }
  }
  return P;
}

int64_t mgpbd_pattern_nnz(const mgpbd_pattern* P) { return P ? (int64_t)P->indices.size() : -1; }

int mgpbd_pattern_copy(const mgpbd_pattern* P, int32_t* indptr, int32_t* indices) {
  if (!P) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_pattern_copy: null pattern");
  std::memcpy(indptr, P->indptr.data(), P->indptr.size() * sizeof(int32_t));
  std::memcpy(indices, P->indices.data(), P->indices.size() * sizeof(int32_t));
  return 0;
}

void mgpbd_pattern_free(mgpbd_pattern* P) { delete P; }

// ---------------------------------------------------------------------------------------
// Face adjacency (reference: cons/tet_aggregation.py:11-37). Faces are keyed by their sorted
// vertex triple; the reference walks faces in first-seen order (dict insertion order: tet
// ascending, dropped vertex 0..3) and, for every face seen exactly twice, adds each tet to
// the other's set. adj[4t..4t+3] receives t's neighbours in that set's iteration order
// (-1 padded); deg[t] their count.
// ---------------------------------------------------------------------------------------
int mgpbd_face_adjacency(int n_tets, const int32_t* t_i, int32_t* adj, int32_t* deg) {
  if (n_tets < 0 || (n_tets > 0 && (!t_i || !adj || !deg)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_face_adjacency: invalid arguments");
  struct Face { int32_t a, b, c; int32_t tet; int32_t seen; };  // seen = 4*t + skip
  std::vector<Face> faces((size_t)4 * n_tets);
  for (int t = 0; t < n_tets; ++t) {
    int32_t v[4];
    for (int l = 0; l < 4; ++l) v[l] = t_i[4LL * t + l];
    for (int skip = 0; skip < 4; ++skip) {
      int32_t f[3]; int m = 0;
      for (int j = 0; j < 4; ++j) if (j != skip) f[m++] = v[j];
      std::sort(f, f + 3);
      faces[4LL * t + skip] = Face{f[0], f[1], f[2], t, 4 * t + skip};
    }
  }
  std::stable_sort(faces.begin(), faces.end(), [](const Face& x, const Face& y) {
    if (x.a != y.a) return x.a < y.a;
    if (x.b != y.b) return x.b < y.b;
    return x.c < y.c;
  });
  // groups of exactly two entries, ordered by the first-seen position of the face
  struct Pair { int32_t seen; int32_t t0, t1; };
  std::vector<Pair> pairs;
  pairs.reserve((size_t)2 * n_tets);
  size_t i = 0, nf = faces.size();
  while (i < nf) {
    size_t j = i + 1;
    while (j < nf && faces[j].a == faces[i].a && faces[j].b == faces[i].b &&
           faces[j].c == faces[i].c) ++j;
    if (j - i == 2) {
      // entries of one group keep first-seen order (stable sort), so [i] was seen first
      pairs.push_back(Pair{faces[i].seen, faces[i].tet, faces[i + 1].tet});
    }
    i = j;
  }
  std::sort(pairs.begin(), pairs.end(), [](const Pair& x, const Pair& y) { return x.seen < y.seen; });
  std::vector<PySet8> sets((size_t)n_tets);
  for (const Pair& p : pairs) {
    sets[p.t0].add(p.t1);
    sets[p.t1].add(p.t0);
  }
  for (int t = 0; t < n_tets; ++t) {
    int32_t tmp[8];
    int m = sets[t].items(tmp);
    if (m > 4) return mgpbd_fail(MGPBD_EINTERNAL, "mgpbd_face_adjacency: >4 face neighbours");
    for (int k = 0; k < 4; ++k) adj[4LL * t + k] = k < m ? tmp[k] : -1;
    deg[t] = m;
  }
  return 0;
}

// ---------------------------------------------------------------------------------------
// Greedy face-adjacent aggregation (reference: cons/tet_aggregation.py:40-127, pass 1; its
// pass 2 never runs because every unvisited tet becomes a seed).
// ---------------------------------------------------------------------------------------
int mgpbd_tet_aggregate(int n_tets, const int32_t* adj, const int32_t* deg, int max_agg,
                        int32_t* agg_ids, int32_t* n_agg_out) {
  if (n_tets < 0 || max_agg < 1 || (n_tets > 0 && (!adj || !deg || !agg_ids)) || !n_agg_out)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_tet_aggregate: invalid arguments");
  // order = sorted(range(n), key=len(adj[t])): stable counting sort by degree 0..4
  std::vector<int32_t> order;
  order.reserve(n_tets);
  for (int d = 0; d <= 4; ++d)
    for (int t = 0; t < n_tets; ++t)
      if (deg[t] == d) order.push_back(t);
  if ((int)order.size() != n_tets)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_tet_aggregate: degree outside 0..4");
  std::vector<uint8_t> visited((size_t)n_tets, 0);
  for (int t = 0; t < n_tets; ++t) agg_ids[t] = -1;
  int32_t n_agg = 0;
  for (int32_t seed : order) {
    if (visited[seed]) continue;
    agg_ids[seed] = n_agg;
    visited[seed] = 1;
    int size = 1;
    // adj[seed] - set(visited ids): iterate adj[seed], insert survivors into a fresh set
    PySet8 cs;
    for (int k = 0; k < deg[seed]; ++k) {
      int32_t nb = adj[4LL * seed + k];
      if (!visited[nb]) cs.add(nb);
    }
    int32_t cand[8];
    int nc = cs.items(cand);
    std::stable_sort(cand, cand + nc, [&](int32_t x, int32_t y) { return deg[x] < deg[y]; });
    for (int c = 0; c < nc; ++c) {
      int32_t nb = cand[c];
      if (visited[nb]) continue;
      if (size >= max_agg) break;
      agg_ids[nb] = n_agg;
      visited[nb] = 1;
      ++size;
      if (size < max_agg) {
        for (int k = 0; k < deg[nb]; ++k) {
          int32_t nb2 = adj[4LL * nb + k];
          if (!visited[nb2] && size < max_agg) {
            agg_ids[nb2] = n_agg;
            visited[nb2] = 1;
            ++size;
          }
        }
      }
    }
    ++n_agg;
  }
  *n_agg_out = n_agg;
  return 0;
}

// ---------------------------------------------------------------------------------------
// Greedy aggregation of a general graph given as sorted CSR adjacency (no self loops), with
// the same growth rule as mgpbd_tet_aggregate: seeds by ascending degree, candidates the
// unvisited neighbours by ascending degree, plus their unvisited neighbours. Used for the
// deeper levels of the extended hierarchy, on the aggregate graph induced by face adjacency.
// ---------------------------------------------------------------------------------------
int mgpbd_graph_aggregate(int n, const int32_t* ptr, const int32_t* idx, int max_agg,
                          int32_t* agg_ids, int32_t* n_agg_out) {
  if (n < 0 || max_agg < 1 || (n > 0 && (!ptr || !agg_ids)) || !n_agg_out)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_graph_aggregate: invalid arguments");
  std::vector<int32_t> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  auto degf = [&](int32_t t) { return ptr[t + 1] - ptr[t]; };
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return degf(x) < degf(y); });
  std::vector<uint8_t> visited((size_t)n, 0);
  for (int i = 0; i < n; ++i) agg_ids[i] = -1;
  int32_t n_agg = 0;
  std::vector<int32_t> cand;
  for (int32_t seed : order) {
    if (visited[seed]) continue;
    agg_ids[seed] = n_agg;
    visited[seed] = 1;
    int size = 1;
    cand.clear();
    for (int32_t k = ptr[seed]; k < ptr[seed + 1]; ++k)
      if (!visited[idx[k]]) cand.push_back(idx[k]);
    std::stable_sort(cand.begin(), cand.end(), [&](int32_t x, int32_t y) { return degf(x) < degf(y); });
    for (int32_t nb : cand) {
      if (visited[nb]) continue;
      if (size >= max_agg) break;
      agg_ids[nb] = n_agg;
      visited[nb] = 1;
      ++size;
      if (size < max_agg)
        for (int32_t k = ptr[nb]; k < ptr[nb + 1]; ++k) {
          int32_t nb2 = idx[k];
          if (!visited[nb2] && size < max_agg) {
            agg_ids[nb2] = n_agg;
            visited[nb2] = 1;
            ++size;
          }
        }
    }
    ++n_agg;
  }
  *n_agg_out = n_agg;
  return 0;
}

// pyamg's standard aggregation (the `aggregation='pyamg'` branch, cons/mgpbd_gpu.py:124-134,
// calls pyamg.smoothed_aggregation_solver with its default 'standard' aggregate on the
// 'symmetric' strength graph; pyamg is not vendored: this restates its published three-pass
// algorithm, Vanek-Mandel-Brezina 1996, in pyamg's node order):
//  1. in node order, a node none of whose neighbours is aggregated yet forms an aggregate
//     with all its neighbours; a node without off-diagonal neighbours stays out (isolated);
//  2. in node order, every node still free joins the aggregate of its first neighbour (in
//     adjacency order) that was aggregated in pass 1;
//  3. in node order, a node still free starts a new aggregate with its still-free neighbours.
// agg_ids[i] = aggregate of i, -1 for isolated nodes; aggregates numbered in creation order.
int mgpbd_standard_aggregate(int n, const int32_t* ptr, const int32_t* idx, int32_t* agg_ids,
                             int32_t* n_agg_out) {
  if (n < 0 || !n_agg_out || (n > 0 && (!ptr || !idx || !agg_ids)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_standard_aggregate: invalid arguments");
  // x: 0 free, k > 0 member of pass-1 aggregate k-1, -k joined pass-1 aggregate k-1 in pass 2,
  // -n-1 isolated (kept distinct from every -k)
  std::vector<int64_t> x((size_t)n, 0);
  const int64_t iso = -(int64_t)n - 1;
  int64_t next = 1;
  for (int i = 0; i < n; ++i) {
    if (x[i]) continue;
    bool has_nbr = false, has_agg_nbr = false;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int32_t j = idx[k];
      if (j == i) continue;
      has_nbr = true;
      if (x[j]) { has_agg_nbr = true; break; }
    }
    if (!has_nbr) {
      x[i] = iso;
    } else if (!has_agg_nbr) {
      x[i] = next;
      for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) x[idx[k]] = next;
      ++next;
    }
  }
  for (int i = 0; i < n; ++i) {
    if (x[i]) continue;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int64_t xj = x[idx[k]];
      if (xj > 0) { x[i] = -xj; break; }
    }
  }
  // final ids of passes 1-2 (0-based, in creation order), -1 isolated, -2 still free
  for (int i = 0; i < n; ++i) {
    const int64_t xi = x[i];
    x[i] = xi > 0 ? xi - 1 : xi == iso ? -1 : xi < 0 ? -xi - 1 : -2;
  }
  --next;   // aggregates made so far; pass 3 numbers on from here
  for (int i = 0; i < n; ++i) {
    if (x[i] != -2) continue;
    x[i] = next;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k)
      if (x[idx[k]] == -2) x[idx[k]] = next;
    ++next;
  }
  for (int i = 0; i < n; ++i) agg_ids[i] = (int32_t)x[i];
  *n_agg_out = (int32_t)next;
  return 0;
}

// Singleton merge for the graph levels of the extended hierarchy (no reference counterpart):
// in node order, a node that is alone in its aggregate joins the smallest aggregate among its
// neighbours' (ties: the lowest id; sizes as updated so far); aggregates are then renumbered
// in ascending order of their old ids, empty ones dropped. Without it the greedy growth leaves
// most nodes of the coarse graphs of irregular meshes as singletons and coarsening stalls.
int mgpbd_graph_merge_singletons(int n, const int32_t* ptr, const int32_t* idx,
                                 int32_t* agg_ids, int32_t* n_agg_inout) {
  if (n < 0 || !n_agg_inout || *n_agg_inout < 0 || (n > 0 && (!ptr || !idx || !agg_ids)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_graph_merge_singletons: invalid arguments");
  const int32_t na = *n_agg_inout;
  std::vector<int32_t> size((size_t)na, 0);
  for (int i = 0; i < n; ++i) {
    if (agg_ids[i] < 0 || agg_ids[i] >= na)
      return mgpbd_fail(MGPBD_EINVAL, "mgpbd_graph_merge_singletons: aggregate id out of range");
    ++size[agg_ids[i]];
  }
  for (int i = 0; i < n; ++i) {
    const int32_t a = agg_ids[i];
    if (size[a] != 1) continue;
    int32_t best = -1;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int32_t c = agg_ids[idx[k]];
      if (c == a) continue;
      if (best < 0 || size[c] < size[best] || (size[c] == size[best] && c < best)) best = c;
    }
    if (best < 0) continue;
    size[a] = 0;
    ++size[best];
    agg_ids[i] = best;
  }
  std::vector<int32_t> renum((size_t)na, -1);
  int32_t m = 0;
  for (int32_t a = 0; a < na; ++a)
    if (size[a] > 0) renum[a] = m++;
  for (int i = 0; i < n; ++i) agg_ids[i] = renum[agg_ids[i]];
  *n_agg_inout = m;
  return 0;
}

// Aggregate graph induced by a fine graph: I ~ J (I != J) when some fine edge joins them.
struct mgpbd_graph {
  std::vector<int32_t> ptr, idx;
};

mgpbd_graph* mgpbd_graph_coarsen(int n_f, const int32_t* ptr, const int32_t* idx,
                                 const int32_t* agg, int n_c) {
  if (n_f < 0 || n_c < 0 || (n_f > 0 && (!ptr || !agg))) {
    mgpbd_set_error("mgpbd_graph_coarsen: invalid arguments");
    return nullptr;
  }
  std::vector<std::vector<int32_t>> rows((size_t)n_c);
  for (int i = 0; i < n_f; ++i) {
    int32_t I = agg[i];
    if (I < 0 || I >= n_c) {
      mgpbd_set_error("mgpbd_graph_coarsen: aggregate id out of range");
      return nullptr;
    }
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      int32_t J = agg[idx[k]];
      if (J != I) rows[I].push_back(J);
    }
  }
  auto* G = new mgpbd_graph;
  G->ptr.assign((size_t)n_c + 1, 0);
  for (int I = 0; I < n_c; ++I) {
    auto& r = rows[I];
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    G->ptr[I + 1] = G->ptr[I] + (int32_t)r.size();
  }
  G->idx.resize((size_t)G->ptr[n_c]);
  for (int I = 0; I < n_c; ++I) std::copy(rows[I].begin(), rows[I].end(), G->idx.begin() + G->ptr[I]);
  return G;
}

int64_t mgpbd_graph_nnz(const mgpbd_graph* G) { return G ? (int64_t)G->idx.size() : -1; }
int mgpbd_graph_copy(const mgpbd_graph* G, int32_t* ptr, int32_t* idx) {
  if (!G) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_graph_copy: null graph");
  std::memcpy(ptr, G->ptr.data(), G->ptr.size() * sizeof(int32_t));
  std::memcpy(idx, G->idx.data(), G->idx.size() * sizeof(int32_t));
  return 0;
}
void mgpbd_graph_free(mgpbd_graph* G) { delete G; }

// ---------------------------------------------------------------------------------------
// Galerkin coarse level for a piecewise-constant P with unit entries (P[i, agg[i]] = 1), as
// the reference builds it in cons/mgpcg_gpu.py:46-99:
//  * pattern: Ac = (P^T A) P computed by scipy, whose sparse products drop entries that sum
//    to exactly 0.0 -- both in P^T A and in (P^T A) P. With `values` given (the FP32 A at
//    solver init) this value-dependent pattern is reproduced (FP64 sums); with
//    values == NULL the structural pattern (every reachable (I, J)) is built instead;
//  * rows sorted by column (scipy CSR from a product);
//  * contribution map: for each coarse entry (I, J) in row-major order, the fine CSR
//    positions a with row(a) in I (ascending) and agg[col(a)] == J, in fine CSR order.
//    Every other fine entry contributes to no coarse entry.
// ---------------------------------------------------------------------------------------
struct mgpbd_galerkin {
  std::vector<int32_t> c_indptr, c_indices, c_diag, contrib_idx, contrib_off;
};

mgpbd_galerkin* mgpbd_galerkin_build(int n_f, const int32_t* indptr, const int32_t* indices,
                                     const float* values, const int32_t* agg, int n_c) {
  if (n_f < 0 || n_c < 0 || (n_f > 0 && (!indptr || !indices || !agg))) {
    mgpbd_set_error("mgpbd_galerkin_build: invalid arguments");
    return nullptr;
  }
  for (int i = 0; i < n_f; ++i)
    if (agg[i] < 0 || agg[i] >= n_c) {
      mgpbd_set_error("mgpbd_galerkin_build: aggregate id out of range");
      return nullptr;
    }
  // members of each aggregate, ascending (rows of P^T)
  std::vector<int32_t> mptr((size_t)n_c + 1, 0), mem((size_t)n_f);
  for (int i = 0; i < n_f; ++i) mptr[agg[i] + 1]++;
  for (int I = 0; I < n_c; ++I) mptr[I + 1] += mptr[I];
  {
    std::vector<int32_t> f(mptr.begin(), mptr.end() - 1);
    for (int i = 0; i < n_f; ++i) mem[f[agg[i]]++] = i;
  }
  auto* G = new mgpbd_galerkin;
  G->c_indptr.assign((size_t)n_c + 1, 0);
  std::vector<std::vector<int32_t>> rowcols((size_t)n_c);
  std::vector<std::vector<int32_t>> rowcnt((size_t)n_c);

#pragma omp parallel
  {
    std::vector<double> tsum((size_t)n_f, 0.0);   // T[I, l] accumulator (value mode)
    std::vector<uint8_t> tmark((size_t)n_f, 0);
    std::vector<int32_t> tlist;
    std::vector<double> csum((size_t)n_c, 0.0);
    std::vector<int32_t> cpos((size_t)n_c, -1);
    std::vector<int32_t> clist;
#pragma omp for schedule(dynamic, 256)
    for (int I = 0; I < n_c; ++I) {
      clist.clear();
      if (values) {
        tlist.clear();
        for (int32_t m = mptr[I]; m < mptr[I + 1]; ++m) {
          int32_t k = mem[m];
          for (int32_t a = indptr[k]; a < indptr[k + 1]; ++a) {
            int32_t l = indices[a];
            if (!tmark[l]) { tmark[l] = 1; tsum[l] = 0.0; tlist.push_back(l); }
            tsum[l] += (double)values[a];
          }
        }
        for (int32_t l : tlist) {
          tmark[l] = 0;
          if (tsum[l] == 0.0) continue;   // dropped from P^T A
          int32_t J = agg[l];
          if (cpos[J] < 0) { cpos[J] = 1; csum[J] = 0.0; clist.push_back(J); }
          csum[J] += tsum[l];
        }
        std::vector<int32_t> kept;
        for (int32_t J : clist) {
          cpos[J] = -1;
          if (csum[J] != 0.0) kept.push_back(J);   // dropped from (P^T A) P
        }
        clist.swap(kept);
      } else {
        for (int32_t m = mptr[I]; m < mptr[I + 1]; ++m) {
          int32_t k = mem[m];
          for (int32_t a = indptr[k]; a < indptr[k + 1]; ++a) {
            int32_t J = agg[indices[a]];
            if (cpos[J] < 0) { cpos[J] = 1; clist.push_back(J); }
          }
        }
        for (int32_t J : clist) cpos[J] = -1;
      }
      std::sort(clist.begin(), clist.end());
      rowcols[I] = clist;
      // contribution counts per coarse entry of this row
      for (size_t p = 0; p < clist.size(); ++p) cpos[clist[p]] = (int32_t)p;
      std::vector<int32_t> cnt(clist.size(), 0);
      for (int32_t m = mptr[I]; m < mptr[I + 1]; ++m) {
        int32_t k = mem[m];
        for (int32_t a = indptr[k]; a < indptr[k + 1]; ++a) {
          int32_t p = cpos[agg[indices[a]]];
          if (p >= 0) cnt[p]++;
        }
      }
      for (int32_t J : clist) cpos[J] = -1;
      rowcnt[I].swap(cnt);
    }
  }
  int64_t cnnz = 0, ncontrib = 0;
  for (int I = 0; I < n_c; ++I) {
    cnnz += (int64_t)rowcols[I].size();
    for (int32_t c : rowcnt[I]) ncontrib += c;
    G->c_indptr[I + 1] = (int32_t)cnnz;
  }
  if (cnnz > INT32_MAX || ncontrib > INT32_MAX) {
    delete G;
    mgpbd_set_error("mgpbd_galerkin_build: size exceeds int32");
    return nullptr;
  }
  G->c_indices.resize((size_t)cnnz);
  G->c_diag.assign((size_t)n_c, -1);
  G->contrib_off.assign((size_t)cnnz + 1, 0);
  {
    int64_t c = 0;
    for (int I = 0; I < n_c; ++I)
      for (size_t p = 0; p < rowcols[I].size(); ++p, ++c) {
        G->c_indices[c] = rowcols[I][p];
        if (rowcols[I][p] == I) G->c_diag[I] = (int32_t)c;
        G->contrib_off[c + 1] = G->contrib_off[c] + rowcnt[I][p];
      }
  }
  G->contrib_idx.resize((size_t)ncontrib);
#pragma omp parallel
  {
    std::vector<int32_t> cpos((size_t)n_c, -1);
    std::vector<int32_t> fillp;
#pragma omp for schedule(dynamic, 256)
    for (int I = 0; I < n_c; ++I) {
      int32_t base = G->c_indptr[I];
      int32_t len = G->c_indptr[I + 1] - base;
      fillp.resize(len);
      for (int32_t p = 0; p < len; ++p) {
        cpos[G->c_indices[base + p]] = p;
        fillp[p] = G->contrib_off[base + p];
      }
      for (int32_t m = mptr[I]; m < mptr[I + 1]; ++m) {
        int32_t k = mem[m];
        for (int32_t a = indptr[k]; a < indptr[k + 1]; ++a) {
          int32_t p = cpos[agg[indices[a]]];
          if (p >= 0) G->contrib_idx[fillp[p]++] = a;
        }
      }
      for (int32_t p = 0; p < len; ++p) cpos[G->c_indices[base + p]] = -1;
    }
  }
  return G;
}

int mgpbd_galerkin_sizes(const mgpbd_galerkin* G, int64_t* c_nnz, int64_t* n_contrib) {
  if (!G) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_galerkin_sizes: null handle");
  *c_nnz = (int64_t)G->c_indices.size();
  *n_contrib = (int64_t)G->contrib_idx.size();
  return 0;
}

int mgpbd_galerkin_copy(const mgpbd_galerkin* G, int32_t* c_indptr, int32_t* c_indices,
                        int32_t* c_diag, int32_t* contrib_idx, int32_t* contrib_off) {
  if (!G) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_galerkin_copy: null handle");
  std::memcpy(c_indptr, G->c_indptr.data(), G->c_indptr.size() * 4);
  std::memcpy(c_indices, G->c_indices.data(), G->c_indices.size() * 4);
  std::memcpy(c_diag, G->c_diag.data(), G->c_diag.size() * 4);
  std::memcpy(contrib_idx, G->contrib_idx.data(), G->contrib_idx.size() * 4);
  std::memcpy(contrib_off, G->contrib_off.data(), G->contrib_off.size() * 4);
  return 0;
}

void mgpbd_galerkin_free(mgpbd_galerkin* G) { delete G; }

// ---------------------------------------------------------------------------------------
// Vertex incidence for the deterministic scatter of dlambda into vertex corrections: for
// every vertex, the (tet, slot) pairs that reference it, encoded 4*tet + slot, ascending.
// ---------------------------------------------------------------------------------------
int mgpbd_vertex_incidence(int n_tets, int n_verts, const int32_t* t_i, int32_t* vptr,
                           int32_t* vinc) {
  if (n_tets < 0 || n_verts < 0 || (n_tets > 0 && (!t_i || !vptr || !vinc)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_vertex_incidence: invalid arguments");
  for (int v = 0; v <= n_verts; ++v) vptr[v] = 0;
  for (int64_t k = 0; k < 4LL * n_tets; ++k) {
    if (t_i[k] < 0 || t_i[k] >= n_verts)
      return mgpbd_fail(MGPBD_EINVAL, "mgpbd_vertex_incidence: vertex id out of range");
    vptr[t_i[k] + 1]++;
  }
  for (int v = 0; v < n_verts; ++v) vptr[v + 1] += vptr[v];
  std::vector<int32_t> f(vptr, vptr + n_verts);
  for (int64_t k = 0; k < 4LL * n_tets; ++k) vinc[f[t_i[k]]++] = (int32_t)k;
  return 0;
}

// ---------------------------------------------------------------------------------------
// SELL-64x4 layout: rows binned 64 per slice (one wavefront); a slice's width is its longest
// row rounded up to a multiple of 4; entry k of row 64s+l sits at
//   slice_ptr[s] + 256*(k/4) + 4*l + (k%4)
// so each lane reads its row four entries (16 bytes) at a time and a wavefront's load is one
// contiguous 1 KiB segment.
// ---------------------------------------------------------------------------------------
static inline int64_t sell_pos_of(const int32_t* slice_ptr, int r, int k) {
  return (int64_t)slice_ptr[r / 64] + 256LL * (k / 4) + 4 * (r % 64) + (k % 4);
}

int mgpbd_sell_layout(int n_rows, const int32_t* indptr, int32_t* slice_ptr) {
  if (n_rows < 0 || (n_rows > 0 && !indptr) || !slice_ptr)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_sell_layout: invalid arguments");
  int n_slices = (n_rows + 63) / 64;
  int64_t acc = 0;
  slice_ptr[0] = 0;
  for (int s = 0; s < n_slices; ++s) {
    int w = 0;
    for (int r = 64 * s; r < std::min(n_rows, 64 * s + 64); ++r) w = std::max(w, indptr[r + 1] - indptr[r]);
    w = (w + 3) / 4 * 4;
    acc += 64LL * w;
    if (acc > INT32_MAX) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_sell_layout: padded size exceeds int32");
    slice_ptr[s + 1] = (int32_t)acc;
  }
  return 0;
}

int mgpbd_sell_map(int n_rows, const int32_t* indptr, const int32_t* slice_ptr, int32_t* sell_pos) {
  if (n_rows < 0 || (n_rows > 0 && (!indptr || !slice_ptr || !sell_pos)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_sell_map: invalid arguments");
  for (int r = 0; r < n_rows; ++r)
    for (int32_t a = indptr[r]; a < indptr[r + 1]; ++a)
      sell_pos[a] = (int32_t)sell_pos_of(slice_ptr, r, a - indptr[r]);
  return 0;
}

// Column array of the layout: real entries at their positions, padding entries of a row
// point at the row itself, rows past n_rows at column 0.
int mgpbd_sell_columns(int n_rows, const int32_t* indptr, const int32_t* indices,
                       const int32_t* slice_ptr, int32_t* sell_idx) {
  if (n_rows < 0 || (n_rows > 0 && (!indptr || !indices || !slice_ptr || !sell_idx)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_sell_columns: invalid arguments");
  int n_slices = (n_rows + 63) / 64;
#pragma omp parallel for schedule(static)
  for (int s = 0; s < n_slices; ++s) {
    int w = (slice_ptr[s + 1] - slice_ptr[s]) / 64;
    for (int l = 0; l < 64; ++l) {
      int r = 64 * s + l;
      for (int k = 0; k < w; ++k) {
        int32_t c;
        if (r >= n_rows) c = 0;
        else if (k < indptr[r + 1] - indptr[r]) c = indices[indptr[r] + k];
        else c = r;
        sell_idx[sell_pos_of(slice_ptr, 64 * s + l, k)] = c;
      }
    }
  }
  return 0;
}

// LDS windows of the windowed packed sweeps: for each block of MGPBD_WINDOW_ROWS rows, the
// sorted ids of the MGPBD_WINDOW_SEG-column segments holding its rows' columns and its rows'
// own indices (the column of a real row's SELL padding). One pass over the entries with a
// per-thread stamp array, then a sort of each block's (short) list.
int mgpbd_sell_window(int n_rows, const int32_t* indptr, const int32_t* indices,
                      int32_t* seg_ptr, int32_t* segs, int64_t cap) {
  if (n_rows <= 0 || !indptr || !indices || !seg_ptr)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_sell_window: invalid arguments");
  const int rows = MGPBD_WINDOW_ROWS, seg = MGPBD_WINDOW_SEG;
  const int64_t n_slices = (n_rows + 63) / 64;
  const int nb = (int)((64 * n_slices + rows - 1) / rows);
  int n_cols = n_rows;
  for (int64_t e = 0; e < indptr[n_rows]; ++e) {
    if (indices[e] < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_sell_window: negative column");
    n_cols = std::max(n_cols, indices[e] + 1);
  }
  const int n_seg = (n_cols + seg - 1) / seg;
  std::vector<std::vector<int32_t>> lists(nb);
#pragma omp parallel
  {
    std::vector<int32_t> stamp(n_seg, -1);
#pragma omp for schedule(static)
    for (int b = 0; b < nb; ++b) {
      std::vector<int32_t>& out = lists[b];
      const int r1 = std::min(n_rows, (b + 1) * rows);
      for (int r = b * rows; r < r1; ++r) {
        const int own = r / seg;
        if (stamp[own] != b) { stamp[own] = b; out.push_back(own); }
        for (int32_t e = indptr[r]; e < indptr[r + 1]; ++e) {
          const int sg = indices[e] / seg;
          if (stamp[sg] != b) { stamp[sg] = b; out.push_back(sg); }
        }
      }
      if (out.empty()) out.push_back(0);   // blocks of padding rows only: one dummy segment
      std::sort(out.begin(), out.end());
    }
  }
  int64_t total = 0;
  for (int b = 0; b < nb; ++b) {
    seg_ptr[b] = (int32_t)total;
    total += (int64_t)lists[b].size();
  }
  if (total > INT32_MAX) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_sell_window: too many segments");
  seg_ptr[nb] = (int32_t)total;
  if (!segs) return 0;
  if (cap < total) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_sell_window: segs capacity below seg_ptr[n_blocks]");
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b)
    std::copy(lists[b].begin(), lists[b].end(), segs + seg_ptr[b]);
  return 0;
}

}  // extern "C"
