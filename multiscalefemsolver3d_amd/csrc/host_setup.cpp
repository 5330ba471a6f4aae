// Host setup of the FEM solve path: synthetic structured meshes and fields,
// the mechanical node renumbering, the node-adjacency portrait, gather lists
// for deterministic device assembly, SELL-C layouts, and a Gmsh writer.
// Everything here is CPU-only and runs once per mesh.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <numeric>
#include <string>
#include <vector>
#include <omp.h>

#include "../../include/mfs.h"
#include "mfs_host.hpp"

namespace mfs {

// ---------------------------------------------------------------------------
// Counter-based generator: splitmix64 finaliser, keyed by (seed, stream, idx).
// This is the definition of the synthetic inputs (the reference has none).
// ---------------------------------------------------------------------------
static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline double rand_uniform(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t key = mix64(seed * 0x100000001B3ull + stream);
  uint64_t h = mix64(key ^ (idx * 0xD6E8FEB86659FD93ull));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
static inline double rand_normal(uint64_t seed, uint64_t stream, uint64_t idx) {
  double u1 = rand_uniform(seed, stream, 2 * idx);
  double u2 = rand_uniform(seed, stream, 2 * idx + 1);
  if (u1 < 1e-300) u1 = 1e-300;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// ---------------------------------------------------------------------------
// Portrait: row r lists every node sharing a tet with r (r included), sorted.
// Built from the node->tet adjacency, rows in parallel.
// ---------------------------------------------------------------------------
void node_tets(int64_t nn, int64_t nt, const int32_t* tets, std::vector<int64_t>& ptr, std::vector<int32_t>& list) {
  ptr.assign(nn + 1, 0);
  for (int64_t k = 0; k < 4 * nt; k++) ptr[tets[k] + 1]++;
  for (int64_t i = 0; i < nn; i++) ptr[i + 1] += ptr[i];
  list.resize(4 * nt);
  std::vector<int64_t> cur(ptr.begin(), ptr.end() - 1);
  for (int64_t e = 0; e < nt; e++)  // ascending tet order within each node
    for (int v = 0; v < 4; v++) list[cur[tets[4 * e + v]]++] = (int32_t)e;
}

// Distinct tets touching the nodes [q*C, (q+1)*C) of each slice q, in
// ascending order.
void build_slice_tets(int64_t nn, int64_t nt, const int32_t* tets, int C, std::vector<int64_t>& sptr,
                      std::vector<int32_t>& stet) {
  std::vector<int64_t> nptr;
  std::vector<int32_t> nlist;
  node_tets(nn, nt, tets, nptr, nlist);
  const int64_t ns = (nn + C - 1) / C;
  std::vector<std::vector<int32_t>> per(ns);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < ns; q++) {
    std::vector<int32_t>& v = per[q];
    v.assign(nlist.begin() + nptr[q * C], nlist.begin() + nptr[std::min<int64_t>(nn, (q + 1) * C)]);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  sptr.assign(ns + 1, 0);
  for (int64_t q = 0; q < ns; q++) sptr[q + 1] = sptr[q] + (int64_t)per[q].size();
  stet.resize(sptr[ns]);
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < ns; q++) std::copy(per[q].begin(), per[q].end(), stet.begin() + sptr[q]);
}

void build_pattern(int64_t n, int64_t nt, const int32_t* tets, Pattern& P) {
  std::vector<int64_t> nptr;
  std::vector<int32_t> nlist;
  node_tets(n, nt, tets, nptr, nlist);
  P.n = n;
  P.rowptr.assign(n + 1, 0);
  std::vector<int32_t> len(n);
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; r++) {
      buf.clear();
      for (int64_t k = nptr[r]; k < nptr[r + 1]; k++) {
        const int32_t* t = tets + 4 * (int64_t)nlist[k];
        for (int v = 0; v < 4; v++) buf.push_back(t[v]);
      }
      if (nptr[r] == nptr[r + 1]) buf.push_back((int32_t)r);  // isolated node keeps its diagonal
      std::sort(buf.begin(), buf.end());
      len[r] = (int32_t)(std::unique(buf.begin(), buf.end()) - buf.begin());
    }
  }
  for (int64_t r = 0; r < n; r++) P.rowptr[r + 1] = P.rowptr[r] + len[r];
  P.col.resize(P.rowptr[n]);
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; r++) {
      buf.clear();
      for (int64_t k = nptr[r]; k < nptr[r + 1]; k++) {
        const int32_t* t = tets + 4 * (int64_t)nlist[k];
        for (int v = 0; v < 4; v++) buf.push_back(t[v]);
      }
      if (nptr[r] == nptr[r + 1]) buf.push_back((int32_t)r);
      std::sort(buf.begin(), buf.end());
      auto end = std::unique(buf.begin(), buf.end());
      std::copy(buf.begin(), end, P.col.begin() + P.rowptr[r]);
    }
  }
}

static inline int64_t find_slot(const Pattern& P, int32_t r, int32_t c) {
  const int32_t* b = P.col.data() + P.rowptr[r];
  const int32_t* e = P.col.data() + P.rowptr[r + 1];
  const int32_t* p = std::lower_bound(b, e, c);
  return (p != e && *p == c) ? (int64_t)(p - P.col.data()) : -1;
}

void build_gather_lists(const Pattern& P, int64_t nt, const int32_t* tets, GatherLists& G) {
  const int64_t n = P.n;
  std::vector<int64_t> nptr;
  std::vector<int32_t> nlist;
  node_tets(n, nt, tets, nptr, nlist);
  const int64_t nnz = P.nnz();
  std::vector<int32_t> cnt(nnz, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < n; r++)
    for (int64_t k = nptr[r]; k < nptr[r + 1]; k++) {
      const int32_t* t = tets + 4 * (int64_t)nlist[k];
      for (int i = 0; i < 4; i++) {
        if (t[i] != r) continue;
        for (int j = 0; j < 4; j++) cnt[find_slot(P, (int32_t)r, t[j])]++;
      }
    }
  G.ptr.assign(nnz + 1, 0);
  for (int64_t s = 0; s < nnz; s++) G.ptr[s + 1] = G.ptr[s] + cnt[s];
  G.entry.resize(G.ptr[nnz]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < n; r++) {
    // rows own their slots, so the per-slot cursors are private to this row
    for (int64_t s = P.rowptr[r]; s < P.rowptr[r + 1]; s++) cnt[s] = 0;
    for (int64_t k = nptr[r]; k < nptr[r + 1]; k++) {
      int64_t e = nlist[k];
      const int32_t* t = tets + 4 * e;
      for (int i = 0; i < 4; i++) {
        if (t[i] != r) continue;
        for (int j = 0; j < 4; j++) {
          int64_t s = find_slot(P, (int32_t)r, t[j]);
          G.entry[G.ptr[s] + cnt[s]++] = (uint32_t)((e << 4) | (i << 2) | j);
        }
      }
    }
  }
}

int64_t renumber_mech(int64_t nn, int64_t nt, const int32_t* tets, const int32_t* mat, int64_t nfix,
                      const int32_t* fixed, int32_t* map) {
  std::fill(map, map + nn, -1);
  int32_t next = 0;
  for (int64_t i = 0; i < nfix; i++)
    if (map[fixed[i]] < 0) map[fixed[i]] = next++;
  int32_t mmax = 0;
  if (mat)
    for (int64_t e = 0; e < nt; e++) mmax = std::max(mmax, mat[e]);
  for (int32_t m = 0; m <= mmax; m++)
    for (int64_t e = 0; e < nt; e++) {
      if ((mat ? mat[e] : 0) != m) continue;
      for (int v = 0; v < 4; v++)
        if (map[tets[4 * e + v]] < 0) map[tets[4 * e + v]] = next++;
    }
  for (int64_t i = 0; i < nn; i++)  // nodes in no tet (not met in the reference's meshes)
    if (map[i] < 0) map[i] = next++;
  return next;
}

void build_sell(const Pattern& P, int C, SellLayout& S) {
  S.C = C;
  S.n_slices = (P.n + C - 1) / C;
  S.slice_base.assign(S.n_slices + 1, 0);
  S.slice_width.assign(S.n_slices, 0);
  for (int64_t q = 0; q < S.n_slices; q++) {
    int32_t w = 0;
    for (int64_t r = q * C; r < std::min<int64_t>(P.n, (q + 1) * C); r++) w = std::max(w, P.rowptr[r + 1] - P.rowptr[r]);
    S.slice_width[q] = w;
    S.slice_base[q + 1] = S.slice_base[q] + (int64_t)w * C;
  }
  S.col.assign(S.slice_base[S.n_slices], 0);
  S.pos.assign(P.nnz(), 0);
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < S.n_slices; q++) {
    for (int l = 0; l < C; l++) {
      int64_t r = q * C + l;
      int32_t self = (int32_t)std::min<int64_t>(r, P.n - 1);
      int32_t len = r < P.n ? P.rowptr[r + 1] - P.rowptr[r] : 0;
      for (int32_t j = 0; j < S.slice_width[q]; j++) {
        int64_t at = S.slice_base[q] + (int64_t)j * C + l;
        if (j < len) {
          S.col[at] = P.col[P.rowptr[r] + j];
          S.pos[P.rowptr[r] + j] = at;
        } else {
          S.col[at] = self;
        }
      }
    }
  }
}

}  // namespace mfs

using namespace mfs;

// ---------------------------------------------------------------------------
// C ABI: host utilities
// ---------------------------------------------------------------------------
static const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

extern "C" int mfs_mesh_box_counts(int nx, int ny, int nz, int64_t* n_nodes, int64_t* n_tets) {
  if (nx < 1 || ny < 1 || nz < 1 || !n_nodes || !n_tets) {
    set_error("mfs_mesh_box_counts: bad argument");
    return MFS_ERR_ARG;
  }
  *n_nodes = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  *n_tets = 6ll * nx * ny * nz;
  return MFS_OK;
}

extern "C" int mfs_mesh_box(int nx, int ny, int nz, double x0, double y0, double z0, double hx, double hy,
                            double hz, double* xyz, int32_t* tets) {
  if (nx < 1 || ny < 1 || nz < 1 || !xyz || !tets) {
    set_error("mfs_mesh_box: bad argument");
    return MFS_ERR_ARG;
  }
  const int64_t sx = nx + 1, sy = ny + 1;
#pragma omp parallel for schedule(static)
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        int64_t id = i + sx * (j + sy * k);
        xyz[3 * id] = x0 + hx * i;
        xyz[3 * id + 1] = y0 + hy * j;
        xyz[3 * id + 2] = z0 + hz * k;
      }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int64_t hex = i + (int64_t)nx * (j + (int64_t)ny * k);
        for (int p = 0; p < 6; p++) {
          int c[3] = {i, j, k};
          int32_t* t = tets + 4 * (6 * hex + p);
          t[0] = (int32_t)(c[0] + sx * (c[1] + sy * c[2]));
          for (int s = 0; s < 3; s++) {
            c[kPerm[p][s]]++;
            t[1 + s] = (int32_t)(c[0] + sx * (c[1] + sy * c[2]));
          }
        }
      }
  return MFS_OK;
}

extern "C" int mfs_mesh_box_face_nodes(int nx, int ny, int nz, int face, int32_t* nodes, int64_t* count) {
  if (face < 0 || face > 5 || !count) {
    set_error("mfs_mesh_box_face_nodes: bad argument");
    return MFS_ERR_ARG;
  }
  const int n[3] = {nx, ny, nz};
  const int axis = face / 2, at = (face & 1) ? n[axis] : 0;
  int64_t c = 0;
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        int idx[3] = {i, j, k};
        if (idx[axis] != at) continue;
        if (nodes) nodes[c] = (int32_t)(i + (int64_t)(nx + 1) * (j + (int64_t)(ny + 1) * k));
        c++;
      }
  *count = c;
  return MFS_OK;
}

extern "C" int mfs_mesh_box_face_tris(int nx, int ny, int nz, int face, int32_t* tris, int64_t* count) {
  if (face < 0 || face > 5 || !count) {
    set_error("mfs_mesh_box_face_tris: bad argument");
    return MFS_ERR_ARG;
  }
  const int n[3] = {nx, ny, nz};
  const int axis = face / 2, plane = (face & 1) ? n[axis] : 0;
  const int64_t sx = nx + 1, sy = ny + 1;
  int64_t c = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int h[3] = {i, j, k};
        if (h[axis] != plane - ((face & 1) ? 1 : 0)) continue;
        for (int p = 0; p < 6; p++) {
          int pts[4][3];
          int cc[3] = {i, j, k};
          std::memcpy(pts[0], cc, sizeof(cc));
          for (int s = 0; s < 3; s++) {
            cc[kPerm[p][s]]++;
            std::memcpy(pts[1 + s], cc, sizeof(cc));
          }
          for (int drop = 0; drop < 4; drop++) {
            bool on = true;
            for (int v = 0; v < 4; v++)
              if (v != drop && pts[v][axis] != plane) on = false;
            if (!on) continue;
            if (tris) {
              int q = 0;
              for (int v = 0; v < 4; v++)
                if (v != drop) tris[3 * c + q++] = (int32_t)(pts[v][0] + sx * (pts[v][1] + sy * pts[v][2]));
            }
            c++;
          }
        }
      }
  *count = c;
  return MFS_OK;
}

extern "C" int mfs_field_inclusions(int64_t nt, const int32_t* tets, const double* xyz, int count, double rmin,
                                    double rmax, uint64_t seed, double v_bg, double v_in, double* out) {
  if (nt < 0 || !tets || !xyz || !out || count < 0) {
    set_error("mfs_field_inclusions: bad argument");
    return MFS_ERR_ARG;
  }
  std::vector<double> c(3 * (size_t)count), r2(count);
  for (int s = 0; s < count; s++) {
    for (int a = 0; a < 3; a++) c[3 * s + a] = rand_uniform(seed, 1, 4 * (uint64_t)s + a);
    double r = rmin + (rmax - rmin) * rand_uniform(seed, 1, 4 * (uint64_t)s + 3);
    r2[s] = r * r;
  }
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < nt; t++) {
    double g[3] = {0, 0, 0};
    for (int v = 0; v < 4; v++)
      for (int a = 0; a < 3; a++) g[a] += xyz[3 * (int64_t)tets[4 * t + v] + a];
    for (int a = 0; a < 3; a++) g[a] *= 0.25;
    double val = v_bg;
    for (int s = 0; s < count; s++) {
      double dx = g[0] - c[3 * s], dy = g[1] - c[3 * s + 1], dz = g[2] - c[3 * s + 2];
      if (dx * dx + dy * dy + dz * dz < r2[s]) {
        val = v_in;
        break;
      }
    }
    out[t] = val;
  }
  return MFS_OK;
}

extern "C" int mfs_field_lognormal(int64_t nt, uint64_t seed, uint64_t stream, int64_t index0, double mu,
                                   double sigma, double* out) {
  if (nt < 0 || !out) {
    set_error("mfs_field_lognormal: bad argument");
    return MFS_ERR_ARG;
  }
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < nt; t++) out[t] = std::exp(mu + sigma * rand_normal(seed, stream, (uint64_t)(index0 + t)));
  return MFS_OK;
}

extern "C" int mfs_field_grf(int64_t nn, int64_t nt, const int32_t* tets, const double* xyz, int modes, double ell,
                             uint64_t seed, double v_lo, double v_hi, double* out, int32_t* material) {
  if (nn < 0 || nt < 0 || !tets || !xyz || !out || modes < 1 || !(ell > 0)) {
    set_error("mfs_field_grf: bad argument");
    return MFS_ERR_ARG;
  }
  std::vector<double> kap(3 * (size_t)modes), ph(modes), g(nn);
  for (int m = 0; m < modes; m++) {
    for (int a = 0; a < 3; a++) kap[3 * m + a] = rand_normal(seed, 7, 3 * (uint64_t)m + a) / ell;
    ph[m] = 6.283185307179586 * rand_uniform(seed, 8, (uint64_t)m);
  }
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < nn; v++) {
    double s = 0.0;
    for (int m = 0; m < modes; m++)
      s += std::cos(kap[3 * m] * xyz[3 * v] + kap[3 * m + 1] * xyz[3 * v + 1] + kap[3 * m + 2] * xyz[3 * v + 2] + ph[m]);
    g[v] = s;
  }
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < nt; t++) {
    double s = 0.0;
    for (int v = 0; v < 4; v++) s += g[tets[4 * t + v]];
    s *= 0.25;
    out[t] = s >= 0.0 ? v_hi : v_lo;
    if (material) material[t] = s >= 0.0 ? 1 : 0;
  }
  return MFS_OK;
}

extern "C" int mfs_renumber(int64_t nn, int64_t nt, const int32_t* tets, const int32_t* material, int64_t nfix,
                            const int32_t* fixed_nodes, int32_t* map) {
  if (nn <= 0 || nt < 0 || !tets || !map || (nfix > 0 && !fixed_nodes)) {
    set_error("mfs_renumber: bad argument");
    return MFS_ERR_ARG;
  }
  for (int64_t i = 0; i < nfix; i++)
    if (fixed_nodes[i] < 0 || fixed_nodes[i] >= nn) {
      set_error("mfs_renumber: fixed node out of range");
      return MFS_ERR_ARG;
    }
  renumber_mech(nn, nt, tets, material, nfix, fixed_nodes, map);
  return MFS_OK;
}

extern "C" int mfs_portrait(int64_t n_rows, int64_t nt, const int32_t* tets, int32_t* rowptr, int32_t* col) {
  if (n_rows <= 0 || nt < 0 || !tets || !rowptr) {
    set_error("mfs_portrait: bad argument");
    return MFS_ERR_ARG;
  }
  for (int64_t k = 0; k < 4 * nt; k++)
    if (tets[k] < 0 || tets[k] >= n_rows) {
      set_error("mfs_portrait: tet node out of range");
      return MFS_ERR_ARG;
    }
  Pattern P;
  build_pattern(n_rows, nt, tets, P);
  std::memcpy(rowptr, P.rowptr.data(), (n_rows + 1) * sizeof(int32_t));
  if (col) std::memcpy(col, P.col.data(), P.col.size() * sizeof(int32_t));
  return MFS_OK;
}

extern "C" int mfs_write_gmsh(const char* path, int64_t nn, const double* xyz, int64_t nt, const int32_t* tets,
                              const int32_t* material, int64_t ntri, const int32_t* tris, const int32_t* tri_tags) {
  if (!path || !xyz || !tets || (ntri > 0 && (!tris || !tri_tags))) {
    set_error("mfs_write_gmsh: bad argument");
    return MFS_ERR_ARG;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("mfs_write_gmsh: cannot open ") + path);
    return MFS_ERR_ARG;
  }
  // Parent tet of each boundary triangle: the "modified" layout the reader
  // accepts without its O(faces x tets) search (Library/Math.h:455-511).
  std::vector<int64_t> parent(ntri, -1);
  if (ntri > 0) {
    std::vector<int64_t> nptr;
    std::vector<int32_t> nlist;
    node_tets(nn, nt, tets, nptr, nlist);
    for (int64_t t = 0; t < ntri; t++) {
      int32_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
      for (int64_t k = nptr[a]; k < nptr[a + 1] && parent[t] < 0; k++) {
        const int32_t* e = tets + 4 * (int64_t)nlist[k];
        int hits = 0;
        for (int v = 0; v < 4; v++) hits += (e[v] == b) + (e[v] == c);
        if (hits == 2) parent[t] = nlist[k];
      }
    }
  }
  std::fprintf(f, "$MeshFormat\n2 0 8 1\n$EndMeshFormat\n$Nodes\n\t%lld\n", (long long)nn);
  for (int64_t i = 0; i < nn; i++)
    std::fprintf(f, "\t\t%lld\t%.16lf\t%.16lf\t%.16lf\n", (long long)(i + 1), xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  std::fprintf(f, "$EndNodes\n$Elements\n\t%lld\n", (long long)(ntri + nt));
  int64_t id = 1;
  for (int64_t t = 0; t < ntri; t++, id++)
    std::fprintf(f, "\t\t%lld\t2\t3\t%d\t%lld\t0\t%d\t%d\t%d\n", (long long)id, tri_tags[t], (long long)(parent[t] + 1),
                 tris[3 * t] + 1, tris[3 * t + 1] + 1, tris[3 * t + 2] + 1);
  for (int64_t e = 0; e < nt; e++, id++) {
    int m = material ? material[e] : 0;
    std::fprintf(f, "\t\t%lld\t4\t3\t%d\t%d\t0\t%d\t%d\t%d\t%d\n", (long long)id, m, m, tets[4 * e] + 1,
                 tets[4 * e + 1] + 1, tets[4 * e + 2] + 1, tets[4 * e + 3] + 1);
  }
  std::fprintf(f, "$EndElements\n");
  std::fclose(f);
  return MFS_OK;
}

// Gmsh 2.2 ASCII reader with SimpleGrid::ReadFromMSH's rules (Library/
// Math.h:327-520): element rows are read as integer lists; 9 integers of
// type 2 are a boundary triangle (physical tag in the 4th field, parent tet
// in the 5th when the header line holds 4 numbers, the "modified" layout),
// 10 integers of type 4 a tet (4th field = material), 8 integers of type 1
// an edge (skipped).  Node ids are 1-based in the file.
namespace {
struct GmshMesh {
  std::vector<double> xyz;
  std::vector<int32_t> tets, mat, tris, tags, parent;
};

bool read_line(FILE* f, std::string& line) {
  line.clear();
  int c;
  while ((c = std::fgetc(f)) != EOF) {
    if (c == '\n') return true;
    line.push_back((char)c);
  }
  return !line.empty();
}

void ints_of(const std::string& line, std::vector<long long>& out) {
  out.clear();
  const char* p = line.c_str();
  while (*p) {
    while (*p == ' ' || *p == '\t' || *p == '\r') p++;
    if (!*p) break;
    char* end = nullptr;
    const long long v = std::strtoll(p, &end, 10);
    if (end == p) break;
    out.push_back(v);
    p = end;
  }
}

int parse_gmsh(const char* path, GmshMesh& M) {
  // closed on every exit, exceptions included
  std::unique_ptr<FILE, int (*)(FILE*)> file(std::fopen(path, "r"), &std::fclose);
  FILE* f = file.get();
  if (!f) {
    set_error(std::string("mfs_read_gmsh: cannot open ") + path);
    return MFS_ERR_ARG;
  }
  std::string line;
  std::vector<long long> iv;
  bool modified = false, ok = true;
  int section = 0;  // 1 format, 2 nodes, 3 elements
  int64_t expect = -1;
  while (ok && read_line(f, line)) {
    if (line.rfind("$MeshFormat", 0) == 0) {
      section = 1;
      continue;
    }
    if (line.rfind("$Nodes", 0) == 0) {
      section = 2;
      expect = -1;
      continue;
    }
    if (line.rfind("$Elements", 0) == 0) {
      section = 3;
      expect = -1;
      continue;
    }
    if (line.rfind("$End", 0) == 0) {
      section = 0;
      continue;
    }
    if (section == 1) {
      ints_of(line, iv);
      // "2 0 8 1" parses as 4 numbers in the modified layout, "2.2 0 8" as fewer
      modified = iv.size() == 4;
    } else if (section == 2) {
      if (expect < 0) {
        ints_of(line, iv);
        // the count only sizes a reservation: a negative or implausible one
        // (more nodes than a 2^31-entry index space holds) marks the file
        // malformed instead of reaching the allocator
        if (iv.empty() || iv[0] < 0 || iv[0] > (long long)INT32_MAX) {
          ok = false;
          break;
        }
        expect = iv[0];
        M.xyz.reserve(3 * (size_t)std::min<int64_t>(expect, 1 << 24));
        continue;
      }
      long long id;
      double x, y, z;
      if (std::sscanf(line.c_str(), "%lld %lf %lf %lf", &id, &x, &y, &z) != 4 ||
          id != (long long)M.xyz.size() / 3 + 1) {
        ok = false;
        break;
      }
      M.xyz.push_back(x);
      M.xyz.push_back(y);
      M.xyz.push_back(z);
    } else if (section == 3) {
      ints_of(line, iv);
      if (expect < 0) {
        expect = iv.empty() ? 0 : iv[0];
        continue;
      }
      if (iv.size() == 9 && iv[1] == 2) {
        for (int k = 6; k < 9; k++) M.tris.push_back((int32_t)(iv[k] - 1));
        M.tags.push_back((int32_t)iv[3]);
        M.parent.push_back(modified ? (int32_t)(iv[4] - 1) : -1);
      } else if (iv.size() == 10 && iv[1] == 4) {
        for (int k = 6; k < 10; k++) M.tets.push_back((int32_t)(iv[k] - 1));
        M.mat.push_back((int32_t)iv[3]);
      }
    }
  }
  const int64_t nn = (int64_t)M.xyz.size() / 3;
  for (int32_t v : M.tets) ok = ok && v >= 0 && v < nn;
  for (int32_t v : M.tris) ok = ok && v >= 0 && v < nn;
  if (!ok) {
    set_error(std::string("mfs_read_gmsh: malformed file ") + path);
    return MFS_ERR_ARG;
  }
  // parent tets missing from the file: from the node-tet adjacency
  const int64_t nt = (int64_t)M.mat.size();
  bool need = false;
  for (int32_t pt : M.parent) need = need || pt < 0 || pt >= nt;
  if (need && nt > 0) {
    std::vector<int64_t> nptr;
    std::vector<int32_t> nlist;
    node_tets(nn, nt, M.tets.data(), nptr, nlist);
    for (size_t t = 0; t < M.parent.size(); t++) {
      if (M.parent[t] >= 0 && M.parent[t] < nt) continue;
      M.parent[t] = -1;
      const int32_t a = M.tris[3 * t], b = M.tris[3 * t + 1], c = M.tris[3 * t + 2];
      for (int64_t k = nptr[a]; k < nptr[a + 1] && M.parent[t] < 0; k++) {
        const int32_t* e = M.tets.data() + 4 * (int64_t)nlist[k];
        int hits = 0;
        for (int v = 0; v < 4; v++) hits += (e[v] == b) + (e[v] == c);
        if (hits == 2) M.parent[t] = nlist[k];
      }
    }
  }
  return MFS_OK;
}
}  // namespace

extern "C" int mfs_read_gmsh(const char* path, int64_t* n_nodes, double* xyz, int64_t* n_tets, int32_t* tets,
                             int32_t* material, int64_t* n_tris, int32_t* tris, int32_t* tri_tags,
                             int32_t* tri_parent) {
  if (!path || !n_nodes || !n_tets || !n_tris) {
    set_error("mfs_read_gmsh: bad argument");
    return MFS_ERR_ARG;
  }
  GmshMesh M;
  int st;
  try {
    st = parse_gmsh(path, M);
  } catch (const std::exception& e) {  // allocation failures of a huge file
    set_error(std::string("mfs_read_gmsh: ") + e.what());
    return MFS_ERR_OOM;
  }
  if (st != MFS_OK) return st;
  const int64_t nn = (int64_t)M.xyz.size() / 3, nt = (int64_t)M.mat.size(), nf = (int64_t)M.tags.size();
  const bool fill = xyz || tets || material || tris || tri_tags || tri_parent;
  if (fill && (*n_nodes != nn || *n_tets != nt || *n_tris != nf)) {
    set_error("mfs_read_gmsh: array sizes differ from the file's counts (call with NULL arrays first)");
    return MFS_ERR_ARG;
  }
  *n_nodes = nn;
  *n_tets = nt;
  *n_tris = nf;
  if (xyz) std::copy(M.xyz.begin(), M.xyz.end(), xyz);
  if (tets) std::copy(M.tets.begin(), M.tets.end(), tets);
  if (material) std::copy(M.mat.begin(), M.mat.end(), material);
  if (tris) std::copy(M.tris.begin(), M.tris.end(), tris);
  if (tri_tags) std::copy(M.tags.begin(), M.tags.end(), tri_tags);
  if (tri_parent) std::copy(M.parent.begin(), M.parent.end(), tri_parent);
  return MFS_OK;
}

// Property file of the Mechanical driver (param_for_solver_%d.txt, read by
// MechanicalMain.cpp:859-1201): one item per line, no comments (the first
// line is read with fscanf and a comment there would be taken as the next
// line).  Dirichlet rows: the tag as one character, then per component a
// value followed by a space or '*' for a free component; the reader only
// takes a value when a space follows it (:963-1000).  Neumann rows: tag,
// value, direction (0 0 0: the value is a normal pressure, :1040-1080).
extern "C" int mfs_write_solver_params(const char* path, int log_flag, int threads, int mesh_format,
                                       const char* mesh_path, double critical_residual, int n_dirichlet,
                                       const mfs_dirichlet_spec* dirichlet, int n_neumann,
                                       const mfs_neumann_spec* neumann, int n_materials,
                                       const mfs_material_spec* materials, const char* result_dir) {
  if (!path || !mesh_path || !result_dir || threads < 1 || (mesh_format != 9 && mesh_format != 10) ||
      n_dirichlet < 0 || n_neumann < 0 || n_materials < 1 || (n_dirichlet > 0 && !dirichlet) ||
      (n_neumann > 0 && !neumann) || !materials) {
    set_error("mfs_write_solver_params: bad argument");
    return MFS_ERR_ARG;
  }
  for (int i = 0; i < n_dirichlet; i++)
    if (dirichlet[i].tag < 0 || dirichlet[i].tag > 9) {
      set_error("mfs_write_solver_params: boundary tags are one digit (0..9) in this format");
      return MFS_ERR_ARG;
    }
  for (int i = 0; i < n_neumann; i++)
    if (neumann[i].tag < 0 || neumann[i].tag > 9) {
      set_error("mfs_write_solver_params: boundary tags are one digit (0..9) in this format");
      return MFS_ERR_ARG;
    }
  for (const char* q : {mesh_path, result_dir})
    if (std::strchr(q, '#') || std::strchr(q, '\n')) {
      set_error("mfs_write_solver_params: paths may not contain '#' or a newline");
      return MFS_ERR_ARG;
    }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("mfs_write_solver_params: cannot open ") + path);
    return MFS_ERR_ARG;
  }
  std::fprintf(f, "%d\n%d\n%d\n%s\n%.17g\n%d\n", log_flag ? 1 : 0, threads, mesh_format, mesh_path,
               critical_residual, n_dirichlet);
  for (int i = 0; i < n_dirichlet; i++) {
    std::fprintf(f, "%d", dirichlet[i].tag);
    for (int a = 0; a < 3; a++) {
      if (dirichlet[i].fixed[a])
        std::fprintf(f, " %.17g ", dirichlet[i].value[a]);
      else
        std::fprintf(f, " *");
    }
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "%d\n", n_neumann);
  for (int i = 0; i < n_neumann; i++)
    std::fprintf(f, "%d %.17g %.17g %.17g %.17g\n", neumann[i].tag, neumann[i].value, neumann[i].direction[0],
                 neumann[i].direction[1], neumann[i].direction[2]);
  std::fprintf(f, "%d\n", n_materials);
  for (int i = 0; i < n_materials; i++) std::fprintf(f, "%.17g %.17g\n", materials[i].E, materials[i].nu);
  std::fprintf(f, "%s\n", result_dir);
  std::fclose(f);
  return MFS_OK;
}
