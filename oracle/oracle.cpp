// FP64 CPU oracle for the FEM solve path of MultiscaleFEMSolver3D.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
// product (libmfs.so) never links, calls or falls back to it.
//
// Every function restates one piece of the reference algorithm and cites the
// reference file:line it follows.  Parity pins (what fixes each function
// independently of this restatement) are listed in the test files; functions
// with no independent pin say "parity unpinned" in their comment.
//
// Conventions kept from the reference:
//   * CSSD storage: per row a sorted "down" list (columns < row), the diagonal,
//     and a sorted "up" list (columns > row)            CSSD_Matrix.h:10-17
//   * elastic element block A[I][J] = sum_q w_q B_J^T D B_I (note the order:
//     the reference puts the J-gradient on the left)     FEM_Grid.h:2125-2191
//   * block -> scalar expansion, row 3i+a, col 3j+b      Math.h:1963-2054
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>
#include <omp.h>

typedef int64_t i64;
typedef uint64_t u64;

#define EXPORT extern "C" __attribute__((visibility("default")))

// ---------------------------------------------------------------------------
// Counter-based RNG (splitmix64 finaliser).  Inputs are synthetic: the
// reference ships no mesh or field generator, so this defines them.
// ---------------------------------------------------------------------------
static inline u64 orc_mix(u64 z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline double orc_uniform(u64 seed, u64 stream, u64 idx) {
  u64 h = orc_mix(orc_mix(seed * 0x100000001B3ull + stream) ^ (idx * 0xD6E8FEB86659FD93ull));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
static inline double orc_normal(u64 seed, u64 stream, u64 idx) {
  double u1 = orc_uniform(seed, stream, 2 * idx);
  double u2 = orc_uniform(seed, stream, 2 * idx + 1);
  if (u1 < 1e-300) u1 = 1e-300;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
EXPORT double orc_rng_uniform(u64 seed, u64 stream, u64 idx) { return orc_uniform(seed, stream, idx); }
EXPORT double orc_rng_normal(u64 seed, u64 stream, u64 idx) { return orc_normal(seed, stream, idx); }

// ---------------------------------------------------------------------------
// Structured Freudenthal mesh of a box: nx*ny*nz hexes, 6 tets per hex split
// along the (0,0,0)-(1,1,1) diagonal; tet t of a hex follows the axis
// permutation perm[t] in the order xyz,xzy,yxz,yzx,zxy,zyx.  Node id is
// i + (nx+1)(j + (ny+1)k).  Hexes are visited x fastest.
// ---------------------------------------------------------------------------
static const int ORC_PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

EXPORT void orc_mesh_box(int nx, int ny, int nz, double x0, double y0, double z0, double hx, double hy, double hz,
                         double* xyz, int* tets) {
  const i64 sx = nx + 1, sy = ny + 1;
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        i64 id = i + sx * (j + sy * k);
        xyz[3 * id + 0] = x0 + hx * i;
        xyz[3 * id + 1] = y0 + hy * j;
        xyz[3 * id + 2] = z0 + hz * k;
      }
  i64 t = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int p = 0; p < 6; p++) {
          int c[3] = {i, j, k};
          tets[4 * t + 0] = (int)(c[0] + sx * (c[1] + sy * c[2]));
          for (int s = 0; s < 3; s++) {
            c[ORC_PERM[p][s]] += 1;
            tets[4 * t + 1 + s] = (int)(c[0] + sx * (c[1] + sy * c[2]));
          }
          t++;
        }
}

// Spherical inclusions: centre U[0,1)^3, radius U(rmin,rmax); a tet takes
// value v_in when its centroid lies inside any sphere (strictly), else v_bg.
EXPORT void orc_field_inclusions(i64 nt, const int* tets, const double* xyz, int count, double rmin, double rmax,
                                 u64 seed, double v_bg, double v_in, double* out) {
  std::vector<double> c(3 * count), r(count);
  for (int s = 0; s < count; s++) {
    for (int a = 0; a < 3; a++) c[3 * s + a] = orc_uniform(seed, 1, 4 * (u64)s + a);
    r[s] = rmin + (rmax - rmin) * orc_uniform(seed, 1, 4 * (u64)s + 3);
  }
  for (i64 t = 0; t < nt; t++) {
    double g[3] = {0, 0, 0};
    for (int v = 0; v < 4; v++)
      for (int a = 0; a < 3; a++) g[a] += xyz[3 * (i64)tets[4 * t + v] + a];
    for (int a = 0; a < 3; a++) g[a] *= 0.25;
    double val = v_bg;
    for (int s = 0; s < count; s++) {
      double dx = g[0] - c[3 * s], dy = g[1] - c[3 * s + 1], dz = g[2] - c[3 * s + 2];
      if (dx * dx + dy * dy + dz * dz < r[s] * r[s]) { val = v_in; break; }
    }
    out[t] = val;
  }
}

// Lognormal per-tet field exp(mu + sigma*N(0,1)) with the normal drawn from
// (seed, stream, global tet index).
EXPORT void orc_field_lognormal(i64 nt, u64 seed, u64 stream, i64 index0, double mu, double sigma, double* out) {
  for (i64 t = 0; t < nt; t++) out[t] = std::exp(mu + sigma * orc_normal(seed, stream, (u64)(index0 + t)));
}

// Thresholded Gaussian random field: g(x) = sum_m cos(kappa_m.x + phi_m),
// kappa_m = n_m / ell with n_m ~ N(0,I), phi_m ~ U[0,2pi), evaluated at the
// nodes.  A tet takes v_hi where the mean of its 4 nodal values is >= 0, else
// v_lo; the phase id (0 / 1) is written to mat.
EXPORT void orc_field_grf(i64 nn, i64 nt, const int* tets, const double* xyz, int modes, double ell, u64 seed,
                          double v_lo, double v_hi, double* out, int* mat) {
  std::vector<double> kap(3 * modes), ph(modes), g(nn);
  for (int m = 0; m < modes; m++) {
    for (int a = 0; a < 3; a++) kap[3 * m + a] = orc_normal(seed, 7, 3 * (u64)m + a) / ell;
    ph[m] = 6.283185307179586 * orc_uniform(seed, 8, (u64)m);
  }
  for (i64 v = 0; v < nn; v++) {
    double s = 0.0;
    for (int m = 0; m < modes; m++)
      s += std::cos(kap[3 * m] * xyz[3 * v] + kap[3 * m + 1] * xyz[3 * v + 1] + kap[3 * m + 2] * xyz[3 * v + 2] + ph[m]);
    g[v] = s;
  }
  for (i64 t = 0; t < nt; t++) {
    double s = 0.0;
    for (int v = 0; v < 4; v++) s += g[tets[4 * t + v]];
    s *= 0.25;
    out[t] = s >= 0.0 ? v_hi : v_lo;
    if (mat) mat[t] = s >= 0.0 ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// Tetrahedron geometry.  GetVolume = |det|/6 (GeometryShape.h:1674-1705);
// SolveAlphaMatrix gives phi_i = a_i0 + a_i1 x + a_i2 y + a_i3 z, i.e. the
// columns of the inverse of [1 x y z] (GeometryShape.h:1960-2086).  Restated
// here by cofactors of the 4x4 matrix.
// ---------------------------------------------------------------------------
static double orc_tet_alpha(const double* P /*4x3*/, double alpha[4][4]) {
  double M[4][4];
  for (int r = 0; r < 4; r++) {
    M[r][0] = 1.0;
    M[r][1] = P[3 * r];
    M[r][2] = P[3 * r + 1];
    M[r][3] = P[3 * r + 2];
  }
  // Gauss-Jordan with partial pivoting on [M | I]
  double A[4][8];
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 8; c++) A[r][c] = c < 4 ? M[r][c] : (c - 4 == r ? 1.0 : 0.0);
  double det = 1.0;
  for (int c = 0; c < 4; c++) {
    int piv = c;
    for (int r = c + 1; r < 4; r++)
      if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
    if (piv != c) {
      for (int k = 0; k < 8; k++) std::swap(A[c][k], A[piv][k]);
      det = -det;
    }
    det *= A[c][c];
    double inv = 1.0 / A[c][c];
    for (int k = 0; k < 8; k++) A[c][k] *= inv;
    for (int r = 0; r < 4; r++)
      if (r != c) {
        double f = A[r][c];
        for (int k = 0; k < 8; k++) A[r][k] -= f * A[c][k];
      }
  }
  // inverse N = A[:,4:8]; phi_i(x) = sum_m [1 x y z]_m N[m][i]
  for (int i = 0; i < 4; i++)
    for (int m = 0; m < 4; m++) alpha[i][m] = A[m][4 + i];
  return std::fabs(det) / 6.0;
}

EXPORT double orc_tet_geometry(const double* P, double* alpha16) {
  double a[4][4];
  double V = orc_tet_alpha(P, a);
  for (int i = 0; i < 4; i++)
    for (int m = 0; m < 4; m++) alpha16[4 * i + m] = a[i][m];
  return V;
}

// Isotropic D in Voigt order (xx,yy,zz,xy,yz,xz): Domain.h:53-78.
static void orc_isotropic_D(double E, double nu, double D[6][6]) {
  double k = E * (1. - nu) / (1. + nu) / (1. - 2. * nu);
  double b = (1. - 2. * nu) / (2. * (1. - nu));
  double a = nu / (1. - nu);
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) D[i][j] = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) D[i][j] = (i == j ? 1.0 : a) * k;
  for (int i = 3; i < 6; i++) D[i][i] = b * k;
}
EXPORT void orc_isotropic_D_export(double E, double nu, double* D36) {
  double D[6][6];
  orc_isotropic_D(E, nu, D);
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) D36[6 * i + j] = D[i][j];
}

// Strain-displacement block of one node, 6x3 (FEM_Grid.h:2159-2168):
// rows xx,yy,zz,xy,yz,xz; derivative.x = derivative.y = derivative.z = grad phi
// (FEM_Grid.h:2914-2928).
static void orc_B(const double g[3], double B[6][3]) {
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 3; j++) B[i][j] = 0.0;
  B[0][0] = g[0];
  B[1][1] = g[1];
  B[2][2] = g[2];
  B[4][2] = g[1];
  B[4][1] = g[2];
  B[5][2] = g[0];
  B[5][0] = g[2];
  B[3][1] = g[0];
  B[3][0] = g[1];
}

// Element stiffness of a linear elastic tet, 12x12 row-major with scalar
// index 3*I+a.  convention 0 = reference: block(I,J) = sum_q w_q B_J^T D B_I
// computed for I<=J and block(J,I) = block(I,J)^T (FEM_Grid.h:2125-2191,
// 4-point rule GeometryShape.h:1754-1766 with weights V/4).
// convention 1 = textbook block(I,J) = V B_I^T D B_J (offered for users whose
// driver expects the symmetric-gradient operator; not a reference path).
EXPORT double orc_elem_elastic(const double* P, double E, double nu, int convention, double* K) {
  double a[4][4];
  double V = orc_tet_alpha(P, a);
  double D[6][6];
  orc_isotropic_D(E, nu, D);
  double B[4][6][3];
  for (int I = 0; I < 4; I++) {
    double g[3] = {a[I][1], a[I][2], a[I][3]};
    orc_B(g, B[I]);
  }
  double w = V / 4.0;
  for (int I = 0; I < 4; I++)
    for (int J = I; J < 4; J++) {
      const double(*Bl)[3] = convention == 0 ? B[J] : B[I];  // left factor, transposed
      const double(*Br)[3] = convention == 0 ? B[I] : B[J];  // right factor
      // gradT_left (3x6) * D (6x6)  ->  T (3x6)  (Math.h:1473-1549 order)
      double T[3][6];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 6; c++) {
          double s = 0;
          for (int k = 0; k < 6; k++) s += Bl[k][r] * D[k][c];
          T[r][c] = s;
        }
      double G[3][3];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          double s = 0;
          for (int k = 0; k < 6; k++) s += T[r][k] * Br[k][c];
          G[r][c] = s;
        }
      // SolveIntegral: result = 0; result += f(x_q) * w_q over 4 points
      double R[3][3];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          double s = 0.0;
          for (int q = 0; q < 4; q++) s += G[r][c] * w;
          R[r][c] = s;
        }
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          K[(3 * I + r) * 12 + 3 * J + c] = R[r][c];
          K[(3 * J + c) * 12 + 3 * I + r] = R[r][c];  // block(J,I) = block(I,J)^T
        }
    }
  return V;
}

// Scalar P1 stiffness sigma*V*grad phi_i . grad phi_j with the 1-point rule
// (FEM_Grid.h:103-151, weight |det|/6 from GeometryShape.h:1760-1761).
EXPORT double orc_elem_scalar(const double* P, double sigma, double* K) {
  double a[4][4];
  double V = orc_tet_alpha(P, a);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double d = a[i][1] * a[j][1] + a[i][2] * a[j][2] + a[i][3] * a[j][3];
      double s = 0.0;
      s += (sigma * d) * V;
      K[4 * i + j] = s;
    }
  return V;
}

// ---------------------------------------------------------------------------
// CSSD matrix (CSSD_Matrix.h:7-17) with double entries, rows in the
// reference's vector-of-vectors form.
// ---------------------------------------------------------------------------
struct Cssd {
  std::vector<double> Diag, X, F;
  std::vector<std::vector<double>> A_up, A_down;
  std::vector<std::vector<int>> up, down;
  int n() const { return (int)Diag.size(); }
  void init(std::vector<std::vector<int>>& u, std::vector<std::vector<int>>& d) {  // CSSD_Matrix.h:79-99
    int n = (int)u.size();
    up = u;
    down = d;
    Diag.assign(n, 0.0);
    X.assign(n, 0.0);
    F.assign(n, 0.0);
    A_up.resize(n);
    A_down.resize(n);
    for (int i = 0; i < n; i++) {
      A_up[i].assign(up[i].size(), 0.0);
      A_down[i].assign(down[i].size(), 0.0);
    }
  }
  // MultiplicationMatrixVector: CSSD_Matrix.h:395-423, schedule(dynamic)
  void mult(const std::vector<double>& B, std::vector<double>& res, int threads) const {
    int n = this->n();
    res.resize(n);
#pragma omp parallel for schedule(dynamic) num_threads(threads)
    for (int i = 0; i < n; i++) {
      double tmp = 0.0;
      for (size_t j = 0; j < A_up[i].size(); j++) tmp += A_up[i][j] * B[up[i][j]];
      for (size_t j = 0; j < A_down[i].size(); j++) tmp += A_down[i][j] * B[down[i][j]];
      tmp += Diag[i] * B[i];
      res[i] = tmp;
    }
  }
};

// Portrait from element node lists (FEM_Grid.h:3236-3287 / :665-716):
// for local pairs i<j push (max,min) to down and (min,max) to up, then
// sort and remove duplicates (Math.h:1674-1727).
static void orc_portrait(int nrows, i64 nt, int npe, const int* elem, std::vector<std::vector<int>>& up,
                         std::vector<std::vector<int>>& down) {
  std::vector<std::vector<int>> tu(nrows), td(nrows);
  for (i64 e = 0; e < nt; e++)
    for (int i = 0; i < npe; i++) {
      int gi = elem[npe * e + i];
      for (int j = i + 1; j < npe; j++) {
        int gj = elem[npe * e + j];
        if (gi > gj) {
          td[gi].push_back(gj);
          tu[gj].push_back(gi);
        } else {
          td[gj].push_back(gi);
          tu[gi].push_back(gj);
        }
      }
    }
  up.assign(nrows, {});
  down.assign(nrows, {});
  for (int r = 0; r < nrows; r++) {
    std::sort(tu[r].begin(), tu[r].end());
    std::sort(td[r].begin(), td[r].end());
    for (int c : tu[r])
      if (up[r].empty() || up[r].back() != c) up[r].push_back(c);
    for (int c : td[r])
      if (down[r].empty() || down[r].back() != c) down[r].push_back(c);
  }
}

// SummPartOfMatrix with the reference's linear column search
// (CSSD_Matrix.h:307-363).  K is local, row-major, size m x m; rows is the
// global index of each local row.
static int orc_summ_part(Cssd& A, int m, const double* K, const int* rows) {
  int missing = 0;
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++) {
      int s = rows[i], c = rows[j];
      bool find = false;
      if (s > c) {
        for (size_t jj = 0; jj < A.down[s].size(); jj++)
          if (A.down[s][jj] == c) { A.A_down[s][jj] += K[i * m + j]; find = true; }
      } else if (s < c) {
        for (size_t jj = 0; jj < A.up[s].size(); jj++)
          if (A.up[s][jj] == c) { A.A_up[s][jj] += K[i * m + j]; find = true; }
      } else {
        A.Diag[s] += K[i * m + i];
        find = true;
      }
      if (!find) missing++;
    }
  return missing;
}

// ---------------------------------------------------------------------------
// Assembled system handle
// ---------------------------------------------------------------------------
struct OrcSystem {
  Cssd A;                  // scalar CSSD after BCs
  std::vector<int> node_map;  // original node -> DOF node (elastic renumbering)
  int missing = 0;
};

static void orc_to_csr(const Cssd& A, std::vector<int>& rowptr, std::vector<int>& col, std::vector<double>& val) {
  int n = A.n();
  rowptr.assign(n + 1, 0);
  for (int i = 0; i < n; i++) rowptr[i + 1] = rowptr[i] + (int)(A.down[i].size() + 1 + A.up[i].size());
  col.resize(rowptr[n]);
  val.resize(rowptr[n]);
  for (int i = 0; i < n; i++) {
    int p = rowptr[i];
    for (size_t j = 0; j < A.down[i].size(); j++) { col[p] = A.down[i][j]; val[p++] = A.A_down[i][j]; }
    col[p] = i;
    val[p++] = A.Diag[i];
    for (size_t j = 0; j < A.up[i].size(); j++) { col[p] = A.up[i][j]; val[p++] = A.A_up[i][j]; }
  }
}

// Dirichlet by elimination on a scalar CSSD (FEM.h:505-619): fixed rows become
// identity with F=X=g; then every row subtracts A_rc*g_c for fixed columns c
// and zeroes A_rc.  `fixed` is the ascending list of fixed DOFs (the monotone
// scan at FEM.h:569-613 assumes ascending order).
static void orc_eliminate(Cssd& A, const std::vector<int>& fixed, const std::vector<double>& g) {
  int n = A.n();
  std::vector<char> isfix(n, 0);
  std::vector<double> gval(n, 0.0);
  for (size_t k = 0; k < fixed.size(); k++) {
    int d = fixed[k];
    isfix[d] = 1;
    gval[d] = g[k];
    A.X[d] = g[k];
    A.F[d] = g[k];
    A.Diag[d] = 1.0;
    std::fill(A.A_down[d].begin(), A.A_down[d].end(), 0.0);
    std::fill(A.A_up[d].begin(), A.A_up[d].end(), 0.0);
  }
  for (int r = 0; r < n; r++) {
    for (size_t jj = 0; jj < A.up[r].size(); jj++) {
      int c = A.up[r][jj];
      if (isfix[c]) { A.F[r] -= A.A_up[r][jj] * A.F[c]; A.A_up[r][jj] = 0.0; }
    }
    for (size_t jj = 0; jj < A.down[r].size(); jj++) {
      int c = A.down[r][jj];
      if (isfix[c]) { A.F[r] -= A.A_down[r][jj] * A.F[c]; A.A_down[r][jj] = 0.0; }
    }
  }
}

// Scalar elliptic system (FEM.h:398-619): DOF = node id, element matrices in
// element order, serial SummPartOfMatrix, Dirichlet elimination.
EXPORT void* orc_scalar_system(int nn, i64 nt, const double* xyz, const int* tets, const double* sigma, int nfix,
                               const int* fixed_nodes, const double* fixed_vals) {
  OrcSystem* S = new OrcSystem();
  std::vector<std::vector<int>> up, down;
  orc_portrait(nn, nt, 4, tets, up, down);
  S->A.init(up, down);
  for (i64 e = 0; e < nt; e++) {
    double P[12];
    for (int v = 0; v < 4; v++)
      for (int a = 0; a < 3; a++) P[3 * v + a] = xyz[3 * (i64)tets[4 * e + v] + a];
    double K[16];
    orc_elem_scalar(P, sigma[e], K);
    S->missing += orc_summ_part(S->A, 4, K, tets + 4 * e);
  }
  // sort fixed list ascending, values with it
  std::vector<int> idx(nfix);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return fixed_nodes[a] < fixed_nodes[b]; });
  std::vector<int> fx;
  std::vector<double> gv;
  for (int k : idx) {
    if (!fx.empty() && fx.back() == fixed_nodes[k]) continue;
    fx.push_back(fixed_nodes[k]);
    gv.push_back(fixed_vals[k]);
  }
  orc_eliminate(S->A, fx, gv);
  S->node_map.resize(nn);
  std::iota(S->node_map.begin(), S->node_map.end(), 0);
  return S;
}

// Mechanical renumbering (FEM.h:1990-2081): Dirichlet vertices first, in list
// order (each tag list sorted unique, MechanicalMain.cpp:1012-1024), then
// remaining vertices by first appearance scanning elements of material
// 0..max in element order.
EXPORT int orc_renumber(int nn, i64 nt, const int* tets, const int* mat, int nfix, const int* fixed_nodes, int* map) {
  for (int i = 0; i < nn; i++) map[i] = -1;
  int k = 0;
  for (int i = 0; i < nfix; i++)
    if (map[fixed_nodes[i]] == -1) map[fixed_nodes[i]] = k++;
  int mmax = 0;
  for (i64 e = 0; e < nt; e++) mmax = std::max(mmax, mat ? mat[e] : 0);
  for (int m = 0; m <= mmax; m++)
    for (i64 e = 0; e < nt; e++)
      if ((mat ? mat[e] : 0) == m)
        for (int j = 0; j < 4; j++)
          if (map[tets[4 * e + j]] == -1) map[tets[4 * e + j]] = k++;
  return k;
}

// Elastic system (FEM.h:1963-2223 + Math.h:1963-2054): renumber, block
// portrait, Neumann loads, element matrices (reference convention unless
// convention==1), serial block assembly, block->scalar copy, then Dirichlet
// elimination of every fixed component (the build replaces the reference's
// 1e20 penalty, FEM.h:2178-2221, by elimination; the solution agrees to
// O(1e-20) relative).  fixed nodes: all 3 components fixed to fixed_vals[3k..].
// Neumann: triangles (3 original node ids each) with traction t[3] per tri:
// f_I += t * area/3, the exact value of the reference's 4-point triangle rule
// on P1 (FEM_Grid.h:2622-2661).
EXPORT void* orc_elastic_system(int nn, i64 nt, const double* xyz, const int* tets, const int* mat, const double* E,
                                double nu, int nfix, const int* fixed_nodes, const double* fixed_vals, int ntri,
                                const int* tris, const double* traction, int convention, int renumber) {
  OrcSystem* S = new OrcSystem();
  S->node_map.resize(nn);
  if (renumber)
    orc_renumber(nn, nt, tets, mat, nfix, fixed_nodes, S->node_map.data());
  else
    std::iota(S->node_map.begin(), S->node_map.end(), 0);
  const std::vector<int>& map = S->node_map;
  std::vector<int> rt(4 * nt);
  for (i64 i = 0; i < 4 * nt; i++) rt[i] = map[tets[i]];
  std::vector<std::vector<int>> up, down;
  orc_portrait(nn, nt, 4, rt.data(), up, down);
  // block matrix held as 9 scalar CSSD-like arrays per entry
  struct Blk { double v[9]; };
  std::vector<Blk> bd(nn);
  std::vector<std::vector<Blk>> bu(nn), bdn(nn);
  for (int i = 0; i < nn; i++) {
    bu[i].assign(up[i].size(), Blk{});
    bdn[i].assign(down[i].size(), Blk{});
    std::memset(bd[i].v, 0, sizeof(bd[i].v));
    for (auto& b : bu[i]) std::memset(b.v, 0, sizeof(b.v));
    for (auto& b : bdn[i]) std::memset(b.v, 0, sizeof(b.v));
  }
  std::vector<double> Fb(3 * (size_t)nn, 0.0);
  // Neumann (SummPartOfVector, CSSD_Matrix.h:364-379)
  for (int t = 0; t < ntri; t++) {
    const double* p0 = xyz + 3 * (i64)tris[3 * t];
    const double* p1 = xyz + 3 * (i64)tris[3 * t + 1];
    const double* p2 = xyz + 3 * (i64)tris[3 * t + 2];
    double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double v[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
    double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    for (int q = 0; q < 3; q++) {
      int d = map[tris[3 * t + q]];
      for (int a = 0; a < 3; a++) Fb[3 * (size_t)d + a] += traction[3 * t + a] * area / 3.0;
    }
  }
  // assembly in element order with the linear column search
  for (i64 e = 0; e < nt; e++) {
    double P[12];
    for (int v = 0; v < 4; v++)
      for (int a = 0; a < 3; a++) P[3 * v + a] = xyz[3 * (i64)tets[4 * e + v] + a];
    double K[144];
    orc_elem_elastic(P, E[e], nu, convention, K);
    const int* rows = rt.data() + 4 * e;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        int s = rows[i], c = rows[j];
        Blk* dst = nullptr;
        if (s > c) {
          for (size_t jj = 0; jj < down[s].size(); jj++)
            if (down[s][jj] == c) dst = &bdn[s][jj];
        } else if (s < c) {
          for (size_t jj = 0; jj < up[s].size(); jj++)
            if (up[s][jj] == c) dst = &bu[s][jj];
        } else {
          dst = &bd[s];
        }
        if (!dst) { S->missing++; continue; }
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) dst->v[3 * a + b] += K[(3 * i + a) * 12 + 3 * j + b];
      }
  }
  // block -> scalar (Math.h:1963-2054): row 3i+ii: down = {3c+jj : c in down(i)}
  // ++ [3i .. 3i+ii-1]; up = [3i+ii+1 .. 3i+2] ++ {3c+jj : c in up(i)}.
  int n = 3 * nn;
  std::vector<std::vector<int>> su(n), sd(n);
  for (int i = 0; i < nn; i++)
    for (int ii = 0; ii < 3; ii++) {
      int r = 3 * i + ii;
      for (int jj = 0; jj < 2 - ii; jj++) su[r].push_back(3 * i + ii + jj + 1);
      for (int c : up[i])
        for (int jj = 0; jj < 3; jj++) su[r].push_back(3 * c + jj);
      for (int c : down[i])
        for (int jj = 0; jj < 3; jj++) sd[r].push_back(3 * c + jj);
      for (int jj = 0; jj < ii; jj++) sd[r].push_back(3 * i + jj);
    }
  S->A.init(su, sd);
  for (int i = 0; i < nn; i++)
    for (int ii = 0; ii < 3; ii++) {
      int r = 3 * i + ii;
      S->A.Diag[r] = bd[i].v[3 * ii + ii];
      S->A.F[r] = Fb[r];
      int p = 0;
      for (int jj = 0; jj < 2 - ii; jj++) S->A.A_up[r][p++] = bd[i].v[3 * ii + ii + jj + 1];
      for (size_t c = 0; c < up[i].size(); c++)
        for (int jj = 0; jj < 3; jj++) S->A.A_up[r][p++] = bu[i][c].v[3 * ii + jj];
      p = 0;
      for (size_t c = 0; c < down[i].size(); c++)
        for (int jj = 0; jj < 3; jj++) S->A.A_down[r][p++] = bdn[i][c].v[3 * ii + jj];
      for (int jj = 0; jj < ii; jj++) S->A.A_down[r][p++] = bd[i].v[3 * ii + jj];
    }
  // Dirichlet elimination on scalar DOFs, ascending
  std::vector<std::pair<int, double>> fx;
  for (int k = 0; k < nfix; k++) {
    int d = map[fixed_nodes[k]];
    for (int a = 0; a < 3; a++) fx.push_back({3 * d + a, fixed_vals[3 * k + a]});
  }
  std::sort(fx.begin(), fx.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
  std::vector<int> fi;
  std::vector<double> gv;
  for (auto& p : fx) {
    if (!fi.empty() && fi.back() == p.first) continue;
    fi.push_back(p.first);
    gv.push_back(p.second);
  }
  orc_eliminate(S->A, fi, gv);
  return S;
}

// Block-only assembly (no BCs, no renumbering): the raw block stiffness as a
// scalar CSR, used to check invariants (rigid translations in the null space).
EXPORT i64 orc_sys_n(void* h) { return ((OrcSystem*)h)->A.n(); }
EXPORT i64 orc_sys_nnz(void* h) {
  const Cssd& A = ((OrcSystem*)h)->A;
  i64 s = 0;
  for (int i = 0; i < A.n(); i++) s += A.up[i].size() + A.down[i].size() + 1;
  return s;
}
EXPORT int orc_sys_missing(void* h) { return ((OrcSystem*)h)->missing; }
EXPORT void orc_sys_csr(void* h, int* rowptr, int* col, double* val, double* rhs, double* x0) {
  OrcSystem* S = (OrcSystem*)h;
  std::vector<int> rp, c;
  std::vector<double> v;
  orc_to_csr(S->A, rp, c, v);
  std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int));
  std::memcpy(col, c.data(), c.size() * sizeof(int));
  std::memcpy(val, v.data(), v.size() * sizeof(double));
  if (rhs) std::memcpy(rhs, S->A.F.data(), S->A.F.size() * sizeof(double));
  if (x0) std::memcpy(x0, S->A.X.data(), S->A.X.size() * sizeof(double));
}
EXPORT void orc_sys_node_map(void* h, int* map) {
  OrcSystem* S = (OrcSystem*)h;
  std::memcpy(map, S->node_map.data(), S->node_map.size() * sizeof(int));
}
EXPORT void orc_sys_free(void* h) { delete (OrcSystem*)h; }

// ---------------------------------------------------------------------------
// Solvers on a CSR matrix (row = down ++ diag ++ up, i.e. the CSSD rows).
// ---------------------------------------------------------------------------
static void csr_mv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y, int threads) {
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int p = rp[i]; p < rp[i + 1]; p++) s += v[p] * x[ci[p]];
    y[i] = s;
  }
}
static double dotp(int n, const double* a, const double* b) {  // Math.h:1633-1654 (serial)
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

EXPORT void orc_csr_mv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y) {
  csr_mv(n, rp, ci, v, x, y, 1);
}

// CG (MSG, CSSD_Matrix.h:1496-1576) when precond==0; Jacobi-preconditioned CG
// when precond==1 (the build's "diagonal preconditioner", BiCG_Diag's
// preconditioner applied to CG for SPD systems, CSSD_Matrix.h:1166-1300).
// Stops when the recursive ||r||/||b|| <= tol.  Returns iterations; writes the
// true relative residual ||b-Ax||/||b||.
EXPORT int orc_pcg(int n, const int* rp, const int* ci, const double* v, const double* b, double* x, double tol,
                   int maxit, int precond, double* true_relres) {
  std::vector<double> r(n), z(n), p(n), q(n), dinv(n, 1.0);
  if (precond)
    for (int i = 0; i < n; i++)
      for (int k = rp[i]; k < rp[i + 1]; k++)
        if (ci[k] == i) dinv[i] = 1.0 / v[k];
  csr_mv(n, rp, ci, v, x, q.data(), 1);
  for (int i = 0; i < n; i++) r[i] = b[i] - q[i];
  double bb = dotp(n, b, b);
  if (bb == 0) bb = 1;
  for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
  p = z;
  double rz = dotp(n, r.data(), z.data());
  int k = 0;
  while (k < maxit && std::sqrt(dotp(n, r.data(), r.data()) / bb) > tol) {
    csr_mv(n, rp, ci, v, p.data(), q.data(), 1);
    double alpha = rz / dotp(n, p.data(), q.data());
    for (int i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    double rz2 = dotp(n, r.data(), z.data());
    double beta = rz2 / rz;
    rz = rz2;
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    k++;
  }
  csr_mv(n, rp, ci, v, x, q.data(), 1);
  double t = 0;
  for (int i = 0; i < n; i++) t += (b[i] - q[i]) * (b[i] - q[i]);
  if (true_relres) *true_relres = std::sqrt(t / bb);
  return k;
}

// BiCGStab with right Jacobi preconditioning when precond==1: the
// reference's BiCG_Stab_Preconditioning (CSSD_Matrix.h:1058-1163) with its
// inner M-solve replaced by M = diag(A): p^ = M^-1 p, v = A p^, alpha =
// (r,r0)/(r0,v), s = r - alpha v, s^ = M^-1 s, t = A s^, omega = (t,s)/(t,t),
// x += alpha p^ + omega s^, r = s - omega t, beta = (rho'/rho)(alpha/omega),
// p = r + beta (p - omega v) (CSSD_Matrix.h:1094-1115).  Stops when the
// recursive ||r|| <= tol ||b|| (the reference checks the true residual every
// 10 steps, :1117-1126).  Returns iterations; true relative residual out.
EXPORT int orc_bicgstab(int n, const int* rp, const int* ci, const double* v, const double* b, double* x, double tol,
                        int maxit, int precond, double* true_relres) {
  std::vector<double> r(n), r0(n), p(n), vv(n), sv(n), t(n), ph(n), sh(n), q(n), dinv(n, 1.0);
  if (precond)
    for (int i = 0; i < n; i++)
      for (int k = rp[i]; k < rp[i + 1]; k++)
        if (ci[k] == i && v[k] != 0) dinv[i] = 1.0 / v[k];
  csr_mv(n, rp, ci, v, x, q.data(), 1);
  for (int i = 0; i < n; i++) r[i] = b[i] - q[i];
  r0 = r;
  p = r;
  const double bb = dotp(n, b, b) > 0 ? dotp(n, b, b) : 1.0;
  double rho = dotp(n, r0.data(), r.data());
  int k = 0;
  double rr = dotp(n, r.data(), r.data());
  while (k < maxit && rr > tol * tol * bb) {
    for (int i = 0; i < n; i++) ph[i] = dinv[i] * p[i];
    csr_mv(n, rp, ci, v, ph.data(), vv.data(), 1);
    const double alpha = rho / dotp(n, r0.data(), vv.data());
    for (int i = 0; i < n; i++) sv[i] = r[i] - alpha * vv[i];
    for (int i = 0; i < n; i++) sh[i] = dinv[i] * sv[i];
    csr_mv(n, rp, ci, v, sh.data(), t.data(), 1);
    const double omega = dotp(n, t.data(), sv.data()) / dotp(n, t.data(), t.data());
    for (int i = 0; i < n; i++) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = sv[i] - omega * t[i];
    }
    const double rho_new = dotp(n, r0.data(), r.data());
    rr = dotp(n, r.data(), r.data());
    k++;
    if (!std::isfinite(rr) || rho_new == 0.0 || omega == 0.0) break;
    const double beta = (rho_new / rho) * (alpha / omega);
    rho = rho_new;
    for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * vv[i]);
  }
  csr_mv(n, rp, ci, v, x, q.data(), 1);
  double tr = 0;
  for (int i = 0; i < n; i++) tr += (b[i] - q[i]) * (b[i] - q[i]);
  if (true_relres) *true_relres = std::sqrt(tr / bb);
  return k;
}

// Restarted GMRES(m) with modified Gram-Schmidt and Givens rotations, right
// Jacobi preconditioning when precond==1 (GMRES, CSSD_Matrix.h:1727-1989 with
// M generalised from 5; the reference's row scaling is a preconditioner and is
// replaced by Jacobi here).  Stops when the true relative residual after a
// cycle, or the estimated one inside a cycle, is <= tol.  Returns the total
// number of Arnoldi steps.
EXPORT int orc_gmres(int n, const int* rp, const int* ci, const double* v, const double* b, double* x, int m,
                     double tol, int maxit, int precond, double* true_relres) {
  std::vector<double> dinv(n, 1.0);
  if (precond)
    for (int i = 0; i < n; i++)
      for (int k = rp[i]; k < rp[i + 1]; k++)
        if (ci[k] == i) dinv[i] = 1.0 / v[k];
  std::vector<std::vector<double>> V(m + 1, std::vector<double>(n));
  std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1), w(n), r(n), t(n);
  double bn = std::sqrt(dotp(n, b, b));
  if (bn == 0) bn = 1;
  int steps = 0;
  double rel = 0;
  while (true) {
    csr_mv(n, rp, ci, v, x, t.data(), 1);
    for (int i = 0; i < n; i++) r[i] = b[i] - t[i];
    double beta = std::sqrt(dotp(n, r.data(), r.data()));
    rel = beta / bn;
    if (rel <= tol || steps >= maxit) break;
    for (int i = 0; i < n; i++) V[0][i] = r[i] / beta;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m && steps < maxit; j++) {
      for (int i = 0; i < n; i++) t[i] = dinv[i] * V[j][i];
      csr_mv(n, rp, ci, v, t.data(), w.data(), 1);
      for (int i = 0; i <= j; i++) {
        double h = dotp(n, w.data(), V[i].data());
        H[i * m + j] = h;
        for (int l = 0; l < n; l++) w[l] -= h * V[i][l];
      }
      double hn = std::sqrt(dotp(n, w.data(), w.data()));
      H[(j + 1) * m + j] = hn;
      if (hn != 0)
        for (int l = 0; l < n; l++) V[j + 1][l] = w[l] / hn;
      for (int i = 0; i < j; i++) {
        double a = H[i * m + j], c = H[(i + 1) * m + j];
        H[i * m + j] = cs[i] * a + sn[i] * c;
        H[(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
      }
      double a = H[j * m + j], c = H[(j + 1) * m + j];
      double d = std::sqrt(a * a + c * c);
      cs[j] = a / d;
      sn[j] = c / d;
      H[j * m + j] = d;
      H[(j + 1) * m + j] = 0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      steps++;
      if (std::fabs(g[j + 1]) / bn <= tol || hn == 0) { j++; break; }
    }
    std::vector<double> y(j);
    for (int i = j - 1; i >= 0; i--) {
      double s = g[i];
      for (int k2 = i + 1; k2 < j; k2++) s -= H[i * m + k2] * y[k2];
      y[i] = s / H[i * m + i];
    }
    std::fill(t.begin(), t.end(), 0.0);
    for (int i = 0; i < j; i++)
      for (int l = 0; l < n; l++) t[l] += y[i] * V[i][l];
    for (int l = 0; l < n; l++) x[l] += dinv[l] * t[l];
  }
  if (true_relres) *true_relres = rel;
  return steps;
}

// Dense Cholesky solve for tiny SPD systems (pins CG on small meshes).
EXPORT int orc_dense_solve(int n, const int* rp, const int* ci, const double* v, const double* b, double* x) {
  std::vector<double> A((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++)
    for (int k = rp[i]; k < rp[i + 1]; k++) A[(size_t)i * n + ci[k]] = v[k];
  // LU with partial pivoting (works for the indefinite reference operator too)
  std::vector<int> piv(n);
  std::vector<double> y(b, b + n);
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)p * n + c])) p = r;
    if (A[(size_t)p * n + c] == 0) return -1;
    if (p != c) {
      for (int k = 0; k < n; k++) std::swap(A[(size_t)p * n + k], A[(size_t)c * n + k]);
      std::swap(y[p], y[c]);
    }
    for (int r = c + 1; r < n; r++) {
      double f = A[(size_t)r * n + c] / A[(size_t)c * n + c];
      if (f == 0) continue;
      for (int k = c; k < n; k++) A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
      y[r] -= f * y[c];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = y[i];
    for (int k = i + 1; k < n; k++) s -= A[(size_t)i * n + k] * x[k];
    x[i] = s / A[(size_t)i * n + i];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Reference CPU path for the CPU baseline: CSSD rows, OpenMP schedule(dynamic)
// SpMV and serial dot products.
// ---------------------------------------------------------------------------
static Cssd cssd_from_csr(int n, const int* rp, const int* ci, const double* v) {
  Cssd A;
  std::vector<std::vector<int>> up(n), down(n);
  for (int i = 0; i < n; i++)
    for (int k = rp[i]; k < rp[i + 1]; k++) {
      if (ci[k] < i) down[i].push_back(ci[k]);
      else if (ci[k] > i) up[i].push_back(ci[k]);
    }
  A.init(up, down);
  for (int i = 0; i < n; i++) {
    int pu = 0, pd = 0;
    for (int k = rp[i]; k < rp[i + 1]; k++) {
      if (ci[k] < i) A.A_down[i][pd++] = v[k];
      else if (ci[k] > i) A.A_up[i][pu++] = v[k];
      else A.Diag[i] = v[k];
    }
  }
  return A;
}

// SSOR "preconditioner" as the reference builds it (CSSD_Matrix.h:458-513):
// off-diagonals stored as -w*a and the diagonal as 1/(a_ii w (2-w)); the solve
// (CSSD_Matrix.h:514-547) then SUBTRACTS the stored -w*a, i.e. it applies SSOR
// of (2D - A).  Restated exactly, sign included.
struct Ssor {
  Cssd P;
};
static void ssor_build(const Cssd& A, double w, Cssd& P) {
  P = A;
  for (int I = 0; I < A.n(); I++) {
    P.Diag[I] = 1.0 / A.Diag[I] / w / (2 - w);
    for (auto& a : P.A_up[I]) a = -1 * w * a;
    for (auto& a : P.A_down[I]) a = -1 * w * a;
  }
}
static void ssor_apply(const Cssd& P, const std::vector<double>& free_diag, std::vector<double>& X,
                       const std::vector<double>& Z) {
  int n = P.n();
  X.resize(n);
  for (int I = n - 1; I >= 0; I--) {
    X[I] = Z[I];
    for (size_t j = 0; j < P.up[I].size(); j++) X[I] -= X[P.up[I][j]] * P.A_up[I][j];
    X[I] /= free_diag[I];
  }
  for (int I = 0; I < n; I++) X[I] = X[I] / P.Diag[I];
  for (int I = 0; I < n; I++) {
    for (size_t j = 0; j < P.down[I].size(); j++) X[I] -= X[P.down[I][j]] * P.A_down[I][j];
    X[I] /= free_diag[I];
  }
}

// MSG_PreconditioningSSOR (CSSD_Matrix.h:1302-1398) restated: PCG with the
// sign-flipped SSOR, stop on the unpreconditioned ||r||/||F|| (the loop
// re-evaluates both dot products every iteration, as the reference does).
// mode 0: the reference's own SSOR-PCG; mode 1: plain CG (MSG,
// CSSD_Matrix.h:1496-1576).  Returns iterations; true relative residual out.
// threads = OpenMP threads for the SpMV (the reference's only parallel part).
EXPORT int orc_refpath_solve(int n, const int* rp, const int* ci, const double* v, const double* b, double* x,
                             double tol, int maxit, int mode, int threads, double* true_relres, double* seconds) {
  Cssd A = cssd_from_csr(n, rp, ci, v);
  A.F.assign(b, b + n);
  A.X.assign(x, x + n);
  double t0 = omp_get_wtime();
  std::vector<double> r(n), z(n), p(n), tm(n);
  Cssd P;
  if (mode == 0) ssor_build(A, 0.75, P);
  A.mult(A.X, tm, threads);
  for (int i = 0; i < n; i++) r[i] = A.F[i] - tm[i];
  if (mode == 0) ssor_apply(P, A.Diag, z, r);
  else z = r;
  p = z;
  double scal_r_z = dotp(n, r.data(), z.data());
  int k;
  for (k = 1; k < maxit && std::sqrt(dotp(n, r.data(), r.data()) / dotp(n, A.F.data(), A.F.data())) > tol; k++) {
    A.mult(p, tm, threads);
    double alpha = scal_r_z / dotp(n, tm.data(), p.data());
    for (int i = 0; i < n; i++) {
      A.X[i] += alpha * p[i];
      r[i] -= alpha * tm[i];
    }
    if (mode == 0) ssor_apply(P, A.Diag, z, r);
    else z = r;
    double betta = dotp(n, r.data(), z.data()) / scal_r_z;
    for (int i = 0; i < n; i++) p[i] = z[i] + betta * p[i];
    scal_r_z = dotp(n, r.data(), z.data());
  }
  A.mult(A.X, tm, threads);
  double t1 = 0, t2 = 0;
  for (int i = 0; i < n; i++) {
    t1 += (tm[i] - A.F[i]) * (tm[i] - A.F[i]);
    t2 += A.F[i] * A.F[i];
  }
  if (seconds) *seconds = omp_get_wtime() - t0;
  if (true_relres) *true_relres = std::sqrt(t1 / t2);
  std::memcpy(x, A.X.data(), n * sizeof(double));
  return k;
}

// BiCG_Diag (CSSD_Matrix.h:1181-1300) restated: Jacobi-preconditioned BiCG
// with 2 SpMVs per iteration, stop on the ABSOLUTE preconditioned recursive
// norm; returns that norm (the reference's return value) and iterations.
EXPORT double orc_bicg_diag(int n, const int* rp, const int* ci, const double* v, const double* b, double* x,
                            double E, int maxiter, const double* diag, int* iters) {
  Cssd A = cssd_from_csr(n, rp, ci, v);
  std::vector<double> X(x, x + n), F(b, b + n);
  std::vector<double> r(n), p(n), r_(n), p_(n), vec(n), vec_(n), vh(n);
  A.mult(X, vh, 1);
  for (int i = 0; i < n; i++) vh[i] = F[i] - vh[i];
  for (int i = 0; i < n; i++) r[i] = vh[i] / diag[i];
  double res = std::sqrt(dotp(n, r.data(), r.data()));
  p = r;
  r_ = r;
  p_ = r_;
  bool flag = false;
  int it = 0;
  while (!flag && it < maxiter) {
    double sc1 = dotp(n, r.data(), r_.data());
    A.mult(p, vh, 1);
    for (int i = 0; i < n; i++) vec[i] = vh[i] / diag[i];
    double sc2 = dotp(n, vec.data(), p_.data());
    double alpha = sc1 / sc2;
    res = 0;
    for (int i = 0; i < n; i++) {
      X[i] += alpha * p[i];
      r[i] -= alpha * vec[i];
      res += r[i] * r[i];
    }
    res = std::sqrt(res);
    for (int i = 0; i < n; i++) vh[i] = p_[i] / diag[i];
    A.mult(vh, vec_, 1);
    for (int i = 0; i < n; i++) r_[i] -= alpha * vec_[i];
    sc2 = dotp(n, r.data(), r_.data());
    double beta = sc2 / sc1;
    if (res < E) flag = true;
    if (!flag)
      for (int i = 0; i < n; i++) {
        p[i] = r[i] + beta * p[i];
        p_[i] = r_[i] + beta * p_[i];
      }
    it++;
  }
  std::memcpy(x, X.data(), n * sizeof(double));
  if (iters) *iters = it;
  return res;
}

// ---------------------------------------------------------------------------
// MsFEM local problems on one coarse hex cell (three-direction formulation:
// three unit directions per coarse node; the boundary trace is the trilinear hat of
// the coarse node times e_d on the fine boundary nodes; zero body force).  The
// reference solves one vector basis per node with edge->face->volume traces
// (MsFEM_Grid.h:4795-5760, FEM.h:2969-3795); by linearity its (1,1,1) basis is
// the sum over d of these when the same traces are used.
//
// Fine mesh: n^3 hexes of the cell [x0,x0+h]^3 (orc_mesh_box), element matrices
// with the reference convention, per-tet modulus E[t], Poisson nu.  Output:
// basis[(node*3 + d) * nfine*3 + 3*fine_node + a] for the 8 coarse nodes (x
// fastest), full fine vectors including boundary values; and the 24x24
// Galerkin matrix Kc = Phi^T K Phi.  Solves with Jacobi-PCG to tol.
// ---------------------------------------------------------------------------
EXPORT int orc_msfem_cell(int n, double x0, double y0, double z0, double h, const double* E, double nu, double tol,
                          double* basis, double* Kc, int* iters_out) {
  int s = n + 1;
  int nn = s * s * s;
  i64 nt = 6 * (i64)n * n * n;
  std::vector<double> xyz(3 * (size_t)nn);
  std::vector<int> tets(4 * nt);
  orc_mesh_box(n, n, n, x0, y0, z0, h / n, h / n, h / n, xyz.data(), tets.data());
  // raw (no BC) elastic matrix in natural node order
  void* raw = orc_elastic_system(nn, nt, xyz.data(), tets.data(), nullptr, E, nu, 0, nullptr, nullptr, 0, nullptr,
                                 nullptr, 0, 0);
  OrcSystem* S = (OrcSystem*)raw;
  std::vector<int> rp, ci;
  std::vector<double> vv;
  orc_to_csr(S->A, rp, ci, vv);
  int N = 3 * nn;
  std::vector<char> bnd(nn, 0);
  for (int k = 0; k < s; k++)
    for (int j = 0; j < s; j++)
      for (int i = 0; i < s; i++)
        if (i == 0 || j == 0 || k == 0 || i == n || j == n || k == n) bnd[i + s * (j + s * k)] = 1;
  // interior system: rows/cols of interior DOFs
  std::vector<int> loc(N, -1);
  int ni = 0;
  for (int d = 0; d < N; d++)
    if (!bnd[d / 3]) loc[d] = ni++;
  std::vector<int> irp(ni + 1, 0), ici;
  std::vector<double> iv;
  for (int d = 0; d < N; d++) {
    if (loc[d] < 0) continue;
    for (int k = rp[d]; k < rp[d + 1]; k++)
      if (loc[ci[k]] >= 0) { ici.push_back(loc[ci[k]]); iv.push_back(vv[k]); }
    irp[loc[d] + 1] = (int)ici.size();
  }
  int maxit_seen = 0;
  for (int c = 0; c < 8; c++) {
    int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    for (int d = 0; d < 3; d++) {
      double* phi = basis + (size_t)(3 * c + d) * N;
      for (int q = 0; q < N; q++) phi[q] = 0.0;
      for (int nd = 0; nd < nn; nd++) {
        if (!bnd[nd]) continue;
        double X = (xyz[3 * nd] - x0) / h, Y = (xyz[3 * nd + 1] - y0) / h, Z = (xyz[3 * nd + 2] - z0) / h;
        double hat = (cx ? X : 1 - X) * (cy ? Y : 1 - Y) * (cz ? Z : 1 - Z);
        phi[3 * nd + d] = hat;
      }
      // rhs = -K_IB g
      std::vector<double> rhs(ni, 0.0), xi(ni, 0.0);
      for (int dd = 0; dd < N; dd++) {
        if (loc[dd] < 0) continue;
        double sacc = 0;
        for (int k = rp[dd]; k < rp[dd + 1]; k++)
          if (loc[ci[k]] < 0) sacc += vv[k] * phi[ci[k]];
        rhs[loc[dd]] = -sacc;
      }
      double rr;
      int it = orc_pcg(ni, irp.data(), ici.data(), iv.data(), rhs.data(), xi.data(), tol, 100000, 1, &rr);
      maxit_seen = std::max(maxit_seen, it);
      for (int dd = 0; dd < N; dd++)
        if (loc[dd] >= 0) phi[dd] = xi[loc[dd]];
    }
  }
  // Kc = Phi^T K Phi over the full cell matrix
  std::vector<double> KP((size_t)24 * N);
  for (int b2 = 0; b2 < 24; b2++) csr_mv(N, rp.data(), ci.data(), vv.data(), basis + (size_t)b2 * N, KP.data() + (size_t)b2 * N, 1);
  for (int a = 0; a < 24; a++)
    for (int b2 = 0; b2 < 24; b2++) Kc[a * 24 + b2] = dotp(N, basis + (size_t)a * N, KP.data() + (size_t)b2 * N);
  orc_sys_free(raw);
  if (iters_out) *iters_out = maxit_seen;
  return N;
}

// BCG_Stab2 (CSSD_Matrix.h:634-729) restated: unpreconditioned BiCGStab with
// the shadow residual r~ = r0; stops when the recursive ||r||/||F|| < E,
// when |beta| <= 1e-16 (the reference's "NO SOLUTION") or after maxiter
// iterations.  Returns iterations; true relative residual out.
static int orc_bcg_stab2(const Cssd& A, const std::vector<double>& F, std::vector<double>& X, double E, int maxiter,
                         double* true_relres) {
  const int N = A.n();
  std::vector<double> r(N), rt(N), p(N), s(N), q(N), h(N);
  A.mult(X, q, 1);
  for (int i = 0; i < N; i++) {
    r[i] = F[i] - q[i];
    rt[i] = r[i];
    p[i] = r[i];
  }
  const double f_norma = std::sqrt(dotp(N, F.data(), F.data()));
  int i;
  for (i = 0; i < maxiter; i++) {
    const double a1 = dotp(N, r.data(), rt.data());
    A.mult(p, q, 1);
    const double alfa = a1 / dotp(N, q.data(), rt.data());
    for (int j = 0; j < N; j++) s[j] = r[j] - alfa * q[j];
    A.mult(s, h, 1);
    const double w = dotp(N, s.data(), h.data()) / dotp(N, h.data(), h.data());
    for (int j = 0; j < N; j++) {
      X[j] += alfa * p[j] + w * s[j];
      r[j] = s[j] - w * h[j];
    }
    const double beta = dotp(N, r.data(), rt.data()) / a1 * alfa / w;
    for (int j = 0; j < N; j++) p[j] = r[j] + beta * (p[j] - w * q[j]);
    const double a_norma = std::sqrt(dotp(N, r.data(), r.data())) / f_norma;
    if (a_norma < E) break;
    if (std::fabs(beta) <= 1e-16) break;
  }
  A.mult(X, q, 1);
  double t = 0;
  for (int j = 0; j < N; j++) t += (F[j] - q[j]) * (F[j] - q[j]);
  if (true_relres) *true_relres = std::sqrt(t) / f_norma;
  return i;
}

EXPORT int orc_bcg_stab2_csr(int n, const int* rp, const int* ci, const double* v, const double* b, double* x,
                             double E, int maxiter, double* true_relres) {
  Cssd A = cssd_from_csr(n, rp, ci, v);
  std::vector<double> F(b, b + n), X(x, x + n);
  const int it = orc_bcg_stab2(A, F, X, E, maxiter, true_relres);
  std::memcpy(x, X.data(), sizeof(double) * n);
  return it;
}

// The reference's volume-problem solve (FEM.h:3699-3770 on SLAE_in_old_DOF):
// needed = max(10^-(ii+1), min_residual); SSOR-PCG (MSG_PreconditioningSSOR,
// w = 0.75) from the current X with max(N/10, 1000) iterations; on success
// ii++ and the step counter restarts, otherwise BCG_Stab2 with the same
// budget; the loop ends when the residual exceeds 100x the best, reaches
// min_residual, or after 5 steps without success.  The best X is kept,
// exactly as the reference does: when the residual reaches min_residual the
// loop breaks before that X is recorded as the best, so the X returned is
// the best one of the earlier steps.  Returns the total iterations.
static int orc_reference_continuation(int N, const int* rp, const int* ci, const double* v, const double* F, double* X,
                                      double min_residual) {
  const int max_size = std::max(N / 10, 1000);
  std::vector<double> best(X, X + N);
  double best_res = 1e25, cur = 1;
  int ii = 1, total = 0;
  Cssd A;
  bool have_cssd = false;
  for (int i = 0; i <= 4; i++) {
    double needed = std::pow(10., -1 * (ii + 1));
    needed = needed < min_residual ? min_residual : needed;
    double rr = 1, sec = 0;
    total += orc_refpath_solve(N, rp, ci, v, F, X, needed, max_size, 0, 1, &rr, &sec);
    cur = std::fabs(rr);
    if (cur < needed) {
      i = 0;
      ii++;
    } else {
      if (!have_cssd) {
        A = cssd_from_csr(N, rp, ci, v);
        have_cssd = true;
      }
      std::vector<double> Fv(F, F + N), Xv(X, X + N);
      total += orc_bcg_stab2(A, Fv, Xv, needed, max_size, &rr);
      std::memcpy(X, Xv.data(), sizeof(double) * N);
      cur = std::fabs(rr);
    }
    if (cur > best_res * 100) break;
    if (cur <= min_residual) {
      best_res = cur;
      break;
    }
    if (cur < best_res) {
      best_res = cur;
      std::memcpy(best.data(), X, sizeof(double) * N);
    }
  }
  std::memcpy(X, best.data(), sizeof(double) * N);
  return total;
}

// MsFEM local problems of one cell solved on the reference's CPU path, for
// the CPU baseline: the cell system of every basis with all fine DOFs, the
// boundary ones eliminated (identity rows, FEM.h:3382-3473), renumbered
// component-major (DOF node + a * nodes, FEM.h:3474-3500), then
// orc_reference_continuation; the Galerkin product as in orc_msfem_cell.
// Returns N; *iters_out = total iterations over the 24 problems.
EXPORT int orc_msfem_cell_refpath(int n, double x0, double y0, double z0, double h, const double* E, double nu,
                                  double min_residual, double* basis, double* Kc, int* iters_out) {
  const int s = n + 1, nn = s * s * s;
  const i64 nt = 6 * (i64)n * n * n;
  std::vector<double> xyz(3 * (size_t)nn);
  std::vector<int> tets(4 * nt);
  orc_mesh_box(n, n, n, x0, y0, z0, h / n, h / n, h / n, xyz.data(), tets.data());
  void* raw = orc_elastic_system(nn, nt, xyz.data(), tets.data(), nullptr, E, nu, 0, nullptr, nullptr, 0, nullptr,
                                 nullptr, 0, 0);
  OrcSystem* S = (OrcSystem*)raw;
  std::vector<int> rp, ci;
  std::vector<double> vv;
  orc_to_csr(S->A, rp, ci, vv);
  const int N = 3 * nn;
  std::vector<char> bnd(nn, 0);
  for (int k = 0; k < s; k++)
    for (int j = 0; j < s; j++)
      for (int i = 0; i < s; i++)
        if (i == 0 || j == 0 || k == 0 || i == n || j == n || k == n) bnd[i + s * (j + s * k)] = 1;
  // component-major permutation: new = node + a * nn
  std::vector<int> perm(N);
  for (int d = 0; d < N; d++) perm[d] = d / 3 + (d % 3) * nn;
  int total = 0;
  for (int c = 0; c < 8; c++) {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    for (int d = 0; d < 3; d++) {
      double* phi = basis + (size_t)(3 * c + d) * N;
      std::vector<double> g(N, 0.0);
      for (int nd = 0; nd < nn; nd++) {
        if (!bnd[nd]) continue;
        const double X = (xyz[3 * nd] - x0) / h, Y = (xyz[3 * nd + 1] - y0) / h, Z = (xyz[3 * nd + 2] - z0) / h;
        g[3 * nd + d] = (cx ? X : 1 - X) * (cy ? Y : 1 - Y) * (cz ? Z : 1 - Z);
      }
      // eliminated system in component-major order: boundary rows identity
      // with F = g, interior rows keep their interior columns, F -= A_IB g
      std::vector<std::vector<std::pair<int, double>>> rows(N);
      std::vector<double> F(N, 0.0), Xs(N, 0.0);
      for (int dd = 0; dd < N; dd++) {
        const int r = perm[dd];
        if (bnd[dd / 3]) {
          rows[r].push_back({r, 1.0});
          F[r] = g[dd];
          Xs[r] = g[dd];
          continue;
        }
        double f = 0;
        for (int k = rp[dd]; k < rp[dd + 1]; k++) {
          const int cc = ci[k];
          if (bnd[cc / 3]) {
            f -= vv[k] * g[cc];
            rows[r].push_back({perm[cc], 0.0});  // eliminated column (kept in the portrait)
          } else {
            rows[r].push_back({perm[cc], vv[k]});
          }
        }
        F[r] = f;
      }
      std::vector<int> prp(N + 1, 0), pci;
      std::vector<double> pv;
      for (int r = 0; r < N; r++) {
        std::sort(rows[r].begin(), rows[r].end());
        for (auto& e : rows[r]) {
          pci.push_back(e.first);
          pv.push_back(e.second);
        }
        prp[r + 1] = (int)pci.size();
      }
      total += orc_reference_continuation(N, prp.data(), pci.data(), pv.data(), F.data(), Xs.data(), min_residual);
      for (int dd = 0; dd < N; dd++) phi[dd] = Xs[perm[dd]];
    }
  }
  std::vector<double> KP((size_t)24 * N);
  for (int b2 = 0; b2 < 24; b2++)
    csr_mv(N, rp.data(), ci.data(), vv.data(), basis + (size_t)b2 * N, KP.data() + (size_t)b2 * N, 1);
  for (int a = 0; a < 24; a++)
    for (int b2 = 0; b2 < 24; b2++) Kc[a * 24 + b2] = dotp(N, basis + (size_t)a * N, KP.data() + (size_t)b2 * N);
  orc_sys_free(raw);
  if (iters_out) *iters_out = total;
  return N;
}

// ---------------------------------------------------------------------------
// MsFEM reference-trace mode: the edge -> face -> volume hierarchy of
// Grid_forMech_Poly_Order1::CreateFunctionalSpaces (MsFEM_Grid.h:5026-5506).
// ---------------------------------------------------------------------------
// Dense LU with partial pivoting, nrhs right-hand sides (columns of B,
// row-major n x nrhs), overwritten by the solutions.
static int gepp_multi(int n, std::vector<double>& A, int nrhs, std::vector<double>& B) {
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)p * n + c])) p = r;
    if (A[(size_t)p * n + c] == 0) return -1;
    if (p != c) {
      for (int k = 0; k < n; k++) std::swap(A[(size_t)p * n + k], A[(size_t)c * n + k]);
      for (int k = 0; k < nrhs; k++) std::swap(B[(size_t)p * nrhs + k], B[(size_t)c * nrhs + k]);
    }
    for (int r = c + 1; r < n; r++) {
      double f = A[(size_t)r * n + c] / A[(size_t)c * n + c];
      if (f == 0) continue;
      for (int k = c; k < n; k++) A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
      for (int k = 0; k < nrhs; k++) B[(size_t)r * nrhs + k] -= f * B[(size_t)c * nrhs + k];
    }
  }
  for (int i = n - 1; i >= 0; i--)
    for (int k = 0; k < nrhs; k++) {
      double s = B[(size_t)i * nrhs + k];
      for (int j = i + 1; j < n; j++) s -= A[(size_t)i * n + j] * B[(size_t)j * nrhs + k];
      B[(size_t)i * nrhs + k] = s / A[(size_t)i * n + i];
    }
  return 0;
}

// Block (I,J) of B_J^T D B_I (convention 0) or B_I^T D B_J (convention 1)
// for gradients gI, gJ, times the weight w: the integrand of every
// SolveLocalMatrix (FEM_Grid.h:3649-3735 segments, :4429-4500 triangles,
// :2125-2191 tets) with the basis derivative x = y = z = grad phi.
static void orc_block(const double gI[3], const double gJ[3], const double D[6][6], int convention, double w,
                      double R[3][3]) {
  double BI[6][3], BJ[6][3];
  orc_B(gI, BI);
  orc_B(gJ, BJ);
  const double(*Bl)[3] = convention == 0 ? BJ : BI;
  const double(*Br)[3] = convention == 0 ? BI : BJ;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      double s = 0;
      for (int k = 0; k < 6; k++)
        for (int l = 0; l < 6; l++) s += Bl[k][r] * D[k][l] * Br[l][c];
      R[r][c] = s * w;
    }
}

// Segment element (FiniteElement_1D_forMech, FEM_Grid.h:3649-3735 with the
// basis of :4011-4101): local frame with X along node0 -> node1, hat
// derivative -+1/len along X mapped back by the reverse basis, 4-point
// rule whose weights add up to the length.  K is 6x6, index 3*I+a.
EXPORT double orc_elem_seg(const double* P, double E, double nu, int convention, double* K) {
  double t[3], L = 0;
  for (int a = 0; a < 3; a++) {
    t[a] = P[3 + a] - P[a];
    L += t[a] * t[a];
  }
  L = std::sqrt(L);
  double g[2][3];
  for (int a = 0; a < 3; a++) {
    g[0][a] = -t[a] / L / L;
    g[1][a] = t[a] / L / L;
  }
  double D[6][6];
  orc_isotropic_D(E, nu, D);
  for (int I = 0; I < 2; I++)
    for (int J = 0; J < 2; J++) {
      double R[3][3];
      orc_block(g[I], g[J], D, convention, L, R);
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) K[(3 * I + r) * 6 + 3 * J + c] = R[r][c];
    }
  return L;
}

// Triangle element (FiniteElement_2D_forMech, FEM_Grid.h:4429-4500 with the
// basis of :4879-4950): orthonormal local frame (X along node0 -> node1, Y
// in the plane, GeometryShape.h:979-1060), phi_i = alpha_i0 + alpha_i1 x +
// alpha_i2 y from the inverse of [1 x y] in local coordinates, gradient
// alpha_i1 X + alpha_i2 Y; area weight (GetVolume, GeometryShape.h:1337-1364).
// K is 9x9, index 3*I+a.
EXPORT double orc_elem_tri(const double* P, double E, double nu, int convention, double* K) {
  double e1[3], e2[3], N[3];
  for (int a = 0; a < 3; a++) {
    e1[a] = P[3 + a] - P[a];
    e2[a] = P[6 + a] - P[a];
  }
  N[0] = e1[1] * e2[2] - e1[2] * e2[1];
  N[1] = e1[2] * e2[0] - e1[0] * e2[2];
  N[2] = e1[0] * e2[1] - e1[1] * e2[0];
  const double nN = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
  const double l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
  double X[3], Y[3];
  for (int a = 0; a < 3; a++) X[a] = e1[a] / l1;
  // Y = (N/|N|) x X
  Y[0] = (N[1] * X[2] - N[2] * X[1]) / nN;
  Y[1] = (N[2] * X[0] - N[0] * X[2]) / nN;
  Y[2] = (N[0] * X[1] - N[1] * X[0]) / nN;
  double loc[3][2];
  for (int i = 0; i < 3; i++) {
    double d[3];
    for (int a = 0; a < 3; a++) d[a] = P[3 * i + a] - P[a];
    loc[i][0] = d[0] * X[0] + d[1] * X[1] + d[2] * X[2];
    loc[i][1] = d[0] * Y[0] + d[1] * Y[1] + d[2] * Y[2];
  }
  // alpha = inverse of M = [1 x_i y_i] (rows i); phi_i coefficients are column i
  double M[3][3] = {{1, loc[0][0], loc[0][1]}, {1, loc[1][0], loc[1][1]}, {1, loc[2][0], loc[2][1]}};
  double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
               M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  double inv[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
      inv[r][c] = (M[r1][c1] * M[r2][c2] - M[r1][c2] * M[r2][c1]) / det;
    }
  double g[3][3];
  for (int i = 0; i < 3; i++)
    for (int a = 0; a < 3; a++) g[i][a] = inv[1][i] * X[a] + inv[2][i] * Y[a];
  const double area = 0.5 * nN;
  double D[6][6];
  orc_isotropic_D(E, nu, D);
  for (int I = 0; I < 3; I++)
    for (int J = 0; J < 3; J++) {
      double R[3][3];
      orc_block(g[I], g[J], D, convention, area, R);
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) K[(3 * I + r) * 9 + 3 * J + c] = R[r][c];
    }
  return area;
}

// Solve a small elastic system with 3 DOFs per node given element lists;
// nodes with known[] set take the values in U (nrhs per DOF), the rest are
// solved (Dirichlet rows replaced and moved to the right side as in
// FEM.h:4126-4200).  U is [node][3][nrhs].
static int orc_small_elastic(int nn, int ne, int npe, const int* elem, const double* Ke /*ne x (3npe)^2*/,
                             const char* known, int nrhs, double* U) {
  std::vector<int> loc(3 * nn, -1);
  int nu_ = 0;
  for (int v = 0; v < nn; v++)
    if (!known[v])
      for (int a = 0; a < 3; a++) loc[3 * v + a] = nu_++;
  std::vector<double> A((size_t)nu_ * nu_, 0.0), B((size_t)nu_ * nrhs, 0.0);
  const int m = 3 * npe;
  for (int e = 0; e < ne; e++)
    for (int I = 0; I < npe; I++)
      for (int a = 0; a < 3; a++) {
        const int row = loc[3 * elem[e * npe + I] + a];
        if (row < 0) continue;
        for (int J = 0; J < npe; J++)
          for (int b = 0; b < 3; b++) {
            const double k = Ke[(size_t)e * m * m + (3 * I + a) * m + 3 * J + b];
            const int cj = 3 * elem[e * npe + J] + b;
            if (loc[cj] >= 0)
              A[(size_t)row * nu_ + loc[cj]] += k;
            else
              for (int r = 0; r < nrhs; r++) B[(size_t)row * nrhs + r] -= k * U[(size_t)cj * nrhs + r];
          }
      }
  if (nu_ > 0 && gepp_multi(nu_, A, nrhs, B) != 0) return -1;
  for (int d = 0; d < 3 * nn; d++)
    if (loc[d] >= 0)
      for (int r = 0; r < nrhs; r++) U[(size_t)d * nrhs + r] = B[(size_t)loc[d] * nrhs + r];
  return 0;
}

// 1-D elastic chain (FEM_1D_forElasticDeformation, FEM.h:3796-4365):
// nodes xyz[(nseg+1)*3] in order, segment moduli E[nseg], Dirichlet at both
// ends with u0[3], un[3].  Writes U[(nseg+1)*3].
EXPORT int orc_edge_1d(int nseg, const double* xyz, const double* E, double nu, int convention, const double* u0,
                       const double* un, double* U) {
  const int nn = nseg + 1;
  std::vector<int> el(2 * nseg);
  std::vector<double> Ke((size_t)nseg * 36);
  for (int s = 0; s < nseg; s++) {
    el[2 * s] = s;
    el[2 * s + 1] = s + 1;
    orc_elem_seg(xyz + 3 * s, E[s], nu, convention, Ke.data() + (size_t)s * 36);
  }
  std::vector<char> known(nn, 0);
  known[0] = known[nn - 1] = 1;
  for (int d = 0; d < 3 * nn; d++) U[d] = 0.0;
  for (int a = 0; a < 3; a++) {
    U[a] = u0[a];
    U[3 * (nn - 1) + a] = un[a];
  }
  return orc_small_elastic(nn, nseg, 2, el.data(), Ke.data(), known.data(), 1, U);
}

// Local bases of one structured cell in the reference-trace mode.
//   edges: per coarse edge a 1-D solve with the coarse-node value at one end
//     and 0 at the other (MsFEM_Grid.h:5076-5175); segment modulus = that of
//     the boundary triangle holding the segment in the first face (x, y, z
//     faces in that order) containing the edge (:4880-4920 takes the
//     triangle's material from the first neighbouring face);
//   faces: per coarse face and face corner a 2-D solve with the two edge
//     solutions that are 1 at the corner on the edges meeting there and 0 on
//     the other two (:5213-5325); triangle modulus = that of the tet owning
//     the triangle;
//   volume: per coarse node a 3-D solve with the face solutions of the three
//     faces meeting at the node and 0 on the others (:5421-5506).
// The reference drives all three components together with (1,1,1); here
// basis 3c+d is driven with e_d, so that the reference's basis of node c is
// the sum over d (every step is linear in its boundary data).
// form 0: Kc = Phi^T K Phi of these bases (K in the reference convention);
// form 1: the component-wise coarse functions Psi_{c,a} = e_a (sum_d
// phi_{c,d})_a, the reference's use of a vector basis per coarse DOF, with
// the reference's coarse element block(I,J) = int B_J^T D B_I of their
// gradients (MsFEM_Grid.h:3837-3975) integrated exactly over the fine tets:
// Kc = blockwise transpose of Psi^T K_textbook Psi.  Returns N.
EXPORT int orc_msfem_cell_reftrace(int n, double x0, double y0, double z0, double h, const double* E, double nu,
                                   double tol, int form, double* basis, double* Kc, int* iters_out) {
  const int s = n + 1;
  const int nn = s * s * s;
  const i64 nt = 6 * (i64)n * n * n;
  std::vector<double> xyz(3 * (size_t)nn);
  std::vector<int> tets(4 * nt);
  orc_mesh_box(n, n, n, x0, y0, z0, h / n, h / n, h / n, xyz.data(), tets.data());
  auto ijk = [&](int v, int a) { return a == 0 ? v % s : a == 1 ? (v / s) % s : v / (s * s); };
  auto on_face = [&](int v, int f) { return ijk(v, f / 2) == (f % 2 ? n : 0); };
  // boundary triangles of each face, with the modulus of the owning tet
  struct Tri {
    int v[3];
    double E;
  };
  std::vector<Tri> ftri[6];
  for (i64 t = 0; t < nt; t++)
    for (int omit = 0; omit < 4; omit++) {
      int v[3], q = 0;
      for (int k = 0; k < 4; k++)
        if (k != omit) v[q++] = tets[4 * t + k];
      for (int f = 0; f < 6; f++)
        if (on_face(v[0], f) && on_face(v[1], f) && on_face(v[2], f)) ftri[f].push_back({{v[0], v[1], v[2]}, E[t]});
    }
  std::vector<double> trace((size_t)3 * nn * 24, 0.0);  // [node][a][basis]
  std::vector<char> traced(nn, 0);
  // ---- edges: direction ax, other axes at sides; corners c0 (bit ax = 0), c1
  std::vector<std::vector<double>> E_sol(12);  // [end][node][3][3 drives]
  for (int e = 0; e < 12; e++) {
    const int ax = e / 4, o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
    const int b1 = e & 1, b2 = (e >> 1) & 1;
    std::vector<int> nodes(s);
    for (int t = 0; t < s; t++) {
      int c[3];
      c[ax] = t;
      c[o1] = b1 * n;
      c[o2] = b2 * n;
      nodes[t] = c[0] + s * (c[1] + s * c[2]);
    }
    // segment moduli from the first face containing the edge
    const int f0 = 2 * o1 + b1;
    std::vector<double> Es(n), pts(3 * s);
    for (int t = 0; t < n; t++) {
      bool found = false;
      for (const Tri& T : ftri[f0]) {
        int hit = 0;
        for (int k = 0; k < 3; k++) hit += (T.v[k] == nodes[t]) + (T.v[k] == nodes[t + 1]);
        if (hit == 2) {
          Es[t] = T.E;
          found = true;
          break;
        }
      }
      if (!found) return -1;
    }
    for (int t = 0; t < s; t++)
      for (int a = 0; a < 3; a++) pts[3 * t + a] = xyz[3 * nodes[t] + a];
    E_sol[e].assign((size_t)2 * s * 9, 0.0);
    for (int end = 0; end < 2; end++)
      for (int d = 0; d < 3; d++) {
        double u0[3] = {0, 0, 0}, un[3] = {0, 0, 0};
        (end == 0 ? u0 : un)[d] = 1.0;
        std::vector<double> U(3 * s);
        if (orc_edge_1d(n, pts.data(), Es.data(), nu, 0, u0, un, U.data()) != 0) return -1;
        for (int t = 0; t < s; t++)
          for (int a = 0; a < 3; a++) E_sol[e][((size_t)(end * s + t) * 3 + a) * 3 + d] = U[3 * t + a];
      }
  }
  // ---- faces
  std::vector<double> F_sol[6];  // [node][3][corner 8][3 drives], corners of the face only
  for (int f = 0; f < 6; f++) {
    const int ax = f / 2, side = f % 2;
    std::vector<int> fnodes;
    std::vector<int> floc(nn, -1);
    for (int v = 0; v < nn; v++)
      if (on_face(v, f)) {
        floc[v] = (int)fnodes.size();
        fnodes.push_back(v);
      }
    const int fn = (int)fnodes.size();
    std::vector<int> el;
    std::vector<double> Ke;
    for (const Tri& T : ftri[f]) {
      double P[9];
      for (int k = 0; k < 3; k++) {
        el.push_back(floc[T.v[k]]);
        for (int a = 0; a < 3; a++) P[3 * k + a] = xyz[3 * T.v[k] + a];
      }
      Ke.resize(Ke.size() + 81);
      orc_elem_tri(P, T.E, nu, 0, Ke.data() + Ke.size() - 81);
    }
    // Dirichlet data on the face boundary from the edge solutions: 8 corners x 3 drives
    std::vector<char> known(fn, 0);
    std::vector<double> U((size_t)fn * 3 * 24, 0.0);
    for (int e = 0; e < 12; e++) {
      const int eax = e / 4;
      if (eax == ax) continue;
      const int o1 = eax == 0 ? 1 : 0, o2 = eax == 2 ? 1 : 2;
      const int b1 = e & 1, b2 = (e >> 1) & 1;
      // the edge lies in face f when its fixed coordinate on axis ax equals the face side
      const int bax = (o1 == ax) ? b1 : b2;
      if (bax != side) continue;
      for (int t = 0; t < s; t++) {
        int c[3];
        c[eax] = t;
        c[o1] = b1 * n;
        c[o2] = b2 * n;
        const int v = c[0] + s * (c[1] + s * c[2]);
        const int fl = floc[v];
        known[fl] = 1;
        for (int end = 0; end < 2; end++) {
          int cc[3];
          cc[eax] = end;
          cc[o1] = b1;
          cc[o2] = b2;
          const int corner = cc[0] | (cc[1] << 1) | (cc[2] << 2);
          for (int a = 0; a < 3; a++)
            for (int d = 0; d < 3; d++)
              U[((size_t)fl * 3 + a) * 24 + 3 * corner + d] = E_sol[e][((size_t)(end * s + t) * 3 + a) * 3 + d];
        }
      }
    }
    if (orc_small_elastic(fn, (int)ftri[f].size(), 3, el.data(), Ke.data(), known.data(), 24, U.data()) != 0)
      return -1;
    // only corners of this face carry a trace on it
    for (int fl = 0; fl < fn; fl++) {
      const int v = fnodes[fl];
      traced[v] = 1;
      for (int corner = 0; corner < 8; corner++) {
        if (((corner >> ax) & 1) != side) continue;
        for (int a = 0; a < 3; a++)
          for (int d = 0; d < 3; d++)
            trace[((size_t)v * 3 + a) * 24 + 3 * corner + d] = U[((size_t)fl * 3 + a) * 24 + 3 * corner + d];
      }
    }
  }
  // ---- volume: reference-convention cell matrix, interior solve per basis
  void* raw = orc_elastic_system(nn, nt, xyz.data(), tets.data(), nullptr, E, nu, 0, nullptr, nullptr, 0, nullptr,
                                 nullptr, 0, 0);
  OrcSystem* S = (OrcSystem*)raw;
  std::vector<int> rp, ci;
  std::vector<double> vv;
  orc_to_csr(S->A, rp, ci, vv);
  const int N = 3 * nn;
  std::vector<int> loc(N, -1);
  int ni = 0;
  for (int d = 0; d < N; d++)
    if (!traced[d / 3]) loc[d] = ni++;
  std::vector<int> irp(ni + 1, 0), ici;
  std::vector<double> iv;
  for (int d = 0; d < N; d++) {
    if (loc[d] < 0) continue;
    for (int k = rp[d]; k < rp[d + 1]; k++)
      if (loc[ci[k]] >= 0) {
        ici.push_back(loc[ci[k]]);
        iv.push_back(vv[k]);
      }
    irp[loc[d] + 1] = (int)ici.size();
  }
  int maxit_seen = 0;
  for (int q = 0; q < 24; q++) {
    double* phi = basis + (size_t)q * N;
    for (int d = 0; d < N; d++) phi[d] = loc[d] < 0 ? trace[(size_t)d * 24 + q] : 0.0;
    std::vector<double> rhs(ni, 0.0), xi(ni, 0.0);
    for (int dd = 0; dd < N; dd++) {
      if (loc[dd] < 0) continue;
      double sacc = 0;
      for (int k = rp[dd]; k < rp[dd + 1]; k++)
        if (loc[ci[k]] < 0) sacc += vv[k] * phi[ci[k]];
      rhs[loc[dd]] = -sacc;
    }
    double rr;
    int it = orc_pcg(ni, irp.data(), ici.data(), iv.data(), rhs.data(), xi.data(), tol, 100000, 1, &rr);
    maxit_seen = std::max(maxit_seen, it);
    for (int dd = 0; dd < N; dd++)
      if (loc[dd] >= 0) phi[dd] = xi[loc[dd]];
  }
  if (form == 1) {
    // Psi_{c,a}[node][b] = delta_ab * sum_d phi_{c,d}[node][a]
    std::vector<double> tmp((size_t)24 * N);
    for (int c = 0; c < 8; c++)
      for (int v = 0; v < nn; v++)
        for (int a = 0; a < 3; a++) {
          double sum = 0;
          for (int d = 0; d < 3; d++) sum += basis[(size_t)(3 * c + d) * N + 3 * v + a];
          for (int a2 = 0; a2 < 3; a2++) tmp[(size_t)(3 * c + a2) * N + 3 * v + a] = a2 == a ? sum : 0.0;
        }
    std::copy(tmp.begin(), tmp.end(), basis);
    void* raw_tb = orc_elastic_system(nn, nt, xyz.data(), tets.data(), nullptr, E, nu, 0, nullptr, nullptr, 0,
                                      nullptr, nullptr, 1, 0);
    orc_to_csr(((OrcSystem*)raw_tb)->A, rp, ci, vv);
    orc_sys_free(raw_tb);
  }
  std::vector<double> KP((size_t)24 * N);
  for (int b2 = 0; b2 < 24; b2++)
    csr_mv(N, rp.data(), ci.data(), vv.data(), basis + (size_t)b2 * N, KP.data() + (size_t)b2 * N, 1);
  for (int a = 0; a < 24; a++)
    for (int b2 = 0; b2 < 24; b2++) {
      const int row = form == 1 ? 3 * (a / 3) + b2 % 3 : a, col = form == 1 ? 3 * (b2 / 3) + a % 3 : b2;
      Kc[row * 24 + col] = dotp(N, basis + (size_t)a * N, KP.data() + (size_t)b2 * N);
    }
  orc_sys_free(raw);
  if (iters_out) *iters_out = maxit_seen;
  return N;
}

// ---------------------------------------------------------------------------
// MLP oracle (FP64): layers L, sizes dims[0..L], ReLU between layers, linear
// output, loss = mean over batch of 0.5*||y - t||^2 / out_dim.  Weights W_l are
// row-major [out][in], biases b_l.  Computes loss and gradients (backprop).
// No reference counterpart (the learned initial guess is build-only).
// ---------------------------------------------------------------------------
EXPORT double orc_mlp_loss_grad(int L, const int* dims, const double* const* W, const double* const* bvec, int B,
                                const double* X, const double* T, double* const* gW, double* const* gb) {
  std::vector<std::vector<double>> act(L + 1), pre(L + 1);
  act[0].assign(X, X + (size_t)B * dims[0]);
  for (int l = 0; l < L; l++) {
    int din = dims[l], dout = dims[l + 1];
    pre[l + 1].assign((size_t)B * dout, 0.0);
    act[l + 1].assign((size_t)B * dout, 0.0);
    for (int s = 0; s < B; s++)
      for (int o = 0; o < dout; o++) {
        double acc = bvec[l][o];
        for (int i = 0; i < din; i++) acc += W[l][(size_t)o * din + i] * act[l][(size_t)s * din + i];
        pre[l + 1][(size_t)s * dout + o] = acc;
        act[l + 1][(size_t)s * dout + o] = (l + 1 < L) ? (acc > 0 ? acc : 0.0) : acc;
      }
  }
  int dl = dims[L];
  double loss = 0;
  std::vector<double> delta((size_t)B * dl);
  for (size_t q = 0; q < (size_t)B * dl; q++) {
    double e = act[L][q] - T[q];
    loss += 0.5 * e * e;
    delta[q] = e / ((double)B * dl);
  }
  loss /= (double)B * dl;
  if (!gW) return loss;
  for (int l = L - 1; l >= 0; l--) {
    int din = dims[l], dout = dims[l + 1];
    for (int o = 0; o < dout; o++) {
      double sb = 0;
      for (int s = 0; s < B; s++) sb += delta[(size_t)s * dout + o];
      gb[l][o] = sb;
      for (int i = 0; i < din; i++) {
        double sw = 0;
        for (int s = 0; s < B; s++) sw += delta[(size_t)s * dout + o] * act[l][(size_t)s * din + i];
        gW[l][(size_t)o * din + i] = sw;
      }
    }
    if (l == 0) break;
    std::vector<double> nd((size_t)B * din, 0.0);
    for (int s = 0; s < B; s++)
      for (int i = 0; i < din; i++) {
        double acc = 0;
        for (int o = 0; o < dout; o++) acc += W[l][(size_t)o * din + i] * delta[(size_t)s * dout + o];
        nd[(size_t)s * din + i] = pre[l][(size_t)s * din + i] > 0 ? acc : 0.0;
      }
    delta.swap(nd);
  }
  return loss;
}
