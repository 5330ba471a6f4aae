// Exercises include/mfs_cssd.hpp the way the reference's drivers use
// CSSD_Matrix: portrait from element cliques, SummPartOfMatrix of local
// matrices, direct writes to Diag/A_up/A_down/F for boundary conditions,
// the host utilities (GetValue/SetValue, SummMatrix, MakeDivideOnSqrtFormSumm,
// the SSOR factor builders, RandomMatrix, SolveMaxEigenvalue,
// Minimization_Problem, ...), then the solver methods, for both
// instantiations the drivers use: <double, double> and the 3x3 block type.
// The value types are minimal local stand-ins with the members and operators
// of the reference's Tensor2Rank3D / Point<double> that the class needs.
//
// Usage: test_cssd_dropin cpu                 host methods only (no GPU)
//        test_cssd_dropin gpu                 + scalar solvers on a 1-D chain
//        test_cssd_dropin bsr3 IN OUT         block system from IN (written by
//                                             tests/test_dropin.py), solved by
//                                             every solver method; solutions
//                                             written to OUT
// Prints "OK" on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mfs_cssd.hpp"

struct Pt {  // Point<double>-like
  double x = 0, y = 0, z = 0;
  Pt() {}
  Pt(double a, double b, double c) : x(a), y(b), z(c) {}
  Pt& operator=(double v) {
    x = y = z = v;
    return *this;
  }
  Pt& operator+=(const Pt& o) {
    x += o.x;
    y += o.y;
    z += o.z;
    return *this;
  }
};

struct T3 {  // Tensor2Rank3D-like
  double val[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  T3& operator=(double v) {
    for (auto& r : val)
      for (double& e : r) e = v;
    return *this;
  }
  T3& operator+=(const T3& o) {
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) val[a][b] += o.val[a][b];
    return *this;
  }
  Pt operator*(const Pt& p) {
    return Pt(val[0][0] * p.x + val[0][1] * p.y + val[0][2] * p.z, val[1][0] * p.x + val[1][1] * p.y + val[1][2] * p.z,
              val[2][0] * p.x + val[2][1] * p.y + val[2][2] * p.z);
  }
};

struct Dense {  // DenseMatrix<double,double>-like local matrix
  int n;
  std::vector<double> A, Fv, Xv;
  explicit Dense(int n_) : n(n_), A(n_ * n_, 0.0), Fv(n_, 0.0), Xv(n_, 0.0) {}
  int GetSize() { return n; }
  double GetElementOfA(int i, int j) { return A[i * n + j]; }
  double GetElementOfF(int i) { return Fv[i]; }
  double GetElementOfX(int i) { return Xv[i]; }
};

struct DenseB {  // DenseMatrix<Tensor2Rank3D, Point<double>>-like local matrix
  int n;
  std::vector<T3> A;
  std::vector<Pt> Fv, Xv;
  explicit DenseB(int n_) : n(n_), A(n_ * n_), Fv(n_), Xv(n_) {}
  int GetSize() { return n; }
  T3 GetElementOfA(int i, int j) { return A[i * n + j]; }
  Pt GetElementOfF(int i) { return Fv[i]; }
  Pt GetElementOfX(int i) { return Xv[i]; }
};

static int fail(const char* msg) {
  std::printf("FAIL: %s\n", msg);
  return 1;
}

static bool near(double a, double b, double tol) { return std::fabs(a - b) <= tol * std::max(1.0, std::fabs(b)); }

// 1-D chain with Dirichlet ends (scalar), assembled like the drivers do.
static void chain(CSSD_Matrix<double, double>& A, int n) {
  std::vector<std::vector<int>> up(n), down(n);
  for (int e = 0; e + 1 < n; e++) {
    up[e].push_back(e + 1);
    down[e + 1].push_back(e);
  }
  A.Initialization(up, down);
  for (int e = 0; e + 1 < n; e++) {
    Dense L(2);
    const double k = 1.0 + (e % 7);
    L.A = {k, -k, -k, k};
    std::vector<int> ids = {e, e + 1};
    A.SummPartOfMatrix(L, ids);
  }
}

static int host_scalar() {
  const int n = 200;
  CSSD_Matrix<double, double> A;
  chain(A, n);
  if (A.Diag[0] != 1.0 || A.Diag[1] != 3.0 || A.A_up[1][0] != -2.0 || A.A_down[1][0] != -1.0)
    return fail("SummPartOfMatrix values");
  std::vector<double> one(n, 1.0), y;
  A.MultiplicationMatrixVector(one, y);
  for (double v : y)
    if (std::fabs(v) > 1e-14) return fail("A*1 != 0");
  // GetValue / SetValue (CSSD_Matrix.h:180-241): present, absent, out of range
  double v = 0;
  if (!A.GetValue(5, 6, v) || v != A.A_up[5][0]) return fail("GetValue upper");
  if (!A.GetValue(6, 5, v) || v != A.A_down[6][0]) return fail("GetValue lower");
  if (!A.GetValue(7, 7, v) || v != A.Diag[7]) return fail("GetValue diagonal");
  if (A.GetValue(5, 9, v) || A.GetValue(n, 0, v) || A.GetValue(0, n + 3, v)) return fail("GetValue absent");
  double w = 42.0;
  if (!A.SetValue(8, 9, w) || A.A_up[8][0] != 42.0) return fail("SetValue");
  if (A.SetValue(8, 12, w) || A.SetValue(n, 1, w)) return fail("SetValue absent");
  w = -7.0;
  A.SetValue(8, 9, w);
  // SummMatrix: this = 2A - 3A = -A
  CSSD_Matrix<double, double> S;
  S.SetMatrix(A);
  S.SummMatrix(A, 2.0, A, -3.0);
  if (S.Diag[10] != -A.Diag[10] || S.A_up[10][0] != -A.A_up[10][0] || S.A_down[10][0] != -A.A_down[10][0])
    return fail("SummMatrix");
  // MakeDivideOnSqrtFormSumm: every row gets unit 2-norm, F scaled alike
  S.SetMatrix(A);
  for (int i = 0; i < n; i++) S.F[i] = 1.0 + i;
  S.MakeDivideOnSqrtFormSumm();
  for (int i = 0; i < n; i++) {
    double s = S.Diag[i] * S.Diag[i];
    for (double a : S.A_up[i]) s += a * a;
    for (double a : S.A_down[i]) s += a * a;
    double r = A.Diag[i] * A.Diag[i];
    for (double a : A.A_up[i]) r += a * a;
    for (double a : A.A_down[i]) r += a * a;
    if (!near(s, 1.0, 1e-14) || !near(S.F[i], (1.0 + i) / std::sqrt(r), 1e-14)) return fail("MakeDivideOnSqrtFormSumm");
  }
  // ClearVariables keeps the portrait, zeroes the values
  S.ClearVariables();
  if (S.Diag[3] != 0.0 || S.A_up[3][0] != 0.0 || S.F[3] != 0.0 || S.id_column_for_A_up[3][0] != 4)
    return fail("ClearVariables");

  // SSOR factor (CSSD_Matrix.h:458-547) against the dense formula
  // M^-1 = w(2-w) (D - wL)^-1 D (D - wU)^-1 on a small non-banded SPD matrix
  {
    const int m = 9;
    std::vector<std::vector<int>> up(m), down(m);
    auto link = [&](int i, int j) {
      up[i].push_back(j);
      down[j].push_back(i);
    };
    for (int i = 0; i + 1 < m; i++) link(i, i + 1);
    for (int i = 0; i + 3 < m; i++) link(i, i + 3);
    for (auto& u : up) std::sort(u.begin(), u.end());
    for (auto& d : down) std::sort(d.begin(), d.end());
    CSSD_Matrix<double, double> M;
    M.Initialization(up, down);
    std::vector<std::vector<double>> Dn(m, std::vector<double>(m, 0.0));
    for (int i = 0; i < m; i++) {
      M.Diag[i] = 6.0 + 0.5 * i;
      Dn[i][i] = M.Diag[i];
      for (size_t k = 0; k < up[i].size(); k++) {
        const int j = up[i][k];
        const double a = -0.3 - 0.1 * ((i + j) % 4);
        M.A_up[i][k] = a;
        Dn[i][j] = a;
        Dn[j][i] = a;
        for (size_t q = 0; q < down[j].size(); q++)
          if (down[j][q] == i) M.A_down[j][q] = a;
      }
    }
    const double om = 0.75;
    CSSD_Matrix<double, double> P;
    P.PrecondorSSOR(om, M);
    std::vector<double> z(m), x, fd(M.Diag);
    for (int i = 0; i < m; i++) z[i] = std::sin(1.0 + i);
    P.SolvePrecondorSSOR_SLAE(fd, x, z);
    // dense: t = (D - wU)^-1 z (backward), t *= w(2-w) D, x = (D - wL)^-1 t
    std::vector<double> t(m);
    for (int i = m - 1; i >= 0; i--) {
      double s = z[i];
      for (int j = i + 1; j < m; j++) s += om * Dn[i][j] * t[j];
      t[i] = s / Dn[i][i];
    }
    for (int i = 0; i < m; i++) t[i] *= om * (2 - om) * Dn[i][i];
    std::vector<double> xd(m);
    for (int i = 0; i < m; i++) {
      double s = t[i];
      for (int j = 0; j < i; j++) s += om * Dn[i][j] * xd[j];
      xd[i] = s / Dn[i][i];
    }
    for (int i = 0; i < m; i++)
      if (!near(x[i], xd[i], 1e-13)) return fail("PrecondorSSOR / SolvePrecondorSSOR_SLAE");
    // symmetric factor on a tridiagonal matrix (every upper list is the
    // run {i+1}): diag_i = a_ii + w^2 a_{i,i+1}^2 / a_{i+1,i+1}, up = w a
    CSSD_Matrix<double, double> T, Q;
    chain(T, 12);
    for (int i = 0; i < 12; i++) T.Diag[i] += 1.0;
    Q.PrecondorSSOR_summetric(om, T);
    for (int i = 0; i + 1 < 12; i++) {
      const double a = T.A_up[i][0];
      if (!near(Q.Diag[i], T.Diag[i] + om * om * a * a / T.Diag[i + 1], 1e-14) || !near(Q.A_up[i][0], om * a, 1e-14) ||
          !near(Q.A_down[i + 1][0], om * a, 1e-14))
        return fail("PrecondorSSOR_summetric");
    }
  }
  // RandomMatrix: symmetric portrait, F = A X
  {
    std::srand(5);
    CSSD_Matrix<double, double> R;
    R.RandomMatrix(300, 2.0);
    std::vector<double> ax;
    R.MultiplicationMatrixVector(R.X, ax);
    for (int i = 0; i < 300; i++) {
      if (!near(ax[i], R.F[i], 1e-14)) return fail("RandomMatrix F");
      for (size_t k = 0; k < R.id_column_for_A_up[i].size(); k++) {
        double a = 0;
        if (!R.GetValue(R.id_column_for_A_up[i][k], i, a) || a != R.A_up[i][k]) return fail("RandomMatrix symmetry");
      }
    }
    // SolveMaxEigenvalue: dominant eigenvalue 10 of diag(1, .., 1, 10) + small coupling
    CSSD_Matrix<double, double> E;
    chain(E, 50);
    for (int i = 0; i < 50; i++) {
      E.Diag[i] = 1.0;
      for (auto& a : E.A_up[i]) a = 0.0;
      for (auto& a : E.A_down[i]) a = 0.0;
    }
    E.Diag[20] = 10.0;
    const double lam = E.SolveMaxEigenvalue();
    if (!near(lam, 10.0, 2e-3)) return fail("SolveMaxEigenvalue");
  }
  // Minimization_Problem: Givens least squares vs the normal equations
  {
    const int m = 5;
    std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0)), H0;
    for (int i = 0; i <= m; i++)
      for (int j = 0; j < m; j++)
        if (i <= j + 1) H[i][j] = std::cos(1.0 + 3 * i + 7 * j) + (i == j ? 2.0 : 0.0);
    H0 = H;
    std::vector<double> f;
    const double beta = 1.7;
    const double res = A.Minimization_Problem(m, beta, f, H);
    // normal equations H0^T H0 y = H0^T (beta e1)
    std::vector<std::vector<double>> N(m, std::vector<double>(m + 1, 0.0));
    for (int a = 0; a < m; a++) {
      for (int b = 0; b < m; b++)
        for (int i = 0; i <= m; i++) N[a][b] += H0[i][a] * H0[i][b];
      N[a][m] = H0[0][a] * beta;
    }
    for (int c = 0; c < m; c++)
      for (int r = c + 1; r < m; r++) {
        const double fct = N[r][c] / N[c][c];
        for (int k = c; k <= m; k++) N[r][k] -= fct * N[c][k];
      }
    std::vector<double> yv(m);
    for (int r = m - 1; r >= 0; r--) {
      double s = N[r][m];
      for (int k = r + 1; k < m; k++) s -= N[r][k] * yv[k];
      yv[r] = s / N[r][r];
    }
    double rr = 0;
    for (int i = 0; i <= m; i++) {
      double s = (i == 0 ? beta : 0.0);
      for (int j = 0; j < m; j++) s -= H0[i][j] * yv[j];
      rr += s * s;
    }
    for (int j = 0; j < m; j++)
      if (!near(f[j], yv[j], 1e-10)) return fail("Minimization_Problem solution");
    if (!near(res, std::sqrt(rr), 1e-10)) return fail("Minimization_Problem residual");
  }
  // Start_Diag_Solver: X = F / diag
  {
    std::vector<double> f = {2.0, 9.0}, x(2), d = {4.0, 3.0};
    A.Start_Diag_Solver(f, x, d);
    if (x[0] != 0.5 || x[1] != 3.0) return fail("Start_Diag_Solver");
  }
  return 0;
}

static int host_block() {
  // two-node block chain: local 2x2 block matrix [K -K; -K K] per segment
  const int n = 6;
  std::vector<std::vector<int>> up(n), down(n);
  for (int e = 0; e + 1 < n; e++) {
    up[e].push_back(e + 1);
    down[e + 1].push_back(e);
  }
  CSSD_Matrix<T3, Pt> A;
  A.Initialization(up, down);
  for (int e = 0; e + 1 < n; e++) {
    DenseB L(2);
    T3 K;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) K.val[a][b] = (a == b ? 2.0 + e : 0.0) + 0.1 * (a + b);
    T3 mK = K;
    for (auto& r : mK.val)
      for (double& x : r) x = -x;
    L.A = {K, mK, mK, K};
    L.Fv[1] = Pt(0, 0, -1.0);
    std::vector<int> ids = {e, e + 1};
    A.SummPartOfMatrix(L, ids);
  }
  if (!near(A.Diag[1].val[0][0], 2.0 + 3.0, 1e-15) || !near(A.A_up[1][0].val[1][2], -0.3, 1e-15) ||
      !near(A.F[2].z, -1.0, 1e-15))
    return fail("block SummPartOfMatrix");
  std::vector<Pt> ones(n, Pt(1, 1, 1)), y;
  A.MultiplicationMatrixVector(ones, y);
  for (auto& p : y)
    if (std::fabs(p.x) + std::fabs(p.y) + std::fabs(p.z) > 1e-14) return fail("block A*translation != 0");
  T3 t;
  if (!A.GetValue(2, 3, t) || t.val[0][1] != A.A_up[2][0].val[0][1] || A.GetValue(2, 4, t) || A.GetValue(9, 0, t))
    return fail("block GetValue");
  if (!A.SetValue(3, 2, t) || A.SetValue(0, 5, t)) return fail("block SetValue");
  CSSD_Matrix<T3, Pt> S;
  S.SetMatrix(A);
  S.SummMatrix(A, 0.5, A, 0.5);
  if (!near(S.Diag[2].val[2][2], A.Diag[2].val[2][2], 1e-15)) return fail("block SummMatrix");
  S.MakeDivideOnSqrtFormSumm();
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) {
      double s = 0;
      for (int b = 0; b < 3; b++) s += S.Diag[i].val[a][b] * S.Diag[i].val[a][b];
      for (auto& B : S.A_up[i])
        for (int b = 0; b < 3; b++) s += B.val[a][b] * B.val[a][b];
      for (auto& B : S.A_down[i])
        for (int b = 0; b < 3; b++) s += B.val[a][b] * B.val[a][b];
      if (!near(s, 1.0, 1e-14)) return fail("block MakeDivideOnSqrtFormSumm");
    }
  S.ClearVariables();
  if (S.Diag[1].val[0][0] != 0.0 || S.F[2].z != 0.0) return fail("block ClearVariables");
  S.Symmetrization();  // appends (reference semantics): lower lists double
  if (S.A_down[1].size() != 2) return fail("block Symmetrization appends");
  return 0;
}

static int gpu_scalar() {
  const int n = 200;
  CSSD_Matrix<double, double> A;
  chain(A, n);
  // Dirichlet by elimination as FEM.h:505-619 does: u(0)=0, u(n-1)=1
  for (int d : {0, n - 1}) {
    const double g = d == 0 ? 0.0 : 1.0;
    A.Diag[d] = 1.0;
    A.F[d] = g;
    A.X[d] = g;
    for (auto& v : A.A_up[d]) v = 0.0;
    for (auto& v : A.A_down[d]) v = 0.0;
  }
  A.F[n - 2] -= A.A_up[n - 2][0] * 1.0;
  A.A_up[n - 2][0] = 0.0;
  A.A_down[1][0] = 0.0;
  // exact solution: piecewise linear with slopes 1/k_e, normalised
  std::vector<double> exact(n, 0.0);
  double tot = 0;
  for (int e = 0; e + 1 < n; e++) tot += 1.0 / (1.0 + (e % 7));
  for (int i = 1; i < n; i++) exact[i] = exact[i - 1] + (1.0 / (1.0 + ((i - 1) % 7))) / tot;
  auto err = [&](const std::vector<double>& x) {
    double m = 0;
    for (int i = 0; i < n; i++) m = std::max(m, std::fabs(x[i] - exact[i]));
    return m;
  };
  auto reset = [&] {
    std::fill(A.X.begin(), A.X.end(), 0.0);
    A.X[n - 1] = 1.0;
  };
  CSSD_Matrix<double, double> P;
  P.PrecondorSSOR(0.75, A);
  reset();
  const double r1 = A.MSG_PreconditioningSSOR(10000, 1e-12, P);
  if (!(r1 <= 1e-12) || err(A.X) > 1e-9) return fail("MSG_PreconditioningSSOR");
  reset();
  std::vector<double> rhs(A.F), res(n, 0.0);
  const double r1b = A.MSG_PreconditioningSSOR(rhs, res, 10000, 1e-12, P);
  if (!(r1b <= 1e-12) || err(res) > 1e-9) return fail("MSG_PreconditioningSSOR(RightSide, Result)");
  reset();
  const double r2 = A.MSG(10000, 1e-12);
  if (!(r2 <= 1e-12) || err(A.X) > 1e-9) return fail("MSG");
  reset();
  const double r3 = A.GMRES(10000, 1e-12);
  if (!(r3 <= 1e-12) || err(A.X) > 1e-9) return fail("GMRES");
  reset();
  const double r5 = A.BiCG_Stab(10000, 1e-12);
  if (!(std::fabs(r5) <= 1e-12) || err(A.X) > 1e-9) return fail("BiCG_Stab");
  reset();
  const double r6 = A.BCG_Stab2(10000, 1e-12);
  if (!(r6 <= 1e-12) || err(A.X) > 1e-9) return fail("BCG_Stab2");
  reset();
  const double r7 = A.BiCG_Stab_Preconditioning(10000, 1e-12, P);
  if (!(std::fabs(r7) <= 1e-12) || err(A.X) > 1e-9) return fail("BiCG_Stab_Preconditioning");
  // already converged: the BiCGStab family returns the residual negated
  const double r8 = A.BiCG_Stab(10000, 1e-6);
  if (!(r8 <= 0.0) || std::fabs(r8) > 1e-6) return fail("BiCG_Stab sign when no iteration is made");
  reset();
  std::vector<double> diag(A.Diag);
  const double r4 = A.BiCG_Diag(10000, 1e-10, diag);
  if (!(r4 < 1e-10) || err(A.X) > 1e-8) return fail("BiCG_Diag");
  return 0;
}

// Block system from tests/test_dropin.py: int64 nb, int32 rowptr[nb+1],
// int32 col[nnzb], double val[9 nnzb], double rhs[3 nb], double x0[3 nb].
static int gpu_bsr3(const char* in, const char* out) {
  FILE* f = std::fopen(in, "rb");
  if (!f) return fail("open input");
  long long nb = 0;
  bool ok = std::fread(&nb, 8, 1, f) == 1 && nb > 0 && nb < (1 << 24);
  std::vector<int> rp(ok ? nb + 1 : 0);
  ok = ok && std::fread(rp.data(), 4, nb + 1, f) == (size_t)(nb + 1) && rp[0] == 0 && rp[nb] > 0;
  std::vector<int> col(ok ? rp[nb] : 0);
  std::vector<double> val(ok ? 9 * (size_t)rp[nb] : 0), rhs(ok ? 3 * nb : 0), x0(ok ? 3 * nb : 0);
  ok = ok && std::fread(col.data(), 4, col.size(), f) == col.size() &&
       std::fread(val.data(), 8, val.size(), f) == val.size() && std::fread(rhs.data(), 8, rhs.size(), f) == rhs.size() &&
       std::fread(x0.data(), 8, x0.size(), f) == x0.size();
  std::fclose(f);
  if (!ok) return fail("read input");
  std::vector<std::vector<int>> up(nb), down(nb);
  for (long long i = 0; i < nb; i++)
    for (int k = rp[i]; k < rp[i + 1]; k++) {
      if (col[k] < 0 || col[k] >= nb) return fail("column out of range");
      if (col[k] < i) down[i].push_back(col[k]);
      if (col[k] > i) up[i].push_back(col[k]);
    }
  CSSD_Matrix<T3, Pt> A;
  A.Initialization(up, down);
  for (long long i = 0; i < nb; i++) {
    int du = 0, dd = 0;
    for (int k = rp[i]; k < rp[i + 1]; k++) {
      T3 b;
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) b.val[a][c] = val[9 * (size_t)k + 3 * a + c];
      if (col[k] < i) A.A_down[i][dd++] = b;
      else if (col[k] > i) A.A_up[i][du++] = b;
      else A.Diag[i] = b;
    }
    A.F[i] = Pt(rhs[3 * i], rhs[3 * i + 1], rhs[3 * i + 2]);
  }
  auto reset = [&] {
    for (long long i = 0; i < nb; i++) A.X[i] = Pt(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]);
  };
  FILE* o = std::fopen(out, "wb");
  if (!o) return fail("open output");
  CSSD_Matrix<double, double> P;
  std::vector<double> diag(3 * nb, 1.0);
  const char* names[] = {"MSG", "MSG_PreconditioningSSOR", "BiCG_Diag", "GMRES", "BiCG_Stab"};
  for (int m = 0; m < 5; m++) {
    reset();
    double r = 0;
    switch (m) {
      case 0: r = A.MSG(200000, 1e-12); break;
      case 1: r = A.MSG_PreconditioningSSOR(200000, 1e-12, P); break;
      case 2: r = A.BiCG_Diag(200000, 1e-12, diag); break;
      case 3: r = A.GMRES(200000, 1e-12); break;
      case 4: r = A.BiCG_Stab(200000, 1e-12); break;
    }
    std::printf("%s: returned %.3e, %d iterations, true relres %.3e\n", names[m], r, A.last_report.iterations,
                A.last_report.true_relres);
    for (long long i = 0; i < nb; i++) {
      const double p[3] = {A.X[i].x, A.X[i].y, A.X[i].z};
      std::fwrite(p, 8, 3, o);
    }
  }
  std::fclose(o);
  return 0;
}

int main(int argc, char** argv) {
  const char* mode = argc > 1 ? argv[1] : "cpu";
  if (std::strcmp(mode, "bsr3") == 0) {
    if (argc < 4) return fail("usage: bsr3 IN OUT");
    if (gpu_bsr3(argv[2], argv[3])) return 1;
    std::printf("OK\n");
    return 0;
  }
  if (host_scalar() || host_block()) return 1;
  if (std::strcmp(mode, "gpu") == 0 && gpu_scalar()) return 1;
  std::printf("OK\n");
  return 0;
}
