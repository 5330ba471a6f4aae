// Exercises include/mfs_cssd.hpp the way the reference's drivers use
// CSSD_Matrix: portrait from element cliques, SummPartOfMatrix of local
// matrices, direct writes to Diag/A_up/A_down/F for boundary conditions, then
// the solver methods.  Value types are minimal local stand-ins with the
// operators the class uses (double; a 3-vector; a 3x3 tensor with .val).
// Usage: test_cssd_dropin cpu|gpu     (prints "OK" on success)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <set>
#include <vector>

#include "mfs_cssd.hpp"

struct Dense {  // DenseMatrix<double,double>-like local matrix
  int n;
  std::vector<double> A, Fv, Xv;
  explicit Dense(int n_) : n(n_), A(n_ * n_, 0.0), Fv(n_, 0.0), Xv(n_, 0.0) {}
  int GetSize() { return n; }
  double GetElementOfA(int i, int j) { return A[i * n + j]; }
  double GetElementOfF(int i) { return Fv[i]; }
  double GetElementOfX(int i) { return Xv[i]; }
};

static int fail(const char* msg) {
  std::printf("FAIL: %s\n", msg);
  return 1;
}

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  // 1-D chain of n nodes, 2-node elements [1 -1; -1 1] * k, Dirichlet at both
  // ends: a Laplace problem with the linear solution x_i = i/(n-1).
  const int n = 200;
  std::vector<std::vector<int>> up(n), down(n);
  for (int e = 0; e + 1 < n; e++) {
    up[e].push_back(e + 1);
    down[e + 1].push_back(e);
  }
  CSSD_Matrix<double, double> A;
  A.Initialization(up, down);
  for (int e = 0; e + 1 < n; e++) {
    Dense L(2);
    const double k = 1.0 + (e % 7);
    L.A = {k, -k, -k, k};
    std::vector<int> ids = {e, e + 1};
    A.SummPartOfMatrix(L, ids);
  }
  if (A.Diag[0] != 1.0 || A.Diag[1] != 3.0 || A.A_up[1][0] != -2.0 || A.A_down[1][0] != -1.0)
    return fail("SummPartOfMatrix values");
  // A * 1 = 0 before boundary conditions
  std::vector<double> one(n, 1.0), y;
  A.MultiplicationMatrixVector(one, y);
  for (double v : y)
    if (std::fabs(v) > 1e-14) return fail("A*1 != 0");
  if (A.GetValue(5, 6) != A.A_up[5][0] || A.GetValue(6, 5) != A.A_down[6][0]) return fail("GetValue");
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
  if (!gpu) {
    std::printf("OK\n");
    return 0;
  }
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
  CSSD_Matrix<double, double> P;
  const double r1 = A.MSG_PreconditioningSSOR(10000, 1e-12, P);
  if (!(r1 <= 1e-12) || err(A.X) > 1e-9) return fail("MSG_PreconditioningSSOR");
  std::fill(A.X.begin(), A.X.end(), 0.0);
  const double r2 = A.MSG(10000, 1e-12);
  if (!(r2 <= 1e-12) || err(A.X) > 1e-9) return fail("MSG");
  std::fill(A.X.begin(), A.X.end(), 0.0);
  const double r3 = A.GMRES(10000, 1e-12);
  if (!(r3 <= 1e-12) || err(A.X) > 1e-9) return fail("GMRES");
  std::fill(A.X.begin(), A.X.end(), 0.0);
  const double r5 = A.BiCG_Stab(10000, 1e-12);
  if (!(r5 <= 1e-12) || err(A.X) > 1e-9) return fail("BiCG_Stab");
  std::fill(A.X.begin(), A.X.end(), 0.0);
  std::vector<double> diag(A.Diag);
  const double r4 = A.BiCG_Diag(10000, 1e-10, diag);
  if (!(r4 < 1e-10) || err(A.X) > 1e-8) return fail("BiCG_Diag");
  std::printf("OK\n");
  return 0;
}
