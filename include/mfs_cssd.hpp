// mfs_cssd.hpp - drop-in replacement for Library/CSSD_Matrix.h of
// AnastasiaKutishcheva/MultiscaleFEMSolver3D, backed by libmfs.so.
//
// The class keeps the reference's public data members and every public
// method of CSSD_Matrix<TForMatrix, TForVector> (CSSD_Matrix.h:7-1990) with
// the same signatures, because the drivers write Diag / A_up / A_down / X / F
// directly (FEM.h:517-528, :2202-2216, MsFEM.h:904-915) and call the methods
// by name.  Host-side methods (Initialization, SummPartOfMatrix, GetValue,
// MultiplicationMatrixVector, the SSOR factor builders, ...) run on the CPU
// with the reference's semantics; the solver methods pack the rows into CSR
// (double) or BSR3 (3x3 blocks), solve on the GPU through the mfs_* C-ABI,
// write X back and return the value the reference returns:
//   MSG, MSG_PreconditioningSSOR, GMRES, BCG_Stab2   true ||F-AX||/||F||
//   BiCG_Stab, BiCG_Stab_blocks,
//   BiCG_Stab_Preconditioning                         the same, negated when
//                                                     no iteration was made
//   BiCG_Diag                                         ||diag^-1 (F-AX)|| (absolute)
// Method mapping (the reference's own algorithms are replaced):
//   MSG                      -> CG (FP64)
//   MSG_PreconditioningSSOR  -> Jacobi-PCG (FP64); the SSOR factor is not used
//   BiCG_Diag                -> Jacobi-MINRES (FP64): the reference applies it
//                               to symmetric elastic systems that are
//                               indefinite in its element convention
//   GMRES                    -> GMRES(30) with Jacobi, FP64
//   BiCG_Stab family         -> BiCGStab with right Jacobi preconditioning, FP64
// Set MFS_CSSD_MIXED=1 in the environment to run the mixed-precision path
// (FP32 inner iterations inside FP64 refinement).
//
// Integration: replace the body of the reference's Library/CSSD_Matrix.h by
// an include of this header (FEM.h and MsFEM.h include that file by a
// relative path), add -I<repo>/include and link -lmfs.  The value types are
// the reference's: double, or Tensor2Rank3D (.val[3][3]) with Point<double>
// (.x/.y/.z); this header only needs those members plus assignment from
// double, +=, and Tensor2Rank3D * Point<double>.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mfs.h"

namespace mfs_cssd {

// One context per process (the reference has no device notion).
inline mfs_ctx default_ctx() {
  static mfs_ctx ctx = [] {
    mfs_ctx c = nullptr;
    if (mfs_ctx_create(0, &c) != MFS_OK) {
      std::fprintf(stderr, "mfs_cssd: %s\n", mfs_last_error());
      std::abort();
    }
    return c;
  }();
  return ctx;
}

inline bool use_mixed() {
  const char* e = std::getenv("MFS_CSSD_MIXED");
  return e && e[0] == '1';
}

}  // namespace mfs_cssd

template <typename TForMatrix, typename TForVector>
class CSSD_Matrix {
 public:
  std::vector<TForMatrix> Diag;
  std::vector<std::vector<TForMatrix>> A_up, A_down;
  std::vector<std::vector<int>> id_column_for_A_up, id_column_for_A_down;
  std::vector<TForVector> X, F;
  bool print_logs = false;
  FILE* log_out = nullptr;

  CSSD_Matrix() {}

  int GetMatrixSize() { return (int)Diag.size(); }

  // CSSD_Matrix.h:38-69
  void SetMatrix(CSSD_Matrix<TForMatrix, TForVector>& A) {
    Diag = A.Diag;
    X = A.X;
    F = A.F;
    print_logs = A.print_logs;
    log_out = A.log_out;
    A_up = A.A_up;
    A_down = A.A_down;
    id_column_for_A_up = A.id_column_for_A_up;
    id_column_for_A_down = A.id_column_for_A_down;
  }

  std::vector<int>* GetIDsColumnsInDownString(int r) { return &id_column_for_A_down[r]; }
  std::vector<int>* GetIDsColumnsInUpString(int r) { return &id_column_for_A_up[r]; }

  // CSSD_Matrix.h:79-99: sizes follow the given portrait; like the
  // reference, existing values in rows that keep their length are kept
  // (a fresh matrix starts at zero).
  void Initialization(std::vector<std::vector<int>>& up, std::vector<std::vector<int>>& down) {
    const size_t n = up.size();
    id_column_for_A_up.resize(n);
    id_column_for_A_down.resize(n);
    Diag.resize(n, zero_m());
    A_up.resize(n);
    A_down.resize(n);
    X.resize(n, zero_v());
    F.resize(n, zero_v());
    for (size_t i = 0; i < n; i++) {
      A_up[i].resize(up[i].size(), zero_m());
      A_down[i].resize(down[i].size(), zero_m());
      id_column_for_A_up[i] = up[i];
      id_column_for_A_down[i] = down[i];
    }
  }

  // CSSD_Matrix.h:100-113
  void UpdateEquation(int id, TForMatrix& diag, TForMatrix& non_diag, TForVector& x, TForVector& f) {
    Diag[id] = diag;
    X[id] = x;
    F[id] = f;
    for (auto& v : A_up[id]) v = non_diag;
    for (auto& v : A_down[id]) v = non_diag;
  }

  // CSSD_Matrix.h:114-127: the lower triangle from the upper one, appended
  // to whatever the lower lists already hold (as the reference does).
  void Symmetrization() {
    id_column_for_A_down.resize(Diag.size());
    A_down.resize(Diag.size());
    for (size_t i = 0; i < id_column_for_A_up.size(); i++)
      for (size_t jj = 0; jj < id_column_for_A_up[i].size(); jj++) {
        const int j = id_column_for_A_up[i][jj];
        id_column_for_A_down[j].push_back((int)i);
        A_down[j].push_back(A_up[i][jj]);
      }
  }

  // CSSD_Matrix.h:128-171: random symmetric test matrix with diagonal
  // values in [0, 1000], off-diagonal values in [0, 1], X = X_val and
  // F = A X.  Rows get up to N/10 - 1 upper entries at distinct random
  // columns (none when N < 20).
  void RandomMatrix(int N, double X_val) {
    Diag.resize(N);
    X.resize(N);
    F.resize(N);
    id_column_for_A_up.assign(N, {});
    A_up.assign(N, {});
    id_column_for_A_down.assign(N, {});
    A_down.assign(N, {});
    const int cap = N / 10;
    for (int i = 0; i < N; i++) {
      Diag[i] = std::rand() * 1000.0 / RAND_MAX;
      X[i] = X_val;
      const int count = (cap > 0 && N - i - 1 > 0) ? (std::rand() % (N - i)) % cap : 0;
      std::vector<int> cols;
      while ((int)cols.size() < count) {
        const int c = std::rand() % (N - i - 1) + (i + 1);
        if (std::find(cols.begin(), cols.end(), c) == cols.end()) cols.push_back(c);
      }
      std::sort(cols.begin(), cols.end());
      id_column_for_A_up[i] = cols;
      A_up[i].resize(cols.size());
      for (auto& v : A_up[i]) v = std::rand() * 1.0 / RAND_MAX;
    }
    Symmetrization();
    MultiplicationMatrixVector(X, F);
  }

  // CSSD_Matrix.h:172-179: all values, X and F to zero; the portrait stays.
  void ClearVariables() {
    for (auto& v : Diag) v = zero_m();
    for (auto& r : A_up)
      for (auto& v : r) v = zero_m();
    for (auto& r : A_down)
      for (auto& v : r) v = zero_m();
    for (auto& v : X) v = zero_v();
    for (auto& v : F) v = zero_v();
  }

  // CSSD_Matrix.h:180-210: false when a row or column index is past the
  // matrix or the entry is not in the portrait.
  bool SetValue(int id_string, int id_column, TForMatrix& val) {
    TForMatrix* p = locate(id_string, id_column);
    if (!p) return false;
    *p = val;
    return true;
  }
  // CSSD_Matrix.h:211-241
  bool GetValue(int id_string, int id_column, TForMatrix& val) {
    TForMatrix* p = locate(id_string, id_column);
    if (!p) return false;
    val = *p;
    return true;
  }

  // CSSD_Matrix.h:243-305 (scalar matrices): largest eigenvalue by power
  // iteration from a Gaussian start vector, stopped when two successive
  // Rayleigh quotients agree to 1e-3 relative, or after size - 1 steps.
  double SolveMaxEigenvalue() {
    const int n = GetMatrixSize();
    std::vector<TForVector> x(n), t;
    double s = 0;
    for (int i = 0; i < n; i++) {
      double u = 0, v = 0, r2 = 0;
      do {
        u = 2.0 * std::rand() / RAND_MAX - 1.0;
        v = 2.0 * std::rand() / RAND_MAX - 1.0;
        r2 = u * u + v * v;
      } while (r2 >= 1.0 || r2 == 0.0);
      x[i] = u * std::sqrt(-2.0 * std::log(r2) / r2);
      s += x[i] * x[i];
    }
    for (auto& v : x) v /= std::sqrt(s);
    MultiplicationMatrixVector(x, t);
    double prev = dot(x, t), lambda = prev;
    for (int k = 1; k < n; k++) {
      MultiplicationMatrixVector(x, t);
      const double nrm = std::sqrt(dot(t, t));
      for (int i = 0; i < n; i++) x[i] = t[i] / nrm;
      MultiplicationMatrixVector(x, t);
      lambda = dot(x, t);
      if (std::fabs(lambda - prev) / std::fabs(prev) <= 1e-3) break;
      prev = lambda;
    }
    return lambda;
  }

  // CSSD_Matrix.h:307-363: scatter-add a local matrix (and its local F, X);
  // entries missing from the portrait are reported like the reference does.
  template <typename Dense>
  void SummPartOfMatrix(Dense& matrix, std::vector<int>& id_row) {
    for (int i = 0; i < matrix.GetSize(); i++) {
      F[id_row[i]] += matrix.GetElementOfF(i);
      X[id_row[i]] += matrix.GetElementOfX(i);
      for (int j = 0; j < matrix.GetSize(); j++) {
        const int s = id_row[i], c = id_row[j];
        if (s == c) {
          Diag[s] += matrix.GetElementOfA(i, i);
          continue;
        }
        TForMatrix* p = find_linear(s, c);
        if (p)
          *p += matrix.GetElementOfA(i, j);
        else
          std::printf("We don't find the element [%d][%d].\n", s, c);
      }
    }
  }
  // CSSD_Matrix.h:364-379
  void SummPartOfVector(std::vector<TForVector>& v, std::vector<int>& id_row) {
    for (size_t i = 0; i < v.size(); i++) F[id_row[i]] += v[i];
  }

  // CSSD_Matrix.h:382-394: this = A*coefA + B*coefB over this matrix's
  // portrait (A and B share it).
  void SummMatrix(CSSD_Matrix<TForMatrix, TForVector>& A, double coefA, CSSD_Matrix<TForMatrix, TForVector>& B,
                  double coefB) {
    for (size_t i = 0; i < Diag.size(); i++) {
      Diag[i] = lin(A.Diag[i], coefA, B.Diag[i], coefB);
      for (size_t j = 0; j < A_up[i].size(); j++) A_up[i][j] = lin(A.A_up[i][j], coefA, B.A_up[i][j], coefB);
      for (size_t j = 0; j < A_down[i].size(); j++)
        A_down[i][j] = lin(A.A_down[i][j], coefA, B.A_down[i][j], coefB);
    }
  }

  // CSSD_Matrix.h:395-423 (host)
  void MultiplicationMatrixVector(std::vector<TForVector>& B, std::vector<TForVector>& result) {
    const int n = (int)B.size();
    if ((int)result.size() != n) result.resize(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
      TForVector t = zero_v();
      for (size_t j = 0; j < A_up[i].size(); j++) t += mul(A_up[i][j], B[id_column_for_A_up[i][j]]);
      for (size_t j = 0; j < A_down[i].size(); j++) t += mul(A_down[i][j], B[id_column_for_A_down[i][j]]);
      t += mul(Diag[i], B[i]);
      result[i] = t;
    }
  }

  // CSSD_Matrix.h:424-457: divide every equation by the 2-norm of its row
  // (for 3x3 blocks, each of the three scalar rows by its own norm), X and
  // F included.
  void MakeDivideOnSqrtFormSumm() {
    for (int i = 0; i < GetMatrixSize(); i++) {
      double s[3] = {0, 0, 0};
      row_sq(Diag[i], s);
      for (auto& v : A_up[i]) row_sq(v, s);
      for (auto& v : A_down[i]) row_sq(v, s);
      for (double& v : s) v = std::sqrt(v);
      row_div(Diag[i], s);
      for (auto& v : A_up[i]) row_div(v, s);
      for (auto& v : A_down[i]) row_div(v, s);
      vec_div(X[i], s);
      vec_div(F[i], s);
    }
  }

  // CSSD_Matrix.h:458-513: the reference's SSOR factor of A in this object:
  // A's portrait (lower part built from the upper one when A has none),
  // off-diagonal values -w*a, diagonal 1 / (a_ii w (2 - w)).  The GPU
  // solvers do not use it (MSG_PreconditioningSSOR runs Jacobi-PCG); it is
  // kept for SolvePrecondorSSOR_SLAE.
  void PrecondorSSOR(double w, CSSD_Matrix<double, double>& A) {
    const size_t n = A.Diag.size();
    Diag.resize(n);
    id_column_for_A_up = A.id_column_for_A_up;
    A_up = A.A_up;
    if (A.A_down.empty()) {
      id_column_for_A_down.assign(A.A_up.size(), {});
      A_down.assign(A.A_up.size(), {});
      for (size_t i = 0; i < A.id_column_for_A_up.size(); i++)
        for (size_t jj = 0; jj < A.id_column_for_A_up[i].size(); jj++) {
          const int j = A.id_column_for_A_up[i][jj];
          id_column_for_A_down[j].push_back((int)i);
          A_down[j].push_back(A.A_up[i][jj]);
        }
    } else {
      id_column_for_A_down = A.id_column_for_A_down;
      A_down = A.A_down;
    }
    for (size_t i = 0; i < n; i++) {
      Diag[i] = 1.0 / A.Diag[i] / w / (2 - w);
      for (auto& v : A_up[i]) v = -w * v;
      for (auto& v : A_down[i]) v = -w * v;
    }
  }
  // CSSD_Matrix.h:514-547: apply the factor built by PrecondorSSOR to Z:
  // backward sweep over the upper part, scaling by 1/Diag, forward sweep
  // over the lower part, both sweeps dividing by free_diag.
  void SolvePrecondorSSOR_SLAE(std::vector<double>& free_diag, std::vector<double>& Xv, std::vector<double>& Z) {
    const int n = GetMatrixSize();
    Xv.resize(Z.size());
    for (int I = n - 1; I >= 0; I--) {
      double s = Z[I];
      for (size_t k = 0; k < id_column_for_A_up[I].size(); k++) s -= Xv[id_column_for_A_up[I][k]] * A_up[I][k];
      Xv[I] = s / free_diag[I];
    }
    for (int I = 0; I < n; I++) Xv[I] = Xv[I] / Diag[I];
    for (int I = 0; I < n; I++) {
      double s = Xv[I];
      for (size_t k = 0; k < id_column_for_A_down[I].size(); k++)
        s -= Xv[id_column_for_A_down[I][k]] * A_down[I][k];
      Xv[I] = s / free_diag[I];
    }
  }
  // CSSD_Matrix.h:548-632: the reference's "symmetric" SSOR-like factor.
  // Reproduces its traversal: in row i only the run of upper columns
  // i+1, i+2, ... that are all present (up to the first missing one) is
  // visited; for those, diag_i = a_ii + sum (w^2/a_pp) a_ip^2 and
  // up_ij = w a_ij + sum_{p>j, p in up(i) and up(j)} (w^2/a_pp) a_ip a_jp;
  // entries past the run keep their previous values; then Symmetrization.
  void PrecondorSSOR_summetric(double w, CSSD_Matrix<double, double>& A) {
    const size_t n = A.Diag.size();
    Diag.resize(n);
    A_up.resize(A.A_up.size());
    for (size_t i = 0; i < A_up.size(); i++) A_up[i].resize(A.A_up[i].size());
    id_column_for_A_up = A.id_column_for_A_up;
    for (size_t i = 0; i < n; i++) {
      const auto& ci = id_column_for_A_up[i];
      size_t run = 0;
      while (run < ci.size() && ci[run] == (int)(i + 1 + run)) run++;
      double d = A.Diag[i];
      for (size_t t = 0; t < run; t++) d += (w * w / A.Diag[ci[t]]) * A.A_up[i][t] * A.A_up[i][t];
      Diag[i] = d;
      for (size_t jj = 0; jj < run; jj++) {
        const int j = ci[jj];
        const auto& cj = id_column_for_A_up[j];
        double v = w * A.A_up[i][jj];
        // p > j in both sorted lists
        size_t a = jj + 1, b = 0;
        while (b < cj.size() && cj[b] <= j) b++;
        while (a < ci.size() && b < cj.size()) {
          if (ci[a] < cj[b]) {
            a++;
          } else if (cj[b] < ci[a]) {
            b++;
          } else {
            v += (w * w / A.Diag[ci[a]]) * A.A_up[i][a] * A.A_up[j][b];
            a++;
            b++;
          }
        }
        A_up[i][jj] = v;
      }
    }
    Symmetrization();
  }

  // ---- solvers (GPU) -----------------------------------------------------
  // CSSD_Matrix.h:634-729
  double BCG_Stab2(int maxiter, double E) { return solve_inplace(MFS_BICGSTAB, maxiter, E, X, F); }
  // CSSD_Matrix.h:730-853
  double BiCG_Stab(int maxiter, double E) { return stab_result(solve_inplace(MFS_BICGSTAB, maxiter, E, X, F)); }
  // CSSD_Matrix.h:854-1057
  double BiCG_Stab_blocks(int maxiter, double E) {
    return stab_result(solve_inplace(MFS_BICGSTAB, maxiter, E, X, F));
  }
  // CSSD_Matrix.h:1058-1163 (the preconditioner matrix M is not used: right
  // Jacobi preconditioning on the GPU)
  double BiCG_Stab_Preconditioning(int maxiter, double E, CSSD_Matrix<double, double>&) {
    return stab_result(solve_inplace(MFS_BICGSTAB, maxiter, E, X, F));
  }
  // CSSD_Matrix.h:1166-1178: X = F / diag.  A diagonal entry below 1e-15 in
  // magnitude is reported on stderr (the reference then waits for a key).
  void Start_Diag_Solver(const std::vector<double>& _F, std::vector<double>& _X, std::vector<double>& _Diag) {
    for (size_t i = 0; i < _Diag.size(); i++) {
      if (std::fabs(_Diag[i]) < 1e-15) std::fprintf(stderr, "%g < %g\n", _Diag[i], 1e-15);
      _X[i] = _F[i] / _Diag[i];
    }
  }
  // CSSD_Matrix.h:1181-1300: returns the absolute ||diag^-1 (F - AX)||.
  double BiCG_Diag(int maxiter, double E, std::vector<double>& _Diag) {
    // stop on ||diag^-1 r|| < E: a relative target on ||r|| that implies it
    const double dn = norm_scaled_rhs(_Diag);
    const double rel = dn > 0 ? E / dn : E;
    solve_inplace(MFS_MINRES, maxiter, rel, X, F);
    return precond_residual(_Diag);
  }
  double MSG_PreconditioningSSOR(int maxiter, double E, CSSD_Matrix<double, double>&) {
    return solve_inplace(MFS_PCG_JACOBI, maxiter, E, X, F);
  }
  double MSG_PreconditioningSSOR(std::vector<double>& RightSide, std::vector<double>& Result, int maxiter, double E,
                                 CSSD_Matrix<double, double>&) {
    return solve_inplace(MFS_PCG_JACOBI, maxiter, E, Result, RightSide);
  }
  double MSG(int maxiter, double E) { return solve_inplace(MFS_CG, maxiter, E, X, F); }
  double MSG(int maxiter, double E, std::vector<TForVector>& X_new, std::vector<TForVector>& F_new, bool) {
    return solve_inplace(MFS_CG, maxiter, E, X_new, F_new);
  }
  // CSSD_Matrix.h:1727-1773: least squares of GMRES's (m+1) x m Hessenberg
  // system with right side f = (Norma_r, 0, ..., 0): Givens rotations
  // triangularise H and rotate f in place, back substitution leaves y in
  // f[0..m-1]; returns |f[m]|, the residual norm.
  double Minimization_Problem(int m, double Norma_r, std::vector<double>& f, std::vector<std::vector<double>>& H) {
    f.assign(m + 1, 0.0);
    f[0] = Norma_r;
    for (int i = 0; i < m; i++) {
      const double t = H[i + 1][i] / H[i][i];
      const double c = 1.0 / std::sqrt(1 + t * t), s = t * c;
      for (size_t k = i; k < H[0].size(); k++) {
        const double hi = H[i][k], hn = H[i + 1][k];
        H[i][k] = c * hi + s * hn;
        H[i + 1][k] = c * hn - s * hi;
      }
      const double fi = f[i], fn = f[i + 1];
      f[i] = c * fi + s * fn;
      f[i + 1] = c * fn - s * fi;
    }
    for (int i = m - 1; i >= 0; i--) {
      f[i] /= H[i][i];
      for (int j = i - 1; j >= 0; j--) f[j] -= H[j][i] * f[i];
    }
    return std::fabs(f[m]);
  }
  // CSSD_Matrix.h:1774-1989 (restart 30 instead of the reference's 5; the
  // matrix is not row-scaled in place)
  double GMRES(int maxiter, double E) { return solve_inplace(MFS_GMRES, maxiter, E, X, F); }

  // Diagnostics of the last GPU solve.
  mfs_solve_report last_report{};

 private:
  static constexpr bool kScalar = std::is_arithmetic<TForMatrix>::value;

  static TForMatrix zero_m() {
    TForMatrix z;
    z = 0.0;
    return z;
  }
  static TForVector zero_v() {
    TForVector z;
    z = 0.0;
    return z;
  }
  // Tensor2Rank3D::operator*(Point) (Tensor.h:211-226) for blocks
  static TForVector mul(TForMatrix a, const TForVector& b) { return a * b; }

  static double dot(const std::vector<TForVector>& a, const std::vector<TForVector>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
  }
  static TForMatrix lin(const TForMatrix& a, double ca, const TForMatrix& b, double cb) {
    if constexpr (kScalar) {
      return a * ca + b * cb;
    } else {
      TForMatrix r = a;
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) r.val[p][q] = a.val[p][q] * ca + b.val[p][q] * cb;
      return r;
    }
  }
  static void row_sq(const TForMatrix& a, double* s) {
    if constexpr (kScalar) {
      s[0] += a * a;
    } else {
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) s[p] += a.val[p][q] * a.val[p][q];
    }
  }
  static void row_div(TForMatrix& a, const double* s) {
    if constexpr (kScalar) {
      a /= s[0];
    } else {
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) a.val[p][q] /= s[p];
    }
  }
  static void vec_div(TForVector& v, const double* s) {
    if constexpr (std::is_arithmetic<TForVector>::value) {
      v /= s[0];
    } else {
      v.x /= s[0];
      v.y /= s[1];
      v.z /= s[2];
    }
  }

  // Entry lookup of SetValue/GetValue: the index lists are sorted, so a
  // binary search finds the column (the reference scans the sorted list).
  TForMatrix* locate(int row, int col) {
    const int n = GetMatrixSize();
    if (row >= n || col >= n || row < 0 || col < 0) return nullptr;
    if (row == col) return &Diag[row];
    auto& ids = row < col ? id_column_for_A_up[row] : id_column_for_A_down[row];
    auto& vals = row < col ? A_up[row] : A_down[row];
    auto it = std::lower_bound(ids.begin(), ids.end(), col);
    if (it == ids.end() || *it != col) return nullptr;
    return &vals[it - ids.begin()];
  }
  // Entry lookup of SummPartOfMatrix: a linear scan, as the reference does.
  TForMatrix* find_linear(int row, int col) {
    auto& ids = row < col ? id_column_for_A_up[row] : id_column_for_A_down[row];
    auto& vals = row < col ? A_up[row] : A_down[row];
    for (size_t k = 0; k < ids.size(); k++)
      if (ids[k] == col) return &vals[k];
    return nullptr;
  }

  // CSR (down ++ diag ++ up) or BSR3 packing of the rows.
  void pack(std::vector<int32_t>& rp, std::vector<int32_t>& col, std::vector<double>& val) const {
    const int n = (int)Diag.size();
    rp.assign(n + 1, 0);
    for (int i = 0; i < n; i++) rp[i + 1] = rp[i] + (int32_t)(A_down[i].size() + 1 + A_up[i].size());
    col.resize(rp[n]);
    const int bw = kScalar ? 1 : 9;
    val.resize((size_t)bw * rp[n]);
    for (int i = 0; i < n; i++) {
      int32_t p = rp[i];
      for (size_t j = 0; j < A_down[i].size(); j++, p++) {
        col[p] = id_column_for_A_down[i][j];
        put(val, p, A_down[i][j]);
      }
      col[p] = i;
      put(val, p++, Diag[i]);
      for (size_t j = 0; j < A_up[i].size(); j++, p++) {
        col[p] = id_column_for_A_up[i][j];
        put(val, p, A_up[i][j]);
      }
    }
  }
  static void put(std::vector<double>& val, int32_t p, const TForMatrix& a) {
    if constexpr (kScalar) {
      val[p] = a;
    } else {
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) val[9 * (size_t)p + 3 * r + c] = a.val[r][c];
    }
  }
  static void flat(const std::vector<TForVector>& v, std::vector<double>& out) {
    if constexpr (std::is_arithmetic<TForVector>::value) {
      out.assign(v.begin(), v.end());
    } else {
      out.resize(3 * v.size());
      for (size_t i = 0; i < v.size(); i++) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
      }
    }
  }
  static void unflat(const std::vector<double>& in, std::vector<TForVector>& v) {
    if constexpr (std::is_arithmetic<TForVector>::value) {
      for (size_t i = 0; i < v.size(); i++) v[i] = in[i];
    } else {
      for (size_t i = 0; i < v.size(); i++) {
        v[i].x = in[3 * i];
        v[i].y = in[3 * i + 1];
        v[i].z = in[3 * i + 2];
      }
    }
  }

  static void check(int st, const char* where) {
    if (st != MFS_OK && st != MFS_ERR_NOT_CONVERGED)
      throw std::runtime_error(std::string(where) + ": " + mfs_last_error());
  }

  // BiCGStab family: negative when the loop made no iteration
  // (CSSD_Matrix.h:848-851).
  double stab_result(double r) const { return last_report.iterations == 0 ? -r : r; }

  double solve_inplace(int method, int maxiter, double E, std::vector<TForVector>& Xv, std::vector<TForVector>& Fv) {
    mfs_ctx ctx = mfs_cssd::default_ctx();
    std::vector<int32_t> rp, col;
    std::vector<double> val, x, f;
    pack(rp, col, val);
    flat(Xv, x);
    flat(Fv, f);
    mfs_mat A = nullptr;
    check(kScalar ? mfs_mat_create_csr(ctx, (int64_t)Diag.size(), rp.data(), col.data(), val.data(), &A)
                  : mfs_mat_create_bsr3(ctx, (int64_t)Diag.size(), rp.data(), col.data(), val.data(), &A),
          "CSSD_Matrix: matrix upload");
    mfs_vec b = nullptr, xv = nullptr;
    check(mfs_vec_create(ctx, (int64_t)f.size(), MFS_F64, &b), "CSSD_Matrix: vector");
    check(mfs_vec_create(ctx, (int64_t)x.size(), MFS_F64, &xv), "CSSD_Matrix: vector");
    check(mfs_vec_upload(b, f.data()), "CSSD_Matrix: upload F");
    check(mfs_vec_upload(xv, x.data()), "CSSD_Matrix: upload X");
    mfs_solve_desc d;
    mfs_solve_desc_default(&d);
    d.method = method;
    d.precision = mfs_cssd::use_mixed() ? MFS_PREC_MIXED : MFS_PREC_F64;
    d.tol = E > 0 ? E : 1e-12;
    d.maxit = maxiter > 0 ? maxiter : 1;
    mfs_solve_report rep{};
    const int st = mfs_solve(A, &d, b, xv, &rep);
    check(st == MFS_ERR_NONFINITE ? MFS_OK : st, "CSSD_Matrix: solve");
    check(mfs_vec_download(xv, x.data()), "CSSD_Matrix: download X");
    unflat(x, Xv);
    mfs_vec_destroy(b);
    mfs_vec_destroy(xv);
    mfs_mat_destroy(A);
    last_report = rep;
    if (print_logs)
      std::printf("\tMFS\t>\t%d\t-\t%.5e (resolve residual)\n", rep.iterations, rep.true_relres);
    return rep.true_relres;
  }

  double norm_scaled_rhs(const std::vector<double>& diag) const {
    double s = 0;
    std::vector<double> f;
    flat(F, f);
    for (size_t i = 0; i < f.size(); i++) s += (f[i] / diag[i]) * (f[i] / diag[i]);
    return std::sqrt(s);
  }
  double precond_residual(const std::vector<double>& diag) {
    std::vector<TForVector> ax;
    MultiplicationMatrixVector(X, ax);
    std::vector<double> a, f;
    flat(ax, a);
    flat(F, f);
    double s = 0;
    for (size_t i = 0; i < f.size(); i++) s += ((f[i] - a[i]) / diag[i]) * ((f[i] - a[i]) / diag[i]);
    return std::sqrt(s);
  }
};
