// mfs_cssd.hpp - drop-in replacement for Library/CSSD_Matrix.h of
// AnastasiaKutishcheva/MultiscaleFEMSolver3D, backed by libmfs.so.
//
// The class keeps the reference's public data members and method names
// (CSSD_Matrix.h:7-1990), because the drivers write Diag / A_up / A_down / X
// / F directly (FEM.h:517-528, :2202-2216, MsFEM.h:904-915).  Host-side
// methods (Initialization, SummPartOfMatrix, MultiplicationMatrixVector, ...)
// run on the CPU with the reference's semantics; the solver methods pack the
// rows into CSR (double) or BSR3 (3x3 blocks), solve on the GPU through the
// mfs_* C-ABI, write X back and return the value the reference returns:
//   MSG, MSG_PreconditioningSSOR, GMRES, BiCG_Stab*   true ||F-AX||/||F||
//   BiCG_Diag                                          ||D^-1 (F-AX)|| (absolute)
// Method mapping (the reference's own algorithms are replaced):
//   MSG                      -> CG (FP64)
//   MSG_PreconditioningSSOR  -> Jacobi-PCG (FP64); the SSOR factor is unused
//   BiCG_Diag                -> Jacobi-PCG (the matrices are symmetric)
//   GMRES                    -> GMRES(30) with Jacobi, FP64
//   BiCG_Stab family         -> BiCGStab with right Jacobi preconditioning, FP64
// Set MFS_CSSD_MIXED=1 in the environment to run the mixed-precision path.
//
// Include it where the reference includes Library/CSSD_Matrix.h, after the
// reference's DenseMatrix / Tensor2Rank3D / Point types.  Link -lmfs.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
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

// Scalar access to the reference's value types: double, or a 3x3 tensor with
// .val[a][b] and a point with .x/.y/.z.
template <typename T>
struct is_scalar_value : std::is_arithmetic<T> {};

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

  // CSSD_Matrix.h:79-99
  void Initialization(std::vector<std::vector<int>>& up, std::vector<std::vector<int>>& down) {
    const int n = (int)up.size();
    id_column_for_A_up = up;
    id_column_for_A_down = down;
    Diag.assign(n, zero_m());
    X.assign(n, zero_v());
    F.assign(n, zero_v());
    A_up.resize(n);
    A_down.resize(n);
    for (int i = 0; i < n; i++) {
      A_up[i].assign(up[i].size(), zero_m());
      A_down[i].assign(down[i].size(), zero_m());
    }
  }

  void UpdateEquation(int id, TForMatrix& diag, TForMatrix& non_diag, TForVector& x, TForVector& f) {
    Diag[id] = diag;
    X[id] = x;
    F[id] = f;
    for (auto& v : A_up[id]) v = non_diag;
    for (auto& v : A_down[id]) v = non_diag;
  }

  // CSSD_Matrix.h:172-179 (lower part from the upper one)
  void Symmetrization() {
    id_column_for_A_down.assign(Diag.size(), {});
    A_down.assign(Diag.size(), {});
    for (size_t i = 0; i < id_column_for_A_up.size(); i++)
      for (size_t jj = 0; jj < id_column_for_A_up[i].size(); jj++) {
        const int j = id_column_for_A_up[i][jj];
        id_column_for_A_down[j].push_back((int)i);
        A_down[j].push_back(A_up[i][jj]);
      }
  }

  // CSSD_Matrix.h:180-241
  bool SetValue(int row, int col, TForMatrix v) {
    TForMatrix* p = find(row, col);
    if (!p) return false;
    *p = v;
    return true;
  }
  TForMatrix GetValue(int row, int col) {
    TForMatrix* p = find(row, col);
    return p ? *p : zero_m();
  }

  // CSSD_Matrix.h:307-363: scatter-add a local matrix; entries missing from
  // the portrait are reported like the reference does.
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
        TForMatrix* p = find(s, c);
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

  // CSSD_Matrix.h:395-423 (host)
  void MultiplicationMatrixVector(std::vector<TForVector>& B, std::vector<TForVector>& result) {
    const int n = (int)Diag.size();
    result.resize(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      TForVector t = zero_v();
      for (size_t j = 0; j < A_up[i].size(); j++) t += mul(A_up[i][j], B[id_column_for_A_up[i][j]]);
      for (size_t j = 0; j < A_down[i].size(); j++) t += mul(A_down[i][j], B[id_column_for_A_down[i][j]]);
      t += mul(Diag[i], B[i]);
      result[i] = t;
    }
  }

  // CSSD_Matrix.h:458-513: kept for the call surface; the GPU solvers use a
  // Jacobi preconditioner, so only the matrix copy is made.
  void PrecondorSSOR(double, CSSD_Matrix<double, double>& A) {
    Diag = A.Diag;
    A_up = A.A_up;
    A_down = A.A_down;
    id_column_for_A_up = A.id_column_for_A_up;
    id_column_for_A_down = A.id_column_for_A_down;
  }

  // ---- solvers (GPU) -----------------------------------------------------
  double MSG(int maxiter, double E) { return solve_inplace(MFS_CG, maxiter, E, X, F); }
  double MSG(int maxiter, double E, std::vector<TForVector>& X_new, std::vector<TForVector>& F_new, bool) {
    return solve_inplace(MFS_CG, maxiter, E, X_new, F_new);
  }
  double MSG_PreconditioningSSOR(int maxiter, double E, CSSD_Matrix<double, double>&) {
    return solve_inplace(MFS_PCG_JACOBI, maxiter, E, X, F);
  }
  double MSG_PreconditioningSSOR(std::vector<double>& RightSide, std::vector<double>& Result, int maxiter, double E,
                                 CSSD_Matrix<double, double>&) {
    return solve_inplace(MFS_PCG_JACOBI, maxiter, E, Result, RightSide);
  }
  // CSSD_Matrix.h:1181-1300 returns the absolute preconditioned norm.
  double BiCG_Diag(int maxiter, double E, std::vector<double>& diag) {
    // stop on ||D^-1 r|| < E: use a relative target on ||r|| that implies it
    const double dn = norm_scaled_rhs(diag);
    const double rel = dn > 0 ? E / dn : E;
    solve_inplace(MFS_PCG_JACOBI, maxiter, rel, X, F);
    return precond_residual(diag);
  }
  double GMRES(int maxiter, double E) { return solve_inplace(MFS_GMRES, maxiter, E, X, F); }
  double BiCG_Stab(int maxiter, double E) { return solve_inplace(MFS_BICGSTAB, maxiter, E, X, F); }
  double BCG_Stab2(int maxiter, double E) { return solve_inplace(MFS_BICGSTAB, maxiter, E, X, F); }
  double BiCG_Stab_blocks(int maxiter, double E) { return solve_inplace(MFS_BICGSTAB, maxiter, E, X, F); }

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
  static TForVector mul(TForMatrix a, const TForVector& b) {
    if constexpr (kScalar)
      return a * b;
    else
      return a * b;  // Tensor2Rank3D::operator*(Point) (Tensor.h:211-226)
  }

  TForMatrix* find(int row, int col) {
    if (row == col) return &Diag[row];
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
