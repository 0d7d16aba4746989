// hodlr.hpp -- C++ drop-in facade of the reference's HODLR API
// (Charles-Chao-Chen/fastSolver2: include/hmatrix.hpp, include/lmatrix.hpp,
// include/tree.hpp, include/matrix.hpp, include/ptr_matrix.hpp) over the C
// ABI in hodlr.h.  Class names, method names and argument meaning follow
// the reference; differences:
//   - Legion's (Context, Runtime*) pair is replaced by hodlr::Context (HIP
//     device + stream); LMatrix regions are device buffers.
//   - failures throw hodlr::Error carrying the C-ABI status instead of
//     assert() (e.g. src/tasks/leaf_solve.cc:90,141);
//   - random generators draw their per-partition seeds from a master seed
//     (hodlr::set_master_seed) instead of time(NULL) (src/matrix.cc:56,83,
//     254,275), so runs are reproducible; the fill order and LCG are the
//     reference's (src/ptr_matrix.cc:62-71);
//   - HMatrix::solve works (src/hmatrix.cc:49-96 asserts) and accepts
//     several right-hand sides (src/tree.cc:9 hard-codes one).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hodlr.h"

namespace hodlr {

// Replaces Legion's (Context ctx, Runtime* runtime) parameters.
struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
};

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& where);
  int status;
};

// Master seed for the drand48-based generators (default 0).
void set_master_seed(long seed);
long master_seed();
// The value the reference's generators take from time(NULL)
// (src/matrix.cc:56,83,254,275): while set, every Matrix/Vector::rand call
// seeds its partition stream with srand48_r(t) and draws one lrand48_r
// seed per partition, exactly as the reference does within one second of
// wall-clock time; clear_rand_clock() returns to the master-seed stream.
void set_rand_clock(long t);
void clear_rand_clock();

class Matrix;

// include/matrix.hpp:7-64
class Vector {
 public:
  Vector();
  explicit Vector(int N, bool has_entry = true);
  Vector(int base, int lvl, bool has_entry = true);
  int rows() const;
  int num_partition() const;
  int offset() const;
  double norm() const;
  void rand(int nPart, int offset);
  void rand(int offset = 0);
  long rand_seed(int) const;
  Matrix to_diag_matrix() const;
  double operator[](int i) const;
  double& operator[](int i);
  const double* data() const { return data_.data(); }
  double* data() { return data_.data(); }
  friend Vector operator+(const Vector&, const Vector&);
  friend Vector operator-(const Vector&, const Vector&);
  friend Vector operator*(double, const Vector&);
  void display(const std::string&) const;

 private:
  int nPart_ = -1, mRows_ = -1, mOffset_ = 0;
  std::vector<long> seeds_;
  std::vector<double> data_;
  bool has_entry_ = true;
};

// include/matrix.hpp:66-154 (host, column-major, data[i + j*rows])
class Matrix {
 public:
  Matrix();
  Matrix(int nRow, int nCol, bool has_entry = true);
  Matrix(int base, int level, int nCol, bool has_entry = true);
  int rows() const;
  int cols() const;
  int levels() const;
  double norm() const;
  double* pointer();
  const double* pointer() const;
  int num_partition() const;
  void rand(int nPart);
  void rand();
  long rand_seed(int) const;
  double operator()(int i, int j) const;
  double& operator()(int i, int j);
  Matrix T() const;
  Matrix block(int rlo, int rhi, int clo, int chi) const;
  Matrix row_block(int rlo, int rhi) const;
  // solve A*X = B on the device (batched dgesv with batch 1); B <- X
  void solve(Matrix& B, const Context& ctx = Context()) const;
  friend Matrix operator*(const Matrix&, const Matrix&);
  friend Matrix operator+(const Matrix&, const Matrix&);
  friend Matrix operator-(const Matrix&, const Matrix&);
  friend Matrix operator*(double, const Matrix&);
  void display(const std::string&) const;
  static Matrix identity(int n);
  static Matrix constant(int m, int n, double v);

 private:
  int mLevel_ = 0, nPart_ = 1, mRows_ = 0, mCols_ = 0;
  std::vector<long> seeds_;
  std::vector<double> data_;
  bool has_entry_ = true;
};

// include/ptr_matrix.hpp: non-owning column-major view (host or device).
class PtrMatrix {
 public:
  PtrMatrix() = default;
  PtrMatrix(int r, int c, int ld, double* p, char trans = 'n')
      : mRows(r), mCols(c), leadD(ld), ptr(p), trans(trans) {}
  int rows() const { return trans == 't' ? mCols : mRows; }
  int cols() const { return trans == 't' ? mRows : mCols; }
  int LD() const { return leadD; }
  double* pointer() const { return ptr; }
  double* pointer(int r, int c) const { return ptr + r + (int64_t)c * leadD; }
  void set_trans(char t) { trans = t; }
  char get_trans() const { return trans; }

 private:
  int mRows = -1, mCols = -1, leadD = -1;
  double* ptr = nullptr;
  char trans = 'n';
};

struct DevBuf;  // device allocation shared by LMatrix views

// include/lmatrix.hpp: a device matrix (rows x total columns, ld = rows)
// with a column window [colIdx, colIdx + cols) and a uniform row partition
// of 2^level parts; two_level_partition splits each part in two.
class LMatrix {
 public:
  LMatrix();
  LMatrix(int rows, int cols, int level, const Context& ctx);
  int rows() const { return mRows; }
  int cols() const { return mCols; }
  int rowBlk() const { return rblock; }
  int column_begin() const { return colIdx; }
  int num_partition() const { return nPart; }
  int partition_level() const { return plevel; }
  int small_block_parts() const { return smallblk; }
  void set_column_size(int n) { mCols = n; }
  void set_column_begin(int c) { colIdx = c; }
  void create(int rows, int cols, const Context& ctx);
  void clear(double value, const Context& ctx);
  void scale(double alpha, const Context& ctx);
  void init_data(const Matrix& mat, const Context& ctx);
  void init_data(int col0, int col1, const Matrix& mat, const Context& ctx);
  void init_dense_blocks(const Matrix& U, const Matrix& V, const Vector& D, const Context& ctx);
  Matrix to_matrix(const Context& ctx) const;
  Matrix to_matrix(int col0, int col1, const Context& ctx) const;
  Matrix to_matrix(int rlo, int rhi, int clo, int chi, const Context& ctx) const;
  void partition(int level, const Context& ctx);
  void two_level_partition(const Context& ctx);
  // KTree::solve -> LeafSolveTask per partition: b <- hsolve(...)
  void solve(LMatrix& b, LMatrix& V, const Context& ctx);
  // VTu.node_solve(VTd): VTd <- eta (node_solve.cc:60-81)
  void node_solve(LMatrix& b, const Context& ctx);
  static void gemmRed(char transa, char transb, double alpha, const LMatrix& A,
                      const LMatrix& B, double beta, LMatrix& C, const Context& ctx);
  static void gemmBro(char transa, char transb, double alpha, const LMatrix& A,
                      const LMatrix& B, double beta, LMatrix& C, const Context& ctx);
  void clear(const Context& ctx);  // destroy the region (lmatrix.cc:886-890)
  // device pointer to (row 0, column colIdx) and the leading dimension
  double* pointer() const;
  int64_t LD() const;

 private:
  friend class UTree;
  std::shared_ptr<DevBuf> buf;
  int mRows = 0, mCols = 0, colIdx = 0;
  int nPart = -1, rblock = 0, plevel = 1, smallblk = 1;
};

// include/tree.hpp
class UTree {
 public:
  void init(const Matrix& U, int nRhs = 1);
  void init(int level, const Matrix& U, const Context& ctx, int nRhs = 1);
  void init_rhs(const Matrix& b, const Context& ctx);
  void partition(int level, const Context& ctx);
  void horizontal_partition(int task_level, const Context& ctx);
  Matrix solution(const Context& ctx);
  LMatrix& uMat_level_new(int i);
  LMatrix& dMat_level_new(int i);
  // tree.cc:103-111: 1-based level accessors of the test drivers
  LMatrix& uMat_level(int i);
  LMatrix& dMat_level(int i);
  LMatrix& leaf() { return U; }
  void clear(const Context& ctx) { U.clear(ctx); }
  int num_rhs() const { return nRhs; }

 private:
  int mLevel = 0, rank = 0, nRhs = 1;
  Matrix UMat;
  LMatrix U;
  std::vector<LMatrix> uMat_vec, dMat_vec;
};

class VTree {
 public:
  void init(const Matrix& V);
  void init(int level, const Matrix& V, const Context& ctx);
  void partition(int level, const Context& ctx);
  void horizontal_partition(int task_level, const Context& ctx);
  LMatrix& leaf() { return V; }
  LMatrix& level(int) { return V; }
  LMatrix& level_new(int) { return V; }
  void clear(const Context& ctx) { V.clear(ctx); }

 private:
  int mLevel = 0;
  Matrix VMat;
  LMatrix V;
};

class KTree {
 public:
  void init(const Matrix& U, const Matrix& V, const Vector& D);
  void init(int level, const Matrix& U, const Matrix& V, const Vector& D, const Context& ctx);
  void partition(int level, const Context& ctx);
  void horizontal_partition(int task_level, const Context& ctx);
  void solve(LMatrix& U, LMatrix& V, const Context& ctx) { K.solve(U, V, ctx); }
  void clear(const Context& ctx) { K.clear(ctx); }

 private:
  int mLevel = 0;
  Matrix UMat, VMat;
  Vector DVec;
  LMatrix K;
};

// include/hmatrix.hpp: A = U V^T + D on a balanced tree of 2^level leaves.
class HMatrix {
 public:
  HMatrix();
  HMatrix(int nProc, int level);
  ~HMatrix();
  // build U * V' + D (hmatrix.cc:22-47)
  void init(const Matrix& U, const Matrix& V, const Vector& D, const Context& ctx = Context());
  // fast solver (hmatrix.cc:49-96): the solution of A x = b, one column
  Vector solve(const Matrix& b, const Context& ctx = Context());
  // every column of b
  Matrix solve_all(const Matrix& b, const Context& ctx = Context());
  void destroy(const Context& ctx = Context());

 private:
  int nProc = 1, level = 0, rank = 0, leaf = 0;
  int64_t N = 0;
  std::shared_ptr<DevBuf> dK, dK0, dU, dV;
  hodlr_plan plan = nullptr;
  bool factored = false;
};

}  // namespace hodlr
