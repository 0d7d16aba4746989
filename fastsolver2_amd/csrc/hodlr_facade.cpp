// hodlr_facade.cpp -- the reference's C++ HODLR API (HMatrix, LMatrix,
// UTree/VTree/KTree, Matrix/Vector) on the device, without Legion.
// Each method cites the reference method it stands in for.
#include "../../include/hodlr/hodlr.hpp"

#include <hip/hip_runtime.h>
#include <stdlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "hodlr_internal.hpp"

namespace hodlr {

// --------------------------------------------------------------- basics
Error::Error(int st, const std::string& where)
    : std::runtime_error(where + ": " + hodlr_strerror(st) +
                         (hodlr_last_error_detail()[0] ? std::string(" - ") + hodlr_last_error_detail()
                                                       : std::string())),
      status(st) {}

static void check(int st, const char* where) {
  if (st != HODLR_OK) throw Error(st, where);
}
static void hcheck(hipError_t e, const char* where) {
  if (e != hipSuccess) throw Error(HODLR_ERR_HIP, std::string(where) + " (" + hipGetErrorString(e) + ")");
}
static void require(bool ok, const char* what) {
  if (!ok) throw Error(HODLR_ERR_ARG, what);
}

struct DevBuf {
  double* p = nullptr;
  size_t n = 0;
  int device = 0;
  DevBuf(size_t count, int dev) : n(count), device(dev) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(dev);
    hipError_t e = hipMalloc(&p, sizeof(double) * (count ? count : 1));
    (void)hipSetDevice(prev);
    hcheck(e, "hipMalloc");
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

static std::shared_ptr<DevBuf> upload(const double* h, size_t n, const Context& ctx) {
  auto b = std::make_shared<DevBuf>(n, ctx.device);
  hcheck(hipMemcpyAsync(b->p, h, sizeof(double) * n, hipMemcpyHostToDevice, ctx.stream), "upload");
  hcheck(hipStreamSynchronize(ctx.stream), "upload sync");
  return b;
}

// The reference seeds every generator from time(NULL) (src/matrix.cc:56,83,
// 254,275); here one drand48 stream seeded by the master seed supplies the
// per-call base seed, so runs are reproducible.
static std::mutex g_seed_mu;
static long g_master = 0;
static struct drand48_data g_stream;
static bool g_stream_init = false;

void set_master_seed(long seed) {
  std::lock_guard<std::mutex> lk(g_seed_mu);
  g_master = seed;
  srand48_r(seed, &g_stream);
  g_stream_init = true;
}
long master_seed() { return g_master; }

static long next_base_seed() {
  std::lock_guard<std::mutex> lk(g_seed_mu);
  if (!g_stream_init) {
    srand48_r(g_master, &g_stream);
    g_stream_init = true;
  }
  long v = 0;
  lrand48_r(&g_stream, &v);
  return g_master + v;
}

static bool g_clock_set = false;
static long g_clock = 0;
void set_rand_clock(long t) {
  std::lock_guard<std::mutex> lk(g_seed_mu);
  g_clock = t;
  g_clock_set = true;
}
void clear_rand_clock() {
  std::lock_guard<std::mutex> lk(g_seed_mu);
  g_clock_set = false;
}
static long call_seed() {
  {
    std::lock_guard<std::mutex> lk(g_seed_mu);
    if (g_clock_set) return g_clock;
  }
  return next_base_seed();
}

// matrix.cc:254-262: srand48_r(time) then one lrand48_r per partition.
static std::vector<long> partition_seeds(int nPart) {
  struct drand48_data buf;
  srand48_r(call_seed(), &buf);
  std::vector<long> s(nPart);
  for (int i = 0; i < nPart; ++i) lrand48_r(&buf, &s[i]);
  return s;
}

// PtrMatrix::rand (ptr_matrix.cc:62-71): rows outer, columns inner.
static void ptr_rand(long seed, int offset, int rows, int cols, int64_t ld, double* a) {
  struct drand48_data buf;
  srand48_r(seed, &buf);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      drand48_r(&buf, &a[i + (int64_t)j * ld]);
      a[i + (int64_t)j * ld] += offset;
    }
}

// --------------------------------------------------------------- Vector
Vector::Vector() = default;
Vector::Vector(int N, bool has) : nPart_(-1), mRows_(N), has_entry_(has) {
  require(N > 0, "Vector(N): N > 0");
  if (has_entry_) data_.assign(N, 0.0);
}
Vector::Vector(int base, int lvl, bool has) : nPart_(1 << lvl), mRows_(base << lvl), has_entry_(has) {
  require(base > 0 && lvl >= 0, "Vector(base, lvl)");
  if (has_entry_) data_.assign(mRows_, 0.0);
}
int Vector::rows() const { return mRows_; }
int Vector::num_partition() const { return nPart_; }
int Vector::offset() const { return mOffset_; }
double Vector::norm() const {
  double s = 0.0;
  for (double v : data_) s += v * v;
  return std::sqrt(s);
}
// matrix.cc:51-77 / :79-102
void Vector::rand(int nPart, int offset) {
  require(nPart > 0 && mRows_ % nPart == 0, "Vector::rand: rows % nPart");
  nPart_ = nPart;
  rand(offset);
}
void Vector::rand(int offset) {
  require(nPart_ > 0 && mRows_ % nPart_ == 0, "Vector::rand: partition");
  mOffset_ = offset;
  seeds_ = partition_seeds(nPart_);
  if (!has_entry_) return;
  const int blk = mRows_ / nPart_;
  for (int k = 0; k < nPart_; ++k) ptr_rand(seeds_[k], offset, blk, 1, blk, data_.data() + (int64_t)k * blk);
}
long Vector::rand_seed(int i) const {
  require(i >= 0 && i < (int)seeds_.size(), "Vector::rand_seed");
  return seeds_[i];
}
Matrix Vector::to_diag_matrix() const {
  Matrix m = Matrix::constant(mRows_, mRows_, 0.0);
  for (int i = 0; i < mRows_; ++i) m(i, i) = data_[i];
  return m;
}
double Vector::operator[](int i) const { return data_.at(i); }
double& Vector::operator[](int i) { return data_.at(i); }
Vector operator+(const Vector& a, const Vector& b) {
  require(a.rows() == b.rows(), "Vector +");
  Vector c(a.rows());
  for (int i = 0; i < a.rows(); ++i) c[i] = a[i] + b[i];
  return c;
}
Vector operator-(const Vector& a, const Vector& b) {
  require(a.rows() == b.rows(), "Vector -");
  Vector c(a.rows());
  for (int i = 0; i < a.rows(); ++i) c[i] = a[i] - b[i];
  return c;
}
Vector operator*(double s, const Vector& a) {
  Vector c(a.rows());
  for (int i = 0; i < a.rows(); ++i) c[i] = s * a[i];
  return c;
}
void Vector::display(const std::string& name) const {
  printf("%s:\n", name.c_str());
  for (double v : data_) printf("%g\n", v);
}

// --------------------------------------------------------------- Matrix
Matrix::Matrix() = default;
Matrix::Matrix(int r, int c, bool has) : mLevel_(0), nPart_(1), mRows_(r), mCols_(c), has_entry_(has) {
  require(r > 0 && c > 0, "Matrix(rows, cols)");
  if (has_entry_) data_.assign((size_t)r * c, 0.0);
}
// matrix.cc:217-230 (asserts base > col)
Matrix::Matrix(int base, int level, int col, bool has)
    : mLevel_(level), nPart_(1 << level), mRows_(base << level), mCols_(col), has_entry_(has) {
  require(base > 0 && col > 0 && level >= 0, "Matrix(base, level, col)");
  require(base > col, "Matrix(base, level, col): base > col (matrix.cc:225)");
  if (has_entry_) data_.assign((size_t)mRows_ * col, 0.0);
}
int Matrix::rows() const { return mRows_; }
int Matrix::cols() const { return mCols_; }
int Matrix::levels() const { return mLevel_; }
double Matrix::norm() const {
  double s = 0.0;
  for (double v : data_) s += v * v;
  return std::sqrt(s);
}
double* Matrix::pointer() { return data_.data(); }
const double* Matrix::pointer() const { return data_.data(); }
int Matrix::num_partition() const { return nPart_; }
// matrix.cc:250-297
void Matrix::rand(int nPart) {
  require(nPart > 0 && mRows_ % nPart == 0, "Matrix::rand: rows % nPart");
  nPart_ = nPart;
  rand();
}
void Matrix::rand() {
  require(nPart_ > 0 && mRows_ % nPart_ == 0, "Matrix::rand: partition");
  seeds_ = partition_seeds(nPart_);
  if (!has_entry_) return;
  const int nrow = mRows_ / nPart_;
  for (int k = 0; k < nPart_; ++k) ptr_rand(seeds_[k], 0, nrow, mCols_, mRows_, data_.data() + (int64_t)k * nrow);
}
long Matrix::rand_seed(int i) const {
  require(i >= 0 && i < (int)seeds_.size(), "Matrix::rand_seed");
  return seeds_[i];
}
double Matrix::operator()(int i, int j) const { return data_.at(i + (size_t)j * mRows_); }
double& Matrix::operator()(int i, int j) { return data_.at(i + (size_t)j * mRows_); }
Matrix Matrix::T() const {
  Matrix t(mCols_, mRows_);
  for (int j = 0; j < mCols_; ++j)
    for (int i = 0; i < mRows_; ++i) t(j, i) = (*this)(i, j);
  return t;
}
Matrix Matrix::block(int rlo, int rhi, int clo, int chi) const {
  require(0 <= rlo && rlo < rhi && rhi <= mRows_ && 0 <= clo && clo < chi && chi <= mCols_, "Matrix::block");
  Matrix b(rhi - rlo, chi - clo);
  for (int j = clo; j < chi; ++j)
    for (int i = rlo; i < rhi; ++i) b(i - rlo, j - clo) = (*this)(i, j);
  return b;
}
Matrix Matrix::row_block(int rlo, int rhi) const { return block(rlo, rhi, 0, mCols_); }
void Matrix::solve(Matrix& B, const Context& ctx) const {
  require(mRows_ == mCols_ && B.rows() == mRows_, "Matrix::solve shapes");
  auto dA = upload(data_.data(), data_.size(), ctx);
  auto dB = upload(B.pointer(), (size_t)B.rows() * B.cols(), ctx);
  int* ipiv = nullptr;
  hcheck(hipMalloc(&ipiv, sizeof(int) * (mRows_ + 1)), "hipMalloc ipiv");
  int* dinfo = ipiv + mRows_;
  int st = hodlr_dgesv_batched(mRows_, B.cols(), dA->p, mRows_, 0, 1, ipiv, dinfo, dB->p, mRows_, 0,
                               ctx.stream);
  int info = 0;
  if (st == HODLR_OK) {
    hcheck(hipStreamSynchronize(ctx.stream), "Matrix::solve sync");
    hcheck(hipMemcpy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost), "Matrix::solve info");
  }
  (void)hipFree(ipiv);
  check(st, "Matrix::solve (dgesv, matrix.cc:340-350)");
  check(info ? HODLR_SINGULAR_LEAF : HODLR_OK, "Matrix::solve: singular (matrix.cc:343)");
  hcheck(hipMemcpy(B.pointer(), dB->p, sizeof(double) * B.rows() * B.cols(), hipMemcpyDeviceToHost),
         "download");
}
Matrix operator*(const Matrix& a, const Matrix& b) {
  require(a.cols() == b.rows(), "Matrix *");
  Matrix c(a.rows(), b.cols());
  for (int j = 0; j < b.cols(); ++j)
    for (int k = 0; k < a.cols(); ++k) {
      const double bk = b(k, j);
      for (int i = 0; i < a.rows(); ++i) c(i, j) += a(i, k) * bk;
    }
  return c;
}
Matrix operator+(const Matrix& a, const Matrix& b) {
  require(a.rows() == b.rows() && a.cols() == b.cols(), "Matrix +");
  Matrix c(a.rows(), a.cols());
  for (int j = 0; j < a.cols(); ++j)
    for (int i = 0; i < a.rows(); ++i) c(i, j) = a(i, j) + b(i, j);
  return c;
}
Matrix operator-(const Matrix& a, const Matrix& b) {
  require(a.rows() == b.rows() && a.cols() == b.cols(), "Matrix -");
  Matrix c(a.rows(), a.cols());
  for (int j = 0; j < a.cols(); ++j)
    for (int i = 0; i < a.rows(); ++i) c(i, j) = a(i, j) - b(i, j);
  return c;
}
Matrix operator*(double s, const Matrix& a) {
  Matrix c(a.rows(), a.cols());
  for (int j = 0; j < a.cols(); ++j)
    for (int i = 0; i < a.rows(); ++i) c(i, j) = s * a(i, j);
  return c;
}
void Matrix::display(const std::string& name) const {
  printf("%s:\n", name.c_str());
  for (int i = 0; i < mRows_; ++i) {
    for (int j = 0; j < mCols_; ++j) printf("%g\t", (*this)(i, j));
    printf("\n");
  }
}
Matrix Matrix::identity(int n) {
  Matrix m(n, n);
  for (int i = 0; i < n; ++i) m(i, i) = 1.0;
  return m;
}
Matrix Matrix::constant(int m, int n, double v) {
  Matrix a(m, n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) a(i, j) = v;
  return a;
}

// --------------------------------------------------------------- LMatrix
LMatrix::LMatrix() = default;
LMatrix::LMatrix(int rows, int cols, int level, const Context& ctx) {
  create(rows, cols, ctx);
  partition(level, ctx);
}
double* LMatrix::pointer() const { return buf->p + (int64_t)colIdx * mRows; }
int64_t LMatrix::LD() const { return mRows; }

// lmatrix.cc:78-97 (region creation); the region starts zeroed
void LMatrix::create(int rows, int cols, const Context& ctx) {
  require(rows > 0 && cols > 0, "LMatrix::create");
  mRows = rows; mCols = cols; colIdx = 0;
  buf = std::make_shared<DevBuf>((size_t)rows * cols, ctx.device);
  hcheck(launch_fill_scale(rows, cols, buf->p, rows, 0, 0.0, ctx.stream), "LMatrix::create fill");
}
// LMatrix::clear(value) / ClearMatrixTask
void LMatrix::clear(double value, const Context& ctx) {
  hcheck(launch_fill_scale(mRows, mCols, pointer(), mRows, 0, value, ctx.stream), "LMatrix::clear");
}
// LMatrix::scale (lmatrix.cc:117-136): multiplies, alpha == 1 is a no-op
void LMatrix::scale(double alpha, const Context& ctx) {
  if (std::fabs(alpha - 1.0) < 1e-10) return;
  hcheck(launch_fill_scale(mRows, mCols, pointer(), mRows, 1, alpha, ctx.stream), "LMatrix::scale");
}
void LMatrix::init_data(const Matrix& mat, const Context& ctx) {
  require(mCols >= mat.cols(), "LMatrix::init_data cols");
  init_data(0, mat.cols(), mat, ctx);
}
// lmatrix.cc:144-176 with InitMatrixTask (init_matrix.cc:57-72): the matrix
// is written into every mat.cols()-wide column block of [col0, col1).
void LMatrix::init_data(int col0, int col1, const Matrix& mat, const Context& ctx) {
  require(col0 >= 0 && col1 <= mCols && col1 - col0 >= mat.cols() && (col1 - col0) % mat.cols() == 0,
          "LMatrix::init_data column range (lmatrix.cc:146-149)");
  require(mat.rows() == mRows, "LMatrix::init_data rows");
  require(nPart > 0 && mat.num_partition() % nPart == 0, "LMatrix::init_data partitions");
  smallblk = mat.num_partition() / nPart;
  auto dm = upload(mat.pointer(), (size_t)mat.rows() * mat.cols(), ctx);
  for (int c = col0; c + mat.cols() <= col1; c += mat.cols())
    hcheck(launch_copy_cols(mRows, mat.cols(), dm->p, mRows, pointer() + (int64_t)c * mRows, mRows,
                            ctx.stream),
           "LMatrix::init_data copy");
  hcheck(hipStreamSynchronize(ctx.stream), "LMatrix::init_data sync");
}
// lmatrix.cc:273-297 + DenseBlockTask (dense_block.cc:54-75)
void LMatrix::init_dense_blocks(const Matrix& U, const Matrix& V, const Vector& D, const Context& ctx) {
  require(U.cols() == V.cols() && U.rows() == V.rows() && U.rows() == D.rows() && U.rows() == mRows,
          "LMatrix::init_dense_blocks shapes");
  const int leaf = mCols;
  require(mRows % leaf == 0, "LMatrix::init_dense_blocks: rows % leaf");
  auto dU = upload(U.pointer(), (size_t)U.rows() * U.cols(), ctx);
  auto dV = upload(V.pointer(), (size_t)V.rows() * V.cols(), ctx);
  auto dD = upload(D.data(), (size_t)D.rows(), ctx);
  check(hodlr_dense_blocks(mRows, leaf, U.cols(), dU->p, mRows, dV->p, mRows, dD->p, 0.0, pointer(),
                           mRows, ctx.stream),
        "LMatrix::init_dense_blocks");
  hcheck(hipStreamSynchronize(ctx.stream), "init_dense_blocks sync");
}
Matrix LMatrix::to_matrix(const Context& ctx) const { return to_matrix(0, mRows, 0, mCols, ctx); }
Matrix LMatrix::to_matrix(int col0, int col1, const Context& ctx) const {
  return to_matrix(0, mRows, col0, col1, ctx);
}
Matrix LMatrix::to_matrix(int rlo, int rhi, int clo, int chi, const Context& ctx) const {
  require(0 <= rlo && rlo < rhi && rhi <= mRows && 0 <= clo && clo < chi && chi <= mCols, "LMatrix::to_matrix");
  hcheck(hipStreamSynchronize(ctx.stream), "to_matrix sync");
  Matrix m(rhi - rlo, chi - clo);
  for (int j = clo; j < chi; ++j)
    hcheck(hipMemcpy(m.pointer() + (int64_t)(j - clo) * (rhi - rlo), pointer() + rlo + (int64_t)j * mRows,
                     sizeof(double) * (rhi - rlo), hipMemcpyDeviceToHost),
           "to_matrix copy");
  return m;
}
// lmatrix.cc:436-446: uniform row partition into 2^level parts
void LMatrix::partition(int level, const Context&) {
  require(level >= 0 && mRows % (1 << level) == 0, "LMatrix::partition");
  nPart = 1 << level;
  rblock = mRows / nPart;
  plevel = 1;
}
// lmatrix.cc:495-523
void LMatrix::two_level_partition(const Context&) {
  require(nPart > 0 && rblock % 2 == 0, "LMatrix::two_level_partition");
  plevel = 2;
  rblock /= 2;
  nPart *= 2;
}
// LMatrix::solve (lmatrix.cc:449-493) + LeafSolveTask (leaf_solve.cc:38-71)
void LMatrix::solve(LMatrix& b, LMatrix& V, const Context& ctx) {
  require(mRows == b.rows() && mRows == V.rows() && b.cols() > 0, "LMatrix::solve shapes");
  require(b.num_partition() == nPart, "LMatrix::solve partitions");
  const int nps = V.small_block_parts();
  int lvl = 0;
  while ((1 << lvl) < nps) ++lvl;
  require((1 << lvl) == nps, "LMatrix::solve: leaves per task must be a power of two");
  const int rank = V.cols();
  for (int p = 0; p < nPart; ++p) {
    const int64_t rlo = (int64_t)p * rblock;
    int st = hodlr_hsolve_ref(rblock, b.cols() - lvl * rank, rank, nps, mRows, pointer() + rlo,
                              b.pointer() + rlo, V.pointer() + rlo, ctx.stream);
    check(st, "LMatrix::solve (leaf_solve.cc:69)");
  }
}
// LMatrix::node_solve (lmatrix.cc:534-581)
void LMatrix::node_solve(LMatrix& b, const Context& ctx) {
  require(nPart == b.num_partition(), "LMatrix::node_solve partitions");
  const int rowBlk = rblock * plevel;
  require(rowBlk / 2 == mCols, "LMatrix::node_solve: rowBlk/2 == cols (lmatrix.cc:551)");
  const int nnode = nPart / plevel;
  int* dinfo = nullptr;
  hcheck(hipMalloc(&dinfo, sizeof(int) * nnode), "hipMalloc info");
  int st = hodlr_node_solve(nnode, mCols, pointer(), mRows, b.pointer(), b.mRows, b.cols(), dinfo,
                            ctx.stream);
  std::vector<int> info(nnode, 0);
  if (st == HODLR_OK) {
    hcheck(hipStreamSynchronize(ctx.stream), "LMatrix::node_solve sync");
    hcheck(hipMemcpy(info.data(), dinfo, sizeof(int) * nnode, hipMemcpyDeviceToHost),
           "LMatrix::node_solve info");
  }
  (void)hipFree(dinfo);
  check(st, "LMatrix::node_solve");
  for (int k = 0; k < nnode; ++k)
    check(info[k] ? HODLR_SINGULAR_CORE : HODLR_OK, "LMatrix::node_solve: singular core (node_solve.cc:78)");
}
// LMatrix::gemmRed (lmatrix.cc:713-761): only ('t','n', beta = 0) is used
void LMatrix::gemmRed(char ta, char tb, double alpha, const LMatrix& A, const LMatrix& B, double beta,
                      LMatrix& C, const Context& ctx) {
  require((ta == 't' || ta == 'T') && (tb == 'n' || tb == 'N'), "gemmRed: op must be ('t','n')");
  require(std::fabs(beta) < 1e-10, "gemmRed: beta == 0 (lmatrix.cc:718)");
  require(A.num_partition() == B.num_partition() && A.num_partition() % C.num_partition() == 0,
          "gemmRed partitions (lmatrix.cc:723-724)");
  require(C.rowBlk() == A.cols() && C.cols() == B.cols(), "gemmRed slot shape");
  int st = hodlr_gemm_red(A.rows(), A.num_partition(), C.num_partition(), alpha, A.pointer(), A.LD(),
                          A.cols(), B.pointer(), B.LD(), B.cols(), C.pointer(), C.LD(), ctx.stream);
  check(st, "LMatrix::gemmRed");
}
// LMatrix::gemmBro (lmatrix.cc:814-865): C is A's region, beta = 1
void LMatrix::gemmBro(char ta, char tb, double alpha, const LMatrix& A, const LMatrix& B, double beta,
                      LMatrix& C, const Context& ctx) {
  require((ta == 'n' || ta == 'N') && (tb == 'n' || tb == 'N'), "gemmBro: op must be ('n','n')");
  require(std::fabs(beta - 1.0) < 1e-10, "gemmBro: beta == 1 (lmatrix.cc:820)");
  require(A.num_partition() == C.num_partition() && A.num_partition() % B.num_partition() == 0,
          "gemmBro partitions (lmatrix.cc:824-825)");
  require(B.rowBlk() == A.cols(), "gemmBro slot shape");
  int st = hodlr_gemm_bro(A.rows(), A.num_partition(), B.num_partition(), alpha, A.pointer(), A.LD(),
                          A.cols(), B.pointer(), B.LD(), C.pointer(), C.LD(), C.cols(), ctx.stream);
  check(st, "LMatrix::gemmBro");
}
void LMatrix::clear(const Context&) { buf.reset(); }

// --------------------------------------------------------------- trees
// UTree (tree.cc:5-121): region N x (nRhs + r*L); cols [0,nRhs) rhs, then
// one r-wide copy of U per level; d_i = cols [0, nRhs + r*i),
// u_i = cols [nRhs + r*i, nRhs + r*(i+1)).
void UTree::init(const Matrix& U, int nr) {
  UMat = U; rank = U.cols(); nRhs = nr;
}
void UTree::init(int level, const Matrix& U, const Context& ctx, int nr) {
  mLevel = level; UMat = U; rank = U.cols(); nRhs = nr;
  this->U.create(U.rows(), nRhs + rank * mLevel, ctx);
}
void UTree::init_rhs(const Matrix& b, const Context& ctx) {
  require(b.cols() == nRhs, "UTree::init_rhs: b.cols() == nRhs");
  U.init_data(0, b.cols(), b, ctx);
}
static void build_views(LMatrix& U, int mLevel, int nRhs, int rank, std::vector<LMatrix>& dv,
                        std::vector<LMatrix>& uv) {
  dv.clear(); uv.clear();
  for (int i = 0; i < mLevel; ++i) {
    const int ncol = nRhs + rank * i;
    LMatrix d = U;
    d.set_column_size(ncol);
    dv.push_back(d);
    LMatrix u = U;
    u.set_column_size(rank);
    u.set_column_begin(ncol);
    uv.push_back(u);
  }
}
// tree.cc:42-76: the region holds one u copy per matrix level
// (UMat.levels()); the level views cover the top `level` levels, the leaf
// tasks' hsolve the rest.
void UTree::partition(int level, const Context& ctx) {
  require(UMat.rows() > 0 && UMat.cols() > 0, "UTree::partition");
  require(level >= 0 && level <= UMat.levels(), "UTree::partition: level <= matrix levels");
  mLevel = level;
  U.create(UMat.rows(), nRhs + rank * UMat.levels(), ctx);
  U.partition(mLevel, ctx);
  U.init_data(nRhs, U.cols(), UMat, ctx);
  build_views(U, mLevel, nRhs, rank, dMat_vec, uMat_vec);
}
void UTree::horizontal_partition(int task_level, const Context& ctx) {
  U.partition(task_level, ctx);
  U.init_data(nRhs, U.cols(), UMat, ctx);
  build_views(U, mLevel, nRhs, rank, dMat_vec, uMat_vec);
}
LMatrix& UTree::uMat_level_new(int i) {
  require(0 <= i && i < mLevel, "UTree::uMat_level_new");
  return uMat_vec[i];
}
LMatrix& UTree::dMat_level_new(int i) {
  require(0 <= i && i < mLevel, "UTree::dMat_level_new");
  return dMat_vec[i];
}
LMatrix& UTree::uMat_level(int i) {
  require(0 < i && i <= mLevel, "UTree::uMat_level");
  return uMat_vec[i - 1];
}
LMatrix& UTree::dMat_level(int i) {
  require(0 < i && i <= mLevel, "UTree::dMat_level");
  return dMat_vec[i - 1];
}
Matrix UTree::solution(const Context& ctx) { return U.to_matrix(0, nRhs, ctx); }

void VTree::init(const Matrix& V) { VMat = V; }
void VTree::init(int level, const Matrix& V, const Context& ctx) {
  mLevel = level; VMat = V;
  this->V.create(V.rows(), V.cols(), ctx);
}
void VTree::partition(int level, const Context& ctx) {
  mLevel = level;
  V.create(VMat.rows(), VMat.cols(), ctx);
  V.partition(level, ctx);
  V.init_data(VMat, ctx);
}
void VTree::horizontal_partition(int task_level, const Context& ctx) {
  V.partition(task_level, ctx);
  V.init_data(VMat, ctx);
}

void KTree::init(const Matrix& U, const Matrix& V, const Vector& D) {
  UMat = U; VMat = V; DVec = D;
  require(U.rows() == V.rows() && U.cols() == V.cols() && U.rows() == D.rows(), "KTree::init shapes");
}
void KTree::init(int level, const Matrix& U, const Matrix& V, const Vector& D, const Context& ctx) {
  init(U, V, D);
  mLevel = level;
  const int nblk = 1 << U.levels();
  K.create(U.rows(), U.rows() / nblk, ctx);
}
void KTree::partition(int level, const Context& ctx) {
  const int nblk = 1 << UMat.levels();
  const int ncol = DVec.rows() / nblk;
  require(ncol > 0, "KTree::partition");
  K.create(DVec.rows(), ncol, ctx);
  mLevel = level;
  K.partition(level, ctx);
  K.init_dense_blocks(UMat, VMat, DVec, ctx);
}
void KTree::horizontal_partition(int task_level, const Context& ctx) {
  K.partition(task_level, ctx);
  K.init_dense_blocks(UMat, VMat, DVec, ctx);
}

// --------------------------------------------------------------- HMatrix
HMatrix::HMatrix() = default;
// hmatrix.cc:7-20
HMatrix::HMatrix(int nProc_, int level_) : nProc(nProc_), level(level_) {
  require(nProc > 0 && level >= 0 && (1 << level) % nProc == 0, "HMatrix: 2^level % nProc == 0");
}
HMatrix::~HMatrix() {
  if (plan) (void)hodlr_plan_destroy(plan);
}
// hmatrix.cc:22-47: A = U V^T + diag(D) with dense leaf blocks
void HMatrix::init(const Matrix& U, const Matrix& V, const Vector& D, const Context& ctx) {
  require(U.rows() == V.rows() && U.rows() == D.rows() && U.cols() == V.cols() && U.cols() > 0,
          "HMatrix::init shapes (hmatrix.cc:27-31)");
  N = U.rows();
  rank = U.cols();
  require(N % (1 << level) == 0, "HMatrix::init: rows % 2^level");
  leaf = (int)(N >> level);
  require(level == 0 || leaf > rank, "HMatrix::init: leaf > rank (matrix.cc:225)");
  dU = upload(U.pointer(), (size_t)N * rank, ctx);
  dV = upload(V.pointer(), (size_t)N * rank, ctx);
  auto dD = upload(D.data(), (size_t)N, ctx);
  dK0 = std::make_shared<DevBuf>((size_t)N * leaf, ctx.device);
  dK = std::make_shared<DevBuf>((size_t)N * leaf, ctx.device);
  check(hodlr_dense_blocks(N, leaf, rank, dU->p, N, dV->p, N, dD->p, 0.0, dK0->p, N, ctx.stream),
        "HMatrix::init dense blocks");
  hcheck(hipStreamSynchronize(ctx.stream), "HMatrix::init sync");
  factored = false;
}
Matrix HMatrix::solve_all(const Matrix& b, const Context& ctx) {
  require(dK0 != nullptr, "HMatrix::solve before init");
  require(b.rows() == N && b.cols() > 0, "HMatrix::solve: b rows");
  if (!plan) {
    hodlr_desc d{N, leaf, level, rank, 16, HODLR_FP64, ctx.device, 0};
    check(hodlr_plan_create(&plan, &d), "HMatrix plan");
  }
  if (!factored) {
    hcheck(hipMemcpyAsync(dK->p, dK0->p, sizeof(double) * (size_t)N * leaf, hipMemcpyDeviceToDevice,
                          ctx.stream),
           "copy K");
    check(hodlr_factor(plan, dK->p, N, dU->p, N, dV->p, N, ctx.stream), "HMatrix factor");
    int where = -1;
    check(hodlr_plan_status(plan, ctx.stream, &where, nullptr), "HMatrix factor status");
    factored = true;
  }
  auto dB = upload(b.pointer(), (size_t)N * b.cols(), ctx);
  check(hodlr_solve(plan, dB->p, N, b.cols(), ctx.stream), "HMatrix::solve");
  Matrix x(b.rows(), b.cols());
  hcheck(hipStreamSynchronize(ctx.stream), "HMatrix::solve sync");
  hcheck(hipMemcpy(x.pointer(), dB->p, sizeof(double) * (size_t)N * b.cols(), hipMemcpyDeviceToHost),
         "HMatrix::solve download");
  return x;
}
Vector HMatrix::solve(const Matrix& b, const Context& ctx) {
  Matrix x = solve_all(b, ctx);
  Vector v((int)N);
  for (int64_t i = 0; i < N; ++i) v[(int)i] = x((int)i, 0);
  return v;
}
void HMatrix::destroy(const Context&) {
  if (plan) (void)hodlr_plan_destroy(plan);
  plan = nullptr;
  dK.reset(); dK0.reset(); dU.reset(); dV.reset();
  factored = false;
}

}  // namespace hodlr
