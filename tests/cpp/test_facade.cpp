// C++ facade tests, after the reference's tests/test.cc: the same checks on
// the same operators, through hodlr::LMatrix / HMatrix (include/hodlr/hodlr.hpp)
// in place of Legion.  Exit status 0 when every test passes; each prints
// "Test for ... passed!" as the reference's tests do.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "hodlr/hodlr.hpp"

using hodlr::Context;
using hodlr::HMatrix;
using hodlr::LMatrix;
using hodlr::Matrix;
using hodlr::Vector;

static int g_fail = 0;

static void report(bool ok, const std::string& what, double res) {
  if (ok) {
    std::printf("Test for %s passed! (residual %.2e)\n", what.c_str(), res);
  } else {
    std::printf("Test for %s FAILED (residual %.2e)\n", what.c_str(), res);
    ++g_fail;
  }
}

// test.cc:215-262 (test_lmatrix_init)
static void test_lmatrix_init(const Context& ctx) {
  const int treelvl = 5, launchlvl = 3, base = 3;
  const int m = base << treelvl, n = 2;
  Matrix mat0(base, treelvl, n);
  mat0.rand();
  LMatrix lmat0(m, n, launchlvl, ctx);
  lmat0.init_data(mat0, ctx);
  const double r0 = (lmat0.to_matrix(ctx) - mat0).norm();

  Matrix UMat(base, treelvl, n);
  UMat.rand();
  const int nRhs = 1, cols = nRhs + treelvl * n;
  LMatrix lgUmat(m, cols, launchlvl, ctx);
  Matrix Rhs(base, treelvl, nRhs);
  Rhs.rand();
  lgUmat.init_data(Rhs, ctx);
  lgUmat.init_data(nRhs, cols, UMat, ctx);
  const double r1 = (lgUmat.to_matrix(0, nRhs, ctx) - Rhs).norm();
  const double r2 = (lgUmat.to_matrix(nRhs, nRhs + n, ctx) - UMat).norm();

  Matrix U(base, treelvl, n), V(base, treelvl, n);
  U.rand();
  V.rand();
  Vector D(base, treelvl);
  D.rand(100);
  const int nrow = D.rows(), nblk = 1 << treelvl, ncol = nrow / nblk;
  LMatrix lmat(nrow, ncol, launchlvl, ctx);
  lmat.init_dense_blocks(U, V, D, ctx);
  Matrix KMat = (U * V.T()) + D.to_diag_matrix();
  const double r3 = (lmat.to_matrix(0, n, 0, n, ctx) - KMat.block(0, n, 0, n)).norm();
  const double worst = std::fmax(std::fmax(r0, r1), std::fmax(r2, r3));
  report(worst < 1e-13, "legion matrix initialization", worst);
}

// test.cc:314-342 (test_gemm_reduce)
static void test_gemm_reduce(const Context& ctx) {
  const int m = 16, n = 3, nProc = 4, level = 2;
  Matrix UMat(m, n), VMat(m, n);
  UMat.rand(nProc);
  VMat.rand(nProc);
  Matrix W0 = VMat.block(0, m / 2, 0, n).T() * UMat.block(0, m / 2, 0, n);
  Matrix W1 = VMat.block(m / 2, m, 0, n).T() * UMat.block(m / 2, m, 0, n);
  LMatrix U(m, n, level, ctx), V(m, n, level, ctx);
  U.init_data(UMat, ctx);
  V.init_data(VMat, ctx);
  LMatrix W(2 * n, n, 1, ctx);
  LMatrix::gemmRed('t', 'n', 1.0, V, U, 0.0, W, ctx);
  const double r0 = (W.to_matrix(0, n, 0, n, ctx) - W0).norm();
  const double r1 = (W.to_matrix(n, 2 * n, 0, n, ctx) - W1).norm();
  report(r0 < 1e-13 && r1 < 1e-13, "gemm reduce", std::fmax(r0, r1));
}

// test.cc:344-371 (test_gemm_broadcast), with C a column window of A's
// region and beta = 1, as LMatrix::gemmBro requires (lmatrix.cc:819,834)
static void test_gemm_broadcast(const Context& ctx) {
  const int m = 16, n = 3, nProc = 8, level = 3;
  Matrix UMat(m, n), VMat(2 * n, n);
  UMat.rand(nProc);
  VMat.rand(2);
  Matrix W0 = UMat.block(0, 8, 0, n) * VMat.block(0, n, 0, n);
  Matrix W1 = UMat.block(8, 16, 0, n) * VMat.block(n, 2 * n, 0, n);
  LMatrix UW(m, 2 * n, level, ctx);  // [U | W], W = 0
  UW.clear(0.0, ctx);
  UW.init_data(0, n, UMat, ctx);
  LMatrix U = UW, W = UW;
  U.set_column_begin(0);
  U.set_column_size(n);
  W.set_column_begin(n);
  W.set_column_size(n);
  LMatrix V(2 * n, n, 1, ctx);
  V.init_data(VMat, ctx);
  LMatrix::gemmBro('n', 'n', 1.0, U, V, 1.0, W, ctx);
  const double r0 = (W.to_matrix(0, m / 2, 0, n, ctx) - W0).norm();
  const double r1 = (W.to_matrix(m / 2, m, 0, n, ctx) - W1).norm();
  report(r0 < 1e-13 && r1 < 1e-13, "gemm broadcast", std::fmax(r0, r1));
}

// test.cc:373-424 (test_node_solve)
static void test_node_solve(const Context& ctx) {
  const int level = 2, r = 3, nProc = 2;
  Matrix VuMat(level * 2 * r, r), VdMat(level * 2 * r, 1);
  VuMat.rand(nProc);
  VdMat.rand(nProc);
  Matrix sln[2];
  for (int l = 0; l < level; ++l) {
    const int ofs = l * 2 * r;
    Matrix S = Matrix::identity(2 * r);
    for (int i = 0; i < r; ++i)
      for (int j = 0; j < r; ++j) {
        S(r + i, j) = VuMat(ofs + i, j);
        S(i, r + j) = VuMat(ofs + r + i, j);
      }
    Matrix rhs(2 * r, 1);
    for (int i = 0; i < r; ++i) {
      rhs(i, 0) = VdMat(ofs + r + i, 0);
      rhs(r + i, 0) = VdMat(ofs + i, 0);
    }
    S.solve(rhs, ctx);
    sln[l] = rhs;
  }
  const int nlevel = 1;
  LMatrix VTu(level * 2 * r, r, nlevel, ctx), VTd(level * 2 * r, 1, nlevel, ctx);
  VTu.init_data(VuMat, ctx);
  VTd.init_data(VdMat, ctx);
  VTu.node_solve(VTd, ctx);
  const double r0 = (VTd.to_matrix(0, 2 * r, 0, 1, ctx) - sln[0]).norm();
  const double r1 = (VTd.to_matrix(2 * r, 4 * r, 0, 1, ctx) - sln[1]).norm();
  report(r0 < 1e-13 && r1 < 1e-13, "node solve", std::fmax(r0, r1));
}

// test.cc:427-456 (test_two_level_reduce): the 2-slot result region is
// created with one partition and split by two_level_partition.
static void test_two_level_reduce(const Context& ctx) {
  const int m = 16, n = 3, nProc = 4, level = 2;
  Matrix UMat(m, n), VMat(m, n);
  UMat.rand(nProc);
  VMat.rand(nProc);
  Matrix W0 = VMat.block(0, m / 2, 0, n).T() * UMat.block(0, m / 2, 0, n);
  Matrix W1 = VMat.block(m / 2, m, 0, n).T() * UMat.block(m / 2, m, 0, n);
  LMatrix U(m, n, level, ctx), V(m, n, level, ctx);
  U.init_data(UMat, ctx);
  V.init_data(VMat, ctx);
  LMatrix W(2 * n, n, 0, ctx);
  W.two_level_partition(ctx);
  LMatrix::gemmRed('t', 'n', 1.0, V, U, 0.0, W, ctx);
  const double r0 = (W.to_matrix(0, n, 0, n, ctx) - W0).norm();
  const double r1 = (W.to_matrix(n, 2 * n, 0, n, ctx) - W1).norm();
  report(r0 < 1e-13 && r1 < 1e-13 && W.num_partition() == 2 && W.partition_level() == 2,
         "two-level gemm reduce", std::fmax(r0, r1));
}

// test.cc:458-486 (test_two_level_broadcast): eta split by
// two_level_partition; C is a column window of A's region with beta = 1,
// as LMatrix::gemmBro requires (lmatrix.cc:819,834).
static void test_two_level_broadcast(const Context& ctx) {
  const int m = 16, n = 3, nProc = 8, level = 3;
  Matrix UMat(m, n), VMat(2 * n, n);
  UMat.rand(nProc);
  VMat.rand(1);
  Matrix W0 = UMat.block(0, 8, 0, n) * VMat.block(0, n, 0, n);
  Matrix W1 = UMat.block(8, 16, 0, n) * VMat.block(n, 2 * n, 0, n);
  LMatrix UW(m, 2 * n, level, ctx);
  UW.clear(0.0, ctx);
  UW.init_data(0, n, UMat, ctx);
  LMatrix U = UW, W = UW;
  U.set_column_begin(0);
  U.set_column_size(n);
  W.set_column_begin(n);
  W.set_column_size(n);
  LMatrix V(2 * n, n, 0, ctx);
  V.init_data(VMat, ctx);
  V.two_level_partition(ctx);
  LMatrix::gemmBro('n', 'n', 1.0, U, V, 1.0, W, ctx);
  const double r0 = (W.to_matrix(0, m / 2, 0, n, ctx) - W0).norm();
  const double r1 = (W.to_matrix(m / 2, m, 0, n, ctx) - W1).norm();
  report(r0 < 1e-13 && r1 < 1e-13, "two-level gemm broadcast", std::fmax(r0, r1));
}

// test.cc:488-541 (test_two_level_node_solve): two cores, the slot regions
// partitioned at level 1 and split by two_level_partition.
static void test_two_level_node_solve(const Context& ctx) {
  const int level = 2, r = 3, nProc = 2;
  Matrix VuMat(level * 2 * r, r), VdMat(level * 2 * r, 1);
  VuMat.rand(nProc);
  VdMat.rand(nProc);
  Matrix sln[2];
  double worst_res = 0.0;
  for (int l = 0; l < level; ++l) {
    const int ofs = l * 2 * r;
    Matrix S = Matrix::identity(2 * r);
    for (int i = 0; i < r; ++i)
      for (int j = 0; j < r; ++j) {
        S(r + i, j) = VuMat(ofs + i, j);
        S(i, r + j) = VuMat(ofs + r + i, j);
      }
    Matrix rhs(2 * r, 1);
    for (int i = 0; i < r; ++i) {
      rhs(i, 0) = VdMat(ofs + r + i, 0);
      rhs(r + i, 0) = VdMat(ofs + i, 0);
    }
    Matrix rhs_copy = rhs;
    S.solve(rhs, ctx);
    sln[l] = rhs;
    worst_res = std::fmax(worst_res, (rhs_copy - S * rhs).norm());
  }
  const int nlevel = 1;
  LMatrix VTu(level * 2 * r, r, nlevel, ctx), VTd(level * 2 * r, 1, nlevel, ctx);
  VTu.init_data(VuMat, ctx);
  VTd.init_data(VdMat, ctx);
  VTu.two_level_partition(ctx);
  VTd.two_level_partition(ctx);
  VTu.node_solve(VTd, ctx);
  const double r0 = (VTd.to_matrix(0, 2 * r, 0, 1, ctx) - sln[0]).norm();
  const double r1 = (VTd.to_matrix(2 * r, 4 * r, 0, 1, ctx) - sln[1]).norm();
  report(r0 < 1e-13 && r1 < 1e-13 && worst_res < 1e-13, "two-level node solve",
         std::fmax(r0, r1));
}

// test.cc:543-660 (test_solver): trees partitioned at launchlvl, the leaf
// tasks (KTree::solve -> LMatrix::solve -> hsolve) solve the levels below
// it, the driver loop (gemmRed x2 -> node_solve -> gemmBro) the levels
// above; the residual check that test.cc:651-658 leaves disabled is on.
static void test_solver(int rank, int treelvl, int launchlvl, const Context& ctx) {
  const int base = 64, n = rank;
  Matrix VMat(base, treelvl, n), UMat(base, treelvl, n), Rhs(base, treelvl, 1);
  VMat.rand();
  UMat.rand();
  Rhs.rand();
  Vector DVec(base, treelvl);
  DVec.rand(1000);
  const int nProc = 1 << launchlvl;
  hodlr::UTree uTree;
  uTree.init(UMat);
  hodlr::VTree vTree;
  vTree.init(VMat);
  hodlr::KTree kTree;
  kTree.init(UMat, VMat, DVec);
  (void)nProc;
  uTree.partition(launchlvl, ctx);
  vTree.partition(launchlvl, ctx);
  kTree.partition(launchlvl, ctx);
  uTree.init_rhs(Rhs, ctx);
  kTree.solve(uTree.leaf(), vTree.leaf(), ctx);
  for (int i = launchlvl; i > 0; --i) {
    LMatrix& V = vTree.level(i);
    LMatrix& u = uTree.uMat_level(i);
    LMatrix& d = uTree.dMat_level(i);
    const int rows = (1 << i) * V.cols();
    LMatrix VTu(rows, u.cols(), i - 1, ctx), VTd(rows, d.cols(), i - 1, ctx);
    VTu.two_level_partition(ctx);
    VTd.two_level_partition(ctx);
    LMatrix::gemmRed('t', 'n', 1.0, V, u, 0.0, VTu, ctx);
    LMatrix::gemmRed('t', 'n', 1.0, V, d, 0.0, VTd, ctx);
    VTu.node_solve(VTd, ctx);
    LMatrix::gemmBro('n', 'n', -1.0, u, VTd, 1.0, d, ctx);
  }
  Matrix x = uTree.solution(ctx);
  Matrix A = (UMat * VMat.T()) + DVec.to_diag_matrix();
  const double res = (Rhs - A * x).norm() / Rhs.norm();
  report(res < 1e-13, "solver (rank " + std::to_string(rank) + ", tree level " +
                          std::to_string(treelvl) + ", launch level " +
                          std::to_string(launchlvl) + ")", res);
}

// Reference-mode fixture through the facade's generators: with the
// generator clock standing for time(NULL), U, V, D and b are the drand48
// streams of seeds 11..14 (8 partitions of 64 rows, rank 8, 2 right-hand
// sides, D offset 640), solved by HMatrix.  The raw arrays go to `dir` for
// the caller to compare with the oracle's generators and the fixture
// solution.
static void test_reference_fixture(const std::string& dir, const Context& ctx) {
  const int base = 64, level = 3, rank = 8, nrhs = 2, offset = 10 * base;
  const long seed = 11;
  Matrix U(base, level, rank), V(base, level, rank), b(base, level, nrhs);
  Vector D(base, level);
  hodlr::set_rand_clock(seed);
  U.rand();
  hodlr::set_rand_clock(seed + 1);
  V.rand();
  hodlr::set_rand_clock(seed + 2);
  D.rand(offset);
  hodlr::set_rand_clock(seed + 3);
  b.rand();
  hodlr::clear_rand_clock();
  HMatrix H(1, level);
  H.init(U, V, D, ctx);
  Matrix x = H.solve_all(b, ctx);
  H.destroy(ctx);
  auto dump = [&](const std::string& name, const double* p, size_t count) {
    const std::string path = dir + "/" + name + ".f64";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, sizeof(double), count, f) != count) {
      std::printf("cannot write %s\n", path.c_str());
      ++g_fail;
    }
    if (f) std::fclose(f);
  };
  dump("U", U.pointer(), (size_t)U.rows() * U.cols());
  dump("V", V.pointer(), (size_t)V.rows() * V.cols());
  dump("D", D.data(), (size_t)D.rows());
  dump("b", b.pointer(), (size_t)b.rows() * b.cols());
  dump("x", x.pointer(), (size_t)x.rows() * x.cols());
  std::printf("reference fixture written to %s\n", dir.c_str());
}

// The end-to-end check that test.cc:651-658 leaves disabled: HMatrix::solve
// against the dense matrix U V^T + diag(D).
static void test_hmatrix_solve(const Context& ctx) {
  const int base = 64, level = 4, rank = 8;
  Matrix U(base, level, rank), V(base, level, rank), b(base, level, 2);
  U.rand();
  V.rand();
  b.rand();
  Vector D(base, level);
  D.rand(1000);
  HMatrix H(1, level);
  H.init(U, V, D, ctx);
  Matrix x = H.solve_all(b, ctx);
  Matrix A = (U * V.T()) + D.to_diag_matrix();
  const double res = (b - A * x).norm() / b.norm();
  Vector x0 = H.solve(b, ctx);  // factor reuse, first column
  double d0 = 0.0;
  for (int i = 0; i < x0.rows(); ++i) d0 = std::fmax(d0, std::fabs(x0[i] - x(i, 0)));
  H.destroy(ctx);
  report(res < 1e-13 && d0 == 0.0, "HMatrix solve", res);
}

int main(int argc, char** argv) {
  Context ctx;  // device 0, null stream
  hodlr::set_master_seed(0);
  try {
    test_lmatrix_init(ctx);
    test_gemm_reduce(ctx);
    test_gemm_broadcast(ctx);
    test_node_solve(ctx);
    test_two_level_reduce(ctx);
    test_two_level_broadcast(ctx);
    test_two_level_node_solve(ctx);
    test_solver(8, 4, 2, ctx);
    test_solver(5, 3, 3, ctx);
    test_solver(6, 3, 0, ctx);
    test_hmatrix_solve(ctx);
    if (argc > 1) test_reference_fixture(argv[1], ctx);
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 2;
  }
  return g_fail ? 1 : 0;
}
