// Exercises the drop-in C++ interface (include/nabo/nabo.h) end to end on the
// GPU, the way libnabo's examples/usage.cpp and tests/knnvalidate.cpp use
// the reference: create, matrix and vector knn, flags, radii, statistics,
// argument errors. Prints "API OK" on success, exits non-zero otherwise.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "nabo/nabo.h"

using namespace Nabo;

#define EXPECT(cond)                                                     \
  do {                                                                   \
    if (!(cond)) {                                                       \
      std::cerr << "FAILED " #cond " at line " << __LINE__ << std::endl; \
      return 1;                                                          \
    }                                                                    \
  } while (0)

template <typename F>
bool throws(F f) {
  try {
    f();
  } catch (const std::runtime_error&) {
    return true;
  }
  return false;
}

int main() {
  std::srand(7);
  const int n = 4000, m = 600;
  Eigen::MatrixXf M = Eigen::MatrixXf::Random(3, n);
  Eigen::MatrixXf q = Eigen::MatrixXf::Random(3, m);

  NNSearchF* kd = NNSearchF::createKDTreeLinearHeap(M);
  NNSearchF* hip = NNSearchF::create(M, 3, NNSearchF::KDTREE_HIP, NNSearchF::TOUCH_STATISTICS,
                                     Parameters("bucketSize", 8u));
  NNSearchF* bf = NNSearchF::createBruteForce(M);
  EXPECT(kd->dim == 3 && hip->dim == 3);
  for (int d = 0; d < 3; ++d) EXPECT(kd->minBound(d) >= -1.f && kd->maxBound(d) <= 1.f);

  // matrix knn, sorted, compared with brute force
  Eigen::MatrixXi i1(5, m), i2(5, m), i3(5, m);
  Eigen::MatrixXf d1(5, m), d2(5, m), d3(5, m);
  kd->knn(q, i1, d1, 5, 0, NNSearchF::SORT_RESULTS);
  const unsigned long touched = hip->knn(q, i2, d2, 5, 0, NNSearchF::SORT_RESULTS);
  bf->knn(q, i3, d3, 5, 0, NNSearchF::SORT_RESULTS);
  EXPECT(touched > 0);
  for (int c = 0; c < m; ++c)
    for (int r = 0; r < 5; ++r) {
      EXPECT(i1(r, c) == i2(r, c) && d1(r, c) == d2(r, c));
      EXPECT(d1(r, c) == d3(r, c));
      if (r) EXPECT(d1(r - 1, c) <= d1(r, c));
    }

  // KDTREE_TREE_HEAP: IndexHeapSTL results, a max-heap unless SORT_RESULTS
  NNSearchF* th = NNSearchF::create(M, 3, NNSearchF::KDTREE_TREE_HEAP);
  Eigen::MatrixXi ti(5, m), tsi(5, m);
  Eigen::MatrixXf td(5, m), tsd(5, m);
  th->knn(q, ti, td, 5, 0, 0);
  th->knn(q, tsi, tsd, 5, 0, NNSearchF::SORT_RESULTS);
  for (int c = 0; c < m; ++c)
    for (int r = 0; r < 5; ++r) {
      EXPECT(tsd(r, c) == d1(r, c));
      EXPECT(td(0, c) >= td(r, c));
      if (r) EXPECT(td((r - 1) / 2, c) >= td(r, c));
    }
  delete th;

  // self match (examples/usage.cpp:37-57)
  Eigen::MatrixXi si(3, n);
  Eigen::MatrixXf sd(3, n);
  kd->knn(M, si, sd, 3, 0, NNSearchF::SORT_RESULTS | NNSearchF::ALLOW_SELF_MATCH);
  for (int c = 0; c < n; ++c) EXPECT(si(0, c) == c);
  Eigen::MatrixXi ni(2, n);
  Eigen::MatrixXf nd(2, n);
  kd->knn(M, ni, nd, 2, 0.1f, 0);
  for (int c = 0; c < n; ++c) EXPECT(ni(0, c) != c);

  // single-vector knn (nabo/nabo.cpp:82-100)
  Eigen::VectorXf v = Eigen::VectorXf::Random(3);
  Eigen::VectorXi vi(4);
  Eigen::VectorXf vd(4);
  kd->knn(v, vi, vd, 4);
  Eigen::MatrixXf vq(3, 1);
  for (int d = 0; d < 3; ++d) vq(d, 0) = v(d);
  Eigen::MatrixXi vmi(4, 1);
  Eigen::MatrixXf vmd(4, 1);
  bf->knn(vq, vmi, vmd, 4, 0, NNSearchF::SORT_RESULTS);
  for (int r = 0; r < 4; ++r) EXPECT(vd(r) == vmd(r, 0));

  // radius: unfilled slots are (InvalidIndex, InvalidValue)
  Eigen::MatrixXi ri(8, m);
  Eigen::MatrixXf rd(8, m);
  kd->knn(q, ri, rd, 8, 0, 0, 0.05f);
  int empty = 0;
  for (int c = 0; c < m; ++c)
    for (int r = 0; r < 8; ++r) {
      if (ri(r, c) == NNSearchF::InvalidIndex) {
        EXPECT(std::isinf(rd(r, c)));
        ++empty;
      } else {
        EXPECT(rd(r, c) <= 0.05f * 0.05f);
      }
    }
  EXPECT(empty > 0);
  Eigen::VectorXf radii = Eigen::VectorXf::Constant(m, 0.05f);
  Eigen::MatrixXi ri2(8, m);
  Eigen::MatrixXf rd2(8, m);
  kd->knn(q, ri2, rd2, radii, 8, 0, 0);
  for (int c = 0; c < m; ++c)
    for (int r = 0; r < 8; ++r) EXPECT(ri2(r, c) == ri(r, c));

  // argument errors (nabo/nabo.cpp:102-131)
  Eigen::MatrixXi bad_i(4, m);
  Eigen::MatrixXf bad_d(5, m);
  EXPECT(throws([&] { kd->knn(q, bad_i, bad_d, 5); }));
  EXPECT(throws([&] { kd->knn(q, i1, d1, 5, 0, 4); }));
  Eigen::MatrixXi big_i(n, m);
  Eigen::MatrixXf big_d(n, m);
  EXPECT(throws([&] { kd->knn(q, big_i, big_d, n); }));
  EXPECT(throws([&] { NNSearchF::create(M, 0); }));
  EXPECT(throws([&] { NNSearchF::create(M, 3, NNSearchF::KDTREE_LINEAR_HEAP, 0,
                                         Parameters("bucketSize", 1u)); }));
  EXPECT(throws([&] { NNSearchF::create(M, 3, NNSearchF::BRUTE_FORCE_CL); }));

  delete kd;
  delete hip;
  delete bf;
  std::printf("API OK\n");
  return 0;
}
