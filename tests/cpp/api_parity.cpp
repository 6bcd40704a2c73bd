// Runs libnabo's C++ interface (include/nabo/nabo.h) on the GPU for every
// cloud type the reference instantiates (nabo/nabo.cpp:181-186: MatrixXf,
// Matrix3Xf, Map<const Matrix3Xf, Aligned>) and every factory entry
// (create with each search type, createBruteForce, createKDTreeLinearHeap,
// createKDTreeTreeHeap), and writes the raw results for a Python test that
// compares them with the oracle bit for bit.
//
// Usage: api_parity IN_DIR OUT_DIR K RADIUS
//   IN_DIR/cloud.bin, IN_DIR/query.bin: int64 count, int64 rows, then
//   count x rows float32 (one point per row, i.e. the column-major
//   rows x count Eigen matrix).
//   OUT_DIR/<cloud type>_<factory>_<flags>.bin: int64 k, int64 m, int64
//   touched, then the k x m indices (int32) and dists2 (float32) in
//   column-major order, i.e. query i's k results at [i*k, i*k+k).
//   Also <...>_radii.bin (the maxRadii overload with radii RADIUS * (1 + i % 3)
//   / 2) and <...>_vector.bin (the single-vector knn for the first 64 queries).
// Prints "PARITY DUMP OK" on success.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "nabo/nabo.h"

using namespace Nabo;

namespace {

struct Points {
  int64_t count = 0, rows = 0;
  std::vector<float> v;
};

Points read_points(const std::string& path) {
  Points p;
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot read " + path);
  in.read(reinterpret_cast<char*>(&p.count), 8);
  in.read(reinterpret_cast<char*>(&p.rows), 8);
  p.v.resize(size_t(p.count * p.rows));
  in.read(reinterpret_cast<char*>(p.v.data()), std::streamsize(p.v.size() * 4));
  if (!in) throw std::runtime_error("truncated " + path);
  return p;
}

template <typename M>
void fill(M& m, const Points& p) {
  for (int64_t c = 0; c < p.count; ++c)
    for (int64_t r = 0; r < p.rows; ++r) m(r, c) = p.v[size_t(c * p.rows + r)];
}

void write_result(const std::string& path, int64_t k, int64_t m, int64_t touched,
                  const int* idx, const float* d2) {
  std::ofstream out(path, std::ios::binary);
  out.write(reinterpret_cast<const char*>(&k), 8);
  out.write(reinterpret_cast<const char*>(&m), 8);
  out.write(reinterpret_cast<const char*>(&touched), 8);
  out.write(reinterpret_cast<const char*>(idx), std::streamsize(k * m * 4));
  out.write(reinterpret_cast<const char*>(d2), std::streamsize(k * m * 4));
  if (!out) throw std::runtime_error("cannot write " + path);
}

template <typename Cloud>
void run_cloud(const std::string& tag, const Cloud& cloud, const Points& qp,
               const std::string& out_dir, int k, float radius) {
  typedef NearestNeighbourSearch<float, Cloud> NNS;
  typename NNS::Matrix q(qp.rows, qp.count);
  fill(q, qp);
  const int64_t m = qp.count;
  struct Entry {
    const char* name;
    NNS* nns;
  };
  const Parameters bucket("bucketSize", 8u);
  Entry entries[] = {
      {"create_bruteforce", NNS::create(cloud, int(cloud.rows()), NNS::BRUTE_FORCE, 0)},
      {"create_linear", NNS::create(cloud, int(cloud.rows()), NNS::KDTREE_LINEAR_HEAP,
                                    NNS::TOUCH_STATISTICS, bucket)},
      {"create_treeheap", NNS::create(cloud, int(cloud.rows()), NNS::KDTREE_TREE_HEAP,
                                      NNS::TOUCH_STATISTICS, bucket)},
      {"create_kdhip", NNS::create(cloud, int(cloud.rows()), NNS::KDTREE_HIP, 0, bucket)},
      {"create_bfhip", NNS::create(cloud, int(cloud.rows()), NNS::BRUTE_FORCE_HIP, 0)},
      {"helper_bruteforce", NNS::createBruteForce(cloud)},
      {"helper_linear", NNS::createKDTreeLinearHeap(cloud)},
      {"helper_treeheap", NNS::createKDTreeTreeHeap(cloud)},
  };
  for (const Entry& e : entries) {
    for (unsigned flags : {0u, unsigned(NNS::SORT_RESULTS)}) {
      typename NNS::IndexMatrix idx(k, m);
      typename NNS::Matrix d2(k, m);
      const unsigned long touched = e.nns->knn(q, idx, d2, k, 0, flags, radius);
      write_result(out_dir + "/" + tag + "_" + e.name + "_f" + std::to_string(flags) + ".bin", k,
                   m, int64_t(touched), &idx.coeffRef(0, 0), &d2.coeffRef(0, 0));
    }
    // the maxRadii overload (nabo/nabo.h:351)
    typename NNS::Vector radii(m);
    for (int64_t i = 0; i < m; ++i) radii.coeffRef(i) = radius * float(1 + i % 3) / 2.f;
    typename NNS::IndexMatrix ridx(k, m);
    typename NNS::Matrix rd2(k, m);
    const unsigned long rt = e.nns->knn(q, ridx, rd2, radii, k, 0, 0);
    write_result(out_dir + "/" + tag + "_" + e.name + "_radii.bin", k, m, int64_t(rt),
                 &ridx.coeffRef(0, 0), &rd2.coeffRef(0, 0));
    // the single-vector knn (nabo/nabo.cpp:82-100) on the first queries
    const int64_t vm = std::min<int64_t>(m, 64);
    std::vector<int> vi(size_t(k * vm));
    std::vector<float> vd(size_t(k * vm));
    for (int64_t i = 0; i < vm; ++i) {
      typename NNS::Vector v(qp.rows);
      for (int64_t r = 0; r < qp.rows; ++r) v.coeffRef(r) = q(r, i);
      typename NNS::IndexVector oi(k);
      typename NNS::Vector od(k);
      e.nns->knn(v, oi, od, k, 0, 0, radius);
      for (int j = 0; j < k; ++j) {
        vi[size_t(i * k + j)] = oi.coeff(j);
        vd[size_t(i * k + j)] = od.coeff(j);
      }
    }
    write_result(out_dir + "/" + tag + "_" + e.name + "_vector.bin", k, vm, 0, vi.data(),
                 vd.data());
    delete e.nns;
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::cerr << "usage: " << argv[0] << " IN_DIR OUT_DIR K RADIUS" << std::endl;
    return 2;
  }
  try {
    const std::string in_dir = argv[1], out_dir = argv[2];
    const int k = std::atoi(argv[3]);
    const float radius = float(std::atof(argv[4]));
    const Points cp = read_points(in_dir + "/cloud.bin");
    const Points qp = read_points(in_dir + "/query.bin");
    if (cp.rows != 3 || qp.rows != 3) throw std::runtime_error("3-D points expected");

    Eigen::MatrixXf dyn(cp.rows, cp.count);
    fill(dyn, cp);
    run_cloud("matrixxf", dyn, qp, out_dir, k, radius);

    Eigen::Matrix3Xf fixed(3, cp.count);
    fill(fixed, cp);
    run_cloud("matrix3xf", fixed, qp, out_dir, k, radius);

    // Map<const Matrix3Xf, Aligned> over the same storage (Eigen's default
    // allocation is 16-byte aligned).
    const Eigen::Map<const Eigen::Matrix3Xf, Eigen::Aligned> mapped(fixed.data(), 3, cp.count);
    run_cloud("map3xf", mapped, qp, out_dir, k, radius);

    // NNSearchD is refused at creation, with a message naming the gap.
    Eigen::MatrixXd dcloud(cp.rows, cp.count);
    for (int64_t c = 0; c < cp.count; ++c)
      for (int64_t r = 0; r < cp.rows; ++r) dcloud(r, c) = cp.v[size_t(c * cp.rows + r)];
    bool refused = false;
    try {
      delete NNSearchD::create(dcloud);
    } catch (const std::runtime_error& e) {
      refused = std::string(e.what()).find("NNSearchD") != std::string::npos;
    }
    if (!refused) {
      std::cerr << "NNSearchD was not refused" << std::endl;
      return 1;
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  std::printf("PARITY DUMP OK\n");
  return 0;
}
