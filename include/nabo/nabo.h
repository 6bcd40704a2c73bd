// nabo/nabo.h - libnabo's public interface, served by the MI355X library.
//
// Drop-in for libnabo's nabo/nabo.h (nabo/nabo.h:204-451 of libnabo 1.1.2):
// the same Nabo::NearestNeighbourSearch<T, CloudType> with its typedefs,
// members, SearchType / CreationOptionFlags / SearchOptionFlags values, the
// three knn overloads and the create* factories, plus two search types
// appended after the reference's: KDTREE_HIP and BRUTE_FORCE_HIP. Header
// only; link with libnabo_hip.so. Every search runs on the GPU through the
// C ABI of nabo_hip.h:
//   KDTREE_HIP, KDTREE_LINEAR_HEAP -> HIP kd-tree, IndexHeapBruteForceVector
//     results (ascending, a new entry after equal distances)
//   KDTREE_TREE_HEAP               -> HIP kd-tree, IndexHeapSTL results
//   BRUTE_FORCE                    -> HIP brute force, IndexHeapSTL results
//     (the reference's own heap: same order and ties)
//   BRUTE_FORCE_HIP                -> HIP brute force, ascending results,
//     ties by index
//   KDTREE_CL_* / BRUTE_FORCE_CL                      -> error, as the
//     reference without OpenCL (nabo/nabo.cpp:148-151)
// Scalar type float only (NNSearchF, over MatrixXf, Matrix3Xf or
// Map<const Matrix3Xf, Aligned> clouds, the reference's instantiations,
// nabo/nabo.cpp:181-186); creating an NNSearchD throws, the interface's one
// intended gap (the GPU searches compute in float).
//
// Matrices are Eigen's when <Eigen/Core> is available, otherwise the small
// column-major subset of nabo_amd/dense.h (aliased as namespace Eigen).
#ifndef NABO_AMD_NABO_H
#define NABO_AMD_NABO_H

#include <algorithm>
#include <any>
#include <cstdint>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>

#if defined(__has_include)
#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
#define NABO_AMD_HAVE_EIGEN 1
#endif
#endif
#ifndef NABO_AMD_HAVE_EIGEN
#include "../nabo_amd/dense.h"
namespace Eigen = nabo_amd::dense;
#endif

#include "../nabo_hip.h"

namespace Nabo {

#define NABO_VERSION "1.1.2"
#define NABO_VERSION_INT 10102

template <typename IndexType>
inline constexpr IndexType invalidIndex() {
  static_assert(std::is_integral<IndexType>::value, "");
  return std::is_unsigned<IndexType>::value ? std::numeric_limits<IndexType>::max()
                                            : IndexType(-1);
}

template <typename ValueType>
inline constexpr ValueType invalidValue() {
  static_assert(std::is_floating_point<ValueType>::value, "");
  return std::numeric_limits<ValueType>::infinity();
}

// Parameter map (nabo/nabo.h:231-254). Keys read: "bucketSize" (unsigned,
// the reference's key), "querySort" (int), "bfPrecision" (int, one of
// NABO_HIP_BF_*), "bfExtra" (unsigned), "device" (int), "heap" (int,
// NABO_HIP_HEAP_*), and the kd-tree search's schedule: "waveQueries",
// "stackPlace" (NABO_HIP_STACK_*), "stlLog" (int; nabo_hip.h).
struct Parameters : public std::map<std::string, std::any> {
  Parameters() {}
  Parameters(const std::string& key, const std::any& value) { (*this)[key] = value; }
  template <typename T>
  T get(const std::string& key, const T& defaultValue) const {
    const_iterator it(find(key));
    if (it != end()) return std::any_cast<T>(it->second);
    return defaultValue;
  }
};

// False for every type, but only once instantiated: lets the wrong-cloud
// overloads below compile until a caller selects one.
template <typename>
struct wrong_cloud_type : std::false_type {};

template <typename T, typename Cloud_T = Eigen::Matrix<T, Eigen::Dynamic, Eigen::Dynamic> >
struct NearestNeighbourSearch {
  typedef typename Eigen::Matrix<T, Eigen::Dynamic, 1> Vector;
  typedef typename Eigen::Matrix<T, Eigen::Dynamic, Eigen::Dynamic> Matrix;
  typedef Cloud_T CloudType;
  typedef int Index;
  typedef typename Eigen::Matrix<Index, Eigen::Dynamic, 1> IndexVector;
  typedef typename Eigen::Matrix<Index, Eigen::Dynamic, Eigen::Dynamic> IndexMatrix;

  const CloudType& cloud;
  const Index dim;
  const unsigned creationOptionFlags;
  const Vector minBound;
  const Vector maxBound;

  static constexpr Index InvalidIndex = invalidIndex<Index>();
  static constexpr T InvalidValue = invalidValue<T>();

  enum SearchType {
    BRUTE_FORCE = 0,
    KDTREE_LINEAR_HEAP,
    KDTREE_TREE_HEAP,
    KDTREE_CL_PT_IN_NODES,
    KDTREE_CL_PT_IN_LEAVES,
    BRUTE_FORCE_CL,
    KDTREE_HIP,       // GPU kd-tree (NABO_HIP_SEARCH_KDTREE)
    BRUTE_FORCE_HIP,  // GPU brute force (NABO_HIP_SEARCH_BRUTE_FORCE)
    SEARCH_TYPE_COUNT
  };
  enum CreationOptionFlags { TOUCH_STATISTICS = 1 };
  enum SearchOptionFlags { ALLOW_SELF_MATCH = 1, SORT_RESULTS = 2 };

  unsigned long knn(const Vector& query, IndexVector& indices, Vector& dists2, const Index k = 1,
                    const T epsilon = 0, const unsigned optionFlags = 0,
                    const T maxRadius = std::numeric_limits<T>::infinity()) const {
    // nabo/nabo.cpp:82-100: a dim x 1 query matrix, k x 1 outputs.
    Matrix q(dim, 1);
    for (Index d = 0; d < dim; ++d) q(d, 0) = query.coeff(d);
    IndexMatrix im(k, 1);
    Matrix dm(k, 1);
    const unsigned long stats = knn(q, im, dm, k, epsilon, optionFlags, maxRadius);
    indices.resize(k);
    dists2.resize(k);
    for (Index i = 0; i < k; ++i) {
      indices.coeffRef(i) = im.coeff(i, 0);
      dists2.coeffRef(i) = dm.coeff(i, 0);
    }
    return stats;
  }

  virtual unsigned long knn(const Matrix& query, IndexMatrix& indices, Matrix& dists2,
                            const Index k = 1, const T epsilon = 0, const unsigned optionFlags = 0,
                            const T maxRadius = std::numeric_limits<T>::infinity()) const = 0;
  virtual unsigned long knn(const Matrix& query, IndexMatrix& indices, Matrix& dists2,
                            const Vector& maxRadii, const Index k = 1, const T epsilon = 0,
                            const unsigned optionFlags = 0) const = 0;

  static NearestNeighbourSearch* create(const CloudType& cloud,
                                        const Index dim = std::numeric_limits<Index>::max(),
                                        const SearchType preferedType = KDTREE_LINEAR_HEAP,
                                        const unsigned creationOptionFlags = 0,
                                        const Parameters& additionalParameters = Parameters());
  // The helpers of nabo/nabo.cpp:157-179 build the reference's own types:
  // BRUTE_FORCE (IndexHeapSTL results), KDTREE_LINEAR_HEAP and
  // KDTREE_TREE_HEAP (IndexHeapSTL results), each on the GPU.
  static NearestNeighbourSearch* createBruteForce(
      const CloudType& cloud, const Index dim = std::numeric_limits<Index>::max(),
      const unsigned creationOptionFlags = 0) {
    return create(cloud, dim, BRUTE_FORCE, creationOptionFlags);
  }
  static NearestNeighbourSearch* createKDTreeLinearHeap(
      const CloudType& cloud, const Index dim = std::numeric_limits<Index>::max(),
      const unsigned creationOptionFlags = 0,
      const Parameters& additionalParameters = Parameters()) {
    return create(cloud, dim, KDTREE_LINEAR_HEAP, creationOptionFlags, additionalParameters);
  }
  static NearestNeighbourSearch* createKDTreeTreeHeap(
      const CloudType& cloud, const Index dim = std::numeric_limits<Index>::max(),
      const unsigned creationOptionFlags = 0,
      const Parameters& additionalParameters = Parameters()) {
    return create(cloud, dim, KDTREE_TREE_HEAP, creationOptionFlags, additionalParameters);
  }

  // A cloud whose type is not CloudType (a fixed-row matrix given to a search
  // over dynamic matrices, for instance) is refused at compile time, as by
  // nabo/nabo.h:390-424: these overloads take any other argument type and
  // fail when instantiated.
  template <typename WrongMatrixType>
  static NearestNeighbourSearch* create(const WrongMatrixType&,
                                        const Index = std::numeric_limits<Index>::max(),
                                        const SearchType = KDTREE_LINEAR_HEAP,
                                        const unsigned = 0, const Parameters& = Parameters()) {
    static_assert(wrong_cloud_type<WrongMatrixType>::value,
                  "Please make sure that the decltype of the first parameter is equal to the "
                  "Matrix typedef");
    return nullptr;
  }
  template <typename WrongMatrixType>
  static NearestNeighbourSearch* createBruteForce(const WrongMatrixType&,
                                                  const Index = std::numeric_limits<Index>::max(),
                                                  const unsigned = 0) {
    static_assert(wrong_cloud_type<WrongMatrixType>::value,
                  "Please make sure that the decltype of the first parameter is equal to the "
                  "Matrix typedef");
    return nullptr;
  }
  template <typename WrongMatrixType>
  static NearestNeighbourSearch* createKDTreeLinearHeap(
      const WrongMatrixType&, const Index = std::numeric_limits<Index>::max(),
      const unsigned = 0, const Parameters& = Parameters()) {
    static_assert(wrong_cloud_type<WrongMatrixType>::value,
                  "Please make sure that the decltype of the first parameter is equal to the "
                  "Matrix typedef");
    return nullptr;
  }
  template <typename WrongMatrixType>
  static NearestNeighbourSearch* createKDTreeTreeHeap(
      const WrongMatrixType&, const Index = std::numeric_limits<Index>::max(),
      const unsigned = 0, const Parameters& = Parameters()) {
    static_assert(wrong_cloud_type<WrongMatrixType>::value,
                  "Please make sure that the decltype of the first parameter is equal to the "
                  "Matrix typedef");
    return nullptr;
  }

  virtual ~NearestNeighbourSearch() {}

 protected:
  // nabo/nabo.cpp:68-80
  NearestNeighbourSearch(const CloudType& cloud, const Index dim,
                         const unsigned creationOptionFlags)
      : cloud(cloud),
        dim(std::min(dim, Index(cloud.rows()))),
        creationOptionFlags(creationOptionFlags),
        minBound(Vector::Constant(this->dim, std::numeric_limits<T>::max())),
        maxBound(Vector::Constant(this->dim, std::numeric_limits<T>::min())) {
    if (cloud.cols() == 0) throw std::runtime_error("Cloud has no points");
    if (cloud.rows() == 0) throw std::runtime_error("Cloud has 0 dimensions");
  }

  // nabo/nabo.cpp:102-131
  void checkSizesKnn(const Matrix& query, const IndexMatrix& indices, const Matrix& dists2,
                     const Index k, const unsigned optionFlags,
                     const Vector* maxRadii = 0) const {
    std::ostringstream e;
    const bool allowSelfMatch(optionFlags & ALLOW_SELF_MATCH);
    if (allowSelfMatch) {
      if (k > cloud.cols())
        e << "Requesting more points (" << k << ") than available in cloud (" << cloud.cols()
          << ")";
    } else if (k > cloud.cols() - 1) {
      e << "Requesting more points (" << k << ") than available in cloud minus 1 ("
        << cloud.cols() - 1 << ") (as self match is forbidden)";
    }
    if (e.str().empty() && query.rows() < dim)
      e << "Query has less dimensions (" << query.rows() << ") than requested for cloud ("
        << dim << ")";
    if (e.str().empty() && indices.rows() != k)
      e << "Index matrix has a different number of rows (" << indices.rows() << ") than k ("
        << k << ")";
    if (e.str().empty() && indices.cols() != query.cols())
      e << "Index matrix has a different number of columns (" << indices.rows()
        << ") than query (" << query.cols() << ")";
    if (e.str().empty() && dists2.rows() != k)
      e << "Distance matrix has a different number of rows (" << dists2.rows() << ") than k ("
        << k << ")";
    if (e.str().empty() && dists2.cols() != query.cols())
      e << "Distance matrix has a different number of columns (" << dists2.rows()
        << ") than query (" << query.cols() << ")";
    if (e.str().empty() && maxRadii && (maxRadii->size() != query.cols()))
      e << "Maximum radii vector has not the same length (" << maxRadii->size()
        << ") than query has columns (" << k << ")";
    const unsigned maxOptionFlagsValue(ALLOW_SELF_MATCH | SORT_RESULTS);
    if (e.str().empty() && optionFlags > maxOptionFlagsValue)
      e << "OR-ed value of option flags (" << optionFlags
        << ") is larger than maximal valid value (" << maxOptionFlagsValue << ")";
    if (!e.str().empty()) throw std::runtime_error(e.str());
  }
};

typedef NearestNeighbourSearch<float> NNSearchF;
typedef NearestNeighbourSearch<double> NNSearchD;

namespace hip_detail {

inline void check(int rc) {
  if (rc != NABO_HIP_OK) throw std::runtime_error(nabo_hip_last_error());
}

// A search of type KDTREE_HIP or BRUTE_FORCE_HIP held by the HIP library.
template <typename T, typename CloudType>
struct HipSearch : public NearestNeighbourSearch<T, CloudType> {
  typedef NearestNeighbourSearch<T, CloudType> Base;
  typedef typename Base::Vector Vector;
  typedef typename Base::Matrix Matrix;
  typedef typename Base::Index Index;
  typedef typename Base::IndexMatrix IndexMatrix;

  nabo_hip_index* handle = nullptr;

  HipSearch(const CloudType& cloud, const Index dim, const unsigned creationOptionFlags,
            int hipType, int heap, const Parameters& params)
      : Base(cloud, dim, creationOptionFlags) {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "");
    // NNSearchD (T = double) is the interface's one intended gap: every GPU
    // search step computes in float, and a double search would need an FP64
    // search path, which this library does not have. Creating one throws,
    // before anything is allocated.
    if (!std::is_same<T, float>::value)
      throw std::runtime_error(
          "NNSearchD is not offered: the HIP searches compute in float (use NNSearchF)");
    nabo_hip_params p;
    nabo_hip_params_default(&p);
    p.bucket_size = params.get<unsigned>("bucketSize", 8u);
    p.query_sort = params.get<int>("querySort", 1);
    p.bf_precision = params.get<int>("bfPrecision", NABO_HIP_BF_FP32);
    p.bf_extra = int(params.get<unsigned>("bfExtra", 32u));
    p.device = params.get<int>("device", -1);
    p.heap = params.get<int>("heap", heap);
    p.wave_queries = params.get<int>("waveQueries", 0);
    p.stack_place = params.get<int>("stackPlace", NABO_HIP_STACK_AUTO);
    p.stl_log = params.get<int>("stlLog", 0);
    const float* data = reinterpret_cast<const float*>(&cloud.coeff(0, 0));
    check(nabo_hip_create(data, int64_t(cloud.cols()), int32_t(cloud.rows()),
                          int64_t(cloud.outerStride()), int32_t(dim), hipType,
                          creationOptionFlags, &p, &handle));
    Vector& lo = const_cast<Vector&>(this->minBound);
    Vector& hi = const_cast<Vector&>(this->maxBound);
    check(nabo_hip_get_bounds(handle, reinterpret_cast<float*>(&lo.coeffRef(0)),
                              reinterpret_cast<float*>(&hi.coeffRef(0))));
  }
  ~HipSearch() override { nabo_hip_destroy(handle); }

  unsigned long knn(const Matrix& query, IndexMatrix& indices, Matrix& dists2, const Index k,
                    const T epsilon, const unsigned optionFlags,
                    const T maxRadius) const override {
    this->checkSizesKnn(query, indices, dists2, k, optionFlags);
    return run(query, indices, dists2, nullptr, k, epsilon, optionFlags, maxRadius);
  }
  unsigned long knn(const Matrix& query, IndexMatrix& indices, Matrix& dists2,
                    const Vector& maxRadii, const Index k, const T epsilon,
                    const unsigned optionFlags) const override {
    this->checkSizesKnn(query, indices, dists2, k, optionFlags, &maxRadii);
    return run(query, indices, dists2, reinterpret_cast<const float*>(&maxRadii.coeff(0)), k,
               epsilon, optionFlags, 0);
  }

 private:
  unsigned long run(const Matrix& query, IndexMatrix& indices, Matrix& dists2,
                    const float* radii, const Index k, const T epsilon,
                    const unsigned optionFlags, const T maxRadius) const {
    uint64_t touched = 0;
    if (query.cols() == 0 || k == 0) return 0;
    check(nabo_hip_knn(handle, reinterpret_cast<const float*>(&query.coeff(0, 0)),
                       int32_t(query.rows()), int64_t(query.rows()), int64_t(query.cols()),
                       int32_t(k), float(epsilon), optionFlags, float(maxRadius), radii,
                       &indices.coeffRef(0, 0), reinterpret_cast<float*>(&dists2.coeffRef(0, 0)),
                       &touched));
    return (unsigned long)touched;
  }
};

}  // namespace hip_detail

// nabo/nabo.cpp:134-155, with the HIP types.
template <typename T, typename CloudType>
NearestNeighbourSearch<T, CloudType>* NearestNeighbourSearch<T, CloudType>::create(
    const CloudType& cloud, const Index dim, const SearchType preferedType,
    const unsigned creationOptionFlags, const Parameters& additionalParameters) {
  if (dim <= 0) throw std::runtime_error("Your space must have at least one dimension");
  switch (preferedType) {
    case BRUTE_FORCE:
    case BRUTE_FORCE_HIP:
      return new hip_detail::HipSearch<T, CloudType>(
          cloud, dim, creationOptionFlags, NABO_HIP_SEARCH_BRUTE_FORCE,
          preferedType == BRUTE_FORCE ? NABO_HIP_HEAP_STL : NABO_HIP_HEAP_LINEAR,
          additionalParameters);
    case KDTREE_LINEAR_HEAP:
    case KDTREE_TREE_HEAP:
    case KDTREE_HIP:
      return new hip_detail::HipSearch<T, CloudType>(
          cloud, dim, creationOptionFlags, NABO_HIP_SEARCH_KDTREE,
          preferedType == KDTREE_TREE_HEAP ? NABO_HIP_HEAP_STL : NABO_HIP_HEAP_LINEAR,
          additionalParameters);
    case KDTREE_CL_PT_IN_NODES:
    case KDTREE_CL_PT_IN_LEAVES:
    case BRUTE_FORCE_CL:
      throw std::runtime_error("OpenCL not found during compilation");
    default:
      throw std::runtime_error("Unknown search type");
  }
}

}  // namespace Nabo

#endif  // NABO_AMD_NABO_H
