// Compile-time refusal of clouds whose type is not the search's CloudType
// (nabo/nabo.h:390-424). Compiled by tests/test_cpp_api.py with one of
// CLOUD_DYNAMIC / CLOUD_FIXED_ROWS and one of the FACTORY_* macros: the
// dynamic matrix must compile, the fixed-row one must not.
#include "nabo/nabo.h"

template <typename Cloud>
int make() {
  Cloud cloud = Cloud::Random(3, 50);
#if defined(FACTORY_CREATE)
  Nabo::NNSearchF* nns = Nabo::NNSearchF::create(cloud);
#elif defined(FACTORY_BRUTE_FORCE)
  Nabo::NNSearchF* nns = Nabo::NNSearchF::createBruteForce(cloud);
#elif defined(FACTORY_LINEAR_HEAP)
  Nabo::NNSearchF* nns = Nabo::NNSearchF::createKDTreeLinearHeap(cloud);
#else
  Nabo::NNSearchF* nns = Nabo::NNSearchF::createKDTreeTreeHeap(cloud);
#endif
  delete nns;
  return 0;
}

int main() {
#if defined(CLOUD_DYNAMIC)
  return make<Eigen::MatrixXf>();
#else
  return make<Eigen::Matrix<float, 4, Eigen::Dynamic> >();
#endif
}
