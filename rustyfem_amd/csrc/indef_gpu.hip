// Placeholder until the GPU indefinite / LU kernels land: refuses loudly.
#include <stdexcept>

#include "indef.hpp"

namespace rf {

std::unique_ptr<IndefGpu> make_indef_gpu(Symbolic, const std::vector<int64_t>&, const std::vector<int64_t>&, bool) {
  throw std::runtime_error("indefinite/LU factorization is not available on the GPU in this build");
}

}  // namespace rf
