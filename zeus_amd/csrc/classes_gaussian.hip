// classes_gaussian.hip -- instantiations of the gaussian classes.
// Dimensions offered through the C ABI: 2, 5, 32, 150.
#include "ops.hpp"

namespace zeus::rt {

const ClassOps*
ops_gaussian(int dim)
{
  switch (dim) {
    case 2: return Ops<obj::Gaussian<2>, 2>::table();
    case 5: return Ops<obj::Gaussian<5>, 5>::table();
    case 32: return Ops<obj::Gaussian<32>, 32>::table();
    case 150: return Ops<obj::Gaussian<150>, 150>::table();
    default: return nullptr;
  }
}

} // namespace zeus::rt
