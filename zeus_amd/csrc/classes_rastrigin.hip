// classes_rastrigin.hip -- instantiations of the rastrigin classes.
// Dimensions offered through the C ABI: 1, 2, 20.
#include "ops.hpp"

namespace zeus::rt {

const ClassOps*
ops_rastrigin(int dim)
{
  switch (dim) {
    case 1: return Ops<obj::Rastrigin<1>, 1>::table();
    case 2: return Ops<obj::Rastrigin<2>, 2>::table();
    case 20: return Ops<obj::Rastrigin<20>, 20>::table();
    default: return nullptr;
  }
}

} // namespace zeus::rt
