// classes_ackley.hip -- instantiations of the ackley classes.
// Dimensions offered through the C ABI: 2, 10.
#include "ops.hpp"

namespace zeus::rt {

const ClassOps*
ops_ackley(int dim)
{
  switch (dim) {
    case 2: return Ops<obj::Ackley<2>, 2>::table();
    case 10: return Ops<obj::Ackley<10>, 10>::table();
    default: return nullptr;
  }
}

} // namespace zeus::rt
