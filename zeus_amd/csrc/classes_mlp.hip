// classes_mlp.hip -- instantiations of the mlp classes.
// Dimensions offered through the C ABI: 321.
#include "ops.hpp"

namespace zeus::rt {

const ClassOps*
ops_mlp(int dim)
{
  switch (dim) {
    case 321: return Ops<obj::Mlp, 321>::table();
    default: return nullptr;
  }
}

} // namespace zeus::rt
