// classes_rosenbrock.hip -- instantiations of the rosenbrock classes.
// Dimensions offered through the C ABI: 2, 4, 10, 50.
#include "ops.hpp"

namespace zeus::rt {

const ClassOps*
ops_rosenbrock(int dim)
{
  switch (dim) {
    case 2: return Ops<obj::Rosenbrock<2>, 2>::table();
    case 4: return Ops<obj::Rosenbrock<4>, 4>::table();
    case 10: return Ops<obj::Rosenbrock<10>, 10>::table();
    case 50: return Ops<obj::Rosenbrock<50>, 50>::table();
    default: return nullptr;
  }
}

} // namespace zeus::rt
