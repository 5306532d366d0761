// classes_sphere.hip -- instantiations of the sphere classes.
// Dimensions offered through the C ABI: 1, 2, 3.
#include "ops.hpp"

namespace zeus::rt {

const ClassOps*
ops_sphere(int dim)
{
  switch (dim) {
    case 1: return Ops<obj::Sphere<1>, 1>::table();
    case 2: return Ops<obj::Sphere<2>, 2>::table();
    case 3: return Ops<obj::Sphere<3>, 3>::table();
    default: return nullptr;
  }
}

} // namespace zeus::rt
