// classes_loglik.hip -- instantiations of the loglik classes.
// Dimensions offered through the C ABI: 4.
#include "ops.hpp"

namespace zeus::rt {

const ClassOps*
ops_loglik(int dim)
{
  switch (dim) {
    case 4: return Ops<obj::LogLikelihood, 4>::table();
    default: return nullptr;
  }
}

} // namespace zeus::rt
