// FP32 shifted-storage and FP16 shifted-storage (FP32 collide) instantiations.
#include "cgl_impl.hpp"

template struct cgl::Impl<float, float, float, true>;
template struct cgl::Impl<_Float16, float, float, true>;
