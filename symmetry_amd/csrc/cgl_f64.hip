// FP64 ("plain") instantiation.  Built with -ffp-contract=off so that every
// multiply and add rounds separately, as in the FP64 CPU path.
#include "cgl_impl.hpp"

template struct cgl::Impl<double, double, double, false>;
