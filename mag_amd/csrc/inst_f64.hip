// Ops instantiation for FP64 links stored as 18 reals (the parity path).
#include "ops_impl.hpp"

template struct mag::Ops<double, 18>;
