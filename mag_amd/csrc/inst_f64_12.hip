// Ops instantiation for FP64 links stored as 12 reals (rows 0 and 1; row 2
// rebuilt as conj(row0 x row1) on load): the FP64 path with two thirds of
// the link bytes and load instructions of the 18-real form.
#include "ops_impl.hpp"

template struct mag::Ops<double, 12>;
