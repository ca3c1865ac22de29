// Ops instantiation for FP32 links stored as 12 reals (the FP32
// molecular-dynamics path; row 2 is rebuilt as conj(row0 x row1) on load).
#include "ops_impl.hpp"

template struct mag::Ops<float, 12>;
