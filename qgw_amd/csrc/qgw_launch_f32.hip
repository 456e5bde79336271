// qgw_launch_f32.hip -- FP32 instantiation of the step kernels (the FP32 and
// compensated-FP32 forms).
#define QGW_T float
#include "launch_impl.hpp"
