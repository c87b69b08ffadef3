// Instantiates the col passes for float.
#define VFD_REAL float
#define VFD_COL_TU
#include "vfd_launch_impl.hpp"
