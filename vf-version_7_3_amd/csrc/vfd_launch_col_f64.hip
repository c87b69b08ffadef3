// Instantiates the col passes for double.
#define VFD_REAL double
#define VFD_COL_TU
#include "vfd_launch_impl.hpp"
