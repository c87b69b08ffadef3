// Instantiates the row passes for double.
#define VFD_REAL double
#define VFD_ROW_TU
#include "vfd_launch_impl.hpp"
