// Instantiates the row passes for float.
#define VFD_REAL float
#define VFD_ROW_TU
#include "vfd_launch_impl.hpp"
