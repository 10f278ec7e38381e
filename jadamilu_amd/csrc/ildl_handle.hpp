// ildl_handle.hpp -- the opaque jdhip_ildl handle of the C-ABI: one host
// multilevel factor (see include/jadamilu_amd.h).
#pragma once
#include "ildl.hpp"

struct jdhip_ildl {
    jdhip::IldlHost h;
};
