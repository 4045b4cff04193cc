// placeholder, replaced below
#include "bs_internal.hpp"
