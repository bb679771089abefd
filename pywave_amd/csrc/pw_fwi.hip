// pw_fwi.hip -- FWI discretisation kernels (filled in below).
#include "pw_internal.h"
