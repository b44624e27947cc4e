// mg_kernels.hip - multigrid kernels (see ffr_mg.hpp).
#include "../ffr_kernels.hpp"
