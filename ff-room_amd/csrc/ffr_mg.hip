// ffr_mg.hip - multigrid-preconditioned CG (see ffr_mg.hpp).
#include "ffr_scene.hpp"

namespace ffr {

size_t mg_bytes(int, int, int, const ffr_params &) { return 0; }
int mg_create(ffr_scene *) { return FFR_OK; }
void mg_destroy(ffr_scene *) {}
int mg_solve(ffr_scene *s, int *, int *, double *)
{
    s->ctx->fail("MG-PCG not available yet");
    return FFR_E_UNSUPPORTED;
}

}  // namespace ffr
