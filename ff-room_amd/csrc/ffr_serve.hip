// ffr_serve.hip - batching and serving of many scenes on one device.
#include "ffr_scene.hpp"

struct ffr_ctx::Server {};

void ffr_ctx::serve_shutdown()
{
    delete server;
    server = nullptr;
}

extern "C" {
int ffr_submit(ffr_ctx *c, ffr_scene *, int, uint64_t *)
{
    if (c) c->fail("serving not available yet");
    return FFR_E_UNSUPPORTED;
}
int ffr_poll(ffr_ctx *c, uint64_t, ffr_step_result *, int *)
{
    if (c) c->fail("serving not available yet");
    return FFR_E_UNSUPPORTED;
}
int ffr_drain(ffr_ctx *c)
{
    if (c) c->fail("serving not available yet");
    return FFR_E_UNSUPPORTED;
}
}
