// qk_serving.cpp -- serving loop (placeholder until the scheduler lands).
#include "qkrylov_amd.h"

extern "C" int qkrylov_serve_run(const qkrylov_serve_request_t*, int, const qkrylov_serve_options_t*,
                                 qkrylov_serve_result_t*, double*, double*, int, double*)
{
    return QKRYLOV_ERROR_EXCEPTION;
}
