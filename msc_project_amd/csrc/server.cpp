// Continuous-batching flow server (see include/lgf_flow.h).
#include "../../include/lgf_flow.h"

extern "C" int lgf_server_create(int, size_t, int, lgf_server** out) {
    if (out) *out = nullptr;
    return LGF_E_UNSUPPORTED;
}
extern "C" int lgf_server_destroy(lgf_server*) { return LGF_E_UNSUPPORTED; }
extern "C" int lgf_server_submit(lgf_server*, const lgf_request*, uint64_t*) { return LGF_E_UNSUPPORTED; }
extern "C" int lgf_server_wait(lgf_server*, uint64_t, double, lgf_result*) { return LGF_E_UNSUPPORTED; }
extern "C" int lgf_server_now(lgf_server*, double*) { return LGF_E_UNSUPPORTED; }
extern "C" int lgf_server_stats(lgf_server*, uint64_t*, uint64_t*, double*, double*) { return LGF_E_UNSUPPORTED; }
extern "C" const char* lgf_server_last_error(const lgf_server*) { return "not built"; }
