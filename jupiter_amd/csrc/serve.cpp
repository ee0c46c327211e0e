// serve.cpp — continuous-batching request loop (placeholder until the
// persistent serving kernel lands).
#include "jgpm.h"

extern "C" int jgpm_serve_start(jgpm_ctx_t, const jgpm_serve_config*, jgpm_serve_t*) { return JGPM_E_STATE; }
extern "C" int jgpm_serve_submit(jgpm_serve_t, jgpm_graph_t, int, int64_t, int64_t, uint64_t*) { return JGPM_E_STATE; }
extern "C" int jgpm_serve_poll(jgpm_serve_t, uint64_t, uint64_t*, double*) { return JGPM_E_STATE; }
extern "C" int jgpm_serve_wait(jgpm_serve_t, uint64_t, uint64_t*, double*) { return JGPM_E_STATE; }
extern "C" int jgpm_serve_stop(jgpm_serve_t) { return JGPM_E_STATE; }
