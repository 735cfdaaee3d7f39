// Entry points whose implementation lands in later files of this build step.
#include "runtime.h"

extern "C" {
static int unsup(const char* what) { (void)what; return OME_ERR_UNSUPPORTED; }
int ome_server_create(ome_ctx*, size_t, int, ome_server**) { return unsup("server"); }
int ome_server_submit(ome_server*, ome_tree*, const void* const*, void*, int64_t*) { return unsup("server"); }
int ome_server_poll(ome_server*, int64_t, int*) { return unsup("server"); }
int ome_server_wait(ome_server*, int64_t, double*) { return unsup("server"); }
int ome_server_destroy(ome_server*) { return unsup("server"); }
}
