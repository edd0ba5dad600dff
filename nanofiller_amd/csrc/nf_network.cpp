// nf_network.cpp - host orchestration of network extraction (placeholder
// until the GPU union-find and CSR kernels land).
#include "nf_host.h"

using namespace nf;

extern "C" int nf_network(nf_handle* h, nf_network_info* info) {
  (void)info;
  return fail(h, NF_E_STATE, "network extraction not built yet");
}

extern "C" int nf_get_network(nf_handle* h, uint32_t* labels, uint32_t* row_ptr, uint32_t* col,
                              double* val, uint64_t cap) {
  (void)labels; (void)row_ptr; (void)col; (void)val; (void)cap;
  return fail(h, NF_E_STATE, "network extraction not built yet");
}
