// Entry points not built yet: they fail loudly.
#include "vnm_internal.h"
using namespace vnm;
extern "C" {
int vnm_retained_nnz(vnm_graph *, int, int, int64_t *) { return set_error(VNM_E_ARG, "not built"); }
int vnm_pack(vnm_graph *, int, int, vnm_packed **) { return set_error(VNM_E_ARG, "not built"); }
int vnm_packed_info_get(const vnm_packed *, vnm_packed_info *) { return set_error(VNM_E_ARG, "not built"); }
void vnm_packed_destroy(vnm_packed *) {}
int vnm_spmm(vnm_packed *, const void *, int, int, int, const float *, float *, int) { return set_error(VNM_E_ARG, "not built"); }
}
