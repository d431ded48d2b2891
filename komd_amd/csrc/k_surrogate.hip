// Learned three-body surrogate -- placeholder until the descriptor/MLP kernels land.
#include "komd_internal.hpp"

namespace komd {
struct Surrogate {};
hipError_t launch_surrogate_forces(komd_ctx*, bool, int) { return hipErrorNotSupported; }
hipError_t hybrid_forces(komd_ctx*, bool) { return hipErrorNotSupported; }
void surrogate_free(komd_ctx* c) { c->sur = nullptr; }
}  // namespace komd

extern "C" {
komd_status komd_surrogate_set_params(komd_ctx*, const float*, int64_t) { return KOMD_ERR_STATE; }
komd_status komd_surrogate_get_params(komd_ctx*, float*, int64_t) { return KOMD_ERR_STATE; }
komd_status komd_surrogate_descriptors(komd_ctx*, float*) { return KOMD_ERR_STATE; }
komd_status komd_surrogate_energies(komd_ctx*, float*) { return KOMD_ERR_STATE; }
komd_status komd_atm_site_energies(komd_ctx*, double*) { return KOMD_ERR_STATE; }
komd_status komd_train_add_snapshot(komd_ctx*) { return KOMD_ERR_STATE; }
komd_status komd_train_clear(komd_ctx*) { return KOMD_ERR_STATE; }
komd_status komd_train_size(komd_ctx*, int64_t*) { return KOMD_ERR_STATE; }
komd_status komd_train_standardize(komd_ctx*) { return KOMD_ERR_STATE; }
komd_status komd_train_epoch(komd_ctx*, int32_t, float, uint64_t, double*) { return KOMD_ERR_STATE; }
komd_status komd_set_hybrid(komd_ctx*, int32_t, double, double, double) { return KOMD_ERR_STATE; }
}
