// Entry points whose kernels are not in this build yet: they report
// GPGMP_E_UNSUPPORTED so that every symbol of include/gpgmp_hip.h is exported.
#include "host.h"

extern "C" {
#define GPGMP_PENDING(sig) \
    int sig { return GPGMP_E_UNSUPPORTED; }
GPGMP_PENDING(gpgmp_min_timestep_device(const float *, const float *, int32_t, float, float *, void *))
GPGMP_PENDING(gpgmp_sort_pairs_device(uint32_t *, uint32_t *, uint32_t *, uint32_t *, int32_t, int32_t, void *))
GPGMP_PENDING(gpgmp_cell_offsets_device(const uint32_t *, int32_t, int32_t, uint32_t *, void *))
GPGMP_PENDING(gpgmp_cell_left_right_device(const uint32_t *, int32_t, int32_t, uint32_t *, uint32_t *, void *))
GPGMP_PENDING(gpgmp_particles_create(const gpgmp_particle_params *, int32_t, void *, gpgmp_particles **))
GPGMP_PENDING(gpgmp_particles_destroy(gpgmp_particles *))
GPGMP_PENDING(gpgmp_particles_set(gpgmp_particles *, const float *, const float *, const uint32_t *, int32_t))
GPGMP_PENDING(gpgmp_particles_get(gpgmp_particles *, float *, float *, uint32_t *, uint32_t *, uint32_t *, int32_t *))
GPGMP_PENDING(gpgmp_particles_count(gpgmp_particles *, int32_t *))
GPGMP_PENDING(gpgmp_particles_step(gpgmp_particles *, int32_t, int32_t))
}
