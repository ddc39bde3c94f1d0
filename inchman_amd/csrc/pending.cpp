// Entry points whose kernels are not in this build yet: they report
// GPGMP_E_UNSUPPORTED so that every symbol of include/gpgmp_hip.h is exported.
#include "host.h"

extern "C" {
#define GPGMP_PENDING(sig) \
    int sig { return GPGMP_E_UNSUPPORTED; }
GPGMP_PENDING(gpgmp_min_timestep_device(const float *, const float *, int32_t, float, float *, void *))
GPGMP_PENDING(gpgmp_rdd_create(const gpgmp_rdd_params *, int32_t, int32_t, void *, gpgmp_rdd **))
GPGMP_PENDING(gpgmp_rdd_destroy(gpgmp_rdd *))
GPGMP_PENDING(gpgmp_rdd_forward(gpgmp_rdd *, const float *, float *, float *, int32_t))
GPGMP_PENDING(gpgmp_rdd_backward(gpgmp_rdd *, const float *, float *, float *, float *, float *))
GPGMP_PENDING(gpgmp_det_hom_diffuse_device(float *, int32_t, int32_t, int32_t, const float *, float, float, void *))
GPGMP_PENDING(gpgmp_det_inhom_sweep_device(float *, int32_t, int32_t, int32_t, int32_t, float, const float *,
                                           const float *, float, void *))
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
