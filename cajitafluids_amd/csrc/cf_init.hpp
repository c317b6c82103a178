// cf_init.hpp -- seeded initial-state generators.
#pragma once

#include <hip/hip_runtime.h>

#include "cf_geometry.hpp"
#include "cf_kernels.hpp"

namespace cfmi
{

constexpr int kMaxBlobs = 32;

struct InitParams
{
    int periodic[3];
    double box_lo[3], box_hi[3];
    int ntube;
    int tube_axis[3];
    double tube_a[3], tube_b[3], tube_sigma[3], tube_gamma[3];
    int nblob;
    double blob_c[kMaxBlobs][3], blob_sigma[kMaxBlobs], blob_amp[kMaxBlobs];
};

double host_rand_uniform(uint64_t seed, uint64_t stream, uint64_t idx);
void make_tube_params(const Geo& g, const cf_config& c, uint64_t seed,
                      InitParams& P);
void make_blob_params(const Geo& g, const cf_config& c, uint64_t seed,
                      int nblob, InitParams& P);
void launch_init_tubes(const Geo& g, const InitParams& P, Fields cur,
                       unsigned long long* vmax, double peak, hipStream_t s);
void launch_init_blobs(const Geo& g, const InitParams& P, double* q,
                       hipStream_t s);
void launch_init_random(const Geo& g, uint64_t seed, Fields cur,
                        hipStream_t s);

} // namespace cfmi
