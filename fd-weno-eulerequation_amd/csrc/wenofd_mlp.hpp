// wenofd_mlp.hpp — the switch MLP trainer handle (internal to the library):
// parameters, Adam moments and a device-resident sample set.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

struct wenofd_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* P = nullptr;  // parameters (kNParams floats)
    float* m = nullptr;  // Adam moments
    float* v = nullptr;
    long long adam_t = 0;
    // samples: features [cap][10], labels [cap]
    float* X = nullptr;
    unsigned char* Y = nullptr;
    long long n = 0, cap = 0, positives = 0;
    // scratch
    float* grad = nullptr;
    float* partial = nullptr;
    float* lpart = nullptr;
    int partial_blocks = 0;
    int* idx = nullptr;
    long long idx_cap = 0;
    float* losses = nullptr;
    long long losses_cap = 0;
    float* dense_X = nullptr;  // one direction's harvest, [faces][4][10]
    unsigned char* dense_Y = nullptr;
    long long dense_cap = 0;
    int* counts = nullptr;
    long long* offsets = nullptr;
    long long counts_cap = 0;
};

namespace wenofd {
// Grows the sample set to hold at least `need` samples (contents kept).
int mlp_reserve(wenofd_mlp* m, long long need);
// Makes room for one direction's dense harvest of `faces` faces.
int mlp_reserve_dense(wenofd_mlp* m, long long faces);
// Appends the kept samples of the dense harvest (nsamp = faces * 4 samples).
int mlp_append_dense(wenofd_mlp* m, long long nsamp, unsigned long long seed, double p_pos, double p_neg,
                     long long* added);
}  // namespace wenofd
