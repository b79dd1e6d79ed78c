// Launch descriptors of the rigid-body scan kernels (rigid.hip).
#pragma once

#include "device_api.hpp"

namespace ausaxs::gpu {

// One slot per body point in the plan's class-sorted order.
struct TransformLaunch {
    const double* xyz;          // [n_points][3] FP64 body coordinates, original order
    const double* T;            // [n_conf][n_bodies][12] = R (row-major 3x3), t
    const int32_t* slot_src;    // point index in xyz
    const int32_t* slot_body;   // body of the slot
    const int32_t* slot_dst;    // position in pts (shared region or replica 0)
    const int32_t* slot_dsto;   // position in orig (original-order copy)
    const int32_t* slot_tag;    // tag: original index within the body
    const int32_t* slot_cls;    // weight class
    const int32_t* body_moving; // [n_bodies] 1 if the body moves between conformations
    int n_slots, n_bodies, n_conf;
    int64_t rep_stride;         // points per replica
    float4* pts;
    float4* orig;
    // weighted scans: the weight pass's copies of pts and orig, with the
    // point's weight (slot_w) in the fourth slot
    const float* slot_w = nullptr;
    float4* pts_w = nullptr;
    float4* orig_w = nullptr;
};
hipError_t launch_transform(const TransformLaunch& p, hipStream_t s);

struct FinalizeLaunch {
    const uint64_t* cnt;   // [(n_conf + 1) * R][nbins + 1]: conformation rows, then the shared rows
    const int64_t* dsum;
    int R, nbins, n_conf;
    const int32_t* row_kind;  // [R] 0 aa, 1 aw, 2 ww (waters that change per conformation)
    const double* row_w;      // [R] float(w_a * w_b) of the row
    const double* base;       // [6][base_nb]: aa value, count, sum d; ww value, count, sum d
    int base_nb;
    double inv_w;             // float inverse width, widened
    double width;             // axis bin width ((bin_count * bin_width) / bin_count)
    double cw;                // water scaling
    int weighted;
    double* H;                // [n_conf][nbins]
    double* dbar;             // [n_conf][nbins]
    int* max_bin;             // [n_conf]
    int* flags;               // [n_conf] 1: a pair fell beyond nbins
    // weighted scans (more than 64 distinct weights): values from the weight
    // pass's sums instead of count * row_w (see weight_value)
    const int64_t* wsum = nullptr;    // same layout as cnt: items' FP64 sums (bits)
    const int64_t* wstrip = nullptr;  // strips' integer sums
    double wscale = 1;
    // per-conformation waters: [n_conf][2] self term of each water set (sum of
    // w^2 and its unsigned cast, SimpleCalculator.h:144-160), added at bin 0
    // after the ww rows; the base then holds no water terms
    const double* ww_self = nullptr;
};
hipError_t launch_finalize(const FinalizeLaunch& p, hipStream_t s);

struct DebyeBatchLaunch {
    const double* H;
    const double* dbar;
    const int* max_bin;
    const double* q;
    int nbins, n_conf, nq;
    double* I;  // [n_conf][nq]
};
hipError_t launch_debye_batch(const DebyeBatchLaunch& p, hipStream_t s);

}  // namespace ausaxs::gpu
