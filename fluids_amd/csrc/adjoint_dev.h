// Device-side records of the WCSPH adjoint: one step's tape (the forward
// values its backward pass reads, sorted order) and the per-step work
// arrays of adjoints.
#pragma once

#include "kernels.h"

namespace sph {

struct AdjTape {
    const float4* prho;     // sorted positions, density in w
    const float4* vel_s;    // sorted velocities
    const float4* vh_s;     // sorted half-step velocities
    const float* pressure;
    const float4* normal;
    const float4* accel;
    const int* nl;          // lists, column-major with stride n
    const int* nc;
    const int* perm;        // sorted index -> stored index of the step's input
};

struct AdjWork {
    float4 *abar, *xbar, *vbar, *vhbar, *nbar;
    float *rbar, *pbar, *part_a, *part_b;
};

int adj_blocks(int n);
// workgroups of the adjoint kernels that split each particle's list over lanes
int adj_split_blocks(int n);
// Backward pass of one WCSPH step: (gx, gv, gvh) are the adjoints of the
// step's outputs (stored order after the step = sorted order of the step),
// (gx_out, gv_out, gvh_out) those of its inputs (stored order before the
// step). params accumulates d/dk_viscosity, d/dsurface_tension and
// d/dgas_stiffness (in that order); capped counts particles whose list
// reached the cap.
void launch_adjoint_step(int n, const AdjTape& T, const WcParams& P, float cn, float nH2, float c_rho, float h2,
                         int fix, float rho0, float k_gas, const float4* gx, const float4* gv, const float4* gvh,
                         AdjWork& W, float4* gx_out, float4* gv_out, float4* gvh_out, double* params, int* capped,
                         hipStream_t s);

}  // namespace sph
