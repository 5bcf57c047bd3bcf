// Launches of the single-system pair-once kernels (direct_pairs.hip, compiled
// with the ILP-first instruction scheduler, which interleaves the pairs of
// the Oseen body: 140 VGPRs and three waves per SIMD, but no VALU
// instruction waiting on the one before it).
#pragma once
#include <hip/hip_runtime.h>

#include "direct_core.h"
#include "direct_sym.h"
#include "image32.h"

namespace dem {

// Kernel addresses (occupancy queries).
const void* force_sym_kernel();
const void* conv_sym_kernel(bool pos32);

// Force sums of blocks x chunks (grid B x S, kSymB / kSymTplForce threads):
// target block b, partner entries [s ec, (s + 1) ec); with img.slot the
// workgroups also insert their block's high words into the boundary-pair
// tables.
void launch_force_sym(dim3 grid, hipStream_t st, const Rec* rec, const RecC* recc, int n, int B,
                      int ec, int tile, double* part, PartJ* part_j, ImgArgs img);

// Oseen sums on 32-bit (pos32) or 64-bit separations; with a flag the kernel
// runs only when (*flag != 0) == run_if.
void launch_conv_sym(bool pos32, dim3 grid, hipStream_t st, const RecC* rec, int n, int B, int ec,
                     int tile, double* part, PartJ* part_j, const int* flag, int run_if);

}  // namespace dem
