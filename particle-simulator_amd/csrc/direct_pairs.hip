// Single-system pair-once direct kernels (Newton's third law): the force
// sums with the boundary-pair high-word inserts, and the Oseen sums on 32-bit
// or 64-bit separations.  Reference: direct_droplet_force_calculator.py
// _calculate_forces (:202-253), _calculate_convection_velocities (:354-468).
// This file is compiled with -mllvm -amdgpu-sched-strategy=max-ilp (build.py):
// the default scheduler left 96 of the Oseen loop's 258 VALU instructions
// waiting on the instruction before them (1.77-1.79 ms per launch at
// N = 65536); the ILP-first schedule has none (1.69-1.71 ms, same results).
#include "direct_pairs.h"

namespace {

using dem::ImgArgs;
using dem::img_insert;
using dem::kSymB;
using dem::kSymTplConv;
using dem::kSymTplForce;
using dem::rec_hi;
using dem::Rec;
using dem::RecC;
using dem::sym_partners;

template <int TPL>
__global__ __launch_bounds__(kSymB / TPL) void k_force_sym(const Rec* __restrict__ rec,
                                                           const RecC* __restrict__ recc, int n,
                                                           int B, int ec, int tile,
                                                           double* __restrict__ part,
                                                           dem::PartJ* __restrict__ part_j,
                                                           ImgArgs img) {
  const int b = blockIdx.x, s = blockIdx.y;
  if (img.slot) {
    // the S workgroups of block b share its 3 x 256 (droplet, axis) keys
    for (int u = s + (int)gridDim.y * (int)threadIdx.x; u < 3 * kSymB;
         u += (int)gridDim.y * (int)blockDim.x) {
      const int i = b * kSymB + u / 3, a = u % 3;
      if (i < n)
        img_insert(img.slot + ((size_t)a << img.log2m), img.log2m, rec_hi(recc[i], a), i);
    }
  }
  const int e0 = s * ec, e1 = min(1 + sym_partners(b, B), e0 + ec);
  dem::force_sym_block<TPL>(rec, n, B, b, s, e0, e1, tile, part, part_j);
}

// POS32: 32-bit separations in a 32-bit LDS layout (more workgroups per
// compute unit).  With a flag, the kernel runs only when (*flag != 0) equals
// run_if: the 32-bit kernel while the boundary-pair lookup holds, the 64-bit
// one as its fallback (both are launched; the other leaves at once).
template <int TPL, bool POS32>
__global__ __launch_bounds__(kSymB / TPL) void k_conv_sym(const RecC* __restrict__ rec, int n,
                                                          int B, int ec, int tile,
                                                          double* __restrict__ part,
                                                          dem::PartJ* __restrict__ part_j,
                                                          const int* __restrict__ flag,
                                                          int run_if) {
  if (flag && ((*flag != 0) != (run_if != 0))) return;
  __shared__ dem::ConvSymLds<kSymB / TPL / 64, POS32> L;
  const int b = blockIdx.x, s = blockIdx.y;
  const int e0 = s * ec, e1 = min(1 + sym_partners(b, B), e0 + ec);
  dem::conv_sym_block<TPL, POS32>(L, rec, n, B, b, s, e0, e1, tile, part, part_j);
}

}  // namespace

namespace dem {

const void* force_sym_kernel() { return (const void*)k_force_sym<kSymTplForce>; }

const void* conv_sym_kernel(bool pos32) {
  return pos32 ? (const void*)k_conv_sym<kSymTplConv, true>
               : (const void*)k_conv_sym<kSymTplConv, false>;
}

void launch_force_sym(dim3 grid, hipStream_t st, const Rec* rec, const RecC* recc, int n, int B,
                      int ec, int tile, double* part, PartJ* part_j, ImgArgs img) {
  hipLaunchKernelGGL(k_force_sym<kSymTplForce>, grid, dim3(kSymB / kSymTplForce), 0, st, rec,
                     recc, n, B, ec, tile, part, part_j, img);
}

void launch_conv_sym(bool pos32, dim3 grid, hipStream_t st, const RecC* rec, int n, int B, int ec,
                     int tile, double* part, PartJ* part_j, const int* flag, int run_if) {
  if (pos32)
    hipLaunchKernelGGL((k_conv_sym<kSymTplConv, true>), grid, dim3(kSymB / kSymTplConv), 0, st,
                       rec, n, B, ec, tile, part, part_j, flag, run_if);
  else
    hipLaunchKernelGGL((k_conv_sym<kSymTplConv, false>), grid, dim3(kSymB / kSymTplConv), 0, st,
                       rec, n, B, ec, tile, part, part_j, flag, run_if);
}

}  // namespace dem
