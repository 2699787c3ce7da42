// Arguments and helpers shared by the frame-batched delay-and-sum kernels.
#pragma once

#include "common.hpp"

namespace rcb {

struct EnsembleArgs {
  const void* rc;  // packed (channels, angles, samples, frames) RC channels
  const void* cr;  // same for CR
  const float* scan;
  const float* xe;
  const float* ye;
  const float* angles;
  const float* starts;
  float* pd;
  float2* signal;
  float* weights;
  float2* workspace;
  int ns, nc, na, nt;
  long long nv;
  double c, fs, fd, fn;
};

enum Method { kOpw = 0, kXdoppler = 1, kRcFmas = 2, kDmas = 3, kStsw = 4, kApertures = 5 };

// How the LDS-tiled kernel multiplies samples by tap coefficients.
enum Coefficients {
  kCoefFloat = 0,   // samples widened to float32, lerp + phasor rotation in float32
  kCoefSplit = 1,   // 16-bit samples: dot2 with hi + lo 16-bit coefficients (float32-accurate)
  kCoefSingle = 2,  // 16-bit samples: dot2 with one 16-bit coefficient (mixed-precision fast form)
};

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X; give every
// XCD a contiguous run of tiles so that concurrent neighbouring tiles share
// that XCD's L2.  The grid is padded to a multiple of 8.
__device__ __forceinline__ long long xcd_tile(long long tiles) {
  long long per_xcd = (tiles + 7) / 8;
  long long b = blockIdx.x;
  return (b % 8) * per_xcd + b / 8;
}

// LDS-tiled OPW / separate-aperture IQ (das_tile.hip).  Returns
// hipErrorNotSupported when the shape is outside what the tiled kernel
// handles (the caller then uses the angle-major kernel), hipErrorInvalidValue
// for invalid storage codes.
hipError_t launch_opw_tile(const EnsembleArgs& p, int channel_storage, int coefficients, int signal_storage,
                           bool apertures);

// The same on a Cartesian grid (axes x, z, y; voxel id (ix * nz + iz) * ny +
// iy), tiled in 8 x 8 lateral patches; p.scan is not read.
hipError_t launch_opw_grid(const EnsembleArgs& p, const float* x, const float* z, const float* y, int nx, int nz,
                           int ny, int channel_storage, int coefficients, int signal_storage, bool apertures);

// OPW on a Cartesian grid on MFMA for FP16 / BF16 channels (das_mfma.hip):
// kCoefSingle (FP16 or BF16) or kCoefSplit (FP16).  Same outputs as
// launch_opw_grid up to float32 summation order.  hipErrorNotSupported for
// shapes or storages it does not handle (na > 16, nt > 224, complex64).
// RCABEAM_STORAGE_FP16_SPLIT channels (scaled FP16 hi + lo pairs) take the
// float32-accurate three-product form; absmax is then required (the scale).
hipError_t launch_opw_grid_mfma(const EnsembleArgs& p, const float* x, const float* z, const float* y, int nx, int nz,
                                int ny, int channel_storage, int coefficients, int signal_storage,
                                const unsigned* absmax = nullptr);

// OPW on a Cartesian grid for complex64 channels, float32 VALU
// interpolation (das_grid.hip): analytic windows, live-channel skipping.
// Same results as launch_opw_grid up to float32 summation order;
// hipErrorNotSupported for na > 16, nt > 208, odd nt or unaligned channels.
// apertures: separate RC and CR IQ into signal (voxels, 2, nt) complex64.
// q_begin / q_end (q_end >= 0) restrict the sums to channels [q_begin, q_end)
// of the 2 nc (RC then CR); accumulate adds them to the complex64 signal
// already there (the power is that of the total).
hipError_t launch_opw_grid_valu(const EnsembleArgs& p, const float* x, const float* z, const float* y, int nx,
                                int nz, int ny, int signal_storage, bool apertures, int q_begin = 0, int q_end = -1,
                                bool accumulate = false);

// Adjoint of the grid OPW for complex64 channels (das_tile.hip): accumulates
// the packed channel gradients grc / gcr and dL/dc (*gc) for the IQ gradient
// giq (nv, nt); p.rc / p.cr hold the forward's packed channels.
// workspace (optional, adjoint_workspace_bytes): the deterministic form
// (fixed-point gradient sums, fixed-order dL/dc), bit-identical run to run.
hipError_t launch_opw_adjoint_grid(const EnsembleArgs& p, const float* x, const float* z, const float* y, int nx,
                                   int nz, int ny, const float2* giq, float2* grc, float2* gcr, double* gc,
                                   void* workspace = nullptr);
size_t adjoint_workspace_bytes(const EnsembleArgs& p, int nx, int nz, int ny);

}  // namespace rcb
