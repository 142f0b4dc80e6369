// rtm_dev.h -- data layout shared by the HIP kernels and the host engine.
//
// One "slot" is one shot in flight on the GPU.  Its descriptor (ShotDev) lives
// in device memory and is read by every kernel through scalar loads; the tile
// table of a launch names (slot, tile origin) pairs, so one launch can carry
// many shots of different window widths (ragged batches) and, in the served
// form, shots at different time steps and phases.
//
// Wavefield layout (per slot, all fp32, row pitch `pitch` floats; field
// pointers address cell (0,0), with RTM_MZT zero rows above, RTM_MXL zero
// columns to the left and at least 16 rows / 25 columns of zeros beyond the
// tiled extent, so radius-10 tile staging never needs bounds checks):
//   sp, svx, svz : source wavefield  (forward propagation, then reverse
//                  reconstruction in place; rtm_model.cu:322-437)
//   sp1, sp2     : split pressure pair of the source wavefield in PML cells
//   rp, rvx, rvz, rp1, rp2 : receiver wavefield (rtm_model.cu:439-450)
//   vx(iz,ix) sits at x=(ix+1/2)dx and uses columns 0..nxp-2; vz(iz,ix) at
//   z=(iz+1/2)dz uses rows 0..nzp-2 (kernel.cu:215-216 array shapes).
// Imaging accumulators g1, illum: nzi x nxl, pitch nxl.
// Boundary store: nt steps x strip_len values, packed per step as
//   vx lr [2][nzi][10] | vz tb [2][10][nxl] | p lr [2][nzi][10] | p tb [2][10][nxl]
// (the packing of kernel.cu:1186-1293), fp32 / fp16 / bf16, with one fp32
// scale per (step, block) for the 16-bit forms.
#pragma once
#include <stdint.h>

#define RTM_R 10            // stencil radius (N/2, common.h:57)
#define RTM_TX 64           // tile width (one wave64 row)
#define RTM_MZT 16          // zero margin rows above row 0 of every field
#define RTM_MXL 16          // zero margin columns left of column 0
#define RTM_ZALIGN 32       // field rows are allocated for nzp rounded up to this + 16 below

// tile flags (Tile::kind): which cell classes a tile contains
enum RtmTileFlag : int { TF_STRIP = 1, TF_SRC = 2, TF_REC = 4 };

enum RtmPhase : int { PH_IDLE = 0, PH_FWD = 1, PH_REV = 2 };
enum RtmStoreFmt : int { ST_F32 = 0, ST_F16 = 1, ST_BF16 = 2 };

struct ShotDev {
  // geometry
  int nxp, nzp, npml, pitch;   // pitch: row stride of every field (floats)
  int nxl, nzi, nt, phase;
  int frows, forg;             // rows allocated per field; offset of (0,0) in it
  int it_off, it_dir;          // time index = it_off + it_dir * step
  int record, rcol0, nrec, _pad0;
  float dt, dtx, dtz, _pad1;
  // model window: value(iz,ix) = m[clamp(iz-mz0,0,mzn-1)*mpitch + mx0 + clamp(ix-mxs,0,mxn-1)]
  const float* rho;
  const float* vv;             // velocity squared (extractvel1, Velocity_function.cpp:2059)
  int mpitch, mz0, mzn, mx0, mxs, mxn, _pad2, _pad3;
  // wavefields
  float *sp, *svx, *svz, *sp1, *sp2;
  float *rp, *rvx, *rvz, *rp1, *rp2;
  float *g1, *illum;
  // per-column / per-row damping factors c1 = 1-dt*d/2, c2 = 2-c1
  const float *c1x, *c2x, *c1xv, *c2xv, *c1z, *c2z, *c1zv, *c2zv;
  const float* coef;           // [100], row r-1 = order 2r
  // source: a = dt*src inside the box [sx0,sx1] x [sz0,sz1]
  int sx0, sx1, sz0, sz1;
  const float* sa;
  const float* wav;
  // receivers
  float* rec;                  // nt x nrec
  const int* cut;              // mute: samples it <= cut[ir] are zero
  // boundary store
  void* store;
  float* sscale;               // nt x 4 scales (16-bit formats)
  unsigned int* smax;          // nt x 4 running max-abs bit patterns (16-bit formats)
  long long strip_len;         // values per step
  int store_fmt, _pad4;
};

struct Tile {
  int slot, x0, z0, kind;
};

struct LaunchArgs {
  const ShotDev* shots;
  const Tile* tiles;
  int ntiles;
  int step;
};
