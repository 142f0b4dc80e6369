// rtm_dev.h -- data layout shared by the HIP kernels and the host engine.
//
// One "slot" is one shot in flight on the GPU.  Its descriptor (ShotDev) lives
// in device memory and is read by every kernel through scalar loads; the tile
// table of a launch names (slot, tile origin) pairs, so one launch can carry
// many shots of different window widths (ragged batches) and, in the served
// form, shots at different time steps and phases.
//
// Field layout (per slot, fp32, row pitch `pitch` floats).  Field pointers
// address cell (0,0); every allocation has RTM_MZT rows above row 0, RTM_MXL
// columns left of column 0 and at least 32 rows / 32 columns beyond the tiled
// extent, so the radius-20 reads of a tile (velocities of the halo need the
// pressure 10 cells further out) never need bounds checks.
//   sp[2], svx[2], svz[2] : source wavefield, two time levels (step `it`
//                           reads set it&1 and writes set (it+1)&1)
//   rp[2], rvx[2], rvz[2] : receiver wavefield, same scheme
//   q1, q2                : split pressure pair in PML cells (in place; the
//                           source field's in the forward phase, the
//                           receiver field's in the reverse phase)
//   vx(iz,ix) sits at x=(ix+1/2)dx and uses columns 0..nxp-2; vz(iz,ix) at
//   z=(iz+1/2)dz uses rows 0..nzp-2 (kernel.cu:215-216 array shapes).
// The model is not copied per slot: the kernels read 1/rho and rho*v^2 (the
// factors of the velocity and pressure updates) from the model the slot
// names -- the resident model, shared by every shot of a batch, or a shot's
// own smoothed window -- at row clamp(iz-mz0, 0, mzn-1) and column
// mx0 + clamp(ix-mxs, 0, mxn-1): the window cut and edge replication of
// extractvel1 / extendvel1 (Velocity_function.cpp:2011-2073) as addressing,
// continued into the field margins.
// Damping factors are stored with the same margins (index -RTM_MXL.. of the
// column arrays, -RTM_MZT.. of the row arrays; 1.0 outside the grid) as
// a = c1/c2 and b = 1/c2 of the reference's c1 = 1 - dt*d/2, c2 = 2 - c1,
// so that the damped update (c1*v - u)/c2 is fma(a, v, -(u*b)); a = b = 1
// off the absorbing layer.
// Imaging accumulators g1, illum: nzi x nxl, pitch nxl.
// Boundary store: nt steps x strip_len values, packed per step as
//   vx lr [2][nzi][10] | vz tb [2][10][nxl] | p lr [2][nzi][10] | p tb [2][10][nxl]
// (the packing of kernel.cu:1186-1293), fp32 / fp16 / bf16, with one fp32
// scale per (step, block) for the 16-bit forms.
#pragma once
#include <stdint.h>

#define RTM_R 10            // stencil radius (N/2, common.h:57)
#define RTM_TX 64           // tile width (one wave64 row)
#define RTM_MZT 32          // margin rows above row 0 of every field
#define RTM_MXL 32          // margin columns left of column 0
#define RTM_ZALIGN 64       // field rows are allocated for nzp rounded up to this
#define RTM_MZB 32          // margin rows below the rounded extent

// tile flags (Tile::kind)
enum RtmTileFlag : int {
  TF_STRIP = 1,   // the tile or its halo holds boundary-strip cells (restored by the reconstruction)
  TF_SRC = 2,     // the tile holds source cells
  TF_REC = 4,     // the tile holds receiver cells
  TF_GEN = 8,     // the tile or its halo leaves the undamped interior (per-cell rules)
  TF_INT = 16,    // reverse phase: the tile holds reconstructed cells (interior or strips)
};

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
  // model window source (resident model or smoothed window): value(iz,ix) =
  //   m[clamp(iz-mz0,0,mzn-1)*mpitch + mx0 + clamp(ix-mxs,0,mxn-1)]
  const float* rho;
  const float* vv;             // velocity squared (extractvel1, Velocity_function.cpp:2059)
  const float* mib;            // 1/rho (rounded once)
  const float* mkk;            // rho*v^2 (rounded once)
  int mpitch, mz0, mzn, mx0, mxs, mxn, _pad2, _pad3;
  // wavefields
  float *sp[2], *svx[2], *svz[2];
  float *rp[2], *rvx[2], *rvz[2];
  float *q1, *q2;
  float *g1, *illum;
  // per-column / per-row damping factors a = c1/c2, b = 1/c2 at p nodes
  // (x, z) and at the velocity half nodes (xv, zv)
  const float *ax, *bx, *axv, *bxv, *az, *bz, *azv, *bzv;
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
  float* sscale;               // nt x 4 scales (16-bit formats), then the wavefield guard word
  long long strip_len;         // values per step
  int store_fmt, _pad4;
  // the other imaging conditions (null when off): ud du lr rl g4 (nzi x nxl
  // each, pitch nxl), then 61 bins of g31 and 61 of g32
  float* ic;
  float dx, dz;                // grid spacings (central differences of the energy norm)
  // decimated snapshots of the source wavefield (null when off): snapshot
  // q (step k = snap_every * (snap_k0 + q)) holds the interior cells of p
  // as forward step k leaves them and of vx, vz as forward step k+1 leaves
  // them ([q][3][nzi][nxl], store_fmt codes of snap_fmt), restored into the
  // reconstructed field after reverse step k; snap_scale [q][3] fp32
  // scales, then [q][3] max-abs words
  void* snap;
  float* snap_scale;
  int snap_every, snap_k0, nsnap, snap_fmt;
};

struct Tile {
  int slot, x0, z0, kind;
  // table positions of the 8 tiles around this one in the same slot (-1:
  // none), for the zero-tile skip of the step kernels
  int nbr[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
};

struct LaunchArgs {
  const ShotDev* shots;
  const Tile* tiles;
  int ntiles;
  int step;
  // zero-tile skip of k_fwd (nullptr: every tile steps): per tile and field
  // level one word whose byte g says that wave g stored a non-zero value of
  // the forward field at that level; two levels of ntiles words
  unsigned long long* act = nullptr;
};
