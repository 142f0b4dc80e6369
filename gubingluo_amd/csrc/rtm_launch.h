// rtm_launch.h -- host-side launchers of the kernels in rtm_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "rtm_dev.h"

// one time step of the tiles in `a` (one launch): phase PH_FWD (source
// modelling) or PH_REV (reconstruction, receiver wavefield, imaging); tz the
// tile height (16, 32 or 64)
hipError_t rtm_launch_step(const LaunchArgs& a, int phase, int tz, hipStream_t st);
// after a forward step: its boundary strips into the store of the listed
// slots (fp32 copy, or 16-bit with per-step block scales; one workgroup per
// slot and strip block)
hipError_t rtm_launch_save(const ShotDev* shots, const int* slots, int nslots, int step, int store_fmt,
                           hipStream_t st);
hipError_t rtm_launch_phase_init(const ShotDev* shots, int slot, int phase, int nblocks, hipStream_t st);
// 1/rho and rho v^2 of n model cells (ib = 1/rho, kk = rho v^2, each rounded once)
hipError_t rtm_launch_model_prep(const float* rho, const float* vv, float* ib, float* kk, long long n, hipStream_t st);
hipError_t rtm_launch_seed(const ShotDev* shots, int slot, int nrec, hipStream_t st);
hipError_t rtm_launch_stack(const void* stack_shots, int nsh, int nx, int nz, double* i1, double* i2, double* il,
                            hipStream_t st);
hipError_t rtm_launch_round(const double* a, float* o, long long n, hipStream_t st);
// the other imaging conditions after reverse step `step` of every slot listed
hipError_t rtm_launch_ic(const ShotDev* shots, const int* slots, int nslots, int step, int max_cells, hipStream_t st);
// their FP64 stacks: i7 = image3s, image4s, image5s, image3m, image4m, image5m, imagetrue
struct StackIcHost { const float* ic; const float* illum; const float* taper; int nx1, nxl; };
hipError_t rtm_launch_stack_ic(const void* stack_shot, int nx, int nz, double* i7, hipStream_t st);
hipError_t rtm_launch_laplace(const float* im, float* out, int nx, int nz, int dxi, int dzi, hipStream_t st);
hipError_t rtm_launch_norm_multi(const float* sg, const float* ss, float* out, long long n, hipStream_t st);
hipError_t rtm_launch_zero_f64(double* a, long long n, hipStream_t st);
hipError_t rtm_launch_bump(ShotDev* shots, const int* slots, int nslots, int k, hipStream_t st);
// pmlvelsmooth1d of n windows in place, one workgroup each (wins: device
// array of SmoothWinHost)
struct SmoothWinHost { float* w; float* v; int nx, nz, npml, _pad; };
hipError_t rtm_launch_pml_smooth(const void* wins, int n, hipStream_t st);
// velsmooth1d of an nz x nx window (out-of-place): weights wt [(2nsp+1)^2]
// (device, double, the host's table) and their sum `norm`
hipError_t rtm_launch_gauss_smooth(const float* in, float* out, int nz, int nx, int nsp, const double* wt, double norm,
                                   hipStream_t st);
// decimated snapshots after step `step` of the listed slots: FWD saves the
// snapshot of time it+1 (if any), REV restores the one of time it (if any)
hipError_t rtm_launch_snap(const ShotDev* shots, const int* slots, int nslots, int phase, int step, int snap_fmt,
                           hipStream_t st);
// wavefield guard of the listed slots after step `step` (largest |p| as float
// bits into each slot's guard word and into *flag; flag may be null)
hipError_t rtm_launch_guard(const ShotDev* shots, const int* slots, int nslots, int step, unsigned int* flag,
                            hipStream_t st);
// a slot's guard word (device address)
inline unsigned int* rtm_guard_word(const ShotDev& d) { return (unsigned int*)(d.sscale + (size_t)d.nt * 4); }

// per-shot entry of the stacking kernel (k_stack)
struct StackShotHost { const float* g1; const float* illum; const float* taper; int nx1, nxl; };
