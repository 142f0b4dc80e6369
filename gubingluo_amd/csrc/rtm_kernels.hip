// rtm_kernels.hip -- gfx950 kernels of the RTM hot path.
//
// Per time step two phases are launched over a tile table that may span many
// shots (ragged windows):
//   velocity phase : vx, vz from p       (kernel.cu:430-500, 542-713 forward;
//                                          kernel.cu:318-388 reverse)
//   pressure phase : p from vx, vz, PML split pair, source, record, strips,
//                    receiver insertion, imaging (kernel.cu:501-541, 714-1185,
//                    36-95, 1186-1416)
// A tile is 64 columns (one wave64 row: every array is x-fastest, so each
// wave touches one contiguous 256-byte row segment) by TZ rows.  The operand
// field is staged once in LDS with its radius-10 halo; x-derivatives read LDS
// rows, z-derivatives a register window of the thread's own column (each
// thread owns TZ/4 consecutive rows).  Fields carry zero margins (MZT rows
// above, MXL columns left, more below/right), so staging needs no bounds
// checks.  Per-tile flags computed on the host keep the strip / source /
// receiver logic out of tiles that do not touch them.
//
// The per-cell arithmetic follows the reference operation order with the
// fused multiply-adds written explicitly (the oracle's canonical FP32 form),
// so results match the CPU restatement bit for bit.  Compiled with
// -ffp-contract=off; float division is IEEE as HIP compiles it.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "rtm_dev.h"

#define LW 83              // 64 + 19: staged width incl. the x halo
#define NTHR 256

// Pointers read from the descriptor are generic to the compiler, which would
// then emit flat loads/stores (they count against the LDS wait counter too,
// so every LDS wait would also wait for outstanding memory loads).  All
// device-memory accesses go through these global-address-space views.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) int gint;
typedef __attribute__((address_space(1))) unsigned int guint;
typedef __attribute__((address_space(1))) unsigned short gushort;
#define GF(p) ((gfloat*)(p))
#define GCF(p) ((const gfloat*)(p))

namespace {

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

// blockIdx -> tile: workgroups are dealt to the 8 XCDs round-robin, so give
// each XCD a contiguous run of the tile table (neighbouring tiles of one shot
// share their halos through that XCD's L2).
__device__ __forceinline__ int xcd_tile(int ntiles)
{
  const int per = gridDim.x >> 3;          // launched as a multiple of 8
  const int b = blockIdx.x;
  const int t = (b & 7) * per + (b >> 3);
  return t < ntiles ? t : -1;
}

__device__ __forceinline__ float model_at(const float* __restrict__ m, const ShotDev& s, int iz, int ix)
{
  const int r = imin(imax(iz - s.mz0, 0), s.mzn - 1);
  const int c = s.mx0 + imin(imax(ix - s.mxs, 0), s.mxn - 1);
  return GCF(m)[(size_t)r * s.mpitch + c];
}

// stage f rows zs..zs+H-1, columns xs..xs+W-1 (W <= 128) into L[H][W]; all
// global loads are issued before the LDS stores.
template <int W, int H>
__device__ __forceinline__ void stage(float* __restrict__ L, const float* __restrict__ f, int pitch, int zs, int xs)
{
  constexpr int K = (H + 3) / 4;
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  float a[K], b[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int r = ly + 4 * k;
    const gfloat* src = GCF(f) + (ptrdiff_t)(zs + imin(r, H - 1)) * pitch + xs;
    a[k] = src[lx];
    b[k] = (W > 64 && lx < W - 64) ? src[64 + lx] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int r = ly + 4 * k;
    if (r < H) {
      L[r * W + lx] = a[k];
      if (W > 64 && lx < W - 64) L[r * W + 64 + lx] = b[k];
    }
  }
}

// ---------------------------------------------------------------- strips --
// store offsets of the boundary strips (kernel.cu:1186-1293 packing); -1 if
// the cell is not part of the block.
__device__ __forceinline__ long long strip_vx(const ShotDev& s, int iz, int ix)
{
  if (iz < s.npml || iz >= s.nzp - s.npml) return -1;
  const int r = iz - s.npml;
  if (ix >= s.npml - RTM_R && ix < s.npml) return (long long)r * RTM_R + ix - (s.npml - RTM_R);
  const int c0 = s.nxp - s.npml - 1;
  if (ix >= c0 && ix < c0 + RTM_R) return (long long)RTM_R * s.nzi + (long long)r * RTM_R + ix - c0;
  return -1;
}
__device__ __forceinline__ long long strip_vz(const ShotDev& s, int iz, int ix)
{
  if (ix < s.npml || ix >= s.nxp - s.npml) return -1;
  const long long base = 2LL * RTM_R * s.nzi;
  const int c = ix - s.npml;
  if (iz >= s.npml - RTM_R && iz < s.npml) return base + (long long)(iz - (s.npml - RTM_R)) * s.nxl + c;
  const int r0 = s.nzp - s.npml - 1;
  if (iz >= r0 && iz < r0 + RTM_R) return base + (long long)RTM_R * s.nxl + (long long)(iz - r0) * s.nxl + c;
  return -1;
}
__device__ __forceinline__ long long strip_p(const ShotDev& s, int iz, int ix)
{
  const long long base_lr = 2LL * RTM_R * s.nzi + 2LL * RTM_R * s.nxl;
  const long long base_tb = base_lr + 2LL * RTM_R * s.nzi;
  if (iz >= s.npml && iz < s.nzp - s.npml) {
    const int r = iz - s.npml;
    if (ix >= s.npml - RTM_R && ix < s.npml) return base_lr + (long long)r * RTM_R + ix - (s.npml - RTM_R);
    const int c0 = s.nxp - s.npml;
    if (ix >= c0 && ix < c0 + RTM_R) return base_lr + (long long)RTM_R * s.nzi + (long long)r * RTM_R + ix - c0;
    return -1;
  }
  if (ix < s.npml || ix >= s.nxp - s.npml) return -1;
  const int c = ix - s.npml;
  if (iz >= s.npml - RTM_R && iz < s.npml) return base_tb + (long long)(iz - (s.npml - RTM_R)) * s.nxl + c;
  const int r0 = s.nzp - s.npml;
  if (iz >= r0 && iz < r0 + RTM_R) return base_tb + (long long)RTM_R * s.nxl + (long long)(iz - r0) * s.nxl + c;
  return -1;
}

// 16-bit store: block of a strip offset (0 vx, 1 vz, 2 p-lr, 3 p-tb)
__device__ __forceinline__ int strip_block(const ShotDev& s, long long off)
{
  const long long a = 2LL * RTM_R * s.nzi, b = a + 2LL * RTM_R * s.nxl, c = b + 2LL * RTM_R * s.nzi;
  return off < a ? 0 : (off < b ? 1 : (off < c ? 2 : 3));
}

__device__ __forceinline__ void store_put(const ShotDev& s, int it, long long off, float v)
{
  if (s.store_fmt == ST_F32) {
    GF(s.store)[(long long)it * s.strip_len + off] = v;
  } else {
    // fp32 staging of step it (by parity), quantised by k_quant after the step
    gfloat* stg = GF(s.store) + (long long)s.nt * s.strip_len / 2 + (long long)(it & 1) * s.strip_len;
    stg[off] = v;
    __atomic_fetch_max((guint*)&s.smax[(long long)it * 4 + strip_block(s, off)], __float_as_uint(fabsf(v)),
                       __ATOMIC_RELAXED);
  }
}

__device__ __forceinline__ float store_get(const ShotDev& s, int it, long long off)
{
  const long long k = (long long)it * s.strip_len + off;
  if (s.store_fmt == ST_F32) return GCF(s.store)[k];
  const float sc = GCF(s.sscale)[(long long)it * 4 + strip_block(s, off)];
  if (s.store_fmt == ST_F16) return __half2float(__ushort_as_half(((const gushort*)s.store)[k])) * sc;
  const unsigned short h = ((const gushort*)s.store)[k];
  return __uint_as_float(((unsigned int)h) << 16) * sc;
}

// radius-r staggered difference a[i] - b[i] with the coefficient row of order
// 2r; the fma chain runs i = 1..r as the reference loops (kernel.cu:459-460)
#define STENCIL_SUM(acc, r, c10, cf, A, B)                                                 \
  do {                                                                                     \
    if ((r) == RTM_R) {                                                                    \
      _Pragma("unroll") for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], (A) - (B), acc); \
    } else {                                                                               \
      const gfloat* cr_ = GCF(cf) + ((r) - 1) * RTM_R;                                     \
      for (int i = 1; i <= (r); i++) acc = __builtin_fmaf(cr_[i - 1], (A) - (B), acc);     \
    }                                                                                      \
  } while (0)

// ------------------------------------------------------- velocity phase --
// Forward velocity update of one wavefield over a tile of the padded grid:
//   vx(iz,ix), ix <= nxp-2, radius min(10, ix+1, nxp-1-ix)   (norderx, kernel.cu:299-307)
//   vz(iz,ix), iz <= nzp-2, radius min(10, iz+1, nzp-1-iz)   (norderz, kernel.cu:309-317)
//   v = (c1*v - D)/c2, D = (sum c_i (p[+i] - p[-i+1]) * dt/h) / rho  (kernel.cu:588-594);
//   c1 = c2 = 1 off the absorbing layer reduces it to v - D (kernel.cu:462).
// L holds p rows z0-9 .. z0+TZ+9, columns x0-9 .. x0+73.
template <int TZ>
__device__ __forceinline__ void vel_fwd(const float* __restrict__ L, float* __restrict__ vx, float* __restrict__ vz,
                                        const ShotDev& s, int x0, int z0, int save_it, int flags,
                                        const float (&vxo)[TZ / 4], const float (&vzo)[TZ / 4],
                                        const float (&rho)[TZ / 4])
{
  constexpr int RPT = TZ / 4;
  const int lx = threadIdx.x & 63, j0 = (threadIdx.x >> 6) * RPT;
  const int ix = x0 + lx;
  const int nxp = s.nxp, nzp = s.nzp, P = s.pitch;
  if (ix >= nxp) return;
  const float* __restrict__ cf = s.coef;
  float c10[RTM_R];
#pragma unroll
  for (int i = 0; i < RTM_R; i++) c10[i] = GCF(cf)[90 + i];
  const bool hasvx = ix <= nxp - 2;
  const int rx = imin(RTM_R, imin(ix + 1, nxp - 1 - ix));
  const float c1x = GCF(s.c1xv)[ix], c2x = GCF(s.c2xv)[ix];
  const bool dampx = !(c1x == 1.0f && c2x == 1.0f);
  float col[RPT + 19];                                 // p(z0+j0-9+k, ix)
#pragma unroll
  for (int k = 0; k < RPT + 19; k++) col[k] = L[(j0 + k) * LW + lx + 9];
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const int iz = z0 + j0 + r;
    if (iz >= nzp) break;
    const size_t id = (size_t)iz * P + ix;
    if (hasvx) {
      const float* __restrict__ Lr = L + (j0 + r + 9) * LW + lx + 9;
      float acc = 0.0f;
      STENCIL_SUM(acc, rx, c10, cf, Lr[i], Lr[1 - i]);
      acc = acc * s.dtx;
      acc = acc / rho[r];
      const float v = dampx ? __builtin_fmaf(c1x, vxo[r], -acc) / c2x : vxo[r] - acc;
      GF(vx)[id] = v;
      if (save_it >= 0 && (flags & TF_STRIP)) {
        const long long o = strip_vx(s, iz, ix);
        if (o >= 0) store_put(s, save_it, o, v);
      }
    }
    if (iz <= nzp - 2) {
      const int rz = imin(RTM_R, imin(iz + 1, nzp - 1 - iz));
      const float c1z = GCF(s.c1zv)[iz], c2z = GCF(s.c2zv)[iz];
      float acc = 0.0f;
      STENCIL_SUM(acc, rz, c10, cf, col[r + 9 + i], col[r + 10 - i]);
      acc = acc * s.dtz;
      acc = acc / rho[r];
      const float v = (c1z == 1.0f && c2z == 1.0f) ? vzo[r] - acc : __builtin_fmaf(c1z, vzo[r], -acc) / c2z;
      GF(vz)[id] = v;
      if (save_it >= 0 && (flags & TF_STRIP)) {
        const long long o = strip_vz(s, iz, ix);
        if (o >= 0) store_put(s, save_it, o, v);
      }
    }
  }
}

// Reverse velocity update of the source wavefield (kernel.cu:318-388 with the
// strip restore of kernel.cu:1294-1347, rtm_model.cu:424-429): strips of step
// it+1 are written back, interior cells get v += D.
template <int TZ>
__device__ __forceinline__ void vel_rev(const float* __restrict__ L, float* __restrict__ vx, float* __restrict__ vz,
                                        const ShotDev& s, int x0, int z0, int it, int flags,
                                        const float (&vxo)[TZ / 4], const float (&vzo)[TZ / 4],
                                        const float (&rho)[TZ / 4])
{
  constexpr int RPT = TZ / 4;
  const int lx = threadIdx.x & 63, j0 = (threadIdx.x >> 6) * RPT;
  const int ix = x0 + lx;
  const int nxp = s.nxp, nzp = s.nzp, P = s.pitch, npml = s.npml;
  if (ix >= nxp) return;
  float c10[RTM_R];
#pragma unroll
  for (int i = 0; i < RTM_R; i++) c10[i] = GCF(s.coef)[90 + i];
  const bool xin_vx = ix >= npml && ix <= nxp - npml - 2;
  const bool xin_vz = ix >= npml && ix < nxp - npml;
  float col[RPT + 19];
#pragma unroll
  for (int k = 0; k < RPT + 19; k++) col[k] = L[(j0 + k) * LW + lx + 9];
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const int iz = z0 + j0 + r;
    if (iz >= nzp) break;
    const size_t id = (size_t)iz * P + ix;
    const long long ovx = (flags & TF_STRIP) ? strip_vx(s, iz, ix) : -1;
    if (ovx >= 0) {
      GF(vx)[id] = store_get(s, it + 1, ovx);
    } else if (xin_vx && iz >= npml && iz < nzp - npml) {
      const float* __restrict__ Lr = L + (j0 + r + 9) * LW + lx + 9;
      float acc = 0.0f;
#pragma unroll
      for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], Lr[i] - Lr[1 - i], acc);
      acc = acc * s.dtx;
      acc = acc / rho[r];
      GF(vx)[id] = vxo[r] + acc;
    }
    const long long ovz = (flags & TF_STRIP) ? strip_vz(s, iz, ix) : -1;
    if (ovz >= 0) {
      GF(vz)[id] = store_get(s, it + 1, ovz);
    } else if (xin_vz && iz >= npml && iz <= nzp - npml - 2) {
      float acc = 0.0f;
#pragma unroll
      for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], col[r + 9 + i] - col[r + 10 - i], acc);
      acc = acc * s.dtz;
      acc = acc / rho[r];
      GF(vz)[id] = vzo[r] + acc;
    }
  }
}

// ------------------------------------------------------- pressure phase --
// VX rows z0..z0+TZ-1, columns x0-10..x0+72; VZ rows z0-10..z0+TZ+8, columns x0..x0+63.
struct DivV { float dxv, dzv; };

template <int RPT>
__device__ __forceinline__ DivV div_v(const float* __restrict__ VXr, const float (&col)[RPT + 19], int r, int rad,
                                      const float* c10, const float* __restrict__ cf,
                                      float dtx, float dtz, float rho, float vv)
{
  float sx = 0.0f, sz = 0.0f;
  if (rad == RTM_R) {
#pragma unroll
    for (int i = 1; i <= RTM_R; i++) {
      sx = __builtin_fmaf(c10[i - 1], VXr[i - 1] - VXr[-i], sx);
      sz = __builtin_fmaf(c10[i - 1], col[r + 9 + i] - col[r + 10 - i], sz);
    }
  } else {
    const gfloat* c = GCF(cf) + (rad - 1) * RTM_R;
    for (int i = 1; i <= rad; i++) {
      sx = __builtin_fmaf(c[i - 1], VXr[i - 1] - VXr[-i], sx);
      sz = __builtin_fmaf(c[i - 1], col[r + 9 + i] - col[r + 10 - i], sz);
    }
  }
  DivV d;
  d.dxv = sx * dtx; d.dxv = d.dxv * rho; d.dxv = d.dxv * vv;     // kernel.cu:536
  d.dzv = sz * dtz; d.dzv = d.dzv * rho; d.dzv = d.dzv * vv;     // kernel.cu:537
  return d;
}

// forward pressure update of one cell.  Interior: p -= dxv + dzv
// (kernel.cu:539).  PML: split pair, p = p1 + p2 (kernel.cu:780-783).
__device__ __forceinline__ float pres_cell(float p, float* __restrict__ p1, float* __restrict__ p2, size_t id,
                                           bool inner, const DivV& d, float c1x, float c2x, float c1z, float c2z)
{
  if (inner) return p - (d.dxv + d.dzv);
  float q1 = GCF(p1)[id], q2 = GCF(p2)[id];
  q1 = __builtin_fmaf(c1x, q1, -d.dxv) / c2x;
  q2 = __builtin_fmaf(c1z, q2, -d.dzv) / c2z;
  GF(p1)[id] = q1; GF(p2)[id] = q2;
  return q1 + q2;
}

template <int TZ>
__device__ __forceinline__ void stage_v(float* __restrict__ VX, float* __restrict__ VZ, const float* __restrict__ vx,
                                        const float* __restrict__ vz, int pitch, int x0, int z0)
{
  stage<LW, TZ>(VX, vx, pitch, z0, x0 - 10);
  stage<64, TZ + 19>(VZ, vz, pitch, z0 - 10, x0);
}

template <int RPT>
__device__ __forceinline__ void load_col64(float (&col)[RPT + 19], const float* __restrict__ VZ, int j0, int lx)
{
#pragma unroll
  for (int k = 0; k < RPT + 19; k++) col[k] = VZ[(j0 + k) * 64 + lx];
}

// forward pressure phase of the source wavefield: FD update, source
// injection (kernel.cu:36-47), recording (kernel.cu:48-71), p strips.
template <int TZ>
__device__ __forceinline__ void pres_fwd_src(const float* __restrict__ VX, const float* __restrict__ VZ,
                                             const ShotDev& s, int x0, int z0, int it, int save, int flags,
                                             const float (&po)[TZ / 4], const float (&rho)[TZ / 4],
                                             const float (&vv)[TZ / 4])
{
  constexpr int RPT = TZ / 4;
  const int lx = threadIdx.x & 63, j0 = (threadIdx.x >> 6) * RPT;
  const int ix = x0 + lx;
  const int nxp = s.nxp, nzp = s.nzp, P = s.pitch, npml = s.npml;
  if (ix >= nxp) return;
  float c10[RTM_R];
#pragma unroll
  for (int i = 0; i < RTM_R; i++) c10[i] = GCF(s.coef)[90 + i];
  const float c1x = GCF(s.c1x)[ix], c2x = GCF(s.c2x)[ix];
  const bool xin = ix >= npml && ix < nxp - npml;
  const bool xring = ix == 0 || ix == nxp - 1;
  float col[RPT + 19];
  load_col64<RPT>(col, VZ, j0, lx);
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const int iz = z0 + j0 + r;
    if (iz >= nzp) break;
    const size_t id = (size_t)iz * P + ix;
    float p = po[r];
    if (!xring && iz > 0 && iz < nzp - 1) {
      const int rad = imin(imin(imin(RTM_R, ix), imin(iz, nxp - 1 - ix)), nzp - 1 - iz);
      const DivV d = div_v<RPT>(VX + (j0 + r) * LW + lx + 10, col, r, rad, c10, s.coef, s.dtx, s.dtz, rho[r], vv[r]);
      const bool inner = xin && iz >= npml && iz < nzp - npml;
      p = pres_cell(p, s.sp1, s.sp2, id, inner, d, c1x, c2x, GCF(s.c1z)[iz], GCF(s.c2z)[iz]);
    }
    if ((flags & TF_SRC) && iz >= s.sz0 && iz <= s.sz1 && ix >= s.sx0 && ix <= s.sx1)
      p = __builtin_fmaf(GCF(s.sa)[(iz - s.sz0) * (s.sx1 - s.sx0 + 1) + ix - s.sx0], GCF(s.wav)[it], p);
    GF(s.sp)[id] = p;
    if ((flags & TF_REC) && s.record && iz == npml) {
      const int ir = ix - s.rcol0;
      if (ir >= 0 && ir < s.nrec) GF(s.rec)[(size_t)it * s.nrec + ir] = p;
    }
    if (save && (flags & TF_STRIP)) {
      const long long o = strip_p(s, iz, ix);
      if (o >= 0) store_put(s, it, o, p);
    }
  }
}

// reverse pressure phase: source reconstruction (kernel.cu:389-429, strip
// restore kernel.cu:1348-1400, source removal rtm_model.cu:437), receiver
// forward step + Dirichlet insertion (kernel.cu:72-95) and imaging
// (kernel.cu:1401-1416).
template <int TZ>
__device__ __forceinline__ void pres_rev(const float* __restrict__ SX, const float* __restrict__ SZ,
                                         const float* __restrict__ RX, const float* __restrict__ RZ,
                                         const ShotDev& s, int x0, int z0, int it, int flags,
                                         const float (&pso)[TZ / 4], const float (&pro)[TZ / 4],
                                         const float (&rho)[TZ / 4], const float (&vv)[TZ / 4])
{
  constexpr int RPT = TZ / 4;
  const int lx = threadIdx.x & 63, j0 = (threadIdx.x >> 6) * RPT;
  const int ix = x0 + lx;
  const int nxp = s.nxp, nzp = s.nzp, P = s.pitch, npml = s.npml;
  if (ix >= nxp) return;
  float c10[RTM_R];
#pragma unroll
  for (int i = 0; i < RTM_R; i++) c10[i] = GCF(s.coef)[90 + i];
  const float c1x = GCF(s.c1x)[ix], c2x = GCF(s.c2x)[ix];
  const bool xin = ix >= npml && ix < nxp - npml;
  const bool xring = ix == 0 || ix == nxp - 1;
  const bool xsrc = (flags & TF_SRC) && ix >= s.sx0 && ix <= s.sx1;
  float cs[RPT + 19], cr[RPT + 19];
  load_col64<RPT>(cs, SZ, j0, lx);
  load_col64<RPT>(cr, RZ, j0, lx);
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const int iz = z0 + j0 + r;
    if (iz >= nzp) break;
    const size_t id = (size_t)iz * P + ix;
    const bool inner = xin && iz >= npml && iz < nzp - npml;
    // source wavefield (interior + strips only; the rest is never read)
    float ps = 0.0f;
    {
      const long long o = (flags & TF_STRIP) ? strip_p(s, iz, ix) : -1;
      bool wr = false;
      if (o >= 0) {
        ps = store_get(s, it, o);
        wr = true;
      } else if (inner) {
        const DivV d = div_v<RPT>(SX + (j0 + r) * LW + lx + 10, cs, r, RTM_R, c10, s.coef, s.dtx, s.dtz, rho[r], vv[r]);
        ps = pso[r] + (d.dxv + d.dzv);
        wr = true;
      }
      if (wr) {
        if (xsrc && iz >= s.sz0 && iz <= s.sz1)
          ps = __builtin_fmaf(-GCF(s.sa)[(iz - s.sz0) * (s.sx1 - s.sx0 + 1) + ix - s.sx0], GCF(s.wav)[it], ps);
        GF(s.sp)[id] = ps;
      }
    }
    // receiver wavefield
    float p = pro[r];
    if (!xring && iz > 0 && iz < nzp - 1) {
      const int rad = imin(imin(imin(RTM_R, ix), imin(iz, nxp - 1 - ix)), nzp - 1 - iz);
      const DivV d = div_v<RPT>(RX + (j0 + r) * LW + lx + 10, cr, r, rad, c10, s.coef, s.dtx, s.dtz, rho[r], vv[r]);
      p = pres_cell(p, s.rp1, s.rp2, id, inner, d, c1x, c2x, GCF(s.c1z)[iz], GCF(s.c2z)[iz]);
    }
    if ((flags & TF_REC) && iz == npml) {
      const int ir = ix - s.rcol0;
      if (ir >= 0 && ir < s.nrec) {
        const float rv = (it <= ((const gint*)s.cut)[ir]) ? 0.0f : GCF(s.rec)[(size_t)it * s.nrec + ir];
        p = s.dt * rv;
      }
    }
    GF(s.rp)[id] = p;
    if (inner) {
      const size_t ii = (size_t)(iz - npml) * s.nxl + ix - npml;
      GF(s.g1)[ii] = __builtin_fmaf(ps, p, GCF(s.g1)[ii]);
      GF(s.illum)[ii] = __builtin_fmaf(ps, ps, GCF(s.illum)[ii]);
    }
  }
}

// own-cell loads of a thread (rows j0..j0+RPT-1 of its column), issued before
// the tile barrier so they overlap the staging; rows / columns past the grid
// read the (zero) margin
template <int RPT>
__device__ __forceinline__ void own_load(float (&o)[RPT], const float* __restrict__ f, const ShotDev& s, int x0, int z0)
{
  const int lx = threadIdx.x & 63, j0 = (threadIdx.x >> 6) * RPT;
#pragma unroll
  for (int r = 0; r < RPT; r++) o[r] = GCF(f)[(ptrdiff_t)(z0 + j0 + r) * s.pitch + x0 + lx];
}
template <int RPT>
__device__ __forceinline__ void own_model(float (&o)[RPT], const float* __restrict__ m, const ShotDev& s, int x0, int z0)
{
  const int lx = threadIdx.x & 63, j0 = (threadIdx.x >> 6) * RPT;
#pragma unroll
  for (int r = 0; r < RPT; r++) o[r] = model_at(m, s, z0 + j0 + r, x0 + lx);
}

}  // namespace

// ================================================================ kernels ==
template <int TZ>
__global__ __launch_bounds__(NTHR) void k_vel_fwd(LaunchArgs a, int save)
{
  constexpr int RPT = TZ / 4;
  __shared__ float L[(TZ + 19) * LW];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const Tile tl = a.tiles[t];
  const ShotDev s = a.shots[tl.slot];
  const int it = s.it_off + s.it_dir * a.step;
  float vxo[RPT], vzo[RPT], rho[RPT];
  stage<LW, TZ + 19>(L, s.sp, s.pitch, tl.z0 - 9, tl.x0 - 9);
  own_load<RPT>(vxo, s.svx, s, tl.x0, tl.z0);
  own_load<RPT>(vzo, s.svz, s, tl.x0, tl.z0);
  own_model<RPT>(rho, s.rho, s, tl.x0, tl.z0);
  __syncthreads();
  vel_fwd<TZ>(L, s.svx, s.svz, s, tl.x0, tl.z0, save ? it : -1, tl.kind, vxo, vzo, rho);
}

template <int TZ>
__global__ __launch_bounds__(NTHR) void k_pres_fwd(LaunchArgs a, int save)
{
  constexpr int RPT = TZ / 4;
  __shared__ float VX[TZ * LW];
  __shared__ float VZ[(TZ + 19) * 64];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const Tile tl = a.tiles[t];
  const ShotDev s = a.shots[tl.slot];
  const int it = s.it_off + s.it_dir * a.step;
  float po[RPT], rho[RPT], vv[RPT];
  stage_v<TZ>(VX, VZ, s.svx, s.svz, s.pitch, tl.x0, tl.z0);
  own_load<RPT>(po, s.sp, s, tl.x0, tl.z0);
  own_model<RPT>(rho, s.rho, s, tl.x0, tl.z0);
  own_model<RPT>(vv, s.vv, s, tl.x0, tl.z0);
  __syncthreads();
  pres_fwd_src<TZ>(VX, VZ, s, tl.x0, tl.z0, it, save, tl.kind, po, rho, vv);
}

template <int TZ>
__global__ __launch_bounds__(NTHR) void k_vel_rev(LaunchArgs a)
{
  constexpr int RPT = TZ / 4;
  __shared__ float LS[(TZ + 19) * LW];
  __shared__ float LR[(TZ + 19) * LW];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const Tile tl = a.tiles[t];
  const ShotDev s = a.shots[tl.slot];
  const int it = s.it_off + s.it_dir * a.step;
  float svx[RPT], svz[RPT], rvx[RPT], rvz[RPT], rho[RPT];
  stage<LW, TZ + 19>(LS, s.sp, s.pitch, tl.z0 - 9, tl.x0 - 9);
  stage<LW, TZ + 19>(LR, s.rp, s.pitch, tl.z0 - 9, tl.x0 - 9);
  own_load<RPT>(svx, s.svx, s, tl.x0, tl.z0);
  own_load<RPT>(svz, s.svz, s, tl.x0, tl.z0);
  own_load<RPT>(rvx, s.rvx, s, tl.x0, tl.z0);
  own_load<RPT>(rvz, s.rvz, s, tl.x0, tl.z0);
  own_model<RPT>(rho, s.rho, s, tl.x0, tl.z0);
  __syncthreads();
  vel_rev<TZ>(LS, s.svx, s.svz, s, tl.x0, tl.z0, it, tl.kind, svx, svz, rho);
  vel_fwd<TZ>(LR, s.rvx, s.rvz, s, tl.x0, tl.z0, -1, tl.kind, rvx, rvz, rho);
}

template <int TZ>
__global__ __launch_bounds__(NTHR) void k_pres_rev(LaunchArgs a)
{
  constexpr int RPT = TZ / 4;
  __shared__ float SX[TZ * LW];
  __shared__ float SZ[(TZ + 19) * 64];
  __shared__ float RX[TZ * LW];
  __shared__ float RZ[(TZ + 19) * 64];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const Tile tl = a.tiles[t];
  const ShotDev s = a.shots[tl.slot];
  const int it = s.it_off + s.it_dir * a.step;
  float pso[RPT], pro[RPT], rho[RPT], vv[RPT];
  stage_v<TZ>(SX, SZ, s.svx, s.svz, s.pitch, tl.x0, tl.z0);
  stage_v<TZ>(RX, RZ, s.rvx, s.rvz, s.pitch, tl.x0, tl.z0);
  own_load<RPT>(pso, s.sp, s, tl.x0, tl.z0);
  own_load<RPT>(pro, s.rp, s, tl.x0, tl.z0);
  own_model<RPT>(rho, s.rho, s, tl.x0, tl.z0);
  own_model<RPT>(vv, s.vv, s, tl.x0, tl.z0);
  __syncthreads();
  pres_rev<TZ>(SX, SZ, RX, RZ, s, tl.x0, tl.z0, it, tl.kind, pso, pro, rho, vv);
}

// 16-bit boundary store: quantise the fp32 staging of step `it` of each slot
// with the max-abs of its block: q = x / scale, scale = max(maxabs / 32768, FLT_MIN).
__global__ __launch_bounds__(NTHR) void k_quant(const ShotDev* __restrict__ shots, const int* __restrict__ slots,
                                                 int nslots, int step)
{
  for (int k = 0; k < nslots; k++) {
    const ShotDev s = shots[slots[k]];
    const int it = s.it_off + s.it_dir * step;
    const float* stg = (const float*)s.store + (long long)s.nt * s.strip_len / 2 + (long long)(it & 1) * s.strip_len;
    float sc[4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
      // a denormal max-abs (source tails before the wave arrives) must not
      // underflow the scale to zero: |x| <= 32768 * FLT_MIN then anyway
      const float m = __uint_as_float(s.smax[(long long)it * 4 + b]);
      sc[b] = m > 0.0f ? fmaxf(m * (1.0f / 32768.0f), 1.17549435e-38f) : 1.0f;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int b = 0; b < 4; b++) s.sscale[(long long)it * 4 + b] = sc[b];
    const long long n = s.strip_len;
    const long long a = 2LL * RTM_R * s.nzi, b = a + 2LL * RTM_R * s.nxl, c = b + 2LL * RTM_R * s.nzi;
    for (long long e = (long long)blockIdx.x * NTHR + threadIdx.x; e < n; e += (long long)gridDim.x * NTHR) {
      const int blk = e < a ? 0 : (e < b ? 1 : (e < c ? 2 : 3));
      const float q = stg[e] / sc[blk];
      const long long k2 = (long long)it * n + e;
      if (s.store_fmt == ST_F16) {
        ((__half*)s.store)[k2] = __float2half_rn(q);
      } else {
        unsigned int u = __float_as_uint(q);          // bf16, round to nearest even
        u += 0x7FFFu + ((u >> 16) & 1u);
        ((unsigned short*)s.store)[k2] = (unsigned short)(u >> 16);
      }
    }
  }
}

// zero the whole field allocations (margins included) that the starting
// phase uses: FWD both sets; REV the source velocities (rtm_model.cu:366-372),
// the receiver set and the imaging accumulators.
__global__ __launch_bounds__(NTHR) void k_phase_init(const ShotDev* __restrict__ shots, int slot, int phase)
{
  const ShotDev s = shots[slot];
  const long long n = (long long)s.frows * s.pitch;
  const long long ni = (long long)s.nzi * s.nxl;
  const long long g0 = (long long)blockIdx.x * NTHR + threadIdx.x, gs = (long long)gridDim.x * NTHR;
  float *sp = s.sp - s.forg, *svx = s.svx - s.forg, *svz = s.svz - s.forg, *sp1 = s.sp1 - s.forg, *sp2 = s.sp2 - s.forg;
  float *rp = s.rp - s.forg, *rvx = s.rvx - s.forg, *rvz = s.rvz - s.forg, *rp1 = s.rp1 - s.forg, *rp2 = s.rp2 - s.forg;
  if (phase == PH_FWD) {
    for (long long e = g0; e < n; e += gs) {
      sp[e] = 0.f; svx[e] = 0.f; svz[e] = 0.f; sp1[e] = 0.f; sp2[e] = 0.f;
      rp[e] = 0.f; rvx[e] = 0.f; rvz[e] = 0.f; rp1[e] = 0.f; rp2[e] = 0.f;
    }
    if (s.store_fmt != ST_F32)
      for (long long e = g0; e < 4LL * s.nt; e += gs) s.smax[e] = 0u;
  } else {
    for (long long e = g0; e < n; e += gs) {
      svx[e] = 0.f; svz[e] = 0.f;
      rp[e] = 0.f; rvx[e] = 0.f; rvz[e] = 0.f; rp1[e] = 0.f; rp2[e] = 0.f;
    }
    for (long long e = g0; e < ni; e += gs) { s.g1[e] = 0.f; s.illum[e] = 0.f; }
  }
}

__global__ __launch_bounds__(NTHR) void k_seed_receivers(const ShotDev* __restrict__ shots, int slot)
{
  const ShotDev s = shots[slot];
  const int ir = blockIdx.x * NTHR + threadIdx.x;
  if (ir >= s.nrec) return;
  const int it = s.nt - 1;
  const float rv = (it <= s.cut[ir]) ? 0.0f : s.rec[(size_t)it * s.nrec + ir];
  s.rp[(size_t)s.npml * s.pitch + s.rcol0 + ir] = s.dt * rv;
}

// taper and stack shots into the FP64 images, one image cell per thread,
// shots in the given order (kernel.cu:1556-1615, rtm_model.cu:462-488)
struct StackShot { const float* g1; const float* illum; const float* taper; int nx1, nxl; };

__global__ __launch_bounds__(NTHR) void k_stack(const StackShot* __restrict__ sh, int nsh, int nx, int nz,
                                                 double* __restrict__ img1, double* __restrict__ img2s,
                                                 double* __restrict__ ill)
{
  const long long e = (long long)blockIdx.x * NTHR + threadIdx.x;
  if (e >= (long long)nx * nz) return;
  const int iz = (int)(e / nx), ix = (int)(e - (long long)iz * nx);
  double a1 = img1[e], a2 = img2s[e], a3 = ill[e];
  for (int k = 0; k < nsh; k++) {
    const StackShot q = sh[k];
    const int c = ix - q.nx1;
    if (c < 0 || c >= q.nxl) continue;
    const float t = q.taper[c];
    const float g = q.g1[(size_t)iz * q.nxl + c] * t;
    const float il = q.illum[(size_t)iz * q.nxl + c] * t;
    a2 += (double)g / ((double)il + 1.0e-6);
    a1 += (double)g;
    a3 += (double)il;
  }
  img1[e] = a1; img2s[e] = a2; ill[e] = a3;
}

// final products: fp32 images from the FP64 stacks, the Laplacian-filtered
// image (laplacefd1d, image_function.cpp:44-79, integer spacings) and the
// multi-shot normalised image (kernel.cu:1580-1590)
__global__ __launch_bounds__(NTHR) void k_round(const double* __restrict__ a, float* __restrict__ o, long long n)
{
  const long long e = (long long)blockIdx.x * NTHR + threadIdx.x;
  if (e < n) o[e] = (float)a[e];
}

__global__ __launch_bounds__(NTHR) void k_laplace(const float* __restrict__ im, float* __restrict__ out, int nx, int nz,
                                                   int dxi, int dzi)
{
  const long long e = (long long)blockIdx.x * NTHR + threadIdx.x;
  if (e >= (long long)nx * nz) return;
  const int iz = (int)(e / nx), ix = (int)(e - (long long)iz * nx);
  if (ix == 0 || iz == 0 || ix == nx - 1 || iz == nz - 1) { out[e] = 0.0f; return; }
  const float c = im[e];
  const float t1 = (float)(((double)(im[e + 1] + im[e - 1]) - 2.0 * (double)c) / (double)(dxi * dxi));
  const float t2 = (float)(((double)(im[e + nx] + im[e - nx]) - 2.0 * (double)c) / (double)(dzi * dzi));
  out[e] = t1 + t2;
}

__global__ __launch_bounds__(NTHR) void k_norm_multi(const float* __restrict__ sg, const float* __restrict__ ss,
                                                      float* __restrict__ out, long long n)
{
  const long long e = (long long)blockIdx.x * NTHR + threadIdx.x;
  if (e < n) out[e] = (float)((double)sg[e] / ((double)ss[e] + 1.0e-6));
}

__global__ __launch_bounds__(NTHR) void k_zero_f64(double* __restrict__ a, long long n)
{
  const long long e = (long long)blockIdx.x * NTHR + threadIdx.x;
  if (e < n) a[e] = 0.0;
}

// advance the time origin of the slots of a captured chunk of `k` steps so
// that replaying the same graph continues where the last replay stopped
__global__ void k_bump(ShotDev* __restrict__ shots, const int* __restrict__ slots, int nslots, int k)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots) {
    ShotDev* s = &shots[slots[i]];
    s->it_off += s->it_dir * k;
  }
}

// =========================================================== launchers ==
#include "rtm_launch.h"

static inline int grid8(int ntiles) { return ((ntiles + 7) / 8) * 8; }

template <int TZ>
static hipError_t launch_phase_t(const LaunchArgs& a, int phase, int save, int which, hipStream_t st)
{
  const dim3 g(grid8(a.ntiles)), b(NTHR);
  if (phase == PH_FWD) {
    if (which == 0) hipLaunchKernelGGL(k_vel_fwd<TZ>, g, b, 0, st, a, save);
    else hipLaunchKernelGGL(k_pres_fwd<TZ>, g, b, 0, st, a, save);
  } else {
    if (which == 0) hipLaunchKernelGGL(k_vel_rev<TZ>, g, b, 0, st, a);
    else hipLaunchKernelGGL(k_pres_rev<TZ>, g, b, 0, st, a);
  }
  return hipGetLastError();
}

hipError_t rtm_launch_phase(const LaunchArgs& a, int phase, int tz, int save, int which, hipStream_t st)
{
  switch (tz) {
    case 8: return launch_phase_t<8>(a, phase, save, which, st);
    case 16: return launch_phase_t<16>(a, phase, save, which, st);
    case 32: return launch_phase_t<32>(a, phase, save, which, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t rtm_launch_step(const LaunchArgs& a, int phase, int tz, int save, hipStream_t st)
{
  hipError_t e = rtm_launch_phase(a, phase, tz, save, 0, st);
  if (e != hipSuccess) return e;
  return rtm_launch_phase(a, phase, tz, save, 1, st);
}

hipError_t rtm_launch_quant(const ShotDev* shots, const int* slots, int nslots, int step, int nblocks, hipStream_t st)
{
  hipLaunchKernelGGL(k_quant, dim3(nblocks), dim3(NTHR), 0, st, shots, slots, nslots, step);
  return hipGetLastError();
}

hipError_t rtm_launch_phase_init(const ShotDev* shots, int slot, int phase, int nblocks, hipStream_t st)
{
  hipLaunchKernelGGL(k_phase_init, dim3(nblocks), dim3(NTHR), 0, st, shots, slot, phase);
  return hipGetLastError();
}

hipError_t rtm_launch_seed(const ShotDev* shots, int slot, int nrec, hipStream_t st)
{
  if (nrec <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_seed_receivers, dim3((nrec + NTHR - 1) / NTHR), dim3(NTHR), 0, st, shots, slot);
  return hipGetLastError();
}

hipError_t rtm_launch_stack(const void* sh, int nsh, int nx, int nz, double* i1, double* i2, double* il, hipStream_t st)
{
  const long long n = (long long)nx * nz;
  hipLaunchKernelGGL(k_stack, dim3((unsigned)((n + NTHR - 1) / NTHR)), dim3(NTHR), 0, st,
                     (const StackShot*)sh, nsh, nx, nz, i1, i2, il);
  return hipGetLastError();
}

hipError_t rtm_launch_round(const double* a, float* o, long long n, hipStream_t st)
{
  hipLaunchKernelGGL(k_round, dim3((unsigned)((n + NTHR - 1) / NTHR)), dim3(NTHR), 0, st, a, o, n);
  return hipGetLastError();
}

hipError_t rtm_launch_laplace(const float* im, float* out, int nx, int nz, int dxi, int dzi, hipStream_t st)
{
  const long long n = (long long)nx * nz;
  hipLaunchKernelGGL(k_laplace, dim3((unsigned)((n + NTHR - 1) / NTHR)), dim3(NTHR), 0, st, im, out, nx, nz, dxi, dzi);
  return hipGetLastError();
}

hipError_t rtm_launch_norm_multi(const float* sg, const float* ss, float* out, long long n, hipStream_t st)
{
  hipLaunchKernelGGL(k_norm_multi, dim3((unsigned)((n + NTHR - 1) / NTHR)), dim3(NTHR), 0, st, sg, ss, out, n);
  return hipGetLastError();
}

hipError_t rtm_launch_zero_f64(double* a, long long n, hipStream_t st)
{
  hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n + NTHR - 1) / NTHR)), dim3(NTHR), 0, st, a, n);
  return hipGetLastError();
}

hipError_t rtm_launch_bump(ShotDev* shots, const int* slots, int nslots, int k, hipStream_t st)
{
  hipLaunchKernelGGL(k_bump, dim3((nslots + 63) / 64), dim3(64), 0, st, shots, slots, nslots, k);
  return hipGetLastError();
}
