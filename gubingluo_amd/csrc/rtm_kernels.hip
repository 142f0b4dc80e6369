// rtm_kernels.hip -- gfx950 kernels of the RTM hot path.
//
// One launch per time step.  A workgroup owns a tile of 64 columns (one
// wave64 row: every array is x-fastest, so a wave touches one contiguous
// 256-byte row segment) by TZ rows of one shot, and does the whole step for
// it: the velocity update (kernel.cu:430-500, 542-713 forward; 318-388
// reverse) of the tile and of the radius-10 halo that the pressure stencil
// reads, then the pressure update (kernel.cu:501-541, 714-1185 forward;
// 389-429 reverse) with source (kernel.cu:36-47), recording (48-71),
// receiver insertion (72-95), boundary strips (1186-1400) and imaging
// (1401-1416).  The halo velocities are computed again by the neighbouring
// tiles instead of being read back, so a step reads p, vx, vz and the model
// once and writes p, vx, vz once per cell.  Fields have two time levels
// (step it reads level it&1 and writes level (it+1)&1), so a tile's reads of
// its neighbours' cells never race with their writes.
//
// LDS: the pressure of the tile and its halo (columns x0-20..x0+83 of the
// tile rows, rows z0-19..z0+TZ+18 of the tile columns), then -- reusing the
// same bytes -- the new velocities (vx: columns x0-10..x0+73; vz: rows
// z0-10..z0+TZ+8).  x-derivatives read LDS rows; z-derivatives read a
// register window of the thread's own column.  The reverse step runs the
// source reconstruction and then the receiver wavefield through the same
// LDS; the imaging products use the reconstructed pressure from registers.
//
// Arithmetic follows the reference operation order with explicit fused
// multiply-adds (the oracle's canonical FP32 form), so every cell matches the
// CPU restatement bit for bit; the redundant halo velocities are the same
// operations on the same inputs.  Compiled with -ffp-contract=off; float
// division is IEEE as HIP compiles it.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "rtm_dev.h"

#define NTHR 256   // threads of the auxiliary kernels

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

// ---------------------------------------------------------------- strips --
// store offsets of the boundary strips (kernel.cu:1186-1293 packing); -1 if
// the cell is not part of the block.
__device__ __forceinline__ int strip_vx(const ShotDev& s, int iz, int ix)
{
  if (iz < s.npml || iz >= s.nzp - s.npml) return -1;
  const int r = iz - s.npml;
  if (ix >= s.npml - RTM_R && ix < s.npml) return r * RTM_R + ix - (s.npml - RTM_R);
  const int c0 = s.nxp - s.npml - 1;
  if (ix >= c0 && ix < c0 + RTM_R) return RTM_R * s.nzi + r * RTM_R + ix - c0;
  return -1;
}
__device__ __forceinline__ int strip_vz(const ShotDev& s, int iz, int ix)
{
  if (ix < s.npml || ix >= s.nxp - s.npml) return -1;
  const int base = 2 * RTM_R * s.nzi;
  const int c = ix - s.npml;
  if (iz >= s.npml - RTM_R && iz < s.npml) return base + (iz - (s.npml - RTM_R)) * s.nxl + c;
  const int r0 = s.nzp - s.npml - 1;
  if (iz >= r0 && iz < r0 + RTM_R) return base + RTM_R * s.nxl + (iz - r0) * s.nxl + c;
  return -1;
}
__device__ __forceinline__ int strip_p(const ShotDev& s, int iz, int ix)
{
  const int base_lr = 2 * RTM_R * s.nzi + 2 * RTM_R * s.nxl;
  const int base_tb = base_lr + 2 * RTM_R * s.nzi;
  if (iz >= s.npml && iz < s.nzp - s.npml) {
    const int r = iz - s.npml;
    if (ix >= s.npml - RTM_R && ix < s.npml) return base_lr + r * RTM_R + ix - (s.npml - RTM_R);
    const int c0 = s.nxp - s.npml;
    if (ix >= c0 && ix < c0 + RTM_R) return base_lr + RTM_R * s.nzi + r * RTM_R + ix - c0;
    return -1;
  }
  if (ix < s.npml || ix >= s.nxp - s.npml) return -1;
  const int c = ix - s.npml;
  if (iz >= s.npml - RTM_R && iz < s.npml) return base_tb + (iz - (s.npml - RTM_R)) * s.nxl + c;
  const int r0 = s.nzp - s.npml;
  if (iz >= r0 && iz < r0 + RTM_R) return base_tb + RTM_R * s.nxl + (iz - r0) * s.nxl + c;
  return -1;
}

// 16-bit store: block of a strip offset (0 vx, 1 vz, 2 p-lr, 3 p-tb)
__device__ __forceinline__ int strip_block(const ShotDev& s, int off)
{
  const int a = 2 * RTM_R * s.nzi, b = a + 2 * RTM_R * s.nxl, c = b + 2 * RTM_R * s.nzi;
  return off < a ? 0 : (off < b ? 1 : (off < c ? 2 : 3));
}

__device__ __forceinline__ void store_put(const ShotDev& s, int it, int off, float v)
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

__device__ __forceinline__ float store_get(const ShotDev& s, int it, int off)
{
  const long long k = (long long)it * s.strip_len + off;
  if (s.store_fmt == ST_F32) return GCF(s.store)[k];
  const float sc = GCF(s.sscale)[(long long)it * 4 + strip_block(s, off)];
  if (s.store_fmt == ST_F16) return __half2float(__ushort_as_half(((const gushort*)s.store)[k])) * sc;
  const unsigned short h = ((const gushort*)s.store)[k];
  return __uint_as_float(((unsigned int)h) << 16) * sc;
}

// ------------------------------------------------------------ tile shape --
// G waves of 64 lanes; each thread owns RPT consecutive rows of one column.
template <int G_, int RPT_>
struct Geo {
  static constexpr int G = G_, RPT = RPT_;
  static constexpr int TZ = G * RPT;                   // tile rows
  static constexpr int NT = 64 * G;                    // threads
  static constexpr int PCW = 104;                      // p row: columns x0-20 .. x0+83
  static constexpr int VXW = 84;                       // vx row: columns x0-10 .. x0+73
  static constexpr int ZR = TZ + 19;                   // vz rows z0-10 .. z0+TZ+8
  static constexpr int KZ = (ZR + G - 1) / G;          // vz rows per thread (consecutive)
  static constexpr int PZR = G * KZ + 19;              // p column rows held (>= TZ + 38)
  static constexpr int KP = (TZ + 38 + G - 1) / G;     // p column rows staged per thread
  static constexpr int KX = (TZ * VXW + NT - 1) / NT;  // vx cells per thread
  static constexpr int LDS_A = PZR * 64 + TZ * PCW;    // PZ | PC
  static constexpr int LDS_B = TZ * VXW + ZR * 64;     // VXL | VZL
  static constexpr int LDS = LDS_A > LDS_B ? LDS_A : LDS_B;
};

// radius-r staggered difference a[i] - b[i] with the coefficient row of
// order 2r; the fma chain runs i = 1..r as the reference loops
// (kernel.cu:459-460).  The loop is unrolled with a predicate so that the
// operand windows stay in registers (a runtime index would put them in
// scratch memory).
#define STENCIL_SUM(acc, r, c10, cf, A, B)                                                        \
  do {                                                                                            \
    if ((r) == RTM_R) {                                                                           \
      _Pragma("unroll") for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], (A) - (B), acc); \
    } else {                                                                                      \
      const gfloat* cr_ = GCF(cf) + ((r) - 1) * RTM_R;                                            \
      _Pragma("unroll") for (int i = 1; i <= RTM_R; i++)                                          \
        if (i <= (r)) acc = __builtin_fmaf(cr_[i - 1], (A) - (B), acc);                           \
    }                                                                                             \
  } while (0)

// field kinds of one pass
enum : int { FK_SRC = 0, FK_RCV = 1, FK_RCN = 2 };

// Stage p rows z0-19 .. z0+TZ+18 of the tile columns into PZ (pitch 64) and
// the tile rows with their x halo into PC (pitch 104, column c = x-(x0-20)).
template <class Q>
__device__ __forceinline__ void stage_p(float* __restrict__ PZ, float* __restrict__ PC, const float* __restrict__ p,
                                        int pitch, int x0, int z0, int lx, int g)
{
  float a[Q::KP], w[Q::RPT];
  const gfloat* base = GCF(p) + (ptrdiff_t)(z0 - 19) * pitch + x0;
  const int wc = lx < 20 ? lx - 20 : lx + 44;          // wing columns (lanes < 40)
#pragma unroll
  for (int m = 0; m < Q::KP; m++) {
    const int k = g + Q::G * m;
    a[m] = 0.0f;
    if (k < Q::TZ + 38) a[m] = base[k * pitch + lx];
  }
#pragma unroll
  for (int m = 0; m < Q::RPT; m++) {
    const int j = g + Q::G * m;                        // tile row
    w[m] = 0.0f;
    if (lx < 40) w[m] = base[(j + 19) * pitch + wc];
  }
#pragma unroll
  for (int m = 0; m < Q::KP; m++) {
    const int k = g + Q::G * m;
    if (k < Q::TZ + 38) {
      PZ[k * 64 + lx] = a[m];
      if (k >= 19 && k < 19 + Q::TZ) PC[(k - 19) * Q::PCW + 20 + lx] = a[m];
    }
  }
#pragma unroll
  for (int m = 0; m < Q::RPT; m++) {
    const int j = g + Q::G * m;
    if (lx < 40) PC[j * Q::PCW + (lx < 20 ? lx : lx + 64)] = w[m];
  }
}

// ---------------------------------------------------------- velocities --
// vz of rows z0-10+kz (kz = g*KZ + k) of the thread's column, vx of the
// cells c = k*NT + tid of the TZ x 84 block.  Forward-type fields (source in
// the forward phase, receiver in the reverse phase):
//   v = (c1*v - D)/c2, D = (sum c_i (p[+i] - p[-i+1]) * dt/h) / rho
//   (kernel.cu:588-594); c1 = c2 = 1 off the absorbing layer reduces it to
//   v - D (kernel.cu:462); radius min(10, distance to the grid edge)
//   (cuda_norderx/z, kernel.cu:299-317).
// Reconstruction: strip cells restored from the store of step it+1
// (kernel.cu:1294-1347), interior cells v + D (kernel.cu:318-388).
template <class Q>
struct VelOut {
  float vz[Q::KZ];
  float vx[Q::KX];
};

template <class Q, int FK>
__device__ __forceinline__ void velocities(VelOut<Q>& o, const float* __restrict__ PZ, const float* __restrict__ PC,
                                           const ShotDev& s, const Tile& tl, const float* __restrict__ vxc,
                                           const float* __restrict__ vzc, float* __restrict__ vxn,
                                           float* __restrict__ vzn, int it, int save, int lx, int g, int tid)
{
  const int x0 = tl.x0, z0 = tl.z0, P = s.pitch, nxp = s.nxp, nzp = s.nzp, npml = s.npml;
  const bool gen = (tl.kind & TF_GEN) != 0;
  const bool strips = (tl.kind & TF_STRIP) != 0;
  float c10[RTM_R];
#pragma unroll
  for (int i = 0; i < RTM_R; i++) c10[i] = GCF(s.coef)[90 + i];
  const float* __restrict__ cf = s.coef;
  const float dtx = s.dtx, dtz = s.dtz;

  // ---- loads of the old velocities and density (issued with the staging)
  float vzo[Q::KZ], rz[Q::KZ], vxo[Q::KX], rx[Q::KX];
  const int ix = x0 + lx;
#pragma unroll
  for (int k = 0; k < Q::KZ; k++) {
    const int kz = g * Q::KZ + k;
    vzo[k] = 0.0f;
    rz[k] = 1.0f;
    if (kz < Q::ZR) {
      const int id = (z0 - 10 + kz) * P + ix;
      vzo[k] = GCF(vzc)[id];
      rz[k] = GCF(s.mrho)[id];
    }
  }
#pragma unroll
  for (int k = 0; k < Q::KX; k++) {
    const int c = k * Q::NT + tid;
    vxo[k] = 0.0f;
    rx[k] = 1.0f;
    if (c < Q::TZ * Q::VXW) {
      const int row = c / Q::VXW, col = c - row * Q::VXW;
      const int id = (z0 + row) * P + x0 - 10 + col;
      vxo[k] = GCF(vxc)[id];
      rx[k] = GCF(s.mrho)[id];
    }
  }
  __syncthreads();                                     // p staged

  // ---- vz
  {
    float cw[Q::KZ + 19];                              // p rows z0-19+g*KZ+m of the column
#pragma unroll
    for (int m = 0; m < Q::KZ + 19; m++) cw[m] = PZ[(g * Q::KZ + m) * 64 + lx];
#pragma unroll
    for (int k = 0; k < Q::KZ; k++) {
      const int kz = g * Q::KZ + k;
      const int iz = z0 - 10 + kz;
      float v = 0.0f;
      if (kz < Q::ZR) {
        if (FK == FK_RCN) {
          const int so = strips ? strip_vz(s, iz, ix) : -1;
          if (so >= 0) {
            v = store_get(s, it + 1, so);
          } else if (!gen || (ix >= npml && ix < nxp - npml && iz >= npml && iz <= nzp - npml - 2)) {
            float acc = 0.0f;
#pragma unroll
            for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], cw[k + 9 + i] - cw[k + 10 - i], acc);
            acc = acc * dtz;
            acc = acc / rz[k];
            v = vzo[k] + acc;
          }
        } else if (!gen) {
          float acc = 0.0f;
#pragma unroll
          for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], cw[k + 9 + i] - cw[k + 10 - i], acc);
          acc = acc * dtz;
          acc = acc / rz[k];
          v = vzo[k] - acc;
        } else if (iz >= 0 && iz <= nzp - 2 && ix < nxp) {
          const int rr = imin(RTM_R, imin(iz + 1, nzp - 1 - iz));
          float acc = 0.0f;
          STENCIL_SUM(acc, rr, c10, cf, cw[k + 9 + i], cw[k + 10 - i]);
          acc = acc * dtz;
          acc = acc / rz[k];
          const float c1 = GCF(s.c1zv)[iz], c2 = GCF(s.c2zv)[iz];
          v = (c1 == 1.0f && c2 == 1.0f) ? vzo[k] - acc : __builtin_fmaf(c1, vzo[k], -acc) / c2;
        }
        // own rows: write the new level (cells that exist for vz)
        if (kz >= 10 && kz < 10 + Q::TZ && ix < nxp && iz <= nzp - 2) {
          GF(vzn)[iz * P + ix] = v;
          if (FK == FK_SRC && save && strips) {
            const int so = strip_vz(s, iz, ix);
            if (so >= 0) store_put(s, it, so, v);
          }
        }
      }
      o.vz[k] = v;
    }
  }
  // ---- vx
#pragma unroll
  for (int k = 0; k < Q::KX; k++) {
    const int c = k * Q::NT + tid;
    float v = 0.0f;
    if (c < Q::TZ * Q::VXW) {
      const int row = c / Q::VXW, col = c - row * Q::VXW;
      const int iz = z0 + row, jx = x0 - 10 + col;
      const float* __restrict__ Pr = PC + row * Q::PCW + col + 10;   // p at column jx
      if (FK == FK_RCN) {
        const int so = strips ? strip_vx(s, iz, jx) : -1;
        if (so >= 0) {
          v = store_get(s, it + 1, so);
        } else if (!gen || (jx >= npml && jx <= nxp - npml - 2 && iz >= npml && iz < nzp - npml)) {
          float acc = 0.0f;
#pragma unroll
          for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], Pr[i] - Pr[1 - i], acc);
          acc = acc * dtx;
          acc = acc / rx[k];
          v = vxo[k] + acc;
        }
      } else if (!gen) {
        float acc = 0.0f;
#pragma unroll
        for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], Pr[i] - Pr[1 - i], acc);
        acc = acc * dtx;
        acc = acc / rx[k];
        v = vxo[k] - acc;
      } else if (jx >= 0 && jx <= nxp - 2 && iz < nzp) {
        const int rr = imin(RTM_R, imin(jx + 1, nxp - 1 - jx));
        float acc = 0.0f;
        STENCIL_SUM(acc, rr, c10, cf, Pr[i], Pr[1 - i]);
        acc = acc * dtx;
        acc = acc / rx[k];
        const float c1 = GCF(s.c1xv)[jx], c2 = GCF(s.c2xv)[jx];
        v = (c1 == 1.0f && c2 == 1.0f) ? vxo[k] - acc : __builtin_fmaf(c1, vxo[k], -acc) / c2;
      }
      if (col >= 10 && col < 74 && iz < nzp && jx <= nxp - 2) {
        GF(vxn)[iz * P + jx] = v;
        if (FK == FK_SRC && save && strips) {
          const int so = strip_vx(s, iz, jx);
          if (so >= 0) store_put(s, it, so, v);
        }
      }
    }
    o.vx[k] = v;
  }
}

// ------------------------------------------------------------ pressure --
struct DivV { float dxv, dzv; };

// VXr points at vx of the cell's column in its LDS row; cz is the column
// window of vz rows; r the thread-local row
template <int NW>
__device__ __forceinline__ DivV div_v(const float* __restrict__ VXr, const float (&cz)[NW], int r, int rad,
                                      const float* c10, const float* __restrict__ cf, float dtx, float dtz,
                                      float rho, float vv)
{
  float sx = 0.0f, sz = 0.0f;
  if (rad == RTM_R) {
#pragma unroll
    for (int i = 1; i <= RTM_R; i++) {
      sx = __builtin_fmaf(c10[i - 1], VXr[i - 1] - VXr[-i], sx);
      sz = __builtin_fmaf(c10[i - 1], cz[r + 9 + i] - cz[r + 10 - i], sz);
    }
  } else {
    const gfloat* c = GCF(cf) + (rad - 1) * RTM_R;
#pragma unroll
    for (int i = 1; i <= RTM_R; i++) {
      if (i <= rad) {
        sx = __builtin_fmaf(c[i - 1], VXr[i - 1] - VXr[-i], sx);
        sz = __builtin_fmaf(c[i - 1], cz[r + 9 + i] - cz[r + 10 - i], sz);
      }
    }
  }
  DivV d;
  d.dxv = sx * dtx; d.dxv = d.dxv * rho; d.dxv = d.dxv * vv;     // kernel.cu:536
  d.dzv = sz * dtz; d.dzv = d.dzv * rho; d.dzv = d.dzv * vv;     // kernel.cu:537
  return d;
}

// One field through one time step on a tile: p of the current level pc (with
// velocities vxc, vzc) to the next level (pn, vxn, vzn).  On return pnew
// holds the new pressure of the thread's own cells (rows g*RPT+r); ps is the
// reconstructed source pressure of the same cells (receiver pass imaging).
template <class Q, int FK>
__device__ __forceinline__ void field_step(float* __restrict__ lds, const ShotDev& s, const Tile& tl, int it, int save,
                                           const float* __restrict__ pc, const float* __restrict__ vxc,
                                           const float* __restrict__ vzc, float* __restrict__ pn,
                                           float* __restrict__ vxn, float* __restrict__ vzn,
                                           const float (&ps)[Q::RPT], float (&pnew)[Q::RPT])
{
  float* PZ = lds;
  float* PC = lds + Q::PZR * 64;
  float* VXL = lds;
  float* VZL = lds + Q::TZ * Q::VXW;
  const int tid = threadIdx.x, lx = tid & 63;
  const int g = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int x0 = tl.x0, z0 = tl.z0, P = s.pitch, nxp = s.nxp, nzp = s.nzp, npml = s.npml;
  const int ix = x0 + lx;
  const bool gen = (tl.kind & TF_GEN) != 0;

  stage_p<Q>(PZ, PC, pc, P, x0, z0, lx, g);
  VelOut<Q> vo;
  velocities<Q, FK>(vo, PZ, PC, s, tl, vxc, vzc, vxn, vzn, it, save, lx, g, tid);

  // own old pressure out of LDS, model of the own cells
  float po[Q::RPT], rho[Q::RPT], vv[Q::RPT];
#pragma unroll
  for (int r = 0; r < Q::RPT; r++) {
    const int j = g * Q::RPT + r;
    po[r] = PC[j * Q::PCW + 20 + lx];
    const int id = (z0 + j) * P + ix;
    rho[r] = GCF(s.mrho)[id];
    vv[r] = GCF(s.mvv)[id];
  }
  __syncthreads();                                     // PZ/PC consumed
#pragma unroll
  for (int k = 0; k < Q::KZ; k++) {
    const int kz = g * Q::KZ + k;
    if (kz < Q::ZR) VZL[kz * 64 + lx] = vo.vz[k];
  }
#pragma unroll
  for (int k = 0; k < Q::KX; k++) {
    const int c = k * Q::NT + tid;
    if (c < Q::TZ * Q::VXW) VXL[c] = vo.vx[k];
  }
  __syncthreads();                                     // new velocities staged

  float c10[RTM_R];
#pragma unroll
  for (int i = 0; i < RTM_R; i++) c10[i] = GCF(s.coef)[90 + i];
  float cz[Q::RPT + 19];
#pragma unroll
  for (int m = 0; m < Q::RPT + 19; m++) cz[m] = VZL[(g * Q::RPT + m) * 64 + lx];
  const bool xin = ix >= npml && ix < nxp - npml;
  const bool xring = ix == 0 || ix == nxp - 1;
  float c1x = 1.0f, c2x = 1.0f;
  if (FK != FK_RCN && gen && ix < nxp) { c1x = GCF(s.c1x)[ix]; c2x = GCF(s.c2x)[ix]; }
  const bool xsrc = (tl.kind & TF_SRC) && ix >= s.sx0 && ix <= s.sx1;
  const bool strips = (tl.kind & TF_STRIP) != 0;

#pragma unroll
  for (int r = 0; r < Q::RPT; r++) {
    const int j = g * Q::RPT + r;
    const int iz = z0 + j;
    float p = po[r];
    const float* __restrict__ VXr = VXL + j * Q::VXW + lx + 10;
    const bool incell = ix < nxp && iz < nzp;
    if (FK == FK_RCN) {
      const bool inner = xin && iz >= npml && iz < nzp - npml;
      const int so = strips ? strip_p(s, iz, ix) : -1;
      bool wr = false;
      if (so >= 0) {
        p = store_get(s, it, so);                      // kernel.cu:1348-1400
        wr = true;
      } else if (!gen || inner) {
        const DivV d = div_v(VXr, cz, r, RTM_R, c10, s.coef, s.dtx, s.dtz, rho[r], vv[r]);
        p = p + (d.dxv + d.dzv);                       // kernel.cu:389-429
        wr = true;
      }
      if (wr && xsrc && iz >= s.sz0 && iz <= s.sz1)    // source removal (rtm_model.cu:437)
        p = __builtin_fmaf(-GCF(s.sa)[(iz - s.sz0) * (s.sx1 - s.sx0 + 1) + ix - s.sx0], GCF(s.wav)[it], p);
      if (wr) GF(pn)[iz * P + ix] = p;
    } else {
      if (!gen) {
        const DivV d = div_v(VXr, cz, r, RTM_R, c10, s.coef, s.dtx, s.dtz, rho[r], vv[r]);
        p = p - (d.dxv + d.dzv);                       // kernel.cu:539
      } else if (incell && !xring && iz > 0 && iz < nzp - 1) {
        const int rad = imin(imin(imin(RTM_R, ix), imin(iz, nxp - 1 - ix)), nzp - 1 - iz);
        const DivV d = div_v(VXr, cz, r, rad, c10, s.coef, s.dtx, s.dtz, rho[r], vv[r]);
        if (xin && iz >= npml && iz < nzp - npml) {
          p = p - (d.dxv + d.dzv);
        } else {                                       // split pair, p = p1 + p2 (kernel.cu:780-783)
          const int id = iz * P + ix;
          float q1 = GCF(s.q1)[id], q2 = GCF(s.q2)[id];
          q1 = __builtin_fmaf(c1x, q1, -d.dxv) / c2x;
          q2 = __builtin_fmaf(GCF(s.c1z)[iz], q2, -d.dzv) / GCF(s.c2z)[iz];
          GF(s.q1)[id] = q1; GF(s.q2)[id] = q2;
          p = q1 + q2;
        }
      }
      if (FK == FK_SRC) {
        if (xsrc && iz >= s.sz0 && iz <= s.sz1)        // kernel.cu:36-47
          p = __builtin_fmaf(GCF(s.sa)[(iz - s.sz0) * (s.sx1 - s.sx0 + 1) + ix - s.sx0], GCF(s.wav)[it], p);
        if (incell) {
          GF(pn)[iz * P + ix] = p;
          if ((tl.kind & TF_REC) && s.record && iz == npml) {   // kernel.cu:48-71
            const int ir = ix - s.rcol0;
            if (ir >= 0 && ir < s.nrec) GF(s.rec)[(size_t)it * s.nrec + ir] = p;
          }
          if (save && strips) {
            const int so = strip_p(s, iz, ix);
            if (so >= 0) store_put(s, it, so, p);
          }
        }
      } else {                                         // receiver wavefield
        if ((tl.kind & TF_REC) && iz == npml) {        // Dirichlet insertion (kernel.cu:72-95)
          const int ir = ix - s.rcol0;
          if (ir >= 0 && ir < s.nrec) {
            const float rv = (it <= ((const gint*)s.cut)[ir]) ? 0.0f : GCF(s.rec)[(size_t)it * s.nrec + ir];
            p = s.dt * rv;
          }
        }
        if (incell) {
          GF(pn)[iz * P + ix] = p;
          if (xin && iz >= npml && iz < nzp - npml) {  // imaging (kernel.cu:1401-1416)
            const int ii = (iz - npml) * s.nxl + ix - npml;
            GF(s.g1)[ii] = __builtin_fmaf(ps[r], p, GCF(s.g1)[ii]);
            GF(s.illum)[ii] = __builtin_fmaf(ps[r], ps[r], GCF(s.illum)[ii]);
          }
        }
      }
    }
    pnew[r] = p;
  }
}

}  // namespace

// ================================================================ kernels ==
// forward time step of the source wavefield (modelling, recording, strips)
template <int G, int RPT>
__global__ __launch_bounds__(64 * G) void k_fwd(LaunchArgs a, int save)
{
  using Q = Geo<G, RPT>;
  __shared__ float lds[Q::LDS];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const Tile tl = a.tiles[t];
  const ShotDev s = a.shots[tl.slot];
  const int it = s.it_off + s.it_dir * a.step;
  const bool odd = it & 1;                             // (a runtime index into s would spill it to scratch)
  float none[Q::RPT], pnew[Q::RPT];
#pragma unroll
  for (int r = 0; r < Q::RPT; r++) none[r] = 0.0f;
  field_step<Q, FK_SRC>(lds, s, tl, it, save, odd ? s.sp[1] : s.sp[0], odd ? s.svx[1] : s.svx[0],
                        odd ? s.svz[1] : s.svz[0], odd ? s.sp[0] : s.sp[1], odd ? s.svx[0] : s.svx[1],
                        odd ? s.svz[0] : s.svz[1], none, pnew);
}

// reverse time step: source reconstruction, then the receiver wavefield and
// the imaging products of the tile
template <int G, int RPT>
__global__ __launch_bounds__(64 * G) void k_rev(LaunchArgs a)
{
  using Q = Geo<G, RPT>;
  __shared__ float lds[Q::LDS];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const Tile tl = a.tiles[t];
  const ShotDev s = a.shots[tl.slot];
  const int it = s.it_off + s.it_dir * a.step;
  const bool odd = it & 1;
  float ps[Q::RPT], pr[Q::RPT], none[Q::RPT];
#pragma unroll
  for (int r = 0; r < Q::RPT; r++) ps[r] = none[r] = 0.0f;
  if (tl.kind & TF_INT) {
    field_step<Q, FK_RCN>(lds, s, tl, it, 0, odd ? s.sp[1] : s.sp[0], odd ? s.svx[1] : s.svx[0],
                          odd ? s.svz[1] : s.svz[0], odd ? s.sp[0] : s.sp[1], odd ? s.svx[0] : s.svx[1],
                          odd ? s.svz[0] : s.svz[1], none, ps);
    __syncthreads();                                   // LDS reused by the receiver pass
  }
  field_step<Q, FK_RCV>(lds, s, tl, it, 0, odd ? s.rp[1] : s.rp[0], odd ? s.rvx[1] : s.rvx[0],
                        odd ? s.rvz[1] : s.rvz[0], odd ? s.rp[0] : s.rp[1], odd ? s.rvx[0] : s.rvx[1],
                        odd ? s.rvz[0] : s.rvz[1], ps, pr);
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

// Phase start of one slot.  FWD: zero both levels of both wavefields and the
// PML pair, and copy the shot's model window into the slot (edge replication
// into the PML, extendvel1 Velocity_function.cpp:2011-2030, and on into the
// margins).  REV: zero both levels of the source velocities (rtm_model.cu:
// 366-372), the receiver set, the PML pair and the imaging accumulators.
__global__ __launch_bounds__(NTHR) void k_phase_init(const ShotDev* __restrict__ shots, int slot, int phase)
{
  const ShotDev s = shots[slot];
  const long long n = (long long)s.frows * s.pitch;
  const long long ni = (long long)s.nzi * s.nxl;
  const long long g0 = (long long)blockIdx.x * NTHR + threadIdx.x, gs = (long long)gridDim.x * NTHR;
  float* f[14] = {s.sp[0], s.sp[1], s.svx[0], s.svx[1], s.svz[0], s.svz[1], s.rp[0], s.rp[1],
                  s.rvx[0], s.rvx[1], s.rvz[0], s.rvz[1], s.q1, s.q2};
  const int first = phase == PH_FWD ? 0 : 2;
  for (long long e = g0; e < n; e += gs)
#pragma unroll
    for (int k = 0; k < 14; k++)
      if (k >= first) (f[k] - s.forg)[e] = 0.0f;
  if (phase == PH_FWD) {
    for (long long e = g0; e < n; e += gs) {
      const int row = (int)(e / s.pitch), col = (int)(e - (long long)row * s.pitch);
      const int iz = row - RTM_MZT, ix = col - RTM_MXL;
      const int r = imin(imax(iz - s.mz0, 0), s.mzn - 1);
      const int c = s.mx0 + imin(imax(ix - s.mxs, 0), s.mxn - 1);
      const size_t m = (size_t)r * s.mpitch + c;
      (s.mrho - s.forg)[e] = s.rho[m];
      (s.mvv - s.forg)[e] = s.vv[m];
    }
    if (s.store_fmt != ST_F32)
      for (long long e = g0; e < 4LL * s.nt; e += gs) s.smax[e] = 0u;
  } else {
    for (long long e = g0; e < ni; e += gs) { s.g1[e] = 0.f; s.illum[e] = 0.f; }
  }
}

// receiver wavefield at it = nt-1 (rtm_model.cu:366-372 with the insertion
// of kernel.cu:72-95), written to the level the first reverse step reads
__global__ __launch_bounds__(NTHR) void k_seed_receivers(const ShotDev* __restrict__ shots, int slot)
{
  const ShotDev s = shots[slot];
  const int ir = blockIdx.x * NTHR + threadIdx.x;
  if (ir >= s.nrec) return;
  const int it = s.nt - 1;
  const float rv = (it <= s.cut[ir]) ? 0.0f : s.rec[(size_t)it * s.nrec + ir];
  float* rp = ((s.nt - 2) & 1) ? s.rp[1] : s.rp[0];
  rp[(size_t)s.npml * s.pitch + s.rcol0 + ir] = s.dt * rv;
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

template <int G, int RPT>
static hipError_t launch_step_t(const LaunchArgs& a, int phase, int save, hipStream_t st)
{
  const dim3 g(grid8(a.ntiles)), b(64 * G);
  if (phase == PH_FWD) hipLaunchKernelGGL((k_fwd<G, RPT>), g, b, 0, st, a, save);
  else hipLaunchKernelGGL((k_rev<G, RPT>), g, b, 0, st, a);
  return hipGetLastError();
}

hipError_t rtm_launch_step(const LaunchArgs& a, int phase, int tz, int save, hipStream_t st)
{
  if (a.ntiles <= 0) return hipSuccess;
  switch (tz) {
    case 16: return launch_step_t<4, 4>(a, phase, save, st);
    case 32: return launch_step_t<4, 8>(a, phase, save, st);
    case 64: return launch_step_t<8, 8>(a, phase, save, st);
    default: return hipErrorInvalidValue;
  }
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
