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
// tiles instead of being read back, so a step reads p, vx, vz once and
// writes them once per cell.  The model (1/rho and rho v^2) is read from one
// array shared by every shot of a launch (the resident model; clamped
// indices give each shot its padded window), which stays in the caches.  Fields have two time levels
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

// The slot descriptor of a step kernel, read through the constant address
// space: no kernel writes a descriptor while a step runs (k_bump advances the
// time origin between launches), so its fields may come from scalar loads and
// be loaded again where needed.  A by-value copy instead has to stay live in
// scalar registers across the whole step (later stores could alias the
// source), which overflows them into vector lanes: about a thousand
// v_readlane reloads in the reverse step.
typedef __attribute__((address_space(4))) const ShotDev const_shot_t;
__device__ __forceinline__ const ShotDev& shot_of(const ShotDev* shots, int slot)
{
  return *(const ShotDev*)((const_shot_t*)shots + slot);
}

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
// store offset of a vx strip cell (kernel.cu:1186-1293 packing), -1 off the
// strip (the halo columns of a reconstruction tile)
__device__ __forceinline__ int strip_vx(const ShotDev& s, int iz, int ix)
{
  if (iz < s.npml || iz >= s.nzp - s.npml) return -1;
  const int r = iz - s.npml;
  if (ix >= s.npml - RTM_R && ix < s.npml) return r * RTM_R + ix - (s.npml - RTM_R);
  const int c0 = s.nxp - s.npml - 1;
  if (ix >= c0 && ix < c0 + RTM_R) return RTM_R * s.nzi + r * RTM_R + ix - c0;
  return -1;
}
// The same offsets split into a row part (wave-uniform: the lanes of a wave
// share their rows) and a column part (per lane, computed once per tile):
// own-column cells then cost one add and a compare.
struct StripCols {
  int vx;    // vx lr column part, -1 off the vx strip columns
  int plr;   // p lr column part, -1 off the p strip columns
  int tb;    // ix - npml on the interior columns (vz / p tb blocks), -1 elsewhere
};
__device__ __forceinline__ StripCols strip_cols(const ShotDev& s, int ix)
{
  const int n = s.npml;
  const int blr = 2 * RTM_R * s.nzi + 2 * RTM_R * s.nxl;
  StripCols c;
  c.vx = (ix >= n - RTM_R && ix < n) ? ix - (n - RTM_R)
       : ((ix >= s.nxp - n - 1 && ix < s.nxp - n - 1 + RTM_R) ? RTM_R * s.nzi + ix - (s.nxp - n - 1) : -1);
  c.plr = (ix >= n - RTM_R && ix < n) ? blr + ix - (n - RTM_R)
        : ((ix >= s.nxp - n && ix < s.nxp - n + RTM_R) ? blr + RTM_R * s.nzi + ix - (s.nxp - n) : -1);
  c.tb = (ix >= n && ix < s.nxp - n) ? ix - n : -1;
  return c;
}
// row part of the vz block offset (rows n-10..n-1 and nzp-n-1..nzp-n+8), -1
__device__ __forceinline__ int strip_row_vz(const ShotDev& s, int iz)
{
  const int n = s.npml, b = 2 * RTM_R * s.nzi;
  if (iz >= n - RTM_R && iz < n) return b + (iz - (n - RTM_R)) * s.nxl;
  const int r0 = s.nzp - n - 1;
  if (iz >= r0 && iz < r0 + RTM_R) return b + RTM_R * s.nxl + (iz - r0) * s.nxl;
  return -1;
}
// row part of the lr blocks (rows n..nzp-n-1), -1
__device__ __forceinline__ int strip_row_lr(const ShotDev& s, int iz)
{
  return (iz >= s.npml && iz < s.nzp - s.npml) ? (iz - s.npml) * RTM_R : -1;
}
// row part of the p tb block (rows n-10..n-1 and nzp-n..nzp-n+9), -1
__device__ __forceinline__ int strip_row_ptb(const ShotDev& s, int iz)
{
  const int n = s.npml, b = 4 * RTM_R * s.nzi + 2 * RTM_R * s.nxl;
  if (iz >= n - RTM_R && iz < n) return b + (iz - (n - RTM_R)) * s.nxl;
  const int r0 = s.nzp - n;
  if (iz >= r0 && iz < r0 + RTM_R) return b + RTM_R * s.nxl + (iz - r0) * s.nxl;
  return -1;
}
__device__ __forceinline__ int soff_vz(int row, const StripCols& c) { return (row >= 0 && c.tb >= 0) ? row + c.tb : -1; }
__device__ __forceinline__ int soff_vx(int lr, const StripCols& c) { return (lr >= 0 && c.vx >= 0) ? lr + c.vx : -1; }
__device__ __forceinline__ int soff_p(int lr, int tb, const StripCols& c)
{
  return (lr >= 0 && c.plr >= 0) ? lr + c.plr : ((tb >= 0 && c.tb >= 0) ? tb + c.tb : -1);
}

// A stored strip value is fetched with the first loads of a pass (raw: the
// fp32 value, or the 16-bit code in the low bits) and decoded where it is
// used, with the four block scales of its step from a small LDS table: the
// compute phases then wait for no store reads.
__device__ __forceinline__ float store_raw(const ShotDev& s, int it, int off)
{
  const long long k = (long long)it * s.strip_len + off;
  if (s.store_fmt == ST_F32) return GCF(s.store)[k];
  return __uint_as_float((unsigned int)((const gushort*)s.store)[k]);
}
// blk: the strip block of the value (0 vx, 1 vz, 2 p-lr, 3 p-tb)
__device__ __forceinline__ float store_decode(const ShotDev& s, int blk, float raw, const float* sc4)
{
  if (s.store_fmt == ST_F32) return raw;
  const float sc = sc4[blk];
  const unsigned int h = __float_as_uint(raw);
  if (s.store_fmt == ST_F16) return __half2float(__ushort_as_half((unsigned short)h)) * sc;
  return __uint_as_float(h << 16) * sc;
}

// ------------------------------------------------------------ tile shape --
// G waves of 64 lanes; each thread owns RPT consecutive rows of one column.
template <int G_, int RPT_>
struct Geo {
  static constexpr int G = G_, RPT = RPT_;
  static constexpr int TZ = G * RPT;                   // tile rows
  static constexpr int NT = 64 * G;                    // threads
  static constexpr int PW = 128;                       // p row: columns x0-20 .. x0+83 (pitch a
                                                       // multiple of 64 for stride-64 paired reads)
  static constexpr int VXW = 84;                       // vx row: columns x0-10 .. x0+73
  static constexpr int ZR = TZ + 19;                   // vz rows z0-10 .. z0+TZ+8
  static constexpr int KZ = (ZR + G - 1) / G;          // vz rows per thread (consecutive)
  static constexpr int PZR = G * KZ + 19;              // p rows held (>= TZ + 38; the last
                                                       // thread's window may run past TZ+38)
  static constexpr int KP = (TZ + 38 + G - 1) / G;     // p rows staged per thread
  static constexpr int KH = (20 * TZ + NT - 1) / NT;   // halo vx cells per thread
  static_assert((20 * TZ) % 64 == 0, "halo cells fill whole waves");
  static constexpr int LDS_A = PZR * PW;               // p rows z0-19.. (x halo on the tile rows)
  static constexpr int LDS_B = TZ * VXW + ZR * 64 + TZ * 64;   // VXL | VZL | SXL
  static constexpr int LDS = LDS_A > LDS_B ? LDS_A : LDS_B;
  // damping factors of the tile (after LDS; written once per pass): a, b at
  // the vz rows, at the vx columns x0-10..x0+73, and at the own rows
  static constexpr int LDS_D = 2 * ZR + 2 * VXW + 2 * TZ;
};

// ------------------------------------------------------- buffer access --
// Every field of a slot is one allocation of frows x pitch floats.  Tile
// accesses are raw buffer loads / stores with the row offset in a scalar
// register and the lane's column offset in a vector register, so the
// addressing costs no vector instructions per access.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t field_rsrc(const float* f00, const ShotDev& s)
{
  return __builtin_amdgcn_make_buffer_rsrc((void*)(f00 - s.forg), (short)0, s.frows * s.pitch * 4, 0x00020000);
}
// imaging accumulators: nzi x nxl, pitch nxl, addressed from interior cell (0,0)
__device__ __forceinline__ rsrc_t image_rsrc(float* a, const ShotDev& s)
{
  return __builtin_amdgcn_make_buffer_rsrc((void*)a, (short)0, s.nzi * s.nxl * 4, 0x00020000);
}
// the model a slot reads (mzn rows of mpitch floats)
__device__ __forceinline__ rsrc_t model_rsrc(const float* m, const ShotDev& s)
{
  return __builtin_amdgcn_make_buffer_rsrc((void*)m, (short)0, s.mzn * s.mpitch * 4, 0x00020000);
}
// byte offsets into the model of padded row iz (wave-uniform) and column ix:
// the window cut with edge replication (extractvel1 + extendvel1,
// Velocity_function.cpp:2011-2073) as a clamp
__device__ __forceinline__ int model_row(const ShotDev& s, int iz)
{
  return imin(imax(iz - s.mz0, 0), s.mzn - 1) * s.mpitch * 4;
}
__device__ __forceinline__ int model_col(const ShotDev& s, int ix)
{
  return (s.mx0 + imin(imax(ix - s.mxs, 0), s.mxn - 1)) * 4;
}
// A load offset past any buffer's end: a buffer load whose lane offset is
// out of range returns 0 for that lane without fetching memory (the same
// holds for the loads that write LDS).  Field cells outside the grid's
// columns -- the margins, and the lanes of a window's last tile column past
// nxp-1 -- are never written and hold 0, so their loads take this offset:
// same values, and the part of the last tile column's lines past the window
// is not fetched.
#define RTM_OOB ((int)0x80000000)
__device__ __forceinline__ int col_load(int voff, int col, int nxp)
{
  return (col >= 0 && col < nxp) ? voff : (voff | RTM_OOB);
}
// a lane's offset, or the out-of-range one where the lane takes no part:
// its load returns 0 and its store is dropped, with no exec-mask branch
__device__ __forceinline__ int pred_off(int voff, bool on) { return on ? voff : (voff | RTM_OOB); }
__device__ __forceinline__ float bld(rsrc_t r, int voff, int soff)
{
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
// Field stores carry the non-temporal hint (cache-policy bit 1, "nt" on
// gfx950): a value written in one step is next read in the following step,
// after the whole batch (far more than the L2 and the Infinity Cache hold)
// has streamed through the caches.  Measured +0.4-0.8 % on the 200-shot
// survey (three alternations with the cached form on one device).
__device__ __forceinline__ void bst(rsrc_t r, float v, int voff, int soff)
{
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 2);
}
// An LDS pointer the compiler must materialise as one register: the taps of
// a stencil then use the small immediate offsets of that base instead of a
// separate address add per access (large constant row offsets do not fit
// the 8-bit offsets of the paired LDS reads)
typedef __attribute__((address_space(3))) const float lds_cfloat_t;
__device__ __forceinline__ const float* lds_base(const float* p)
{
  unsigned a = (unsigned)(uintptr_t)(lds_cfloat_t*)p;
  asm volatile("" : "+v"(a));
  return (const float*)(lds_cfloat_t*)(uintptr_t)a;
}

// load one dword per lane into LDS: lane i writes lds_row[i] (lds_row is
// wave-uniform)
typedef __attribute__((address_space(3))) void lds_void_t;
__device__ __forceinline__ void bld_lds(rsrc_t r, float* lds_row, int voff, int soff)
{
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void_t*)lds_row, 4, voff, soff, 0, 0);
}

// true for every lane of the wave if any active lane has c: branches on it
// are uniform, so the rare cases (reduced radius near the grid edge, damped
// cells) cost their instructions only in the waves that hold them
__device__ __forceinline__ bool any_lane(bool c) { return __builtin_amdgcn_ballot_w64(c) != 0; }

// radius-r staggered difference a[i] - b[i] with the coefficient row of
// order 2r; the fma chain runs i = 1..r as the reference loops
// (kernel.cu:459-460).  cf is the coefficient table in LDS with each row
// zero beyond its order, so a reduced-radius lane runs the same ten steps
// with its own row and no predicate: the extra steps add 0*(a-b) = +-0,
// which leaves the sum unchanged (the operands past the radius are finite
// cells of the staged halo or the zero margins).  Row r-1 = 9 is c10, so
// lanes at full radius give the same sums in either branch.  The reduced
// branch writes the difference as -(b - a), the same value: were it the
// same expression as the full branch's, the compiler would hoist the ten
// differences above the branch and hold them in registers across it.
#define STENCIL_SUM(acc, r, c10, cf, A, B)                                                        \
  do {                                                                                            \
    if (!any_lane((r) != RTM_R)) {                                                                \
      _Pragma("unroll") for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], (A) - (B), acc); \
    } else {                                                                                      \
      const float* cr_ = lds_base((cf) + ((r) - 1) * RTM_R); /* read per use, not hoisted */      \
      _Pragma("unroll") for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(cr_[i - 1], -((B) - (A)), acc); \
    }                                                                                             \
  } while (0)

// entry t of the coefficient table as the kernels hold it: row t/10 (order
// 2(t/10+1)) zero beyond its t/10+1 coefficients
__device__ __forceinline__ float ctab_entry(const ShotDev& s, int t)
{
  return (t % RTM_R) <= (t / RTM_R) ? GCF(s.coef)[t] : 0.0f;
}

// field kinds of one pass
enum : int { FK_SRC = 0, FK_RCV = 1, FK_RCN = 2 };

// New velocity of one cell from the p window Pr (Pr[i] = p at +i along the
// derivative), old value vo and reciprocal density ib.  Forward-type fields
// (source in the forward phase, receiver in the reverse phase):
//   v = (c1*v - u)/c2, u = (sum c_i (p[+i] - p[-i+1]) * dt/h) / rho
//   (kernel.cu:588-594), evaluated as u = (s*dt/h)*ib, v = fma(a, v, -(u*b))
//   with a = c1/c2, b = 1/c2; off the absorbing layer (a = b = 1) that is
//   v - u (kernel.cu:462); radius min(10, distance to the grid edge)
//   (cuda_norderx/z, kernel.cu:299-317).  VALID false: the velocity does
//   not exist there (its value is never read); the damping arrays carry 1
//   outside the grid.
// Reconstruction: strip cells restored from the store of step it+1
// (kernel.cu:1294-1347; VO then holds the raw stored value), interior cells
// v + u as fma(s*dt/h, ib, v) (kernel.cu:318-388).
#define VEL_CELL(FK, gen, strips, SOFF, VALID, RAD, INTERIOR, P_AT, VO, IB, DTH, AV, BV, OUT)               \
  do {                                                                                                     \
    float v_ = 0.0f;                                                                                       \
    if (FK == FK_RCN) {                                                                                    \
      /* computed for every cell and then selected (no divergent branch): */                               \
      /* cells off the interior stay 0, strip cells take the stored value */                               \
      float acc = 0.0f;                                                                                    \
      _Pragma("unroll") for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], P_AT(i) - P_AT(1 - i), acc); \
      v_ = __builtin_fmaf(acc * (DTH), (IB), (VO));                                                        \
      if ((gen) && !(INTERIOR)) v_ = 0.0f;                                                                 \
      if (strips) {                                                                                        \
        const int sb_ = (SOFF);  /* strip block of a strip cell, -1 elsewhere */                            \
        if (sb_ >= 0) v_ = store_decode(s, sb_, (VO), DMP);                                                \
      }                                                                                                    \
    } else if (!(gen)) {                                                                                   \
      float acc = 0.0f;                                                                                    \
      _Pragma("unroll") for (int i = 1; i <= RTM_R; i++) acc = __builtin_fmaf(c10[i - 1], P_AT(i) - P_AT(1 - i), acc); \
      v_ = (VO) - (acc * (DTH)) * (IB);                                                                    \
    } else {                                                                                               \
      const int rr_ = (VALID) ? (RAD) : RTM_R;                                                             \
      float acc = 0.0f;                                                                                    \
      STENCIL_SUM(acc, rr_, c10, ctab, P_AT(i), P_AT(1 - i));                                              \
      const float u_ = (acc * (DTH)) * (IB);                                                               \
      /* one form for every cell of the tile: a = b = 1 (undamped) gives */                                \
      /* fma(1, v, -u) == v - u exactly, cheaper than a branch per cell */                                 \
      v_ = __builtin_fmaf((AV), (VO), -(u_ * (BV)));                                                       \
    }                                                                                                      \
    OUT = v_;                                                                                              \
  } while (0)

// ------------------------------------------------------------ pressure --
struct DivV { float sx, sz; };

// Staggered difference of vz along z (kernel.cu:520-534) of thread-local row
// r from the column window cz of vz rows; the x half (sx) comes from the row
// segments of the pressure phase
template <int NW>
__device__ __forceinline__ float div_z(const float (&cz)[NW], int r, int rad, const float* c10,
                                       const float* __restrict__ cf)
{
  float sz = 0.0f;
  if (rad == RTM_R) {
#pragma unroll
    for (int i = 1; i <= RTM_R; i++) sz = __builtin_fmaf(c10[i - 1], cz[r + 9 + i] - cz[r + 10 - i], sz);
  } else {
    const float* c = lds_base(cf + (rad - 1) * RTM_R);
#pragma unroll
    for (int i = 1; i <= RTM_R; i++) sz = __builtin_fmaf(c[i - 1], cz[r + 9 + i] - cz[r + 10 - i], sz);
  }
  return sz;
}

// p - rho v^2 (sx dt/dx + sz dt/dz) (kernel.cu:536-539), kk = rho v^2
__device__ __forceinline__ float pres_sum(const DivV& d, float dtx, float dtz)
{
  return __builtin_fmaf(d.sx, dtx, d.sz * dtz);
}

// One field through one time step on a tile: p of the current level pc (with
// velocities vxc, vzc) to the next level (pn, vxn, vzn).  On return pnew
// holds the new pressure of the thread's own cells (rows g*RPT+r); ps is the
// reconstructed source pressure of the same cells (receiver pass imaging).
//
// Thread roles (lane lx, wave g):
//   p staging  rows g+G*m of the tile columns (PZ) and their x wings (PC)
//   vz         rows z0-10 + g*KZ + k of column x0+lx (z window from PZ)
//   vx         own rows z0 + g*RPT + r of column x0+lx, and the halo
//              columns x0-10..x0-1, x0+64..x0+73 dealt over all threads
//   p          own rows z0 + g*RPT + r of column x0+lx
template <class Q, int FK>
__device__ __forceinline__ void field_step(float* __restrict__ lds, const ShotDev& s, const Tile& tl, int it,
                                           const float* pc, const float* vxc, const float* vzc, float* pn, float* vxn,
                                           float* vzn, const float* __restrict__ ctab, const float (&ps)[Q::RPT],
                                           float (&pnew)[Q::RPT], unsigned char* actw = nullptr)
{
  constexpr int G = Q::G, RPT = Q::RPT, TZ = Q::TZ, NT = Q::NT, KZ = Q::KZ, KP = Q::KP, KH = Q::KH;
  // the OR of the bit patterns of the lane's stored field values (-0 counts
  // as non-zero): written to actw[wave] at the end for the zero-tile skip
  unsigned int nzb = 0u;
  constexpr int PW = Q::PW, VXW = Q::VXW, ZR = Q::ZR;
  float* PS = lds;                                     // p rows z0-19.., columns x0-20..x0+83
  float* DMP = lds + Q::LDS;                           // damping table (damped passes of gen tiles)
  constexpr int DZA = 0, DZB = ZR, DXA = 2 * ZR, DXB = 2 * ZR + VXW, DPA = 2 * ZR + 2 * VXW, DPB = DPA + TZ;
  float* VXL = lds;                                    // after the second barrier
  float* VZL = lds + TZ * VXW;
  const int tid = threadIdx.x, lx = tid & 63;
  const int g = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int x0 = tl.x0, z0 = tl.z0, nxp = s.nxp, nzp = s.nzp, npml = s.npml;
  // halo vx cell k*NT + tid exists for the whole wave or none of it (20*TZ
  // is a multiple of 64): a wave-uniform test
  auto halo_on = [&](int k) { return k * NT + g * 64 < 20 * TZ; };
  const int ix = x0 + lx;
  const bool gen = (tl.kind & TF_GEN) != 0;
  // strip cells that the reconstruction restores (the tile's and its halo's);
  // the forward pass saves its strips in k_save after the step
  const bool strips = FK == FK_RCN && (tl.kind & TF_STRIP) != 0;
  const int pr4 = s.pitch * 4;
  const int srow0 = (z0 + RTM_MZT) * pr4;             // byte offset of tile row 0
  const int vcol = (ix + RTM_MXL) * 4;                // byte offset of the lane's column
  const int vld = col_load(vcol, ix, nxp);             // ... for loads
  const int icol = (ix - npml) * 4;                    // ... in the imaging accumulators
  const rsrc_t Rp = field_rsrc(pc, s), Rvx = field_rsrc(vxc, s), Rvz = field_rsrc(vzc, s);
  const rsrc_t Rib = model_rsrc(s.mib, s);
  const int mcol = model_col(s, ix);                   // the lane's model column
  float c10[RTM_R];
#pragma unroll
  for (int i = 0; i < RTM_R; i++) c10[i] = GCF(s.coef)[90 + i];
  const float dtx = s.dtx, dtz = s.dtz;

  // ---- 1. loads: p with its halo straight into LDS (buffer loads with the
  //         lds modifier: lane i of a wave writes its dword to base + 4 i),
  //         old velocities and density of every velocity cell into registers
  {
    // tile rows: columns x0-20..x0+83 as two full-wave loads (x0-20..x0+43
    // and x0+20..x0+83; the middle 24 columns arrive twice with the same
    // values), so that no load is issued by part of a wave; halo rows:
    // columns x0..x0+63
    const int ca = col_load((x0 - 20 + lx + RTM_MXL) * 4, x0 - 20 + lx, nxp);
    const int cb = col_load((x0 + 20 + lx + RTM_MXL) * 4, x0 + 20 + lx, nxp);
#pragma unroll
    for (int m = 0; m < KP; m++) {
      const int k = g + G * m;
      if (k < TZ + 38) {
        const int so = srow0 + (k - 19) * pr4;
        if (k >= 19 && k < 19 + TZ) {
          bld_lds(Rp, PS + k * PW, ca, so);
          bld_lds(Rp, PS + k * PW + 40, cb, so);
        } else {
          bld_lds(Rp, PS + k * PW + 20, vld, so);
        }
      }
    }
  }
  float vzo[KZ], ibz[KZ];
#pragma unroll
  for (int k = 0; k < KZ; k++) {
    const int kz = g * KZ + k;
    vzo[k] = 0.0f;
    ibz[k] = 1.0f;
    if (kz < ZR) {
      const int so = srow0 + (kz - 10) * pr4;
      vzo[k] = bld(Rvz, vld, so);
      ibz[k] = bld(Rib, mcol, model_row(s, z0 - 10 + kz));
    }
  }
  // reconstruction of a strip tile: the stored value of a strip cell
  // replaces its old velocity, fetched now with the rest (after the old
  // value's load, which it overwrites) and decoded where it is used; which
  // cells are strip cells is kept as bits of sm (vz rows, vx rows, halo vx,
  // p rows, p rows in the top/bottom block), so the offsets are computed once
  StripCols sc{-1, -1, -1};
  if (strips) sc = strip_cols(s, ix);
  const bool rstrips = FK == FK_RCN && strips;
  constexpr int SMX = KZ, SMH = KZ + RPT, SMP = KZ + RPT + KH, SMT = KZ + RPT + KH + RPT;
  static_assert(SMT + RPT <= 64, "strip bits");
  unsigned long long sm = 0;
  if (rstrips) {
#pragma unroll
    for (int k = 0; k < KZ; k++) {
      const int kz = g * KZ + k;
      const int st = kz < ZR ? soff_vz(strip_row_vz(s, z0 - 10 + kz), sc) : -1;
      if (st >= 0) { vzo[k] = store_raw(s, it + 1, st); sm |= 1ull << k; }
    }
  }
  auto sbit = [&](int b) { return ((sm >> b) & 1ull) != 0; };
  // per-lane column data of the generic path: existence, radius, damping;
  // the damping of the tile's rows and halo columns goes to an LDS table, so
  // that the compute phases wait for no small loads
  bool vxv = true;
  int rrx = RTM_R;
  float axv = 1.0f, bxv = 1.0f, ax = 1.0f, bx = 1.0f;
  if (gen) {
    vxv = ix <= nxp - 2;
    rrx = vxv ? imin(RTM_R, imin(ix + 1, nxp - 1 - ix)) : RTM_R;
  }
  if (FK != FK_RCN && gen) {
    axv = GCF(s.axv)[ix];
    bxv = GCF(s.bxv)[ix];
    if (ix < nxp) { ax = GCF(s.ax)[ix]; bx = GCF(s.bx)[ix]; }
    // rows z0-10.. and columns x0-10.. lie within the arrays' margins
    if (tid < ZR) { DMP[DZA + tid] = GCF(s.azv)[z0 - 10 + tid]; DMP[DZB + tid] = GCF(s.bzv)[z0 - 10 + tid]; }
    if (tid < VXW) { DMP[DXA + tid] = GCF(s.axv)[x0 - 10 + tid]; DMP[DXB + tid] = GCF(s.bxv)[x0 - 10 + tid]; }
    if (tid < TZ) { DMP[DPA + tid] = GCF(s.az)[z0 + tid]; DMP[DPB + tid] = GCF(s.bz)[z0 + tid]; }
  }
  // 16-bit store: block scales of the steps the reconstruction reads (it+1:
  // velocities, it: pressure) in the damping table's place (unused there)
  if (rstrips && s.store_fmt != ST_F32 && tid < 8)
    DMP[tid] = GCF(s.sscale)[(long long)(tid < 4 ? it + 1 : it) * 4 + (tid & 3)];
  __builtin_amdgcn_s_waitcnt(0);                       // the LDS writes of the staging loads
  __syncthreads();                                     // p staged
  float vxo[RPT], ibx[RPT];
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const int so = srow0 + (g * RPT + r) * pr4;
    vxo[r] = bld(Rvx, vld, so);
    ibx[r] = bld(Rib, mcol, model_row(s, z0 + g * RPT + r));
  }
  float hvo[KH], hib[KH];
  int hrow[KH], hcol[KH];
#pragma unroll
  for (int k = 0; k < KH; k++) {
    const int h = k * NT + tid;
    hvo[k] = 0.0f;
    hib[k] = 1.0f;
    hrow[k] = h / 20;
    const int hc = h - hrow[k] * 20;
    hcol[k] = hc < 10 ? hc : hc + 64;                  // VXL column: x = x0 - 10 + hcol
    if (halo_on(k)) {
      const int off = col_load(srow0 + hrow[k] * pr4 + (x0 - 10 + hcol[k] + RTM_MXL) * 4, x0 - 10 + hcol[k], nxp);
      hvo[k] = bld(Rvx, off, 0);
      hib[k] = bld(Rib, model_col(s, x0 - 10 + hcol[k]) +
                            imin(imax(z0 + hrow[k] - s.mz0, 0), s.mzn - 1) * s.mpitch * 4, 0);
    }
  }
  if (rstrips) {
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int st = soff_vx(strip_row_lr(s, z0 + g * RPT + r), sc);
      if (st >= 0) { vxo[r] = store_raw(s, it + 1, st); sm |= 1ull << (SMX + r); }
    }
#pragma unroll
    for (int k = 0; k < KH; k++) {
      const int st = halo_on(k) ? strip_vx(s, z0 + hrow[k], x0 - 10 + hcol[k]) : -1;
      if (st >= 0) { hvo[k] = store_raw(s, it + 1, st); sm |= 1ull << (SMH + k); }
    }
  }

  // ---- 2. new velocities
  float vzn_[KZ], vxn_[RPT], hvn[KH];
  {
    const rsrc_t Rvzn = field_rsrc(vzn, s);
    float cw[KZ + 19];                                 // p rows z0-19+g*KZ+m of the column
#pragma unroll
    for (int m = 0; m < KZ + 19; m++) cw[m] = PS[(g * KZ + m) * PW + 20 + lx];
#pragma unroll
    for (int k = 0; k < KZ; k++) {
      const int kz = g * KZ + k;
      const int iz = z0 - 10 + kz;                     // wave-uniform
#define PZA(i) cw[k + 9 + (i)]
      VEL_CELL(FK, gen, strips, (sbit(k) ? 1 : -1), (iz >= 0 && iz <= nzp - 2), imin(RTM_R, imin(iz + 1, nzp - 1 - iz)),
               (ix >= npml && ix < nxp - npml && iz >= npml && iz <= nzp - npml - 2), PZA, vzo[k], ibz[k], dtz,
               DMP[DZA + kz], DMP[DZB + kz], vzn_[k]);
#undef PZA
      if (kz >= 10 && kz < 10 + TZ) {                  // own rows (wave-uniform)
        if (!gen) bst(Rvzn, vzn_[k], vcol, srow0 + (kz - 10) * pr4);   // every lane's cell exists
        else bst(Rvzn, vzn_[k], pred_off(vcol, ix < nxp && iz <= nzp - 2), srow0 + (kz - 10) * pr4);
        nzb |= __float_as_uint(vzn_[k]);
      }
    }
  }

  // operands of the pressure phase, fetched now so that one memory round trip
  // for all rows overlaps the vx computation (inside the row loop each row
  // would wait for its own): rho v^2 of the own cells, and per cell either
  // the PML split pair (forward-type passes, damped cells) or the imaging
  // accumulators (receiver pass, interior cells) -- never both
  const bool xin = ix >= npml && ix < nxp - npml;
  const bool xring = ix == 0 || ix == nxp - 1;
  float kk[RPT], qa[RPT], qb[RPT];
  {
    const rsrc_t Rkk = model_rsrc(s.mkk, s);
    const rsrc_t R1 = field_rsrc(s.q1, s), R2 = field_rsrc(s.q2, s);
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int j = g * RPT + r, iz = z0 + j;
      const int so = srow0 + j * pr4;
      kk[r] = bld(Rkk, mcol, model_row(s, iz));
      qa[r] = qb[r] = 0.0f;
      const bool incell = ix < nxp && iz < nzp;
      const bool inner = xin && iz >= npml && iz < nzp - npml;
      const int io = imax(iz - npml, 0) * s.nxl * 4;   // wave-uniform row offset (rows >= npml used)
      if (FK == FK_RCV && !gen) {                      // every lane's cell is interior
        qa[r] = bld(image_rsrc(s.g1, s), icol, io);
        qb[r] = bld(image_rsrc(s.illum, s), icol, io);
      } else if (FK != FK_RCN && gen) {
        // the split pair of damped cells, the accumulators of interior cells:
        // loads with predicated offsets (0 where a lane takes no part), each
        // pair issued only by waves that hold such a cell
        const bool dm = incell && iz > 0 && iz < nzp - 1 && !xring && !inner;
        if (any_lane(dm)) {
          const int pq = pred_off(vcol, dm);
          qa[r] = bld(R1, pq, so);
          qb[r] = bld(R2, pq, so);
        }
        if (FK == FK_RCV) {
          const bool im = incell && inner;
          if (any_lane(im)) {
            const int pi = pred_off(icol, im);
            const float a2 = bld(image_rsrc(s.g1, s), pi, io), b2 = bld(image_rsrc(s.illum, s), pi, io);
            qa[r] = im ? a2 : qa[r];
            qb[r] = im ? b2 : qb[r];
          }
        }
      }
      if (rstrips) {                                   // reconstruction: stored p of strip cells
        const int lr = strip_row_lr(s, iz);
        const int sp_ = soff_p(lr, strip_row_ptb(s, iz), sc);
        if (sp_ >= 0) {
          qa[r] = store_raw(s, it, sp_);
          sm |= 1ull << (SMP + r);
          if (lr < 0) sm |= 1ull << (SMT + r);         // top/bottom block (rows outside n..nzp-n-1)
        }
      }
    }
  }
  {
    const rsrc_t Rvxn = field_rsrc(vxn, s);
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int j = g * RPT + r, iz = z0 + j;
      const float* __restrict__ Pr = lds_base(PS + (j + 19) * PW + lx + 11) + 9;   // p at column ix
#define PCA(i) Pr[(i)]
      VEL_CELL(FK, gen, strips, (sbit(SMX + r) ? 0 : -1), vxv, rrx,
               (ix >= npml && ix <= nxp - npml - 2 && iz >= npml && iz < nzp - npml), PCA, vxo[r], ibx[r], dtx,
               axv, bxv, vxn_[r]);
#undef PCA
      if (!gen) bst(Rvxn, vxn_[r], vcol, srow0 + j * pr4);
      else bst(Rvxn, vxn_[r], pred_off(vcol, iz < nzp && ix <= nxp - 2), srow0 + j * pr4);
      nzb |= __float_as_uint(vxn_[r]);
    }
#pragma unroll
    for (int k = 0; k < KH; k++) {
      const int iz = z0 + hrow[k], jx = x0 - 10 + hcol[k];
      const float* __restrict__ Pr = lds_base(PS + (hrow[k] + 19) * PW + hcol[k] + 1) + 9;   // p at column jx
      hvn[k] = 0.0f;
      if (halo_on(k)) {
        const bool hv = jx >= 0 && jx <= nxp - 2;
#define PCA(i) Pr[(i)]
        VEL_CELL(FK, gen, strips, (sbit(SMH + k) ? 0 : -1), hv, (hv ? imin(RTM_R, imin(jx + 1, nxp - 1 - jx)) : RTM_R),
                 (jx >= npml && jx <= nxp - npml - 2 && iz >= npml && iz < nzp - npml), PCA, hvo[k], hib[k], dtx,
                 DMP[DXA + hcol[k]], DMP[DXB + hcol[k]], hvn[k]);
#undef PCA
      }
    }
  }

  // own old pressure out of LDS
  float po[RPT];
#pragma unroll
  for (int r = 0; r < RPT; r++) po[r] = PS[(g * RPT + r + 19) * PW + 20 + lx];
  __syncthreads();                                     // PS consumed
#pragma unroll
  for (int k = 0; k < KZ; k++) {
    const int kz = g * KZ + k;
    if (kz < ZR) VZL[kz * 64 + lx] = vzn_[k];
  }
#pragma unroll
  for (int r = 0; r < RPT; r++) VXL[(g * RPT + r) * VXW + lx + 10] = vxn_[r];
#pragma unroll
  for (int k = 0; k < KH; k++)
    if (halo_on(k)) VXL[hrow[k] * VXW + hcol[k]] = hvn[k];
  __syncthreads();                                     // new velocities staged

  // ---- 3. x-differences of the new vx by row segments: thread tid takes
  //         RPT consecutive cells of one tile row and reads their vx window
  //         once (RPT + 19 values instead of 20 per cell), and parks the sums
  //         in SXL for the pressure update.  The radius is the cell's own
  //         (the same rule as the pressure update's); a segment of a wave
  //         with any reduced-radius cell takes the zero-padded coefficient
  //         rows, which give the radius-10 sums bit for bit.
  float* SXL = lds + TZ * VXW + ZR * 64;
  {
    constexpr int NSEG = 64 / RPT;                     // NSEG * TZ == NT
    const int j = tid / NSEG, c0 = (tid - j * NSEG) * RPT;
    const int iz = z0 + j;
    float w[RPT + 19];                                 // vx at columns x0 + c0 - 10 ..
#pragma unroll
    for (int m = 0; m < RPT + 19; m++) w[m] = VXL[j * VXW + c0 + m];
    int radc[RPT];
    bool red = false;
#pragma unroll
    for (int c = 0; c < RPT; c++) {
      radc[c] = RTM_R;
      if (FK != FK_RCN && gen) {
        const int jx = x0 + c0 + c;
        const bool upd = iz > 0 && iz < nzp - 1 && jx < nxp && jx != 0 && jx != nxp - 1;
        if (upd) radc[c] = imin(imin(RTM_R, imin(jx, nxp - 1 - jx)), imin(iz, nzp - 1 - iz));
        red = red || radc[c] != RTM_R;
      }
    }
    if (FK == FK_RCN || !gen || !any_lane(red)) {
#pragma unroll
      for (int c = 0; c < RPT; c++) {
        float sx = 0.0f;
#pragma unroll
        for (int i = 1; i <= RTM_R; i++) sx = __builtin_fmaf(c10[i - 1], w[c + 9 + i] - w[c + 10 - i], sx);
        SXL[j * 64 + c0 + c] = sx;
      }
    } else {
#pragma unroll
      for (int c = 0; c < RPT; c++) {
        const float* cr = lds_base(ctab + (radc[c] - 1) * RTM_R);
        float sx = 0.0f;
#pragma unroll
        for (int i = 1; i <= RTM_R; i++) sx = __builtin_fmaf(cr[i - 1], w[c + 9 + i] - w[c + 10 - i], sx);
        SXL[j * 64 + c0 + c] = sx;
      }
    }
  }
  __syncthreads();                                     // x-differences staged

  // ---- 4. new pressure of the own cells
  float cz[RPT + 19];
#pragma unroll
  for (int m = 0; m < RPT + 19; m++) cz[m] = VZL[(g * RPT + m) * 64 + lx];
  const int radx = imin(RTM_R, imin(ix, nxp - 1 - ix));
  const bool xsrc = (tl.kind & TF_SRC) && ix >= s.sx0 && ix <= s.sx1;
  const rsrc_t Rpn = field_rsrc(pn, s);

#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const int j = g * RPT + r;
    const int iz = z0 + j;                             // wave-uniform
    const int so = srow0 + j * pr4;
    float p = po[r];
    const float sxl = SXL[j * 64 + lx];
    const bool incell = ix < nxp && iz < nzp;
    if (FK == FK_RCN) {
      const bool inner = xin && iz >= npml && iz < nzp - npml;
      // computed for every cell and then selected (no divergent branch):
      // interior cells take the update, strip cells the stored value
      DivV d;
      d.sx = sxl;
      d.sz = div_z(cz, r, RTM_R, c10, ctab);
      const float pc = __builtin_fmaf(kk[r], pres_sum(d, dtx, dtz), p);   // kernel.cu:389-429
      bool wr = !gen || inner;
      if (wr) p = pc;
      if (strips && sbit(SMP + r)) {
        p = store_decode(s, sbit(SMT + r) ? 3 : 2, qa[r], DMP + 4);   // kernel.cu:1348-1400
        wr = true;
      }
      if ((tl.kind & TF_SRC) && iz >= s.sz0 && iz <= s.sz1 && wr && xsrc)   // source removal (rtm_model.cu:437)
        p = __builtin_fmaf(-GCF(s.sa)[(iz - s.sz0) * (s.sx1 - s.sx0 + 1) + ix - s.sx0], GCF(s.wav)[it], p);
      if (!gen) bst(Rpn, p, vcol, so);                 // every lane's cell is interior
      else bst(Rpn, p, pred_off(vcol, wr), so);
    } else {
      if (!gen) {
        DivV d;
        d.sx = sxl;
        d.sz = div_z(cz, r, RTM_R, c10, ctab);
        p = __builtin_fmaf(-kk[r], pres_sum(d, dtx, dtz), p);  // kernel.cu:539
      } else {
        // the outer ring keeps its value (only the source is added there);
        // cells off the grid are computed with radius 10 and dropped
        const bool zupd = iz > 0 && iz < nzp - 1;
        const bool upd = zupd && incell && !xring;
        const int rad = upd ? imin(radx, imin(iz, nzp - 1 - iz)) : RTM_R;
        DivV d;
        d.sx = sxl;
        if (!any_lane(rad != RTM_R)) d.sz = div_z(cz, r, RTM_R, c10, ctab);
        else d.sz = div_z(cz, r, rad, c10, ctab);
        const bool split = upd && !(xin && iz >= npml && iz < nzp - npml);
        float pu = __builtin_fmaf(-kk[r], pres_sum(d, dtx, dtz), p);
        if (any_lane(split)) {                         // split pair, p = p1 + p2 (kernel.cu:780-783)
          const rsrc_t R1 = field_rsrc(s.q1, s), R2 = field_rsrc(s.q2, s);
          float q1 = qa[r], q2 = qb[r];
          const float dxv = (d.sx * dtx) * kk[r], dzv = (d.sz * dtz) * kk[r];
          q1 = __builtin_fmaf(ax, q1, -(dxv * bx));   // (c1 q - d)/c2 (kernel.cu:774-779)
          q2 = __builtin_fmaf(DMP[DPA + j], q2, -(dzv * DMP[DPB + j]));
          const int pq = pred_off(vcol, split);
          bst(R1, q1, pq, so);
          bst(R2, q2, pq, so);
          if (split) nzb |= __float_as_uint(q1) | __float_as_uint(q2);
          if (split) pu = q1 + q2;
        }
        if (upd) p = pu;
      }
      if (FK == FK_SRC) {
        if ((tl.kind & TF_SRC) && iz >= s.sz0 && iz <= s.sz1 && xsrc)   // kernel.cu:36-47
          p = __builtin_fmaf(GCF(s.sa)[(iz - s.sz0) * (s.sx1 - s.sx0 + 1) + ix - s.sx0], GCF(s.wav)[it], p);
        auto record = [&](bool on) {
          if ((tl.kind & TF_REC) && s.record && iz == npml) {   // kernel.cu:48-71
            const int ir = ix - s.rcol0;
            if (on && ir >= 0 && ir < s.nrec) GF(s.rec)[(size_t)it * s.nrec + ir] = p;
          }
        };
        if (!gen) {
          bst(Rpn, p, vcol, so);
          record(true);
        } else {
          bst(Rpn, p, pred_off(vcol, incell), so);
          record(incell);
        }
      } else {                                         // receiver wavefield
        if ((tl.kind & TF_REC) && iz == npml) {        // Dirichlet insertion (kernel.cu:72-95)
          const int ir = ix - s.rcol0;
          if (ir >= 0 && ir < s.nrec) {
            const float rv = (it <= ((const gint*)s.cut)[ir]) ? 0.0f : GCF(s.rec)[(size_t)it * s.nrec + ir];
            p = s.dt * rv;
          }
        }
        auto image = [&]() {                           // imaging (kernel.cu:1401-1416)
          const int io = (iz - npml) * s.nxl * 4;
          bst(image_rsrc(s.g1, s), __builtin_fmaf(ps[r], p, qa[r]), icol, io);
          bst(image_rsrc(s.illum, s), __builtin_fmaf(ps[r], ps[r], qb[r]), icol, io);
        };
        if (!gen) {                                    // every lane's cell is interior
          bst(Rpn, p, vcol, so);
          image();
        } else {
          bst(Rpn, p, pred_off(vcol, incell), so);
          const bool im = incell && xin && iz >= npml && iz < nzp - npml;
          if (any_lane(im)) {
            const int io = imax(iz - npml, 0) * s.nxl * 4;
            const int pi = pred_off(icol, im);
            bst(image_rsrc(s.g1, s), __builtin_fmaf(ps[r], p, qa[r]), pi, io);
            bst(image_rsrc(s.illum, s), __builtin_fmaf(ps[r], ps[r], qb[r]), pi, io);
          }
        }
      }
      nzb |= __float_as_uint(p);
    }
    pnew[r] = p;
  }
  if (FK == FK_SRC && actw) {
    const bool any = any_lane(nzb != 0u);
    if (lx == 0) actw[g] = any ? 1 : 0;
  }
}

// Zero-tile skip of the forward step: a tile skips the step when the forward
// field is zero at the level the step reads, on the tile and its 8
// neighbours (all the step's inputs are then +0 and so would be its
// outputs: the step reads 19 rows and columns out, within the adjacent
// tiles for tiles of 32 or 64 rows; the host passes no words for 16-row
// tiles), and at the level it writes, on the tile itself (the buffer
// already holds those +0).  The wavefield starts at +0 and spreads from the
// source, so far tiles skip until it comes near.  Words of both levels are
// fetched beside the tile's other first loads; `it` picks the level.  (The
// same skip of the reverse step's receiver field cost more than it saved:
// that field rarely has a zero neighbourhood.)
struct SkipWords { unsigned long long self0, self1, nb0, nb1; };   // (no arrays: a runtime index would spill)
__device__ __forceinline__ SkipWords skip_words(const unsigned long long* act, int ntiles, int t, const Tile& tl)
{
  SkipWords w{act[t], act[ntiles + t], 0ull, 0ull};
#pragma unroll
  for (int j = 0; j < 8; j++)
    if (tl.nbr[j] >= 0) {
      w.nb0 |= act[tl.nbr[j]];
      w.nb1 |= act[ntiles + tl.nbr[j]];
    }
  return w;
}
__device__ __forceinline__ bool skip_of(const SkipWords& w, int it)
{
  const bool odd = it & 1;
  const unsigned long long in = odd ? (w.self1 | w.nb1) : (w.self0 | w.nb0);
  const unsigned long long out = odd ? w.self0 : w.self1;
  return (in | out) == 0ull;
}

}  // namespace

// ================================================================ kernels ==
// forward time step of the source wavefield (modelling, recording, strips)
template <int G, int RPT>
__global__ __launch_bounds__(64 * G) void k_fwd(LaunchArgs a)
{
  using Q = Geo<G, RPT>;
  __shared__ float lds[Q::LDS + Q::LDS_D];
  __shared__ float ctab[100];                          // coefficient table (reduced radii near the edges)
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const Tile tl = a.tiles[t];
  if (tl.slot < 0) return;                             // padding of a shorter XCD run
  const bool track = a.act && !(tl.kind & (TF_SRC | TF_REC));
  SkipWords sw{};
  if (track) sw = skip_words(a.act, a.ntiles, t, tl);
  const ShotDev& s = shot_of(a.shots, tl.slot);
  const int it = s.it_off + s.it_dir * a.step;
  if (track && skip_of(sw, it)) return;                // the forward field is +0 around the tile
  if (threadIdx.x < 100) ctab[threadIdx.x] = ctab_entry(s, threadIdx.x);
  const bool odd = it & 1;                             // (a runtime index into s would spill it to scratch)
  float none[Q::RPT], pnew[Q::RPT];
#pragma unroll
  for (int r = 0; r < Q::RPT; r++) none[r] = 0.0f;
  unsigned char* actw = a.act ? (unsigned char*)(a.act + (size_t)((it + 1) & 1) * a.ntiles + t) : nullptr;
  field_step<Q, FK_SRC>(lds, s, tl, it, odd ? s.sp[1] : s.sp[0], odd ? s.svx[1] : s.svx[0],
                        odd ? s.svz[1] : s.svz[0], odd ? s.sp[0] : s.sp[1], odd ? s.svx[0] : s.svx[1],
                        odd ? s.svz[0] : s.svz[1], ctab, none, pnew, actw);
}

// reverse time step: source reconstruction, then the receiver wavefield and
// the imaging products of the tile
template <int G, int RPT>
__global__ __launch_bounds__(64 * G) __attribute__((amdgpu_waves_per_eu(G == 4 && RPT == 4 ? 5 : 1))) void k_rev(LaunchArgs a)
{
  using Q = Geo<G, RPT>;
  __shared__ float lds[Q::LDS + Q::LDS_D];
  __shared__ float ctab[100];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const Tile tl = a.tiles[t];
  if (tl.slot < 0) return;                             // padding of a shorter XCD run
  const ShotDev& s = shot_of(a.shots, tl.slot);
  if (threadIdx.x < 100) ctab[threadIdx.x] = ctab_entry(s, threadIdx.x);
  const int it = s.it_off + s.it_dir * a.step;
  const bool odd = it & 1;
  float ps[Q::RPT], pr[Q::RPT], none[Q::RPT];
#pragma unroll
  for (int r = 0; r < Q::RPT; r++) ps[r] = none[r] = 0.0f;
  if (tl.kind & TF_INT) {
    field_step<Q, FK_RCN>(lds, s, tl, it, odd ? s.sp[1] : s.sp[0], odd ? s.svx[1] : s.svx[0],
                          odd ? s.svz[1] : s.svz[0], odd ? s.sp[0] : s.sp[1], odd ? s.svx[0] : s.svx[1],
                          odd ? s.svz[0] : s.svz[1], ctab, none, ps);
    __syncthreads();                                   // LDS reused by the receiver pass
  }
  field_step<Q, FK_RCV>(lds, s, tl, it, odd ? s.rp[1] : s.rp[0], odd ? s.rvx[1] : s.rvx[0],
                        odd ? s.rvz[1] : s.rvz[0], odd ? s.rp[0] : s.rp[1], odd ? s.rvx[0] : s.rvx[1],
                        odd ? s.rvz[0] : s.rvz[1], ctab, ps, pr);
}

// Boundary strips of forward step `it` -- the new fields of that step -- into
// the store, packed as kernel.cu:1186-1293 per step:
//   vx lr [2][nzi][10] | vz tb [2][10][nxl] | p lr [2][nzi][10] | p tb [2][10][nxl]
// fp32 is a copy (k_save); the 16-bit forms store each block with a scale of
// its own (k_save16).
struct StripBlock {
  const float* f;    // field (cell (0,0)), new level
  int lr;            // 1: [2][nzi][10] column strips, 0: [2][10][nxl] row strips
  int lo, n;         // offset of the block in the step's record, element count
  int a0, a1;        // first column (lr) / row (tb) of side 0 and side 1
};
__device__ __forceinline__ StripBlock strip_block_of(const ShotDev& s, int b, int lev)
{
  const int n = s.npml, e1 = 2 * RTM_R * s.nzi, e2 = e1 + 2 * RTM_R * s.nxl, e3 = e2 + 2 * RTM_R * s.nzi;
  // (levels selected, not indexed: a run-time index would put s in scratch)
  const float* vx = lev ? s.svx[1] : s.svx[0];
  const float* vz = lev ? s.svz[1] : s.svz[0];
  const float* p = lev ? s.sp[1] : s.sp[0];
  StripBlock k;
  if (b == 0) k = {vx, 1, 0, e1, n - RTM_R, s.nxp - n - 1};                  // strip_vx
  else if (b == 1) k = {vz, 0, e1, e2 - e1, n - RTM_R, s.nzp - n - 1};        // strip_vz
  else if (b == 2) k = {p, 1, e2, e3 - e2, n - RTM_R, s.nxp - n};             // strip_p lr
  else k = {p, 0, e3, (int)s.strip_len - e3, n - RTM_R, s.nzp - n};            // strip_p tb
  return k;
}
// element e of a block: its field value
__device__ __forceinline__ float strip_elem(const ShotDev& s, const StripBlock& k, int e)
{
  if (k.lr) {                                          // e = (side*nzi + r)*10 + c
    const int q = e / RTM_R, c = e - q * RTM_R;
    const int side = q >= s.nzi, r = q - side * s.nzi;
    return k.f[(size_t)(s.npml + r) * s.pitch + (side ? k.a1 : k.a0) + c];
  }
  const int q = e / s.nxl, c = e - q * s.nxl;          // e = (side*10 + r)*nxl + c
  const int side = q >= RTM_R, r = q - side * RTM_R;
  return k.f[(size_t)((side ? k.a1 : k.a0) + r) * s.pitch + s.npml + c];
}
// A step's strip record is small (2*40(nz+nxl) values per shot), so a save
// is latency-bound: every thread keeps SV loads in flight, and each block is
// split over gridDim.y workgroups with lane-contiguous elements (every load
// and store of a wave covers 64 consecutive values of the record).
constexpr int SV = 8;
__device__ __forceinline__ int save_it(const ShotDev& s, int step)
{
  return step < 0 ? -step - 1 : s.it_off + s.it_dir * step;   // step < 0: absolute -step-1
}
// the wavefield guard word (k_guard) after the nt x 4 scales
__device__ __forceinline__ unsigned int* guard_word(const ShotDev& s) { return (unsigned int*)(s.sscale + (size_t)s.nt * 4); }

// fp32: copy block b of the slot's step into the store
__global__ __launch_bounds__(NTHR) void k_save(const ShotDev* __restrict__ shots, const int* __restrict__ slots,
                                                int nslots, int step)
{
  const int k = blockIdx.x >> 2, b = blockIdx.x & 3;
  if (k >= nslots) return;
  const ShotDev s = shots[slots[k]];
  const int it = save_it(s, step);
  const StripBlock blk = strip_block_of(s, b, (it + 1) & 1);
  const long long base = (long long)it * s.strip_len + blk.lo;
  const int stride = gridDim.y * NTHR * SV;
  for (int e0 = blockIdx.y * NTHR * SV + threadIdx.x; e0 < blk.n; e0 += stride) {
    float v[SV];
#pragma unroll
    for (int u = 0; u < SV; u++) v[u] = e0 + u * NTHR < blk.n ? strip_elem(s, blk, e0 + u * NTHR) : 0.0f;
#pragma unroll
    for (int u = 0; u < SV; u++)
      if (e0 + u * NTHR < blk.n) GF(s.store)[base + e0 + u * NTHR] = v[u];
  }
}

// 16-bit forms: one workgroup of SQT threads per (slot, block) takes the
// block's max-abs, keeps the scale max(maxabs / 32768, FLT_MIN) (1 for an
// all-zero block; a denormal max-abs, e.g. source tails before the wave
// arrives, must not underflow the scale to zero: |x| <= 32768 * FLT_MIN then
// anyway) and stores x / scale rounded to nearest even.  Blocks of up to
// SQV x SQT values stay in registers between the two steps (a block holds
// 20 x nxl or 20 x nz values: windows up to 2457 cells wide and deep); a
// larger one is read twice.  (Against a max-abs launch with an
// atomic per workgroup followed by a quantising launch: 20.2 against 22.5 us
// per step on 16 precision slots, one launch fewer.)
constexpr int SQT = 1024, SQV = 48;
__global__ __launch_bounds__(SQT) void k_save16(const ShotDev* __restrict__ shots, const int* __restrict__ slots,
                                                 int nslots, int step)
{
  __shared__ float red[SQT / 64];
  const int k = blockIdx.x >> 2, b = blockIdx.x & 3;
  if (k >= nslots) return;
  const ShotDev s = shots[slots[k]];
  const int it = save_it(s, step);
  const StripBlock blk = strip_block_of(s, b, (it + 1) & 1);
  const long long base = (long long)it * s.strip_len + blk.lo;
  const int tid = threadIdx.x;
  const bool held = blk.n <= SQV * SQT;               // block-uniform
  float v[SQV];
  float m = 0.0f;
  if (held) {
#pragma unroll
    for (int u = 0; u < SQV; u++) {
      const int e = tid + u * SQT;
      v[u] = e < blk.n ? strip_elem(s, blk, e) : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < SQV; u++) m = fmaxf(m, fabsf(v[u]));
  } else {
    for (int e = tid; e < blk.n; e += SQT) m = fmaxf(m, fabsf(strip_elem(s, blk, e)));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  m = red[0];
#pragma unroll
  for (int w = 1; w < SQT / 64; w++) m = fmaxf(m, red[w]);
  const float sc = m > 0.0f ? fmaxf(m * (1.0f / 32768.0f), 1.17549435e-38f) : 1.0f;
  if (tid == 0) s.sscale[(long long)it * 4 + b] = sc;
  auto put = [&](int e, float x) {
    const float q = x / sc;
    if (s.store_fmt == ST_F16) {
      ((__half*)s.store)[base + e] = __float2half_rn(q);
    } else {
      unsigned int w = __float_as_uint(q);             // bf16, round to nearest even
      w += 0x7FFFu + ((w >> 16) & 1u);
      ((unsigned short*)s.store)[base + e] = (unsigned short)(w >> 16);
    }
  };
  if (held) {
#pragma unroll
    for (int u = 0; u < SQV; u++) {
      const int e = tid + u * SQT;
      if (e < blk.n) put(e, v[u]);
    }
  } else {
    for (int e = tid; e < blk.n; e += SQT) put(e, strip_elem(s, blk, e));
  }
}

// 1/rho and rho v^2 of a model, rounded once (the factors the step kernels
// read): the resident model at its creation, a shot's smoothed window
// (pmlvelsmooth1d / velsmooth1d) whenever it changes
__global__ __launch_bounds__(NTHR) void k_model_prep(const float* __restrict__ rho, const float* __restrict__ vv,
                                                      float* __restrict__ ib, float* __restrict__ kk, long long n)
{
  for (long long e = (long long)blockIdx.x * NTHR + threadIdx.x; e < n; e += (long long)gridDim.x * NTHR) {
    const float rh = rho[e];
    ib[e] = 1.0f / rh;
    kk[e] = rh * vv[e];
  }
}

// ------------------------------------------------------------ snapshots --
// Decimated snapshots of the source wavefield.  The reverse pass runs the
// leapfrog backwards, velocities first: after reverse step k it holds the
// pressure of forward step k and the velocities of forward step k+1 (its
// strips are read the same way, rtm_model.cu:424-434).  Snapshot k (k =
// snap_every * (snap_k0 + q)) therefore keeps the interior of p as forward
// step k leaves it and of vx, vz as forward step k+1 leaves them, and the
// reverse pass writes them back after its step k, so that the steps below k
// start from the stored field.  Block (q, f), f = 0 p, 1 vx, 2 vz: nzi x nxl
// codes; 16-bit forms with one max-abs scale per block (the strips' rule:
// max(maxabs / 32768, FLT_MIN), 1 for an all-zero block).
__device__ __forceinline__ int snap_index(const ShotDev& s, int k)
{
  if (!s.snap || k < 0 || k % s.snap_every != 0) return -1;
  const int q = k / s.snap_every - s.snap_k0;
  return (q >= 0 && q < s.nsnap) ? q : -1;
}
// the snapshot a forward step `it` contributes field f to: p of step k = it,
// velocities of step it = k + 1
__device__ __forceinline__ int snap_of_fwd(const ShotDev& s, int it, int f)
{
  return snap_index(s, f == 0 ? it : it - 1);
}
__device__ __forceinline__ const float* snap_field(const ShotDev& s, int f, bool lev)
{
  return f == 0 ? (lev ? s.sp[1] : s.sp[0]) : (f == 1 ? (lev ? s.svx[1] : s.svx[0]) : (lev ? s.svz[1] : s.svz[0]));
}
__device__ __forceinline__ unsigned int* snap_acc(const ShotDev& s) { return (unsigned int*)(s.snap_scale + 3 * s.nsnap); }

// The snapshot kernels run after every step of a pass (a captured graph
// holds the same launches for every step) and return at once unless the step
// takes or restores a snapshot.  When it does, a block is the nz x nxl
// interior of one field: one wave per row, rows dealt over the waves of
// gridDim.y workgroups, a lane taking SU columns 64 apart per iteration
// (independent loads in flight, no index division).
constexpr int SU = 4, SWB = NTHR / 64;

// forward, pass 1 (16-bit forms): max-abs of the workgroup's share of block
// (q, f) after forward step `step` (its new level).  Grid (3 slots, GY).
__global__ __launch_bounds__(NTHR) void k_snap_max(const ShotDev* __restrict__ shots, const int* __restrict__ slots,
                                                    int step)
{
  __shared__ float red[NTHR / 64];
  const ShotDev& s = shot_of(shots, slots[blockIdx.x / 3]);
  const int it = s.it_off + s.it_dir * step, f = blockIdx.x % 3;
  const int q = snap_of_fwd(s, it, f);
  if (q < 0) return;
  const float* fld = snap_field(s, f, (it + 1) & 1);
  const int nxl = s.nxl, lane = threadIdx.x & 63;
  float m = 0.0f;
  for (int r = blockIdx.y * SWB + (threadIdx.x >> 6); r < s.nzi; r += gridDim.y * SWB) {
    const gfloat* row = GCF(fld) + (long long)(s.npml + r) * s.pitch + s.npml;
    for (int c0 = lane; c0 < nxl; c0 += SU * 64) {
      float a[SU];
#pragma unroll
      for (int u = 0; u < SU; u++) a[u] = c0 + u * 64 < nxl ? fabsf(row[c0 + u * 64]) : 0.0f;
#pragma unroll
      for (int u = 0; u < SU; u++) m = fmaxf(m, a[u]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NTHR / 64; w++) m = fmaxf(m, red[w]);
    atomicMax(snap_acc(s) + q * 3 + f, __float_as_uint(m));
  }
}

// forward, pass 2: block (q, f) stored as codes (fp32: a copy) with its scale
__global__ __launch_bounds__(NTHR) void k_snap_q(const ShotDev* __restrict__ shots, const int* __restrict__ slots,
                                                  int step)
{
  const ShotDev& s = shot_of(shots, slots[blockIdx.x / 3]);
  const int it = s.it_off + s.it_dir * step, f = blockIdx.x % 3;
  const int q = snap_of_fwd(s, it, f);
  if (q < 0) return;
  const float* fld = snap_field(s, f, (it + 1) & 1);
  const int nxl = s.nxl, n = s.nzi * nxl, lane = threadIdx.x & 63;
  const long long base = (long long)(q * 3 + f) * n;
  float sc = 1.0f;
  if (s.snap_fmt != ST_F32) {
    const float m = __uint_as_float(snap_acc(s)[q * 3 + f]);
    sc = m > 0.0f ? fmaxf(m * (1.0f / 32768.0f), 1.17549435e-38f) : 1.0f;
  }
  if (blockIdx.y == 0 && threadIdx.x == 0) s.snap_scale[q * 3 + f] = sc;
  for (int r = blockIdx.y * SWB + (threadIdx.x >> 6); r < s.nzi; r += gridDim.y * SWB) {
    const gfloat* row = GCF(fld) + (long long)(s.npml + r) * s.pitch + s.npml;
    const long long ob = base + (long long)r * nxl;
    for (int c0 = lane; c0 < nxl; c0 += SU * 64) {
      float v[SU];
#pragma unroll
      for (int u = 0; u < SU; u++) v[u] = c0 + u * 64 < nxl ? row[c0 + u * 64] : 0.0f;
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int c = c0 + u * 64;
        if (c >= nxl) continue;
        if (s.snap_fmt == ST_F32) {
          GF(s.snap)[ob + c] = v[u];
        } else if (s.snap_fmt == ST_F16) {
          ((__half*)s.snap)[ob + c] = __float2half_rn(v[u] / sc);
        } else {
          unsigned int w = __float_as_uint(v[u] / sc);   // bf16, round to nearest even
          w += 0x7FFFu + ((w >> 16) & 1u);
          ((unsigned short*)s.snap)[ob + c] = (unsigned short)(w >> 16);
        }
      }
    }
  }
}

// reverse: after reverse step `step` = k (its new level (k+1)&1), snapshot
// k replaces the reconstructed interior cells: p on the
// interior, vx on its columns npml..nxp-npml-2, vz on its rows
// npml..nzp-npml-2 (the cells the reconstruction computes; the strips come
// from the boundary store as on every step)
__global__ __launch_bounds__(NTHR) void k_snap_load(const ShotDev* __restrict__ shots, const int* __restrict__ slots,
                                                     int step)
{
  const ShotDev& s = shot_of(shots, slots[blockIdx.x / 3]);
  const int it = s.it_off + s.it_dir * step;
  const int q = snap_index(s, it);
  if (q < 0) return;
  const int f = blockIdx.x % 3;
  float* fld = (float*)snap_field(s, f, (it + 1) & 1);
  const int nxl = s.nxl, n = s.nzi * nxl, lane = threadIdx.x & 63;
  const int ncol = f == 1 ? nxl - 1 : nxl, nrow = f == 2 ? s.nzi - 1 : s.nzi;
  const long long base = (long long)(q * 3 + f) * n;
  const float sc = s.snap_scale[q * 3 + f];
  for (int r = blockIdx.y * SWB + (threadIdx.x >> 6); r < nrow; r += gridDim.y * SWB) {
    gfloat* row = GF(fld) + (long long)(s.npml + r) * s.pitch + s.npml;
    const long long ib = base + (long long)r * nxl;
    for (int c0 = lane; c0 < ncol; c0 += SU * 64) {
      float v[SU];
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int c = c0 + u * 64;
        v[u] = 0.0f;
        if (c >= ncol) continue;
        if (s.snap_fmt == ST_F32) v[u] = GCF(s.snap)[ib + c];
        else if (s.snap_fmt == ST_F16) v[u] = __half2float(((const __half*)s.snap)[ib + c]) * sc;
        else v[u] = __uint_as_float((unsigned int)((const unsigned short*)s.snap)[ib + c] << 16) * sc;
      }
#pragma unroll
      for (int u = 0; u < SU; u++)
        if (c0 + u * 64 < ncol) row[c0 + u * 64] = v[u];
    }
  }
}

// ------------------------------------------------------- model smoothing --
// velsmooth1d (Velocity_function.cpp:199-313) of an nz x nx window: every
// output cell is the sum over the (2 nsp + 1)^2 window around it of
// weight x value, in double, rows then columns as the reference loops, on
// the window extended by its edge values (a clamp), rounded to float and
// then divided (in double) by the sum of the weights.  The weights are
// tabulated on the host with the reference's expression (wt, row-major),
// so every term is the reference's.
__global__ __launch_bounds__(NTHR) void k_gauss_smooth(const float* __restrict__ in, float* __restrict__ out, int nz,
                                                        int nx, int nsp, const double* __restrict__ wt, double norm)
{
  const long long e = (long long)blockIdx.x * NTHR + threadIdx.x;
  if (e >= (long long)nz * nx) return;
  const int iz = (int)(e / nx), ix = (int)(e - (long long)iz * nx);
  const int k = 2 * nsp + 1;
  double acc = 0.0;
  for (int i = 0; i < k; i++) {
    const float* row = in + (size_t)imin(imax(iz + i - nsp, 0), nz - 1) * nx;
    const double* w = wt + (size_t)i * k;
    for (int j = 0; j < k; j++) acc += w[j] * (double)row[imin(imax(ix + j - nsp, 0), nx - 1)];
  }
  out[e] = (float)((double)(float)acc / norm);
}

// pmlvelsmooth1d (Velocity_function.cpp:9-103) of one nz x nx window per
// workgroup (wins[blockIdx.x]: the window w, in place, and a scratch window
// v of the same size).  The reference mirrors the interior into the
// absorbing layer, then relaxes the layer in place -- each cell becomes the
// average of its four neighbours, taking those already updated in the sweep
// as updated -- until a sweep changes no cell by more than 5e-4 (relative)
// or 1001 sweeps ran.  The in-place order is kept exactly:
//   * left and right layers (columns 1..npml-1, rows 1..nz-2 in order, the
//     reference's ix-outer / iz-inner loop): lane t of one wave owns column
//     t+1 and runs one row behind lane t-1, which hands it the new value of
//     its left neighbour by a lane shuffle; the cell above is the lane's own
//     previous result; the cells right and below are not updated yet in this
//     sweep and come from memory.  Left and right are independent (two waves).
//   * top and bottom layers (rows 1..npml-1, columns 1..nx-2) the same with
//     lanes over rows: the new value above from the lane before, the one on
//     the left the lane's own.
//   * the edge columns and rows: four independent recurrences, one lane
//     each, on copies staged in LDS; then the four corners.
// Each sweep ends with the largest relative change over the layer ring (the
// rest of the window never changes) and copies the ring back into w.  Sums
// are formed in the reference's operand order, as floats, scaled by the
// double literal 0.25.  Needs npml - 1 <= 64, nx, nz >= 3 npml (the
// mirrors read only interior cells) and nx, nz <= RTM_SMOOTH_MAXN.
#define RTM_SMOOTH_MAXN 4096
struct SmoothWin { float* w; float* v; int nx, nz, npml, _pad; };

__device__ __forceinline__ float avg4(float a, float b, float c, float d)
{
  return (float)(0.25 * (double)(a + b + c + d));
}

// one side of the layer, one wave: lanes over `lanes` lines starting at
// line0, positions 1..len along each line.  LR: lines are columns and
// positions rows (sum E + W + S + N with W from the lane before, N the
// lane's own); otherwise lines are rows and positions columns (W the lane's
// own, N from the lane before).
template <bool LR>
__device__ void smooth_side(float* __restrict__ v, int nx, int line0, int lanes, int len)
{
  const int t = threadIdx.x & 63;
  float prev = 0.0f;                                 // the lane's new value at the previous position
  for (int k = 0; k < len + lanes - 1; k++) {
    const float before = __shfl_up(prev, 1, 64);     // lane t-1's new value at this lane's position
    const int j = 1 + k - t;
    if (t < lanes && j >= 1 && j <= len) {
      const int ln = line0 + t;
      const size_t c = LR ? (size_t)j * nx + ln : (size_t)ln * nx + j;
      const float E = v[c + 1], S = v[c + nx];
      float W, N;
      if (LR) { W = t > 0 ? before : v[c - 1]; N = j > 1 ? prev : v[c - nx]; }
      else { W = j > 1 ? prev : v[c - 1]; N = t > 0 ? before : v[c - nx]; }
      prev = avg4(E, W, S, N);
      v[c] = prev;
    }
  }
}

__global__ __launch_bounds__(NTHR) void k_pml_smooth(const SmoothWin* __restrict__ wins)
{
  __shared__ float ed[8 * RTM_SMOOTH_MAXN];          // col 0, 1, nx-1, nx-2 (nz each), row 0, 1, nz-1, nz-2 (nx each)
  __shared__ float red[NTHR / 64];
  __shared__ float err_s;
  const SmoothWin sw = wins[blockIdx.x];
  float* __restrict__ w = sw.w;
  float* __restrict__ v = sw.v;
  const int nx = sw.nx, nz = sw.nz, npml = sw.npml, tid = threadIdx.x, wave = tid >> 6;
  // mirror the interior into the layer: columns (rows 1..nz-2), then rows
  // (columns 1..nx-2, which include the mirrored columns)
  for (int e = tid; e < npml * (nz - 2); e += NTHR) {
    const int ix = e % npml, iz = 1 + e / npml;
    w[(size_t)iz * nx + ix] = w[(size_t)iz * nx + 2 * npml - ix - 1];
    w[(size_t)iz * nx + nx - npml + ix] = w[(size_t)iz * nx + nx - npml - ix - 1];
  }
  __syncthreads();
  for (int e = tid; e < npml * (nx - 2); e += NTHR) {
    const int iz = e / (nx - 2), ix = 1 + e % (nx - 2);
    w[(size_t)iz * nx + ix] = w[(size_t)(2 * npml - iz - 1) * nx + ix];
    w[(size_t)(nz - npml + iz) * nx + ix] = w[(size_t)(nz - npml - iz - 1) * nx + ix];
  }
  __syncthreads();
  for (long long e = tid; e < (long long)nx * nz; e += NTHR) v[e] = w[e];
  __syncthreads();
  float* c0 = ed;
  float* c1 = ed + nz;
  float* cN = ed + 2 * nz;
  float* cM = ed + 3 * nz;
  float* r0 = ed + 4 * nz;
  float* r1 = r0 + nx;
  float* rN = r0 + 2 * nx;
  float* rM = r0 + 3 * nx;
  const int ring = 2 * npml * nx + 2 * npml * (nz - 2 * npml);
  float err = 1.0f;
  for (int num = 0; num <= 1000 && (double)err > 0.0005; num++) {
    // left and right layers
    if (wave == 0) smooth_side<true>(v, nx, 1, npml - 1, nz - 2);
    else if (wave == 1) smooth_side<true>(v, nx, nx - npml, npml - 1, nz - 2);
    __syncthreads();
    // top and bottom layers
    if (wave == 0) smooth_side<false>(v, nx, 1, npml - 1, nx - 2);
    else if (wave == 1) smooth_side<false>(v, nx, nz - npml, npml - 1, nx - 2);
    __syncthreads();
    // edges: stage, four recurrences, corners, write back
    for (int i = tid; i < nz; i += NTHR) {
      c0[i] = v[(size_t)i * nx]; c1[i] = v[(size_t)i * nx + 1];
      cN[i] = v[(size_t)i * nx + nx - 1]; cM[i] = v[(size_t)i * nx + nx - 2];
    }
    for (int i = tid; i < nx; i += NTHR) {
      r0[i] = v[i]; r1[i] = v[nx + i];
      rN[i] = v[(size_t)(nz - 1) * nx + i]; rM[i] = v[(size_t)(nz - 2) * nx + i];
    }
    __syncthreads();
    if (tid == 0) {                                  // column 0: C + E + S + N
      float up = c0[0];
      for (int iz = 1; iz < nz - 1; iz++) { up = avg4(c0[iz], c1[iz], c0[iz + 1], up); c0[iz] = up; }
    } else if (tid == 64) {                          // column nx-1
      float up = cN[0];
      for (int iz = 1; iz < nz - 1; iz++) { up = avg4(cN[iz], cM[iz], cN[iz + 1], up); cN[iz] = up; }
    } else if (tid == 128) {                         // row 0: C + S + W + E
      float lf = r0[0];
      for (int ix = 1; ix < nx - 1; ix++) { lf = avg4(r0[ix], r1[ix], lf, r0[ix + 1]); r0[ix] = lf; }
    } else if (tid == 192) {                         // row nz-1
      float lf = rN[0];
      for (int ix = 1; ix < nx - 1; ix++) { lf = avg4(rN[ix], rM[ix], lf, rN[ix + 1]); rN[ix] = lf; }
    }
    __syncthreads();
    if (tid == 0) {                                  // corners (none reads another)
      const float a = avg4(c0[0], c0[1], r0[1], c1[1]);
      const float b = avg4(c0[nz - 1], c0[nz - 2], rN[1], c1[nz - 2]);
      const float c = avg4(cN[0], r0[nx - 2], cN[1], cM[1]);
      const float d = avg4(cN[nz - 1], cN[nz - 2], rN[nx - 2], cM[nz - 2]);
      c0[0] = a; r0[0] = a; c0[nz - 1] = b; rN[0] = b; cN[0] = c; r0[nx - 1] = c; cN[nz - 1] = d; rN[nx - 1] = d;
    }
    __syncthreads();
    for (int i = tid; i < nz; i += NTHR) { v[(size_t)i * nx] = c0[i]; v[(size_t)i * nx + nx - 1] = cN[i]; }
    for (int i = tid; i < nx; i += NTHR) { v[i] = r0[i]; v[(size_t)(nz - 1) * nx + i] = rN[i]; }
    __syncthreads();
    // largest relative change over the ring, then the ring back into w
    float m = -1.0e+5f;
    for (int e = tid; e < ring; e += NTHR) {
      int iz, ix;
      if (e < 2 * npml * nx) { const int r = e / nx; iz = r < npml ? r : nz - 2 * npml + r; ix = e - r * nx; }
      else { const int f = e - 2 * npml * nx, r = f / (2 * npml), c = f - r * 2 * npml;
             iz = npml + r; ix = c < npml ? c : nx - 2 * npml + c; }
      const size_t o = (size_t)iz * nx + ix;
      const float a = w[o], b = v[o];
      m = fmaxf(m, fabsf(a - b) / fabsf(a));
      w[o] = b;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((tid & 63) == 0) red[wave] = m;
    __syncthreads();
    if (tid == 0) {
      float x = red[0];
      for (int i = 1; i < NTHR / 64; i++) x = fmaxf(x, red[i]);
      err_s = x;
    }
    __syncthreads();
    err = err_s;
  }
}

// Phase start of one slot.  FWD: zero both levels of both wavefields and the
// PML pair.  REV: zero both levels of the source velocities (rtm_model.cu:
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
    if (g0 == 0) guard_word(s)[0] = 0u;
    if (s.snap)
      for (long long e = g0; e < 3LL * s.nsnap; e += gs) snap_acc(s)[e] = 0u;   // snapshot max-abs words
  } else {
    for (long long e = g0; e < ni; e += gs) { s.g1[e] = 0.f; s.illum[e] = 0.f; }
    if (s.ic)
      for (long long e = g0; e < ni * (5 + 2 * 61); e += gs) s.ic[e] = 0.f;
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

// ------------------------------------------------- other imaging conditions --
// After reverse step `step` (kernel.cu:1417-1529, one thread per interior
// cell of a slot): wavefield decomposition, Poynting-vector angle bins and the
// energy norm, from the new fields (level (it+1)&1) and the fields before the
// step (level it&1).  The reference's quirks are kept: vx and vxs are read at
// the p index (stride nxp in an array of rows nxp-1, kernel.cu:1431,1468) and
// the source Poynting vector takes the receiver's vz (kernel.cu:1469).  Each
// expression is evaluated as nvcc contracts it by default (a*b + c*d ->
// fma(a, b, c*d); accumulations -> fma), with the float / double mixing of
// the reference's double literals.  Off the hot path: launched only when the
// job asks for these images.
__global__ __launch_bounds__(NTHR) void k_ic(const ShotDev* __restrict__ shots, const int* __restrict__ slots, int step)
{
  const ShotDev& s = shot_of(shots, slots[blockIdx.y]);
  if (!s.ic) return;
  const int nxl = s.nxl, nn = s.nzi * nxl;
  const int e = blockIdx.x * NTHR + threadIdx.x;
  if (e >= nn) return;
  const int iz = e / nxl + s.npml, ix = e - (e / nxl) * nxl + s.npml;   // padded cell
  const int nxp = s.nxp, pitch = s.pitch;
  const int it = s.it_off + s.it_dir * step;
  const bool odd = it & 1;                              // the step read level it&1, wrote the other
  const float* psn = odd ? s.sp[0] : s.sp[1];
  const float* pso = odd ? s.sp[1] : s.sp[0];
  const float* prn = odd ? s.rp[0] : s.rp[1];
  const float* pro = odd ? s.rp[1] : s.rp[0];
  const float* vxs_f = odd ? s.svx[0] : s.svx[1];
  const float* vzs_f = odd ? s.svz[0] : s.svz[1];
  const float* vxr_f = odd ? s.rvx[0] : s.rvx[1];
  const float* vzr_f = odd ? s.rvz[0] : s.rvz[1];
  const long long id = (long long)iz * pitch + ix;
  // the p index read in an array of rows nxp-1: row q/(nxp-1), column
  // q%(nxp-1); past that array's end (windows taller than about npml*nxp, where
  // the reference reads out of bounds) the value is taken as 0
  const int q = iz * nxp + ix, qr = q / (nxp - 1), qc = q - qr * (nxp - 1);
  const bool qin = qr < s.nzp;
  const long long iq = qin ? (long long)qr * pitch + qc : 0;
  const float ps = psn[id], p = prn[id];
  const float vxs = qin ? vxs_f[iq] : 0.0f, vzs = vzs_f[id], vxr = qin ? vxr_f[iq] : 0.0f, vzr = vzr_f[id];
  const int mr = imin(imax(iz - s.mz0, 0), s.mzn - 1), mc = s.mx0 + imin(imax(ix - s.mxs, 0), s.mxn - 1);
  const float vv = s.vv[(size_t)mr * s.mpitch + mc];
  float* ic = s.ic;
  // cuda_wavefield_decomposition (kernel.cu:1417-1450)
  const float vtxs = -ps * vxs, vtzs = -ps * vzs, vtxr = -p * vxr, vtzr = -p * vzr;
  const float psu = vtzs < 0 ? ps : 0.0f, psd = vtzs > 0 ? ps : 0.0f, psl = vtxs < 0 ? ps : 0.0f, psr = vtxs > 0 ? ps : 0.0f;
  const float pu = vtzr < 0 ? p : 0.0f, pd = vtzr > 0 ? p : 0.0f, pl = vtxr < 0 ? p : 0.0f, pr = vtxr > 0 ? p : 0.0f;
  // An accumulator plane is read and written only where neither factor of its
  // product is zero.  There the reference's fma(a, b, c) adds +-0 and returns
  // c unchanged, except that it can turn a -0 into +0.  A plane holds -0 only
  // after an fma whose exact product was a negative value too small for a
  // float, so the skipped form can differ from the reference's only in the
  // sign of a zero.  The stacks sum the planes in FP64 from +0
  // (+0 + -0 = +0), so the images are the same bit for bit.  The up/down
  // (left/right) conditions exclude each other, and a wave travelling one way
  // keeps them constant over a region, so whole cache lines are skipped.
  if (psu != 0.0f && pd != 0.0f) ic[e] = __builtin_fmaf(psu, pd, ic[e]);
  if (psd != 0.0f && pu != 0.0f) ic[nn + e] = __builtin_fmaf(psd, pu, ic[nn + e]);
  if (psl != 0.0f && pr != 0.0f) ic[2 * nn + e] = __builtin_fmaf(psl, pr, ic[2 * nn + e]);
  if (psr != 0.0f && pl != 0.0f) ic[3 * nn + e] = __builtin_fmaf(psr, pl, ic[3 * nn + e]);
  // cuda_poynting (kernel.cu:1452-1488).  Both bins add a multiple of ps*p:
  // where that product is zero (either wavefield zero, or the product
  // underflows) they add +-0, so the bins keep their value up to the sign of
  // a zero, as above.  The angle, its cosine and sine and the FP64 quotient
  // are then not needed.
  const float ksx = ps * vxs, ksz = ps * vzr, krx = p * vxr, krz = p * vzr;
  if (ps * p != 0.0f) {
    const float qd = sqrtf(__builtin_fmaf(ksx, ksx, ksz * ksz) * __builtin_fmaf(krx, krx, krz * krz));
    float w = (float)((double)__builtin_fmaf(ksx, krx, ksz * krz) / (qd > 0 ? (double)qd : 1.0e-6));
    if (w >= -1.0 && w <= 1.0) {
      const double pi = 3.1415926535897932;
      w = (float)(0.5 * (double)acosf(w));
      w = (float)(pi / 2.0 - (double)w);
      if (w > 0 && w <= pi / 3.0) {
        const float dw = (float)(pi / 180);
        const size_t ib = (size_t)(int)(w / dw) * nn + e;
        float* g31 = ic + 5 * (size_t)nn;
        float* g32 = ic + 66 * (size_t)nn;
        g31[ib] = __builtin_fmaf(ps * p, cosf(w), g31[ib]);
        g32[ib] = (float)((double)g32[ib] + (double)(ps * p) * 16.0 * pi * (double)sqrtf(vv) / (double)sinf(w));
      }
    }
  }
  // cuda_energynorm (kernel.cu:1489-1529)
  const float psx = (float)(0.5 * (double)(psn[id + 1] - psn[id - 1]) / (double)s.dx);
  const float psz = (float)(0.5 * (double)(psn[id + pitch] - psn[id - pitch]) / (double)s.dz);
  const float px = (float)(0.5 * (double)(prn[id + 1] - prn[id - 1]) / (double)s.dx);
  const float pz = (float)(0.5 * (double)(prn[id + pitch] - prn[id - pitch]) / (double)s.dz);
  const float t = -(ps - pso[id]) * (p - pro[id]) / (s.dt * s.dt * vv);
  const float en = __builtin_fmaf(psz, pz, __builtin_fmaf(psx, px, t));
  // a sum of floats is +0 on exact cancellation and never underflows to 0,
  // so this plane never holds -0 and adding +-0 leaves it bit for bit
  if (en != 0.0f) ic[4 * nn + e] = ic[4 * nn + e] + en;
}

// one shot's other images into the FP64 stacks (kernel.cu:1530-1579): g2 as
// the four decomposed products summed, g3 / g3_true as the 61 bins summed,
// each tapered, then normalised by the tapered illumination (image3s..5s)
// or not (image3m..5m, imagetrue)
struct StackIc { const float* ic; const float* illum; const float* taper; int nx1, nxl; };
__global__ __launch_bounds__(NTHR) void k_stack_ic(const StackIc* __restrict__ sh, int nx, int nz, double* __restrict__ i7)
{
  const long long e = (long long)blockIdx.x * NTHR + threadIdx.x;
  if (e >= (long long)nx * nz) return;
  const int iz = (int)(e / nx), ix = (int)(e - (long long)iz * nx);
  const StackIc q = *sh;
  const int c = ix - q.nx1;
  if (c < 0 || c >= q.nxl) return;
  const size_t nn = (size_t)nz * q.nxl, ip = (size_t)iz * q.nxl + c;
  const float* ic = q.ic;
  const float t = q.taper[c];
  float g3 = 0.0f, g3t = 0.0f;
  for (int iw = 0; iw <= 60; iw++) { g3 += ic[(5 + iw) * nn + ip]; g3t += ic[(66 + iw) * nn + ip]; }
  const float g2 = ((ic[ip] + ic[nn + ip]) + ic[2 * nn + ip]) + ic[3 * nn + ip];
  const float il = q.illum[ip] * t;
  const float g[3] = {g2 * t, g3 * t, ic[4 * nn + ip] * t};
  const size_t n = (size_t)nx * nz;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    i7[k * n + e] += (double)g[k] / ((double)il + 1.0e-6);
    i7[(3 + k) * n + e] += (double)g[k];
  }
  i7[6 * n + e] += (double)(g3t * t);
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

// Wavefield guard: the device form of the reference's max|p| print every
// ntsnap steps (rtm_model.cu:271-273, 324, 415).  The largest |p| of each
// listed slot's pressure as written by step `step` (it = it_off + it_dir *
// step wrote level (it+1)&1; the source field, and in the reverse phase also
// the receiver field) goes into the slot's guard word and into *flag as float
// bits: non-negative floats order as their bits, and an infinity (0x7F800000)
// or a NaN (0x7FC00000 and up) compares above every finite value, so the host
// detects a blow-up (unstable dt, bad model) from one word.  Grid: (slots, GW).
__global__ __launch_bounds__(NTHR) void k_guard(const ShotDev* __restrict__ shots, const int* __restrict__ slots,
                                                 int step, unsigned int* __restrict__ flag)
{
  __shared__ unsigned int red[NTHR / 64];
  const ShotDev& s = shot_of(shots, slots[blockIdx.x]);
  const int it = s.it_off + s.it_dir * step;
  const bool lev = (it + 1) & 1;
  const float* ps = lev ? s.sp[1] : s.sp[0];
  const float* pr = lev ? s.rp[1] : s.rp[0];
  const bool rev = s.phase == PH_REV;
  // one wave per row, rows dealt over the waves of gridDim.y workgroups; a
  // lane takes GU columns 64 apart per iteration (independent loads in
  // flight, no index division, lanes busy on windows of any width)
  constexpr int GU = 4, WPB = NTHR / 64;
  const int lane = threadIdx.x & 63, nxp = s.nxp;
  unsigned int m = 0u;
  for (int iz = blockIdx.y * WPB + (threadIdx.x >> 6); iz < s.nzp; iz += gridDim.y * WPB) {
    const gfloat* rs = GCF(ps) + (long long)iz * s.pitch;
    const gfloat* rr = GCF(pr) + (long long)iz * s.pitch;
    for (int x0 = lane; x0 < nxp; x0 += GU * 64) {
      unsigned int a[GU];
#pragma unroll
      for (int u = 0; u < GU; u++) {
        const int ix = x0 + u * 64;
        a[u] = ix < nxp ? __float_as_uint(rs[ix]) & 0x7FFFFFFFu : 0u;
        if (rev && ix < nxp) a[u] = max(a[u], __float_as_uint(rr[ix]) & 0x7FFFFFFFu);
      }
#pragma unroll
      for (int u = 0; u < GU; u++) m = max(m, a[u]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NTHR / 64; w++) m = max(m, red[w]);
    atomicMax(guard_word(s), m);
    if (flag) atomicMax(flag, m);
  }
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
static hipError_t launch_step_t(const LaunchArgs& a, int phase, hipStream_t st)
{
  const dim3 g(grid8(a.ntiles)), b(64 * G);
  if (phase == PH_FWD) hipLaunchKernelGGL((k_fwd<G, RPT>), g, b, 0, st, a);
  else hipLaunchKernelGGL((k_rev<G, RPT>), g, b, 0, st, a);
  return hipGetLastError();
}

hipError_t rtm_launch_step(const LaunchArgs& a, int phase, int tz, hipStream_t st)
{
  if (a.ntiles <= 0) return hipSuccess;
  switch (tz) {
    case 16: return launch_step_t<4, 4>(a, phase, st);
    // 32 rows: the forward step as 8 waves x 4 rows (76 registers, 6 waves
    // per SIMD), the reverse step as 4 waves x 8 rows (its 8 x 4 form needs
    // 91 registers and loses 15 %)
    case 32: return phase == PH_FWD ? launch_step_t<8, 4>(a, phase, st) : launch_step_t<4, 8>(a, phase, st);
    case 64: return launch_step_t<8, 8>(a, phase, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t rtm_launch_save(const ShotDev* shots, const int* slots, int nslots, int step, int store_fmt,
                           hipStream_t st)
{
  if (nslots <= 0) return hipSuccess;
  // blocks of up to 2*10*nz values: 4 workgroups each (fp32); the 16-bit
  // passes use 8 (they run on the few-shot launches of the served form too)
  if (store_fmt == ST_F32) {
    hipLaunchKernelGGL(k_save, dim3(4 * nslots, 4), dim3(NTHR), 0, st, shots, slots, nslots, step);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_save16, dim3(4 * nslots), dim3(SQT), 0, st, shots, slots, nslots, step);
  return hipGetLastError();
}

hipError_t rtm_launch_phase_init(const ShotDev* shots, int slot, int phase, int nblocks, hipStream_t st)
{
  hipLaunchKernelGGL(k_phase_init, dim3(nblocks), dim3(NTHR), 0, st, shots, slot, phase);
  return hipGetLastError();
}

hipError_t rtm_launch_model_prep(const float* rho, const float* vv, float* ib, float* kk, long long n, hipStream_t st)
{
  const long long nb = (n + NTHR - 1) / NTHR;
  hipLaunchKernelGGL(k_model_prep, dim3((unsigned)(nb < 4096 ? nb : 4096)), dim3(NTHR), 0, st, rho, vv, ib, kk, n);
  return hipGetLastError();
}

hipError_t rtm_launch_seed(const ShotDev* shots, int slot, int nrec, hipStream_t st)
{
  if (nrec <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_seed_receivers, dim3((nrec + NTHR - 1) / NTHR), dim3(NTHR), 0, st, shots, slot);
  return hipGetLastError();
}

hipError_t rtm_launch_ic(const ShotDev* shots, const int* slots, int nslots, int step, int max_cells, hipStream_t st)
{
  if (nslots <= 0 || max_cells <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ic, dim3((unsigned)((max_cells + NTHR - 1) / NTHR), (unsigned)nslots), dim3(NTHR), 0, st,
                     shots, slots, step);
  return hipGetLastError();
}

hipError_t rtm_launch_stack_ic(const void* sh, int nx, int nz, double* i7, hipStream_t st)
{
  const long long n = (long long)nx * nz;
  hipLaunchKernelGGL(k_stack_ic, dim3((unsigned)((n + NTHR - 1) / NTHR)), dim3(NTHR), 0, st, (const StackIc*)sh, nx, nz, i7);
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

hipError_t rtm_launch_snap(const ShotDev* shots, const int* slots, int nslots, int phase, int step, int snap_fmt,
                           hipStream_t st)
{
  if (nslots <= 0) return hipSuccess;
  // about 2048 workgroups when a snapshot is taken (each returns at once on
  // the other steps): 42 per block for 16 slots, 8 for 200
  const int gy = std::min(64, std::max(8, 2048 / (3 * nslots)));
  const dim3 g(3 * nslots, gy), b(NTHR);
  if (phase == PH_FWD) {
    if (snap_fmt != ST_F32) hipLaunchKernelGGL(k_snap_max, g, b, 0, st, shots, slots, step);
    hipLaunchKernelGGL(k_snap_q, g, b, 0, st, shots, slots, step);
  } else {
    hipLaunchKernelGGL(k_snap_load, g, b, 0, st, shots, slots, step);
  }
  return hipGetLastError();
}

hipError_t rtm_launch_gauss_smooth(const float* in, float* out, int nz, int nx, int nsp, const double* wt, double norm,
                                   hipStream_t st)
{
  const long long n = (long long)nz * nx;
  hipLaunchKernelGGL(k_gauss_smooth, dim3((unsigned)((n + NTHR - 1) / NTHR)), dim3(NTHR), 0, st, in, out, nz, nx, nsp,
                     wt, norm);
  return hipGetLastError();
}

hipError_t rtm_launch_pml_smooth(const void* wins, int n, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pml_smooth, dim3(n), dim3(NTHR), 0, st, (const SmoothWin*)wins);
  return hipGetLastError();
}

hipError_t rtm_launch_guard(const ShotDev* shots, const int* slots, int nslots, int step, unsigned int* flag,
                            hipStream_t st)
{
  if (nslots <= 0) return hipSuccess;
  // about 2048 workgroups in all: 10 per slot for 200 survey slots, 64 for
  // a few wide ones
  const int gy = nslots >= 256 ? 8 : std::min(64, std::max(8, 2048 / nslots));
  hipLaunchKernelGGL(k_guard, dim3(nslots, gy), dim3(NTHR), 0, st, shots, slots, step, flag);
  return hipGetLastError();
}

hipError_t rtm_launch_bump(ShotDev* shots, const int* slots, int nslots, int k, hipStream_t st)
{
  hipLaunchKernelGGL(k_bump, dim3((nslots + 63) / 64), dim3(64), 0, st, shots, slots, nslots, k);
  return hipGetLastError();
}
