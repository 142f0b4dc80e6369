// rtm_engine.cpp -- host runtime of the RTM path: shot set-up, slot memory,
// batched forward / reverse passes (plain launches or replayed hipGraph
// chunks), FP64 stacking, outputs and the reference's file side effects.
#include "rtm_engine.h"

#include <math.h>
#include <cmath>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <thread>
#include <tuple>

#include "host_io.h"
#include "rtm_launch.h"

namespace rtm {

static thread_local std::string g_err;
void set_error(const std::string& s) { g_err = s; }
const char* get_error() { return g_err.c_str(); }

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                           \
      return -10;                                                                          \
    }                                                                                      \
  } while (0)

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// ------------------------------------------------------------ geometry ----
bool shot_geometry(const rtmh_params& P, int ishot, ShotSetup& s, const float* field)
{
  const float dx = P.dx, dz = P.dz;
  const int nx = P.nx, npml = P.npml;
  int nx1, nx2, is = 0, noffset = 0;
  float spx, spz;
  int layout = P.flag_layout;
  if (field) {
    // a recorded gather (rtm_real.cu:121-136): source at sx, receivers from
    // offset to offsetmax on its right (the receiver rows of layout 3); the
    // reference does not clamp the window to the model, so a gather reaching
    // past it is refused here
    const float sx = field[0], offset = field[1], offsetmax = field[2];
    if (sx < 0 || offset < 0 || offsetmax < offset) return false;
    noffset = (int)(offset / dx);
    nx1 = (int)(sx / dx);
    nx2 = (int)((sx + offsetmax) / dx);
    if (nx2 > nx - 1) return false;
    spx = sx - nx1 * dx;
    layout = 3;
  } else if (P.flag_layout == 1) {                                  // rtm_model.cu:124-134
    const int NX = (int)(P.offsetmax / dx + 1);
    const int sp1 = (int)(P.spx0 / dx), np = (int)(P.dspx / dx), nr = (NX - 1) / 2;
    is = sp1 + (ishot - 1) * np;
    nx1 = imax(0, is - nr);
    nx2 = imin(nx - 1, is + nr);
    spx = (is - nx1) * dx;
  } else if (P.flag_layout == 2) {                           // rtm_model.cu:627-630
    nx1 = 0; nx2 = nx - 1;
    spx = P.spx0 + (ishot - 1) * P.dspx;
  } else if (P.flag_layout == 3 || P.flag_layout == 4) {     // rtm_model.cu:1126-1138, 1633-1645
    const int NX = (int)(P.offsetmax / dx + 1);
    const int sp1 = (int)(P.spx0 / dx), np = (int)(P.dspx / dx);
    is = sp1 + (ishot - 1) * np;
    noffset = (int)(P.offset / dx);
    const int nr = NX - 1;
    if (P.flag_layout == 3) { nx1 = imax(0, is); nx2 = imin(nx - 1, is + nr); }
    else { nx1 = imax(0, is - nr); nx2 = imin(nx - 1, is); }
    spx = (is - nx1) * dx;
  } else {
    return false;
  }
  const int nxl = nx2 - nx1 + 1;
  if (nxl <= 0) return false;
  spz = P.spz0;
  rtmh_geom& g = s.g;
  memset(&g, 0, sizeof(g));
  g.ishot = ishot; g.nx1 = nx1; g.nx2 = nx2; g.nxl = nxl; g.noffset = noffset;
  s.hk_nt = (short)(unsigned short)(P.nt);
  s.hk_dt = (short)(unsigned short)(P.dt * 1000000);
  s.hk_dx = dx;
  if (layout <= 2) {
    s.hk_itraceall = (ishot - 1) * nxl + 1;
    s.hk_offset = (int)(nx1 * dx - spx);
    s.hk_sx = (int)spx;
    s.hk_gx = (int)(nx1 * dx);
    s.hk_ntrace = nxl;
  } else {
    s.hk_itraceall = (ishot - 1) * (nxl - noffset) + 1;
    s.hk_offset = (int)P.offset;
    s.hk_sx = (int)spx;
    s.hk_gx = (int)(P.offset + nx1 * dx);
    s.hk_ntrace = nxl - noffset;
  }
  spx = spx + npml * dx;
  spz = spz + npml * dz;
  g.nxp = nxl + 2 * npml;
  g.nzp = P.nz + 2 * npml;
  g.nsx = (int)(spx / dx);
  g.nsz = (int)(spz / dz);
  if (layout <= 2) { g.nrec = nxl; g.rcol0 = npml; }
  else if (layout == 3) { g.nrec = nxl - noffset; g.rcol0 = npml + noffset; }
  else { g.nrec = nxl - noffset; g.rcol0 = npml; }
  if (g.nrec < 0) g.nrec = 0;
  return true;
}

// padded-window value of a model at padded (iz, ix): the edge replication of
// extendvel1 (Velocity_function.cpp:2011-2030) as a clamp into the window
static inline float win_at(const float* m, int nx, int nz, int nx1, int nxl, int npml, int iz, int ix)
{
  const int r = imin(imax(iz - npml, 0), nz - 1);
  const int c = nx1 + imin(imax(ix - npml, 0), nxl - 1);
  return m[(size_t)r * nx + c];
}

void ricker(std::vector<float>& w, int nt, float fdom, float dt)
{
  const double pi = 3.1415926535897932;                       // common.h:11
  w.resize(nt);
  for (int it = 0; it < nt; it++) {
    float t = (float)(pi * (double)fdom * fabs((double)(it * dt) - 1.0 / (double)fdom));
    t *= t;
    w[it] = (float)((1.0 - 2.0 * (double)t) * (double)expf(-t));
  }
}

void taper_row(std::vector<float>& t, int nxl, int nwin)
{
  t.resize(nxl);
  const float alpha = 0.06f;                                  // rtm_model.cu:71
  for (int ix = 0; ix < nxl; ix++) {
    float a = -alpha * alpha, v = 1.0f;
    if (ix < nwin) { const float d = (float)(nwin - ix); a = a * d; a = a * d; v = expf(a); }
    else if (ix >= nxl - nwin) { const float d = (float)(ix - nxl + nwin + 1); a = a * d; a = a * d; v = expf(a); }
    t[ix] = v;
  }
}

// PML profile (kernel.cu:174-269) and the damping factors of kernel.cu:592-593,
// c1 = 1 - dt*d/2 and c2 = 2 - c1, kept as a = c1/c2 and b = 1/c2 (each
// rounded once): the kernels evaluate (c1*v - u)/c2 as fma(a, v, -(u*b)),
// which off the layer (a = b = 1) is v - u
static void damping(std::vector<float>& a, std::vector<float>& b, int n, int npml, float vpmax, float h,
                    bool half, float dt)
{
  const float R = 1.0e-10f;
  const float d = npml * h;
  const float d0 = (float)(1.5 * (double)vpmax * (double)logf(1 / R) / (double)d);
  a.resize(n); b.resize(n);
  for (int i = 0; i < n; i++) {
    float dd = 0.0f;
    if (i < npml) {
      float t = half ? (float)((double)(float)(npml - i) - 0.5) : (float)(npml - i);
      t = t / (float)npml; t = t * t; dd = d0 * t;
    } else if (i >= n - npml) {
      float t = half ? (float)((double)(float)(i - n + npml + 1) - 0.5) : (float)(i - n + npml + 1);
      t = t / (float)npml; t = t * t; dd = d0 * t;
    }
    const float c1 = (float)(1.0 - 0.5 * (double)dt * (double)dd);
    const float c2 = (float)(2.0 - (double)c1);
    a[i] = c1 / c2;
    b[i] = 1.0f / c2;
  }
}

// the velsmooth1d weights (Velocity_function.cpp:277-303): exp of minus the
// squared offset over 2 (nsp/3)^2, row-major over the (2 nsp + 1)^2 window,
// and their sum in the reference's order
void gauss_weights(int nsp, std::vector<double>& wt, double& norm)
{
  const int k = 2 * nsp + 1;
  const double den = 2.0 * nsp / 3.0 * nsp / 3.0;
  wt.resize((size_t)k * k);
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) {
      const double d1 = i - nsp, d2 = j - nsp;
      wt[(size_t)i * k + j] = exp(-(d1 * d1 + d2 * d2) / den);
    }
  norm = 0;
  for (double x : wt) norm += x;
}

void shot_windows(const rtmh_params& P, const float* vp0, const float* rho0, ShotSetup& s, bool smooth_here)
{
  const rtmh_geom& g = s.g;
  const int nxp = g.nxp, nzp = g.nzp;
  s.wvv.resize((size_t)nxp * nzp);
  s.wrho.resize((size_t)nxp * nzp);
  for (int iz = 0; iz < nzp; iz++)
    for (int ix = 0; ix < nxp; ix++) {
      const float v = win_at(vp0, P.nx, P.nz, g.nx1, g.nxl, P.npml, iz, ix);
      s.wvv[(size_t)iz * nxp + ix] = v * v;                    // v^2 (Velocity_function.cpp:2059)
      s.wrho[(size_t)iz * nxp + ix] = win_at(rho0, P.nx, P.nz, g.nx1, g.nxl, P.npml, iz, ix);
    }
  if (P.medium_flag && smooth_here) {                          // rtm_model.cu:204-206
    rtmh_pml_smooth(s.wvv.data(), nxp, nzp, P.npml);
    rtmh_pml_smooth(s.wrho.data(), nxp, nzp, P.npml);
  }
}

static void parallel_for(int n, const std::function<void(int)>& fn);

bool pml_smooth_on_gpu(int nx, int nz, int npml)
{
  return npml >= 2 && npml - 1 <= 64 && nx >= 3 * npml && nz >= 3 * npml && nx <= 4096 && nz <= 4096;
}

// pmlvelsmooth1d of host windows on the GPU (k_pml_smooth, one workgroup per
// window), in groups of at most ~1 GiB of windows and scratch
int Engine::pml_smooth_windows(std::vector<float*>& wins, const std::vector<int>& nx, const std::vector<int>& nz,
                               int npml)
{
  HIPCHK(hipSetDevice(device_));
  const size_t n = wins.size();
  size_t k0 = 0;
  while (k0 < n) {
    size_t k1 = k0, bytes = 0;
    while (k1 < n && (k1 == k0 || bytes + 2 * align256((size_t)nx[k1] * nz[k1] * 4) <= ((size_t)1 << 30))) {
      bytes += 2 * align256((size_t)nx[k1] * nz[k1] * 4);
      k1++;
    }
    const size_t m = k1 - k0;
    char* buf = (char*)pool_.get(bytes + align256(m * sizeof(SmoothWinHost)));
    if (!buf) { set_error("out of device memory (model smoothing)"); return -3; }
    std::vector<SmoothWinHost> d(m);
    size_t o = 0;
    for (size_t i = 0; i < m; i++) {
      const size_t wb = (size_t)nx[k0 + i] * nz[k0 + i] * 4;
      d[i].w = (float*)(buf + o); o += align256(wb);
      d[i].v = (float*)(buf + o); o += align256(wb);
      d[i].nx = nx[k0 + i]; d[i].nz = nz[k0 + i]; d[i].npml = npml; d[i]._pad = 0;
      HIPCHK(hipMemcpyAsync(d[i].w, wins[k0 + i], wb, hipMemcpyHostToDevice, st_));
    }
    HIPCHK(hipMemcpyAsync(buf + o, d.data(), m * sizeof(SmoothWinHost), hipMemcpyHostToDevice, st_));
    HIPCHK(rtm_launch_pml_smooth(buf + o, (int)m, st_));
    for (size_t i = 0; i < m; i++)
      HIPCHK(hipMemcpyAsync(wins[k0 + i], d[i].w, (size_t)nx[k0 + i] * nz[k0 + i] * 4, hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    pool_.put(buf);
    k0 = k1;
  }
  return 0;
}

// velsmooth1d of host windows on the GPU (k_gauss_smooth), one at a time
int Engine::gauss_smooth_windows(std::vector<float*>& wins, const std::vector<int>& nx, const std::vector<int>& nz,
                                 int nsp)
{
  if (nsp <= 0 || wins.empty()) return 0;
  HIPCHK(hipSetDevice(device_));
  std::vector<double> wt;
  double norm = 0;
  gauss_weights(nsp, wt, norm);
  size_t big = 0;
  for (size_t k = 0; k < wins.size(); k++) big = std::max(big, (size_t)nx[k] * nz[k] * 4);
  char* buf = (char*)pool_.get(2 * align256(big) + align256(wt.size() * 8));
  if (!buf) { set_error("out of device memory (model smoothing)"); return -3; }
  float* a = (float*)buf;
  float* b = (float*)(buf + align256(big));
  double* dwt = (double*)(buf + 2 * align256(big));
  HIPCHK(hipMemcpyAsync(dwt, wt.data(), wt.size() * 8, hipMemcpyHostToDevice, st_));
  for (size_t k = 0; k < wins.size(); k++) {
    const size_t wb = (size_t)nx[k] * nz[k] * 4;
    HIPCHK(hipMemcpyAsync(a, wins[k], wb, hipMemcpyHostToDevice, st_));
    HIPCHK(rtm_launch_gauss_smooth(a, b, nz[k], nx[k], nsp, dwt, norm, st_));
    HIPCHK(hipMemcpyAsync(wins[k], b, wb, hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
  }
  pool_.put(buf);
  return 0;
}

// the padded windows of every shot of a job, with medium_flag their layers
// smoothed: on the GPU where the window allows, on the host otherwise
int Engine::job_windows(const rtmh_params* pk, const float* vp0, const float* rho0, std::vector<ShotSetup>& setup)
{
  const int nsh = (int)setup.size();
  std::vector<char> gpu(nsh, 0);
  for (int k = 0; k < nsh; k++)
    gpu[k] = pk[k].medium_flag && pml_smooth_on_gpu(setup[k].g.nxp, setup[k].g.nzp, pk[k].npml);
  parallel_for(nsh, [&](int k) { shot_windows(pk[k], vp0, rho0, setup[k], !gpu[k]); });
  std::vector<float*> w;
  std::vector<int> nx, nz;
  for (int k = 0; k < nsh; k++)
    if (gpu[k]) {
      w.push_back(setup[k].wvv.data()); w.push_back(setup[k].wrho.data());
      for (int i = 0; i < 2; i++) { nx.push_back(setup[k].g.nxp); nz.push_back(setup[k].g.nzp); }
    }
  return w.empty() ? 0 : pml_smooth_windows(w, nx, nz, pk[0].npml);
}

// direct-wave mute (kernel.cu:96-173) as the last muted sample per receiver,
// from the v^2 window `wvv` (or the resident model's window when null)
void mute_cuts(const rtmh_params& P, const float* vp0, ShotSetup& s, const float* wvv)
{
  const rtmh_geom& g = s.g;
  const int npml = P.npml, nxp = g.nxp;
  const float dx2 = P.dx * P.dx, dz2 = P.dz * P.dz;
  const float dt = P.dt;
  const float _dt = (float)(1.0 / (double)dt);
  const int nw = (int)(P.direct / (P.fdom * dt));
  const int tlength = (int)(P.direct0 / dt);
  const int coff = (g.rcol0 > npml) ? g.noffset : 0;   // layout 3 and recorded gathers (mute2)
  // v^2 of the padded window the forward pass uses (the mute kernels read d_vp)
  auto vv_at = [&](int iz, int ix) {
    if (wvv) return wvv[(size_t)iz * nxp + ix];
    const float v = win_at(vp0, P.nx, P.nz, g.nx1, g.nxl, npml, iz, ix);
    return v * v;
  };
  const float vvsrc = vv_at(g.nsz, g.nsx);
  s.cut.assign(g.nrec > 0 ? g.nrec : 1, -1);
  for (int ir = 0; ir < g.nrec; ir++) {
    const int col = ir + npml + coff;
    const float vvr = vv_at(npml, col);
    const float vtemp = (float)(0.482 * (1.0 / (double)sqrtf(vvr) + 1.0 / (double)sqrtf(vvsrc)));
    float dist = (col - g.nsx) * (col - g.nsx) * dx2 + (npml - g.nsz) * (npml - g.nsz) * dz2;
    dist = sqrtf(dist);
    const int temp = tlength < P.nt ? tlength : P.nt;
    const int c0 = (int)(dist * vtemp * _dt) + nw;
    int c = c0 < P.nt ? c0 : P.nt;
    c = c > temp ? c : temp;
    c = c < P.nt ? c : P.nt;
    s.cut[ir] = c;
  }
}

void shot_setup(const rtmh_params& P, const float* vp0, const float* colmax_vv, ShotSetup& s, const float* wvv)
{
  const rtmh_geom& g = s.g;
  const int npml = P.npml, nxp = g.nxp, nzp = g.nzp;
  // per-shot vmax over the padded window (rtm_model.cu:207): edge replication
  // adds no new values, so the window's column maxima of v^2 suffice; a
  // smoothed window gives its own maximum
  float m = colmax_vv[g.nx1];
  for (int ix = g.nx1; ix <= g.nx2; ix++) m = colmax_vv[ix] > m ? colmax_vv[ix] : m;
  if (wvv) {
    m = wvv[0];
    for (size_t i = 0; i < (size_t)nxp * nzp; i++) m = wvv[i] > m ? wvv[i] : m;
  }
  s.vpmax = sqrtf(m);
  damping(s.ax, s.bx, nxp, npml, s.vpmax, P.dx, false, P.dt);
  damping(s.axv, s.bxv, nxp - 1, npml, s.vpmax, P.dx, true, P.dt);
  damping(s.az, s.bz, nzp, npml, s.vpmax, P.dz, false, P.dt);
  damping(s.azv, s.bzv, nzp - 1, npml, s.vpmax, P.dz, true, P.dt);
  s.axv.push_back(1.0f); s.bxv.push_back(1.0f);              // pad to nxp / nzp
  s.azv.push_back(1.0f); s.bzv.push_back(1.0f);
  // Gaussian source (kernel.cu:23-35): bounding box of its non-zero values.
  // expf(-a) is 0 in float for a > 104 (below half the smallest denormal), so
  // only the rectangle around the source where alp^2 dist <= 120 (plus two
  // cells) can hold non-zero values; evaluating the whole window instead
  // (245 K expf for a 401 x 612 window) dominated the host set-up of a batch
  const float dx2 = P.dx * P.dx, dz2 = P.dz * P.dz;
  int x0 = nxp, x1 = -1, z0 = nzp, z1 = -1;
  int bx0 = 0, bx1 = nxp - 1, bz0 = 0, bz1 = nzp - 1;
  const double a2 = (double)P.alp * (double)P.alp;
  if (a2 > 0.0) {
    const double r2 = 120.0 / a2;
    const double rxd = std::ceil(std::sqrt(r2 / dx2)) + 2.0, rzd = std::ceil(std::sqrt(r2 / dz2)) + 2.0;
    if (rxd < nxp) { bx0 = std::max(0, g.nsx - (int)rxd); bx1 = std::min(nxp - 1, g.nsx + (int)rxd); }
    if (rzd < nzp) { bz0 = std::max(0, g.nsz - (int)rzd); bz1 = std::min(nzp - 1, g.nsz + (int)rzd); }
  }
  const int bw = std::max(0, bx1 - bx0 + 1), bh = std::max(0, bz1 - bz0 + 1);
  std::vector<float> src((size_t)bw * bh);
  for (int iz = bz0; iz <= bz1; iz++)
    for (int ix = bx0; ix <= bx1; ix++) {
      const float dist = (ix - g.nsx) * (ix - g.nsx) * dx2 + (iz - g.nsz) * (iz - g.nsz) * dz2;
      const float v = P.amp * expf(-P.alp * P.alp * dist);
      src[(size_t)(iz - bz0) * bw + ix - bx0] = v;
      if (v != 0.0f) { x0 = imin(x0, ix); x1 = imax(x1, ix); z0 = imin(z0, iz); z1 = imax(z1, iz); }
    }
  s.sa.clear();
  if (x1 >= x0) {
    s.sx0 = x0; s.sx1 = x1; s.sz0 = z0; s.sz1 = z1;
    for (int iz = z0; iz <= z1; iz++)
      for (int ix = x0; ix <= x1; ix++) s.sa.push_back(P.dt * src[(size_t)(iz - bz0) * bw + ix - bx0]);
  } else {
    s.sx0 = 0; s.sx1 = -1; s.sz0 = 0; s.sz1 = -1;
    s.sa.push_back(0.0f);
  }
  mute_cuts(P, vp0, s, wvv);
}

// ------------------------------------------------------------ checkpoint ----
// File: "RTMCKPT1", job fingerprint, nx, nz, shots done, then the three FP64
// stacks (image1, image2s, illum accumulators), nz*nx each.
struct CkptHead {
  char magic[8];
  uint64_t fp;
  int32_t nx, nz, done, version;
};

static uint64_t fnv1a(uint64_t h, const void* p, size_t n)
{
  const unsigned char* b = (const unsigned char*)p;
  for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ULL; }
  return h;
}

// everything that changes a shot's image except which shots run: parameters
// but nshot, the model, the coefficients, the boundary precision, the apertures
uint64_t Engine::job_fingerprint(const rtmh_params& P, const Model& M, const float* coef, const rtmh_options& opt,
                                 int nsh)
{
  rtmh_params q = P;
  q.nshot = 0;
  uint64_t h = 1469598103934665603ULL;
  h = fnv1a(h, &q, sizeof(q));
  h = fnv1a(h, &M.nx, sizeof(M.nx));
  h = fnv1a(h, &M.nz, sizeof(M.nz));
  h = fnv1a(h, M.vp0.data(), M.vp0.size() * 4);
  h = fnv1a(h, M.rho0.data(), M.rho0.size() * 4);
  h = fnv1a(h, coef, 100 * 4);
  h = fnv1a(h, &opt.store_fmt, sizeof(opt.store_fmt));
  h = fnv1a(h, &opt.ic_extra, sizeof(opt.ic_extra));
  h = fnv1a(h, &opt.snap_every, sizeof(opt.snap_every));
  h = fnv1a(h, &opt.snap_fmt, sizeof(opt.snap_fmt));
  if (opt.shot_offsetmax) h = fnv1a(h, opt.shot_offsetmax, (size_t)nsh * 4);
  if (opt.shot_field) h = fnv1a(h, opt.shot_field, (size_t)nsh * 12);
  return h;
}

// 1: read, 0: no file, -1: another job's or damaged
int Engine::read_checkpoint(const char* path, uint64_t fp, int nx, int nz, int nstk, int& done, std::vector<double>& st3)
{
  FILE* f = fopen(path, "rb");
  if (!f) return 0;
  CkptHead h;
  const size_t n = (size_t)nx * nz * nstk;
  st3.resize(n);
  const bool ok = fread(&h, sizeof(h), 1, f) == 1 && memcmp(h.magic, "RTMCKPT1", 8) == 0 && h.fp == fp &&
                  h.nx == nx && h.nz == nz && h.done >= 0 && fread(st3.data(), 8, n, f) == n;
  fclose(f);
  if (!ok) return -1;
  done = h.done;
  return 1;
}

int Engine::resume_point(const char* path, uint64_t fp, int nx, int nz, int nstk, int nsh, int& done,
                         std::vector<double>& st3, std::string& why)
{
  done = 0;
  const int cr = read_checkpoint(path, fp, nx, nz, nstk, done, st3);
  if (cr < 0) {
    why = std::string("checkpoint ") + path + " belongs to another job or is damaged";
    return -5;
  }
  if (cr == 1 && done > nsh) {
    // its stacks hold shots past this job's end
    why = std::string("checkpoint ") + path + " holds " + std::to_string(done) + " shots, more than this job's " +
          std::to_string(nsh);
    return -5;
  }
  return cr;
}

int Engine::write_checkpoint(const char* path, uint64_t fp, int nx, int nz, int done, const std::vector<double>& st3)
{
  CkptHead h;
  memcpy(h.magic, "RTMCKPT1", 8);
  h.fp = fp; h.nx = nx; h.nz = nz; h.done = done; h.version = 1;
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return -1;
  bool ok = fwrite(&h, sizeof(h), 1, f) == 1 && fwrite(st3.data(), 8, st3.size(), f) == st3.size();
  ok = (fflush(f) == 0) && ok;
  ok = (fclose(f) == 0) && ok;
  if (!ok || rename(tmp.c_str(), path) != 0) { remove(tmp.c_str()); return -1; }
  return 0;
}

// fn(0..n-1) on up to 16 host threads (independent items)
static void parallel_for(int n, const std::function<void(int)>& fn)
{
  const int nth = std::max(1, std::min({n, 16, (int)std::thread::hardware_concurrency()}));
  if (nth <= 1) { for (int k = 0; k < n; k++) fn(k); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nth; t++)
    th.emplace_back([&, t]() { for (int k = t; k < n; k += nth) fn(k); });
  for (auto& x : th) x.join();
}

// ------------------------------------------------------------------ pool ----
DevPool::~DevPool()
{
  trim();
  if (arena_) (void)hipFree(arena_);
}

bool DevPool::reserve(size_t bytes)
{
  if (arena_ || bytes < 256) return false;
  // only a pool with no block in use (every block it holds is cached) takes
  // an arena: put() maps each pointer to the arena, so no block from before
  // it may come back there
  if (size_.size() != free_.size()) return false;
  bytes &= ~(size_t)255;
  void* p = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  nalloc_++;
  const bool ok = hipMalloc(&p, bytes) == hipSuccess;
  if (!ok) (void)hipGetLastError();                  // handled: the pool stays lazy
  alloc_s_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (!ok) return false;
  trim();
  arena_ = (char*)p;
  arena_bytes_ = bytes;
  ext_.reset(bytes);
  return true;
}

void* DevPool::get(size_t bytes)
{
  if (arena_) {
    const size_t off = ext_.get(bytes);
    return off == Extents::kNone ? nullptr : arena_ + off;   // none: the caller waits or refuses
  }
  auto it = free_.lower_bound(bytes);
  if (it != free_.end() && it->first <= bytes + bytes / 4) {
    void* p = it->second;
    free_.erase(it);
    return p;
  }
  void* p = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  struct Clock {   // the host time of this call's device allocations
    DevPool* pool; std::chrono::steady_clock::time_point t0;
    ~Clock() { pool->alloc_s_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  } clock{this, t0};
  nalloc_++;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    // handled here (cached blocks go back, then a second try; the caller
    // reports a null return): the failure must not stay the thread's last
    // HIP error, which the next launch check would otherwise report
    (void)hipGetLastError();
    trim();
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  size_[p] = bytes;
  held_ += bytes;
  return p;
}

void DevPool::put(void* p)
{
  if (!p) return;
  if (arena_) {
    (void)ext_.put((size_t)((char*)p - arena_));
    return;
  }
  free_.insert({size_[p], p});
}

void DevPool::trim()
{
  const auto t0 = std::chrono::steady_clock::now();
  for (auto& kv : free_) {
    (void)hipFree(kv.second);
    held_ -= size_[kv.second];
    size_.erase(kv.second);
  }
  free_.clear();
  alloc_s_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------- engine ----
bool GraphKey::operator<(const GraphKey& o) const
{
  if (ntiles != o.ntiles) return ntiles < o.ntiles;
  if (tz != o.tz) return tz < o.tz;
  if (phase != o.phase) return phase < o.phase;
  if (save != o.save) return save < o.save;
  if (fmt != o.fmt) return fmt < o.fmt;
  if (k != o.k) return k < o.k;
  if (ic != o.ic) return ic < o.ic;
  if (nslots != o.nslots) return nslots < o.nslots;
  if (guard != o.guard) return guard < o.guard;
  if (snap != o.snap) return snap < o.snap;
  if (skip != o.skip) return skip < o.skip;
  if (shots != o.shots) return shots < o.shots;
  if (tiles != o.tiles) return tiles < o.tiles;
  return slots < o.slots;
}

Engine::Engine(int device) : device_(device)
{
  (void)hipSetDevice(device_);
  (void)hipStreamCreateWithFlags(&st_, hipStreamNonBlocking);
}

Engine::~Engine()
{
  drop_graphs();
  for (auto& kv : models_) {
    (void)hipFree(kv.second.d_rho); (void)hipFree(kv.second.d_vv);
    (void)hipFree(kv.second.d_ib); (void)hipFree(kv.second.d_kk);
  }
  models_.clear();
  if (cb_) (void)hipFree(cb_);
  pool_.trim();
  if (st_) (void)hipStreamDestroy(st_);
}

void Engine::drop_graphs()
{
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
}

static size_t store_elem(int fmt) { return fmt == ST_F32 ? 4 : 2; }

// fields of a slot: two levels of p, vx, vz for the source and the receiver
// wavefield, the PML pair (the model is read from the shared one)
static constexpr int kSlotFields = 14;

// small per-slot arrays, in this order, each 256-byte aligned
struct SmallLayout {
  size_t c[8], sa, cut, taper, total;
};
static SmallLayout small_layout(const ShotSetup& s)
{
  SmallLayout L;
  size_t o = 0;
  // damping arrays carry the field margins: columns -RTM_MXL .. pitch-RTM_MXL-1,
  // rows -RTM_MZT .. frows-RTM_MZT-1 (value 1 outside the grid)
  const size_t nx = Engine::pitch_of(s.g.nxp), nz = Engine::frows_of(s.g.nzp);
  const size_t n8[8] = {nx, nx, nx, nx, nz, nz, nz, nz};
  for (int i = 0; i < 8; i++) { L.c[i] = o; o += align256(n8[i] * 4); }
  L.sa = o; o += align256(s.sa.size() * 4);
  L.cut = o; o += align256(s.cut.size() * 4);
  L.taper = o; o += align256(s.taper.size() * 4 + 4);
  L.total = o;
  return L;
}

SnapPlan snap_plan(const rtmh_params& P, int every, int fmt)
{
  SnapPlan sp;
  if (every <= 0) return sp;
  // past the wavelet (the Ricker's delay 1/fdom and its width): there the
  // reference's reconstruction carries the one-step source-removal lag
  const int t0 = (int)std::ceil(4.0 / ((double)P.fdom * (double)P.dt));
  sp.every = every;
  sp.fmt = fmt;
  sp.k0 = std::max(1, (t0 + every - 1) / every);
  const int kmax = (P.nt - 2) / every;
  sp.n = std::max(0, kmax - sp.k0 + 1);
  return sp;
}

static size_t snap_bytes(const ShotSetup& s, int nz, const SnapPlan* sp)
{
  if (!sp || !sp->on()) return 0;
  const size_t e = sp->fmt == ST_F32 ? 4 : 2;
  return align256((size_t)sp->n * 3 * nz * s.g.nxl * e) + align256((size_t)sp->n * 3 * 8);
}

size_t Engine::slot_bytes(const ShotSetup& s, int nt, int store_fmt, bool ic, const SnapPlan* snap)
{
  const size_t pitch = pitch_of(s.g.nxp), nzp = s.g.nzp;
  const size_t F = align256((size_t)frows_of((int)nzp) * pitch * 4);
  const size_t nzi = nzp - 2 * (size_t)((s.g.nxp - s.g.nxl) / 2);
  const size_t I = align256(nzi * s.g.nxl * 4);
  const size_t strip = (size_t)40 * (nzi + s.g.nxl);
  size_t b = small_layout(s).total + kSlotFields * F + 2 * I;
  b += align256((size_t)nt * (s.g.nrec > 0 ? s.g.nrec : 1) * 4);
  b += align256((size_t)nt * strip * store_elem(store_fmt));
  b += align256((size_t)nt * 4 * 4 + 16 * 4);        // scales, 16-bit max-abs accumulators, guard word
  if (!s.wvv.empty()) b += 4 * align256((size_t)s.g.nxp * s.g.nzp * 4);   // smoothed windows: v^2, rho, 1/rho, rho v^2
  if (ic) b += align256((size_t)(5 + 2 * 61) * nzi * s.g.nxl * 4);         // other imaging conditions
  b += snap_bytes(s, (int)nzi, snap);
  return b;
}

void Engine::layout_slot(const rtmh_params& P, const ShotSetup& s, int store_fmt, char* base, ShotDev& d,
                         std::vector<std::pair<void*, std::vector<char>>>& uploads, const float* coef_dev,
                         const float* wav_dev, const ModelDev& model, bool ic, const SnapPlan* snap)
{
  memset(&d, 0, sizeof(d));
  d.dx = P.dx; d.dz = P.dz;
  const rtmh_geom& g = s.g;
  d.nxp = g.nxp; d.nzp = g.nzp; d.npml = P.npml; d.pitch = pitch_of(g.nxp);
  d.nxl = g.nxl; d.nzi = P.nz; d.nt = P.nt; d.phase = PH_IDLE;
  d.frows = frows_of(g.nzp); d.forg = RTM_MZT * d.pitch + RTM_MXL;
  d.it_off = 0; d.it_dir = 1;
  d.record = 1; d.rcol0 = g.rcol0; d.nrec = g.nrec;
  d.dt = P.dt; d.dtx = P.dt / P.dx; d.dtz = P.dt / P.dz;   // rtm_model.cu:73-74
  d.rho = model.rho; d.vv = model.vv; d.mib = model.ib; d.mkk = model.kk; d.mpitch = model.pitch;
  d.mz0 = P.npml; d.mzn = P.nz; d.mx0 = g.nx1; d.mxs = P.npml; d.mxn = g.nxl;
  // small arrays
  const SmallLayout L = small_layout(s);
  std::vector<char> img(L.total, 0);
  const std::vector<float>* cv[8] = {&s.ax, &s.bx, &s.axv, &s.bxv, &s.az, &s.bz, &s.azv, &s.bzv};
  const int pad[8] = {RTM_MXL, RTM_MXL, RTM_MXL, RTM_MXL, RTM_MZT, RTM_MZT, RTM_MZT, RTM_MZT};
  const int len[8] = {d.pitch, d.pitch, d.pitch, d.pitch, d.frows, d.frows, d.frows, d.frows};
  for (int i = 0; i < 8; i++) {
    float* dst = (float*)(img.data() + L.c[i]);
    for (int k = 0; k < len[i]; k++) dst[k] = 1.0f;
    memcpy(dst + pad[i], cv[i]->data(), cv[i]->size() * 4);
  }
  memcpy(img.data() + L.sa, s.sa.data(), s.sa.size() * 4);
  memcpy(img.data() + L.cut, s.cut.data(), s.cut.size() * 4);
  memcpy(img.data() + L.taper, s.taper.data(), s.taper.size() * 4);
  uploads.push_back({base, std::move(img)});
  const float** cp[8] = {&d.ax, &d.bx, &d.axv, &d.bxv, &d.az, &d.bz, &d.azv, &d.bzv};
  for (int i = 0; i < 8; i++) *cp[i] = (const float*)(base + L.c[i]) + pad[i];
  d.sx0 = s.sx0; d.sx1 = s.sx1; d.sz0 = s.sz0; d.sz1 = s.sz1;
  d.sa = (const float*)(base + L.sa);
  d.cut = (const int*)(base + L.cut);
  d.coef = coef_dev; d.wav = wav_dev;
  size_t o = L.total;
  const size_t F = align256((size_t)d.frows * d.pitch * 4);
  float** fp[kSlotFields] = {&d.sp[0], &d.sp[1], &d.svx[0], &d.svx[1], &d.svz[0], &d.svz[1],
                             &d.rp[0], &d.rp[1], &d.rvx[0], &d.rvx[1], &d.rvz[0], &d.rvz[1],
                             &d.q1, &d.q2};
  for (int i = 0; i < kSlotFields; i++) { *fp[i] = (float*)(base + o) + d.forg; o += F; }
  const size_t I = align256((size_t)d.nzi * d.nxl * 4);
  d.g1 = (float*)(base + o); o += I;
  d.illum = (float*)(base + o); o += I;
  d.rec = (float*)(base + o); o += align256((size_t)P.nt * (g.nrec > 0 ? g.nrec : 1) * 4);
  d.strip_len = (long long)40 * (d.nzi + d.nxl);
  d.store_fmt = store_fmt;
  d.store = base + o;
  o += align256((size_t)P.nt * d.strip_len * store_elem(store_fmt));
  d.sscale = (float*)(base + o); o += align256((size_t)P.nt * 4 * 4 + 16 * 4);
  if (!s.wvv.empty()) {
    // the shot's own (smoothed) padded windows as the model the slot reads;
    // 1/rho and rho v^2 of them by k_model_prep once they are uploaded
    const size_t wb = (size_t)g.nxp * g.nzp * 4;
    float* wv = (float*)(base + o); o += align256(wb);
    float* wr = (float*)(base + o); o += align256(wb);
    float* wi = (float*)(base + o); o += align256(wb);
    float* wk = (float*)(base + o); o += align256(wb);
    uploads.push_back({wv, std::vector<char>((const char*)s.wvv.data(), (const char*)s.wvv.data() + wb)});
    uploads.push_back({wr, std::vector<char>((const char*)s.wrho.data(), (const char*)s.wrho.data() + wb)});
    d.vv = wv; d.rho = wr; d.mib = wi; d.mkk = wk;
    d.mpitch = g.nxp; d.mz0 = 0; d.mzn = g.nzp; d.mx0 = 0; d.mxs = 0; d.mxn = g.nxp;
  }
  if (ic) { d.ic = (float*)(base + o); o += align256((size_t)(5 + 2 * 61) * d.nzi * d.nxl * 4); }
  if (snap && snap->on()) {
    const size_t e = snap->fmt == ST_F32 ? 4 : 2;
    d.snap = base + o;
    o += align256((size_t)snap->n * 3 * d.nzi * d.nxl * e);
    d.snap_scale = (float*)(base + o);
    o += align256((size_t)snap->n * 3 * 8);
    d.snap_every = snap->every; d.snap_k0 = snap->k0; d.nsnap = snap->n; d.snap_fmt = snap->fmt;
  }
}

hipGraphExec_t Engine::graph_for(const GraphKey& k, const ShotDev* dshots, const Tile* dtiles, const int* dslots,
                                 int nslots, int store_fmt)
{
  auto it = graphs_.find(k);
  if (it != graphs_.end()) return it->second;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  LaunchArgs a{dshots, dtiles, k.ntiles, 0};
  if (k.skip) a.act = d_act_;
  for (int s = 0; s < k.k; s++) {
    a.step = s;
    (void)rtm_launch_step(a, k.phase, k.tz, st_);
    if (k.save) (void)rtm_launch_save(dshots, dslots, nslots, s, store_fmt, st_);
    if (k.ic) (void)rtm_launch_ic(dshots, dslots, nslots, s, k.ic, st_);
    if (k.snap) (void)rtm_launch_snap(dshots, dslots, nslots, k.phase, s, k.snap - 1, st_);
  }
  (void)rtm_launch_bump((ShotDev*)dshots, dslots, nslots, k.k, st_);
  // the chunk's last step, relative to the advanced time origin
  if (k.guard) (void)rtm_launch_guard(dshots, dslots, nslots, -1, d_flag_, st_);
  if (hipStreamEndCapture(st_, &g) != hipSuccess || !g) return nullptr;
  hipGraphExec_t ex = nullptr;
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
  (void)hipGraphDestroy(g);
  if (ex) graphs_[k] = ex;
  (void)store_fmt;
  return ex;
}

int Engine::run_pass(int phase, int nsteps, const ShotDev* dshots, const Tile* dtiles, int ntiles,
                     const std::vector<int>& slots, const int* dslots, int tz, int store_fmt, int chunk,
                     rtmh_stats* prof, int guard_every, bool skip)
{
  // zero-tile skip words: every tile active until its first step has stepped
  // (the words then follow the fields), whatever the fields hold now
  unsigned long long* act = skip ? d_act_ : nullptr;
  if (act) HIPCHK(hipMemsetAsync(act, 0xFF, 2 * (size_t)ntiles * sizeof(unsigned long long), st_));
  const int save = phase == PH_FWD ? 1 : 0;
  const int ic = phase == PH_REV ? ic_cells_ : 0;          // the other imaging conditions after each reverse step
  const int nslots = (int)slots.size();
  int done = 0;
  if (prof) {
    // every step kernel (and the strip save after it) bracketed by its own
    // event pair on the launch stream
    hipEvent_t ev[3];
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    LaunchArgs a{dshots, dtiles, ntiles, 0};
    a.act = act;
    const int kb = phase == PH_FWD ? 0 : 2;
    for (int s = 0; s < nsteps; s++) {
      a.step = s;
      HIPCHK(hipEventRecord(ev[0], st_));
      HIPCHK(rtm_launch_step(a, phase, tz, st_));
      HIPCHK(hipEventRecord(ev[1], st_));
      if (save) HIPCHK(rtm_launch_save(dshots, dslots, nslots, s, store_fmt, st_));
      if (ic) HIPCHK(rtm_launch_ic(dshots, dslots, nslots, s, ic, st_));
      if (snap_.on()) HIPCHK(rtm_launch_snap(dshots, dslots, nslots, phase, s, snap_.fmt, st_));
      HIPCHK(hipEventRecord(ev[2], st_));
      HIPCHK(hipEventSynchronize(ev[2]));
      float m0 = 0, m1 = 0;
      HIPCHK(hipEventElapsedTime(&m0, ev[0], ev[1]));
      HIPCHK(hipEventElapsedTime(&m1, ev[1], ev[2]));
      prof->kern_ms[kb] += m0; prof->kern_calls[kb]++;
      if (save) { prof->kern_ms[kb + 1] += m1; prof->kern_calls[kb + 1]++; }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return 0;
  }
  if (chunk > 0 && nsteps >= chunk) {
    GraphKey key{ntiles, tz, phase, save, store_fmt, chunk, ic, nslots, guard_every > 0 ? 1 : 0,
                 snap_.on() ? 1 + snap_.fmt : 0, act ? 1 : 0, dshots, dtiles, dslots};
    hipGraphExec_t ex = graph_for(key, dshots, dtiles, dslots, nslots, store_fmt);
    if (!ex) { set_error("hipGraph capture/instantiate failed"); return -11; }
    const int nch = nsteps / chunk;
    for (int c = 0; c < nch; c++) HIPCHK(hipGraphLaunch(ex, st_));
    done = nch * chunk;
  }
  LaunchArgs a{dshots, dtiles, ntiles, 0};
  a.act = act;
  const int rest = nsteps - done;
  for (int s = 0; s < rest; s++) {
    a.step = s;
    HIPCHK(rtm_launch_step(a, phase, tz, st_));
    if (save) HIPCHK(rtm_launch_save(dshots, dslots, nslots, s, store_fmt, st_));
    if (ic) HIPCHK(rtm_launch_ic(dshots, dslots, nslots, s, ic, st_));
    if (snap_.on()) HIPCHK(rtm_launch_snap(dshots, dslots, nslots, phase, s, snap_.fmt, st_));
    if (guard_every > 0 && ((s + 1) % guard_every == 0 || s == rest - 1))
      HIPCHK(rtm_launch_guard(dshots, dslots, nslots, s, d_flag_, st_));
  }
  return 0;
}

int Engine::guard_clear()
{
  static const unsigned int zero = 0u;
  HIPCHK(hipMemcpyAsync(d_flag_, &zero, 4, hipMemcpyHostToDevice, st_));
  return 0;
}

int Engine::guard_check(const std::vector<ShotDev>& ds, const std::vector<int>& shots, int k0, const char* pass)
{
  unsigned int f = 0;
  HIPCHK(hipMemcpy(&f, d_flag_, 4, hipMemcpyDeviceToHost));
  if (f < 0x7F800000u) return 0;
  std::string bad;
  for (size_t j = 0; j < ds.size(); j++) {
    unsigned int w = 0;
    HIPCHK(hipMemcpy(&w, rtm_guard_word(ds[j]), 4, hipMemcpyDeviceToHost));
    if (w >= 0x7F800000u) bad += (bad.empty() ? "" : ", ") + std::to_string(shots[k0 + j]);
  }
  set_error(std::string("the ") + pass + " wavefield of shot(s) " + bad +
            " is not finite (blow-up: dt above the stability bound for the model, or non-finite model values)");
  return -7;
}

// tile table of a batch with the per-tile flags (TF_*) of the cell classes
// the tile and the halo its step reads hold.  A slot's tiles are listed in
// bands of two tile rows, column by column (z0, x0), (z0 + tz, x0), (z0,
// x0 + 64), ...: a tile's pressure halo reaches 19-20 rows into the tiles
// above and below, and with the lower neighbour next in the XCD's run those
// rows are still in L2.  Measured on the 200-shot survey against row-by-row
// order, alternated on two boxes: +1.0 to +1.5 % and +0.5 % (bands of 3 and
// 4 rows: the same as 2; whole columns: no gain).
void build_tiles(const std::vector<ShotDev>& ds, int tz, std::vector<Tile>& tiles, const std::vector<int>* slot_ids)
{
  tiles.clear();
  for (int k = 0; k < (int)ds.size(); k++) {
    const int sid = slot_ids ? (*slot_ids)[k] : k;
    const ShotDev& d = ds[k];
    const int n = d.npml, R = RTM_R, nxp = d.nxp, nzp = d.nzp;
    for (int band = 0; band < nzp; band += 2 * tz)
      for (int x0 = 0; x0 < nxp; x0 += RTM_TX)
        for (int z0 = band; z0 < std::min(band + 2 * tz, nzp); z0 += tz) {
        const int z1 = z0 + tz - 1, x1 = x0 + RTM_TX - 1;
        // rectangle [za,zb] x [xa,xb] against the tile grown by (hz, hx)
        auto hit = [&](int hz, int hx, int za, int zb, int xa, int xb) {
          return z0 - hz <= zb && z1 + hz >= za && x0 - hx <= xb && x1 + hx >= xa;
        };
        int f = 0;
        // strips (vx, p: columns n-R..n-1 and nxp-n-1..nxp-n+9 over interior
        // rows; vz, p: rows n-R..n-1 and nzp-n-1..nzp-n+9 over interior
        // columns), against the velocity halo of the tile
        if (hit(10, 10, n, nzp - n - 1, n - R, n - 1) || hit(10, 10, n, nzp - n - 1, nxp - n - 1, nxp - n + R - 1) ||
            hit(10, 10, n - R, n - 1, n, nxp - n - 1) || hit(10, 10, nzp - n - 1, nzp - n + R - 1, n, nxp - n - 1))
          f |= TF_STRIP;
        if (d.sx1 >= d.sx0 && hit(0, 0, d.sz0, d.sz1, d.sx0, d.sx1)) f |= TF_SRC;
        if (d.nrec > 0 && hit(0, 0, n, n, d.rcol0, d.rcol0 + d.nrec - 1)) f |= TF_REC;
        // undamped full-radius interior for every velocity the step computes
        // (vx columns x0-10..x0+73, vz rows z0-10..z0+tz+8)
        const bool inside = x0 - 10 >= n && x1 + 10 <= nxp - n - 2 && z0 - 10 >= n && z1 + 9 <= nzp - n - 2;
        if (!inside) f |= TF_GEN;
        // reverse phase: reconstructed cells (interior and strips)
        if (hit(0, 0, n - R, nzp - n + R - 1, n - R, nxp - n + R - 1)) f |= TF_INT;
        tiles.push_back(Tile{sid, x0, z0, f});
      }
  }
}

// Runs of the tile table per XCD.  The step kernels deal workgroups to the 8
// XCDs round-robin and give XCD c the entries [c*per, (c+1)*per) of the
// table in order (xcd_tile; per = launched workgroups / 8).  The table is cut
// into 8 consecutive runs of equal estimated cost -- the slowest XCD sets the
// length of a launch -- and each run is padded to the longest with empty
// entries (slot -1), whose workgroups return at once.  Within a run the long
// tiles go first, so that a launch ends on short ones: in the reverse step the
// tiles holding reconstructed cells (two field steps against one), in the
// forward step those of the absorbing layer (the split pair and the per-cell
// rules; about 1.55 times an interior tile).  Measured on the 200-shot
// survey, alternated on one box: long tiles first, reverse pass -1.0 to
// -1.3 % (four pairs) and forward pass -4 % (three pairs); runs of equal cost
// instead of equal count, a further -0.3 to -0.5 % (three pairs).
void layout_runs(std::vector<Tile>& tiles, int phase)
{
  const size_t n = tiles.size();
  if (n == 0) return;
  const int flag = phase == PH_REV ? TF_INT : TF_GEN;
  const double heavy = phase == PH_REV ? 2.0 : 1.55;
  auto cost = [&](const Tile& t) { return (t.kind & flag) ? heavy : 1.0; };
  double total = 0;
  for (const Tile& t : tiles) total += cost(t);
  size_t cut[9];
  cut[0] = 0;
  int c = 1;
  double acc = 0;
  for (size_t i = 0; i < n && c < 8; i++) {
    acc += cost(tiles[i]);
    while (c < 8 && acc >= total * c / 8.0) cut[c++] = i + 1;
  }
  while (c <= 8) cut[c++] = n;
  size_t lmax = 0;
  for (int k = 0; k < 8; k++) lmax = std::max(lmax, cut[k + 1] - cut[k]);
  const size_t per = (n + 7) / 8;
  if (lmax > per + per / 4) {                          // runs far from equal in count: equal counts
    for (int k = 0; k <= 8; k++) cut[k] = std::min(n, (size_t)k * per);
    lmax = per;
  }
  std::vector<Tile> out(8 * lmax, Tile{-1, 0, 0, 0});
  for (int k = 0; k < 8; k++) {
    auto b = out.begin() + (size_t)k * lmax;
    std::copy(tiles.begin() + cut[k], tiles.begin() + cut[k + 1], b);
    std::stable_partition(b, b + (cut[k + 1] - cut[k]), [flag](const Tile& t) { return (t.kind & flag) != 0; });
  }
  tiles.swap(out);
}

// Tile::nbr from the final table: tiles are keyed by (slot, z0, x0); the
// neighbours of (z0, x0) are at z0 +- tz and x0 +- 64 in the same slot
void link_neighbours(std::vector<Tile>& tiles, int tz)
{
  std::map<std::tuple<int, int, int>, int> at;
  for (int i = 0; i < (int)tiles.size(); i++)
    if (tiles[i].slot >= 0) at[std::make_tuple(tiles[i].slot, tiles[i].z0, tiles[i].x0)] = i;
  for (Tile& t : tiles) {
    int n = 0;
    for (int dz = -1; dz <= 1; dz++)
      for (int dx = -1; dx <= 1; dx++) {
        if (dz == 0 && dx == 0) continue;
        int k = -1;
        if (t.slot >= 0) {
          auto it = at.find(std::make_tuple(t.slot, t.z0 + dz * tz, t.x0 + dx * RTM_TX));
          if (it != at.end()) k = it->second;
        }
        t.nbr[n++] = k;
      }
  }
}

int params_check(const rtmh_params& P, std::string& why);
static bool params_ok(const rtmh_params& P, std::string& why) { return params_check(P, why) == 0; }

int params_check(const rtmh_params& P, std::string& why)
{
  if (P.nx < 2 || P.nz < 2) { why = "nx, nz must be >= 2"; return -1; }
  if (P.nt < 2) { why = "nt must be >= 2"; return -1; }
  if (P.npml < 10) { why = "npml must be >= 10 (radius-10 strips)"; return -1; }
  if (!(P.dx > 0 && P.dz > 0 && P.dt > 0)) { why = "dx, dz, dt must be > 0"; return -1; }
  if (P.flag_layout < 1 || P.flag_layout > 4) { why = "flag_layout must be 1..4"; return -1; }
  if (P.dshot <= 0) { why = "dshot must be > 0"; return -1; }
  if (P.ismth < 0) { why = "ismth must be >= 0"; return -1; }
  return 0;
}

float dt_stable(const rtmh_params& P, const float* coef, float vmax)
{
  float sum = 0.0f;                                           // sum1abs(diffcoef, N/2)
  for (int i = 0; i < RTM_R; i++) sum += fabsf(coef[(RTM_R - 1) * RTM_R + i]);
  const float hmin = P.dx < P.dz ? P.dx : P.dz;
  return hmin / (sqrtf(2.0f) * sum * vmax);
}

int stability_check(const rtmh_params& P, const float* coef, float vmax, std::string& why)
{
  if (!(vmax > 0.0f) || !std::isfinite(vmax)) { why = "the model's largest velocity is not a positive number"; return -1; }
  const float dtm = dt_stable(P, coef, vmax);
  if (P.dt > dtm) {
    char b[256];
    snprintf(b, sizeof(b),
             "dt = %g s exceeds the stability bound min(dx,dz)/(sqrt(2) sum|c| vmax) = %g s for vmax = %g m/s "
             "(the driver's get_constant refines dt to meet it)", (double)P.dt, (double)dtm, (double)vmax);
    why = b;
    return -1;
  }
  return 0;
}

int Engine::model_create(int nx, int nz, const float* vp0, const float* rho0, int* id)
{
  if (nx < 2 || nz < 2 || !vp0 || !rho0) { set_error("bad model"); return -1; }
  for (size_t i = 0; i < (size_t)nx * nz; i++)
    if (!(vp0[i] >= 0.0f) || !(rho0[i] > 0.0f) || !std::isfinite(vp0[i]) || !std::isfinite(rho0[i])) {
      set_error("bad model: velocity must be finite and >= 0, density finite and > 0 (cell " + std::to_string(i) + ")");
      return -1;
    }
  HIPCHK(hipSetDevice(device_));
  Model m;
  m.nx = nx; m.nz = nz;
  const size_t nm = (size_t)nx * nz;
  m.vp0.assign(vp0, vp0 + nm);
  m.rho0.assign(rho0, rho0 + nm);
  std::vector<float> vv(nm);
  // v^2 as extractvel1 stores it (Velocity_function.cpp:2059)
  for (size_t i = 0; i < nm; i++) vv[i] = vp0[i] * vp0[i];
  m.colmax.assign(nx, 0.0f);
  for (int ix = 0; ix < nx; ix++) {
    float mx = vv[ix];
    for (int iz = 0; iz < nz; iz++) mx = vv[(size_t)iz * nx + ix] > mx ? vv[(size_t)iz * nx + ix] : mx;
    m.colmax[ix] = mx;
  }
  if (hipMalloc(&m.d_rho, nm * 4) != hipSuccess || hipMalloc(&m.d_vv, nm * 4) != hipSuccess ||
      hipMalloc(&m.d_ib, nm * 4) != hipSuccess || hipMalloc(&m.d_kk, nm * 4) != hipSuccess) {
    (void)hipFree(m.d_rho); (void)hipFree(m.d_vv); (void)hipFree(m.d_ib); (void)hipFree(m.d_kk);
    (void)hipGetLastError();                         // reported here, not by a later launch
    set_error("out of device memory (model)");
    return -3;
  }
  HIPCHK(hipMemcpy(m.d_rho, rho0, nm * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m.d_vv, vv.data(), nm * 4, hipMemcpyHostToDevice));
  HIPCHK(rtm_launch_model_prep(m.d_rho, m.d_vv, m.d_ib, m.d_kk, (long long)nm, st_));
  HIPCHK(hipStreamSynchronize(st_));
  m.bytes = 4 * nm * 4;
  std::lock_guard<std::mutex> lk(models_mu_);
  const int k = next_model_++;
  models_[k] = std::move(m);
  *id = k;
  return 0;
}

int Engine::model_destroy(int id)
{
  std::lock_guard<std::mutex> lk(models_mu_);
  auto it = models_.find(id);
  if (it == models_.end()) { set_error("no such model"); return -1; }
  (void)hipFree(it->second.d_rho);
  (void)hipFree(it->second.d_vv);
  (void)hipFree(it->second.d_ib);
  (void)hipFree(it->second.d_kk);
  models_.erase(it);
  return 0;
}

size_t Engine::taper_offset(const ShotSetup& s) { return small_layout(s).taper; }

const Engine::Model* Engine::model(int id) const
{
  std::lock_guard<std::mutex> lk(models_mu_);
  auto it = models_.find(id);
  return it == models_.end() ? nullptr : &it->second;
}

int Engine::run(const rtmh_params& P, const float* vp0, const float* rho0, const float* coef, const rtmh_options& opt,
                rtmh_outputs* out, rtmh_stats* stats)
{
  int id = -1;
  int rc = model_create(P.nx, P.nz, vp0, rho0, &id);
  if (rc) return rc;
  rc = run_model(id, P, coef, opt, out, stats);
  model_destroy(id);
  return rc;
}

// make the persistent control block (descriptors, tiles, slot list, stack
// entries) large enough; graphs captured against the old block are dropped
int Engine::ensure_cb(int nslots, size_t ntiles)
{
  if (nslots <= cb_slots_ && ntiles <= cb_tiles_) return 0;
  drop_graphs();
  if (cb_) (void)hipFree(cb_);
  cb_ = nullptr;
  cb_slots_ = nslots > cb_slots_ ? nslots : cb_slots_;
  if (cb_slots_ < 64) cb_slots_ = 64;
  cb_tiles_ = ntiles > cb_tiles_ ? ntiles * 2 : cb_tiles_;
  const size_t bytes = align256(cb_slots_ * sizeof(ShotDev)) + align256(cb_tiles_ * sizeof(Tile)) +
                       align256(cb_slots_ * 4) + align256(cb_slots_ * sizeof(StackShotHost)) + 256 +
                       align256(2 * cb_tiles_ * sizeof(unsigned long long));
  HIPCHK(hipMalloc(&cb_, bytes));
  d_ds_ = (ShotDev*)cb_;
  d_tiles_ = (Tile*)((char*)cb_ + align256(cb_slots_ * sizeof(ShotDev)));
  d_slots_ = (int*)((char*)d_tiles_ + align256(cb_tiles_ * sizeof(Tile)));
  d_stk_ = (StackShotHost*)((char*)d_slots_ + align256(cb_slots_ * 4));
  d_flag_ = (unsigned int*)((char*)d_stk_ + align256(cb_slots_ * sizeof(StackShotHost)));
  d_act_ = (unsigned long long*)((char*)d_flag_ + 256);
  return 0;
}

int Engine::run_model(int model_id, const rtmh_params& P, const float* coef, const rtmh_options& opt,
                      rtmh_outputs* out, rtmh_stats* stats)
{
  auto t0 = std::chrono::steady_clock::now();
  std::string why;
  if (!params_ok(P, why)) { set_error(why); return -1; }
  if (opt.store_fmt < 0 || opt.store_fmt > 2) { set_error("store_fmt must be 0, 1 or 2"); return -1; }
  if (opt.snap_every < 0 || opt.snap_fmt < 0 || opt.snap_fmt > 2) {
    set_error("snap_every must be >= 0 and snap_fmt 0, 1 or 2");
    return -1;
  }
  if (opt.tile_z != 0 && opt.tile_z != 16 && opt.tile_z != 32 && opt.tile_z != 64) {
    set_error("tile_z must be 0 (autotuned), 16, 32 or 64");
    return -1;
  }
  int tz = opt.tile_z > 0 ? opt.tile_z : 16;
  const Model* M = model(model_id);
  if (!M) { set_error("no such model"); return -1; }
  if (M->nx != P.nx || M->nz != P.nz) { set_error("model size differs from nx, nz"); return -1; }
  {
    float mv = 0.0f;
    for (float c : M->colmax) mv = c > mv ? c : mv;
    if (stability_check(P, coef, sqrtf(mv), why) != 0) { set_error(why); return -1; }
  }
  // wavefield guard interval: the reference's ntsnap (rtm_model.cu:21)
  const int guard_every = std::max(1, (int)(P.tsnap / P.dt));
  HIPCHK(hipSetDevice(device_));
  const float* vp0 = M->vp0.data();

  const int nx = P.nx, nz = P.nz, nt = P.nt;
  std::vector<int> shots;
  for (int is = P.startshot; is <= P.nshot; is += P.dshot) shots.push_back(is);
  const int nsh = (int)shots.size();
  const size_t nm = (size_t)nx * nz;
  // per-shot set-up, taper window state carried in shot order (rtm_model.cu:70, 465-466)
  std::vector<ShotSetup> setup(nsh);
  std::vector<rtmh_params> pk(nsh, P);
  for (int k = 0; k < nsh; k++) {
    if (opt.shot_offsetmax) pk[k].offsetmax = opt.shot_offsetmax[k];
    if (!shot_geometry(pk[k], shots[k], setup[k], opt.shot_field ? opt.shot_field + 3 * k : nullptr)) {
      set_error("shot window outside the model");
      return -2;
    }
  }
  // model smoothing: each shot gets its own padded windows (smoothed on the
  // host as the reference does, shots in parallel); the slots then read them
  // instead of the resident model
  const bool smooth = P.medium_flag != 0 || P.ismth > 0;
  const bool icx = opt.ic_extra != 0;
  snap_ = snap_plan(P, opt.snap_every, opt.snap_fmt);
  const SnapPlan* snp = snap_.on() ? &snap_ : nullptr;
  if (smooth) {
    const int wr = job_windows(pk.data(), vp0, M->rho0.data(), setup);
    if (wr != 0) return wr;
  }
  int nwin = 50;
  for (int k = 0; k < nsh; k++) {
    shot_setup(pk[k], vp0, M->colmax.data(), setup[k], smooth ? setup[k].wvv.data() : nullptr);
    while (2 * nwin >= setup[k].g.nxl) nwin -= 3;
    setup[k].nwin = nwin;
    taper_row(setup[k].taper, setup[k].g.nxl, nwin);
  }
  std::vector<float> wav;
  ricker(wav, nt, P.fdom, P.dt);

  // device: model, coefficients, wavelet, FP64 stacks
  const size_t common_bytes = align256(100 * 4) + align256((size_t)nt * 4) +
                              3 * align256(nm * 8) + 3 * align256(nm * 4) +
                              (icx ? 7 * align256(nm * 8) + 7 * align256(nm * 4) + align256(nsh * sizeof(StackIcHost)) : 0);
  char* common = (char*)pool_.get(common_bytes);
  if (!common) { set_error("out of device memory (job)"); return -3; }
  const ModelDev mdev = M->dev();
  float* d_coef = (float*)common;
  float* d_wav = (float*)((char*)d_coef + align256(100 * 4));
  double* d_i1 = (double*)((char*)d_wav + align256((size_t)nt * 4));
  double* d_i2 = (double*)((char*)d_i1 + align256(nm * 8));
  double* d_il = (double*)((char*)d_i2 + align256(nm * 8));
  float* d_f1 = (float*)((char*)d_il + align256(nm * 8));
  float* d_f2 = (float*)((char*)d_f1 + align256(nm * 4));
  float* d_f3 = (float*)((char*)d_f2 + align256(nm * 4));
  // the other imaging conditions: 7 FP64 stacks, their fp32 images, stack entries
  double* d_i7 = icx ? (double*)((char*)d_f3 + align256(nm * 4)) : nullptr;
  float* d_f7 = icx ? (float*)((char*)d_i7 + 7 * align256(nm * 8)) : nullptr;
  StackIcHost* d_stkic = icx ? (StackIcHost*)((char*)d_f7 + 7 * align256(nm * 4)) : nullptr;
  auto i7 = [&](int k) { return d_i7 + (size_t)k * nm; };     // contiguous: k_stack_ic indexes k*nx*nz
  auto f7 = [&](int k) { return d_f7 + (size_t)k * nm; };
  if (icx)
    for (int k = 0; k < 7; k++) HIPCHK(rtm_launch_zero_f64(i7(k), (long long)nm, st_));
  // every FP64 stack (checkpoint order)
  std::vector<double*> stacks = {nullptr, nullptr, nullptr};
  HIPCHK(hipMemcpyAsync(d_coef, coef, 100 * 4, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(d_wav, wav.data(), (size_t)nt * 4, hipMemcpyHostToDevice, st_));
  HIPCHK(rtm_launch_zero_f64(d_i1, (long long)nm, st_));
  HIPCHK(rtm_launch_zero_f64(d_i2, (long long)nm, st_));
  HIPCHK(rtm_launch_zero_f64(d_il, (long long)nm, st_));
  HIPCHK(hipStreamSynchronize(st_));
  stacks = {d_i1, d_i2, d_il};
  if (icx)
    for (int k = 0; k < 7; k++) stacks.push_back(i7(k));
  const int nstk = (int)stacks.size();

  // checkpoint / resume: the stacks of the shots done so far, in job order
  const uint64_t job_fp = job_fingerprint(P, *M, coef, opt, nsh);
  int k_start = 0;
  if (opt.checkpoint && opt.resume) {
    std::vector<double> st3;
    int done = 0;
    const int cr = resume_point(opt.checkpoint, job_fp, nx, nz, nstk, nsh, done, st3, why);
    if (cr < 0) {
      set_error(why);
      pool_.put(common);
      return cr;
    }
    if (cr == 1) {
      k_start = done;
      for (int i = 0; i < nstk; i++)
        HIPCHK(hipMemcpy(stacks[i], st3.data() + (size_t)i * nm, nm * 8, hipMemcpyHostToDevice));
    }
  }

  // memory budget and batch size
  size_t freeb = 0, totalb = 0;
  HIPCHK(hipMemGetInfo(&freeb, &totalb));
  double budget = opt.mem_budget > 0 ? opt.mem_budget : 0.8 * (double)(freeb + pool_.held());
  const int bmax = opt.batch > 0 ? opt.batch : nsh;   // 0: as many as the budget holds

  hipEvent_t e0, e1, e2;
  HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1)); HIPCHK(hipEventCreate(&e2));
  double fwd_ms = 0, rev_ms = 0, cells = 0, algb = 0, step_fwd = 0, step_rev = 0, fwdb = 0, revb = 0;
  int last_tz = tz, last_tz_r = tz;
  if (stats) memset(stats, 0, sizeof(*stats));
  int nbatches = 0, maxb = 0;
  std::vector<double> img_tmp;  // for temp-image checkpoints
  int rc = 0;

  for (int k0 = k_start; k0 < nsh && rc == 0;) {
    // greedy batch of consecutive shots within the budget
    int k1 = k0;
    double used = 0;
    while (k1 < nsh && k1 - k0 < bmax) {
      const double b = (double)slot_bytes(setup[k1], nt, opt.store_fmt, icx, snp);
      if (k1 > k0 && used + b > budget) break;
      used += b;
      k1++;
    }
    const int nb = k1 - k0;
    maxb = nb > maxb ? nb : maxb;
    nbatches++;
    std::vector<void*> mem(nb, nullptr);
    std::vector<ShotDev> ds(nb);
    std::vector<std::pair<void*, std::vector<char>>> uploads;
    for (int j = 0; j < nb; j++) {
      mem[j] = pool_.get(slot_bytes(setup[k0 + j], nt, opt.store_fmt, icx, snp));
      if (!mem[j]) { set_error("out of device memory (shot slots)"); rc = -3; break; }
      layout_slot(P, setup[k0 + j], opt.store_fmt, (char*)mem[j], ds[j], uploads, d_coef, d_wav, mdev, icx, snp);
      ds[j].record = (P.mode == 0) ? 1 : 0;
    }
    if (rc != 0) { for (void* p : mem) pool_.put(p); break; }
    // tile height: given, cached for this geometry class, or tuned on the
    // first forward steps of this batch (every candidate computes the same
    // values, so the tuning steps are part of the pass)
    int maxp = 0, maxz = 0;
    for (int j = 0; j < nb; j++) { maxp = std::max(maxp, ds[j].nxp); maxz = std::max(maxz, ds[j].nzp); }
    // the forward and reverse step kernels differ, so each pass has its own
    // tile height (key bit 62: reverse)
    const long long tkey = ((long long)(maxp / 64) << 40) | ((long long)(maxz / 32) << 20) | nb;
    const long long rkey = tkey | (1LL << 62);
    const int cands[3] = {16, 32, 64};
    bool tune = false, tune_r = false;
    int tz_r = tz;
    if (opt.tile_z == 0) {
      const bool can = nt >= 3 * kTuneSteps + 2 && !opt.profile;
      auto ct = tuned_.find(tkey);
      if (ct != tuned_.end()) tz = ct->second;
      else tune = can;
      auto cr = tuned_.find(rkey);
      if (cr != tuned_.end()) tz_r = cr->second;
      else tune_r = can;
    }
    std::vector<Tile> tiles, tcand[3];
    size_t tcap = 0;
    for (int c = 0; c < 3; c++) {
      build_tiles(ds, cands[c], tcand[c]);
      tcap = std::max(tcap, tcand[c].size());
    }
    build_tiles(ds, tz, tiles);
    layout_runs(tiles, PH_FWD);
    link_neighbours(tiles, tz);
    // room for any height's padded table (runs at most 1.25 x the even share)
    tcap = std::max(tcap, tiles.size()) * 5 / 4 + 16;
    // the first 3 x kTuneSteps steps of a pass, kTuneSteps with each candidate
    // height, timed with events; every candidate computes the same cells, so
    // these are steps of the pass.  Returns the fastest height.
    auto tune_pass = [&](int phase) -> int {
      hipEvent_t te[4];
      for (auto& e : te) HIPCHK(hipEventCreate(&e));
      HIPCHK(hipEventRecord(te[0], st_));
      for (int c = 0; c < 3; c++) {
        HIPCHK(hipMemcpyAsync(d_tiles_, tcand[c].data(), tcand[c].size() * sizeof(Tile), hipMemcpyHostToDevice, st_));
        LaunchArgs a{d_ds_, d_tiles_, (int)tcand[c].size(), 0};
        for (int s2 = 0; s2 < kTuneSteps; s2++) {
          a.step = c * kTuneSteps + s2;
          HIPCHK(rtm_launch_step(a, phase, cands[c], st_));
          if (phase == PH_FWD) HIPCHK(rtm_launch_save(d_ds_, d_slots_, nb, a.step, opt.store_fmt, st_));
          else if (ic_cells_) HIPCHK(rtm_launch_ic(d_ds_, d_slots_, nb, a.step, ic_cells_, st_));
          if (snap_.on()) HIPCHK(rtm_launch_snap(d_ds_, d_slots_, nb, phase, a.step, snap_.fmt, st_));
        }
        HIPCHK(hipEventRecord(te[c + 1], st_));
      }
      HIPCHK(rtm_launch_bump(d_ds_, d_slots_, nb, 3 * kTuneSteps, st_));
      HIPCHK(hipEventSynchronize(te[3]));
      float best = 1e30f;
      int pick = cands[0];
      for (int c = 0; c < 3; c++) {
        float m = 0;
        HIPCHK(hipEventElapsedTime(&m, te[c], te[c + 1]));
        if (m < best) { best = m; pick = cands[c]; }
      }
      for (auto& e : te) (void)hipEventDestroy(e);
      return pick;
    };
    std::vector<int> slots(nb);
    for (int j = 0; j < nb; j++) slots[j] = j;
    if (ensure_cb(nb, tcap) != 0) { for (void* p : mem) pool_.put(p); rc = -3; break; }
    ShotDev* d_ds = d_ds_;
    Tile* d_tiles = d_tiles_;
    int* d_slots = d_slots_;
    StackShotHost* d_stk = d_stk_;
    for (auto& u : uploads) HIPCHK(hipMemcpyAsync(u.first, u.second.data(), u.second.size(), hipMemcpyHostToDevice, st_));
    for (int j = 0; j < nb; j++) { ds[j].phase = PH_FWD; ds[j].it_off = 0; ds[j].it_dir = 1; }
    HIPCHK(hipMemcpyAsync(d_ds, ds.data(), nb * sizeof(ShotDev), hipMemcpyHostToDevice, st_));
    HIPCHK(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, st_));
    HIPCHK(hipMemcpyAsync(d_slots, slots.data(), nb * 4, hipMemcpyHostToDevice, st_));
    for (int j = 0; j < nb; j++) {
      HIPCHK(rtm_launch_phase_init(d_ds, j, PH_FWD, 512, st_));
      if (smooth)                                    // the factors of the shot's smoothed windows
        HIPCHK(rtm_launch_model_prep(ds[j].rho, ds[j].vv, (float*)ds[j].mib, (float*)ds[j].mkk,
                                     (long long)setup[k0 + j].g.nxp * setup[k0 + j].g.nzp, st_));
    }
    if (opt.verbose)
      for (int j = 0; j < nb; j++) printf("N0. %d shot source wavefield calculating......\n", shots[k0 + j]);

    // ---- forward pass: propagation + record + boundary strips
    if (guard_clear() != 0) { for (void* p : mem) pool_.put(p); rc = -10; break; }
    HIPCHK(hipEventRecord(e0, st_));
    int fwd_done = 0;
    if (tune) {
      tz = tune_pass(PH_FWD);
      tuned_[tkey] = tz;
      fwd_done = 3 * kTuneSteps;
      build_tiles(ds, tz, tiles);
      layout_runs(tiles, PH_FWD);
      link_neighbours(tiles, tz);
      HIPCHK(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, st_));
    }
    rc = run_pass(PH_FWD, nt - fwd_done, d_ds, d_tiles, (int)tiles.size(), slots, d_slots, tz, opt.store_fmt,
                  opt.graph_chunk, (opt.profile && stats) ? stats : nullptr, guard_every, fwd_skip_ok(tz));
    if (rc) { for (void* p : mem) pool_.put(p); break; }
    HIPCHK(hipEventRecord(e1, st_));
    HIPCHK(hipStreamSynchronize(st_));
    if ((rc = guard_check(ds, shots, k0, "forward")) != 0) { for (void* p : mem) pool_.put(p); break; }
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    fwd_ms += ms;

    // seismograms: outputs, mode-0 files, mode-1 inputs
    std::vector<float> rec;
    for (int j = 0; j < nb && rc == 0; j++) {
      const ShotSetup& s = setup[k0 + j];
      const size_t nrv = (size_t)nt * s.g.nrec;
      const int ishot = shots[k0 + j];
      if (P.mode == 0) {
        if ((out && out->shot_seis && out->shot_seis[k0 + j]) || opt.write_files) {
          rec.resize(nrv);
          HIPCHK(hipMemcpy(rec.data(), ds[j].rec, nrv * 4, hipMemcpyDeviceToHost));
          if (out && out->shot_seis && out->shot_seis[k0 + j]) memcpy(out->shot_seis[k0 + j], rec.data(), nrv * 4);
          if (opt.write_files) {
            SuKey key;
            key.iTraceAll = s.hk_itraceall; key.iTrace = 1; key.Offset = s.hk_offset; key.sx = s.hk_sx;
            key.gx = s.hk_gx; key.Nt = s.hk_nt; key.Dt = s.hk_dt; key.Dx = s.hk_dx; key.Ntrace = s.hk_ntrace;
            char path[64];
            snprintf(path, sizeof(path), "./seisReal/%dseisreal.su", ishot);
            rtmio_write_seis(rec.data(), key, 0, P.sgyflag, P.endianflag, path);
          }
        }
      } else {
        rec.assign(nrv, 0.0f);
        if (opt.seis_wait && opt.seis_wait(opt.seis_user, k0 + j) != 0) {
          set_error("the input record of shot " + std::to_string(ishot) + " is not available");
          rc = -4;
          break;
        }
        if (out && out->shot_seis_in && out->shot_seis_in[k0 + j]) {
          memcpy(rec.data(), out->shot_seis_in[k0 + j], nrv * 4);
        } else {
          char path[64];
          snprintf(path, sizeof(path), "./seisReal/%dseisreal.su", ishot);
          SuKey key;
          if (rtmio_read_seis(rec.data(), &key, 0, P.sgyflag, path, (int)nrv) != 0) {
            set_error(std::string("cannot read ") + path);
            rc = -4;
            break;
          }
        }
        HIPCHK(hipMemcpy(ds[j].rec, rec.data(), nrv * 4, hipMemcpyHostToDevice));
      }
    }
    if (rc) { for (void* p : mem) pool_.put(p); break; }

    // the other imaging conditions of this batch: interior cells of its widest shot
    ic_cells_ = 0;
    if (icx)
      for (int j = 0; j < nb; j++) ic_cells_ = std::max(ic_cells_, nz * setup[k0 + j].g.nxl);

    // velsmooth1d of each shot's window before its reverse pass (rtm_model.cu:
    // 402-405): reconstruction and receiver propagation use the smoothed
    // model.  On the GPU, out of place through the slot's factor arrays (they
    // are rebuilt from the smoothed windows right after)
    if (P.ismth > 0) {
      std::vector<double> wt;
      double norm = 0;
      gauss_weights(P.ismth, wt, norm);
      double* d_wt = (double*)pool_.get(wt.size() * 8);
      if (!d_wt) { set_error("out of device memory (smoothing weights)"); for (void* p : mem) pool_.put(p); rc = -3; break; }
      HIPCHK(hipMemcpyAsync(d_wt, wt.data(), wt.size() * 8, hipMemcpyHostToDevice, st_));
      for (int j = 0; j < nb; j++) {
        ShotSetup& s = setup[k0 + j];
        const int nzw = s.g.nzp, nxw = s.g.nxp;
        const size_t wb = (size_t)nzw * nxw * 4;
        float* vv = (float*)ds[j].vv;
        float* rh = (float*)ds[j].rho;
        float* t1 = (float*)ds[j].mib;
        float* t2 = (float*)ds[j].mkk;
        HIPCHK(rtm_launch_gauss_smooth(vv, t1, nzw, nxw, P.ismth, d_wt, norm, st_));
        HIPCHK(rtm_launch_gauss_smooth(rh, t2, nzw, nxw, P.ismth, d_wt, norm, st_));
        HIPCHK(hipMemcpyAsync(vv, t1, wb, hipMemcpyDeviceToDevice, st_));
        HIPCHK(hipMemcpyAsync(rh, t2, wb, hipMemcpyDeviceToDevice, st_));
        HIPCHK(rtm_launch_model_prep(rh, vv, t1, t2, (long long)nzw * nxw, st_));
        if (opt.shot_field) {
          // a recorded gather is muted after the smoothing (rtm_real.cu:370-403),
          // so with the smoothed velocities
          HIPCHK(hipMemcpyAsync(s.wvv.data(), vv, wb, hipMemcpyDeviceToHost, st_));
          HIPCHK(hipStreamSynchronize(st_));
          mute_cuts(pk[k0 + j], vp0, s, s.wvv.data());
          HIPCHK(hipMemcpyAsync((void*)ds[j].cut, s.cut.data(), s.cut.size() * 4, hipMemcpyHostToDevice, st_));
          HIPCHK(hipStreamSynchronize(st_));
        }
      }
      HIPCHK(hipStreamSynchronize(st_));
      pool_.put(d_wt);
    }

    // ---- reverse pass: reconstruction + receiver propagation + imaging
    for (int j = 0; j < nb; j++) { ds[j].phase = PH_REV; ds[j].it_off = nt - 2; ds[j].it_dir = -1; }
    HIPCHK(hipMemcpyAsync(d_ds, ds.data(), nb * sizeof(ShotDev), hipMemcpyHostToDevice, st_));
    for (int j = 0; j < nb; j++) {
      HIPCHK(rtm_launch_phase_init(d_ds, j, PH_REV, 512, st_));
      HIPCHK(rtm_launch_seed(d_ds, j, ds[j].nrec, st_));
    }
    if (opt.verbose) printf("source wavefiled prepared over...\nBegin backward......\n");
    if (guard_clear() != 0) { for (void* p : mem) pool_.put(p); rc = -10; break; }
    HIPCHK(hipEventRecord(e1, st_));
    int rev_done = 0;
    if (tune_r) {
      tz_r = tune_pass(PH_REV);
      tuned_[rkey] = tz_r;
      rev_done = 3 * kTuneSteps;
    }
    build_tiles(ds, tz_r, tiles);
    layout_runs(tiles, PH_REV);
    HIPCHK(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, st_));
    rc = run_pass(PH_REV, nt - 1 - rev_done, d_ds, d_tiles, (int)tiles.size(), slots, d_slots, tz_r, opt.store_fmt,
                  opt.graph_chunk, (opt.profile && stats) ? stats : nullptr, guard_every);
    last_tz = tz;
    last_tz_r = tz_r;
    if (rc) { for (void* p : mem) pool_.put(p); break; }
    HIPCHK(hipEventRecord(e2, st_));
    HIPCHK(hipStreamSynchronize(st_));
    if ((rc = guard_check(ds, shots, k0, "reverse")) != 0) { for (void* p : mem) pool_.put(p); break; }
    HIPCHK(hipEventElapsedTime(&ms, e1, e2));
    rev_ms += ms;

    for (int j = 0; j < nb; j++) {
      const ShotSetup& s = setup[k0 + j];
      const size_t ni = (size_t)nz * s.g.nxl;
      if (out && out->shot_g1 && out->shot_g1[k0 + j]) HIPCHK(hipMemcpy(out->shot_g1[k0 + j], ds[j].g1, ni * 4, hipMemcpyDeviceToHost));
      if (out && out->shot_illum && out->shot_illum[k0 + j]) HIPCHK(hipMemcpy(out->shot_illum[k0 + j], ds[j].illum, ni * 4, hipMemcpyDeviceToHost));
      const double nxz = (double)s.g.nxp * s.g.nzp, inn = (double)nz * s.g.nxl;
      cells += 2.0 * nt * nxz + (double)(nt - 1) * inn;
      // algorithmic bytes: FD step 32 B/cell + 16 B per PML cell (forward and
      // receiver passes), reconstruction 32 B and imaging 16 B per interior
      // cell, boundary strips written once and read once
      const double pml = nxz - inn;
      const double strip = 40.0 * (nz + s.g.nxl) * (double)store_elem(opt.store_fmt);
      algb += 2.0 * nt * (32.0 * nxz + 16.0 * pml) + (double)(nt - 1) * inn * (32.0 + 16.0) + 2.0 * nt * strip;
      if (k0 == 0) {
        step_fwd += 32.0 * nxz + 16.0 * pml + strip;
        step_rev += 32.0 * nxz + 16.0 * pml + inn * (32.0 + 16.0) + strip;
      }
      fwdb += (double)nt * (32.0 * nxz + 16.0 * pml + strip);
      revb += (double)(nt - 1) * (32.0 * nxz + 16.0 * pml + inn * (32.0 + 16.0) + strip);
    }
    // ---- stack in shot order (FP64); temp images every 50 shots
    std::vector<StackShotHost> stk(nb);
    for (int j = 0; j < nb; j++) stk[j] = StackShotHost{ds[j].g1, ds[j].illum, (const float*)((const char*)mem[j] + small_layout(setup[k0 + j]).taper), setup[k0 + j].g.nx1, setup[k0 + j].g.nxl};
    HIPCHK(hipMemcpyAsync(d_stk, stk.data(), nb * sizeof(StackShotHost), hipMemcpyHostToDevice, st_));
    if (icx) {
      std::vector<StackIcHost> sic(nb);
      for (int j = 0; j < nb; j++) sic[j] = StackIcHost{ds[j].ic, stk[j].illum, stk[j].taper, stk[j].nx1, stk[j].nxl};
      HIPCHK(hipMemcpyAsync(d_stkic, sic.data(), nb * sizeof(StackIcHost), hipMemcpyHostToDevice, st_));
    }
    for (int j = 0; j < nb; j++) {
      HIPCHK(rtm_launch_stack(d_stk + j, 1, nx, nz, d_i1, d_i2, d_il, st_));
      if (icx) HIPCHK(rtm_launch_stack_ic(d_stkic + j, nx, nz, d_i7, st_));
      const int ishot = shots[k0 + j];
      if (opt.write_files && (ishot - 1) % 50 == 0) {
        std::vector<double> a(nm);
        std::vector<float> f(nm), zero(nm, 0.0f);
        struct { double* d; const char* name; } outs[] = {{d_i2, "Image2s"}, {d_i1, "Image1"}, {d_il, "Illum"}, {d_i1, "Image2m"}};
        for (auto& o : outs) {
          HIPCHK(hipMemcpyAsync(a.data(), o.d, nm * 8, hipMemcpyDeviceToHost, st_));
          HIPCHK(hipStreamSynchronize(st_));
          for (size_t i = 0; i < nm; i++) f[i] = (float)a[i];
          char path[64];
          snprintf(path, sizeof(path), "./output/%stemp.su", o.name);
          rtmio_write_image(f.data(), nz, nx, P.dx, path);
        }
        const char* icn[] = {"Image3s", "Image4s", "Image5s", "Image3m", "Image4m", "Image5m", "Imagetrue"};
        for (int k = 0; k < 7; k++) {
          const float* img = zero.data();
          if (icx) {
            HIPCHK(hipMemcpyAsync(a.data(), i7(k), nm * 8, hipMemcpyDeviceToHost, st_));
            HIPCHK(hipStreamSynchronize(st_));
            for (size_t i = 0; i < nm; i++) f[i] = (float)a[i];
            img = f.data();
          }
          char path[64];
          snprintf(path, sizeof(path), "./output/%stemp.su", icn[k]);
          rtmio_write_image(img, nz, nx, P.dx, path);
        }
      }
    }
    HIPCHK(hipStreamSynchronize(st_));
    if (opt.verbose)
      for (int j = 0; j < nb; j++) printf("%d shot finished: %g (s)\n", shots[k0 + j], 0.0);
    for (void* p : mem) pool_.put(p);
    ic_cells_ = 0;
    if (opt.checkpoint) {
      std::vector<double> st3(nstk * nm);
      for (int i = 0; i < nstk; i++)
        HIPCHK(hipMemcpy(st3.data() + (size_t)i * nm, stacks[i], nm * 8, hipMemcpyDeviceToHost));
      if (write_checkpoint(opt.checkpoint, job_fp, nx, nz, k1, st3) != 0) {
        set_error(std::string("cannot write checkpoint ") + opt.checkpoint);
        rc = -6;
        break;
      }
    }
    k0 = k1;
  }
  if (rc != 0) { pool_.put(common); (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2); return rc; }

  // ---- final images
  HIPCHK(rtm_launch_round(d_i1, d_f1, (long long)nm, st_));
  HIPCHK(rtm_launch_round(d_i2, d_f2, (long long)nm, st_));
  HIPCHK(rtm_launch_round(d_il, d_f3, (long long)nm, st_));
  if (out && icx) {
    float* dst[7] = {out->image3s, out->image4s, out->image5s, out->image3m, out->image4m, out->image5m, out->imagetrue};
    for (int k = 0; k < 7; k++) {
      if (!dst[k]) continue;
      HIPCHK(rtm_launch_round(i7(k), f7(k), (long long)nm, st_));
      HIPCHK(hipMemcpyAsync(dst[k], f7(k), nm * 4, hipMemcpyDeviceToHost, st_));
    }
  }
  if (out) {
    if (out->image1) HIPCHK(hipMemcpyAsync(out->image1, d_f1, nm * 4, hipMemcpyDeviceToHost, st_));
    if (out->image2s) HIPCHK(hipMemcpyAsync(out->image2s, d_f2, nm * 4, hipMemcpyDeviceToHost, st_));
    if (out->illum) HIPCHK(hipMemcpyAsync(out->illum, d_f3, nm * 4, hipMemcpyDeviceToHost, st_));
    if (out->image_lap || out->image_norm) {
      // reuse the FP64 buffer d_i2 as float scratch after image2s is copied out
      HIPCHK(hipStreamSynchronize(st_));
      float* scratch = (float*)d_i1;
      if (out->image_lap) {
        HIPCHK(rtm_launch_laplace(d_f1, scratch, nx, nz, (int)P.dx, (int)P.dz, st_));
        HIPCHK(hipMemcpyAsync(out->image_lap, scratch, nm * 4, hipMemcpyDeviceToHost, st_));
        HIPCHK(hipStreamSynchronize(st_));
      }
      if (out->image_norm) {
        HIPCHK(rtm_launch_norm_multi(d_f1, d_f3, scratch, (long long)nm, st_));
        HIPCHK(hipMemcpyAsync(out->image_norm, scratch, nm * 4, hipMemcpyDeviceToHost, st_));
      }
    }
  }
  HIPCHK(hipStreamSynchronize(st_));
  pool_.put(common);
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2);
  if (stats) {
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->fwd_ms = fwd_ms; stats->rev_ms = rev_ms; stats->gpu_ms = fwd_ms + rev_ms;
    stats->cell_updates = cells; stats->nshots = nsh; stats->batches = nbatches; stats->tile_z = last_tz;
    stats->tile_z_rev = last_tz_r;
    stats->batch = maxb; stats->hbm_bytes_alg = algb;
    stats->step_bytes_fwd = step_fwd; stats->step_bytes_rev = step_rev;
    stats->fwd_bytes_alg = fwdb; stats->rev_bytes_alg = revb;
  }
  return 0;
}

int Engine::copy_probe(size_t bytes, int reps, double* gbps)
{
  HIPCHK(hipSetDevice(device_));
  void* a = pool_.get(bytes);
  void* b = pool_.get(bytes);
  if (!a || !b) {
    if (a) pool_.put(a);
    if (b) pool_.put(b);
    set_error("out of device memory (copy probe)");
    return -3;
  }
  HIPCHK(hipMemsetAsync(a, 0, bytes, st_));
  HIPCHK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, st_));   // warm
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st_));
  for (int r = 0; r < reps; r++) HIPCHK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, st_));
  HIPCHK(hipEventRecord(e1, st_));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  pool_.put(a);
  pool_.put(b);
  *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
  return 0;
}

int Engine::fd_steps(const rtmh_params& P, int ishot, const float* vp0, const float* rho0, const float* coef,
                     int nsteps, int it0, int with_source, float* pf, float* vxf, float* vzf, float* p1f, float* p2f)
{
  std::string why;
  if (!params_ok(P, why)) { set_error(why); return -1; }
  if (it0 < 0 || it0 + nsteps > P.nt) { set_error("it0 + nsteps exceeds nt"); return -1; }
  HIPCHK(hipSetDevice(device_));
  ShotSetup s;
  if (!shot_geometry(P, ishot, s)) { set_error("shot window outside the model"); return -2; }
  const int nx = P.nx, nz = P.nz;
  const size_t nm = (size_t)nx * nz;
  std::vector<float> vv(nm), colmax(nx);
  for (size_t i = 0; i < nm; i++) vv[i] = vp0[i] * vp0[i];
  for (int ix = 0; ix < nx; ix++) {
    float m = vv[ix];
    for (int iz = 0; iz < nz; iz++) m = vv[(size_t)iz * nx + ix] > m ? vv[(size_t)iz * nx + ix] : m;
    colmax[ix] = m;
  }
  shot_setup(P, vp0, colmax.data(), s);
  taper_row(s.taper, s.g.nxl, 1);
  if (!with_source) { s.sx0 = 0; s.sx1 = -1; s.sz0 = 0; s.sz1 = -1; }
  std::vector<float> wav;
  ricker(wav, P.nt, P.fdom, P.dt);
  const size_t cb = 4 * align256(nm * 4) + align256(400) + align256((size_t)P.nt * 4);
  char* common = (char*)pool_.get(cb);
  const size_t sb = slot_bytes(s, P.nt, ST_F32);
  char* slot = (char*)pool_.get(sb);
  std::vector<Tile> tiles;
  const int tz = 32;
  ShotDev d;
  std::vector<std::pair<void*, std::vector<char>>> up;
  float* d_rho = (float*)common;
  float* d_vv = (float*)(common + align256(nm * 4));
  float* d_ib = (float*)(common + 2 * align256(nm * 4));
  float* d_kk = (float*)(common + 3 * align256(nm * 4));
  float* d_coef = (float*)(common + 4 * align256(nm * 4));
  float* d_wav = (float*)((char*)d_coef + align256(400));
  layout_slot(P, s, ST_F32, slot, d, up, d_coef, d_wav, ModelDev{d_rho, d_vv, d_ib, d_kk, nx});
  d.record = 0; d.phase = PH_FWD; d.it_off = it0; d.it_dir = 1;
  std::vector<ShotDev> dv(1, d);
  build_tiles(dv, tz, tiles);
  char* ctl = (char*)pool_.get(align256(sizeof(ShotDev)) + align256(tiles.size() * sizeof(Tile)));
  ShotDev* d_ds = (ShotDev*)ctl;
  Tile* d_tiles = (Tile*)(ctl + align256(sizeof(ShotDev)));
  HIPCHK(hipMemcpy(d_rho, rho0, nm * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_vv, vv.data(), nm * 4, hipMemcpyHostToDevice));
  HIPCHK(rtm_launch_model_prep(d_rho, d_vv, d_ib, d_kk, (long long)nm, st_));
  HIPCHK(hipStreamSynchronize(st_));
  HIPCHK(hipMemcpy(d_coef, coef, 400, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_wav, wav.data(), (size_t)P.nt * 4, hipMemcpyHostToDevice));
  for (auto& u : up) HIPCHK(hipMemcpy(u.first, u.second.data(), u.second.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_ds, &d, sizeof(ShotDev), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
  // fields in: host [nzp*nxp] -> device pitch
  const int nxp = s.g.nxp, nzp = s.g.nzp, pitch = d.pitch;
  // fields in at level it0&1, out of level (it0+nsteps)&1; the model window
  // and the zeroed levels come from the phase-start kernel
  HIPCHK(rtm_launch_phase_init(d_ds, 0, PH_FWD, 512, st_));
  HIPCHK(hipStreamSynchronize(st_));
  float* hf[5] = {pf, vxf, vzf, p1f, p2f};
  const int l0 = it0 & 1, l1 = (it0 + nsteps) & 1;
  float* din[5] = {d.sp[l0], d.svx[l0], d.svz[l0], d.q1, d.q2};
  float* dout[5] = {d.sp[l1], d.svx[l1], d.svz[l1], d.q1, d.q2};
  for (int k = 0; k < 5; k++)
    HIPCHK(hipMemcpy2D(din[k], (size_t)pitch * 4, hf[k], (size_t)nxp * 4, (size_t)nxp * 4, nzp, hipMemcpyHostToDevice));
  LaunchArgs a{d_ds, d_tiles, (int)tiles.size(), 0};
  for (int k = 0; k < nsteps; k++) {
    a.step = k;
    HIPCHK(rtm_launch_step(a, PH_FWD, tz, st_));
  }
  HIPCHK(hipStreamSynchronize(st_));
  for (int k = 0; k < 5; k++)
    HIPCHK(hipMemcpy2D(hf[k], (size_t)nxp * 4, dout[k], (size_t)pitch * 4, (size_t)nxp * 4, nzp, hipMemcpyDeviceToHost));
  pool_.put(ctl); pool_.put(slot); pool_.put(common);
  return 0;
}

}  // namespace rtm
