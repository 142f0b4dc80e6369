// host_driver.cpp -- the drop-in driver of the reference program rtm_gpu
// (RTM_GPU_velstr.cpp): parameter file, grid/time-step guard, model
// interpolation and density, selfdoc banner, coefficient table, dispatch to
// rtm_model() and the .su image outputs.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/rtm_hip.h"
#include "host_io.h"
#include "rtm_engine.h"
#include "rtm_launch.h"

// Liu-Yang optimised staggered-grid coefficients, orders 2..20
// (FD_function.cpp:73-165); row r-1 holds order 2r as assembled at
// RTM_GPU_velstr.cpp:184-196.
extern "C" void rtmh_coef_table(float* out)
{
  static const float t[10][10] = {
      {1.0f},
      {0.1129042e+1f, -0.4301412e-1f},
      {0.1185991e+1f, -0.7249965e-1f, 0.6301572e-2f},
      {0.1217990e+1f, -0.9382142e-1f, 0.1507536e-1f, -0.1700324e-2f},
      {0.1236607e+1f, -0.1082265e+0f, 0.2343440e-1f, -0.5033546e-2f, 0.6817483e-3f},
      {0.1247662e+1f, -0.1175538e+0f, 0.2997970e-1f, -0.8719078e-2f, 0.2215897e-2f, -0.3462075e-3f},
      {0.1254799e+1f, -0.1238928e+0f, 0.3494371e-1f, -0.1208897e-1f, 0.4132531e-2f, -0.1197110e-2f,
       0.2122227e-3f},
      {0.1259312e+1f, -0.1280347e+0f, 0.3841945e-1f, -0.1473229e-1f, 0.5924913e-2f, -0.2248618e-2f,
       0.7179226e-3f, -0.1400855e-3f},
      {0.1262502e+1f, -0.1310244e+0f, 0.4103928e-1f, -0.1686807e-1f, 0.7530520e-2f, -0.3345071e-2f,
       0.1380367e-2f, -0.4808410e-3f, 0.1023759e-3f},
      {0.1264748e+1f, -0.1331606e+0f, 0.4296909e-1f, -0.1851897e-1f, 0.8861071e-2f, -0.4347073e-2f,
       0.2076101e-2f, -0.9164925e-3f, 0.3437446e-3f, -0.7874250e-4f},
  };
  for (int r = 0; r < 10; r++)
    for (int i = 0; i < 10; i++) out[r * 10 + i] = t[r][i];
}

// parameter file reader (RTM_GPU_velstr.cpp:60-120): each value group is
// preceded by a comment line; the reference reads it with fgets/fscanf.
extern "C" int rtmh_read_parfile(const char* path, rtmh_parfile* pf)
{
  memset(pf, 0, sizeof(*pf));
  FILE* fp = fopen(path, "r");
  if (!fp) return -1;
  char buff[1024];
  int ok = 1;
#define SKIP2 do { ok &= fgets(buff, sizeof(buff), fp) != NULL; ok &= fgets(buff, sizeof(buff), fp) != NULL; } while (0)
  ok &= fgets(buff, sizeof(buff), fp) != NULL;
  ok &= fscanf(fp, "%f %d %f", &pf->dt0, &pf->nt0, &pf->tsnap) == 3; SKIP2;
  ok &= fscanf(fp, "%f %f %f %f", &pf->xmax, &pf->zmax, &pf->dx0, &pf->dz0) == 4; SKIP2;
  ok &= fscanf(fp, "%f %f %f", &pf->fdom, &pf->amp, &pf->alp) == 3; SKIP2;
  ok &= fscanf(fp, "%f %f %f", &pf->spx0, &pf->spz0, &pf->dspx) == 3; SKIP2;
  ok &= fscanf(fp, "%f %f", &pf->offset, &pf->offsetmax) == 2; SKIP2;
  ok &= fscanf(fp, "%d %d %d %d", &pf->startshot, &pf->nshot, &pf->dshot, &pf->nshotall) == 4; SKIP2;
  ok &= fscanf(fp, "%d", &pf->npml) == 1; SKIP2;
  ok &= fscanf(fp, "%d %d", &pf->medium_flag, &pf->ismth) == 2; SKIP2;
  ok &= fscanf(fp, "%f %f", &pf->direct, &pf->direct0) == 2; SKIP2;
  ok &= fscanf(fp, "%d", &pf->idevice) == 1; SKIP2;
  ok &= fscanf(fp, "%d", &pf->flag_layout) == 1; SKIP2;
  ok &= fscanf(fp, "%d", &pf->mode) == 1; SKIP2;
  ok &= fscanf(fp, "%d", &pf->dataflag) == 1; SKIP2;
  ok &= fscanf(fp, "%d", &pf->sgyflag) == 1; SKIP2;
  ok &= fscanf(fp, "%d", &pf->endianflag) == 1; SKIP2;
  ok &= fscanf(fp, "%d", &pf->rhoflag) == 1; SKIP2;
  ok &= fscanf(fp, "%255s", pf->vp_file) == 1;
  if (pf->dataflag == 1) { SKIP2; ok &= fscanf(fp, "%255s", pf->shotgather) == 1; }
  if (pf->rhoflag == 1) { SKIP2; ok &= fscanf(fp, "%255s", pf->rho_file) == 1; }
#undef SKIP2
  fclose(fp);
  return ok ? 0 : -2;
}

// get_constant (RTM_GPU_velstr.cpp:319-362): dispersion guard may halve the
// spacing (doubling the sample count), the stability guard may shrink dt.
extern "C" void rtmh_get_constant(const float* vp, float fdom, float dx0, float dz0, int nx0, int nz0, int nt0,
                                  float dt0, rtmh_grid* g)
{
  g->dx = dx0; g->dz = dz0; g->nx = nx0; g->nz = nz0; g->nt = nt0; g->dt = dt0; g->lx = 1; g->lz = 1;
  const int n = nx0 * nz0;
  float vmax = vp[0], vmin = vp[0];
  for (int i = 0; i < n; i++) { vmax = vp[i] > vmax ? vp[i] : vmax; vmin = vp[i] < vmin ? vp[i] : vmin; }
  g->vpmax = vmax; g->vpmin = vmin;
  const float dx_max = (float)((double)(vmin / fdom) / 7.25);
  const float dz_max = dx_max;
  while (dx_max < g->dx) { g->dx = g->dx / 2; g->lx *= 2; g->nx = (g->nx - 1) * 2 + 1; }
  while (dz_max < g->dz) { g->dz = g->dz / 2; g->lz *= 2; g->nz = (g->nz - 1) * 2 + 1; }
  const float hmin = g->dx < g->dz ? g->dx : g->dz;
  float c[100];
  rtmh_coef_table(c);
  float s = 0;
  for (int i = 0; i < 10; i++) s = (float)((double)s + fabs((double)c[90 + i]));   // sum1abs
  const float dt_max = hmin / (sqrtf(2.0f) * s * vmax);
  const float t = nt0 * dt0;
  if (dt_max < dt0) {
    g->dt = (float)((int)(dt_max * 10000.0) * 0.0001);
    g->nt = (int)(t / g->dt);
  }
}

// pmlvelsmooth1d (Velocity_function.cpp:9-103).  The relaxation updates the
// working copy in place, so the sweep order below is part of the result: the
// left/right layers column by column, then the top/bottom layers, then the
// edge rows/columns (their 4-point average includes the cell itself), then the
// four corners.  The change test compares the sweep's result with the window
// before it; fabs of a float is the float overload in the reference's C++.
extern "C" void rtmh_pml_smooth(float* w, int nx, int nz, int npml)
{
  auto at = [nx](std::vector<float>& a, int iz, int ix) -> float& { return a[(size_t)iz * nx + ix]; };
  for (int ix = 0; ix < npml; ix++)
    for (int iz = 1; iz < nz - 1; iz++) {
      w[(size_t)iz * nx + ix] = w[(size_t)iz * nx + 2 * npml - ix - 1];
      w[(size_t)iz * nx + nx - npml + ix] = w[(size_t)iz * nx + nx - npml - ix - 1];
    }
  for (int iz = 0; iz < npml; iz++)
    for (int ix = 1; ix < nx - 1; ix++) {
      w[(size_t)iz * nx + ix] = w[(size_t)(2 * npml - iz - 1) * nx + ix];
      w[(size_t)(nz - npml + iz) * nx + ix] = w[(size_t)(nz - npml - iz - 1) * nx + ix];
    }
  const size_t n = (size_t)nx * nz;
  std::vector<float> v(w, w + n);
  // (float)(0.25 * (double)sum): the reference's double literal
  auto avg = [](float a, float b, float c, float d) { return (float)(0.25 * (double)(a + b + c + d)); };
  float err = 1.0f;
  for (int sweep = 0; sweep <= 1000 && (double)err > 0.0005; sweep++) {
    for (int ix = 1; ix < npml; ix++)
      for (int iz = 1; iz < nz - 1; iz++) {
        const int xr = ix + nx - npml - 1;
        at(v, iz, ix) = avg(at(v, iz, ix + 1), at(v, iz, ix - 1), at(v, iz + 1, ix), at(v, iz - 1, ix));
        at(v, iz, xr) = avg(at(v, iz, xr + 1), at(v, iz, xr - 1), at(v, iz + 1, xr), at(v, iz - 1, xr));
      }
    for (int ix = 1; ix < nx - 1; ix++)
      for (int iz = 1; iz < npml; iz++) {
        const int zb = iz + nz - npml - 1;
        at(v, iz, ix) = avg(at(v, iz, ix + 1), at(v, iz, ix - 1), at(v, iz + 1, ix), at(v, iz - 1, ix));
        at(v, zb, ix) = avg(at(v, zb, ix + 1), at(v, zb, ix - 1), at(v, zb + 1, ix), at(v, zb - 1, ix));
      }
    for (int iz = 1; iz < nz - 1; iz++) {
      at(v, iz, 0) = avg(at(v, iz, 0), at(v, iz, 1), at(v, iz + 1, 0), at(v, iz - 1, 0));
      at(v, iz, nx - 1) = avg(at(v, iz, nx - 1), at(v, iz, nx - 2), at(v, iz + 1, nx - 1), at(v, iz - 1, nx - 1));
    }
    for (int ix = 1; ix < nx - 1; ix++) {
      at(v, 0, ix) = avg(at(v, 0, ix), at(v, 1, ix), at(v, 0, ix - 1), at(v, 0, ix + 1));
      at(v, nz - 1, ix) = avg(at(v, nz - 1, ix), at(v, nz - 2, ix), at(v, nz - 1, ix - 1), at(v, nz - 1, ix + 1));
    }
    at(v, 0, 0) = avg(at(v, 0, 0), at(v, 1, 0), at(v, 0, 1), at(v, 1, 1));
    at(v, nz - 1, 0) = avg(at(v, nz - 1, 0), at(v, nz - 2, 0), at(v, nz - 1, 1), at(v, nz - 2, 1));
    at(v, 0, nx - 1) = avg(at(v, 0, nx - 1), at(v, 0, nx - 2), at(v, 1, nx - 1), at(v, 1, nx - 2));
    at(v, nz - 1, nx - 1) = avg(at(v, nz - 1, nx - 1), at(v, nz - 2, nx - 1), at(v, nz - 1, nx - 2), at(v, nz - 2, nx - 2));
    float worst = -1.0e+5f;
    for (size_t i = 0; i < n; i++) {
      const float rel = fabsf(w[i] - v[i]) / fabsf(w[i]);
      if (rel > worst) worst = rel;
    }
    err = worst;
    std::copy(v.begin(), v.end(), w);
  }
}

// velsmooth1d (Velocity_function.cpp:199-313).  The weights depend only on
// the offset, so they are tabulated once (the same exp arguments, so the same
// values); each cell's weighted sum runs over the window in row order in
// double, as the reference's, and is divided by the sum of the weights.
extern "C" void rtmh_gauss_smooth(float* w, int nz, int nx, int nsp)
{
  if (nsp <= 0) return;
  const int k = 2 * nsp + 1, nxe = nx + 2 * nsp, nze = nz + 2 * nsp;
  std::vector<double> wt;
  double norm = 0;
  rtm::gauss_weights(nsp, wt, norm);                  // row order, as the reference sums them
  // edge-extended copy: each extended cell takes the nearest window cell
  std::vector<float> e((size_t)nze * nxe);
  for (int iz = 0; iz < nze; iz++) {
    const int sz = std::min(std::max(iz - nsp, 0), nz - 1);
    for (int ix = 0; ix < nxe; ix++) e[(size_t)iz * nxe + ix] = w[(size_t)sz * nx + std::min(std::max(ix - nsp, 0), nx - 1)];
  }
  std::vector<float> out((size_t)nz * nx);
  for (int iz = 0; iz < nz; iz++)
    for (int ix = 0; ix < nx; ix++) {
      double s = 0;
      for (int i = 0; i < k; i++) {
        const float* row = &e[(size_t)(iz + i) * nxe + ix];
        const double* wr = &wt[(size_t)i * k];
        for (int j = 0; j < k; j++) s += wr[j] * row[j];
      }
      out[(size_t)iz * nx + ix] = (float)s;
    }
  for (size_t i = 0; i < out.size(); i++) w[i] = (float)(out[i] / norm);
}

// interpvel1 (Velocity_function.cpp:1946-1977): linear refinement by lx, lz
extern "C" void rtmh_interp_model(const float* vp0, float* vp, int nx0, int nz0, int nx, int nz, int lx, int lz)
{
  (void)nz;
  std::vector<float> vel((size_t)nx * nz0);
  for (int iz = 0; iz < nz0; iz++) {
    for (int ix = 0; ix < nx0 - 1; ix++)
      for (int il = 0; il < lx; il++)
        vel[(size_t)iz * nx + ix * lx + il] =
            ((lx - il) * vp0[(size_t)iz * nx0 + ix] + il * vp0[(size_t)iz * nx0 + ix + 1]) / lx;
    vel[(size_t)iz * nx + (nx0 - 1) * lx] = vp0[(size_t)iz * nx0 + nx0 - 1];
  }
  for (int ix = 0; ix < nx; ix++) {
    for (int iz = 0; iz < nz0 - 1; iz++)
      for (int il = 0; il < lz; il++)
        vp[(size_t)(iz * lz + il) * nx + ix] =
            ((lz - il) * vel[(size_t)iz * nx + ix] + il * vel[(size_t)(iz + 1) * nx + ix]) / lz;
    vp[(size_t)(nz0 - 1) * lz * nx + ix] = vel[(size_t)(nz0 - 1) * nx + ix];
  }
}

// selfdoc banner (RTM_GPU_velstr.cpp:155-182)
extern "C" void rtmh_selfdoc(const rtmh_parfile* pf, int nx0, int nz0, const rtmh_grid* g)
{
  printf("=====================================================\n");
  printf("=====================================================\n");
  printf("Prestack Reverse Time Migration by Bingluo Gu from UPC.\n");
  printf("            version 4.0  date 2015.11.25               \n");
  printf("=====================================================\n");
  printf("=====================================================\n");
  printf("================Required parameters:=================\n");
  printf(" velfile=   %s\n", pf->vp_file);
  printf(" nshot  =   %d\t", pf->nshot); printf(" fshot  =   %d\t", pf->startshot); printf(" dshot  =   %d\n", pf->dshot);
  printf(" nz     =   %d\t", nz0); printf(" nx     =   %d\t", nx0); printf(" nt     =   %d\n", pf->nt0);
  printf(" dz     =   %.1f\t", pf->dz0); printf(" dx     =   %.1f\t", pf->dx0); printf(" dt     =   %.4f\n", pf->dt0);
  printf(" t      =   %f\n", pf->nt0 * pf->dt0);
  printf(" vmin   =   %.2f\n", g->vpmin);
  printf(" vmax   =   %.2f\n", g->vpmax);
  printf(" fdom      =  %f  dominant frequency\n", pf->fdom);
  printf(" npml      =  %2d  pml absorb layers number\n", pf->npml);
  printf(" Norder    =  %2d  order of finite difference\n", 20);
  printf(" dataflag  =  %2d  0: model data:                 1: real data\n", pf->dataflag);
  printf(" mode      =  %2d  0: generate seismic profile;   1: seismic profile exist\n", pf->mode);
  printf(" format:      %2d  0: su;                         1: sgy\n", pf->sgyflag);
  printf(" Endian:      %2d  0: little edition;             1: big edition\n", pf->endianflag);
  printf(" layout:      %2d  1: Bilateral;  2: full array;  3: Unilateral(r);  4: Unilateral(l)\n", pf->flag_layout);
  printf("=================Optional parameters:================\n");
  printf(" nz     =   %d\t", g->nz); printf(" nx   =   %d\t", g->nx); printf(" nt   =   %d\n", g->nt);
  printf(" dz     =   %.1f\t", g->dz); printf(" dx   =   %.1f\t", g->dx); printf(" dt   =   %.4f\n", g->dt);
  printf(" lz     =   %d\t", g->lz); printf(" lx   =   %d\n", g->lx);
  printf(" t      =   %f\n", g->nt * g->dt);
  printf("=======================================================\n");
}

// Laplacian-filtered and multi-shot normalised images on the GPU
extern "C" int rtmh_postprocess(int device, const float* image1, const float* illum, int nx, int nz, float dx,
                                float dz, float* image_lap, float* image_norm)
{
  if (hipSetDevice(device) != hipSuccess) { rtm::set_error("hipSetDevice failed"); return -1; }
  const size_t n = (size_t)nx * nz;
  float *a = nullptr, *b = nullptr, *c = nullptr;
  if (hipMalloc(&a, n * 4) != hipSuccess || hipMalloc(&b, n * 4) != hipSuccess || hipMalloc(&c, n * 4) != hipSuccess) {
    (void)hipFree(a); (void)hipFree(b);
    (void)hipGetLastError();                         // reported here, not by a later launch
    rtm::set_error("out of device memory");
    return -3;
  }
  int rc = 0;
  if (hipMemcpy(a, image1, n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b, illum, n * 4, hipMemcpyHostToDevice) != hipSuccess) rc = -10;
  if (!rc && image_lap) {
    if (rtm_launch_laplace(a, c, nx, nz, (int)dx, (int)dz, 0) != hipSuccess ||
        hipMemcpy(image_lap, c, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = -10;
  }
  if (!rc && image_norm) {
    if (rtm_launch_norm_multi(a, b, c, (long long)n, 0) != hipSuccess ||
        hipMemcpy(image_norm, c, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = -10;
  }
  if (rc) rtm::set_error("postprocess kernels failed");
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(c);
  return rc;
}

// main() of RTM_GPU_velstr.cpp:37-317
extern "C" int rtmh_driver_main(int argc, char** argv)
{
  (void)argc; (void)argv;
  char parname[512];
  {
    FILE* fp = fopen("./input/cmd.in", "r");
    if (!fp) { printf("cmd.in is not exist !\n"); return 1; }
    char buff[1024];
    int ok = fgets(buff, sizeof(buff), fp) != NULL;
    ok &= fscanf(fp, "%500s", parname) == 1;
    fclose(fp);
    if (!ok) { printf("cmd.in is malformed !\n"); return 1; }
  }
  char path[600];
  snprintf(path, sizeof(path), "./input/%s", parname);
  rtmh_parfile pf;
  const int prc = rtmh_read_parfile(path, &pf);
  if (prc == -1) { printf("%s is not exist !\n", parname); return 1; }
  if (prc != 0) { printf("%s is malformed !\n", parname); return 1; }
  const int nx0 = (int)(pf.xmax / pf.dx0) + 1;
  const int nz0 = (int)(pf.zmax / pf.dz0) + 1;
  if (pf.spx0 < 0 || pf.spz0 < 0 || pf.spx0 + (pf.nshot - 1) * pf.dspx > (nx0 - 1) * pf.dx0 ||
      pf.spx0 + (pf.nshot - 1) * pf.dspx < 0) {
    printf("the shot position out of the model !\nplease check the parameter !\n");
    return 1;
  }
  std::vector<float> vp0((size_t)nx0 * nz0);
  snprintf(path, sizeof(path), "./input/%s", pf.vp_file);
  if (rtmio_read_velocity(path, nx0, nz0, vp0.data()) != 0) { printf("The file %s open failed !\n", pf.vp_file); return 1; }
  rtmh_grid g;
  rtmh_get_constant(vp0.data(), pf.fdom, pf.dx0, pf.dz0, nx0, nz0, pf.nt0, pf.dt0, &g);
  const int nx = g.nx, nz = g.nz;
  const size_t n = (size_t)nx * nz;
  std::vector<float> vp(n), rho(n);
  rtmh_interp_model(vp0.data(), vp.data(), nx0, nz0, nx, nz, g.lx, g.lz);
  if (pf.rhoflag) {
    snprintf(path, sizeof(path), "./input/%s", pf.rho_file);
    if (rtmio_read_velocity(path, nx0, nz0, vp0.data()) != 0) { printf("The file %s open failed !\n", pf.rho_file); return 1; }
    float m = vp0[0];
    for (float v : vp0) m = v > m ? v : m;
    const int rhomax = (int)m;                                  // int, RTM_GPU_velstr.cpp:141-142
    if (rhomax < 10.0)
      for (float& v : vp0) v *= 1000.0;
    rtmh_interp_model(vp0.data(), rho.data(), nx0, nz0, nx, nz, g.lx, g.lz);
  } else {
    for (float& v : rho) v = 2000.0f;
  }
  rtmh_selfdoc(&pf, nx0, nz0, &g);
  float diffcoef[100];
  rtmh_coef_table(diffcoef);
  std::vector<float> wavelet(g.nt, 0.0f);
  std::vector<float> img[10];
  for (auto& v : img) v.assign(n, 0.0f);
  float *image2m = img[0].data(), *image3m = img[1].data(), *image4m = img[2].data(), *image5m = img[3].data();
  float *image1 = img[4].data(), *image2s = img[5].data(), *image3s = img[6].data(), *image4s = img[7].data();
  float *image5s = img[8].data(), *Illum = img[9].data();
  if (!pf.dataflag) {
    rtm_model(pf.idevice, pf.nshot, pf.startshot, pf.dshot, pf.mode, pf.medium_flag, pf.flag_layout, pf.sgyflag,
              pf.endianflag, pf.ismth, nx, nz, g.nt, g.dx, g.dz, g.dt, pf.npml, pf.direct, pf.direct0, pf.tsnap,
              pf.fdom, pf.spx0, pf.spz0, pf.dspx, pf.offset, pf.offsetmax, pf.amp, pf.alp, vp.data(), rho.data(),
              diffcoef, wavelet.data(), image1, image2s, image3s, image4s, image5s, image2m, image3m, image4m,
              image5m, Illum);
  } else {
    // shot table of the gather file, scanned once and kept in
    // shotprofileId.dat (RTM_GPU_velstr.cpp:234-252)
    const int nall = pf.nshotall > 0 ? pf.nshotall : 1;
    std::vector<int> prof(4 * (size_t)nall, 0);
    int* rows[4] = {prof.data(), prof.data() + nall, prof.data() + 2 * (size_t)nall, prof.data() + 3 * (size_t)nall};
    int endian = pf.endianflag;
    FILE* tf = fopen("shotprofileId.dat", "r");
    if (!tf) {
      snprintf(path, sizeof(path), "./seisReal/%s", pf.shotgather);
      int found = 0;
      if (rtmh_scan_gather(path, pf.sgyflag, &endian, prof.data(), nall, &found) != 0) {
        printf("%s\n", rtmh_last_error());
        return 1;
      }
      tf = fopen("shotprofileId.dat", "w");
      if (tf) {
        for (int i = 0; i < nall; i++) fprintf(tf, "%d %d %d %d\n", rows[0][i], rows[1][i], rows[2][i], rows[3][i]);
        fclose(tf);
      }
    } else {
      for (int i = 0; i < nall; i++)
        if (fscanf(tf, "%d %d %d %d", &rows[0][i], &rows[1][i], &rows[2][i], &rows[3][i]) != 4) break;
      fclose(tf);
    }
    // the mute's width and minimum length in samples (rtm_real.cu:28-29)
    const int nw = (int)(pf.direct / (pf.fdom * g.dt)), tlength = (int)(pf.direct0 / g.dt);
    rtm_real(pf.idevice, pf.nshot, pf.startshot, pf.dshot, pf.medium_flag, pf.flag_layout, pf.sgyflag, endian,
             pf.ismth, nx, nz, g.nt, g.dx, g.dz, g.dt, nx0, nz0, pf.nt0, pf.dx0, pf.dz0, pf.dt0, pf.npml, g.lx, g.lz,
             pf.tsnap, pf.fdom, pf.spz0, pf.amp, pf.alp, nw, tlength, vp.data(), rho.data(), diffcoef,
             wavelet.data(), pf.shotgather, rows, image1, image2s, image3s, image4s, image5s, image2m, image3m,
             image4m, image5m, Illum);
  }
  // image outputs (RTM_GPU_velstr.cpp:259-297)
  struct { const float* im; const char* name; } outs[] = {
      {Illum, "Illum"}, {image2m, "Image2m"}, {image3m, "Image3m"}, {image4m, "Image4m"}, {image5m, "Image5m"},
      {image1, "Image1"}, {image2s, "Image2s"}, {image3s, "Image3s"}, {image4s, "Image4s"}, {image5s, "Image5s"}};
  for (auto& o : outs) {
    snprintf(path, sizeof(path), "./output/%d%s.su", pf.idevice, o.name);
    rtmio_write_image(o.im, nz, nx, g.dz, path);
  }
  // the Laplacian-filtered cross-correlation image and the multi-shot
  // normalised image (image_function.cpp:44-79, kernel.cu:1580-1590)
  std::vector<float> lap(n), nrm(n);
  if (rtmh_postprocess(pf.idevice, image1, Illum, nx, nz, g.dx, g.dz, lap.data(), nrm.data()) == 0) {
    snprintf(path, sizeof(path), "./output/%dImage1lap.su", pf.idevice);
    rtmio_write_image(lap.data(), nz, nx, g.dz, path);
    snprintf(path, sizeof(path), "./output/%dImagenorm.su", pf.idevice);
    rtmio_write_image(nrm.data(), nz, nx, g.dz, path);
  } else {
    fprintf(stderr, "rtm_gpu: %s\n", rtmh_last_error());
    return 1;
  }
  printf("The program calculate over !\n");
  return 0;
}
