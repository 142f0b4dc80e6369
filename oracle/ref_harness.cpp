// ref_harness.cpp -- extern "C" wrappers around the reference's own host
// helpers, compiled together with their unmodified sources into
// oracle/_ref/libref_host.so (see the Makefile's `ref` target).  Used only in
// the build container, by tests/golden/make_golden.py to write fixtures and
// by tests that cross-check the restatements where the library exists.
#include <stdio.h>
#include <string.h>

#include "function.h"   // the reference's header (FD/IO/Velocity/image helpers)
#include "sgyhead.h"    // the reference's header (SU key header, IO)

extern "C" {

void ref_diff_coef(float* data, int order) { Diff_coeffVelStrLiuyang(data, order); }

void ref_ricker(float* w, float fdom, int nt, float dt, int flag) { Ricker(w, fdom, nt, dt, flag); }

void ref_window_pad(float* out, float* in, int nx, int nz, int nx1, int nx2, int npml, int square)
{
  int nxl = nx2 - nx1 + 1;
  float* tmp = new float[(size_t)nxl * nz];
  if (square) extractvel1(tmp, in, nx, nz, nx1, nx2);
  else extractrho1(tmp, in, nx, nz, nx1, nx2);
  extendvel1(out, tmp, nxl, nz, npml);
  delete[] tmp;
}

void ref_interpvel(float* vp0, float* vp, int nx0, int nz0, int nx, int nz, int lx, int lz)
{
  interpvel1(vp0, vp, nx0, nz0, nx, nz, lx, lz);
}

void ref_laplacian(int nx, int nz, int dx, int dz, float* image) { laplacefd1d(nx, nz, dx, dz, image); }

void ref_scaleimage(float* illum, float* grad, int n) { scaleimage1d(illum, grad, n); }

void ref_pmlvelsmooth(float* vp, int nx, int nz, int npml) { pmlvelsmooth1d(vp, nx, nz, npml); }

void ref_velsmooth(float* vp, int n1, int n2, int nsp) { velsmooth1d(vp, n1, n2, nsp); }

float ref_maxval(float* v, int n) { return Maxval1(v, n); }
float ref_sum1abs(float* v, int n) { return sum1abs(v, n); }

int ref_write_seis(float* seis, int itraceall, int itrace, int offset, int sx, int gx, int nt, int dt_us, float dx,
                   int ntrace, int outflag, int sgyflag, int endianflag, const char* path)
{
  sgyheadkey k;
  k.iTraceAll = itraceall; k.iTrace = itrace; k.Offset = offset; k.sx = sx; k.gx = gx;
  k.Nt = (short)nt; k.Dt = (short)dt_us; k.Dx = dx; k.Ntrace = ntrace;
  char buff[40];
  if (strlen(path) >= sizeof(buff)) return -1;
  strcpy(buff, path);
  Outseisfile1d(seis, k, outflag, sgyflag, endianflag, buff);
  return 0;
}

int ref_read_seis(float* seis, int* nt, int* ntrace, int sgyflag, const char* path)
{
  sgyheadkey k;
  char buff[40];
  if (strlen(path) >= sizeof(buff)) return -1;
  strcpy(buff, path);
  Inseisfile1d(seis, &k, 0, sgyflag, buff);
  *nt = k.Nt;
  *ntrace = k.Ntrace;
  return 0;
}

int ref_write_image(float* image, int nz, int nx, float d, const char* path)
{
  char buff[40];
  if (strlen(path) >= sizeof(buff)) return -1;
  strcpy(buff, path);
  Output1d(image, nz, nx, d, buff, 1);
  return 0;
}

}  // extern "C"
