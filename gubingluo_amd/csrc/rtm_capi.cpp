// rtm_capi.cpp -- the C ABI of librtm_hip.so (declared in include/rtm_hip.h).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <memory>
#include <vector>

#include "../../include/rtm_hip.h"
#include "host_io.h"
#include "rtm_engine.h"

using rtm::Engine;

struct rtmh_ctx {
  std::unique_ptr<Engine> eng;
};

rtm::Engine* rtmh_ctx_engine(rtmh_ctx* ctx) { return ctx->eng.get(); }

extern "C" const char* rtmh_last_error(void) { return rtm::get_error(); }
extern "C" const char* rtmh_version(void) { return "rtm-mi355x 0.1 (gfx950)"; }

extern "C" int rtmh_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int rtmh_create(int device, rtmh_ctx** out)
{
  if (!out) return -1;
  int n = rtmh_device_count();
  if (device < 0 || device >= n) {
    rtm::set_error("no such HIP device (" + std::to_string(device) + " of " + std::to_string(n) + ")");
    return -1;
  }
  rtmh_ctx* c = new rtmh_ctx;
  c->eng.reset(new Engine(device));
  *out = c;
  return 0;
}

extern "C" int rtmh_destroy(rtmh_ctx* ctx)
{
  delete ctx;
  return 0;
}

extern "C" int rtmh_shot_geometry(const rtmh_params* p, int ishot, rtmh_geom* g)
{
  rtm::ShotSetup s;
  if (!p || !g || !rtm::shot_geometry(*p, ishot, s)) return -1;
  *g = s.g;
  return 0;
}

extern "C" int rtmh_gather_geometry(const rtmh_params* p, int ishot, const float* field, rtmh_geom* g)
{
  rtm::ShotSetup s;
  if (!p || !g || !field || !rtm::shot_geometry(*p, ishot, s, field)) return -1;
  *g = s.g;
  return 0;
}

extern "C" int rtmh_run(rtmh_ctx* ctx, const rtmh_params* p, const float* vp0, const float* rho0, const float* diffcoef,
                        const rtmh_options* opt, rtmh_outputs* out, rtmh_stats* stats)
{
  if (!ctx || !p || !vp0 || !rho0 || !diffcoef) { rtm::set_error("null argument"); return -1; }
  rtmh_options o;
  memset(&o, 0, sizeof(o));
  o.graph_chunk = 0;
  if (opt) o = *opt;
  return ctx->eng->run(*p, vp0, rho0, diffcoef, o, out, stats);
}

extern "C" int rtmh_model_create(rtmh_ctx* ctx, int nx, int nz, const float* vp0, const float* rho0, int* model_id)
{
  if (!ctx || !model_id) { rtm::set_error("null argument"); return -1; }
  return ctx->eng->model_create(nx, nz, vp0, rho0, model_id);
}

extern "C" int rtmh_model_destroy(rtmh_ctx* ctx, int model_id)
{
  if (!ctx) { rtm::set_error("null argument"); return -1; }
  return ctx->eng->model_destroy(model_id);
}

extern "C" int rtmh_run_model(rtmh_ctx* ctx, int model_id, const rtmh_params* p, const float* diffcoef,
                              const rtmh_options* opt, rtmh_outputs* out, rtmh_stats* stats)
{
  if (!ctx || !p || !diffcoef) { rtm::set_error("null argument"); return -1; }
  rtmh_options o;
  memset(&o, 0, sizeof(o));
  if (opt) o = *opt;
  return ctx->eng->run_model(model_id, *p, diffcoef, o, out, stats);
}

extern "C" int rtmh_smooth_gpu(rtmh_ctx* ctx, float* w, int nx, int nz, int npml, int nsp)
{
  if (!ctx || !w || nx < 2 || nz < 2 || npml < 0 || nsp < 0) { rtm::set_error("bad argument"); return -1; }
  std::vector<float*> ws{w};
  std::vector<int> vx{nx}, vz{nz};
  if (npml > 0) {
    if (!rtm::pml_smooth_on_gpu(nx, nz, npml)) {
      rtm::set_error("window outside the GPU layer smoother's shapes (npml <= 65, nx, nz in [3 npml, 4096])");
      return -1;
    }
    const int rc = ctx->eng->pml_smooth_windows(ws, vx, vz, npml);
    if (rc) return rc;
  }
  return nsp > 0 ? ctx->eng->gauss_smooth_windows(ws, vx, vz, nsp) : 0;
}

extern "C" int rtmh_copy_probe(rtmh_ctx* ctx, double bytes, int reps, double* gbps)
{
  if (!ctx || !gbps || bytes <= 0 || reps <= 0) {
    rtm::set_error("bad argument");
    return -1;
  }
  return ctx->eng->copy_probe((size_t)bytes, reps, gbps);
}

extern "C" int rtmh_fd_steps(rtmh_ctx* ctx, const rtmh_params* p, int ishot, const float* vp0, const float* rho0,
                             const float* diffcoef, int nsteps, int it0, int with_source, float* pf, float* vxf,
                             float* vzf, float* p1f, float* p2f)
{
  if (!ctx || !p || !vp0 || !rho0 || !diffcoef || !pf || !vxf || !vzf || !p1f || !p2f) {
    rtm::set_error("null argument");
    return -1;
  }
  return ctx->eng->fd_steps(*p, ishot, vp0, rho0, diffcoef, nsteps, it0, with_source, pf, vxf, vzf, p1f, p2f);
}

// ------------------------------------------------------------- drop-in ----
extern "C" void rtm_model(int idevice, int nshot, int startshot, int dshot, int mode, int medium_flag,
                          int flag_layout, int sgyflag, int endianflag, int ismth, int nx, int nz, int nt, float dx,
                          float dz, float dt, int npml, float direct, float direct0, float tsnap, float fdom,
                          float spx0, float spz0, float dspx, float offset, float offsetmax, float amp, float alp,
                          float* vp0, float* rho0, float* diffcoef, float* wavelet, float* image1, float* image2s,
                          float* image3s, float* image4s, float* image5s, float* image2m, float* image3m,
                          float* image4m, float* image5m, float* illum)
{
  (void)wavelet;  // regenerated on the device side as kernel.cu:13-22 does
  rtmh_params P;
  P.nshot = nshot; P.startshot = startshot; P.dshot = dshot; P.mode = mode; P.medium_flag = medium_flag;
  P.flag_layout = flag_layout; P.sgyflag = sgyflag; P.endianflag = endianflag; P.ismth = ismth;
  P.nx = nx; P.nz = nz; P.nt = nt; P.dx = dx; P.dz = dz; P.dt = dt; P.npml = npml;
  P.direct = direct; P.direct0 = direct0; P.tsnap = tsnap; P.fdom = fdom; P.spx0 = spx0; P.spz0 = spz0;
  P.dspx = dspx; P.offset = offset; P.offsetmax = offsetmax; P.amp = amp; P.alp = alp;
  if ((int)(tsnap / dt) < 1) {
    // ntsnap = (int)(tsnap/dt) is a modulus in the reference (rtm_model.cu:21, 271)
    fprintf(stderr, "rtm_model: tsnap must be >= dt\n");
    exit(1);
  }
  rtmh_ctx* ctx = nullptr;
  if (rtmh_create(idevice, &ctx) != 0) {
    fprintf(stderr, "rtm_model: %s\n", rtmh_last_error());
    exit(1);
  }
  rtmh_options o;
  memset(&o, 0, sizeof(o));
  o.batch = 0;
  o.store_fmt = RTMH_STORE_F32;
  o.graph_chunk = 50;
  o.write_files = 1;
  o.verbose = 1;
  // checkpoint / resume for the drop-in, whose signature is the reference's:
  // RTMH_CHECKPOINT names the file (written after every batch, resumed from)
  const char* ck = getenv("RTMH_CHECKPOINT");
  if (ck && *ck) { o.checkpoint = ck; o.resume = 1; }
  o.ic_extra = 1;                                      // every image the reference computes
  const size_t n = (size_t)nx * nz;
  std::vector<float> imagetrue(n, 0.0f);
  rtmh_outputs out;
  memset(&out, 0, sizeof(out));
  out.image1 = image1;
  out.image2s = image2s;
  out.illum = illum;
  out.image3s = image3s; out.image4s = image4s; out.image5s = image5s;
  out.image3m = image3m; out.image4m = image4m; out.image5m = image5m;
  out.imagetrue = imagetrue.data();
  rtmh_stats st;
  memset(&st, 0, sizeof(st));
  const int rc = rtmh_run(ctx, &P, vp0, rho0, diffcoef, &o, &out, &st);
  if (rc != 0) {
    fprintf(stderr, "rtm_model: %s\n", rtmh_last_error());
    rtmh_destroy(ctx);
    exit(1);
  }
  rtmh_destroy(ctx);
  memcpy(image2m, image1, n * sizeof(float));          // the reference's normalisation calls are commented out
  char path[64];
  snprintf(path, sizeof(path), "./output/%dImagetrue.su", idevice);
  rtmio_write_image(imagetrue.data(), nz, nx, dx, path);
}

