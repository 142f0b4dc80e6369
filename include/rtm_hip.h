/*
 * rtm_hip.h -- C ABI of librtm_hip.so, the MI355X (gfx950) implementation of
 * the 2-D acoustic prestack RTM path of gubingluo.
 *
 * Two layers:
 *   1. the drop-in entry points the reference's driver binds (same symbol,
 *      same argument list, same meaning);
 *   2. a native API (prefix rtmh_) with status-code returns, used by the
 *      batched / reduced-precision / served forms, the tests and the bench.
 * Plain C types only: host pointers and sizes; device memory never crosses
 * this boundary.
 */
#ifndef RTM_HIP_H
#define RTM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* 1. drop-in                                                          */
/* ------------------------------------------------------------------ */

/* Replaces rtm_model() of rtm_model.cu:12-17, declared by the caller at
 * RTM_GPU_velstr.cpp:23-28 and called at RTM_GPU_velstr.cpp:229-232.
 * Shots ishot = startshot, startshot+dshot, ... <= nshot (nshot is the last
 * index); vp0/rho0 host [nz*nx] (m/s, kg/m^3); diffcoef host [100];
 * wavelet ignored (the Ricker is regenerated, as the reference does on the
 * GPU).  Outputs are host [nz*nx] arrays, overwritten: image1 (cross-
 * correlation stack), image2s (per-shot illumination-normalised stack),
 * image2m (= image1, the reference's normalisation calls are commented out),
 * illum (illumination stack), image3s/4s/5s and image3m/4m/5m (wavefield
 * decomposition, Poynting and energy-norm conditions, normalised by the
 * illumination or not; kernel.cu:1417-1579).  File side effects as the
 * reference: mode 0 writes
 * ./seisReal/<ishot>seisreal.su, mode 1 reads it; ./output/Image*temp.su
 * every 50 shots; ./output/<idevice>Imagetrue.su.  On failure prints the
 * error and exits with status 1 (the reference exits 0 from check_gpu_error,
 * kernel.cu:1616-1624). */
void rtm_model(int idevice, int nshot, int startshot, int dshot, int mode, int medium_flag, int flag_layout,
               int sgyflag, int endianflag, int ismth, int nx, int nz, int nt, float dx, float dz, float dt,
               int npml, float direct, float direct0, float tsnap, float fdom, float spx0, float spz0, float dspx,
               float offset, float offsetmax, float amp, float alp, float *vp0, float *rho0, float *diffcoef,
               float *wavelet, float *image1, float *image2s, float *image3s, float *image4s, float *image5s,
               float *image2m, float *image3m, float *image4m, float *image5m, float *illum);

/* Replaces rtm_real() of rtm_real.cu:17-22 (field data), with the caller's
 * prototype (RTM_GPU_velstr.cpp:30-35), which differs from the callee's in
 * two positions: nw and tlength arrive as ints, and here they are the
 * direct-wave mute's width and minimum length in samples, the values
 * rtm_real.cu:28-29 derives from direct and direct0.  Per shot ishot it reads
 * shotprofileId[0..3][ishot-1] (sx, traces, first trace, offset; the table of
 * rtmh_scan_gather), the gather from ./seisReal/<filename_shotgather>,
 * resamples it onto the modelling grid (rtmh_interp_seis) and migrates it
 * with the receivers from the offset to the far end on the source's right;
 * outputs and file side effects as rtm_model(). */
void rtm_real(int idevice, int nshot, int startshot, int dshot, int medium_flag, int flag_layout, int sgyflag,
              int endianflag, int ismth, int nx, int nz, int nt, float dx, float dz, float dt, int nx0, int nz0,
              int nt0, float dx0, float dz0, float dt0, int npml, int lx, int lz, float tsnap, float fdom,
              float spz0, float amp, float alp, int nw, int tlength, float *vp, float *rho, float *diffcoef,
              float *wavelet, char filename_shotgather[40], int **shotprofileId, float *image1, float *image2s,
              float *image3s, float *image4s, float *image5s, float *image2m, float *image3m, float *image4m,
              float *image5m, float *illum);

/* Shot-gather scan (scanshotgather, RTM_GPU_velstr.cpp:363-454): a new shot
 * wherever the trace header's sx changes.  prof is [4][max_shots] (rows:
 * sx, traces, first trace 1-based, offset of the first trace), the layout of
 * the reference's shotprofileId; *endianflag = 1 when the headers are
 * big-endian (the reference's probe: ns or dt not positive as read, both
 * positive byte-swapped).  0 on success, -3 when the file holds more than
 * max_shots shots. */
int rtmh_scan_gather(const char *path, int sgyflag, int *endianflag, int *prof, int max_shots, int *nshots);

/* The traces of one shot (rtm_real.cu:379-391): ntrace traces of nt0
 * samples from trace itraceall (1-based) into seis[nt0][ncols]; traces past
 * ncols are dropped, columns past ntrace are zero. */
int rtmh_read_gather(const char *path, int sgyflag, int endianflag, int itraceall, int ntrace, int nt0,
                     float *seis, int ncols);

/* Interpseis2d (Interp_function.cpp:245-264): seisobs[nt0][nx0] sampled at
 * dt0 onto seiscal[nt][nx] at dt: f-k trace interpolation by
 * lx = (nx-1)/(nx0-1) when nx > nx0 (GeneralFKinterp), Kaiser-windowed sinc
 * in time when nt > nt0 (Sincinterp), a copy otherwise (the reference leaves
 * the record unset there).  f-k interpolation of an odd nt0 or onto more than
 * lx*nx0 traces reads past the reference's buffers and is refused (-1). */
int rtmh_interp_seis(float *seiscal, const float *seisobs, int nx, int nx0, int nt, int nt0, float dt, float dt0);

/* ------------------------------------------------------------------ */
/* 2. native API                                                       */
/* ------------------------------------------------------------------ */

/* The scalar arguments of rtm_model(), same order and meaning. */
typedef struct rtmh_params {
  int nshot, startshot, dshot, mode, medium_flag, flag_layout, sgyflag, endianflag, ismth;
  int nx, nz, nt;
  float dx, dz, dt;
  int npml;
  float direct, direct0, tsnap, fdom, spx0, spz0, dspx, offset, offsetmax, amp, alp;
} rtmh_params;

enum { RTMH_STORE_F32 = 0, RTMH_STORE_F16 = 1, RTMH_STORE_BF16 = 2 };

typedef struct rtmh_options {
  int batch;          /* shots per launch; 0 = as many as fit the memory budget  */
  int store_fmt;      /* boundary store: RTMH_STORE_F32 / _F16 / _BF16           */
  int graph_chunk;    /* time steps per captured hipGraph; 0 = plain launches    */
  int tile_z;         /* tile height 16/32/64; 0 = autotuned                     */
  int write_files;    /* reproduce the reference's file side effects             */
  int verbose;        /* 0 silent, 1 per-shot lines as the reference prints       */
  double mem_budget;  /* bytes of HBM for shot slots; 0 = 80% of free memory      */
  int profile;        /* 1: plain launches timed one by one with HIP events       */
  const float *shot_offsetmax; /* per shot (job order) aperture overriding offsetmax
                                  (ragged windows); NULL: offsetmax for every shot */
  /* Checkpoint / resume (no counterpart in the reference, whose temp images
   * every 50 shots are fp32 snapshots it never reads back): after every batch
   * the FP64 image stacks and the number of shots done are written to
   * `checkpoint` (atomically: a temporary file renamed over it), with a
   * fingerprint of the job (parameters except nshot, model, coefficients,
   * boundary precision, per-shot apertures).  With resume = 1 a job whose
   * fingerprint matches continues after the shots done, and its images equal
   * those of an uninterrupted run bit for bit; a checkpoint of another job is
   * an error, a missing file starts from the first shot. */
  const char *checkpoint;      /* NULL: no checkpoint */
  int resume;
  /* 1: also the decomposition, Poynting and energy-norm imaging conditions
   * (image3s/4s/5s, image3m/4m/5m, imagetrue; kernel.cu:1417-1555), one more
   * launch per reverse step.  The drop-in rtm_model sets it. */
  int ic_extra;
  /* per shot (job order) [sx, offset, offsetmax] of a recorded gather
   * (rtm_real.cu:121-136): window from sx to sx + offsetmax, receivers from
   * offset on the source's right, direct-wave mute of cuda_mute2; NULL: the
   * geometry of flag_layout.  The drop-in rtm_real sets it. */
  const float *shot_field;
  /* mode 1: called with the shot's job index before its input record
   * (rtmh_outputs.shot_seis_in) is read, which happens after the forward
   * pass of its batch, so that a caller can still be preparing the records
   * meanwhile; non-zero return aborts the job.  NULL: records are ready. */
  int (*seis_wait)(void *user, int shot);
  void *seis_user;
  /* Decimated snapshots of the source wavefield (no counterpart in the
   * reference, which reconstructs the source wavefield from the boundary
   * strips alone): with snap_every = k > 0 the forward pass saves the
   * interior of p, vx and vz at every k-th time step past the wavelet
   * (t >= 4/fdom) in snap_fmt (RTMH_STORE_*; 16-bit with one max-abs scale
   * per field and snapshot), and the reverse pass restores them into the
   * reconstructed field at those steps, as checkpoints that bound how far
   * the reconstruction can drift from the forward field.  Memory per shot:
   * 3 nz nxl values per snapshot.  0: off. */
  int snap_every;
  int snap_fmt;
} rtmh_options;

/* Optional outputs; any pointer may be NULL. Images are [nz*nx]. */
typedef struct rtmh_outputs {
  float *image1, *image2s, *illum;
  float *image_lap;         /* laplacefd1d(image1), image_function.cpp:44-79      */
  float *image_norm;        /* image1/(illum+1e-6), kernel.cu:1580-1590           */
  float **shot_seis;        /* per shot (job order): [nt*nrec] unmuted record     */
  float **shot_g1;          /* per shot: [nz*nxl] cross-correlation, untapered    */
  float **shot_illum;       /* per shot: [nz*nxl] illumination, untapered         */
  const float **shot_seis_in; /* per shot: [nt*nrec] data to migrate (mode 1)     */
  /* with ic_extra: decomposition (3), Poynting (4) and energy-norm (5)
   * images normalised by the illumination (s) or not (m), and imagetrue */
  float *image3s, *image4s, *image5s, *image3m, *image4m, *image5m, *imagetrue;
} rtmh_outputs;

typedef struct rtmh_stats {
  double total_ms;          /* whole job incl. setup and copies                   */
  double gpu_ms;            /* propagation passes (forward + reverse), HIP events */
  double fwd_ms, rev_ms;
  double cell_updates;      /* 2*nt*nxp*nzp + (nt-1)*nz*nxl summed over shots     */
  int nshots, batches, tile_z, batch;
  double hbm_bytes_alg;     /* algorithmic bytes of the propagation passes        */
  /* profile=1 only: per kernel (0 forward step, 1 strip save after it,
   * 2 reverse step, 3 unused) summed launch time and launch count */
  double kern_ms[4];
  int kern_calls[4];
  double step_bytes_fwd;    /* algorithmic bytes of one forward step of the batch */
  double step_bytes_rev;    /* ... of one reverse step                            */
  int tile_z_rev;           /* tile height of the reverse pass (tile_z: forward)  */
  int _pad0;
  double fwd_bytes_alg;     /* algorithmic bytes of all forward steps (all batches) */
  double rev_bytes_alg;     /* ... of all reverse steps                           */
} rtmh_stats;

/* last error message of this thread ("" if none) */
const char *rtmh_last_error(void);
/* library version string */
const char *rtmh_version(void);
/* number of HIP devices (0 without a GPU); does not initialise a context */
int rtmh_device_count(void);

typedef struct rtmh_ctx rtmh_ctx;
int rtmh_create(int device, rtmh_ctx **out);
int rtmh_destroy(rtmh_ctx *ctx);

/* per-shot geometry of a job (window, padded dims, receivers); -1 if the
 * shot lies outside the model */
typedef struct rtmh_geom {
  int ishot, nx1, nx2, nxl, nxp, nzp, nsx, nsz, nrec, rcol0, noffset;
} rtmh_geom;
int rtmh_shot_geometry(const rtmh_params *p, int ishot, rtmh_geom *g);
/* Window of a recorded gather, field = [sx, offset, offsetmax]
 * (rtm_real.cu:121-136); -1 when it leaves the model. */
int rtmh_gather_geometry(const rtmh_params *p, int ishot, const float *field, rtmh_geom *g);

/* Whole job as rtm_model() does it, with options; returns 0 or a negative
 * status (message in rtmh_last_error()): -1 bad parameters, including a dt
 * above the stability bound min(dx,dz)/(sqrt(2) sum|c| vmax) of the model
 * (get_constant, RTM_GPU_velstr.cpp:355-361, where the driver refines dt
 * instead); -2 a shot window outside the model; -3 out of device memory;
 * -4 an input record missing; -5 a checkpoint of another job, or one with
 * more shots done than the job has; -6 a checkpoint not written; -7 a
 * wavefield that is not finite (checked on the device every
 * ntsnap = tsnap/dt steps, or after every replayed graph chunk, the
 * reference's max|p| print of rtm_model.cu:271-273); -10/-11 HIP errors. */
int rtmh_run(rtmh_ctx *ctx, const rtmh_params *p, const float *vp0, const float *rho0, const float *diffcoef,
             const rtmh_options *opt, rtmh_outputs *out, rtmh_stats *stats);

/* Device-resident velocity model (density and v^2 kept in HBM across jobs;
 * the per-shot windows of extractvel1/extendvel1 are read from it by
 * clamped indexing instead of per-shot copies). */
int rtmh_model_create(rtmh_ctx *ctx, int nx, int nz, const float *vp0, const float *rho0, int *model_id);
int rtmh_model_destroy(rtmh_ctx *ctx, int model_id);
/* rtmh_run on a resident model (p->nx, p->nz must match it). */
int rtmh_run_model(rtmh_ctx *ctx, int model_id, const rtmh_params *p, const float *diffcoef,
                   const rtmh_options *opt, rtmh_outputs *out, rtmh_stats *stats);

/* nsteps forward FD steps of one shot's source wavefield from given fields
 * (unit-level parity of the velocity-stress + PML update).  All field arrays
 * are host [nzp*nxp] (vx uses columns 0..nxp-2, vz rows 0..nzp-2; p1/p2 hold
 * the split pressure pair at PML cells).  it0: wavelet index of the first
 * step; with_source adds the source term as the reference does. */
int rtmh_fd_steps(rtmh_ctx *ctx, const rtmh_params *p, int ishot, const float *vp0, const float *rho0,
                  const float *diffcoef, int nsteps, int it0, int with_source, float *pf, float *vxf, float *vzf,
                  float *p1f, float *p2f);

/* Achievable-bandwidth probe (no counterpart in the reference; measurement
 * only): `reps` device-to-device copies of `bytes` on the context's stream,
 * timed with HIP events; *gbps = (read + written bytes) / average time. */
int rtmh_copy_probe(rtmh_ctx *ctx, double bytes, int reps, double *gbps);


/* ------------------------------------------------------------------ */
/* 3. drop-in driver helpers (the host side of RTM_GPU_velstr.cpp)     */
/* ------------------------------------------------------------------ */

/* Liu-Yang staggered coefficient table, out[100], row r-1 = order 2r
 * (FD_function.cpp:73-165 assembled as RTM_GPU_velstr.cpp:184-196). */
void rtmh_coef_table(float *out);

/* Parameter file of the driver (RTM_GPU_velstr.cpp:60-120). */
typedef struct rtmh_parfile {
  float dt0; int nt0; float tsnap;
  float xmax, zmax, dx0, dz0;
  float fdom, amp, alp;
  float spx0, spz0, dspx;
  float offset, offsetmax;
  int startshot, nshot, dshot, nshotall;
  int npml, medium_flag, ismth;
  float direct, direct0;
  int idevice, flag_layout, mode, dataflag, sgyflag, endianflag, rhoflag;
  char vp_file[256], shotgather[256], rho_file[256];
} rtmh_parfile;
/* 0 ok, -1 missing file, -2 malformed */
int rtmh_read_parfile(const char *path, rtmh_parfile *pf);

/* Grid after the dispersion / stability guard get_constant()
 * (RTM_GPU_velstr.cpp:319-362). */
typedef struct rtmh_grid {
  float dx, dz, dt; int nx, nz, nt, lx, lz; float vpmin, vpmax;
} rtmh_grid;
void rtmh_get_constant(const float *vp, float fdom, float dx0, float dz0, int nx0, int nz0, int nt0, float dt0,
                       rtmh_grid *g);

/* interpvel1 (Velocity_function.cpp:1946-1977). */
void rtmh_interp_model(const float *vp0, float *vp, int nx0, int nz0, int nx, int nz, int lx, int lz);

/* Model smoothing of a shot's padded window (v^2 or density; nz rows of nx,
 * x fastest), in place.  rtmh_pml_smooth replaces pmlvelsmooth1d
 * (Velocity_function.cpp:9-103; applied when medium_flag != 0, rtm_model.cu:
 * 204-206): the absorbing layer mirrored from its neighbour cells, then
 * relaxed with 4-point averages until a sweep changes no cell by more than
 * 5e-4 (relative) or 1001 sweeps ran.  rtmh_gauss_smooth replaces
 * velsmooth1d (Velocity_function.cpp:199-313; applied before the reverse
 * pass when ismth > 0, rtm_model.cu:402-405): a (2 nsp + 1)^2 Gaussian
 * window (sigma nsp/3) over the edge-extended window, normalised.  Both give
 * the reference's values bit for bit. */
void rtmh_pml_smooth(float *w, int nx, int nz, int npml);
void rtmh_gauss_smooth(float *w, int nz, int nx, int nsp);

/* The same two smoothers on the GPU of a context, in place on a host window
 * (x fastest, nz rows of nx): pmlvelsmooth1d when npml > 0 (the sweeps in the
 * reference's in-place order, one workgroup per window; windows with
 * npml <= 65 and nx, nz in [3 npml, 4096]), then velsmooth1d when nsp > 0.
 * Bit-identical to rtmh_pml_smooth / rtmh_gauss_smooth.  The engine smooths
 * each shot's windows this way (medium_flag before the forward pass, ismth
 * before the reverse pass, on the slot's device windows); this entry point
 * exposes it for checking and for callers preparing models themselves. */
int rtmh_smooth_gpu(rtmh_ctx *ctx, float *w, int nx, int nz, int npml, int nsp);

/* selfdoc banner (RTM_GPU_velstr.cpp:155-182) on stdout. */
void rtmh_selfdoc(const rtmh_parfile *pf, int nx0, int nz0, const rtmh_grid *g);

/* Laplacian-filtered image (laplacefd1d, image_function.cpp:44-79, integer
 * spacings as its int arguments) and multi-shot normalised image
 * (cuda_applyimagingconditionmulti, kernel.cu:1580-1590) on the GPU. */
int rtmh_postprocess(int device, const float *image1, const float *illum, int nx, int nz, float dx, float dz,
                     float *image_lap, float *image_norm);

/* rtmh_postprocess without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments and results as rtmh_postprocess.  Its staging (image1,
 * illum and one result, nx*nz floats each) lives in a file-scope device
 * array of 3 * RTMH_NOALLOC_MAX_CELLS floats (24 MiB) that every GPU the
 * library runs on holds until the process ends; images of more than
 * RTMH_NOALLOC_MAX_CELLS cells are refused (-3).  Calls on one GPU share
 * that array: calls that overlap on the same GPU race and must be
 * serialised by the caller. */
#define RTMH_NOALLOC_MAX_CELLS (2048 * 1024)
int rtmh_postprocess_noalloc(int device, const float *image1, const float *illum, int nx, int nz, float dx, float dz,
                             float *image_lap, float *image_norm);

/* main() of RTM_GPU_velstr.cpp:37-317: reads ./input/cmd.in, runs, writes
 * ./output/<dev>*.su (plus <dev>Image1lap.su and <dev>Imagenorm.su). */
int rtmh_driver_main(int argc, char **argv);

/* ------------------------------------------------------------------ */
/* 4. served form: continuous batching of RTM requests on one GPU      */
/* ------------------------------------------------------------------ */
/* A worker thread advances every resident shot by one time step per tick:
 * shots of different requests, grid sizes, time steps and phases (forward /
 * reverse) share the same launches.  Queued shots are admitted first come
 * first served as slots and HBM budget free up; each request's shots are
 * stacked in job order, so a request's images do not depend on what else
 * runs beside it.  Requests are mode-0 jobs (data modelled, then migrated)
 * on models created with rtmh_model_create on the same context, which must
 * outlive them.  A context serves either rtmh_run* calls or one server at a
 * time. */
typedef struct rtmh_server rtmh_server;
typedef struct rtmh_server_options {
  double mem_budget;  /* bytes of HBM for slots and request stacks; 0 = 80% of the
                         memory free once the context's cached slot memory is
                         released.  rtmh_server_create takes the whole budget
                         in one device allocation (up to ~21 ms per GB) and
                         carves slots from it; a budget the device cannot
                         give in one piece is allocated as shots need it */
  int max_slots;      /* shots resident at once; 0 = 256                          */
  int tile_z;         /* 16 / 32 / 64; 0 = 32                                     */
  int store_fmt;      /* boundary store precision of every request                */
} rtmh_server_options;
typedef struct rtmh_request_stats {
  double submit_s, start_s, done_s;   /* seconds since the server started          */
  double cell_updates;
  int nshots, status;
} rtmh_request_stats;
int rtmh_server_create(rtmh_ctx *ctx, const rtmh_server_options *opt, rtmh_server **out);
/* queue one job; returns its request id.  Refused (-1) like rtmh_run: bad
 * parameters, dt above the stability bound of the model (get_constant,
 * RTM_GPU_velstr.cpp:355-361); refused (-3) when one shot cannot fit the
 * budget. */
int rtmh_server_submit(rtmh_server *srv, int model_id, const rtmh_params *p, const float *diffcoef,
                       int *request_id);
/* HBM the server holds for a request of this job (no counterpart in the
 * reference; for admission planning): *shot_max = bytes of its largest shot
 * while resident (slot, plus its images until they are stacked), *request =
 * the request's own block (coefficients, wavelet, tapers, FP64 stacks,
 * outputs).  rtmh_server_submit refuses (-3) a job with shot_max + request
 * above the server's budget; admission waits while the budget is held by
 * other shots and fails the request (-3 from rtmh_server_wait) when the
 * device refuses memory that nothing resident will free. */
int rtmh_server_request_bytes(rtmh_server *srv, int model_id, const rtmh_params *p, const float *diffcoef,
                              double *shot_max, double *request);
/* 0 when done (outputs copied, stats filled), 1 on timeout, < 0 on error:
 * the request failed (-3 memory, -7 its wavefield blew up), or the server did */
int rtmh_server_wait(rtmh_server *srv, int request_id, double timeout_s, rtmh_outputs *out,
                     rtmh_request_stats *stats);
/* server-wide counters: ticks, launches, peak resident shots, busy seconds,
 * the HBM budget the server admits against (bytes), and the device
 * allocations it made (count, host seconds in hipMalloc/hipFree; 1 and the
 * time of the reservation when the budget was taken at start) */
typedef struct rtmh_server_stats {
  long long ticks, launches;
  int peak_slots, active_slots, queued_requests;
  double busy_s, cell_updates;
  double budget_bytes;
  long long allocations;
  double alloc_s;
} rtmh_server_stats;
int rtmh_server_query(rtmh_server *srv, rtmh_server_stats *st);
int rtmh_server_destroy(rtmh_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* RTM_HIP_H */
