/*
 * rtm_oracle.c -- CPU oracle for the 2-D acoustic prestack RTM path of the
 * reference (gubingluo: RTM_GPU_velstr.cpp -> rtm_model.cu -> kernel.cu).
 *
 * *** TEST INFRASTRUCTURE ONLY ***
 * This library is the checker.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load or call it, and only to compute the
 * expected answer or the CPU baseline time.  The product library
 * (gubingluo_amd) never links, loads or falls back to it.
 *
 * What is restated (file:line of the reference):
 *   shot window + padded source position .... rtm_model.cu:122-153 (layout 1),
 *                                             :626-652 (2), :1125-1155 (3), :1632-1660 (4)
 *   window extraction, v^2, edge padding .... Velocity_function.cpp:2011-2073
 *   per-shot vmax, PML profiles ............. rtm_model.cu:207, kernel.cu:174-269
 *   Ricker wavelet, Gaussian source ......... kernel.cu:13-35
 *   FD order maps ........................... kernel.cu:271-317
 *   velocity / pressure updates + PML ....... kernel.cu:430-1185
 *   record / insert / mute .................. kernel.cu:48-173
 *   boundary strips save / restore .......... kernel.cu:1186-1400
 *   reverse reconstruction .................. kernel.cu:318-429, rtm_model.cu:413-437
 *   cross-correlation + illumination ........ kernel.cu:1401-1416
 *   taper, stacking ......................... kernel.cu:1556-1615, rtm_model.cu:462-488
 *   Laplacian filter ........................ image_function.cpp:44-79
 *   multi-shot normalisation ................ kernel.cu:1580-1590, image_function.cpp:11-20
 *
 * Three arithmetic forms of the same discrete algorithm:
 *   _f32  reference order: every FP32 operation of the reference's kernels in
 *         source order (the parity yardstick for the seismogram and images);
 *   _k32  kernel order: FP32 with the operation order of the HIP kernels --
 *         velocity u = (s*dt/h) * (1/rho), pressure p -= (rho*v^2)*(sx*dt/dx +
 *         sz*dt/dz) as fma, damping fma(c1/c2, q, -(d * (1/c2))); per-cell 1/rho
 *         and rho*v^2 and per-node c1/c2, 1/c2 are rounded once.  A checking
 *         aid: the GPU is compared with it bit for bit, and it is compared with
 *         _f32 within the stated tolerances;
 *   _f64  the reference order in FP64 (error yardstick of reduced precision).
 *
 * Two pressure/PML forms:
 *   literal      -- the kernels' own thread ranges (valid only for npml == 32,
 *                   where the 32-wide PML blocks tile the absorbing layer);
 *   generalised  -- every cell of the padded grid updated by the same
 *                   per-cell formulas, for any npml >= 10.  Identical to the
 *                   literal form at npml == 32 (checked by the tests).
 *
 * Also restated, not reference semantics: the library's 16-bit boundary
 * store (orc_store16_block: per-step block scales, binary16 / bfloat16 codes
 * rounded to nearest even) and its decimated snapshots (exact or 16-bit)
 * (orc_shot_opts.snap_every), restored into the reconstruction.
 *
 * Checks: tests/test_oracle_*.py and tests/test_golden.py pin each function
 * by golden vectors of the reference's own host code, closed forms, physical
 * properties (Green's function, mirror symmetry, plane-wave attenuation in
 * the absorbing layer, image and traveltime locations) or brute force.
 * Parity unpinned: the index quirks of the other imaging conditions
 * (kernel.cu:1431,1468-1469: vx read at the p index, the source Poynting
 * vector with the receiver's vz) rest on reading the source (position-coded
 * fields show that the oracle reads what that source names), as do the last
 * bits of CUDA's acosf/cosf/sinf there and of its expf/logf in the set-up
 * values (wavelet, source, PML d0); the CUDA path cannot be built here.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))
#define ORC_PI 3.1415926535897932   /* common.h:11 */
#define R10 10                      /* N/2, common.h:57 */

static inline int orc_imin(int a, int b) { return a < b ? a : b; }
static inline int orc_imax(int a, int b) { return a > b ? a : b; }

/* the scalar arguments of rtm_model() (rtm_model.cu:12-17) */
typedef struct {
    int nshot, startshot, dshot, mode, medium_flag, flag_layout, sgyflag, endianflag, ismth;
    int nx, nz, nt;
    float dx, dz, dt;
    int npml;
    float direct, direct0, tsnap, fdom, spx0, spz0, dspx, offset, offsetmax, amp, alp;
    /* field != 0: the shot is a recorded gather (rtm_real.cu:121-136) with
       source at fsx and receivers from foffset to foffsetmax on its right */
    int field;
    float fsx, foffset, foffsetmax;
} orc_params;

/* per-shot geometry */
typedef struct {
    int ishot, nx1, nx2, nxl, nxp, nzp, nsx, nsz;
    int nrec;      /* receivers recorded / inserted                     */
    int rcol0;     /* padded column of receiver 0                       */
    int noffset;   /* layouts 3/4                                       */
    /* SU key header of the shot (rtm_model.cu:137-145) */
    int hk_itraceall, hk_offset, hk_sx, hk_gx, hk_ntrace, hk_nt, hk_dt;
    float hk_dx;
} orc_geom;

typedef struct {
    int literal;          /* 1: literal kernel ranges (npml must be 32)        */
    int forward_only;     /* stop after the forward propagation                */
    int nt_limit;         /* >0: only this many forward steps (timing samples) */
    int snap_it;          /* step whose wavefields are returned (-1: none)     */
    const float *vvpad;   /* optional padded v^2 window (nzp*nxp)              */
    const float *rhopad;  /* optional padded density window                    */
    int exact_reverse;    /* checking aid, not reference semantics: remove the
                             source at it+1 from interior cells only, so the
                             reverse pass inverts the forward one */
    float *ic_out;        /* optional: the other imaging conditions of the shot,
                             4 x nz x nxl (g2 decomposition sum, g3 Poynting
                             theta sum, g4 energy norm, g3_true), untapered */
    int snap_every;       /* >0: decimated snapshots of the library's option
                             snap_every (not reference semantics): the forward
                             pass keeps the interior of p of step k and of the
                             velocities of step k+1 for k = k0*every, ...,
                             k <= nt-2, k0 = max(1, ceil(t0/every)),
                             t0 = ceil(4/(fdom*dt)); after reverse step k
                             (imaging included) they replace the reconstructed
                             source field's interior cells, exactly (no
                             quantisation) */
    int store_fmt;        /* boundary store of the library's option store_fmt
                             (not reference semantics; FP32 forms only):
                             0 exact, 1 binary16, 2 bfloat16 with one scale
                             per step and strip block (orc_store16_block) */
    int snap_fmt;         /* snapshots of snap_every stored the same way
                             (FP32 forms): 0 exact, 1 binary16, 2 bfloat16,
                             one scale per snapshot and field over its
                             nz x nxl interior */
} orc_shot_opts;

/* ------------------------------------------------------------------ */
ORC_EXPORT int orc_shot_geometry(const orc_params *P, int ishot, orc_geom *g)
{
    memset(g, 0, sizeof(*g));
    const float dx = P->dx, dz = P->dz;
    const int nx = P->nx, npml = P->npml;
    int nx1, nx2, is = 0, noffset = 0;
    float spx, spz;
    g->ishot = ishot;
    int layout = P->flag_layout;
    if (P->field) {                                              /* rtm_real.cu:121-136 */
        /* noffset, nx1, nx2 from the gather's sx, offset and offsetmax on
           the modelling grid; the window is not clamped to the model there,
           so a gather that leaves it has no defined result */
        noffset = (int)(P->foffset / dx);
        nx1 = (int)(P->fsx / dx);
        nx2 = (int)((P->fsx + P->foffsetmax) / dx);
        if (P->fsx < 0 || nx2 > nx - 1) return -1;
        spx = P->fsx - nx1 * dx;
        layout = 3;                                              /* record2 / insert_record2 / mute2 */
    } else if (P->flag_layout == 1) {                            /* rtm_model.cu:124-134 */
        int NX = (int)(P->offsetmax / dx + 1);
        int sp1 = (int)(P->spx0 / dx);
        int np = (int)(P->dspx / dx);
        int nr = (NX - 1) / 2;
        is = sp1 + (ishot - 1) * np;
        nx1 = orc_imax(0, is - nr);
        nx2 = orc_imin(nx - 1, is + nr);
        spx = (is - nx1) * dx;
    } else if (P->flag_layout == 2) {                            /* rtm_model.cu:627-630 */
        nx1 = 0; nx2 = nx - 1;
        spx = P->spx0 + (ishot - 1) * P->dspx;
    } else if (P->flag_layout == 3 || P->flag_layout == 4) {     /* :1126-1138, :1633-1645 */
        int NX = (int)(P->offsetmax / dx + 1);
        int sp1 = (int)(P->spx0 / dx);
        int np = (int)(P->dspx / dx);
        is = sp1 + (ishot - 1) * np;
        noffset = (int)(P->offset / dx);
        int nr = NX - 1;
        if (P->flag_layout == 3) { nx1 = orc_imax(0, is); nx2 = orc_imin(nx - 1, is + nr); }
        else { nx1 = orc_imax(0, is - nr); nx2 = orc_imin(nx - 1, is); }
        spx = (is - nx1) * dx;
    } else {
        return -1;
    }
    int nxl = nx2 - nx1 + 1;
    if (nxl <= 0) return -1;
    spz = P->spz0;
    g->nx1 = nx1; g->nx2 = nx2; g->nxl = nxl; g->noffset = noffset;
    /* key header */
    g->hk_nt = (int)(short)(unsigned short)(P->nt);
    g->hk_dt = (int)(short)(unsigned short)(P->dt * 1000000);
    g->hk_dx = dx;
    if (layout <= 2) {
        g->hk_itraceall = (ishot - 1) * nxl + 1;
        g->hk_offset = (int)(nx1 * dx - spx);
        g->hk_sx = (int)(spx);
        g->hk_gx = (int)(nx1 * dx);
        g->hk_ntrace = nxl;
    } else {
        g->hk_itraceall = (ishot - 1) * (nxl - noffset) + 1;
        g->hk_offset = (int)(P->offset);
        g->hk_sx = (int)(spx);
        g->hk_gx = (int)(P->offset + nx1 * dx);
        g->hk_ntrace = nxl - noffset;
    }
    spx = spx + npml * dx;
    spz = spz + npml * dz;
    g->nxp = nxl + 2 * npml;
    g->nzp = P->nz + 2 * npml;
    g->nsx = (int)(spx / dx);
    g->nsz = (int)(spz / dz);
    /* receivers: record1/insert1 (layouts 1,2), record2 (3), record3 (4) */
    if (layout <= 2) { g->nrec = nxl; g->rcol0 = npml; }
    else if (layout == 3) { g->nrec = nxl - noffset; g->rcol0 = npml + noffset; }
    else { g->nrec = nxl - noffset; g->rcol0 = npml; }
    if (g->nrec < 0) g->nrec = 0;
    return 0;
}

/* extractvel1/extractrho1 + extendvel1 (Velocity_function.cpp:2011-2073) */
ORC_EXPORT void orc_window_pad(float *out, const float *in, int nx, int nz, int nx1, int nx2,
                               int npml, int square)
{
    int nxl = nx2 - nx1 + 1, nxp = nxl + 2 * npml, nzp = nz + 2 * npml;
    for (int iz = 0; iz < nz; iz++)
        for (int ix = nx1; ix <= nx2; ix++) {
            float v = in[(size_t)iz * nx + ix];
            out[(size_t)(iz + npml) * nxp + ix - nx1 + npml] = square ? v * v : v;
        }
    for (int iz = npml; iz < nzp - npml; iz++)
        for (int ix = 0; ix < npml; ix++) {
            out[(size_t)iz * nxp + ix] = out[(size_t)iz * nxp + npml];
            out[(size_t)iz * nxp + nxp - 1 - ix] = out[(size_t)iz * nxp + nxp - npml - 1];
        }
    for (int ix = 0; ix < nxp; ix++)
        for (int iz = 0; iz < npml; iz++) {
            out[(size_t)iz * nxp + ix] = out[(size_t)npml * nxp + ix];
            out[(size_t)(nzp - 1 - iz) * nxp + ix] = out[(size_t)(nzp - 1 - npml) * nxp + ix];
        }
}

/* pmlvelsmooth1d (Velocity_function.cpp:9-103), applied to the padded v^2 and
 * density windows when medium_flag is set (rtm_model.cu:204-206): the
 * absorbing layer is mirrored from the cells next to it, then relaxed in
 * place with 4-point averages in the reference's sweep order until the
 * largest relative change of a sweep is <= 5e-4 or 1001 sweeps have run.
 * w is nz x nx, x fastest.  The reference is C++: fabs of a float is the
 * float overload, so the change ratio is a float division. */
ORC_EXPORT void orc_pml_smooth(float *w, int nx, int nz, int npml)
{
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
    float *v = (float *)malloc(n * sizeof(float));
    memcpy(v, w, n * sizeof(float));
#define SV(iz, ix) v[(size_t)(iz) * nx + (ix)]
    float err = 1.0f;
    int num = 0;
    const float maxit = 1000;
    while (num <= maxit && err > 0.0005) {
        for (int ix = 1; ix < npml; ix++)
            for (int iz = 1; iz < nz - 1; iz++) {
                SV(iz, ix) = 0.25 * (SV(iz, ix + 1) + SV(iz, ix - 1) + SV(iz + 1, ix) + SV(iz - 1, ix));
                SV(iz, ix + nx - npml - 1) = 0.25 * (SV(iz, ix + nx - npml) + SV(iz, ix + nx - npml - 2) +
                                                     SV(iz + 1, ix + nx - npml - 1) + SV(iz - 1, ix + nx - npml - 1));
            }
        for (int ix = 1; ix < nx - 1; ix++)
            for (int iz = 1; iz < npml; iz++) {
                SV(iz, ix) = 0.25 * (SV(iz, ix + 1) + SV(iz, ix - 1) + SV(iz + 1, ix) + SV(iz - 1, ix));
                SV(iz + nz - npml - 1, ix) = 0.25 * (SV(iz + nz - npml - 1, ix + 1) + SV(iz + nz - npml - 1, ix - 1) +
                                                     SV(iz + nz - npml, ix) + SV(iz + nz - npml - 2, ix));
            }
        for (int iz = 1; iz < nz - 1; iz++) {
            SV(iz, 0) = 0.25 * (SV(iz, 0) + SV(iz, 1) + SV(iz + 1, 0) + SV(iz - 1, 0));
            SV(iz, nx - 1) = 0.25 * (SV(iz, nx - 1) + SV(iz, nx - 2) + SV(iz + 1, nx - 1) + SV(iz - 1, nx - 1));
        }
        for (int ix = 1; ix < nx - 1; ix++) {
            SV(0, ix) = 0.25 * (SV(0, ix) + SV(1, ix) + SV(0, ix - 1) + SV(0, ix + 1));
            SV(nz - 1, ix) = 0.25 * (SV(nz - 1, ix) + SV(nz - 2, ix) + SV(nz - 1, ix - 1) + SV(nz - 1, ix + 1));
        }
        SV(0, 0) = 0.25 * (SV(0, 0) + SV(1, 0) + SV(0, 1) + SV(1, 1));
        SV(nz - 1, 0) = 0.25 * (SV(nz - 1, 0) + SV(nz - 2, 0) + SV(nz - 1, 1) + SV(nz - 2, 1));
        SV(0, nx - 1) = 0.25 * (SV(0, nx - 1) + SV(0, nx - 2) + SV(1, nx - 1) + SV(1, nx - 2));
        SV(nz - 1, nx - 1) = 0.25 * (SV(nz - 1, nx - 1) + SV(nz - 2, nx - 1) + SV(nz - 1, nx - 2) + SV(nz - 2, nx - 2));
        float m = -1.0e+5f;
        for (size_t i = 0; i < n; i++) {
            const float r = fabsf(w[i] - v[i]) / fabsf(w[i]);
            m = r > m ? r : m;
        }
        err = m;
        memcpy(w, v, n * sizeof(float));
        num++;
    }
#undef SV
    free(v);
}

/* velsmooth1d (Velocity_function.cpp:199-313), applied to the padded v^2 and
 * density windows before the reverse pass when ismth > 0 (rtm_model.cu:
 * 402-405): the window extended by nsp cells (edge values), a (2 nsp + 1)^2
 * Gaussian window with sigma = nsp/3 summed in double in row order, then
 * divided by the sum of the weights.  w is n1 (z) x n2 (x), x fastest. */
ORC_EXPORT void orc_gauss_smooth(float *w, int n1, int n2, int nsp)
{
    const int n1e = n1 + 2 * nsp, n2e = n2 + 2 * nsp;
    float *b = (float *)calloc((size_t)n1e * n2e, sizeof(float));
    float *a = (float *)calloc((size_t)n1 * n2, sizeof(float));
#define SB(i, j) b[(size_t)(i) * n2e + (j)]
#define SA(i, j) a[(size_t)(i) * n2 + (j)]
    for (int i1 = 0; i1 < n1; i1++)
        for (int i2 = 0; i2 < n2; i2++) SB(i1 + nsp, i2 + nsp) = w[(size_t)i1 * n2 + i2];
    for (int i1 = 0; i1 < n1; i1++)
        for (int i2 = 0; i2 < nsp; i2++) {
            SB(i1 + nsp, i2) = w[(size_t)i1 * n2];
            SB(i1 + nsp, i2 + n2 + nsp) = w[(size_t)i1 * n2 + n2 - 1];
        }
    for (int i1 = 0; i1 < nsp; i1++)
        for (int i2 = 0; i2 < n2; i2++) {
            SB(i1, i2 + nsp) = w[i2];
            SB(nsp + n1 + i1, i2 + nsp) = w[(size_t)(n1 - 1) * n2 + i2];
        }
    for (int i1 = 0; i1 < nsp; i1++)
        for (int i2 = 0; i2 < nsp; i2++) {
            SB(i1, i2) = w[0];
            SB(i1, nsp + n2 + i2) = w[n2 - 1];
            SB(i1 + nsp + n1, i2) = w[(size_t)(n1 - 1) * n2];
            SB(i1 + nsp + n1, i2 + nsp + n2) = w[(size_t)(n1 - 1) * n2 + n2 - 1];
        }
    for (int i1 = nsp; i1 < n1 + nsp; i1++)
        for (int i2 = nsp; i2 < n2 + nsp; i2++) {
            double s = 0;
            for (int i11 = i1 - nsp; i11 <= i1 + nsp; i11++)
                for (int i22 = i2 - nsp; i22 <= i2 + nsp; i22++) {
                    const double d1 = i11 - i1, d2 = i22 - i2;
                    const double wt = exp(-(d1 * d1 + d2 * d2) / (2.0 * nsp / 3.0 * nsp / 3.0));
                    s += wt * SB(i11, i22);
                }
            SA(i1 - nsp, i2 - nsp) = (float)s;
        }
    double norm = 0;
    for (int i11 = 0; i11 <= 2.0 * nsp; i11++)
        for (int i22 = 0; i22 <= 2.0 * nsp; i22++) {
            const double d1 = i11 - nsp, d2 = i22 - nsp;
            norm += exp(-(d1 * d1 + d2 * d2) / (2.0 * nsp / 3.0 * nsp / 3.0));
        }
    for (size_t i = 0; i < (size_t)n1 * n2; i++) w[i] = (float)(a[i] / norm);
#undef SA
#undef SB
    free(a); free(b);
}

/* cuda_mute1/2/3 (kernel.cu:96-173): last muted sample per receiver.
 * vvpad: padded v^2 window of the shot (NULL: built from vp0). */
ORC_EXPORT void orc_mute_cuts(const orc_params *P, const orc_geom *g, const float *unused,
                              const float *vp0, const float *vvpad, int *cut)
{
    (void)unused;
    const int nxp = g->nxp, npml = P->npml, nt = P->nt;
    float *vv = NULL;
    if (!vvpad) {
        vv = (float *)malloc((size_t)nxp * g->nzp * sizeof(float));
        orc_window_pad(vv, vp0, P->nx, P->nz, g->nx1, g->nx2, npml, 1);
        vvpad = vv;
    }
    float dt = P->dt;
    float _dt = (float)(1.0 / (double)dt);                      /* rtm_model.cu:75 */
    float dx2 = P->dx * P->dx, dz2 = P->dz * P->dz;
    int nw = (int)(P->direct / (P->fdom * dt));                 /* rtm_model.cu:22-23 */
    int tlength = (int)(P->direct0 / dt);
    int coff = (P->flag_layout == 3 || P->field) ? g->noffset : 0;   /* mute2 shifts by noffset */
    for (int ir = 0; ir < g->nrec; ir++) {
        int col = ir + npml + coff;
        float vtemp = (float)(0.482 * (1.0 / (double)sqrtf(vvpad[(size_t)npml * nxp + col]) +
                                       1.0 / (double)sqrtf(vvpad[(size_t)g->nsz * nxp + g->nsx])));
        float dist = (col - g->nsx) * (col - g->nsx) * dx2 + (npml - g->nsz) * (npml - g->nsz) * dz2;
        dist = sqrtf(dist);
        int temp = tlength < nt ? tlength : nt;
        int c0 = (int)(dist * vtemp * _dt) + nw;
        int c = c0 < nt ? c0 : nt;
        c = c > temp ? c : temp;
        c = c < nt ? c : nt;
        cut[ir] = c;
    }
    free(vv);
}

/* cuda_ricker_wavelet (kernel.cu:13-22) in its float/double mixture */
static float orc_ricker_f32(int it, float fdom, float dt)
{
    float temp = (float)(ORC_PI * (double)fdom * fabs((double)(it * dt) - 1.0 / (double)fdom));
    temp *= temp;
    return (float)((1.0 - 2.0 * (double)temp) * (double)expf(-temp));
}

ORC_EXPORT void orc_ricker(float *w, int nt, float fdom, float dt)
{
    for (int it = 0; it < nt; it++) w[it] = orc_ricker_f32(it, fdom, dt);
}

/* cuda_taper_calculate (kernel.cu:1591-1604), alpha = 0.06 (rtm_model.cu:71) */
static float orc_taper(int ix, int nxl, int nwin)
{
    float alpha = 0.06f;
    float a = -alpha * alpha;
    if (ix < nwin) { float d = (float)(nwin - ix); a = a * d; a = a * d; return expf(a); }
    if (ix >= nxl - nwin && ix < nxl) { float d = (float)(ix - nxl + nwin + 1); a = a * d; a = a * d; return expf(a); }
    return 1.0f;
}

ORC_EXPORT void orc_taper_row(float *out, int nxl, int nwin)
{
    for (int ix = 0; ix < nxl; ix++) out[ix] = orc_taper(ix, nxl, nwin);
}

/* laplacefd1d (image_function.cpp:44-79); dx, dz are int as in the source */
ORC_EXPORT void orc_laplacian(float *image, int nx, int nz, int dx, int dz)
{
    float *c = (float *)malloc((size_t)nx * nz * sizeof(float));
    memcpy(c, image, (size_t)nx * nz * sizeof(float));
    for (int ix = 1; ix < nx - 1; ix++)
        for (int iz = 1; iz < nz - 1; iz++) {
            float t1 = (float)(((double)(c[(size_t)iz * nx + ix + 1] + c[(size_t)iz * nx + ix - 1]) -
                                2.0 * (double)c[(size_t)iz * nx + ix]) / (double)(dx * dx));
            float t2 = (float)(((double)(c[(size_t)(iz + 1) * nx + ix] + c[(size_t)(iz - 1) * nx + ix]) -
                                2.0 * (double)c[(size_t)iz * nx + ix]) / (double)(dz * dz));
            image[(size_t)iz * nx + ix] = t1 + t2;
        }
    for (int ix = 0; ix < nx; ix++) { image[ix] = 0.0f; image[(size_t)(nz - 1) * nx + ix] = 0.0f; }
    for (int iz = 0; iz < nz; iz++) { image[(size_t)iz * nx] = 0.0f; image[(size_t)iz * nx + nx - 1] = 0.0f; }
    free(c);
}

/* cuda_applyimagingconditionmulti (kernel.cu:1580-1590) */
ORC_EXPORT void orc_normalize_multi(float *out, const float *sg, const float *ss, size_t n)
{
    for (size_t i = 0; i < n; i++) out[i] = (float)((double)sg[i] / ((double)ss[i] + 1.0e-6));
}

/* scaleimage1d (image_function.cpp:11-20) */
ORC_EXPORT void orc_scaleimage(const float *illum, float *grad, size_t n)
{
    float m = fabsf(illum[0]);
    for (size_t i = 0; i < n; i++) m = fabsf(illum[i]) > m ? fabsf(illum[i]) : m;
    float mi = (float)(0.0001 * (double)m);
    for (size_t i = 0; i < n; i++) grad[i] /= illum[i] + mi;
}

/* Liu-Yang optimised staggered coefficients (FD_function.cpp:73-165)
 * assembled as RTM_GPU_velstr.cpp:184-196: row r-1 holds order 2r. */
ORC_EXPORT void orc_coef_table(float *out)
{
    static const float t[10][10] = {
        {1.0f},
        {0.1129042e+1f, -0.4301412e-1f},
        {0.1185991e+1f, -0.7249965e-1f, 0.6301572e-2f},
        {0.1217990e+1f, -0.9382142e-1f, 0.1507536e-1f, -0.1700324e-2f},
        {0.1236607e+1f, -0.1082265e+0f, 0.2343440e-1f, -0.5033546e-2f, 0.6817483e-3f},
        {0.1247662e+1f, -0.1175538e+0f, 0.2997970e-1f, -0.8719078e-2f, 0.2215897e-2f, -0.3462075e-3f},
        {0.1254799e+1f, -0.1238928e+0f, 0.3494371e-1f, -0.1208897e-1f, 0.4132531e-2f, -0.1197110e-2f, 0.2122227e-3f},
        {0.1259312e+1f, -0.1280347e+0f, 0.3841945e-1f, -0.1473229e-1f, 0.5924913e-2f, -0.2248618e-2f, 0.7179226e-3f, -0.1400855e-3f},
        {0.1262502e+1f, -0.1310244e+0f, 0.4103928e-1f, -0.1686807e-1f, 0.7530520e-2f, -0.3345071e-2f, 0.1380367e-2f, -0.4808410e-3f, 0.1023759e-3f},
        {0.1264748e+1f, -0.1331606e+0f, 0.4296909e-1f, -0.1851897e-1f, 0.8861071e-2f, -0.4347073e-2f, 0.2076101e-2f, -0.9164925e-3f, 0.3437446e-3f, -0.7874250e-4f},
    };
    for (int r = 0; r < 10; r++)
        for (int i = 0; i < 10; i++) out[r * 10 + i] = t[r][i];
}

/* ------------------------------------------------------------------ */
/* The library's 16-bit boundary store (not reference semantics).  Each step's
 * strips are four blocks (vx lr, vz tb, p lr, p tb); a block is stored as
 * codes of value / scale with scale = max(maxabs / 32768, FLT_MIN) (1 for an
 * all-zero block; NaN values do not raise the max), rounded to nearest even
 * into IEEE binary16 or bfloat16, and read back as code * scale.  The
 * conversions below follow the IEEE 754 definitions of the two formats; the
 * bfloat16 one rounds the top 16 bits of the binary32 pattern (ties to even)
 * as the device code does, NaN patterns included. */
static uint16_t orc_f2h(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u, a = x & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) return (uint16_t)(sign | (a > 0x7F800000u ? 0x7E00u : 0x7C00u));
    if (a >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);     /* 65520 and up: to infinity */
    if (a < 0x38800000u) {                                        /* below 2^-14: subnormal */
        float af;
        memcpy(&af, &a, 4);
        return (uint16_t)(sign | (uint32_t)rintf(af * 16777216.0f));   /* units of 2^-24, exact scaling */
    }
    uint32_t r = a - 0x38000000u;                                 /* exponent bias 127 -> 15 */
    r += 0x0FFFu + ((r >> 13) & 1u);
    return (uint16_t)(sign | (r >> 13));
}
static float orc_h2f(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t x;
    if (e == 0) {
        float f = (float)m * 5.9604644775390625e-8f;            /* m * 2^-24, exact */
        memcpy(&x, &f, 4);
    } else if (e == 31) {
        x = 0x7F800000u | (m << 13);
    } else {
        x = ((e + 112u) << 23) | (m << 13);
    }
    x |= sign;
    float f;
    memcpy(&f, &x, 4);
    return f;
}
static uint16_t orc_f2bf(float f)
{
    uint32_t w;
    memcpy(&w, &f, 4);
    w += 0x7FFFu + ((w >> 16) & 1u);
    return (uint16_t)(w >> 16);
}
static float orc_bf2f(uint16_t h)
{
    const uint32_t w = (uint32_t)h << 16;
    float f;
    memcpy(&f, &w, 4);
    return f;
}
/* one block in place: values replaced by what the store reads back; codes
 * (optional) receive the 16-bit patterns; returns the scale */
ORC_EXPORT float orc_store16_block(float *v, long long n, int fmt, uint16_t *codes)
{
    float mx = 0.0f;
    for (long long i = 0; i < n; i++) {
        const float a = fabsf(v[i]);
        if (a > mx) mx = a;
    }
    const float sc = mx > 0.0f ? fmaxf(mx * (1.0f / 32768.0f), 1.17549435e-38f) : 1.0f;
    for (long long i = 0; i < n; i++) {
        const float q = v[i] / sc;
        const uint16_t c = fmt == 1 ? orc_f2h(q) : orc_f2bf(q);
        if (codes) codes[i] = c;
        v[i] = (fmt == 1 ? orc_h2f(c) : orc_bf2f(c)) * sc;
    }
    return sc;
}

/* ------------------------------------------------------------------ */
#define RT float
#define SFX(n) n##_f32
#define FMA fmaf
#define ORACLE_IS_F64 0
#define ORACLE_KORDER 0
#include "rtm_oracle_impl.h"
#undef SFX
#undef ORACLE_KORDER

#define SFX(n) n##_k32
#define ORACLE_KORDER 1
#include "rtm_oracle_impl.h"
#undef RT
#undef SFX
#undef FMA
#undef ORACLE_IS_F64
#undef ORACLE_KORDER

#define RT double
#define SFX(n) n##_f64
#define FMA fma
#define ORACLE_IS_F64 1
#define ORACLE_KORDER 0
#include "rtm_oracle_impl.h"
#undef RT
#undef SFX
#undef FMA
#undef ORACLE_IS_F64
#undef ORACLE_KORDER
