/*
 * rtm_oracle_impl.h -- body of the CPU oracle, included twice by rtm_oracle.c:
 *   RT = float  (SFX = _f32): the FP32 same-order restatement of the reference;
 *   RT = double (SFX = _f64): the same discrete algorithm in FP64, the error
 *                             yardstick for reduced-precision forms.
 *
 * TEST INFRASTRUCTURE ONLY (see rtm_oracle.c).
 *
 * Canonical FP32 operation order.  Every statement follows the reference's
 * CUDA source literally, one IEEE rounding per C operator, except three
 * patterns that nvcc's default --fmad=true contracts and that this oracle
 * therefore writes as a fused multiply-add:
 *   - stencil accumulation   s += c_i*(a_i - b_i)      -> s = fma(c_i, a_i-b_i, s)
 *   - PML damping            (c1*v - d)/c2              -> fma(c1, v, -d)/c2
 *   - injection / imaging    p += a*w ; g += x*y        -> fma(a, w, p) ; fma(x, y, g)
 * All other products, sums and divisions are rounded separately, in source
 * order.  Double-precision literals of the source (1.0, 0.5, 2.0, eps=1e-6)
 * keep their double arithmetic.  Bit-level parity with the CUDA binary itself
 * is not pinnable (it cannot be built or run here).
 */

/* --------------------------------------------------------------------- */
/* per-shot state                                                         */
/* --------------------------------------------------------------------- */
typedef struct {
    int nxp, nzp, npml, nxl, nzi;      /* padded dims, pml, interior dims */
    RT *p, *vx, *vz;                   /* p: nzp*nxp ; vx: nzp*(nxp-1) ; vz: (nzp-1)*nxp */
    RT *pl1, *pl2, *pr1, *pr2;         /* nzp*npml  (left/right strips incl. corners)    */
    RT *pt1, *pt2, *pb1, *pb2;         /* npml*nxl  (top/bottom strips, interior columns) */
} SFX(wfield);

typedef struct {
    int nxp, nzp, npml, nxl, nzi;
    RT dt, dtx, dtz;
    RT coef[100];                       /* diffcoef[(r-1)*10 + i-1]                  */
    RT *rho, *vv;                       /* padded window: density, v^2 (nzp*nxp)     */
    float *rhf, *vvf;                   /* the same windows in float, as smoothed    */
    RT *ib, *kk;                        /* kernel order: 1/rho, rho*v^2 (nzp*nxp)    */
    RT *ax, *bx, *az, *bz;              /* kernel order: c1/c2, 1/c2 at p nodes      */
    RT *axv, *bxv, *azv, *bzv;          /* ... at vx / vz half nodes                 */
    RT *ddx, *ddz, *ddxv, *ddzv;        /* pml profiles: nxp, nzp, nxp-1, nzp-1      */
    RT *c1x, *c2x, *c1z, *c2z;          /* derived damping factors at p nodes        */
    RT *c1xv, *c2xv, *c1zv, *c2zv;      /* ... at vx / vz half nodes                 */
    int *norder, *norderx, *norderz;    /* literal-mode order maps (kernel.cu:271-317) */
    RT *src;                            /* Gaussian source, full padded grid         */
    RT *wav;                            /* Ricker, nt                                */
} SFX(medium);

static void SFX(wf_alloc)(SFX(wfield) *w, int nxp, int nzp, int npml)
{
    w->nxp = nxp; w->nzp = nzp; w->npml = npml;
    w->nxl = nxp - 2 * npml; w->nzi = nzp - 2 * npml;
    w->p  = (RT *)calloc((size_t)nzp * nxp, sizeof(RT));
    w->vx = (RT *)calloc((size_t)nzp * (nxp - 1), sizeof(RT));
    w->vz = (RT *)calloc((size_t)(nzp - 1) * nxp, sizeof(RT));
    w->pl1 = (RT *)calloc((size_t)nzp * npml, sizeof(RT));
    w->pl2 = (RT *)calloc((size_t)nzp * npml, sizeof(RT));
    w->pr1 = (RT *)calloc((size_t)nzp * npml, sizeof(RT));
    w->pr2 = (RT *)calloc((size_t)nzp * npml, sizeof(RT));
    w->pt1 = (RT *)calloc((size_t)npml * w->nxl, sizeof(RT));
    w->pt2 = (RT *)calloc((size_t)npml * w->nxl, sizeof(RT));
    w->pb1 = (RT *)calloc((size_t)npml * w->nxl, sizeof(RT));
    w->pb2 = (RT *)calloc((size_t)npml * w->nxl, sizeof(RT));
}

static void SFX(wf_zero)(SFX(wfield) *w)
{
    size_t nxp = w->nxp, nzp = w->nzp, npml = w->npml, nxl = w->nxl;
    memset(w->p, 0, nzp * nxp * sizeof(RT));
    memset(w->vx, 0, nzp * (nxp - 1) * sizeof(RT));
    memset(w->vz, 0, (nzp - 1) * nxp * sizeof(RT));
    memset(w->pl1, 0, nzp * npml * sizeof(RT)); memset(w->pl2, 0, nzp * npml * sizeof(RT));
    memset(w->pr1, 0, nzp * npml * sizeof(RT)); memset(w->pr2, 0, nzp * npml * sizeof(RT));
    memset(w->pt1, 0, npml * nxl * sizeof(RT)); memset(w->pt2, 0, npml * nxl * sizeof(RT));
    memset(w->pb1, 0, npml * nxl * sizeof(RT)); memset(w->pb2, 0, npml * nxl * sizeof(RT));
}

static void SFX(wf_free)(SFX(wfield) *w)
{
    free(w->p); free(w->vx); free(w->vz);
    free(w->pl1); free(w->pl2); free(w->pr1); free(w->pr2);
    free(w->pt1); free(w->pt2); free(w->pb1); free(w->pb2);
}

static void SFX(med_free)(SFX(medium) *m)
{
    free(m->rho); free(m->vv); free(m->ib); free(m->kk); free(m->rhf); free(m->vvf);
    free(m->ax); free(m->bx); free(m->az); free(m->bz);
    free(m->axv); free(m->bxv); free(m->azv); free(m->bzv);
    free(m->ddx); free(m->ddz); free(m->ddxv); free(m->ddzv);
    free(m->c1x); free(m->c2x); free(m->c1z); free(m->c2z);
    free(m->c1xv); free(m->c2xv); free(m->c1zv); free(m->c2zv);
    free(m->norder); free(m->norderx); free(m->norderz);
    free(m->src); free(m->wav);
}

/* PML profile, kernel.cu:174-269.  half = 0 at p nodes, 1 at velocity
 * half nodes (the "- 0.5" of cuda_pmlCoeffvx/vz).  n is the array length the
 * kernel was launched with (nxp, nzp, nxp-1 or nzp-1). */
static void SFX(pml_profile)(RT *d, int n, int npml, RT vpmax, float h, int half)
{
#if ORACLE_IS_F64
    double dd = (double)npml * (double)h;
    double d0 = 1.5 * (double)vpmax * log(1.0 / 1.0e-10) / dd;
#else
    float R = 1.0e-10f;
    float dd = npml * h;
    float d0 = (float)(1.5 * (double)vpmax * (double)logf(1 / R) / (double)dd);
#endif
    for (int i = 0; i < n; i++) {
        RT t;
        if (i < npml) {
            t = half ? (RT)((double)(RT)(npml - i) - 0.5) : (RT)(npml - i);
            t = t / (RT)npml; t = t * t; d[i] = (RT)d0 * t;
        } else if (i < n - npml) {
            d[i] = 0;
        } else {
            t = half ? (RT)((double)(RT)(i - n + npml + 1) - 0.5) : (RT)(i - n + npml + 1);
            t = t / (RT)npml; t = t * t; d[i] = (RT)d0 * t;
        }
    }
}

/* c1 = 1 - 0.5*dt*d ; c2 = 2 - c1, evaluated in double as the source's
 * double literals make them (kernel.cu:592-593), then rounded to RT. */
static void SFX(damp_factors)(const RT *d, int n, float dt, RT *c1, RT *c2)
{
    for (int i = 0; i < n; i++) {
        c1[i] = (RT)(1.0 - 0.5 * (double)dt * (double)d[i]);
        c2[i] = (RT)(2.0 - (double)c1[i]);
    }
}

/* --------------------------------------------------------------------- */
/* medium set-up for one shot (rtm_model.cu:156-263)                      */
/* --------------------------------------------------------------------- */
static void SFX(med_setup)(SFX(medium) *m, const orc_geom *g, const orc_params *P,
                           const float *vp0, const float *rho0, const float *diffcoef,
                           const float *vvpad_in, const float *rhopad_in)
{
    int nxp = g->nxp, nzp = g->nzp, npml = P->npml;
    size_t nxz = (size_t)nxp * nzp;
    memset(m, 0, sizeof(*m));
    m->nxp = nxp; m->nzp = nzp; m->npml = npml; m->nxl = g->nxl; m->nzi = P->nz;
    m->dt = (RT)P->dt;
#if ORACLE_IS_F64
    m->dtx = (double)P->dt / (double)P->dx;
    m->dtz = (double)P->dt / (double)P->dz;
#else
    m->dtx = P->dt / P->dx;                     /* rtm_model.cu:73-74 */
    m->dtz = P->dt / P->dz;
#endif
    for (int i = 0; i < 100; i++) m->coef[i] = (RT)diffcoef[i];

    /* extractvel1 / extractrho1 / extendvel1 (Velocity_function.cpp:2011-2073):
     * window [nx1,nx2], v stored squared, edge replication into the PML. */
    float *vvf = (float *)malloc(nxz * sizeof(float));
    float *rhf = (float *)malloc(nxz * sizeof(float));
    if (vvpad_in) {
        memcpy(vvf, vvpad_in, nxz * sizeof(float));
        memcpy(rhf, rhopad_in, nxz * sizeof(float));
    } else {
        orc_window_pad(vvf, vp0, P->nx, P->nz, g->nx1, g->nx2, npml, 1);
        orc_window_pad(rhf, rho0, P->nx, P->nz, g->nx1, g->nx2, npml, 0);
    }
    /* pml layers smooth (rtm_model.cu:204-206), before vpmax */
    if (P->medium_flag) {
        orc_pml_smooth(vvf, nxp, nzp, npml);
        orc_pml_smooth(rhf, nxp, nzp, npml);
    }
    m->vv = (RT *)malloc(nxz * sizeof(RT));
    m->rho = (RT *)malloc(nxz * sizeof(RT));
    float vmax2 = vvf[0];
    for (size_t i = 0; i < nxz; i++) {
        m->vv[i] = (RT)vvf[i]; m->rho[i] = (RT)rhf[i];
        if (vvf[i] > vmax2) vmax2 = vvf[i];     /* Maxval1, function.cpp:130-138 */
    }
#if ORACLE_IS_F64
    RT vpmax = sqrt((double)vmax2);
#else
    RT vpmax = sqrtf(vmax2);                    /* rtm_model.cu:207 */
#endif
    m->vvf = vvf; m->rhf = rhf;

    m->ddx = (RT *)malloc(nxp * sizeof(RT));  m->ddz = (RT *)malloc(nzp * sizeof(RT));
    m->ddxv = (RT *)malloc((nxp - 1) * sizeof(RT)); m->ddzv = (RT *)malloc((nzp - 1) * sizeof(RT));
    SFX(pml_profile)(m->ddx, nxp, npml, vpmax, P->dx, 0);
    SFX(pml_profile)(m->ddz, nzp, npml, vpmax, P->dz, 0);
    SFX(pml_profile)(m->ddxv, nxp - 1, npml, vpmax, P->dx, 1);
    SFX(pml_profile)(m->ddzv, nzp - 1, npml, vpmax, P->dz, 1);
    m->c1x = (RT *)malloc(nxp * sizeof(RT)); m->c2x = (RT *)malloc(nxp * sizeof(RT));
    m->c1z = (RT *)malloc(nzp * sizeof(RT)); m->c2z = (RT *)malloc(nzp * sizeof(RT));
    m->c1xv = (RT *)malloc(nxp * sizeof(RT)); m->c2xv = (RT *)malloc(nxp * sizeof(RT));
    m->c1zv = (RT *)malloc(nzp * sizeof(RT)); m->c2zv = (RT *)malloc(nzp * sizeof(RT));
    SFX(damp_factors)(m->ddx, nxp, P->dt, m->c1x, m->c2x);
    SFX(damp_factors)(m->ddz, nzp, P->dt, m->c1z, m->c2z);
    SFX(damp_factors)(m->ddxv, nxp - 1, P->dt, m->c1xv, m->c2xv);
    SFX(damp_factors)(m->ddzv, nzp - 1, P->dt, m->c1zv, m->c2zv);
    /* kernel order: per-cell reciprocal density and bulk modulus, damping as
     * ratio and reciprocal (c1 = c2 = 1 gives a = b = 1, so the damped form
     * fma(a, v, -(u*b)) is v - u off the absorbing layer) */
    m->ib = (RT *)malloc(nxz * sizeof(RT)); m->kk = (RT *)malloc(nxz * sizeof(RT));
    for (size_t i = 0; i < nxz; i++) { m->ib[i] = (RT)1 / m->rho[i]; m->kk[i] = m->rho[i] * m->vv[i]; }
    m->ax = (RT *)malloc(nxp * sizeof(RT)); m->bx = (RT *)malloc(nxp * sizeof(RT));
    m->az = (RT *)malloc(nzp * sizeof(RT)); m->bz = (RT *)malloc(nzp * sizeof(RT));
    m->axv = (RT *)malloc(nxp * sizeof(RT)); m->bxv = (RT *)malloc(nxp * sizeof(RT));
    m->azv = (RT *)malloc(nzp * sizeof(RT)); m->bzv = (RT *)malloc(nzp * sizeof(RT));
    for (int i = 0; i < nxp; i++) { m->ax[i] = m->c1x[i] / m->c2x[i]; m->bx[i] = (RT)1 / m->c2x[i]; }
    for (int i = 0; i < nzp; i++) { m->az[i] = m->c1z[i] / m->c2z[i]; m->bz[i] = (RT)1 / m->c2z[i]; }
    for (int i = 0; i < nxp - 1; i++) { m->axv[i] = m->c1xv[i] / m->c2xv[i]; m->bxv[i] = (RT)1 / m->c2xv[i]; }
    for (int i = 0; i < nzp - 1; i++) { m->azv[i] = m->c1zv[i] / m->c2zv[i]; m->bzv[i] = (RT)1 / m->c2zv[i]; }

    /* order maps, kernel.cu:271-317 (closed forms; memset 0 elsewhere) */
    m->norder = (int *)calloc(nxz, sizeof(int));
    m->norderx = (int *)calloc(nxp - 1, sizeof(int));
    m->norderz = (int *)calloc(nzp - 1, sizeof(int));
    for (int iz = 1; iz < nzp - 1; iz++)
        for (int ix = 1; ix < nxp - 1; ix++)
            m->norder[(size_t)iz * nxp + ix] = 2 * orc_imin(orc_imin(orc_imin(R10, ix), orc_imin(iz, nxp - 1 - ix)), nzp - 1 - iz);
    for (int i = 0; i < nxp - 1; i++) m->norderx[i] = 2 * orc_imin(R10, orc_imin(i + 1, nxp - 1 - i));
    for (int i = 0; i < nzp - 1; i++) m->norderz[i] = 2 * orc_imin(R10, orc_imin(i + 1, nzp - 1 - i));

    /* cuda_source (kernel.cu:23-35) */
    m->src = (RT *)malloc(nxz * sizeof(RT));
    float dx2 = P->dx * P->dx, dz2 = P->dz * P->dz;
    for (int iz = 0; iz < nzp; iz++)
        for (int ix = 0; ix < nxp; ix++) {
#if ORACLE_IS_F64
            double dist = (double)((ix - g->nsx) * (ix - g->nsx)) * (double)dx2 + (double)((iz - g->nsz) * (iz - g->nsz)) * (double)dz2;
            m->src[(size_t)iz * nxp + ix] = (double)P->amp * exp(-(double)P->alp * (double)P->alp * dist);
#else
            float dist = (ix - g->nsx) * (ix - g->nsx) * dx2 + (iz - g->nsz) * (iz - g->nsz) * dz2;
            m->src[(size_t)iz * nxp + ix] = P->amp * expf(-P->alp * P->alp * dist);
#endif
        }
    /* cuda_ricker_wavelet (kernel.cu:13-22) */
    m->wav = (RT *)malloc((size_t)P->nt * sizeof(RT));
    for (int it = 0; it < P->nt; it++) {
#if ORACLE_IS_F64
        double t = ORC_PI * (double)P->fdom * fabs((double)it * (double)P->dt - 1.0 / (double)P->fdom);
        t = t * t;
        m->wav[it] = (1.0 - 2.0 * t) * exp(-t);
#else
        m->wav[it] = orc_ricker_f32(it, P->fdom, P->dt);
#endif
    }
}

/* --------------------------------------------------------------------- */
/* generalised forward step: every cell of the padded grid, any npml>=10  */
/* --------------------------------------------------------------------- */
#define PX(w, iz, ix)  (w)->p[(size_t)(iz) * nxp + (ix)]
#define VX(w, iz, ix)  (w)->vx[(size_t)(iz) * (nxp - 1) + (ix)]
#define VZ(w, iz, ix)  (w)->vz[(size_t)(iz) * nxp + (ix)]

/* velocity half-step.  vx: all rows, columns 0..nxp-2 (kernel.cu:430-465 for
 * the interior, :542-630 for the pml columns); vz: rows 0..nzp-2, all columns
 * (kernel.cu:466-500, :631-713).  Local radius r = min(10, dist to edge). */
static void SFX(gen_vel)(SFX(wfield) *w, const SFX(medium) *m)
{
    const int nxp = w->nxp, nzp = w->nzp;
    const RT dtx = m->dtx, dtz = m->dtz;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nzp; iz++) {
        for (int ix = 0; ix <= nxp - 2; ix++) {
            int r = orc_imin(R10, orc_imin(ix + 1, nxp - 1 - ix));
            const RT *c = &m->coef[(r - 1) * 10];
            RT s = 0;
            for (int i = 1; i <= r; i++) s = FMA(c[i - 1], PX(w, iz, ix + i) - PX(w, iz, ix - i + 1), s);
#if ORACLE_KORDER
            const RT u = (s * dtx) * m->ib[(size_t)iz * nxp + ix];
            VX(w, iz, ix) = FMA(m->axv[ix], VX(w, iz, ix), -(u * m->bxv[ix]));
#else
            s = s * dtx; s = s / m->rho[(size_t)iz * nxp + ix];
            VX(w, iz, ix) = FMA(m->c1xv[ix], VX(w, iz, ix), -s) / m->c2xv[ix];
#endif
        }
    }
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz <= nzp - 2; iz++) {
        int r = orc_imin(R10, orc_imin(iz + 1, nzp - 1 - iz));
        const RT *c = &m->coef[(r - 1) * 10];
        for (int ix = 0; ix < nxp; ix++) {
            RT s = 0;
            for (int i = 1; i <= r; i++) s = FMA(c[i - 1], PX(w, iz + i, ix) - PX(w, iz - i + 1, ix), s);
#if ORACLE_KORDER
            const RT u = (s * dtz) * m->ib[(size_t)iz * nxp + ix];
            VZ(w, iz, ix) = FMA(m->azv[iz], VZ(w, iz, ix), -(u * m->bzv[iz]));
#else
            s = s * dtz; s = s / m->rho[(size_t)iz * nxp + ix];
            VZ(w, iz, ix) = FMA(m->c1zv[iz], VZ(w, iz, ix), -s) / m->c2zv[iz];
#endif
        }
    }
}

/* storage of the split pressure pair for a pml cell (iz, ix): left/right
 * column strips hold the corners (kernel.cu:780-783, :1028-1031), top/bottom
 * strips the interior columns (kernel.cu:906-909, :957-960). */
static void SFX(split_ptr)(SFX(wfield) *w, int iz, int ix, RT **q1, RT **q2)
{
    const int nxp = w->nxp, nzp = w->nzp, npml = w->npml, nxl = w->nxl;
    size_t k;
    if (ix < npml) { k = (size_t)iz * npml + ix; *q1 = &w->pl1[k]; *q2 = &w->pl2[k]; }
    else if (ix >= nxp - npml) { k = (size_t)iz * npml + ix - nxp + npml; *q1 = &w->pr1[k]; *q2 = &w->pr2[k]; }
    else if (iz < npml) { k = (size_t)iz * nxl + ix - npml; *q1 = &w->pt1[k]; *q2 = &w->pt2[k]; }
    else { k = (size_t)(iz - nzp + npml) * nxl + ix - npml; *q1 = &w->pb1[k]; *q2 = &w->pb2[k]; }
}

/* pressure step.  Interior (kernel.cu:501-541): p -= dxv + dzv.  PML cells
 * (kernel.cu:714-1185): split pair, p = p1 + p2.  Outer ring untouched. */
static void SFX(gen_pres)(SFX(wfield) *w, const SFX(medium) *m)
{
    const int nxp = w->nxp, nzp = w->nzp, npml = w->npml;
    const RT dtx = m->dtx, dtz = m->dtz;
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nzp - 2; iz++) {
        for (int ix = 1; ix <= nxp - 2; ix++) {
            int r = orc_imin(orc_imin(orc_imin(R10, ix), orc_imin(iz, nxp - 1 - ix)), nzp - 1 - iz);
            const RT *c = &m->coef[(r - 1) * 10];
            RT sx = 0, sz = 0;
            for (int i = 1; i <= r; i++) {
                sx = FMA(c[i - 1], VX(w, iz, ix + i - 1) - VX(w, iz, ix - i), sx);
                sz = FMA(c[i - 1], VZ(w, iz + i - 1, ix) - VZ(w, iz - i, ix), sz);
            }
            size_t id = (size_t)iz * nxp + ix;
            int inner = (ix >= npml && ix < nxp - npml && iz >= npml && iz < nzp - npml);
#if ORACLE_KORDER
            if (inner) {
                w->p[id] = FMA(-m->kk[id], FMA(sx, dtx, sz * dtz), w->p[id]);
            } else {
                const RT dxv = (sx * dtx) * m->kk[id], dzv = (sz * dtz) * m->kk[id];
                RT *q1, *q2;
                SFX(split_ptr)(w, iz, ix, &q1, &q2);
                *q1 = FMA(m->ax[ix], *q1, -(dxv * m->bx[ix]));
                *q2 = FMA(m->az[iz], *q2, -(dzv * m->bz[iz]));
                w->p[id] = *q1 + *q2;
            }
#else
            RT dxv = sx * dtx; dxv = dxv * m->rho[id]; dxv = dxv * m->vv[id];
            RT dzv = sz * dtz; dzv = dzv * m->rho[id]; dzv = dzv * m->vv[id];
            if (inner) {
                w->p[id] = w->p[id] - (dxv + dzv);
            } else {
                RT *q1, *q2;
                SFX(split_ptr)(w, iz, ix, &q1, &q2);
                *q1 = FMA(m->c1x[ix], *q1, -dxv) / m->c2x[ix];
                *q2 = FMA(m->c1z[iz], *q2, -dzv) / m->c2z[iz];
                w->p[id] = *q1 + *q2;
            }
#endif
        }
    }
}

/* cuda_add_source (kernel.cu:36-47): p +- dt*src*w[it] on the whole grid */
static void SFX(add_source)(RT *p, const SFX(medium) *m, int it, int sign)
{
    size_t nxz = (size_t)m->nxp * m->nzp;
    RT wv = m->wav[it];
    for (size_t i = 0; i < nxz; i++) {
        RT a = m->dt * m->src[i];
        p[i] = sign > 0 ? FMA(a, wv, p[i]) : FMA(-a, wv, p[i]);
    }
}

/* --------------------------------------------------------------------- */
/* literal restatement of the reference kernels (valid for npml == 32)    */
/* Thread index ranges and guards follow kernel.cu; the shared-memory     */
/* tiles only cache global values, so reads are written against global.  */
/* --------------------------------------------------------------------- */
static void SFX(lit_vel)(SFX(wfield) *w, const SFX(medium) *m)
{
    const int nxp = w->nxp, nzp = w->nzp, npml = w->npml;
    const RT dtx = m->dtx, dtz = m->dtz;
    const RT *c10 = &m->coef[90];
    /* cuda_forward_vx: idz in [0,nzp), idx in [npml, nxp-npml-2] */
    for (int iz = 0; iz < nzp; iz++)
        for (int ix = npml; ix < nxp - npml - 1; ix++) {
            RT s = 0;
            for (int i = 1; i <= R10; i++) s = FMA(c10[i - 1], PX(w, iz, ix + i) - PX(w, iz, ix - i + 1), s);
            s = s * dtx; s = s / m->rho[(size_t)iz * nxp + ix];
            VX(w, iz, ix) -= s;
        }
    /* cuda_forward_vz: idz in [npml, nzp-npml-2], idx in [0,nxp) */
    for (int iz = npml; iz < nzp - npml - 1; iz++)
        for (int ix = 0; ix < nxp; ix++) {
            RT s = 0;
            for (int i = 1; i <= R10; i++) s = FMA(c10[i - 1], PX(w, iz + i, ix) - PX(w, iz - i + 1, ix), s);
            s = s * dtz; s = s / m->rho[(size_t)iz * nxp + ix];
            VZ(w, iz, ix) -= s;
        }
    /* cuda_pml_vxlr: grid (.., 2) of 32x32 blocks */
    for (int by = 0; by < 2; by++)
        for (int iz = 0; iz < nzp; iz++)
            for (int ty = 0; ty < 32; ty++) {
                int ix = ty + 32 * by;
                int ok;
                if (by == 0) ok = ix < npml;
                else { ix = nxp - 2 * npml + ix - 1; ok = (ix >= nxp - npml - 1 && ix < nxp - 1); }
                if (!ok) continue;
                int r = m->norderx[ix] / 2;
                RT s = 0;
                for (int i = 1; i <= r; i++) s = FMA(m->coef[(r - 1) * 10 + i - 1], PX(w, iz, ix + i) - PX(w, iz, ix - i + 1), s);
                s = s * dtx; s = s / m->rho[(size_t)iz * nxp + ix];
                RT c1 = (RT)(1.0 - 0.5 * (double)m->dt * (double)m->ddxv[ix]);
                RT c2 = (RT)(2.0 - (double)c1);
                VX(w, iz, ix) = FMA(c1, VX(w, iz, ix), -s) / c2;
            }
    /* cuda_pml_vztb: grid (2, ..) */
    for (int bx = 0; bx < 2; bx++)
        for (int tx = 0; tx < 32; tx++)
            for (int ix = 0; ix < nxp; ix++) {
                int iz = tx + 32 * bx;
                int ok;
                if (bx == 0) ok = iz < npml;
                else { iz = nzp - 2 * npml + iz - 1; ok = (iz >= nzp - npml - 1 && iz < nzp - 1); }
                if (!ok) continue;
                int r = m->norderz[iz] / 2;
                RT s = 0;
                for (int i = 1; i <= r; i++) s = FMA(m->coef[(r - 1) * 10 + i - 1], PX(w, iz + i, ix) - PX(w, iz - i + 1, ix), s);
                s = s * dtz; s = s / m->rho[(size_t)iz * nxp + ix];
                RT c1 = (RT)(1.0 - 0.5 * (double)m->dt * (double)m->ddzv[iz]);
                RT c2 = (RT)(2.0 - (double)c1);
                VZ(w, iz, ix) = FMA(c1, VZ(w, iz, ix), -s) / c2;
            }
}

/* one split-pressure cell as written in cuda_pml_plr/ptb/pconner */
static void SFX(lit_pcell)(SFX(wfield) *w, const SFX(medium) *m, int iz, int ix,
                           RT *q1, RT *q2, int dampx, int dampz)
{
    const int nxp = w->nxp;
    size_t id = (size_t)iz * nxp + ix;
    int r = m->norder[id] / 2;
    RT sx = 0, sz = 0;
    for (int i = 1; i <= r; i++) {
        sx = FMA(m->coef[(r - 1) * 10 + i - 1], VX(w, iz, ix + i - 1) - VX(w, iz, ix - i), sx);
        sz = FMA(m->coef[(r - 1) * 10 + i - 1], VZ(w, iz + i - 1, ix) - VZ(w, iz - i, ix), sz);
    }
    RT dxv = sx * m->dtx; dxv = dxv * m->rho[id]; dxv = dxv * m->vv[id];
    RT dzv = sz * m->dtz; dzv = dzv * m->rho[id]; dzv = dzv * m->vv[id];
    RT c1x = 1, c2x = 1, c1z = 1, c2z = 1;
    if (dampx) { c1x = (RT)(1.0 - 0.5 * (double)m->dt * (double)m->ddx[ix]); c2x = (RT)(2.0 - (double)c1x); }
    if (dampz) { c1z = (RT)(1.0 - 0.5 * (double)m->dt * (double)m->ddz[iz]); c2z = (RT)(2.0 - (double)c1z); }
    *q1 = FMA(c1x, *q1, -dxv) / c2x;
    *q2 = FMA(c1z, *q2, -dzv) / c2z;
    w->p[id] = *q1 + *q2;
}

static void SFX(lit_pres)(SFX(wfield) *w, const SFX(medium) *m)
{
    const int nxp = w->nxp, nzp = w->nzp, npml = w->npml, nxl = w->nxl;
    const RT *c10 = &m->coef[90];
    /* cuda_forward_p: interior */
    for (int iz = npml; iz < nzp - npml; iz++)
        for (int ix = npml; ix < nxp - npml; ix++) {
            RT sx = 0, sz = 0;
            for (int i = 1; i <= R10; i++) {
                sx = FMA(c10[i - 1], VX(w, iz, ix + i - 1) - VX(w, iz, ix - i), sx);
                sz = FMA(c10[i - 1], VZ(w, iz + i - 1, ix) - VZ(w, iz - i, ix), sz);
            }
            size_t id = (size_t)iz * nxp + ix;
            RT dxv = sx * m->dtx; dxv = dxv * m->rho[id]; dxv = dxv * m->vv[id];
            RT dzv = sz * m->dtz; dzv = dzv * m->rho[id]; dzv = dzv * m->vv[id];
            w->p[id] = w->p[id] - (dxv + dzv);
        }
    /* cuda_pml_plr: rows idz = npml + .., grid (.., 2) */
    for (int iz = npml; iz < nzp - npml; iz++)
        for (int ty = 0; ty < 32; ty++) {
            int ix = ty;                                  /* left  */
            if (ix > 0) SFX(lit_pcell)(w, m, iz, ix, &w->pl1[(size_t)iz * npml + ix], &w->pl2[(size_t)iz * npml + ix], 1, 0);
            ix = nxp - 2 * npml + ty + 32;                /* right */
            if (ix < nxp - 1) {
                size_t k = (size_t)iz * npml + ix - nxp + npml;
                SFX(lit_pcell)(w, m, iz, ix, &w->pr1[k], &w->pr2[k], 1, 0);
            }
        }
    /* cuda_pml_ptb: columns idx = npml + .., grid (2, ..) */
    for (int ix = npml; ix < nxp - npml; ix++)
        for (int tx = 0; tx < 32; tx++) {
            int iz = tx;                                  /* top    */
            if (iz > 0) {
                size_t k = (size_t)iz * nxl + ix - npml;
                SFX(lit_pcell)(w, m, iz, ix, &w->pt1[k], &w->pt2[k], 0, 1);
            }
            iz = nzp - 2 * npml + tx + 32;                /* bottom */
            if (iz < nzp - 1) {
                size_t k = (size_t)(iz - nzp + npml) * nxl + ix - npml;
                SFX(lit_pcell)(w, m, iz, ix, &w->pb1[k], &w->pb2[k], 0, 1);
            }
        }
    /* cuda_pml_pconner: grid (2,2) */
    for (int tx = 0; tx < 32; tx++)
        for (int ty = 0; ty < 32; ty++) {
            int iz, ix; size_t k;
            iz = tx; ix = ty;                                              /* left-top */
            if (ix > 0 && iz > 0) { k = (size_t)iz * npml + ix; SFX(lit_pcell)(w, m, iz, ix, &w->pl1[k], &w->pl2[k], 1, 1); }
            iz = tx; ix = nxp - 2 * npml + ty + 32;                        /* right-top */
            if (ix < nxp - 1 && iz > 0) { k = (size_t)iz * npml + ix - nxp + npml; SFX(lit_pcell)(w, m, iz, ix, &w->pr1[k], &w->pr2[k], 1, 1); }
            iz = nzp - 2 * npml + tx + 32; ix = ty;                        /* left-bottom */
            if (ix > 0 && iz < nzp - 1) { k = (size_t)iz * npml + ix; SFX(lit_pcell)(w, m, iz, ix, &w->pl1[k], &w->pl2[k], 1, 1); }
            iz = nzp - 2 * npml + tx + 32; ix = nxp - 2 * npml + ty + 32;  /* right-bottom */
            if (ix < nxp - 1 && iz < nzp - 1) { k = (size_t)iz * npml + ix - nxp + npml; SFX(lit_pcell)(w, m, iz, ix, &w->pr1[k], &w->pr2[k], 1, 1); }
        }
}

/* one forward time step of a wavefield (rtm_model.cu:327-336) */
static void SFX(fwd_step)(SFX(wfield) *w, const SFX(medium) *m, int literal)
{
    if (literal) { SFX(lit_vel)(w, m); SFX(lit_pres)(w, m); }
    else { SFX(gen_vel)(w, m); SFX(gen_pres)(w, m); }
}

/* --------------------------------------------------------------------- */
/* boundary strips (kernel.cu:1186-1400): per step, 4 packed blocks       */
/*   vx  lr : [2][nzi][10]  cols npml-10..npml-1 | nxp-npml-1..nxp-npml+8 */
/*   vz  tb : [2][10][nxl]  rows npml-10..npml-1 | nzp-npml-1..nzp-npml+8 */
/*   p   lr : [2][nzi][10]  cols npml-10..npml-1 | nxp-npml..nxp-npml+9   */
/*   p   tb : [2][10][nxl]  rows npml-10..npml-1 | nzp-npml..nzp-npml+9   */
/* --------------------------------------------------------------------- */
static size_t SFX(strip_len)(int nzi, int nxl) { return (size_t)40 * (nzi + nxl); }

static void SFX(strips_io)(SFX(wfield) *w, RT *buf, int save, int which /*1=vel,2=p,3=both*/)
{
    const int nxp = w->nxp, nzp = w->nzp, npml = w->npml, nxl = w->nxl, nzi = w->nzi;
    RT *bvx = buf, *bvz = buf + 20 * nzi, *bpl = bvz + 20 * nxl, *bpt = bpl + 20 * nzi;
#define IO(dst, src) do { if (save) (dst) = (src); else (src) = (dst); } while (0)
    if (which & 1) {
        for (int iz = 0; iz < nzi; iz++)
            for (int k = 0; k < 10; k++) {
                IO(bvx[iz * 10 + k], VX(w, iz + npml, k - 10 + npml));
                IO(bvx[10 * nzi + iz * 10 + k], VX(w, iz + npml, k + nxp - npml - 1));
            }
        for (int k = 0; k < 10; k++)
            for (int ix = 0; ix < nxl; ix++) {
                IO(bvz[k * nxl + ix], VZ(w, k - 10 + npml, ix + npml));
                IO(bvz[10 * nxl + k * nxl + ix], VZ(w, k + nzp - npml - 1, ix + npml));
            }
    }
    if (which & 2) {
        for (int iz = 0; iz < nzi; iz++)
            for (int k = 0; k < 10; k++) {
                IO(bpl[iz * 10 + k], PX(w, iz + npml, k - 10 + npml));
                IO(bpl[10 * nzi + iz * 10 + k], PX(w, iz + npml, k + nxp - npml));
            }
        for (int k = 0; k < 10; k++)
            for (int ix = 0; ix < nxl; ix++) {
                IO(bpt[k * nxl + ix], PX(w, k - 10 + npml, ix + npml));
                IO(bpt[10 * nxl + k * nxl + ix], PX(w, k + nzp - npml, ix + npml));
            }
    }
#undef IO
}

/* --------------------------------------------------------------------- */
/* reverse reconstruction of the source wavefield (kernel.cu:318-429,     */
/* rtm_model.cu:424-437)                                                   */
/* --------------------------------------------------------------------- */
static void SFX(rev_vel)(SFX(wfield) *w, const SFX(medium) *m)
{
    const int nxp = w->nxp, nzp = w->nzp, npml = w->npml;
    const RT *c10 = &m->coef[90];
    /* cuda_backward_vx: rows [npml, nzp-npml), cols [npml, nxp-npml-2] */
#pragma omp parallel for schedule(static)
    for (int iz = npml; iz < nzp - npml; iz++)
        for (int ix = npml; ix < nxp - npml - 1; ix++) {
            RT s = 0;
            for (int i = 1; i <= R10; i++) s = FMA(c10[i - 1], PX(w, iz, ix + i) - PX(w, iz, ix - i + 1), s);
#if ORACLE_KORDER
            VX(w, iz, ix) = FMA(s * m->dtx, m->ib[(size_t)iz * nxp + ix], VX(w, iz, ix));
#else
            s = s * m->dtx; s = s / m->rho[(size_t)iz * nxp + ix];
            VX(w, iz, ix) += s;
#endif
        }
    /* cuda_backward_vz: rows [npml, nzp-npml-2], cols [npml, nxp-npml) */
#pragma omp parallel for schedule(static)
    for (int iz = npml; iz < nzp - npml - 1; iz++)
        for (int ix = npml; ix < nxp - npml; ix++) {
            RT s = 0;
            for (int i = 1; i <= R10; i++) s = FMA(c10[i - 1], PX(w, iz + i, ix) - PX(w, iz - i + 1, ix), s);
#if ORACLE_KORDER
            VZ(w, iz, ix) = FMA(s * m->dtz, m->ib[(size_t)iz * nxp + ix], VZ(w, iz, ix));
#else
            s = s * m->dtz; s = s / m->rho[(size_t)iz * nxp + ix];
            VZ(w, iz, ix) += s;
#endif
        }
}

static void SFX(rev_pres)(SFX(wfield) *w, const SFX(medium) *m)
{
    const int nxp = w->nxp, nzp = w->nzp, npml = w->npml;
    const RT *c10 = &m->coef[90];
#pragma omp parallel for schedule(static)
    for (int iz = npml; iz < nzp - npml; iz++)
        for (int ix = npml; ix < nxp - npml; ix++) {
            RT sx = 0, sz = 0;
            for (int i = 1; i <= R10; i++) {
                sx = FMA(c10[i - 1], VX(w, iz, ix + i - 1) - VX(w, iz, ix - i), sx);
                sz = FMA(c10[i - 1], VZ(w, iz + i - 1, ix) - VZ(w, iz - i, ix), sz);
            }
            size_t id = (size_t)iz * nxp + ix;
#if ORACLE_KORDER
            w->p[id] = FMA(m->kk[id], FMA(sx, m->dtx, sz * m->dtz), w->p[id]);
#else
            RT dxv = sx * m->dtx; dxv = dxv * m->rho[id]; dxv = dxv * m->vv[id];
            RT dzv = sz * m->dtz; dzv = dzv * m->rho[id]; dzv = dzv * m->vv[id];
            w->p[id] = w->p[id] + (dxv + dzv);
#endif
        }
}
#undef PX
#undef VX
#undef VZ

/* --------------------------------------------------------------------- */
/* one shot: forward (record + boundary save), mute, reverse + imaging    */
/* --------------------------------------------------------------------- */
/* opts: see orc_shot_opts in rtm_oracle.c */
/* The other imaging conditions of one reverse step on the interior cells
 * (kernel.cu:1417-1529), after the cross-correlation, with the reference's
 * quirks: vx / vxs are read at the p index (stride nxp instead of nxp-1,
 * kernel.cu:1431,1468) and the source Poynting vector takes the receiver's vz
 * (kernel.cu:1469).  Operation order as nvcc contracts it by default
 * (a*b + c*d -> fma(a, b, c*d); accumulations -> fma); the float / double
 * mixing of the double literals (pi, eps, 0.5, 16.0) as written.
 * icb: ud du lr rl g4 (nz x nxl each), then 61 bins of g31, 61 of g32. */
static void SFX(extra_ic)(const SFX(wfield) *S, const SFX(wfield) *R, const SFX(medium) *m, const RT *ps_pre,
                          const RT *p_pre, RT *icb, float dx, float dz, float dt)
{
    const int nxp = m->nxp, nzp = m->nzp, npml = m->npml, nxl = m->nxl;
    const size_t nn = (size_t)m->nzi * nxl;
    RT *ud = icb, *du = icb + nn, *lr = icb + 2 * nn, *rl = icb + 3 * nn, *g4 = icb + 4 * nn;
    RT *g31 = icb + 5 * nn, *g32 = icb + 66 * nn;
    const float dw = (float)(ORC_PI / 180);
    for (int iz = npml; iz < nzp - npml; iz++)
        for (int ix = npml; ix < nxp - npml; ix++) {
            const size_t id = (size_t)iz * nxp + ix, ip = (size_t)(iz - npml) * nxl + ix - npml;
            const RT ps = S->p[id], p = R->p[id];
            const int qin = id < (size_t)nzp * (nxp - 1);   /* past the vx array's end: 0 */
            const RT vxs = qin ? S->vx[id] : 0, vzs = S->vz[id], vxr = qin ? R->vx[id] : 0, vzr = R->vz[id];
            /* cuda_wavefield_decomposition (kernel.cu:1417-1450) */
            const RT vtxs = -ps * vxs, vtzs = -ps * vzs, vtxr = -p * vxr, vtzr = -p * vzr;
            const RT psu = vtzs < 0 ? ps : 0, psd = vtzs > 0 ? ps : 0, psl = vtxs < 0 ? ps : 0, psr = vtxs > 0 ? ps : 0;
            const RT pu = vtzr < 0 ? p : 0, pd = vtzr > 0 ? p : 0, pl = vtxr < 0 ? p : 0, pr = vtxr > 0 ? p : 0;
            ud[ip] = FMA(psu, pd, ud[ip]);
            du[ip] = FMA(psd, pu, du[ip]);
            lr[ip] = FMA(psl, pr, lr[ip]);
            rl[ip] = FMA(psr, pl, rl[ip]);
            /* cuda_poynting (kernel.cu:1452-1488) */
            const RT ksx = ps * vxs, ksz = ps * vzr, krx = p * vxr, krz = p * vzr;
            const RT vv = m->vv[id];
#if ORACLE_IS_F64
            const double q = sqrt(FMA(ksx, ksx, ksz * ksz) * FMA(krx, krx, krz * krz));
            double w = FMA(ksx, krx, ksz * krz) / (q > 0 ? q : 1.0e-6);
            if (w >= -1.0 && w <= 1.0) {
                w = ORC_PI / 2.0 - 0.5 * acos(w);
                if (w > 0 && w <= ORC_PI / 3.0) {
                    const size_t ib = (size_t)(int)(w / (ORC_PI / 180)) * nn + ip;
                    g31[ib] = FMA(ps * p, cos(w), g31[ib]);
                    g32[ib] += ps * p * 16.0 * ORC_PI * sqrt(vv) / sin(w);
                }
            }
#else
            const float q = sqrtf(fmaf(ksx, ksx, ksz * ksz) * fmaf(krx, krx, krz * krz));
            float w = (float)((double)fmaf(ksx, krx, ksz * krz) / (q > 0 ? (double)q : 1.0e-6));
            if (w >= -1.0 && w <= 1.0) {
                w = (float)(0.5 * (double)acosf(w));
                w = (float)(ORC_PI / 2.0 - (double)w);
                if (w > 0 && w <= ORC_PI / 3.0) {
                    const size_t ib = (size_t)(int)(w / dw) * nn + ip;
                    g31[ib] = fmaf(ps * p, cosf(w), g31[ib]);
                    g32[ib] = (float)((double)g32[ib] +
                                      (double)(ps * p) * 16.0 * ORC_PI * (double)sqrtf(vv) / (double)sinf(w));
                }
            }
#endif
            (void)dw;
            /* cuda_energynorm (kernel.cu:1489-1529): central differences of the
             * new fields, time differences against the fields before the step */
            const RT psx = (RT)(0.5 * (double)(S->p[id + 1] - S->p[id - 1]) / (double)dx);
            const RT psz = (RT)(0.5 * (double)(S->p[id + nxp] - S->p[id - nxp]) / (double)dz);
            const RT px = (RT)(0.5 * (double)(R->p[id + 1] - R->p[id - 1]) / (double)dx);
            const RT pz = (RT)(0.5 * (double)(R->p[id + nxp] - R->p[id - nxp]) / (double)dz);
            const RT t = -(ps - ps_pre[id]) * (p - p_pre[id]) / ((RT)dt * (RT)dt * vv);
            g4[ip] = g4[ip] + FMA(psz, pz, FMA(psx, px, t));
        }
}

/* one step of the other imaging conditions on given fields (reference shapes:
 * p nzp*nxp, vx nzp*(nxp-1), vz (nzp-1)*nxp, vv nzp*nxp), for the pins */
ORC_EXPORT void SFX(orc_ic_cells)(int nxp, int nzp, int npml, const RT *ps, const RT *vxs, const RT *vzs, const RT *p,
                                  const RT *vxr, const RT *vzr, const RT *ps_pre, const RT *p_pre, const RT *vv,
                                  float dx, float dz, float dt, RT *icb)
{
    SFX(wfield) S, R;
    SFX(medium) m;
    memset(&S, 0, sizeof(S)); memset(&R, 0, sizeof(R)); memset(&m, 0, sizeof(m));
    S.p = (RT *)ps; S.vx = (RT *)vxs; S.vz = (RT *)vzs;
    R.p = (RT *)p; R.vx = (RT *)vxr; R.vz = (RT *)vzr;
    m.nxp = nxp; m.nzp = nzp; m.npml = npml; m.nxl = nxp - 2 * npml; m.nzi = nzp - 2 * npml; m.vv = (RT *)vv;
    SFX(extra_ic)(&S, &R, &m, ps_pre, p_pre, icb, dx, dz, dt);
}

/* index of the snapshot of step k, -1 if k is not a snapshot step */
static int SFX(snap_q)(int k, int every, int k0, int n)
{
    if (every <= 0 || k < 0 || k % every != 0) return -1;
    const int q = k / every - k0;
    return (q >= 0 && q < n) ? q : -1;
}

static int SFX(run_shot)(const orc_params *P, int ishot, const float *vp0, const float *rho0,
                         const float *diffcoef, const orc_shot_opts *o,
                         float *seis_out, const float *seis_in,
                         RT *g1_out, RT *illum_out,
                         float *snap_fwd, float *snap_rec)
{
    orc_geom g;
    if (orc_shot_geometry(P, ishot, &g) != 0) return -1;
    const int nxp = g.nxp, nzp = g.nzp, npml = P->npml, nt = P->nt;
    const int nxl = g.nxl, nzi = P->nz;
    if (o->literal && (npml != 32 || ORACLE_KORDER)) return -2;
    if (npml < 10) return -3;
    if ((o->store_fmt || o->snap_fmt) && sizeof(RT) != sizeof(float)) return -4;   /* 16-bit: FP32 forms */
    SFX(medium) m;
    SFX(med_setup)(&m, &g, P, vp0, rho0, diffcoef, o->vvpad, o->rhopad);

    SFX(wfield) S, Rw;
    SFX(wf_alloc)(&S, nxp, nzp, npml);
    size_t slen = SFX(strip_len)(nzi, nxl);
    /* decimated snapshots (orc_shot_opts.snap_every): snapshot q of step
     * k = (k0 + q) * every holds p of step k and vx, vz of step k + 1 (what the
     * reverse leapfrog holds after reverse step k), interior cells only */
    int sn_every = o->snap_every, sn_k0 = 0, sn_n = 0;
    if (sn_every > 0) {
        const int t0 = (int)ceil(4.0 / ((double)P->fdom * (double)P->dt));
        sn_k0 = (t0 + sn_every - 1) / sn_every;
        if (sn_k0 < 1) sn_k0 = 1;
        sn_n = (nt - 2) / sn_every - sn_k0 + 1;
        if (sn_n < 0) sn_n = 0;
    }
    const size_t sn_cells = (size_t)nzi * nxl;
    RT *snp = sn_n > 0 ? (RT *)malloc((size_t)sn_n * 3 * sn_cells * sizeof(RT)) : NULL;
    RT *store = (RT *)malloc(slen * (size_t)nt * sizeof(RT));
    RT *rec = (RT *)calloc((size_t)nt * g.nrec, sizeof(RT));
    const size_t rrow = (size_t)npml * nxp;        /* surface row of the padded grid */
    int nt_fwd = o->nt_limit > 0 && o->nt_limit < nt ? o->nt_limit : nt;

    /* forward: source propagation with recording and boundary saving
     * (rtm_model.cu:269-285 and :322-346; fused as rtm_model_pure.cu:278-304) */
    for (int it = 0; it < nt_fwd; it++) {
        SFX(fwd_step)(&S, &m, o->literal);
        SFX(add_source)(S.p, &m, it, +1);
        for (int ir = 0; ir < g.nrec; ir++) rec[(size_t)it * g.nrec + ir] = S.p[rrow + g.rcol0 + ir];
        SFX(strips_io)(&S, store + (size_t)it * slen, 1, 3);
        if (o->store_fmt) {                      /* the library's 16-bit store, block by block */
            RT *st = store + (size_t)it * slen;
            const long long nb[4] = {20LL * nzi, 20LL * nxl, 20LL * nzi, 20LL * nxl};
            for (int b = 0; b < 4; b++) {
                orc_store16_block((float *)st, nb[b], o->store_fmt, NULL);
                st += nb[b];
            }
        }
        if (snap_fwd && it == o->snap_it)
            for (size_t i = 0; i < (size_t)nxp * nzp; i++) snap_fwd[i] = (float)S.p[i];
        if (snp) {
            const int qp = SFX(snap_q)(it, sn_every, sn_k0, sn_n), qv = SFX(snap_q)(it - 1, sn_every, sn_k0, sn_n);
            for (int iz = 0; iz < nzi; iz++)
                for (int ix = 0; ix < nxl; ix++) {
                    const size_t e = (size_t)iz * nxl + ix;
                    if (qp >= 0) snp[(size_t)(3 * qp) * sn_cells + e] = S.p[(size_t)(iz + npml) * nxp + ix + npml];
                    if (qv >= 0) {
                        snp[(size_t)(3 * qv + 1) * sn_cells + e] = S.vx[(size_t)(iz + npml) * (nxp - 1) + ix + npml];
                        snp[(size_t)(3 * qv + 2) * sn_cells + e] = S.vz[(size_t)(iz + npml) * nxp + ix + npml];
                    }
                }
            if (o->snap_fmt) {                   /* the library's 16-bit snapshots: one scale per field */
                if (qp >= 0) orc_store16_block((float *)(snp + (size_t)(3 * qp) * sn_cells), (long long)sn_cells, o->snap_fmt, NULL);
                if (qv >= 0)
                    for (int f = 1; f < 3; f++)
                        orc_store16_block((float *)(snp + (size_t)(3 * qv + f) * sn_cells), (long long)sn_cells, o->snap_fmt, NULL);
            }
        }
    }
    if (seis_out)
        for (size_t i = 0; i < (size_t)nt * g.nrec; i++) seis_out[i] = (float)rec[i];
    if (o->forward_only) {
        SFX(wf_free)(&S); free(store); free(rec); free(snp); SFX(med_free)(&m);
        return 0;
    }
    if (seis_in)
        for (size_t i = 0; i < (size_t)nt * g.nrec; i++) rec[i] = (RT)seis_in[i];
    /* direct-wave mute (kernel.cu:96-173): rtm_model mutes with the model of
     * the forward pass, rtm_real after the smoothing below (rtm_real.cu:370-403) */
    int *cut = (int *)malloc(sizeof(int) * (g.nrec > 0 ? g.nrec : 1));
    if (!P->field) {
        orc_mute_cuts(P, &g, NULL, vp0, m.vvf, cut);
        for (int ir = 0; ir < g.nrec; ir++)
            for (int it = 0; it < nt && it <= cut[ir]; it++) rec[(size_t)it * g.nrec + ir] = 0;
    }

    /* velsmooth1d of the window before the reverse pass (rtm_model.cu:402-405):
     * reconstruction and receiver propagation use the smoothed model */
    if (P->ismth > 0) {
        orc_gauss_smooth(m.vvf, nzp, nxp, P->ismth);
        orc_gauss_smooth(m.rhf, nzp, nxp, P->ismth);
        for (size_t i = 0; i < (size_t)nxp * nzp; i++) {
            m.vv[i] = (RT)m.vvf[i]; m.rho[i] = (RT)m.rhf[i];
            m.ib[i] = (RT)1 / m.rho[i]; m.kk[i] = m.rho[i] * m.vv[i];
        }
    }
    if (P->field) {
        orc_mute_cuts(P, &g, NULL, vp0, m.vvf, cut);
        for (int ir = 0; ir < g.nrec; ir++)
            for (int it = 0; it < nt && it <= cut[ir]; it++) rec[(size_t)it * g.nrec + ir] = 0;
    }
    free(cut);
    /* reverse: ps <- P_{nt-1}; vxs = vzs = 0 (rtm_model.cu:366-372) */
    memset(S.vx, 0, (size_t)nzp * (nxp - 1) * sizeof(RT));
    memset(S.vz, 0, (size_t)(nzp - 1) * nxp * sizeof(RT));
    SFX(wf_alloc)(&Rw, nxp, nzp, npml);
    const size_t nn = (size_t)nzi * nxl;
    RT *icb = NULL, *ps_pre = NULL, *p_pre = NULL;
    if (o->ic_out) {
        icb = (RT *)calloc(nn * (5 + 2 * 61), sizeof(RT));
        ps_pre = (RT *)malloc((size_t)nxp * nzp * sizeof(RT));
        p_pre = (RT *)malloc((size_t)nxp * nzp * sizeof(RT));
    }
    RT *g1 = (RT *)calloc((size_t)nzi * nxl, sizeof(RT));
    RT *il = (RT *)calloc((size_t)nzi * nxl, sizeof(RT));
    /* seed receiver wavefield with the last sample (rtm_model.cu:410) */
    for (int ir = 0; ir < g.nrec; ir++) Rw.p[rrow + g.rcol0 + ir] = m.dt * rec[(size_t)(nt - 1) * g.nrec + ir];
    for (int it = nt - 2; it >= 0; it--) {
        if (icb) {                                     /* d_ps_pre, d_p_pre (rtm_model.cu:419-420) */
            memcpy(ps_pre, S.p, (size_t)nxp * nzp * sizeof(RT));
            memcpy(p_pre, Rw.p, (size_t)nxp * nzp * sizeof(RT));
        }
        /* source wavefield: vel strips of step it+1, interior vel update,
         * p strips of step it, interior p update, source removal */
        SFX(strips_io)(&S, store + (size_t)(it + 1) * slen, 0, 1);
        SFX(rev_vel)(&S, &m);
        SFX(strips_io)(&S, store + (size_t)it * slen, 0, 2);
        SFX(rev_pres)(&S, &m);
        if (o->exact_reverse) {
            RT wv = m.wav[it + 1];
            for (int iz = npml; iz < nzp - npml; iz++)
                for (int ix = npml; ix < nxp - npml; ix++) {
                    size_t id = (size_t)iz * nxp + ix;
                    S.p[id] = FMA(-(m.dt * m.src[id]), wv, S.p[id]);
                }
        } else {
            SFX(add_source)(S.p, &m, it, -1);
        }
        /* receiver wavefield: forward step, Dirichlet insertion */
        SFX(fwd_step)(&Rw, &m, o->literal);
        for (int ir = 0; ir < g.nrec; ir++) Rw.p[rrow + g.rcol0 + ir] = m.dt * rec[(size_t)it * g.nrec + ir];
        /* cross-correlation + illumination (kernel.cu:1401-1416) */
        for (int iz = 0; iz < nzi; iz++)
            for (int ix = 0; ix < nxl; ix++) {
                size_t idp = (size_t)(iz + npml) * nxp + ix + npml, idi = (size_t)iz * nxl + ix;
                g1[idi] = FMA(S.p[idp], Rw.p[idp], g1[idi]);
                il[idi] = FMA(S.p[idp], S.p[idp], il[idi]);
            }
        if (icb) SFX(extra_ic)(&S, &Rw, &m, ps_pre, p_pre, icb, P->dx, P->dz, P->dt);
        if (snap_rec && it == o->snap_it)
            for (size_t i = 0; i < (size_t)nxp * nzp; i++) snap_rec[i] = (float)S.p[i];
        /* snapshot restore after reverse step it: p on the interior, vx on
         * its columns npml..nxp-npml-2, vz on its rows npml..nzp-npml-2 */
        const int qr = snp ? SFX(snap_q)(it, sn_every, sn_k0, sn_n) : -1;
        if (qr >= 0)
            for (int iz = 0; iz < nzi; iz++)
                for (int ix = 0; ix < nxl; ix++) {
                    const size_t e = (size_t)iz * nxl + ix;
                    S.p[(size_t)(iz + npml) * nxp + ix + npml] = snp[(size_t)(3 * qr) * sn_cells + e];
                    if (ix < nxl - 1)
                        S.vx[(size_t)(iz + npml) * (nxp - 1) + ix + npml] = snp[(size_t)(3 * qr + 1) * sn_cells + e];
                    if (iz < nzi - 1)
                        S.vz[(size_t)(iz + npml) * nxp + ix + npml] = snp[(size_t)(3 * qr + 2) * sn_cells + e];
                }
    }
    free(snp);
    if (icb) {
        /* cuda_stack_udlr, cuda_stack_theta (kernel.cu:1530-1555) */
        for (size_t i = 0; i < nn; i++) {
            RT g3 = 0, g3t = 0;
            for (int iw = 0; iw <= 60; iw++) { g3 += icb[(5 + iw) * nn + i]; g3t += icb[(66 + iw) * nn + i]; }
            o->ic_out[i] = (float)(((icb[i] + icb[nn + i]) + icb[2 * nn + i]) + icb[3 * nn + i]);
            o->ic_out[nn + i] = (float)g3;
            o->ic_out[2 * nn + i] = (float)icb[4 * nn + i];
            o->ic_out[3 * nn + i] = (float)g3t;
        }
        free(icb); free(ps_pre); free(p_pre);
    }
    memcpy(g1_out, g1, (size_t)nzi * nxl * sizeof(RT));
    memcpy(illum_out, il, (size_t)nzi * nxl * sizeof(RT));
    free(g1); free(il);
    SFX(wf_free)(&S); SFX(wf_free)(&Rw); free(store); free(rec); SFX(med_free)(&m);
    return 0;
}

/* --------------------------------------------------------------------- */
/* exported entry points                                                   */
/* --------------------------------------------------------------------- */
ORC_EXPORT int SFX(orc_shot)(const orc_params *P, int ishot, const float *vp0, const float *rho0,
                             const float *diffcoef, const orc_shot_opts *o,
                             float *seis_out, const float *seis_in,
                             RT *g1_out, RT *illum_out, float *snap_fwd, float *snap_rec)
{
    return SFX(run_shot)(P, ishot, vp0, rho0, diffcoef, o, seis_out, seis_in, g1_out, illum_out, snap_fwd, snap_rec);
}

/* run a few forward steps from given fields (unit-level parity of the FD
 * update).  Fields use the reference shapes (p nzp*nxp, vx nzp*(nxp-1),
 * vz (nzp-1)*nxp, split pairs as in wfield).  The medium is set up from the
 * shot geometry of `ishot`. */
ORC_EXPORT int SFX(orc_fd_steps)(const orc_params *P, int ishot, const float *vp0, const float *rho0,
                                 const float *diffcoef, int literal, int nsteps, int it0, int with_source,
                                 RT *p, RT *vx, RT *vz, RT *pl1, RT *pl2, RT *pr1, RT *pr2,
                                 RT *pt1, RT *pt2, RT *pb1, RT *pb2)
{
    orc_geom g;
    if (orc_shot_geometry(P, ishot, &g) != 0) return -1;
    if (literal && (P->npml != 32 || ORACLE_KORDER)) return -2;
    SFX(medium) m;
    SFX(med_setup)(&m, &g, P, vp0, rho0, diffcoef, NULL, NULL);
    SFX(wfield) w;
    w.nxp = g.nxp; w.nzp = g.nzp; w.npml = P->npml; w.nxl = g.nxl; w.nzi = P->nz;
    w.p = p; w.vx = vx; w.vz = vz; w.pl1 = pl1; w.pl2 = pl2; w.pr1 = pr1; w.pr2 = pr2;
    w.pt1 = pt1; w.pt2 = pt2; w.pb1 = pb1; w.pb2 = pb2;
    for (int k = 0; k < nsteps; k++) {
        SFX(fwd_step)(&w, &m, literal);
        if (with_source) SFX(add_source)(w.p, &m, it0 + k, +1);
    }
    SFX(med_free)(&m);
    return 0;
}

/* whole rtm_model() job: shots startshot..nshot step dshot, taper, stacks.
 * stack_mode 0: FP32 stacks exactly as rtm_model.cu:477-488;
 * stack_mode 1: FP64 stacks (image sums kept in double, rounded at the end). */
ORC_EXPORT int SFX(orc_rtm)(const orc_params *P, const float *vp0, const float *rho0,
                            const float *diffcoef, int literal, int stack_mode, int snap_every, int store_fmt,
                            int snap_fmt, float *image1, float *image2s, float *illum, int *nshots_done)
{
    const int nx = P->nx, nz = P->nz;
    size_t nimg = (size_t)nx * nz;
    double *a1 = (double *)calloc(nimg, sizeof(double));
    double *a2 = (double *)calloc(nimg, sizeof(double));
    double *a3 = (double *)calloc(nimg, sizeof(double));
    float *f1 = (float *)calloc(nimg, sizeof(float));
    float *f2 = (float *)calloc(nimg, sizeof(float));
    float *f3 = (float *)calloc(nimg, sizeof(float));
    int nwin = 50, done = 0, rc = 0;
    orc_shot_opts o;
    memset(&o, 0, sizeof(o));
    o.literal = literal; o.snap_it = -1; o.snap_every = snap_every; o.store_fmt = store_fmt; o.snap_fmt = snap_fmt;
    for (int ishot = P->startshot; ishot <= P->nshot; ishot += P->dshot) {
        orc_geom g;
        if (orc_shot_geometry(P, ishot, &g) != 0) { rc = -1; break; }
        int nxl = g.nxl;
        RT *g1 = (RT *)malloc((size_t)nz * nxl * sizeof(RT));
        RT *il = (RT *)malloc((size_t)nz * nxl * sizeof(RT));
        rc = SFX(run_shot)(P, ishot, vp0, rho0, diffcoef, &o, NULL, NULL, g1, il, NULL, NULL);
        if (rc != 0) { free(g1); free(il); break; }
        /* taper (rtm_model.cu:465-473, kernel.cu:1591-1615), nwin persists */
        while (2 * nwin >= nxl) nwin -= 3;
        for (int iz = 0; iz < nz; iz++)
            for (int ix = 0; ix < nxl; ix++) {
                RT t = (RT)orc_taper(ix, nxl, nwin);
                g1[(size_t)iz * nxl + ix] *= t;
                il[(size_t)iz * nxl + ix] *= t;
            }
        /* cuda_applyics / cuda_applyic (kernel.cu:1556-1579) */
        for (int iz = 0; iz < nz; iz++)
            for (int ix = 0; ix < nxl; ix++) {
                size_t ip = (size_t)iz * nxl + ix, ifu = (size_t)iz * nx + ix + g.nx1;
                if (stack_mode == 0) {
                    f2[ifu] = (float)((double)f2[ifu] + (double)g1[ip] / ((double)il[ip] + 1.0e-6));
                    f1[ifu] = (float)(f1[ifu] + g1[ip]);
                    f3[ifu] = (float)(f3[ifu] + il[ip]);
                } else {
                    a2[ifu] += (double)g1[ip] / ((double)il[ip] + 1.0e-6);
                    a1[ifu] += (double)g1[ip];
                    a3[ifu] += (double)il[ip];
                }
            }
        free(g1); free(il);
        done++;
    }
    for (size_t i = 0; i < nimg; i++) {
        if (stack_mode == 0) { image1[i] = f1[i]; image2s[i] = f2[i]; illum[i] = f3[i]; }
        else { image1[i] = (float)a1[i]; image2s[i] = (float)a2[i]; illum[i] = (float)a3[i]; }
    }
    if (nshots_done) *nshots_done = done;
    free(a1); free(a2); free(a3); free(f1); free(f2); free(f3);
    return rc;
}
