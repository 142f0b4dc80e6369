// field_data.cpp -- the field-data path of the drop-in: the shot-gather scan,
// gather reading, seismogram resampling onto the modelling grid and
// rtm_real(), which migrates recorded gathers through the same engine as
// rtm_model() (recorded data in place of modelled data, one-sided receiver
// spread per gather).
//
// Everything here is host code run once per shot; the propagation and
// imaging stay in the engine's kernels.
#include "../../include/rtm_hip.h"
#include "host_io.h"
#include "rtm_engine.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <complex>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

typedef std::complex<double> cd;

void set_err(const std::string& s) { rtm::set_error(s); }

// ------------------------------------------------------------------ FFT ----
// Unnormalised complex DFT in double precision, X[k] = sum_j x[j] e^{sign 2 pi i jk/n}:
// sign -1 is FFTW_FORWARD, +1 FFTW_BACKWARD (no 1/n either way), the
// conventions of the FFTW calls at Interp_function.cpp:150-225.  Powers of
// two by an iterative radix-2 transform, lengths with prime factors up to 7
// by mixed-radix Cooley-Tukey, others by Bluestein's chirp-z identity
// jk = (j^2 + k^2 - (k-j)^2)/2 over a power-of-two convolution.
void fft_pow2(cd* a, int n, int sign)
{
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    const double ang = sign * 2.0 * M_PI / len;
    for (int k = 0; k < len / 2; k++) {
      const cd w(cos(ang * k), sin(ang * k));
      for (int i = k; i < n; i += len) {
        const cd u = a[i], v = a[i + len / 2] * w;
        a[i] = u + v;
        a[i + len / 2] = u - v;
      }
    }
  }
}

struct Chirp {
  long long key;
  int m;
  std::vector<cd> w, Bf;     // e^{sign i pi k^2/n}; FFT of the conjugate chirp, length m
};

void fft_bluestein(cd* a, int n, int sign)
{
  // chirp and its transform per (length, direction), kept per thread
  thread_local std::vector<Chirp> cache;
  const long long key = 2LL * n + (sign > 0 ? 1 : 0);
  const Chirp* c = nullptr;
  for (auto& e : cache)
    if (e.key == key) { c = &e; break; }
  if (!c) {
    Chirp e;
    e.key = key;
    e.m = 1;
    while (e.m < 2 * n - 1) e.m <<= 1;
    e.w.resize(n);
    const long long n2 = 2LL * n;
    for (int k = 0; k < n; k++) {
      const long long q = ((long long)k * k) % n2;      // e^{i pi k^2/n} has period 2n in k^2
      const double ang = sign * M_PI * (double)q / n;
      e.w[k] = cd(cos(ang), sin(ang));
    }
    e.Bf.assign(e.m, cd(0, 0));
    e.Bf[0] = std::conj(e.w[0]);
    for (int k = 1; k < n; k++) e.Bf[k] = e.Bf[e.m - k] = std::conj(e.w[k]);
    fft_pow2(e.Bf.data(), e.m, -1);
    if (cache.size() >= 8) cache.erase(cache.begin());
    cache.push_back(std::move(e));
    c = &cache.back();
  }
  const int m = c->m;
  std::vector<cd> A(m, cd(0, 0));
  for (int k = 0; k < n; k++) A[k] = a[k] * c->w[k];
  fft_pow2(A.data(), m, -1);
  for (int i = 0; i < m; i++) A[i] *= c->Bf[i];
  fft_pow2(A.data(), m, 1);
  for (int k = 0; k < n; k++) a[k] = A[k] / (double)m * c->w[k];
}

// Mixed-radix Cooley-Tukey (decimation in time, radices 2, 3, 5, 7) for
// lengths whose prime factors are all small, e.g. trace lengths 2000 or 4000:
// the p sub-transforms of the decimated sequences x[r + p j], then p-point
// butterflies with twiddles W_n^{rk}.  tw holds e^{sign 2 pi i j/N}, j < N.
void fft_mixed(const cd* x, int stride, cd* y, int n, int N, const cd* tw)
{
  if (n == 1) { y[0] = x[0]; return; }
  int p = 2;
  while (n % p) p++;
  const int m = n / p;
  for (int r = 0; r < p; r++) fft_mixed(x + (size_t)r * stride, stride * p, y + (size_t)r * m, m, N, tw);
  const int sn = N / n, sp = N / p;
  cd t[7], o[7];
  for (int k = 0; k < m; k++) {
    for (int r = 0; r < p; r++) t[r] = y[(size_t)r * m + k] * tw[(size_t)((long long)r * k * sn % N)];
    for (int q = 0; q < p; q++) {
      cd acc = t[0];
      for (int r = 1; r < p; r++) acc += t[r] * tw[(size_t)((r * q) % p) * sp];
      o[q] = acc;
    }
    for (int q = 0; q < p; q++) y[(size_t)k + (size_t)q * m] = o[q];
  }
}

bool small_factors(int n)
{
  for (int p : {2, 3, 5, 7})
    while (n % p == 0) n /= p;
  return n == 1;
}

void fft(cd* a, int n, int sign)
{
  if (n <= 1) return;
  if ((n & (n - 1)) == 0) { fft_pow2(a, n, sign); return; }
  if (!small_factors(n)) { fft_bluestein(a, n, sign); return; }
  // twiddle table per (length, direction), kept per thread
  thread_local std::vector<std::pair<long long, std::vector<cd>>> cache;
  const long long key = 2LL * n + (sign > 0 ? 1 : 0);
  const std::vector<cd>* tw = nullptr;
  for (auto& c : cache)
    if (c.first == key) { tw = &c.second; break; }
  if (!tw) {
    std::vector<cd> t(n);
    for (int j = 0; j < n; j++) {
      const double ang = sign * 2.0 * M_PI * (double)j / n;
      t[j] = cd(cos(ang), sin(ang));
    }
    if (cache.size() >= 8) cache.erase(cache.begin());
    cache.emplace_back(key, std::move(t));
    tw = &cache.back().second;
  }
  std::vector<cd> x(a, a + n);
  fft_mixed(x.data(), 1, a, n, n, tw->data());
}

// 2-D transform of a row-major n0 x n1 array (fftw_plan_dft_2d(n0, n1, ...))
void fft2(cd* a, int n0, int n1, int sign)
{
  for (int i = 0; i < n0; i++) {
    const cd* r = a + (size_t)i * n1;
    bool zero = true;
    for (int j = 0; j < n1 && zero; j++) zero = (r[j] == cd(0, 0));
    if (!zero) fft(a + (size_t)i * n1, n1, sign);        // the transform of zeros is zeros
  }
  std::vector<cd> col(n0);
  for (int j = 0; j < n1; j++) {
    for (int i = 0; i < n0; i++) col[i] = a[(size_t)i * n1 + j];
    fft(col.data(), n0, sign);
    for (int i = 0; i < n0; i++) a[(size_t)i * n1 + j] = col[i];
  }
}

// ------------------------------------------------------- interpolation ----
// modZeroBessel (Interp_function.cpp:10-23): I0 by its series, 19 terms
double bessel_i0(double x)
{
  const double h = x / 2;
  double num = 1, fact = 1, r = 1;
  for (int i = 1; i < 20; i++) {
    num *= h * h;
    fact *= i;
    r += num / (fact * fact);
  }
  return r;
}

// Sincinterp (Interp_function.cpp:61-112): 21-point Kaiser-windowed sinc
// (beta 2) in time, trace by trace.  The window of output sample it is
// centred on input sample (int)((it+1)*dt/dt0) while the sinc is evaluated
// at t = it*dt; windows cut by the record start are right-aligned in the
// Kaiser table, windows cut by its end left-aligned.  float arithmetic as
// the reference (the sinc itself in double through the double pi).
void sinc_interp(const float* in, float* out, int nx, int ldin, int ldout, int nt, int nt0, float dt, float dt0)
{
  const int nw = 21, half = 10;
  float wk[21];
  const double m2 = (nw - 1) / 2.0, den = bessel_i0(2.0);
  for (int n = 0; n < nw; n++) {
    double v = (n - m2) / m2;
    v = 1 - v * v;
    wk[n] = (float)(bessel_i0(2.0 * sqrt(v)) / den);
  }
  std::vector<float> tn(nt0);
  for (int it = 0; it < nt0; it++) tn[it] = it * dt0;
  // the window, its taps and sincs depend on it only; each trace then sums
  // in*sinc*w in k order
  float sc[21];
  for (int it = 0; it < nt; it++) {
    const int it0 = (int)((it + 1) * dt / dt0);
    const int s0 = it0 - half > 0 ? it0 - half : 0;
    const int s1 = it0 + half < nt0 - 1 ? it0 + half : nt0 - 1;
    const int len = s1 - s0 + 1;
    const int w0 = (s0 <= half) ? nw - len : 0;
    const float t = it * dt;
    for (int k = 0; k < len; k++) {
      const float x = (t - tn[k + s0]) / dt0;
      sc[k] = (x == 0.0f) ? 1.0f : (float)(sin(3.1415926535897932 * x) / (3.1415926535897932 * x));
    }
    for (int ix = 0; ix < nx; ix++) {
      float sum = 0.0f;
      for (int k = 0; k < len; k++) sum = sum + in[(size_t)(k + s0) * ldin + ix] * sc[k] * wk[k + w0];
      out[(size_t)it * ldout + ix] = sum;
    }
  }
}

// GeneralFKinterp (Interp_function.cpp:121-243): trace interpolation by an
// integer factor lx in the f-k domain.  Steps as the reference takes them:
//  1) K = FFT2(record, nx0 x nt0); C = K repeated lx times along k
//  2) S = FFT2(record zero-padded to lx*nx0 x lx*nt0); Z = the average of S's
//     lx blocks along k, repeated, plus 1e-5 of its largest modulus on the
//     real part; H = S/Z on the first nf = nt0/2+1 frequencies, |H| > lx^2
//     replaced by lx
//  3) G1 = inverse FFT over k of H*C per frequency; the frequencies above nf
//     are the conjugates of G1 rows (nt0+1)/2, (nt0+1)/2 - 1, ... (for even
//     nt0 one row above the mirror image of each); inverse FFT over time,
//     real part, no 1/n scaling.
// For odd nt0 the reference's first mirrored row is one past its buffer, and
// for nx > lx*nx0 its output columns are; such inputs are refused.
int fk_interp(const float* in, int ldin, float* out, int ldout, int nx, int nx0, int nt0, int lx, std::string& err)
{
  if (nt0 % 2 != 0) { err = "x-interpolation of an odd trace length (nt0) is not supported"; return -1; }
  if (nx > lx * nx0) { err = "x-interpolation: more output traces than lx*nx0"; return -1; }
  const int nf = 1 + nt0 / 2, Nx = lx * nx0, Nt = lx * nt0;
  std::vector<cd> B((size_t)nx0 * nt0), K, C((size_t)Nx * nt0), S((size_t)Nx * Nt, cd(0, 0)),
      Z((size_t)Nx * Nt, cd(0, 0)), H((size_t)Nx * nf), G1((size_t)nf * Nx), G((size_t)Nx * nt0);
  for (int ix = 0; ix < nx0; ix++)
    for (int it = 0; it < nt0; it++) B[(size_t)ix * nt0 + it] = cd(in[(size_t)it * ldin + ix], 0.0);
  K = B;
  fft2(K.data(), nx0, nt0, -1);
  for (int il = 0; il < lx; il++)
    for (int ix = 0; ix < nx0; ix++)
      for (int it = 0; it < nt0; it++) C[(size_t)(ix + il * nx0) * nt0 + it] = K[(size_t)ix * nt0 + it];
  for (int ix = 0; ix < nx0; ix++)
    for (int it = 0; it < nt0; it++) S[(size_t)ix * Nt + it] = B[(size_t)ix * nt0 + it];
  fft2(S.data(), Nx, Nt, -1);
  for (int il = 0; il < lx; il++)
    for (int ix = 0; ix < nx0; ix++)
      for (int it = 0; it < Nt; it++) Z[(size_t)ix * Nt + it] += S[(size_t)(ix + il * nx0) * Nt + it];
  for (int ix = 0; ix < nx0; ix++)
    for (int it = 0; it < Nt; it++) Z[(size_t)ix * Nt + it] /= (double)lx;
  for (int il = 1; il < lx; il++)
    for (int ix = 0; ix < nx0; ix++)
      for (int it = 0; it < Nt; it++) Z[(size_t)(ix + il * nx0) * Nt + it] = Z[(size_t)ix * Nt + it];
  // fftw_maxfabs: a float maximum of float moduli (common.h:26)
  float zmax = 0.0f;
  for (const cd& z : Z) {
    const float m = sqrtf((float)(z.real() * z.real() + z.imag() * z.imag()));
    zmax = zmax > m ? zmax : m;
  }
  for (cd& z : Z) z += cd(zmax * 0.00001, 0.0);
  for (int ix = 0; ix < Nx; ix++)
    for (int it = 0; it < nf; it++) {
      const cd s = S[(size_t)ix * Nt + it], z = Z[(size_t)ix * Nt + it];
      const double d = z.real() * z.real() + z.imag() * z.imag();
      cd h((s.real() * z.real() + s.imag() * z.imag()) / d, (s.imag() * z.real() - s.real() * z.imag()) / d);
      if (sqrtf((float)(h.real() * h.real() + h.imag() * h.imag())) > lx * lx) h = cd((float)lx, 0.0);
      H[(size_t)ix * nf + it] = h;
    }
  for (int it = 0; it < nf; it++) {
    cd* row = G1.data() + (size_t)it * Nx;
    for (int ix = 0; ix < Nx; ix++) {
      const cd h = H[(size_t)ix * nf + it], c = C[(size_t)ix * nt0 + it];
      row[ix] = cd(h.real() * c.real() - h.imag() * c.imag(), h.real() * c.imag() + h.imag() * c.real());
    }
    fft(row, Nx, 1);
  }
  for (int ix = 0; ix < Nx; ix++) {
    cd* g = G.data() + (size_t)ix * nt0;
    for (int it = 0; it < nf; it++) g[it] = G1[(size_t)it * Nx + ix];
    for (int it = nf, it1 = (nt0 + 1) / 2; it < nt0; it++, it1--) g[it] = std::conj(G1[(size_t)it1 * Nx + ix]);
    fft(g, nt0, 1);
  }
  for (int ix = 0; ix < nx; ix++)
    for (int it = 0; it < nt0; it++) out[(size_t)it * ldout + ix] = (float)G[(size_t)ix * nt0 + it].real();
  return 0;
}

inline uint32_t bswap32(uint32_t v) { return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24); }
inline uint16_t bswap16(uint16_t v) { return (uint16_t)((v >> 8) | (v << 8)); }

// the key fields scanshotgather reads from a 240-byte trace header
// (fread_sgyhead, sgyhead_function.cpp:299-320): offset @36, sx @72, ns @114, dt @116
struct GKey {
  int32_t offset, sx;
  int16_t ns, dt;
};
bool read_gkey(FILE* fp, GKey& k, int swap)
{
  unsigned char h[240];
  if (fread(h, 1, 240, fp) != 240) return false;
  uint32_t o, s;
  uint16_t n, d;
  memcpy(&o, h + 36, 4); memcpy(&s, h + 72, 4); memcpy(&n, h + 114, 2); memcpy(&d, h + 116, 2);
  if (swap) { o = bswap32(o); s = bswap32(s); n = bswap16(n); d = bswap16(d); }
  k.offset = (int32_t)o; k.sx = (int32_t)s; k.ns = (int16_t)n; k.dt = (int16_t)d;
  return true;
}

}  // namespace

// --------------------------------------------------------------- C ABI ----
extern "C" int rtmh_scan_gather(const char* path, int sgyflag, int* endianflag, int* prof, int max_shots, int* nshots)
{
  if (!path || !endianflag || !prof || max_shots < 1) { set_err("null argument"); return -1; }
  FILE* fp = fopen(path, "rb");
  if (!fp) { set_err(std::string("cannot open ") + path); return -1; }
  fseek(fp, 0, SEEK_END);
  const long length = ftell(fp);
  const long start = sgyflag ? 3600L : 0L;
  fseek(fp, start, SEEK_SET);
  GKey k;
  if (!read_gkey(fp, k, 0)) { fclose(fp); set_err("no trace header"); return -2; }
  // endianness probe of RTM_GPU_velstr.cpp:383-396
  int swap = 0;
  if (k.dt <= 0 || k.ns <= 0) {
    GKey k2;
    fseek(fp, start, SEEK_SET);
    read_gkey(fp, k2, 1);
    if (k2.dt > 0 && k2.ns > 0) { swap = 1; k = k2; }
  }
  *endianflag = swap;
  if (k.ns <= 0) { fclose(fp); set_err("trace length (ns) is not positive"); return -2; }
  const long tr = 240L + 4L * k.ns;
  // shot table (RTM_GPU_velstr.cpp:409-453): a new shot wherever sx changes;
  // rows sx, number of traces, first trace (1-based), offset of the first trace
  int* sxr = prof;
  int* ntr = prof + max_shots;
  int* itr = prof + 2 * (size_t)max_shots;
  int* off = prof + 3 * (size_t)max_shots;
  int ns = 0;
  sxr[0] = k.sx; itr[0] = 1; off[0] = k.offset; ntr[0] = 0;
  const long ntraces = (length - start) / tr;
  int first = 1;
  for (long i = 0; i < ntraces; i++) {
    fseek(fp, start + i * tr, SEEK_SET);
    GKey h;
    if (!read_gkey(fp, h, swap)) break;
    if (h.ns <= 0 || h.dt <= 0) break;
    if (h.sx != sxr[ns]) {
      ntr[ns] = (int)(i + 1) - first;
      if (++ns >= max_shots) { fclose(fp); set_err("more shots in the gather file than the table holds"); return -3; }
      sxr[ns] = h.sx; itr[ns] = (int)(i + 1); off[ns] = h.offset;
      first = (int)(i + 1);
    }
    ntr[ns] = (int)(i + 2) - first;
  }
  fclose(fp);
  if (nshots) *nshots = ns + 1;
  return 0;
}

static int read_gather_impl(const char* path, int sgyflag, int endianflag, int itraceall, int ntrace, int nt0, float* seis,
                            int ncols, std::string& err)
{
  if (!path || !seis || itraceall < 1 || ntrace < 0 || nt0 < 1 || ncols < 0) { err = "bad argument"; return -1; }
  FILE* fp = fopen(path, "rb");
  if (!fp) { err = std::string("cannot open ") + path; return -1; }
  // rtm_real.cu:379-391: trace itraceall onwards, samples byte-swapped when
  // endianflag (float_to_float); traces past the record's width are dropped
  const long long off = (long long)(itraceall - 1) * (240 + (long long)nt0 * 4) + 3600LL * (sgyflag ? 1 : 0);
  if (fseek(fp, (long)off, SEEK_SET) != 0) { fclose(fp); err = "seek past the end of the gather file"; return -2; }
  memset(seis, 0, sizeof(float) * (size_t)nt0 * ncols);
  std::vector<float> t(nt0);
  const int n = ntrace < ncols ? ntrace : ncols;
  for (int ix = 0; ix < n; ix++) {
    fseek(fp, 240L, SEEK_CUR);
    if (fread(t.data(), 4, nt0, fp) != (size_t)nt0) { fclose(fp); err = "short gather file"; return -2; }
    if (endianflag)
      for (float& v : t) { uint32_t u; memcpy(&u, &v, 4); u = bswap32(u); memcpy(&v, &u, 4); }
    for (int it = 0; it < nt0; it++) seis[(size_t)it * ncols + ix] = t[it];
  }
  fclose(fp);
  return 0;
}

extern "C" int rtmh_read_gather(const char* path, int sgyflag, int endianflag, int itraceall, int ntrace, int nt0,
                                float* seis, int ncols)
{
  std::string err;
  const int rc = read_gather_impl(path, sgyflag, endianflag, itraceall, ntrace, nt0, seis, ncols, err);
  if (rc != 0) set_err(err);
  return rc;
}

static int interp_seis_impl(float* seiscal, const float* seisobs, int nx, int nx0, int nt, int nt0, float dt, float dt0,
                            std::string& err)
{
  if (!seiscal || !seisobs || nx < 1 || nx0 < 1 || nt < 1 || nt0 < 1) { err = "bad argument"; return -1; }
  // Interpseis2d (Interp_function.cpp:245-264)
  if (nx > nx0) {
    if (nx0 < 2) { err = "x-interpolation needs nx0 >= 2"; return -1; }
    const int lx = (nx - 1) / (nx0 - 1);
    if (nt > nt0) {
      std::vector<float> tmp((size_t)nt0 * nx);
      if (fk_interp(seisobs, nx0, tmp.data(), nx, nx, nx0, nt0, lx, err) != 0) return -1;
      sinc_interp(tmp.data(), seiscal, nx, nx, nx, nt, nt0, dt, dt0);
    } else {
      // the reference writes all nt0 rows into the nt-row record; the first nt are kept
      std::vector<float> tmp((size_t)nt0 * nx);
      if (fk_interp(seisobs, nx0, tmp.data(), nx, nx, nx0, nt0, lx, err) != 0) return -1;
      memcpy(seiscal, tmp.data(), sizeof(float) * (size_t)nt * nx);
    }
  } else if (nt > nt0) {
    sinc_interp(seisobs, seiscal, nx, nx0, nx, nt, nt0, dt, dt0);
  } else {
    // the reference's `seiscal = seisobs` rebinds its local pointer and
    // leaves the caller's record unset; the samples are copied here
    for (int it = 0; it < nt; it++)
      for (int ix = 0; ix < nx; ix++) seiscal[(size_t)it * nx + ix] = seisobs[(size_t)it * nx0 + ix];
  }
  return 0;
}

extern "C" int rtmh_interp_seis(float* seiscal, const float* seisobs, int nx, int nx0, int nt, int nt0, float dt,
                                float dt0)
{
  std::string err;
  const int rc = interp_seis_impl(seiscal, seisobs, nx, nx0, nt, nt0, dt, dt0, err);
  if (rc != 0) set_err(err);
  return rc;
}

// rtm_real.cu:17-581 with the caller's prototype (RTM_GPU_velstr.cpp:30-35):
// nw and tlength are the direct-wave mute's width and minimum length in
// samples, as rtm_real.cu:28-29 derives them from direct and direct0.
extern "C" void rtm_real(int idevice, int nshot, int startshot, int dshot, int medium_flag, int flag_layout,
                         int sgyflag, int endianflag, int ismth, int nx, int nz, int nt, float dx, float dz, float dt,
                         int nx0, int nz0, int nt0, float dx0, float dz0, float dt0, int npml, int lx, int lz,
                         float tsnap, float fdom, float spz0, float amp, float alp, int nw, int tlength, float* vp,
                         float* rho, float* diffcoef, float* wavelet, char filename_shotgather[40],
                         int** shotprofileId, float* image1, float* image2s, float* image3s, float* image4s,
                         float* image5s, float* image2m, float* image3m, float* image4m, float* image5m,
                         float* illum)
{
  (void)flag_layout; (void)nx0; (void)nz0; (void)dz0; (void)lx; (void)lz; (void)wavelet;
  if ((int)(tsnap / dt) < 1) { fprintf(stderr, "rtm_real: tsnap must be >= dt\n"); exit(1); }
  if (!shotprofileId) { fprintf(stderr, "rtm_real: no shot table (shotprofileId)\n"); exit(1); }
  char path[96];
  snprintf(path, sizeof(path), "./seisReal/%s", filename_shotgather);
  std::vector<int> shots;
  for (int is = startshot; is <= nshot; is += dshot) shots.push_back(is);
  const int nsh = (int)shots.size();
  std::vector<float> field(3 * (size_t)(nsh > 0 ? nsh : 1));
  std::vector<std::vector<float>> rec(nsh);
  std::vector<const float*> recp(nsh);
  // per gather: apertures (here), traces and resampling (on up to 16 host
  // threads, each shot's error kept with it).  The engine reads a record only
  // after the forward pass of its batch, so the preparation runs beside it:
  // seis_wait blocks until the shot's record is done.
  std::vector<std::string> err(nsh);
  std::vector<int> ncol0(nsh), ncol(nsh);
  for (int k = 0; k < nsh; k++) {
    const int j = shots[k] - 1;
    const int sx = shotprofileId[0][j], ntrace = shotprofileId[1][j], off = shotprofileId[3][j];
    // rtm_real.cu:121-136: apertures of the recorded and the modelling grid
    const float offsetmax = (float)(off + (ntrace - 1) * dx0), offset = (float)off, spx = (float)sx;
    const int noffset0 = (int)(offset / dx0), nx10 = (int)(spx / dx0), nx20 = (int)((spx + offsetmax) / dx0);
    const int noffset = (int)(offset / dx), nx1 = (int)(spx / dx), nx2 = (int)((spx + offsetmax) / dx);
    ncol0[k] = nx20 - nx10 + 1 - noffset0;
    ncol[k] = nx2 - nx1 + 1 - noffset;
    if (ncol0[k] < 1 || ncol[k] < 1) { fprintf(stderr, "rtm_real: shot %d has no receivers\n", shots[k]); exit(1); }
    rec[k].assign((size_t)nt * ncol[k], 0.0f);
    recp[k] = rec[k].data();
    field[3 * k] = spx; field[3 * k + 1] = offset; field[3 * k + 2] = offsetmax;
  }
  struct Ready {
    std::mutex mu;
    std::condition_variable cv;
    std::vector<char> done;
    const std::vector<std::string>* err;
  } ready;
  ready.done.assign(nsh, 0);
  ready.err = &err;
  auto prep = [&](int k) {
    const int j = shots[k] - 1;
    std::vector<float> obs((size_t)nt0 * ncol0[k]);
    if (read_gather_impl(path, sgyflag, endianflag, shotprofileId[2][j], shotprofileId[1][j], nt0, obs.data(),
                         ncol0[k], err[k]) == 0)
      interp_seis_impl(rec[k].data(), obs.data(), ncol[k], ncol0[k], nt, nt0, dt, dt0, err[k]);
    {
      std::lock_guard<std::mutex> lk(ready.mu);
      ready.done[k] = 1;
    }
    ready.cv.notify_all();
  };
  const int nth = std::max(1, std::min({nsh, 16, (int)std::thread::hardware_concurrency()}));
  std::vector<std::thread> th;
  for (int t = 0; t < nth; t++)
    th.emplace_back([&, t]() { for (int k = t; k < nsh; k += nth) prep(k); });
  auto wait_cb = [](void* user, int k) -> int {
    Ready* r = (Ready*)user;
    std::unique_lock<std::mutex> lk(r->mu);
    r->cv.wait(lk, [&] { return r->done[k] != 0; });
    return (*r->err)[k].empty() ? 0 : -1;
  };
  rtmh_params P;
  memset(&P, 0, sizeof(P));
  P.nshot = nshot; P.startshot = startshot; P.dshot = dshot; P.mode = 1; P.medium_flag = medium_flag;
  P.flag_layout = 3; P.sgyflag = sgyflag; P.endianflag = endianflag; P.ismth = ismth;
  P.nx = nx; P.nz = nz; P.nt = nt; P.dx = dx; P.dz = dz; P.dt = dt; P.npml = npml;
  // the engine derives the mute from direct, direct0 as (int)(direct/(fdom*dt))
  // and (int)(direct0/dt); half a sample above nw and tlength gives them back
  P.direct = (nw + 0.5f) * fdom * dt;
  P.direct0 = (tlength + 0.5f) * dt;
  P.tsnap = tsnap; P.fdom = fdom; P.spx0 = 0; P.spz0 = spz0; P.dspx = 0; P.offset = 0; P.offsetmax = 0;
  P.amp = amp; P.alp = alp;
  rtmh_ctx* ctx = nullptr;
  if (rtmh_create(idevice, &ctx) != 0) { fprintf(stderr, "rtm_real: %s\n", rtmh_last_error()); exit(1); }
  rtmh_options o;
  memset(&o, 0, sizeof(o));
  o.store_fmt = RTMH_STORE_F32;
  // batches of about a quarter of the job (at least 32 shots): the first
  // batch's reverse pass then needs only the first records, and the host
  // resamples the rest beside the GPU (128 survey gathers: 61 Gcell/s in one
  // batch, 71 in batches of 32)
  o.batch = nsh <= 32 ? 0 : std::max(32, (nsh + 3) / 4);
  o.graph_chunk = 50;
  o.write_files = 1;
  o.verbose = 1;
  const char* ck = getenv("RTMH_CHECKPOINT");
  if (ck && *ck) { o.checkpoint = ck; o.resume = 1; }
  o.ic_extra = 1;
  o.shot_field = field.data();
  o.seis_wait = wait_cb;
  o.seis_user = &ready;
  const size_t n = (size_t)nx * nz;
  std::vector<float> imagetrue(n, 0.0f);
  rtmh_outputs out;
  memset(&out, 0, sizeof(out));
  out.image1 = image1; out.image2s = image2s; out.illum = illum;
  out.image3s = image3s; out.image4s = image4s; out.image5s = image5s;
  out.image3m = image3m; out.image4m = image4m; out.image5m = image5m;
  out.imagetrue = imagetrue.data();
  out.shot_seis_in = recp.data();
  rtmh_stats st;
  memset(&st, 0, sizeof(st));
  const int rc = rtmh_run(ctx, &P, vp, rho, diffcoef, &o, &out, &st);
  for (auto& x : th) x.join();
  for (int k = 0; k < nsh; k++)
    if (!err[k].empty()) { fprintf(stderr, "rtm_real: shot %d: %s\n", shots[k], err[k].c_str()); exit(1); }
  if (rc != 0) {
    fprintf(stderr, "rtm_real: %s\n", rtmh_last_error());
    rtmh_destroy(ctx);
    exit(1);
  }
  rtmh_destroy(ctx);
  memcpy(image2m, image1, n * sizeof(float));
  char ip[64];
  snprintf(ip, sizeof(ip), "./output/%dImagetrue.su", idevice);
  rtmio_write_image(imagetrue.data(), nz, nx, dx, ip);
}
