// host_io.cpp -- SU / SEG-Y I/O of the drop-in host layer.
// Byte layouts follow sgyhead.h:4-103 and IO_function.cpp; see host_io.h.
#include "host_io.h"

#include <stdlib.h>
#include <string.h>

static_assert(sizeof(SuTraceHeader) == 240, "SU trace header must be 240 bytes");

static inline uint16_t bswap16(uint16_t v) { return (uint16_t)((v >> 8) | (v << 8)); }
static inline uint32_t bswap32(uint32_t v)
{
  return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}
static inline void swap16(void* p) { uint16_t v; memcpy(&v, p, 2); v = bswap16(v); memcpy(p, &v, 2); }
static inline void swap32(void* p) { uint32_t v; memcpy(&v, p, 4); v = bswap32(v); memcpy(p, &v, 4); }

// byte-swap every field of a trace header (swap_sgyhead, sgyhead_function.cpp)
static void swap_header(SuTraceHeader* h)
{
  unsigned char* b = (unsigned char*)h;
  // 4-byte fields: 0..28, 36..68, 72..88, 180..208 ; everything else 2-byte
  static const int four[] = {0, 4, 8, 12, 16, 20, 24, 36, 40, 44, 48, 52, 56, 60, 64, 72, 76, 80, 84,
                             180, 184, 188, 192, 196, 200, 204};
  char is4[240] = {0};
  for (int o : four) { is4[o] = 1; swap32(b + o); }
  for (int o = 0; o < 240;) {
    if (is4[o]) { o += 4; continue; }
    swap16(b + o);
    o += 2;
  }
}

extern "C" int rtmio_write_seis(const float* seis, SuKey key, int outflag, int sgyflag, int endianflag, const char* path)
{
  FILE* fp = fopen(path, "wb");
  if (!fp) { printf("The file %s open failed !\n", path); return -1; }
  if (sgyflag) {
    short vol[1800];
    memset(vol, 0, sizeof(vol));
    fwrite(vol, sizeof(short), 1800, fp);
  }
  const int nt = key.Nt, nx = key.Ntrace;
  float* tr = (float*)malloc(sizeof(float) * (nt > 0 ? nt : 1));
  for (int ix = 0; ix < nx; ix++) {
    SuTraceHeader h;
    memset(&h, 0, sizeof(h));
    h.tracl = key.iTraceAll + ix;
    h.tracf = key.iTrace + ix;
    h.offset = key.Offset + (int)(ix * key.Dx);
    h.sx = key.sx;
    h.gx = key.gx + (int)(ix * key.Dx);
    h.ns = key.Nt;
    h.dt = key.Dt;
    h.d2 = key.Dx;
    h.ntr = nx;
    if (endianflag) swap_header(&h);
    fwrite(&h, sizeof(h), 1, fp);
    for (int it = 0; it < nt; it++) tr[it] = outflag ? seis[(size_t)ix * nt + it] : seis[(size_t)it * nx + ix];
    if (endianflag)
      for (int it = 0; it < nt; it++) swap32(&tr[it]);
    fwrite(tr, sizeof(float), nt, fp);
  }
  free(tr);
  fclose(fp);
  return 0;
}

// the nine key fields of a header at the current file position (fread_sgyhead)
static int read_key(FILE* fp, SuKey* k, int swap)
{
  SuTraceHeader h;
  if (fread(&h, sizeof(h), 1, fp) != 1) return -1;
  if (swap) swap_header(&h);
  k->iTraceAll = h.tracl; k->iTrace = h.tracf; k->Offset = h.offset; k->sx = h.sx; k->gx = h.gx;
  k->Nt = h.ns; k->Dt = h.dt; k->Dx = h.d2; k->Ntrace = h.ntr;
  return 0;
}

extern "C" int rtmio_read_seis(float* seis, SuKey* key, int inflag, int sgyflag, const char* path, int max_values)
{
  FILE* fp = fopen(path, "rb");
  if (!fp) { printf("The file %s open failed !\n", path); return -1; }
  long start = sgyflag ? 3600L : 0L;
  fseek(fp, start, SEEK_SET);
  SuKey k;
  if (read_key(fp, &k, 0) != 0) { fclose(fp); return -2; }
  int swap = 0;
  fseek(fp, 0, SEEK_END);
  const long sz = ftell(fp) - start;
  fseek(fp, start, SEEK_SET);
  SuKey k2;
  read_key(fp, &k2, 1);
  // get_parameter's probe (IO_function.cpp:253-259) swaps when dt or ns is
  // not positive; a byte-swapped header can also look positive, so a header
  // whose trace length does not tile the file loses to one that does
  const bool fits = k.Nt > 0 && sz % (240L + 4L * k.Nt) == 0;
  const bool fits2 = k2.Nt > 0 && k2.Dt > 0 && sz % (240L + 4L * k2.Nt) == 0;
  if ((k.Dt <= 0 || k.Nt <= 0 || !fits) && fits2) { swap = 1; k = k2; }
  const int nt = k.Nt;
  if (nt <= 0) { fclose(fp); return -3; }
  const int ntrace = (int)(sz / (240L + 4L * nt));
  k.Ntrace = ntrace;
  *key = k;
  if (!seis) { fclose(fp); return 0; }
  if ((long long)ntrace * nt > (long long)max_values) { fclose(fp); return -4; }
  float* tr = (float*)malloc(sizeof(float) * nt);
  fseek(fp, start, SEEK_SET);
  for (int ix = 0; ix < ntrace; ix++) {
    fseek(fp, 240L, SEEK_CUR);
    if (fread(tr, sizeof(float), nt, fp) != (size_t)nt) { free(tr); fclose(fp); return -5; }
    if (swap)
      for (int it = 0; it < nt; it++) swap32(&tr[it]);
    for (int it = 0; it < nt; it++) {
      if (inflag) seis[(size_t)ix * nt + it] = tr[it];
      else seis[(size_t)it * ntrace + ix] = tr[it];
    }
  }
  free(tr);
  fclose(fp);
  return 0;
}

extern "C" int rtmio_write_image(const float* image, int nz, int nx, float d, const char* path)
{
  FILE* fp = fopen(path, "wb");
  if (!fp) { printf("The file %s open failed !\n", path); return -1; }
  short header[120];
  memset(header, 0, sizeof(header));
  header[57] = (short)nz;
  // (short)(d*1000000.0): the double->short conversion of IO_function.cpp:384
  // keeps the low 16 bits of the integer value on x86.
  header[58] = (short)(int)(d * 1000000.0);
  header[104] = (short)nx;
  float* tr = (float*)malloc(sizeof(float) * (nz > 0 ? nz : 1));
  for (int ix = 0; ix < nx; ix++) {
    header[0] = (short)(ix + 1);
    fwrite(header, 2, 120, fp);
    for (int iz = 0; iz < nz; iz++) tr[iz] = image[(size_t)iz * nx + ix];
    fwrite(tr, sizeof(float), nz, fp);
  }
  free(tr);
  fclose(fp);
  return 0;
}

extern "C" int rtmio_read_velocity(const char* path, int nx, int nz, float* vp)
{
  FILE* fp = fopen(path, "rb");
  if (!fp) { printf("The file %s open failed !\n", path); return -1; }
  float* tr = (float*)malloc(sizeof(float) * nz);
  int rc = 0;
  for (int ix = 0; ix < nx; ix++) {
    fseek(fp, 240L, SEEK_CUR);
    if (fread(tr, sizeof(float), nz, fp) != (size_t)nz) { rc = -2; break; }
    for (int iz = 0; iz < nz; iz++) vp[(size_t)iz * nx + ix] = tr[iz];
  }
  free(tr);
  fclose(fp);
  return rc;
}
