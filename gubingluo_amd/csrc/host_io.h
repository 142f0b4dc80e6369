// host_io.h -- SU / SEG-Y file formats of the reference, written and read by
// the drop-in host layer (sgyhead.h, sgyhead_function.cpp, IO_function.cpp).
#pragma once
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

// 240-byte SU trace header (sgyhead.h:4-89); field order and sizes fixed by
// the format.  Only the nine key fields are ever non-zero.
typedef struct SuTraceHeader {
  int32_t tracl, tracr, fldr, tracf, ep, cdp, cdpt;
  int16_t trid, nvs, nhs, duse;
  int32_t offset, gelev, selev, sdepth, gdel, sdel, swdep, gwdep;
  int16_t scalel, scalco;
  int32_t sx, sy, gx, gy;
  int16_t counit, wevel, swevel, sut, gut, sstat, gstat, tstat, laga, lagb, delrt, muts, mute, ns, dt, gain, igc,
      igi, corr, sfs, sfe, slen, styp, stas, stae, tatyp, afilf, afils, nofilf, nofils, lcf, hcf, lcs, hcs, year, day,
      hour, minute, sec, timbas, trwf, grnors, grnofr, grnlof, gaps, otrav;
  float d1, f1, d2, f2, ungpow, unscale;
  int32_t ntr;
  int16_t mark, shortpad;
  int16_t unass[14];
} SuTraceHeader;

// the key subset the reference reads and writes (sgyhead.h:92-103)
typedef struct SuKey {
  int32_t iTraceAll, iTrace, Offset, sx, gx;
  int16_t Nt, Dt;
  float Dx;
  int32_t Ntrace;
} SuKey;

// Outseisfile1d (IO_function.cpp:131-182): seis[nt][ntrace] (outflag 0) or
// [ntrace][nt] (outflag 1); sgyflag adds a zero 3600-byte volume header;
// endianflag writes big-endian.  Returns 0 on success.
int rtmio_write_seis(const float* seis, SuKey key, int outflag, int sgyflag, int endianflag, const char* path);

// Inseisfile1d (IO_function.cpp:284-324) with get_parameter (:235-283):
// endianness detected from the first header; fills key (Nt, Dt, Ntrace from
// the file).  seis may be NULL to query sizes.  Returns 0 on success.
int rtmio_read_seis(float* seis, SuKey* key, int inflag, int sgyflag, const char* path, int max_values);

// Output1d with Out_flag=1 (IO_function.cpp:366-398): image[nz][nx] written
// as nx traces of nz samples with a 120-short header.
int rtmio_write_image(const float* image, int nz, int nx, float d, const char* path);

// ReadVelMod (IO_function.cpp:9-35): nx traces of nz floats after 240-byte
// headers -> vp[nz][nx].
int rtmio_read_velocity(const char* path, int nx, int nz, float* vp);

#ifdef __cplusplus
}
#endif
