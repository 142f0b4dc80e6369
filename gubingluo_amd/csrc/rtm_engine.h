// rtm_engine.h -- host runtime of the RTM path on one GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rtm_hip.h"
#include "rtm_dev.h"
#include "rtm_launch.h"

namespace rtm {

// ---- per-shot host set-up -------------------------------------------------
struct ShotSetup {
  rtmh_geom g;
  // SU key header of the shot (rtm_model.cu:137-145)
  int hk_itraceall = 0, hk_offset = 0, hk_sx = 0, hk_gx = 0, hk_ntrace = 0;
  short hk_nt = 0, hk_dt = 0;
  float hk_dx = 0.f;
  float vpmax = 0.f;
  // damping a = c1/c2, b = 1/c2 at p nodes (x, z) and velocity nodes (xv, zv)
  std::vector<float> ax, bx, axv, bxv, az, bz, azv, bzv;
  int sx0 = 0, sx1 = -1, sz0 = 0, sz1 = -1;                          // source box
  std::vector<float> sa;                                              // dt*src in the box
  std::vector<int> cut;                                               // mute
  std::vector<float> taper;                                           // nxl
  int nwin = 0;
  // padded v^2 and density windows (nzp x nxp) of a job with model smoothing
  // (medium_flag / ismth); empty otherwise (the slot then reads the resident
  // model)
  std::vector<float> wvv, wrho;
};

// shot window and padded geometry (rtm_model.cu:122-153 and the layout
// variants); returns false if the window is empty or the layout unknown
bool shot_geometry(const rtmh_params& P, int ishot, ShotSetup& s, const float* field = nullptr);
// host set-up of one shot given the job's v^2 column maxima; with a padded v^2
// window (wvv, nzp x nxp) vpmax and the mute come from it instead
void shot_setup(const rtmh_params& P, const float* vp0, const float* colmax_vv, ShotSetup& s,
                const float* wvv = nullptr);
// the padded windows of a shot (extractvel1 / extractrho1 + extendvel1) and,
// with medium_flag, their absorbing layers smoothed (pmlvelsmooth1d)
void shot_windows(const rtmh_params& P, const float* vp0, const float* rho0, ShotSetup& s, bool smooth_here = true);
// windows the GPU layer smoother takes (one wave per layer side, mirrors of
// interior cells only, edge copies in LDS)
bool pml_smooth_on_gpu(int nx, int nz, int npml);
// velsmooth1d weights of half-width nsp (row-major) and their sum
void gauss_weights(int nsp, std::vector<double>& wt, double& norm);
// Ricker of kernel.cu:13-22
void ricker(std::vector<float>& w, int nt, float fdom, float dt);
// taper row (kernel.cu:1591-1604)
void taper_row(std::vector<float>& t, int nxl, int nwin);

// ---- device memory pool ------------------------------------------------------
// First-fit carving of one range [0, bytes) into blocks of whole 256-byte
// units; put() merges a freed block with the free ranges either side of it.
// Offsets only: the device pool lays it over one device allocation.
class Extents {
 public:
  static constexpr size_t kNone = ~(size_t)0;
  void reset(size_t bytes)
  {
    holes_.clear();
    blocks_.clear();
    if (bytes) holes_[0] = bytes;
  }
  // offset of a block of at least `bytes` (at least one unit, so that no two
  // blocks share an offset), kNone when no free range holds it
  size_t get(size_t bytes)
  {
    if (bytes > kNone - 255) return kNone;             // would wrap when rounded up
    const size_t need = bytes ? (bytes + 255) & ~(size_t)255 : 256;
    for (auto it = holes_.begin(); it != holes_.end(); ++it) {
      if (it->second < need) continue;
      const size_t off = it->first, len = it->second;
      holes_.erase(it);
      if (len > need) holes_[off + need] = len - need;
      blocks_[off] = need;
      return off;
    }
    return kNone;
  }
  // false for an offset that is not the start of a block in use
  bool put(size_t off)
  {
    auto bl = blocks_.find(off);
    if (bl == blocks_.end()) return false;
    size_t start = off, len = bl->second;
    blocks_.erase(bl);
    auto next = holes_.lower_bound(off);
    if (next != holes_.end() && next->first == off + len) {
      len += next->second;
      next = holes_.erase(next);
    }
    if (next != holes_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        start = prev->first;
        len += prev->second;
        holes_.erase(prev);
      }
    }
    holes_[start] = len;
    return true;
  }
  const std::map<size_t, size_t>& holes() const { return holes_; }
  const std::map<size_t, size_t>& blocks() const { return blocks_; }
 private:
  std::map<size_t, size_t> holes_;    // offset -> length of the free ranges
  std::map<size_t, size_t> blocks_;   // offset -> length of the blocks in use
};

class DevPool {
 public:
  ~DevPool();
  void* get(size_t bytes);
  void put(void* p);
  void trim();
  // Take `bytes` of device memory in one allocation now and carve every later
  // get() from it (first fit, 256-byte blocks, free neighbours merged on
  // put()): no device allocation call after this one.  False (and the pool
  // stays as it was) if the device refuses or a block of the pool is in use.
  bool reserve(size_t bytes);
  size_t held() const { return arena_ ? arena_bytes_ : held_; }
  // host seconds spent in hipMalloc / hipFree, and the number of hipMalloc calls
  double alloc_seconds() const { return alloc_s_; }
  long long allocations() const { return nalloc_; }
 private:
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> size_;
  size_t held_ = 0;
  double alloc_s_ = 0;
  long long nalloc_ = 0;
  char* arena_ = nullptr;
  size_t arena_bytes_ = 0;
  Extents ext_;
};

// ---- engine -----------------------------------------------------------------
// a model on the device as the slots read it: density, v^2, and the step
// kernels' factors 1/rho and rho v^2, rows of `pitch` floats
struct ModelDev {
  const float *rho, *vv, *ib, *kk;
  int pitch;
};

// decimated snapshots of a job: every `every` steps from k0*every while
// t >= 4/fdom (the wavelet has passed), n of them up to nt-2; fmt the store format
struct SnapPlan {
  int every = 0, k0 = 0, n = 0, fmt = 0;
  bool on() const { return every > 0 && n > 0; }
};
SnapPlan snap_plan(const rtmh_params& P, int every, int fmt);

struct SlotMem {
  void* base = nullptr;
  size_t bytes = 0;
};

// Everything a captured chunk bakes in by value: the launch shapes (tiles,
// tile height, slots: k_save's and k_ic's grids and k_bump's count), the
// kernels chosen (phase, store format, imaging conditions, guard) and the
// control-block addresses.  Descriptors and tiles are read from the control
// block at replay, so batches that agree on all of these share a graph.
struct GraphKey {
  int ntiles, tz, phase, save, fmt, k;     // save: strip save after every (forward) step, fmt its store format
  int ic;                                  // cells of the other imaging conditions after every reverse step (0: off)
  int nslots;                              // slots of the batch (k_save / k_ic grids, k_bump count)
  int guard;                               // wavefield guard after the chunk
  int snap;                                // snapshot launches after every step: 0 off, else 1 + format
  int skip;                                // zero-tile skip of the step kernels
  const void* shots;
  const void* tiles;
  const void* slots;
  bool operator<(const GraphKey& o) const;
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  int device() const { return device_; }
  // device memory the engine keeps for reuse by later jobs, released
  // (a served form sizes its budget from the memory left free)
  void release_cached() { pool_.trim(); }
  hipStream_t stream() const { return st_; }

  // device-resident velocity models (rho and v^2), shared by many jobs
  struct Model {
    int nx = 0, nz = 0;
    float* d_rho = nullptr;
    float* d_vv = nullptr;
    float* d_ib = nullptr;      // 1/rho
    float* d_kk = nullptr;      // rho v^2
    ModelDev dev() const { return ModelDev{d_rho, d_vv, d_ib, d_kk, nx}; }
    std::vector<float> vp0;     // host copy: mute velocities
    std::vector<float> rho0;    // host copy: smoothed windows
    std::vector<float> colmax;  // host: column maxima of v^2 (per-shot vmax)
    size_t bytes = 0;
  };
  int model_create(int nx, int nz, const float* vp0, const float* rho0, int* id);
  int model_destroy(int id);
  const Model* model(int id) const;
  // checkpoint file of a job (FP64 stacks after the shots done)
  static uint64_t job_fingerprint(const rtmh_params& P, const Model& M, const float* coef, const rtmh_options& opt,
                                  int nsh);
  static int read_checkpoint(const char* path, uint64_t fp, int nx, int nz, int nstk, int& done, std::vector<double>& st3);
  static int write_checkpoint(const char* path, uint64_t fp, int nx, int nz, int done, const std::vector<double>& st3);
  // where a job of nsh shots resumes from its checkpoint: 1 with the shots
  // done and the stacks, 0 without a file (start at shot 0), -5 (why set) for
  // another job's or a damaged file, or one with more shots done than nsh
  static int resume_point(const char* path, uint64_t fp, int nx, int nz, int nstk, int nsh, int& done,
                          std::vector<double>& st3, std::string& why);

  int run(const rtmh_params& P, const float* vp0, const float* rho0, const float* coef, const rtmh_options& opt,
          rtmh_outputs* out, rtmh_stats* stats);
  int run_model(int model_id, const rtmh_params& P, const float* coef, const rtmh_options& opt, rtmh_outputs* out,
                rtmh_stats* stats);
  // pmlvelsmooth1d of host windows (nz[k] rows of nx[k]) on the GPU, in place
  int pml_smooth_windows(std::vector<float*>& wins, const std::vector<int>& nx, const std::vector<int>& nz, int npml);
  // every shot's padded windows (and, with medium_flag, their smoothed layers)
  int job_windows(const rtmh_params* pk, const float* vp0, const float* rho0, std::vector<ShotSetup>& setup);
  // velsmooth1d of host windows on the GPU, in place
  int gauss_smooth_windows(std::vector<float*>& wins, const std::vector<int>& nx, const std::vector<int>& nz, int nsp);
  int fd_steps(const rtmh_params& P, int ishot, const float* vp0, const float* rho0, const float* coef, int nsteps,
               int it0, int with_source, float* pf, float* vxf, float* vzf, float* p1f, float* p2f);

  // device-to-device copy bandwidth (read + write bytes per second / 1e9)
  int copy_probe(size_t bytes, int reps, double* gbps);

  // bytes one slot of this geometry needs on the device
  static size_t slot_bytes(const ShotSetup& s, int nt, int store_fmt, bool ic = false,
                           const SnapPlan* snap = nullptr);
  // build the device descriptor of one slot inside `base` (uploads: host
  // images of its small arrays to copy to the device)
  // (the slot reads `model`, or its own smoothed windows when s carries them)
  static void layout_slot(const rtmh_params& P, const ShotSetup& s, int store_fmt, char* base, ShotDev& d,
                          std::vector<std::pair<void*, std::vector<char>>>& uploads, const float* coef_dev,
                          const float* wav_dev, const ModelDev& model, bool ic = false,
                          const SnapPlan* snap = nullptr);
  // offset of the taper row inside a slot block
  static size_t taper_offset(const ShotSetup& s);
  std::mutex& models_mutex() { return models_mu_; }
  // row pitch: RTM_MXL margin columns, the grid rounded up to the tile width,
  // 32 more; rows: RTM_MZT margin rows, the grid rounded up to the tallest
  // tile, RTM_MZB more
  static int pitch_of(int nxp) { return RTM_MXL + ((nxp + 63) / 64) * 64 + 32; }
  static int frows_of(int nzp) { return RTM_MZT + ((nzp + RTM_ZALIGN - 1) / RTM_ZALIGN) * RTM_ZALIGN + RTM_MZB; }

 private:
  int device_;
  hipStream_t st_ = nullptr;
  DevPool pool_;
  std::map<GraphKey, hipGraphExec_t> graphs_;
  std::map<int, Model> models_;
  std::map<long long, int> tuned_;      // geometry class -> tile height
  static constexpr int kTuneSteps = 12;
  mutable std::mutex models_mu_;
  int next_model_ = 1;
  // persistent control block: captured graphs reference these addresses
  void* cb_ = nullptr;
  int cb_slots_ = 0;
  size_t cb_tiles_ = 0;
  ShotDev* d_ds_ = nullptr;
  Tile* d_tiles_ = nullptr;
  int* d_slots_ = nullptr;
  struct StackShotHost* d_stk_ = nullptr;
  unsigned int* d_flag_ = nullptr;      // wavefield guard of the running pass (largest |p| bits)
  unsigned long long* d_act_ = nullptr; // zero-tile skip words of the running pass (2 x tiles)
  int ic_cells_ = 0;                    // the other imaging conditions of the running batch: interior cells (0: off)
  SnapPlan snap_;                       // snapshots of the running job
  int ensure_cb(int nslots, size_t ntiles);

  // guard_every > 0: the wavefield guard after every guard_every steps (plain
  // launches) or after every replayed chunk, and after the last step
  int run_pass(int phase, int nsteps, const ShotDev* dshots, const Tile* dtiles, int ntiles,
               const std::vector<int>& slots, const int* dslots, int tz, int store_fmt, int chunk,
               rtmh_stats* prof, int guard_every = 0, bool skip = false);
  // clears the guard flag before a pass; after it, checks the flag and names
  // the shots whose wavefield is not finite (-7)
  int guard_clear();
  int guard_check(const std::vector<ShotDev>& ds, const std::vector<int>& shots, int k0, const char* pass);
  hipGraphExec_t graph_for(const GraphKey& k, const ShotDev* dshots, const Tile* dtiles, const int* dslots,
                           int nslots, int store_fmt);
  void drop_graphs();
};

// tile table of a list of slot descriptors with the per-tile flags
// (tiles carry the index of the descriptor in `ds` plus `slot_base`)
void build_tiles(const std::vector<ShotDev>& ds, int tz, std::vector<Tile>& tiles, const std::vector<int>* slot_ids = nullptr);
// the table as 8 per-XCD runs of equal estimated cost, long tiles first in
// each, padded with empty entries (slot -1) (phase: PH_FWD / PH_REV)
void layout_runs(std::vector<Tile>& tiles, int phase);
// Tile::nbr of every tile: the table positions of the tiles of height tz
// above, below and beside it in its slot (after the final layout)
void link_neighbours(std::vector<Tile>& tiles, int tz);
// Whether the forward step's zero-tile skip may run on tiles of height tz.
// One step reads the pressure 2*RTM_R - 1 rows above and below a cell (the
// velocity halo recomputed from the pressure RTM_R - 1 / RTM_R rows further
// out), and the skip looks at the 8 adjacent tiles only: that covers the
// reach when tz >= 2*RTM_R - 1 (32 and 64 rows), not for 16-row tiles, where
// a wavefield front in the last rows of the tile two above can reach the
// first rows of a tile whose neighbours are still +0.
inline bool fwd_skip_ok(int tz) { return tz >= 2 * RTM_R - 1; }
int params_check(const rtmh_params& P, std::string& why);
// the stability bound of get_constant (RTM_GPU_velstr.cpp:355-361):
// min(dx, dz) / (sqrt(2) sum|c| vmax), sum over the order-20 coefficient row
// (coef[90..99]); vmax in m/s
float dt_stable(const rtmh_params& P, const float* coef, float vmax);
// 0 when dt is within the bound, -1 with the reason otherwise
int stability_check(const rtmh_params& P, const float* coef, float vmax, std::string& why);

// thread-local error text for rtmh_last_error()
void set_error(const std::string& s);
const char* get_error();

}  // namespace rtm
