// rtm_server.cpp -- the served form: continuous batching of RTM requests on
// one GPU (rtmh_server_* in include/rtm_hip.h).
//
// A worker thread owns a stream and a slot table.  Each tick launches one
// time step for every resident shot: the forward-phase shots in one
// velocity + pressure launch pair, the reverse-phase shots in another, each
// shot at its own time index (it = it_off + it_dir * tick).  Phase changes
// and completions are known in advance from the tick count, so the host
// never waits on the GPU to schedule; it only polls the events that close
// requests.  Queued shots are admitted in arrival order whenever a slot and
// budget are free (per-shot bytes are known before admission), so requests
// that arrive while others run start at the next tick.  A finished shot's
// image and illumination move to a small pending buffer, its slot is
// released, and each request stacks its shots in job order into its own
// FP64 images -- a request's result is the same alone or in any mix.
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/rtm_hip.h"
#include "admission.h"
#include "rtm_engine.h"
#include "rtm_launch.h"

namespace rtm {

static inline size_t a256(size_t n) { return (n + 255) & ~(size_t)255; }
static constexpr long long kGuardTicks = 64;   // served form: wavefield guard interval (a power of two)

struct SrvRequest {
  int id = 0, model_id = 0;
  rtmh_params P;
  std::vector<float> coef;
  std::vector<ShotSetup> setup;
  std::vector<int> shots;
  int next_admit = 0, next_stack = 0;
  std::vector<float*> pend_g1, pend_il;   // finished shots waiting to be stacked
  std::vector<char> finished;
  char* dev = nullptr;                     // coef, wavelet, tapers, stack entries, stacks, outputs
  size_t dev_bytes = 0;
  float *d_coef = nullptr, *d_wav = nullptr, *d_taper = nullptr;
  StackShotHost* d_stk = nullptr;
  double* d_wt = nullptr;                  // velsmooth1d weights (ismth > 0)
  double wt_norm = 0;
  double *d_i1 = nullptr, *d_i2 = nullptr, *d_il = nullptr;
  float *d_f1 = nullptr, *d_f2 = nullptr, *d_f3 = nullptr, *d_lap = nullptr, *d_nrm = nullptr;
  std::vector<size_t> taper_off;
  std::vector<float> image1, image2s, illum, image_lap, image_norm;
  std::vector<std::vector<float>> stage;   // host images uploaded asynchronously (kept alive)
  double t_submit = 0, t_start = -1, t_done = -1, cells = 0;
  hipEvent_t done_ev = nullptr;
  bool finalizing = false, complete = false;
  int status = 0;                          // 0, or the error of a failed request
  std::string err;
  int resident = 0;                        // shots of this request in slots
  std::vector<unsigned int> guard;         // per shot: largest |p| bits (k_guard)
};

struct SrvSlot {
  int req = -1, k = -1, phase = PH_IDLE;
  int steps_left = 0;
  void* mem = nullptr;
  void* pend = nullptr;   // the shot's image and illumination once it finishes, until stacked
  size_t bytes = 0;
  std::vector<std::pair<void*, std::vector<char>>> uploads;
};

class Server {
 public:
  Server(Engine* eng, const rtmh_server_options& o);
  ~Server();
  int start();
  int submit(int model_id, const rtmh_params& P, const float* coef, int* id);
  int request_bytes(int model_id, const rtmh_params& P, const float* coef, double* shot_max, double* request);
  int wait(int id, double timeout, rtmh_outputs* out, rtmh_request_stats* st);
  void query(rtmh_server_stats* st);

 private:
  void loop();
  double now() const
  {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
  }
  Admit admit_one(SrvRequest& r, std::string& why);   // under mu_
  int request_device(SrvRequest& r);       // under mu_
  void fail_request(SrvRequest& r, int status, const std::string& why);   // under mu_
  void complete_request(SrvRequest& r);    // under mu_
  static size_t request_bytes(const SrvRequest& r);
  size_t shot_need(const SrvRequest& r, int k) const;
  void transition(int sid);
  void finish_shot(int sid);
  void stack_ready(SrvRequest& r);
  void finalize(SrvRequest& r);
  void rebuild_tiles();
  void fail(const char* what, hipError_t e);

  Engine* eng_;
  rtmh_server_options opt_;
  hipStream_t st_ = nullptr;
  DevPool pool_;
  std::thread th_;
  std::mutex mu_;
  std::mutex smooth_mu_;                   // submitters preparing smoothed windows on the engine
  std::condition_variable cv_, cv_done_;
  std::deque<int> queue_;
  std::map<int, std::unique_ptr<SrvRequest>> reqs_;
  std::vector<SrvSlot> slots_;
  std::vector<ShotDev> hds_;
  ShotDev* d_ds_ = nullptr;
  Tile *d_tf_ = nullptr, *d_tr_ = nullptr;
  int *d_qs_ = nullptr;
  int *d_as_ = nullptr;                    // every resident slot (periodic guard)
  int *d_id_ = nullptr;                    // 0..max_slots-1 (single-slot launches)
  int na_ = 0;
  std::vector<int> aslots_;
  size_t tile_cap_ = 0;
  int ntf_ = 0, ntr_ = 0, nq_ = 0;
  std::vector<Tile> tf_, tr_;
  std::vector<int> qslots_;
  bool dirty_ = true, stop_ = false;
  double budget_ = 0, used_ = 0;
  // written by the loop only, read by query() from other threads
  std::atomic<long long> tick_{0}, launches_{0};
  std::atomic<double> busy_{0};
  int next_id_ = 1, peak_ = 0, active_ = 0;
  double cells_done_ = 0;
  // The first device error ends the server: its code (atomic: the loop reads
  // it without a lock, submit and wait under mu_) and its message (under
  // err_mu_, which is taken last -- nothing else is locked while it is held --
  // so fail() may run with or without mu_ held).
  std::atomic<int> err_{0};
  std::mutex err_mu_;
  std::string err_msg_;
  void fail_with(int code, std::string msg);
  std::string err_message();
  std::chrono::steady_clock::time_point t0_;
};

Server::Server(Engine* eng, const rtmh_server_options& o) : eng_(eng), opt_(o)
{
  if (opt_.max_slots <= 0) opt_.max_slots = 256;
  // 32-row tiles: on the served stream of 8 models (grids 256^2..2048x1024)
  // they ran 5 % faster than 16 rows and 9 % faster than 64
  if (opt_.tile_z <= 0) opt_.tile_z = 32;
  t0_ = std::chrono::steady_clock::now();
}

int Server::start()
{
  if (hipSetDevice(eng_->device()) != hipSuccess) { set_error("hipSetDevice failed"); return -10; }
  if (hipStreamCreateWithFlags(&st_, hipStreamNonBlocking) != hipSuccess) { set_error("stream"); return -10; }
  // the context runs no rtmh_run* job while it serves: slot memory its
  // earlier jobs left cached goes back to the device before the budget is
  // taken from the free memory
  eng_->release_cached();
  size_t fr = 0, tot = 0;
  (void)hipMemGetInfo(&fr, &tot);
  budget_ = opt_.mem_budget > 0 ? opt_.mem_budget : 0.8 * (double)fr;
  // the whole budget in one device allocation, carved into request blocks and
  // slots as shots come and go: a hipMalloc in the serving loop stalls the
  // stream it feeds (tens of ms each in a process whose device memory is
  // fragmented).  A budget the device cannot hold in one piece is served by
  // allocations on demand, as before.
  if (pool_.reserve((size_t)budget_)) budget_ = (double)pool_.held();
  slots_.assign(opt_.max_slots, SrvSlot());
  hds_.assign(opt_.max_slots, ShotDev());
  if (hipMalloc(&d_ds_, sizeof(ShotDev) * opt_.max_slots) != hipSuccess ||
      hipMalloc(&d_qs_, sizeof(int) * opt_.max_slots) != hipSuccess ||
      hipMalloc(&d_as_, sizeof(int) * opt_.max_slots) != hipSuccess ||
      hipMalloc(&d_id_, sizeof(int) * opt_.max_slots) != hipSuccess) {
    (void)hipGetLastError();
    set_error("out of device memory (server control)");
    return -3;
  }
  {
    std::vector<int> id(opt_.max_slots);
    for (int i = 0; i < opt_.max_slots; i++) id[i] = i;
    if (hipMemcpy(d_id_, id.data(), id.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
      set_error("server control upload");
      return -10;
    }
  }
  th_ = std::thread([this] { loop(); });
  return 0;
}

Server::~Server()
{
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (th_.joinable()) th_.join();
  if (st_) (void)hipStreamSynchronize(st_);
  for (auto& kv : reqs_) {
    if (kv.second->done_ev) (void)hipEventDestroy(kv.second->done_ev);
    if (kv.second->dev) pool_.put(kv.second->dev);
  }
  for (auto& s : slots_) {
    if (s.mem) pool_.put(s.mem);
    if (s.pend) pool_.put(s.pend);
  }
  (void)hipFree(d_ds_); (void)hipFree(d_qs_); (void)hipFree(d_as_); (void)hipFree(d_id_);
  (void)hipFree(d_tf_); (void)hipFree(d_tr_);
  pool_.trim();
  if (st_) (void)hipStreamDestroy(st_);
}

void Server::fail(const char* what, hipError_t e)
{
  fail_with(-10, std::string(what) + ": " + hipGetErrorString(e));
}

void Server::fail_with(int code, std::string msg)
{
  std::lock_guard<std::mutex> lk(err_mu_);
  if (err_.load() != 0) return;                        // keep the first error
  err_msg_ = std::move(msg);
  err_.store(code);
}

std::string Server::err_message()
{
  std::lock_guard<std::mutex> lk(err_mu_);
  return err_msg_;
}

int Server::request_bytes(int model_id, const rtmh_params& P, const float* coef, double* shot_max, double* request)
{
  std::string why;
  if (params_check(P, why) != 0) { set_error(why); return -1; }
  const Engine::Model* M = eng_->model(model_id);
  if (!M || M->nx != P.nx || M->nz != P.nz) { set_error("no such model or size mismatch"); return -1; }
  SrvRequest r;
  r.P = P;
  r.coef.assign(coef, coef + 100);
  for (int is = P.startshot; is <= P.nshot; is += P.dshot) r.shots.push_back(is);
  r.setup.resize(r.shots.size());
  int nwin = 50;
  size_t need = 0;
  for (size_t k = 0; k < r.shots.size(); k++) {
    if (!shot_geometry(P, r.shots[k], r.setup[k])) { set_error("shot window outside the model"); return -2; }
    shot_setup(P, M->vp0.data(), M->colmax.data(), r.setup[k]);
    while (2 * nwin >= r.setup[k].g.nxl) nwin -= 3;
    taper_row(r.setup[k].taper, r.setup[k].g.nxl, nwin);
    need = std::max(need, shot_need(r, (int)k));
  }
  *shot_max = (double)need;
  *request = (double)request_bytes(r);
  return 0;
}

int Server::submit(int model_id, const rtmh_params& P, const float* coef, int* id)
{
  std::string why;
  if (params_check(P, why) != 0) { set_error(why); return -1; }
  if (P.mode != 0) { set_error("served requests model their data (mode 0)"); return -1; }
  const Engine::Model* M = eng_->model(model_id);
  if (!M || M->nx != P.nx || M->nz != P.nz) { set_error("no such model or size mismatch"); return -1; }
  auto r = std::make_unique<SrvRequest>();
  r->model_id = model_id;
  r->P = P;
  r->coef.assign(coef, coef + 100);
  for (int is = P.startshot; is <= P.nshot; is += P.dshot) r->shots.push_back(is);
  const int nsh = (int)r->shots.size();
  if (nsh == 0) { set_error("no shots"); return -1; }
  r->setup.resize(nsh);
  for (int k = 0; k < nsh; k++)
    if (!shot_geometry(P, r->shots[k], r->setup[k])) { set_error("shot window outside the model"); return -2; }
  // model smoothing: each shot's padded windows (pmlvelsmooth1d on the GPU
  // with medium_flag), read by its slot instead of the resident model
  const bool smooth = P.medium_flag != 0 || P.ismth > 0;
  if (smooth) {
    std::vector<rtmh_params> pk(nsh, P);
    std::lock_guard<std::mutex> lk(smooth_mu_);
    const int rc = eng_->job_windows(pk.data(), M->vp0.data(), M->rho0.data(), r->setup);
    if (rc) return rc;
  }
  int nwin = 50;
  for (int k = 0; k < nsh; k++) {
    shot_setup(P, M->vp0.data(), M->colmax.data(), r->setup[k], smooth ? r->setup[k].wvv.data() : nullptr);
    while (2 * nwin >= r->setup[k].g.nxl) nwin -= 3;
    r->setup[k].nwin = nwin;
    taper_row(r->setup[k].taper, r->setup[k].g.nxl, nwin);
    const double nxz = (double)r->setup[k].g.nxp * r->setup[k].g.nzp;
    r->cells += 2.0 * P.nt * nxz + (double)(P.nt - 1) * P.nz * r->setup[k].g.nxl;
  }
  r->pend_g1.assign(nsh, nullptr);
  r->pend_il.assign(nsh, nullptr);
  r->finished.assign(nsh, 0);
  r->guard.assign(nsh, 0u);
  {
    float mv = 0.0f;
    for (float c : M->colmax) mv = c > mv ? c : mv;
    if (stability_check(P, coef, sqrtf(mv), why) != 0) { set_error(why); return -1; }
  }
  // a shot that cannot fit the budget even alone is refused here, not queued
  size_t need = 0;
  for (int k = 0; k < nsh; k++) need = std::max(need, shot_need(*r, k));
  const size_t extra = request_bytes(*r);
  std::lock_guard<std::mutex> lk(mu_);
  if (const int e = err_.load()) { set_error(err_message()); return e; }
  if (admit_rule(budget_, 0.0, 0, false, (double)need, (double)extra) == Admit::kReject) {
    set_error("a shot of this request needs " + std::to_string(need + extra) + " bytes of HBM (slot " +
              std::to_string(need) + ", request block " + std::to_string(extra) + "), above the server's budget of " +
              std::to_string((size_t)budget_));
    return -3;
  }
  r->id = next_id_++;
  r->t_submit = now();
  *id = r->id;
  queue_.push_back(r->id);
  reqs_[r->id] = std::move(r);
  cv_.notify_all();
  return 0;
}

// bytes of a request's own device block (the layout of request_device)
size_t Server::request_bytes(const SrvRequest& r)
{
  const size_t nm = (size_t)r.P.nx * r.P.nz, nsh = r.shots.size();
  size_t o = a256(400) + a256((size_t)r.P.nt * 4);
  for (size_t k = 0; k < nsh; k++) o += a256(r.setup[k].taper.size() * 4);
  const size_t nw = r.P.ismth > 0 ? (size_t)(2 * r.P.ismth + 1) * (2 * r.P.ismth + 1) : 0;
  return o + a256(nsh * sizeof(StackShotHost)) + 3 * a256(nm * 8) + 5 * a256(nm * 4) + a256(nw * 8);
}

// bytes shot k holds while resident: its slot and, once finished, its pending
// image and illumination until it is stacked
size_t Server::shot_need(const SrvRequest& r, int k) const
{
  return Engine::slot_bytes(r.setup[k], r.P.nt, opt_.store_fmt) + 2 * a256((size_t)r.P.nz * r.setup[k].g.nxl * 4);
}

// request-level device block: coefficients, wavelet, tapers, stack entries,
// FP64 stacks and fp32 outputs
int Server::request_device(SrvRequest& r)
{
  const rtmh_params& P = r.P;
  const size_t nm = (size_t)P.nx * P.nz, nsh = r.shots.size();
  size_t o = 0;
  const size_t oc = o; o += a256(400);
  const size_t ow = o; o += a256((size_t)P.nt * 4);
  r.taper_off.resize(nsh);
  const size_t ot = o;
  for (size_t k = 0; k < nsh; k++) { r.taper_off[k] = o - ot; o += a256(r.setup[k].taper.size() * 4); }
  const size_t os = o; o += a256(nsh * sizeof(StackShotHost));
  const size_t oi = o; o += 3 * a256(nm * 8);
  const size_t of = o; o += 5 * a256(nm * 4);
  std::vector<double> wt;
  if (P.ismth > 0) gauss_weights(P.ismth, wt, r.wt_norm);
  const size_t owt = o; o += a256(wt.size() * 8);
  r.dev_bytes = o;
  r.dev = (char*)pool_.get(o);
  if (!r.dev) return -3;
  r.d_coef = (float*)(r.dev + oc);
  r.d_wav = (float*)(r.dev + ow);
  r.d_taper = (float*)(r.dev + ot);
  r.d_stk = (StackShotHost*)(r.dev + os);
  r.d_i1 = (double*)(r.dev + oi);
  r.d_i2 = (double*)(r.dev + oi + a256(nm * 8));
  r.d_il = (double*)(r.dev + oi + 2 * a256(nm * 8));
  float* f = (float*)(r.dev + of);
  r.d_f1 = f; r.d_f2 = (float*)((char*)f + a256(nm * 4)); r.d_f3 = (float*)((char*)f + 2 * a256(nm * 4));
  r.d_lap = (float*)((char*)f + 3 * a256(nm * 4)); r.d_nrm = (float*)((char*)f + 4 * a256(nm * 4));
  std::vector<float> wav;
  ricker(wav, P.nt, P.fdom, P.dt);
  r.stage.push_back(r.coef);
  r.stage.push_back(wav);
  std::vector<float> tap((o - ot) / 4 + 1, 0.0f);
  for (size_t k = 0; k < nsh; k++)
    memcpy(tap.data() + r.taper_off[k] / 4, r.setup[k].taper.data(), r.setup[k].taper.size() * 4);
  r.stage.push_back(tap);
  hipError_t e;
  if ((e = hipMemcpyAsync(r.d_coef, r.stage[0].data(), 400, hipMemcpyHostToDevice, st_)) != hipSuccess ||
      (e = hipMemcpyAsync(r.d_wav, r.stage[1].data(), (size_t)P.nt * 4, hipMemcpyHostToDevice, st_)) != hipSuccess ||
      (e = hipMemcpyAsync(r.d_taper, r.stage[2].data(), (os - ot), hipMemcpyHostToDevice, st_)) != hipSuccess) {
    fail("request upload", e);
    return -10;
  }
  if (!wt.empty()) {                           // velsmooth1d weights (kept alive in stage)
    r.d_wt = (double*)(r.dev + owt);
    r.stage.push_back(std::vector<float>(wt.size() * 2));
    memcpy(r.stage.back().data(), wt.data(), wt.size() * 8);
    if ((e = hipMemcpyAsync(r.d_wt, r.stage.back().data(), wt.size() * 8, hipMemcpyHostToDevice, st_)) != hipSuccess) {
      fail("smoothing weights", e);
      return -10;
    }
  }
  (void)rtm_launch_zero_f64(r.d_i1, (long long)nm, st_);
  (void)rtm_launch_zero_f64(r.d_i2, (long long)nm, st_);
  (void)rtm_launch_zero_f64(r.d_il, (long long)nm, st_);
  return 0;
}

Admit Server::admit_one(SrvRequest& r, std::string& why)
{
  int sid = -1;
  for (int i = 0; i < (int)slots_.size(); i++)
    if (slots_[i].phase == PH_IDLE && !slots_[i].mem) { sid = i; break; }
  if (sid < 0) return Admit::kWait;
  const int k = r.next_admit;
  const ShotSetup& s = r.setup[k];
  const size_t need = Engine::slot_bytes(s, r.P.nt, opt_.store_fmt);
  const size_t extra = r.dev ? 0 : request_bytes(r);
  // pending (finished, not yet stacked) buffers of a request are bounded by
  // its resident shots; count them with the slot
  const size_t pend = 2 * a256((size_t)r.P.nz * s.g.nxl * 4);
  bool releasing = false;
  for (auto& kv : reqs_) releasing |= (kv.second->finalizing && !kv.second->complete);
  const Admit a = admit_rule(budget_, used_, active_, releasing, (double)(need + pend), (double)extra);
  if (a != Admit::kAdmit) {
    if (a == Admit::kReject) why = "the shot does not fit the HBM budget";
    return a;
  }
  if (!r.dev) {
    if (request_device(r) != 0) {
      if (err_) return Admit::kWait;
      why = "the device refused the request block (" + std::to_string(request_bytes(r)) + " bytes)";
      return alloc_failed_rule(active_, releasing);
    }
    used_ += r.dev_bytes;
    r.t_start = now();
  }
  SrvSlot& sl = slots_[sid];
  sl.mem = pool_.get(need);
  sl.pend = sl.mem ? pool_.get(pend) : nullptr;
  if (!sl.pend) {
    if (sl.mem) { pool_.put(sl.mem); sl.mem = nullptr; }
    why = "the device refused a slot of " + std::to_string(need + pend) + " bytes";
    return alloc_failed_rule(active_, releasing);
  }
  sl.bytes = need + pend;
  used_ += sl.bytes;
  sl.req = r.id;
  sl.k = k;
  sl.phase = PH_FWD;
  sl.steps_left = r.P.nt;
  sl.uploads.clear();
  const Engine::Model* M = eng_->model(r.model_id);
  ShotDev& d = hds_[sid];
  Engine::layout_slot(r.P, s, opt_.store_fmt, (char*)sl.mem, d, sl.uploads, r.d_coef, r.d_wav, M->dev());
  d.record = 1;
  d.phase = PH_FWD;
  d.it_off = (int)(-tick_);
  d.it_dir = 1;
  hipError_t e;
  for (auto& u : sl.uploads)
    if ((e = hipMemcpyAsync(u.first, u.second.data(), u.second.size(), hipMemcpyHostToDevice, st_)) != hipSuccess)
      fail("slot upload", e);
  if ((e = hipMemcpyAsync(d_ds_ + sid, &hds_[sid], sizeof(ShotDev), hipMemcpyHostToDevice, st_)) != hipSuccess)
    fail("descriptor upload", e);
  (void)rtm_launch_phase_init(d_ds_, sid, PH_FWD, 256, st_);
  if (!s.wvv.empty())                          // the factors of the shot's (smoothed) windows
    (void)rtm_launch_model_prep(d.rho, d.vv, (float*)d.mib, (float*)d.mkk, (long long)s.g.nxp * s.g.nzp, st_);
  r.next_admit++;
  r.resident++;
  active_++;
  peak_ = active_ > peak_ ? active_ : peak_;
  dirty_ = true;
  return Admit::kAdmit;
}

// a request that cannot run: no more of its shots are admitted; once none is
// resident it completes with `status` (its device block released)
void Server::fail_request(SrvRequest& r, int status, const std::string& why)
{
  if (r.status == 0) { r.status = status; r.err = why; }
  r.next_admit = (int)r.shots.size();
  if (r.resident == 0 && !r.finalizing) complete_request(r);
}

void Server::complete_request(SrvRequest& r)
{
  r.complete = true;
  r.t_done = now();
  r.stage.clear();
  if (r.dev) { pool_.put(r.dev); used_ -= r.dev_bytes; r.dev = nullptr; }
  for (size_t k = 0; k < r.pend_g1.size(); k++)
    if (r.pend_g1[k]) {
      pool_.put(r.pend_g1[k]);
      used_ -= 2 * a256((size_t)r.P.nz * r.setup[k].g.nxl * 4);
      r.pend_g1[k] = r.pend_il[k] = nullptr;
    }
  if (r.status == 0) {
    cells_done_ += r.cells;
    for (size_t k = 0; k < r.guard.size(); k++)
      if (r.guard[k] >= 0x7F800000u) {
        r.status = -7;
        r.err = "the wavefield of shot " + std::to_string(r.shots[k]) +
                " is not finite (blow-up: dt above the stability bound for the model, or non-finite model values)";
        break;
      }
  }
  cv_done_.notify_all();
}

void Server::rebuild_tiles()
{
  std::vector<ShotDev> fw, rv;
  std::vector<int> fid, rid;
  qslots_.clear();
  aslots_.clear();
  for (int i = 0; i < (int)slots_.size(); i++) {
    if (slots_[i].phase != PH_IDLE) aslots_.push_back(i);
    if (slots_[i].phase == PH_FWD) {
      fw.push_back(hds_[i]); fid.push_back(i);
      qslots_.push_back(i);                           // strips saved after each forward step
    } else if (slots_[i].phase == PH_REV) {
      rv.push_back(hds_[i]); rid.push_back(i);
    }
  }
  build_tiles(fw, opt_.tile_z, tf_, &fid);
  build_tiles(rv, opt_.tile_z, tr_, &rid);
  layout_runs(tf_, PH_FWD);
  layout_runs(tr_, PH_REV);
  const size_t need = tf_.size() > tr_.size() ? tf_.size() : tr_.size();
  if (need > tile_cap_) {
    (void)hipStreamSynchronize(st_);
    (void)hipFree(d_tf_); (void)hipFree(d_tr_);
    tile_cap_ = need * 2 + 1024;
    if (hipMalloc(&d_tf_, tile_cap_ * sizeof(Tile)) != hipSuccess ||
        hipMalloc(&d_tr_, tile_cap_ * sizeof(Tile)) != hipSuccess) {
      (void)hipGetLastError();                       // reported through err_, not by a launch
      fail_with(-3, "out of device memory (tiles)");
      return;
    }
  }
  // the previous tables may still be read by launches in flight: the copies
  // are stream-ordered behind them
  hipError_t e;
  if (!tf_.empty() && (e = hipMemcpyAsync(d_tf_, tf_.data(), tf_.size() * sizeof(Tile), hipMemcpyHostToDevice, st_)) != hipSuccess)
    fail("tile upload", e);
  if (!tr_.empty() && (e = hipMemcpyAsync(d_tr_, tr_.data(), tr_.size() * sizeof(Tile), hipMemcpyHostToDevice, st_)) != hipSuccess)
    fail("tile upload", e);
  if (!qslots_.empty() && (e = hipMemcpyAsync(d_qs_, qslots_.data(), qslots_.size() * 4, hipMemcpyHostToDevice, st_)) != hipSuccess)
    fail("slot list upload", e);
  if (!aslots_.empty() && (e = hipMemcpyAsync(d_as_, aslots_.data(), aslots_.size() * 4, hipMemcpyHostToDevice, st_)) != hipSuccess)
    fail("slot list upload", e);
  (void)hipStreamSynchronize(st_);   // host vectors are reused next rebuild
  ntf_ = (int)tf_.size();
  ntr_ = (int)tr_.size();
  nq_ = (int)qslots_.size();
  na_ = (int)aslots_.size();
  dirty_ = false;
}

void Server::transition(int sid)
{
  SrvSlot& sl = slots_[sid];
  SrvRequest& r = *reqs_[sl.req];
  if (sl.phase == PH_FWD) {
    ShotDev& d = hds_[sid];
    if (r.P.ismth > 0) {
      // velsmooth1d of the shot's windows before its reverse pass
      // (rtm_model.cu:402-405), through the factor arrays as scratch
      const ShotSetup& s = r.setup[sl.k];
      const int nzw = s.g.nzp, nxw = s.g.nxp;
      const size_t wb = (size_t)nzw * nxw * 4;
      float *vv = (float*)d.vv, *rh = (float*)d.rho, *t1 = (float*)d.mib, *t2 = (float*)d.mkk;
      (void)rtm_launch_gauss_smooth(vv, t1, nzw, nxw, r.P.ismth, r.d_wt, r.wt_norm, st_);
      (void)rtm_launch_gauss_smooth(rh, t2, nzw, nxw, r.P.ismth, r.d_wt, r.wt_norm, st_);
      hipError_t e1;
      if ((e1 = hipMemcpyAsync(vv, t1, wb, hipMemcpyDeviceToDevice, st_)) != hipSuccess ||
          (e1 = hipMemcpyAsync(rh, t2, wb, hipMemcpyDeviceToDevice, st_)) != hipSuccess)
        fail("smoothing copy", e1);
      (void)rtm_launch_model_prep(rh, vv, t1, t2, (long long)nzw * nxw, st_);
    }
    d.phase = PH_REV;
    d.it_off = (int)(r.P.nt - 2 + tick_);      // it = nt-2 - (tick - tick_start)
    d.it_dir = -1;
    hipError_t e;
    if ((e = hipMemcpyAsync(d_ds_ + sid, &hds_[sid], sizeof(ShotDev), hipMemcpyHostToDevice, st_)) != hipSuccess)
      fail("descriptor upload", e);
    (void)rtm_launch_phase_init(d_ds_, sid, PH_REV, 256, st_);
    (void)rtm_launch_seed(d_ds_, sid, d.nrec, st_);
    sl.phase = PH_REV;
    sl.steps_left = r.P.nt - 1;
    dirty_ = true;
  } else {
    finish_shot(sid);
  }
}

void Server::finish_shot(int sid)
{
  SrvSlot& sl = slots_[sid];
  SrvRequest& r = *reqs_[sl.req];
  const int k = sl.k;
  const size_t n = (size_t)r.P.nz * r.setup[k].g.nxl * 4;
  // the guard after the shot's last step (tick_ - 1), then its word to the host
  (void)rtm_launch_guard(d_ds_, d_id_ + sid, 1, (int)(tick_ - 1), nullptr, st_);
  (void)hipMemcpyAsync(&r.guard[k], rtm_guard_word(hds_[sid]), 4, hipMemcpyDeviceToHost, st_);
  r.resident--;
  float* g = (float*)sl.pend;                 // taken with the slot
  r.pend_g1[k] = g;
  r.pend_il[k] = (float*)((char*)g + a256(n));
  hipError_t e;
  if ((e = hipMemcpyAsync(r.pend_g1[k], hds_[sid].g1, n, hipMemcpyDeviceToDevice, st_)) != hipSuccess ||
      (e = hipMemcpyAsync(r.pend_il[k], hds_[sid].illum, n, hipMemcpyDeviceToDevice, st_)) != hipSuccess)
    fail("pending copy", e);
  r.finished[k] = 1;
  pool_.put(sl.mem);
  used_ -= sl.bytes;
  used_ += 2 * a256(n);
  sl = SrvSlot();
  hds_[sid].phase = PH_IDLE;
  active_--;
  dirty_ = true;
  if (r.status != 0) {                         // a failed request: drop its results
    if (r.resident == 0 && !r.finalizing) complete_request(r);
    return;
  }
  stack_ready(r);
}

void Server::stack_ready(SrvRequest& r)
{
  const int nsh = (int)r.shots.size();
  while (r.next_stack < nsh && r.finished[r.next_stack]) {
    const int k = r.next_stack;
    StackShotHost h{r.pend_g1[k], r.pend_il[k], r.d_taper + r.taper_off[k] / 4, r.setup[k].g.nx1, r.setup[k].g.nxl};
    hipError_t e;
    r.stage.push_back(std::vector<float>(sizeof(StackShotHost) / 4 + 1));
    memcpy(r.stage.back().data(), &h, sizeof(h));
    if ((e = hipMemcpyAsync(r.d_stk + k, r.stage.back().data(), sizeof(h), hipMemcpyHostToDevice, st_)) != hipSuccess)
      fail("stack entry", e);
    (void)rtm_launch_stack(r.d_stk + k, 1, r.P.nx, r.P.nz, r.d_i1, r.d_i2, r.d_il, st_);
    const size_t n = (size_t)r.P.nz * r.setup[k].g.nxl * 4;
    pool_.put(r.pend_g1[k]);          // stream-ordered behind the stacking kernel
    used_ -= 2 * a256(n);
    r.pend_g1[k] = r.pend_il[k] = nullptr;
    r.next_stack++;
  }
  if (r.next_stack == nsh && !r.finalizing) finalize(r);
}

void Server::finalize(SrvRequest& r)
{
  const rtmh_params& P = r.P;
  const size_t nm = (size_t)P.nx * P.nz;
  (void)rtm_launch_round(r.d_i1, r.d_f1, (long long)nm, st_);
  (void)rtm_launch_round(r.d_i2, r.d_f2, (long long)nm, st_);
  (void)rtm_launch_round(r.d_il, r.d_f3, (long long)nm, st_);
  (void)rtm_launch_laplace(r.d_f1, r.d_lap, P.nx, P.nz, (int)P.dx, (int)P.dz, st_);
  (void)rtm_launch_norm_multi(r.d_f1, r.d_f3, r.d_nrm, (long long)nm, st_);
  r.image1.resize(nm); r.image2s.resize(nm); r.illum.resize(nm); r.image_lap.resize(nm); r.image_norm.resize(nm);
  float* dst[5] = {r.image1.data(), r.image2s.data(), r.illum.data(), r.image_lap.data(), r.image_norm.data()};
  float* src[5] = {r.d_f1, r.d_f2, r.d_f3, r.d_lap, r.d_nrm};
  hipError_t e;
  for (int i = 0; i < 5; i++)
    if ((e = hipMemcpyAsync(dst[i], src[i], nm * 4, hipMemcpyDeviceToHost, st_)) != hipSuccess) fail("result copy", e);
  (void)hipEventCreateWithFlags(&r.done_ev, hipEventDisableTiming);
  (void)hipEventRecord(r.done_ev, st_);
  r.finalizing = true;
}

void Server::loop()
{
  (void)hipSetDevice(eng_->device());
  std::deque<std::pair<long long, hipEvent_t>> inflight;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      // admission: first come first served over shots
      while (!queue_.empty() && !err_) {
        SrvRequest& r = *reqs_[queue_.front()];
        if (r.next_admit >= (int)r.shots.size()) { queue_.pop_front(); continue; }
        std::string why;
        const Admit a = admit_one(r, why);
        if (a == Admit::kWait) break;
        if (a == Admit::kReject) {
          fail_request(r, -3, why);
          queue_.pop_front();
        }
      }
      // close requests whose results have landed
      for (auto& kv : reqs_) {
        SrvRequest& r = *kv.second;
        if (r.finalizing && !r.complete && hipEventQuery(r.done_ev) == hipSuccess) complete_request(r);
      }
      bool pending_final = false;
      for (auto& kv : reqs_) pending_final |= (kv.second->finalizing && !kv.second->complete);
      if (stop_) break;
      if (active_ == 0) {
        if (pending_final) {
          lk.unlock();
          std::this_thread::sleep_for(std::chrono::microseconds(200));
          continue;
        }
        cv_.wait_for(lk, std::chrono::milliseconds(20));
        continue;
      }
    }
    const double tb = now();
    if (dirty_) rebuild_tiles();
    if (err_) { cv_done_.notify_all(); std::this_thread::sleep_for(std::chrono::milliseconds(5)); continue; }
    if (ntf_ > 0) {
      LaunchArgs a{d_ds_, d_tf_, ntf_, (int)tick_};
      if (rtm_launch_step(a, PH_FWD, opt_.tile_z, st_) != hipSuccess) fail("forward launch", hipGetLastError());
      if (nq_ > 0) (void)rtm_launch_save(d_ds_, d_qs_, nq_, (int)tick_, opt_.store_fmt, st_);
      launches_ += 2;
    }
    if (ntr_ > 0) {
      LaunchArgs a{d_ds_, d_tr_, ntr_, (int)tick_};
      if (rtm_launch_step(a, PH_REV, opt_.tile_z, st_) != hipSuccess) fail("reverse launch", hipGetLastError());
      launches_ += 2;
    }
    // wavefield guard of every resident shot (the shots' own ntsnap differ;
    // one interval for all, and once more after each shot's last step)
    if (na_ > 0 && (tick_ & (kGuardTicks - 1)) == kGuardTicks - 1)
      (void)rtm_launch_guard(d_ds_, d_as_, na_, (int)tick_, nullptr, st_);
    tick_++;
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (int i = 0; i < (int)slots_.size(); i++) {
        if (slots_[i].phase == PH_IDLE) continue;
        if (--slots_[i].steps_left == 0) transition(i);
      }
    }
    // bound how far the host runs ahead of the GPU
    if ((tick_ & 7) == 0) {
      hipEvent_t ev;
      (void)hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      (void)hipEventRecord(ev, st_);
      inflight.push_back({tick_.load(), ev});
      while (inflight.size() > 4) {
        (void)hipEventSynchronize(inflight.front().second);
        (void)hipEventDestroy(inflight.front().second);
        inflight.pop_front();
      }
    }
    busy_.store(busy_.load() + (now() - tb));
  }
  for (auto& p : inflight) { (void)hipEventSynchronize(p.second); (void)hipEventDestroy(p.second); }
}

int Server::wait(int id, double timeout, rtmh_outputs* out, rtmh_request_stats* st)
{
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) { set_error("no such request"); return -1; }
  SrvRequest& r = *it->second;
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout > 0 ? timeout : 1e9);
  while (!r.complete && !err_) {
    if (cv_done_.wait_until(lk, until) == std::cv_status::timeout && !r.complete) return 1;
  }
  if (const int e = err_.load(); e && !r.complete) { set_error(err_message()); return e; }
  if (r.status != 0) {
    if (st) { st->submit_s = r.t_submit; st->start_s = r.t_start; st->done_s = r.t_done; st->status = r.status; }
    set_error(r.err);
    return r.status;
  }
  if (out) {
    const size_t nm = (size_t)r.P.nx * r.P.nz * 4;
    if (out->image1) memcpy(out->image1, r.image1.data(), nm);
    if (out->image2s) memcpy(out->image2s, r.image2s.data(), nm);
    if (out->illum) memcpy(out->illum, r.illum.data(), nm);
    if (out->image_lap) memcpy(out->image_lap, r.image_lap.data(), nm);
    if (out->image_norm) memcpy(out->image_norm, r.image_norm.data(), nm);
  }
  if (st) {
    st->submit_s = r.t_submit; st->start_s = r.t_start; st->done_s = r.t_done;
    st->cell_updates = r.cells; st->nshots = (int)r.shots.size(); st->status = r.status;
  }
  return 0;
}

void Server::query(rtmh_server_stats* st)
{
  std::lock_guard<std::mutex> lk(mu_);
  st->ticks = tick_; st->launches = launches_; st->peak_slots = peak_; st->active_slots = active_;
  st->queued_requests = (int)queue_.size(); st->busy_s = busy_.load(); st->cell_updates = cells_done_;
  st->budget_bytes = budget_;
  st->allocations = pool_.allocations();
  st->alloc_s = pool_.alloc_seconds();
}

}  // namespace rtm

// ------------------------------------------------------------------ C ABI --
struct rtmh_server {
  std::unique_ptr<rtm::Server> s;
};

// the engine of a context (rtm_capi.cpp)
rtm::Engine* rtmh_ctx_engine(rtmh_ctx* ctx);

extern "C" int rtmh_server_create(rtmh_ctx* ctx, const rtmh_server_options* opt, rtmh_server** out)
{
  if (!ctx || !out) { rtm::set_error("null argument"); return -1; }
  rtmh_server_options o;
  memset(&o, 0, sizeof(o));
  if (opt) o = *opt;
  if (o.store_fmt < 0 || o.store_fmt > 2) { rtm::set_error("store_fmt must be 0, 1 or 2"); return -1; }
  if (o.tile_z != 0 && o.tile_z != 16 && o.tile_z != 32 && o.tile_z != 64) {
    rtm::set_error("tile_z must be 0 (32), 16, 32 or 64");
    return -1;
  }
  auto* h = new rtmh_server;
  h->s.reset(new rtm::Server(rtmh_ctx_engine(ctx), o));
  const int rc = h->s->start();
  if (rc) { delete h; return rc; }
  *out = h;
  return 0;
}

extern "C" int rtmh_server_submit(rtmh_server* srv, int model_id, const rtmh_params* p, const float* diffcoef,
                                  int* request_id)
{
  if (!srv || !p || !diffcoef || !request_id) { rtm::set_error("null argument"); return -1; }
  return srv->s->submit(model_id, *p, diffcoef, request_id);
}

extern "C" int rtmh_server_request_bytes(rtmh_server* srv, int model_id, const rtmh_params* p, const float* diffcoef,
                                         double* shot_max, double* request)
{
  if (!srv || !p || !diffcoef || !shot_max || !request) { rtm::set_error("null argument"); return -1; }
  return srv->s->request_bytes(model_id, *p, diffcoef, shot_max, request);
}

extern "C" int rtmh_server_wait(rtmh_server* srv, int request_id, double timeout_s, rtmh_outputs* out,
                                rtmh_request_stats* stats)
{
  if (!srv) { rtm::set_error("null argument"); return -1; }
  return srv->s->wait(request_id, timeout_s, out, stats);
}

extern "C" int rtmh_server_query(rtmh_server* srv, rtmh_server_stats* st)
{
  if (!srv || !st) { rtm::set_error("null argument"); return -1; }
  srv->s->query(st);
  return 0;
}

extern "C" int rtmh_server_destroy(rtmh_server* srv)
{
  delete srv;
  return 0;
}
