// host_logic.cpp -- CPU checks of the host logic of librtm_hip.so that needs
// no GPU: the served form's admission rules (driven here by a fake device
// allocator), the first-fit arena its slots are carved from, the per-XCD runs
// of the tile table and the neighbour links of the zero-tile skip, and the
// checkpoint resume decision.  Built and run by
// tests/test_host_logic.py; prints one line per check and exits non-zero on
// the first failure.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <deque>
#include <string>
#include <vector>

#include "admission.h"
#include "rtm_engine.h"

using rtm::Admit;

static int g_fail = 0;
#define CHECK(c, ...)                                                  \
  do {                                                                 \
    if (!(c)) {                                                        \
      printf("FAIL %s:%d: %s -- ", __FILE__, __LINE__, #c);            \
      printf(__VA_ARGS__);                                             \
      printf("\n");                                                    \
      g_fail = 1;                                                      \
    }                                                                  \
  } while (0)

// ---- a served-form admission loop over a fake device -------------------------
// Mirrors the server's loop: shots are admitted first come first served while
// admit_rule allows and the fake device grants the memory; resident shots
// finish after 10 ticks and release their bytes; a request whose shot is
// rejected fails at once.
struct FakeDevice {
  double capacity, held = 0;
  bool alloc(double b) { if (held + b > capacity) return false; held += b; return true; }
  void release(double b) { held -= b; }
};
struct Req {
  std::vector<double> need;  // per shot
  double extra;              // request block
  int next = 0, resident = 0, done = 0;
  bool failed = false, has_block = false;
  std::vector<long long> admitted_at;
};
struct Sim {
  double budget, used = 0;
  FakeDevice dev;
  std::deque<int> queue;
  std::vector<Req> reqs;
  struct Res { int r, k; long long end; double bytes; };
  std::vector<Res> res;
  int peak = 0;
  long long tick = 0;
  Sim(double b, double cap) : budget(b), dev{cap} {}
  int active() const { return (int)res.size(); }
  void admission()
  {
    while (!queue.empty()) {
      Req& r = reqs[queue.front()];
      if (r.failed || r.next >= (int)r.need.size()) { queue.pop_front(); continue; }
      const double extra = r.has_block ? 0.0 : r.extra;
      Admit a = rtm::admit_rule(budget, used, active(), false, r.need[r.next], extra);
      if (a == Admit::kAdmit) {
        if (!r.has_block) {
          if (!dev.alloc(r.extra)) a = rtm::alloc_failed_rule(active(), false);
          else { r.has_block = true; used += r.extra; }
        }
        if (a == Admit::kAdmit && !dev.alloc(r.need[r.next])) a = rtm::alloc_failed_rule(active(), false);
      }
      if (a == Admit::kWait) break;
      if (a == Admit::kReject) {                 // the request fails; its block is released
        if (r.has_block) { used -= r.extra; dev.release(r.extra); r.has_block = false; }
        r.failed = true;
        queue.pop_front();
        continue;
      }
      used += r.need[r.next];
      res.push_back({queue.front(), r.next, tick + 10, r.need[r.next]});
      r.admitted_at.push_back(tick);
      r.next++;
      r.resident++;
      peak = active() > peak ? active() : peak;
    }
  }
  // runs until every request is done or failed; false on a stall
  bool run(long long max_ticks = 100000)
  {
    for (int i = 0; i < (int)reqs.size(); i++) queue.push_back(i);
    for (tick = 0; tick < max_ticks; tick++) {
      admission();
      for (size_t i = 0; i < res.size();) {
        if (res[i].end <= tick) {
          Req& r = reqs[res[i].r];
          used -= res[i].bytes;
          dev.release(res[i].bytes);
          r.resident--;
          r.done++;
          if (r.done == (int)r.need.size()) { used -= r.extra; dev.release(r.extra); }
          res.erase(res.begin() + i);
        } else {
          i++;
        }
      }
      bool all = true;
      for (auto& r : reqs) all = all && (r.failed || r.done == (int)r.need.size());
      if (all && res.empty()) return true;
    }
    return false;
  }
};

static void admission_checks()
{
  // a shot that cannot fit even alone: refused before it is queued
  CHECK(rtm::admit_rule(100.0, 0.0, 0, false, 90.0, 20.0) == Admit::kReject, "90+20 > 100");
  CHECK(rtm::admit_rule(100.0, 0.0, 0, false, 80.0, 20.0) == Admit::kAdmit, "80+20 <= 100");
  // does not fit now: waits while something resident or releasing will free memory
  CHECK(rtm::admit_rule(100.0, 50.0, 1, false, 60.0, 0.0) == Admit::kWait, "resident");
  CHECK(rtm::admit_rule(100.0, 50.0, 0, true, 60.0, 0.0) == Admit::kWait, "releasing");
  CHECK(rtm::admit_rule(100.0, 50.0, 0, false, 60.0, 0.0) == Admit::kReject, "nothing will free it");
  CHECK(rtm::alloc_failed_rule(0, false) == Admit::kReject, "device full, nothing resident");
  CHECK(rtm::alloc_failed_rule(2, false) == Admit::kWait, "device full, shots resident");
  printf("ok admission rules\n");

  {  // a budget that holds one shot (and the request block) at a time
    Sim s(130.0, 1e9);
    s.reqs.push_back({{100, 100, 100}, 25});
    s.reqs.push_back({{100, 100}, 25});
    CHECK(s.run(), "stalled");
    CHECK(s.peak == 1, "peak %d", s.peak);
    CHECK(!s.reqs[0].failed && !s.reqs[1].failed, "failed");
    CHECK(s.reqs[0].admitted_at.size() == 3 && s.reqs[1].admitted_at.size() == 2, "admitted");
    CHECK(s.reqs[0].admitted_at.back() < s.reqs[1].admitted_at.front(), "first come first served");
    printf("ok one shot at a time, in order\n");
  }
  {  // the budget holds several shots: they run together
    Sim s(1000.0, 1e9);
    s.reqs.push_back({{100, 100, 100}, 25});
    s.reqs.push_back({{100, 100}, 25});
    CHECK(s.run(), "stalled");
    CHECK(s.peak == 5, "peak %d", s.peak);
    printf("ok shots of several requests resident together\n");
  }
  {  // the budget claims more than the device holds: the first shot that the
     // device refuses with nothing resident fails its request; later requests
     // that fit still run
    Sim s(1000.0, 150.0);
    s.reqs.push_back({{200}, 25});      // never fits the device
    s.reqs.push_back({{100, 100}, 25}); // fits one shot at a time
    CHECK(s.run(), "stalled");
    CHECK(s.reqs[0].failed, "the oversized request must fail");
    CHECK(!s.reqs[1].failed && s.reqs[1].done == 2, "the fitting request must complete");
    CHECK(s.peak == 1, "peak %d", s.peak);
    printf("ok device refusal fails the request instead of queueing it\n");
  }
  {  // the device refuses while shots are resident: wait, then run
    Sim s(1000.0, 260.0);
    s.reqs.push_back({{100, 100}, 25});
    s.reqs.push_back({{100}, 25});
    CHECK(s.run(), "stalled");
    CHECK(!s.reqs[0].failed && !s.reqs[1].failed, "failed");
    CHECK(s.reqs[1].admitted_at.front() >= 10, "second request waits for memory");
    printf("ok device refusal with shots resident waits\n");
  }
}

// ---- checkpoint resume of a shrunken job --------------------------------------
static void checkpoint_checks(const char* dir)
{
  const std::string path = std::string(dir) + "/ckpt.bin";
  const int nx = 7, nz = 5, nstk = 3;
  std::vector<double> st((size_t)nx * nz * nstk);
  for (size_t i = 0; i < st.size(); i++) st[i] = 0.5 * (double)i;
  const uint64_t fp = 0x1234abcdULL;
  CHECK(rtm::Engine::write_checkpoint(path.c_str(), fp, nx, nz, 5, st) == 0, "write");
  int done = -1;
  std::vector<double> back;
  std::string why;
  int rc = rtm::Engine::resume_point(path.c_str(), fp, nx, nz, nstk, 5, done, back, why);
  CHECK(rc == 1 && done == 5 && back == st, "same job rc %d done %d", rc, done);
  rc = rtm::Engine::resume_point(path.c_str(), fp, nx, nz, nstk, 8, done, back, why);
  CHECK(rc == 1 && done == 5, "longer job rc %d done %d", rc, done);
  why.clear();
  rc = rtm::Engine::resume_point(path.c_str(), fp, nx, nz, nstk, 3, done, back, why);
  CHECK(rc == -5 && why.find("more than this job") != std::string::npos, "shrunken job rc %d (%s)", rc, why.c_str());
  rc = rtm::Engine::resume_point(path.c_str(), fp + 1, nx, nz, nstk, 5, done, back, why);
  CHECK(rc == -5, "another job's checkpoint rc %d", rc);
  rc = rtm::Engine::resume_point((path + ".none").c_str(), fp, nx, nz, nstk, 5, done, back, why);
  CHECK(rc == 0 && done == 0, "no file rc %d", rc);
  unlink(path.c_str());
  printf("ok checkpoint resume (same, longer, shrunken, other job, missing)\n");
}

// ---- the server's arena: first fit over one range, neighbours merged ---------
// Random gets and puts against a byte map of the range: blocks never overlap,
// stay inside the range and are 256-byte aligned; the free ranges and the
// blocks in use tile the range exactly with no two free ranges adjacent; the
// lowest-offset hole that fits is the one taken; once everything is returned
// the range is one free range again.
static void extents_checks()
{
  rtm::Extents e;
  const size_t R = 256 * 1000;
  e.reset(R);
  std::vector<int> owner(R / 256, -1);
  std::vector<std::pair<size_t, size_t>> live;   // offset, units
  unsigned rng = 12345;
  auto rnd = [&](unsigned m) { rng = rng * 1103515245u + 12345u; return (rng >> 8) % m; };
  int bad = 0, got = 0, refused = 0;
  for (int it = 0; it < 20000 && !bad; it++) {
    if (live.empty() || rnd(100) < 55) {
      const size_t bytes = 1 + rnd(256 * 40);
      const size_t units = (bytes + 255) / 256;
      // the first hole (lowest offset) of at least `units` in the byte map
      size_t expect = rtm::Extents::kNone, run = 0;
      for (size_t u = 0; u < owner.size(); u++) {
        run = owner[u] < 0 ? run + 1 : 0;
        if (run == units) { expect = (u + 1 - units) * 256; break; }
      }
      const size_t off = e.get(bytes);
      if (off != expect) { bad = 1; CHECK(off == expect, "get(%zu) gave %zu, first fit is %zu", bytes, off, expect); }
      if (off == rtm::Extents::kNone) { refused++; continue; }
      got++;
      for (size_t u = off / 256; u < off / 256 + units; u++) owner[u] = it;
      live.push_back({off, units});
    } else {
      const size_t i = rnd((unsigned)live.size());
      CHECK(e.put(live[i].first), "put of a live block at %zu refused", live[i].first);
      for (size_t u = live[i].first / 256; u < live[i].first / 256 + live[i].second; u++) owner[u] = -1;
      live[i] = live.back();
      live.pop_back();
    }
    // the maps tile [0, R): sorted, gap-free, holes never adjacent
    size_t pos = 0;
    auto h = e.holes().begin();
    auto b = e.blocks().begin();
    bool prev_hole = false;
    while (pos < R && !bad) {
      if (h != e.holes().end() && h->first == pos) {
        if (prev_hole) { bad = 1; CHECK(!prev_hole, "two adjacent free ranges at %zu", pos); }
        for (size_t u = pos / 256; u < (pos + h->second) / 256; u++)
          if (owner[u] >= 0) { bad = 1; CHECK(owner[u] < 0, "free range over a live block at unit %zu", u); break; }
        pos += h->second; ++h; prev_hole = true;
      } else if (b != e.blocks().end() && b->first == pos) {
        pos += b->second; ++b; prev_hole = false;
      } else {
        bad = 1;
        CHECK(false, "offset %zu is neither free nor in use", pos);
      }
    }
    if (!bad && (pos != R || h != e.holes().end() || b != e.blocks().end())) {
      bad = 1;
      CHECK(false, "the maps end at %zu of %zu", pos, R);
    }
  }
  CHECK(got > 1000 && refused > 10, "the sequence exercised too little (%d blocks, %d refusals)", got, refused);
  CHECK(!e.put(1), "put of an offset that starts no block was accepted");
  for (auto& l : live) CHECK(e.put(l.first), "put at %zu refused", l.first);
  CHECK(e.holes().size() == 1 && e.holes().begin()->first == 0 && e.holes().begin()->second == R && e.blocks().empty(),
        "after every block came back: %zu free ranges, %zu blocks", e.holes().size(), e.blocks().size());
  CHECK(e.get(R) == 0 && e.get(1) == rtm::Extents::kNone, "the whole range as one block");
  e.reset(1024);
  const size_t z0 = e.get(0), z1 = e.get(0);
  CHECK(z0 != z1 && z0 != rtm::Extents::kNone && z1 != rtm::Extents::kNone, "zero-byte blocks share offset %zu", z0);
  // sizes that would wrap when rounded up to a unit are refused, not carved as empty blocks
  const size_t big = rtm::Extents::kNone - 100;
  CHECK(e.get(rtm::Extents::kNone) == rtm::Extents::kNone && e.get(big) == rtm::Extents::kNone,
        "a size near the top of size_t was carved");
  CHECK(e.blocks().size() == 2, "%zu blocks after the refusals", e.blocks().size());
  printf("extents: %d blocks, %d refusals, first fit and merging hold\n", got, refused);
}

// ---- neighbour links of the zero-tile skip ----------------------------------
// Two slots of different windows through build_tiles, the per-XCD layout and
// link_neighbours: every link of a tile points at the tile one step away in
// the same slot, and every such tile is linked; padding entries link nothing.
static void neighbour_checks()
{
  std::vector<ShotDev> ds(2);
  memset(ds.data(), 0, ds.size() * sizeof(ShotDev));
  ds[0].npml = 40; ds[0].nxp = 300; ds[0].nzp = 170;
  ds[1].npml = 40; ds[1].nxp = 701; ds[1].nzp = 150;
  for (int tz : {16, 32}) {
    std::vector<Tile> tiles;
    rtm::build_tiles(ds, tz, tiles);
    rtm::layout_runs(tiles, PH_REV);
    rtm::link_neighbours(tiles, tz);
    int links = 0, bad = 0;
    for (size_t i = 0; i < tiles.size(); i++) {
      const Tile& t = tiles[i];
      int n = 0;
      for (int dz = -1; dz <= 1; dz++)
        for (int dx = -1; dx <= 1; dx++) {
          if (dz == 0 && dx == 0) continue;
          const int k = t.nbr[n++];
          int want = -1;
          if (t.slot >= 0)
            for (size_t j = 0; j < tiles.size(); j++)
              if (tiles[j].slot == t.slot && tiles[j].z0 == t.z0 + dz * tz && tiles[j].x0 == t.x0 + dx * 64) want = (int)j;
          if (k != want) bad++;
          links += k >= 0;
        }
    }
    CHECK(bad == 0 && links > 0, "tz %d: %d wrong links of %d", tz, bad, links);
  }
  // the skip only where the adjacent tiles hold the step's 19-row reach
  CHECK(!rtm::fwd_skip_ok(16) && rtm::fwd_skip_ok(32) && rtm::fwd_skip_ok(64), "skip gate by tile height");
  printf("ok neighbour links of the zero-tile skip\n");
}

// ---- per-XCD runs of the tile table -----------------------------------------
// layout_runs on a ragged batch: the table becomes 8 runs of one length (the
// kernels deal run c to XCD c), every tile of the input exactly once, empty
// entries only at the ends of runs, each run's long tiles (reverse: holding
// reconstructed cells; forward: in the absorbing layer) before its short
// ones, each run a contiguous piece of the input (neighbours stay on one
// XCD), and -- when cut by cost -- run costs within one long tile of the
// even share.
static void layout_checks()
{
  // the engine sizes its tile table for n * 5 / 4 + 16 entries: the padded
  // table stays within that however the long tiles are spread (all first,
  // all last, alternating, none), for every table length up to 300
  for (int phase : {PH_FWD, PH_REV}) {
    const int flag = phase == PH_REV ? TF_INT : TF_GEN;
    for (int n = 1; n <= 300; n++)
      for (int pat = 0; pat < 4; pat++) {
        std::vector<Tile> t;
        for (int i = 0; i < n; i++) {
          const bool heavy = pat == 0 ? i < n / 2 : pat == 1 ? i >= n / 2 : pat == 2 ? (i & 1) : false;
          t.push_back(Tile{i % 3, 64 * i, 0, heavy ? flag : 0});
        }
        rtm::layout_runs(t, phase);
        size_t real = 0;
        for (const Tile& x : t) real += x.slot >= 0;
        CHECK(t.size() % 8 == 0 && t.size() <= (size_t)n * 5 / 4 + 16 && real == (size_t)n,
              "phase %d n %d pattern %d: %zu entries", phase, n, pat, t.size());
      }
  }
  std::vector<ShotDev> ds(5);
  memset(ds.data(), 0, ds.size() * sizeof(ShotDev));
  const int w[5] = {300, 701, 450, 1300, 520}, h[5] = {170, 150, 260, 210, 180};
  for (int k = 0; k < 5; k++) { ds[k].npml = 40; ds[k].nxp = w[k]; ds[k].nzp = h[k]; }
  for (int phase : {PH_FWD, PH_REV}) {
    std::vector<Tile> in;
    rtm::build_tiles(ds, 32, in);
    std::vector<Tile> t = in;
    rtm::layout_runs(t, phase);
    const int flag = phase == PH_REV ? TF_INT : TF_GEN;
    const double heavy = phase == PH_REV ? 2.0 : 1.55;
    CHECK(t.size() % 8 == 0, "phase %d: table of %zu entries is not 8 runs", phase, t.size());
    CHECK(t.size() <= in.size() * 5 / 4 + 16, "phase %d: padded table of %zu for %zu tiles", phase, t.size(), in.size());
    const size_t len = t.size() / 8;
    // every input tile exactly once
    auto key = [](const Tile& x) { return ((long long)x.slot << 40) | ((long long)x.z0 << 20) | x.x0; };
    std::vector<long long> a, b;
    for (const Tile& x : in) a.push_back(key(x));
    for (const Tile& x : t) if (x.slot >= 0) b.push_back(key(x));
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    CHECK(a == b, "phase %d: tiles lost or repeated (%zu in, %zu out)", phase, a.size(), b.size());
    double total = 0;
    for (const Tile& x : in) total += (x.kind & flag) ? heavy : 1.0;
    double worst = 0;
    size_t next = 0;                                 // the runs take the input in order, piece by piece
    for (int c = 0; c < 8; c++) {
      bool pad = false, shortseen = false, ok = true;
      double cost = 0;
      std::vector<long long> run;
      for (size_t i = c * len; i < (c + 1) * len; i++) {
        const Tile& x = t[i];
        if (x.slot < 0) { pad = true; continue; }
        if (pad) ok = false;                         // a tile after padding
        const bool lng = (x.kind & flag) != 0;
        if (lng && shortseen) ok = false;            // a long tile after a short one
        shortseen |= !lng;
        cost += lng ? heavy : 1.0;
        run.push_back(key(x));
      }
      CHECK(ok, "phase %d run %d: order (long, then short, then padding) broken", phase, c);
      // the run is the next contiguous piece of the input, reordered
      std::vector<long long> piece;
      for (size_t i = next; i < next + run.size() && i < in.size(); i++) piece.push_back(key(in[i]));
      next += run.size();
      std::sort(run.begin(), run.end());
      std::sort(piece.begin(), piece.end());
      CHECK(run == piece, "phase %d run %d is not a contiguous piece of the input", phase, c);
      worst = std::max(worst, std::abs(cost - total / 8));
    }
    CHECK(worst <= heavy + 1e-9 || len == (in.size() + 7) / 8,
          "phase %d: a run's cost is %.2f away from the even share %.2f", phase, worst, total / 8);
  }
  printf("ok per-XCD runs of the tile table\n");
}

int main(int argc, char** argv)
{
  layout_checks();
  neighbour_checks();
  extents_checks();
  admission_checks();
  checkpoint_checks(argc > 1 ? argv[1] : ".");
  if (g_fail) return 1;
  printf("all host logic checks passed\n");
  return 0;
}
