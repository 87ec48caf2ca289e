// Batched form: many independent MOLOCH domains advanced together.
//
// A step of every member is recorded, not launched: each stage the member's
// launchers emit (the same launchers the single-domain step runs) becomes a
// LaunchRec holding the stage's argument struct.  The i-th records of all
// members are the same stage, so they merge into one launch of the stage's
// batched entry point over the tiles of every member: an argument table
// (one struct per member) and a block map (block -> member, tile).  The
// merged launches of one step are captured once into a hipGraph and the
// graph is replayed each step.  The plan depends only on which buffer of
// each ping-pong pair is live at the start of the step, so at most two
// plans exist (an odd nadv alternates them).
//
// The members can be split into at most kMaxGroups groups
// (mlc_batch_set_streams), one per hardware queue of the process, each with
// its own stream and graphs, so that different stages of different groups
// run concurrently.  With one group (the default) a step is one graph.
//
// Block order: the hardware hands workgroup b of a launch to XCD b mod 8.
// The map gives each XCD a contiguous run of tiles (member by member, level
// by level, row by row), so neighbouring tiles, whose halos overlap, share
// that XCD's L2.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "mlc_internal.h"

namespace mlc {

struct BatchLaunch {
  const void *fn = nullptr;
  dim3 block;
  size_t lds = 0;
  unsigned nblocks = 0;
  size_t tab_off = 0, map_off = 0;
};

struct BatchPlan {
  std::vector<int> key;                  // live buffers at step start
  std::vector<int> post;                 // live buffers after the step
  std::vector<BatchLaunch> launches;
  void *arena = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

constexpr int kXcds = 8;
constexpr int kMaxGroups = 4;   // the hardware queues of a process (GPU_MAX_HW_QUEUES)

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Tile list of a launch in dispatch order for an XCD-contiguous assignment.
static void xcd_order(const std::vector<uint4> &tiles, uint4 *out) {
  const size_t T = tiles.size();
  size_t start[kXcds + 1];
  start[0] = 0;
  for (int x = 0; x < kXcds; ++x) {
    const size_t cnt = T > (size_t)x ? (T - x + kXcds - 1) / kXcds : 0;
    start[x + 1] = start[x] + cnt;
  }
  for (size_t b = 0; b < T; ++b) out[b] = tiles[start[b % kXcds] + b / kXcds];
}

}  // namespace mlc

// A batch runs as one or more groups of members; each group has its own
// stream and step graphs, so the groups' stages overlap on the device.
struct BatchGroup {
  std::vector<mlc_handle *> m;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<mlc::BatchPlan> plans;
  std::vector<unsigned> gens;         // members' config_gen when the plans were built
};

struct mlc_batch {
  std::vector<mlc_handle *> all;
  std::vector<BatchGroup> groups;
  hipStream_t stream = nullptr;       // group 0's stream
  long long npoints = 0;
};

namespace mlc {

static std::vector<int> live_key(const BatchGroup *b) {
  std::vector<int> k;
  for (auto *h : b->m) k.insert(k.end(), h->cur, h->cur + kMaxSlots);
  return k;
}

static void set_live(BatchGroup *b, const std::vector<int> &k) {
  for (size_t i = 0; i < b->m.size(); ++i)
    std::copy(k.begin() + i * kMaxSlots, k.begin() + (i + 1) * kMaxSlots, b->m[i]->cur);
}

static void free_plan(BatchPlan &p) {
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  if (p.arena) (void)hipFree(p.arena);
  p.exec = nullptr;
  p.graph = nullptr;
  p.arena = nullptr;
}

// Record one step of every member and build its merged launches and graph.
static int build_plan(BatchGroup *b, BatchPlan *plan) {
  const size_t n = b->m.size();
  if (n == 0) return fail(MLC_E_STATE, "batch group without members");
  plan->key = live_key(b);
  std::vector<std::vector<LaunchRec>> recs(n);
  int rc = MLC_OK;
  for (size_t i = 0; i < n && rc == MLC_OK; ++i) {
    mlc_handle *h = b->m[i];
    h->rec = &recs[i];
    rc = step_body(h);
    h->rec = nullptr;
  }
  plan->post = live_key(b);
  set_live(b, plan->key);
  if (rc) return rc;
  // lay out the argument tables and block maps in one arena
  std::vector<unsigned char> host;
  auto reserve = [&](size_t bytes) {
    const size_t off = align_up(host.size(), 256);
    host.resize(off + bytes);
    return off;
  };
  plan->launches.clear();
  std::vector<uint4> tiles;
  // Members launch the same stage with one kernel when their records agree;
  // records may differ (another variant of a stage, a pass fused into
  // another on some members only), so the merge walks every member's record
  // in its own order.  Each round launches one kind of stage (entry point,
  // block shape, argument size) for the members whose next launch it is,
  // preferring a kind that no other member still has pending further on, so
  // that the members missing a stage do not split the launches they share.
  auto same = [](const LaunchRec &a, const LaunchRec &b) {
    return a.many == b.many && a.block.x == b.block.x && a.block.y == b.block.y &&
           a.block.z == b.block.z && a.args.size() == b.args.size();
  };
  std::vector<size_t> pos(n, 0);
  for (;;) {
    std::vector<size_t> live;
    for (size_t i = 0; i < n; ++i)
      if (pos[i] < recs[i].size()) live.push_back(i);
    if (live.empty()) break;
    // the kind to launch: the first next launch that every member still
    // holding that kind has as its next one, else the first member's
    const LaunchRec *kind = &recs[live[0]][pos[live[0]]];
    for (size_t a : live) {
      const LaunchRec &cand = recs[a][pos[a]];
      bool ready = true;
      for (size_t i : live) {
        if (same(recs[i][pos[i]], cand)) continue;
        for (size_t q = pos[i] + 1; q < recs[i].size() && ready; ++q)
          if (same(recs[i][q], cand)) ready = false;
        if (!ready) break;
      }
      if (ready) {
        kind = &cand;
        break;
      }
    }
    std::vector<size_t> gr;
    for (size_t i : live)
      if (same(recs[i][pos[i]], *kind)) gr.push_back(i);
    BatchLaunch L;
    L.fn = kind->many;
    L.block = kind->block;
    // tab[m] of the batched entry point: structs packed at sizeof(A)
    const size_t asz = kind->args.size();
    L.tab_off = reserve(asz * gr.size());
    tiles.clear();
    for (size_t m = 0; m < gr.size(); ++m) {
      const LaunchRec &r = recs[gr[m]][pos[gr[m]]];
      std::memcpy(host.data() + L.tab_off + m * asz, r.args.data(), asz);
      L.lds = std::max(L.lds, r.lds);
      for (unsigned z = 0; z < r.grid.z; ++z)
        for (unsigned y = 0; y < r.grid.y; ++y)
          for (unsigned x = 0; x < r.grid.x; ++x)
            tiles.push_back(make_uint4((unsigned)m, x, y, z));
    }
    if (tiles.empty() || tiles.size() > 0x7fffffffULL)
      return fail(MLC_E_STATE, "batched launch has no tiles or too many");
    L.nblocks = (unsigned)tiles.size();
    L.map_off = reserve(tiles.size() * sizeof(uint4));
    xcd_order(tiles, reinterpret_cast<uint4 *>(host.data() + L.map_off));
    plan->launches.push_back(L);
    for (size_t i : gr) ++pos[i];
  }
  MLC_HIP(hipMalloc(&plan->arena, host.size()));
  MLC_HIP(hipMemcpy(plan->arena, host.data(), host.size(), hipMemcpyHostToDevice));
  // capture the step
  MLC_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t err = hipSuccess;
  for (auto &L : plan->launches) {
    const void *tab = static_cast<unsigned char *>(plan->arena) + L.tab_off;
    const void *map = static_cast<unsigned char *>(plan->arena) + L.map_off;
    void *kargs[2] = {&tab, &map};
    err = hipLaunchKernel(L.fn, dim3(L.nblocks), L.block, kargs, L.lds, b->stream);
    if (err != hipSuccess) break;
  }
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
  if (err != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return hip_check(err, "hipLaunchKernel (capture)");
  }
  if (e2 != hipSuccess) return hip_check(e2, "hipStreamEndCapture");
  plan->graph = graph;
  MLC_HIP(hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0));
  return MLC_OK;
}

}  // namespace mlc

using namespace mlc;

extern "C" {

int mlc_batch_create(const mlc_grid_desc *descs, int n, mlc_batch **out) {
  if (!descs || !out || n < 1) return fail(MLC_E_ARG, "null argument or n < 1");
  *out = nullptr;
  // the stages a member records: TKE staggering, the tracer update, and
  // one WAF launch per kMaxAdv bulk fields
  auto chunks = [](const mlc_grid_desc &g) {
    const int nb = A_QX0 + 1 + (g.ipptls > 0 ? g.nqx - 1 : 0) + (g.ibltyp == 2) +
                   (g.ntr > 0 ? g.ntr : 0) - (g.precision == MLC_MIXED ? 2 : 0);
    return (nb + kMaxAdv - 1) / kMaxAdv;
  };
  for (int i = 1; i < n; ++i) {
    if (descs[i].precision != descs[0].precision || descs[i].nadv != descs[0].nadv ||
        descs[i].nsound != descs[0].nsound)
      return fail(MLC_E_ARG, "batch members must share precision, nadv and nsound");
    if ((descs[i].ibltyp == 2) != (descs[0].ibltyp == 2) ||
        (descs[i].ntr > 0) != (descs[0].ntr > 0) || chunks(descs[i]) != chunks(descs[0]))
      return fail(MLC_E_ARG, "batch members must agree on TKE (ibltyp 2), on having tracers, "
                             "and on the number of WAF launches their fields need");
  }
  mlc_batch *b = new (std::nothrow) mlc_batch();
  if (!b) return fail(MLC_E_ARG, "out of host memory");
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(MLC_E_HIP, "hipStreamCreate failed");
  }
  b->groups.emplace_back();
  b->groups[0].stream = b->stream;
  for (int i = 0; i < n; ++i) {
    mlc_handle *h = nullptr;
    const int rc = create_handle(&descs[i], b->stream, &h);
    if (rc) {
      const std::string msg = std::string("member ") + std::to_string(i) + ": " + mlc_last_error();
      mlc_batch_destroy(b);
      return fail(rc, msg);
    }
    h->batch_member = true;
    b->all.push_back(h);
    b->groups[0].m.push_back(h);
    b->npoints += mlc_npoints(h);
  }
  *out = b;
  return MLC_OK;
}

int mlc_batch_member(mlc_batch *b, int m, mlc_handle **h) {
  if (!b || !h) return fail(MLC_E_ARG, "null argument");
  if (m < 0 || m >= (int)b->all.size()) return fail(MLC_E_ARG, "member index out of range");
  *h = b->all[m];
  return MLC_OK;
}

int mlc_batch_step(mlc_batch *b, int nsteps) {
  if (!b) return fail(MLC_E_ARG, "null batch");
  if (nsteps < 0) return fail(MLC_E_ARG, "nsteps < 0");
  for (auto *h : b->all)
    if (!h->inited) return fail(MLC_E_STATE, "mlc_init has not been called on every member");
  int rc;
  for (int s = 0; s < nsteps; ++s) {
    for (auto &gr : b->groups) {
      for (auto *h : gr.m)
        if (h->tend_dirty && (rc = mlc_reset_tendencies(h))) return rc;
      // a member's parameter or launch shape changed since the plans were
      // recorded (their argument structs hold copies): record them again
      std::vector<unsigned> gens;
      for (auto *h : gr.m) gens.push_back(h->config_gen);
      if (gens != gr.gens) {
        MLC_HIP(hipStreamSynchronize(gr.stream));
        for (auto &p : gr.plans) free_plan(p);
        gr.plans.clear();
        gr.gens = gens;
      }
      const std::vector<int> key = live_key(&gr);
      BatchPlan *plan = nullptr;
      for (auto &p : gr.plans)
        if (p.key == key) plan = &p;
      if (!plan) {
        gr.plans.emplace_back();
        plan = &gr.plans.back();
        if ((rc = build_plan(&gr, plan))) {
          free_plan(*plan);
          gr.plans.pop_back();
          return rc;
        }
      }
      MLC_HIP(hipGraphLaunch(plan->exec, gr.stream));
      set_live(&gr, plan->post);
    }
  }
  return MLC_OK;
}

static void drop_plans(mlc_batch *b) {
  for (auto &gr : b->groups) {
    for (auto &p : gr.plans) free_plan(p);
    gr.plans.clear();
  }
}

int mlc_batch_set_streams(mlc_batch *b, int k) {
  if (!b) return fail(MLC_E_ARG, "null batch");
  if (k < 1 || k > (int)b->all.size() || k > kMaxGroups)
    return fail(MLC_E_ARG, "group count must lie in 1..min(members, 4)");
  for (auto &gr : b->groups) MLC_HIP(hipStreamSynchronize(gr.stream));
  // the new streams first, so that a failure leaves the batch as it was
  std::vector<hipStream_t> fresh((size_t)k, nullptr);
  fresh[0] = b->stream;
  for (int g = 1; g < k; ++g)
    if (hipStreamCreateWithFlags(&fresh[g], hipStreamNonBlocking) != hipSuccess) {
      for (int q = 1; q < g; ++q) (void)hipStreamDestroy(fresh[q]);
      return fail(MLC_E_HIP, "hipStreamCreate failed");
    }
  drop_plans(b);
  for (auto &gr : b->groups)
    if (gr.own_stream) (void)hipStreamDestroy(gr.stream);
  b->groups.clear();
  b->groups.resize((size_t)k);
  for (int g = 0; g < k; ++g) {
    b->groups[g].stream = fresh[g];
    b->groups[g].own_stream = g > 0;
  }
  // largest members first onto the group with the fewest points
  std::vector<size_t> order(b->all.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  auto pts = [&](size_t i) { return mlc_npoints(b->all[i]); };
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return pts(a) > pts(c); });
  std::vector<long long> load((size_t)k, 0);
  for (size_t i : order) {
    size_t g = 0;
    for (size_t q = 1; q < load.size(); ++q)
      if (load[q] < load[g]) g = q;
    load[g] += pts(i);
    b->all[i]->stream = b->groups[g].stream;
    b->groups[g].m.push_back(b->all[i]);
  }
  return MLC_OK;
}

int mlc_batch_autotune(mlc_batch *b, int nsteps) {
  if (!b) return fail(MLC_E_ARG, "null batch");
  if (nsteps < 1) return fail(MLC_E_ARG, "nsteps < 1");
  for (auto *h : b->all) {
    const int rc = autotune(h, nsteps, nullptr);
    if (rc) return rc;
  }
  // The knobs that change how many launches a stage records (the mixed
  // form's merged WAF batches, the horizontal WAF form they apply to and its
  // fields per block, which also decides whether the curvature pass runs on
  // its own) are made to agree across members, so that their records line
  // up and every stage merges into shared launches: take the value most
  // members chose.
  auto common = [&](int mlc_handle::Tune::*knob) {
    std::vector<int> vals;
    for (auto *h : b->all) vals.push_back(h->tune.*knob);
    int best = vals[0], best_n = 0;
    for (int v : vals) {
      const int n = (int)std::count(vals.begin(), vals.end(), v);
      if (n > best_n) { best = v; best_n = n; }
    }
    for (auto *h : b->all) h->tune.*knob = best;
  };
  common(&mlc_handle::Tune::mixed_merge);
  common(&mlc_handle::Tune::waf_h);
  common(&mlc_handle::Tune::waf_hgroup);   // 0 makes the fused mixed WAF one launch
  drop_plans(b);                           // the launch shapes changed
  // Then the member grouping: merged launches in one stream fill the GPU
  // with small members, concurrent groups overlap the latency-bound stages
  // of large ones.  Each candidate count is timed over nsteps batch steps
  // (after one step that captures its graphs) from the same state.
  const int n = (int)b->all.size();
  if (n < 2) return MLC_OK;
  std::vector<Snapshot> snaps(b->all.size());
  int rc = MLC_OK;
  auto sync_all = [&]() -> int {
    for (auto &gr : b->groups) MLC_HIP(hipStreamSynchronize(gr.stream));
    return MLC_OK;
  };
  auto finish = [&](int code) {
    (void)sync_all();
    for (auto &sn : snaps) snapshot_free(&sn);
    return code;
  };
  for (size_t i = 0; i < b->all.size() && rc == MLC_OK; ++i)
    rc = snapshot_take(b->all[i], &snaps[i]);
  if (rc || (rc = sync_all())) return finish(rc);
  const int keep = (int)b->groups.size();
  int best_k = keep;
  double best_s = 0.0;
  for (int k = 1; k <= kMaxGroups; k *= 2) {
    if (k > n) break;
    if ((rc = mlc_batch_set_streams(b, k))) return finish(rc);
    for (size_t i = 0; i < b->all.size() && rc == MLC_OK; ++i)
      rc = snapshot_restore(b->all[i], snaps[i]);
    if (rc || (rc = sync_all())) return finish(rc);
    if ((rc = mlc_batch_step(b, 1)) || (rc = sync_all())) return finish(rc);
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = mlc_batch_step(b, nsteps)) || (rc = sync_all())) return finish(rc);
    const double sec =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (best_s == 0.0 || sec < best_s) {
      best_s = sec;
      best_k = k;
    }
  }
  if ((rc = mlc_batch_set_streams(b, best_k))) return finish(rc);
  for (size_t i = 0; i < b->all.size() && rc == MLC_OK; ++i)
    rc = snapshot_restore(b->all[i], snaps[i]);
  return finish(rc);
}

int mlc_batch_streams(const mlc_batch *b) { return b ? (int)b->groups.size() : 0; }

int mlc_batch_synchronize(mlc_batch *b) {
  if (!b) return fail(MLC_E_ARG, "null batch");
  for (auto &gr : b->groups) MLC_HIP(hipStreamSynchronize(gr.stream));
  return MLC_OK;
}

long long mlc_batch_npoints(const mlc_batch *b) { return b ? b->npoints : 0; }

int mlc_batch_destroy(mlc_batch *b) {
  if (!b) return MLC_OK;
  for (auto &gr : b->groups)
    if (gr.stream) (void)hipStreamSynchronize(gr.stream);
  drop_plans(b);
  for (auto *h : b->all) destroy_handle(h);
  for (auto &gr : b->groups)
    if (gr.own_stream) (void)hipStreamDestroy(gr.stream);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MLC_OK;
}

}  // extern "C"
