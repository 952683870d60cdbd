// Batched trajectories (many independent plans advanced together on one GPU).
//
// The reference runs one trajectory per process (a SLURM array, one GPU
// each).  Small grids leave an MI355X mostly idle and launch-bound.  A batch
// runs its members in one of two ways:
//  * per member (modes 0-2): each member's chunk of steps is captured into a
//    HIP graph once and the graphs are replayed on a few streams (mode 1: one
//    joint graph per chunk; mode 2: one graph of every member's whole run);
//  * grouped (modes 3-4): members of one kind (dim, N, precision, viscous
//    mode, kernel variants, particle interpolator) form a group whose kernels
//    run once per pass for all of its members (multi-member launches, grid
//    z-slice = member, arguments from a device table), groups on their own
//    streams; mode 3 launches eagerly, mode 4 captures the whole run of every
//    group into one graph, one launch per run.
// A member's kernels and their arguments are those of qlles_step_advect /
// qlles_step on the same plan, and every reduction in them runs in a fixed
// order, so each trajectory is bit-identical to its solo run whatever the
// mode, stream count, chunk length or interleaving.
#include <algorithm>
#include <chrono>
#include <vector>

#include "capi_common.hpp"

using qlapi::fail;
using qlapi::guarded;

struct qlles_batch {
  struct Member {
    qlles_plan* p;
    qlles_interp* it;
    void* uh;
    double* r;
    int np;
    void* ur;
    size_t uh_bytes, r_bytes, ur_bytes;
    hipGraphExec_t g = nullptr;  // `chunk` steps on this member's arrays
  };
  std::vector<Member> m;
  int lanes = 1;  // streams the members are spread over (member i -> stream i % lanes)
  int chunk = 1;  // steps per graph launch
  int single = 0;  // 1: one graph holds every member's chunk as parallel branches;
                   // 2: one graph holds every member's whole run (its own step count)
  hipGraphExec_t all = nullptr;
  std::vector<int> all_n;  // modes 2, 4: the step counts `all` was captured for
  // grouped modes: members in table order (step counts non-increasing), the
  // engine group (nullptr: one member that cannot be grouped, run alone) and
  // the grouped particle advection (nullptr: per-member interpolation)
  struct Group {
    std::vector<int> mem;
    ql::EngineGroup* eg = nullptr;
    ql::InterpGroup* ig = nullptr;
  };
  std::vector<Group> groups;
  std::vector<int> group_perm;  // member order the groups were made for
  // microseconds per step of each kind's group run alone (in the order the
  // kinds first appear among the members), measured by the tuner; empty:
  // estimated from the grid sizes
  std::vector<double> kind_us;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  bool built = false;
  double tuned_us_per_traj_step = 0.0;
};

namespace {

void drop_groups(qlles_batch* b) {
  if (b->groups.empty()) return;
  // group tables may still be read by launches in flight
  for (hipStream_t s : b->streams) (void)hipStreamSynchronize(s);
  for (auto& g : b->groups) {
    delete g.eg;
    ql::destroy_interp_group(g.ig);
  }
  b->groups.clear();
  b->group_perm.clear();
}

void drop_graphs(qlles_batch* b) {
  // qlles_batch_run returns with its launches queued: a later configure, add
  // or destroy must not free graphs those launches are still executing
  bool live = b->all != nullptr;
  for (const auto& x : b->m) live = live || x.g != nullptr;
  if (live)
    for (hipStream_t s : b->streams) (void)hipStreamSynchronize(s);
  for (auto& x : b->m)
    if (x.g) {
      (void)hipGraphExecDestroy(x.g);
      x.g = nullptr;
    }
  if (b->all) {
    (void)hipGraphExecDestroy(b->all);
    b->all = nullptr;
  }
  b->all_n.clear();
  b->built = false;
  drop_groups(b);
}

void drop_streams(qlles_batch* b) {
  for (hipStream_t s : b->streams) (void)hipStreamSynchronize(s);  // work queued by an earlier run
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  b->streams.clear();
  b->join.clear();
}

void make_streams(qlles_batch* b) {
  if ((int)b->streams.size() == b->lanes) return;
  drop_streams(b);
  for (int l = 0; l < b->lanes; ++l) {
    hipStream_t s = nullptr;
    QL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
    b->streams.push_back(s);
    hipEvent_t e = nullptr;
    QL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->join.push_back(e);
  }
  if (!b->fork) QL_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
}

// capture `steps` steps of member x (on arrays uh/r/ur) into an executable graph
hipGraphExec_t capture(const qlles_batch::Member& x, void* uh, double* r, void* ur, int steps, hipStream_t s) {
  if (x.p->eng->profiling()) throw std::invalid_argument("disable qlles_profile on batch members");
  QL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    qlapi::run_steps(x.p, x.it, nullptr, uh, steps, r, x.np, ur, nullptr, s);
  } catch (...) {
    (void)hipStreamEndCapture(s, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  QL_HIP(hipStreamEndCapture(s, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  QL_HIP(e);
  return exec;
}

// one graph: the first stream forks to the others, member i's n[i] steps are
// captured on its stream, and the streams join back
hipGraphExec_t capture_all(qlles_batch* b, const std::vector<int>& n) {
  for (const auto& x : b->m)
    if (x.p->eng->profiling()) throw std::invalid_argument("disable qlles_profile on batch members");
  hipStream_t o = b->streams[0];
  QL_HIP(hipStreamBeginCapture(o, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    QL_HIP(hipEventRecord(b->fork, o));
    for (int l = 1; l < b->lanes; ++l) QL_HIP(hipStreamWaitEvent(b->streams[l], b->fork, 0));
    for (size_t i = 0; i < b->m.size(); ++i) {
      const auto& x = b->m[i];
      if (n[i] > 0)
        qlapi::run_steps(x.p, x.it, nullptr, x.uh, n[i], x.r, x.np, x.ur, nullptr, b->streams[i % b->lanes]);
    }
    for (int l = 1; l < b->lanes; ++l) {
      QL_HIP(hipEventRecord(b->join[l], b->streams[l]));
      QL_HIP(hipStreamWaitEvent(o, b->join[l], 0));
    }
  } catch (...) {
    (void)hipStreamEndCapture(o, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  QL_HIP(hipStreamEndCapture(o, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  QL_HIP(e);
  return exec;
}

// ---------------- grouped modes ----------------

// how a member's particles are advected: 0 none, 1 a NUFFT that can join a
// grouped one, 2 per member
int adv_class(const qlles_interp* it) {
  if (it == nullptr) return 0;
  return ql::interp_groupable(it, it) ? 1 : 2;
}

// Groups for step counts n: members of one kind together (in the order they
// were added), each group's members by decreasing step count (stable).  The
// groups are made again when that order changes.
std::vector<std::vector<int>> partition(const qlles_batch* b) {
  const int M = (int)b->m.size();
  std::vector<std::vector<int>> parts;
  for (int i = 0; i < M; ++i) {
    const auto& x = b->m[i];
    bool placed = false;
    for (auto& pt : parts) {
      const auto& r = b->m[pt[0]];
      if (!r.p->eng->groupable_with(x.p->eng)) continue;
      const int cr = adv_class(r.it), cx = adv_class(x.it);
      if (cr != cx || (cr == 1 && !ql::interp_groupable(r.it, x.it))) continue;
      pt.push_back(i);
      placed = true;
      break;
    }
    if (!placed) parts.push_back({i});
  }
  return parts;
}

void make_groups(qlles_batch* b, const std::vector<int>& n) {
  for (const auto& x : b->m)
    if (x.p->eng->profiling()) throw std::invalid_argument("disable qlles_profile on batch members");
  std::vector<std::vector<int>> parts = partition(b);
  // the key: the member order, group by group, and every member's kernel
  // variants (a group launches the variants of its first member)
  std::vector<int> perm;
  for (auto& pt : parts) {
    std::stable_sort(pt.begin(), pt.end(), [&](int a, int c) { return n[a] > n[c]; });
    perm.insert(perm.end(), pt.begin(), pt.end());
    perm.push_back(-1);
  }
  for (const auto& x : b->m) {
    int xv = 0, zv = 0;
    x.p->eng->get_variant(&xv, &zv);
    perm.push_back(xv);
    perm.push_back(zv);
  }
  if (!b->groups.empty() && perm == b->group_perm) return;
  if (b->all) {
    // the whole-run graph reads the old groups' tables
    for (hipStream_t s : b->streams) QL_HIP(hipStreamSynchronize(s));
    (void)hipGraphExecDestroy(b->all);
    b->all = nullptr;
    b->all_n.clear();
  }
  drop_groups(b);
  try {
    for (const auto& pt : parts) {
      qlles_batch::Group g;
      g.mem = pt;
      const auto& x0 = b->m[pt[0]];
      if (x0.p->eng->groupable_with(x0.p->eng)) {
        std::vector<ql::Engine*> es;
        std::vector<void*> uh;
        for (int i : pt) {
          es.push_back(b->m[i].p->eng);
          uh.push_back(b->m[i].uh);
        }
        g.eg = x0.p->eng->make_group(es, uh);
        if (adv_class(x0.it) == 1) {
          std::vector<qlles_interp*> its;
          std::vector<ql::SpecRef> cont, last;
          std::vector<double*> r;
          std::vector<int> np;
          std::vector<void*> ur;
          std::vector<double> dt;
          for (size_t j = 0; j < pt.size(); ++j) {
            const auto& x = b->m[pt[j]];
            its.push_back(x.it);
            cont.push_back(g.eg->spec_cont((int)j));
            last.push_back(g.eg->spec_last((int)j));
            r.push_back(x.r);
            np.push_back(x.np);
            ur.push_back(x.ur);
            dt.push_back(x.p->desc.dt);
          }
          try {
            g.ig = ql::make_interp_group(its, x0.p->desc.N, cont, last, r, np, ur, dt);
          } catch (...) {
            delete g.eg;
            throw;
          }
        }
      }
      b->groups.push_back(g);
    }
  } catch (...) {
    drop_groups(b);
    throw;
  }
  b->group_perm = perm;
}

// steps [from, to) of group g, whose members take ns[j] steps, on stream s
void group_steps(qlles_batch* b, qlles_batch::Group& g, const std::vector<int>& ns, int from, int to, hipStream_t s) {
  if (g.eg == nullptr) {  // a member of its own kind, stepped as qlles_step_advect does
    const auto& x = b->m[g.mem[0]];
    const int k = std::min(to, ns[0]) - from;
    if (k > 0) qlapi::run_steps(x.p, x.it, nullptr, x.uh, k, x.r, x.np, x.ur, nullptr, s);
    return;
  }
  std::function<void(hipStream_t, int, int)> after;
  if (g.ig) {
    after = [&g](hipStream_t st, int m, int f) { ql::interp_group_run(g.ig, m, f, st); };
  } else if (b->m[g.mem[0]].it) {
    after = [b, &g](hipStream_t st, int m, int f) {
      for (int j = 0; j < m; ++j) {
        const auto& x = b->m[g.mem[j]];
        const ql::SpecRef sp = j < f ? g.eg->spec_cont(j) : g.eg->spec_last(j);
        ql::interp_hat(x.it, x.p->desc.N, sp.p, sp.ld, sp.q, x.r, x.np, x.ur, x.p->desc.dt, st);
      }
    };
  }
  g.eg->steps(ns.data(), from, to, s, after);
}

// time per step of group g run alone: measured by the tuner, else estimated
// as the longer of its launch chain (about 6 us per kernel) and its bytes at
// 4 TB/s (3-D: 155 B per grid point and byte of precision, the fused step's
// traffic; 2-D: 50)
double group_us(const qlles_batch* b, size_t gi) {
  if (gi < b->kind_us.size() && b->kind_us[gi] > 0) return b->kind_us[gi];
  const auto& g = b->groups[gi];
  const ql::Desc& d = b->m[g.mem[0]].p->desc;
  const double pts = d.dim == 3 ? (double)d.N * d.N * d.N : (double)d.N * d.N;
  const double eb = d.dtype == QLLES_F32 ? 4.0 : 8.0;
  const double bytes = (d.dim == 3 ? 155.0 : 50.0) * eb * pts * (double)g.mem.size();
  return std::max((d.dim == 3 ? 22.0 : 12.0) * 6.0, bytes / 4e6);
}

// the grouped run of step counts n.  Groups go to the lanes longest first,
// each to the lane with the least work so far (a group's work: its time per
// step times its longest member's steps), so the heaviest kinds do not share
// a stream -- and with it a hardware queue, which runs its kernels in order
// -- with others.  The groups are interleaved `chunk` steps at a time so
// every stream has queued work.
void enqueue_grouped(qlles_batch* b, const std::vector<int>& n) {
  const size_t G = b->groups.size();
  std::vector<std::vector<int>> ns(G);
  int maxn = 0;
  for (size_t gi = 0; gi < G; ++gi) {
    for (int i : b->groups[gi].mem) ns[gi].push_back(n[i]);
    maxn = std::max(maxn, ns[gi][0]);
  }
  std::vector<size_t> order(G);
  std::vector<double> work(G);
  for (size_t gi = 0; gi < G; ++gi) {
    order[gi] = gi;
    work[gi] = group_us(b, gi) * ns[gi][0];
  }
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work[x] > work[y]; });
  std::vector<int> lane(G, 0);
  std::vector<double> load(b->lanes, 0.0);
  for (size_t gi : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lane[gi] = l;
    load[l] += work[gi];
  }
  for (int k = 0; k < maxn; k += b->chunk)
    for (size_t gi : order)
      if (ns[gi][0] > k)
        group_steps(b, b->groups[gi], ns[gi], k, std::min(k + b->chunk, ns[gi][0]), b->streams[lane[gi]]);
}

hipGraphExec_t capture_grouped(qlles_batch* b, const std::vector<int>& n) {
  hipStream_t o = b->streams[0];
  QL_HIP(hipStreamBeginCapture(o, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    QL_HIP(hipEventRecord(b->fork, o));
    for (int l = 1; l < b->lanes; ++l) QL_HIP(hipStreamWaitEvent(b->streams[l], b->fork, 0));
    enqueue_grouped(b, n);
    for (int l = 1; l < b->lanes; ++l) {
      QL_HIP(hipEventRecord(b->join[l], b->streams[l]));
      QL_HIP(hipStreamWaitEvent(o, b->join[l], 0));
    }
  } catch (...) {
    (void)hipStreamEndCapture(o, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  QL_HIP(hipStreamEndCapture(o, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  QL_HIP(e);
  return exec;
}

void run_grouped(qlles_batch* b, const std::vector<int>& n, hipStream_t caller) {
  make_groups(b, n);  // outside any capture: it allocates and uploads tables
  if (b->single == 4 && (b->all == nullptr || b->all_n != n)) {
    if (b->all) {
      QL_HIP(hipStreamSynchronize(caller));  // a previous launch may still run
      (void)hipGraphExecDestroy(b->all);
      b->all = nullptr;
    }
    b->all = capture_grouped(b, n);
    b->all_n = n;
  }
  QL_HIP(hipEventRecord(b->fork, caller));
  for (hipStream_t s : b->streams) QL_HIP(hipStreamWaitEvent(s, b->fork, 0));
  if (b->single == 4) {
    QL_HIP(hipGraphLaunch(b->all, b->streams[0]));
    QL_HIP(hipEventRecord(b->join[0], b->streams[0]));
    QL_HIP(hipStreamWaitEvent(caller, b->join[0], 0));
    return;
  }
  enqueue_grouped(b, n);
  for (int l = 0; l < b->lanes; ++l) {
    QL_HIP(hipEventRecord(b->join[l], b->streams[l]));
    QL_HIP(hipStreamWaitEvent(caller, b->join[l], 0));
  }
}

void build(qlles_batch* b) {
  if (b->built) return;
  make_streams(b);
  drop_graphs(b);
  if (b->single < 2) {
    for (size_t i = 0; i < b->m.size(); ++i) {
      auto& x = b->m[i];
      x.g = capture(x, x.uh, x.r, x.ur, b->chunk, b->streams[i % b->lanes]);
    }
  }
  if (b->single == 1) b->all = capture_all(b, std::vector<int>(b->m.size(), b->chunk));
  b->built = true;
}

// advance member i by n_i steps: whole chunks by graph replay, the rest
// eagerly; in mode 2, one launch of the graph of the whole run (captured
// again when the step counts change)
void run(qlles_batch* b, const std::vector<int>& n, hipStream_t caller) {
  build(b);
  if (b->single >= 3) {
    run_grouped(b, n, caller);
    return;
  }
  if (b->single == 2 && (b->all == nullptr || b->all_n != n)) {
    if (b->all) {
      QL_HIP(hipStreamSynchronize(caller));  // a previous launch may still run
      (void)hipGraphExecDestroy(b->all);
      b->all = nullptr;
    }
    b->all = capture_all(b, n);
    b->all_n = n;
  }
  QL_HIP(hipEventRecord(b->fork, caller));
  for (hipStream_t s : b->streams) QL_HIP(hipStreamWaitEvent(s, b->fork, 0));
  if (b->single == 2) {
    QL_HIP(hipGraphLaunch(b->all, b->streams[0]));
    QL_HIP(hipEventRecord(b->join[0], b->streams[0]));
    QL_HIP(hipStreamWaitEvent(caller, b->join[0], 0));
    return;
  }
  std::vector<int> left(n);
  if (b->single == 1 && b->all) {
    // whole chunks that every member still needs: one launch of the joint graph
    int common = *std::min_element(left.begin(), left.end()) / b->chunk;
    for (int c = 0; c < common; ++c) QL_HIP(hipGraphLaunch(b->all, b->streams[0]));
    for (int& v : left) v -= common * b->chunk;
    // the other streams continue after the joint launches
    if (common > 0 && b->lanes > 1) {
      QL_HIP(hipEventRecord(b->join[0], b->streams[0]));
      for (int l = 1; l < b->lanes; ++l) QL_HIP(hipStreamWaitEvent(b->streams[l], b->join[0], 0));
    }
  }
  // round-robin over members so that every stream always has queued work
  for (bool any = true; any;) {
    any = false;
    for (size_t i = 0; i < b->m.size(); ++i) {
      if (left[i] >= b->chunk) {
        QL_HIP(hipGraphLaunch(b->m[i].g, b->streams[i % b->lanes]));
        left[i] -= b->chunk;
        any = true;
      }
    }
  }
  for (size_t i = 0; i < b->m.size(); ++i) {
    if (left[i] > 0) {
      const auto& x = b->m[i];
      qlapi::run_steps(x.p, x.it, nullptr, x.uh, left[i], x.r, x.np, x.ur, nullptr, b->streams[i % b->lanes]);
    }
  }
  for (int l = 0; l < b->lanes; ++l) {
    QL_HIP(hipEventRecord(b->join[l], b->streams[l]));
    QL_HIP(hipStreamWaitEvent(caller, b->join[l], 0));
  }
}

size_t elem_bytes(const qlles_plan* p) { return p->desc.dtype == QLLES_F32 ? 4 : 8; }

}  // namespace

extern "C" {

int qlles_batch_create(qlles_batch** out) {
  if (!out) return fail(QLLES_ERR_ARG, "null argument");
  *out = nullptr;
  return guarded([&] { *out = new qlles_batch(); });
}

int qlles_batch_destroy(qlles_batch* b) {
  if (!b) return QLLES_OK;
  return guarded([&] {
    drop_graphs(b);
    drop_streams(b);
    if (b->fork) (void)hipEventDestroy(b->fork);
    delete b;
  });
}

int qlles_batch_add(qlles_batch* b, qlles_plan* p, qlles_interp* it, void* u_hat, double* r, int np, void* u_r,
                    int* index) {
  if (!b || !p || !u_hat) return fail(QLLES_ERR_ARG, "null argument");
  if (it && (!r || !u_r || np < 0)) return fail(QLLES_ERR_ARG, "particles need r, u_r and np >= 0");
  if (it && (ql::interp_dim(it) != p->desc.dim || ql::interp_dtype(it) != p->desc.dtype ||
             ql::interp_n(it) > p->desc.N))
    return fail(QLLES_ERR_ARG, "interpolator does not match the plan (dim, dtype, n <= N)");
  for (const auto& x : b->m) {
    if (x.p == p) return fail(QLLES_ERR_ARG, "each batch member needs its own plan (plans hold workspace)");
    if (it && x.it == it) return fail(QLLES_ERR_ARG, "each batch member needs its own interpolator");
  }
  return guarded([&] {
    const int N = p->desc.N, dim = p->desc.dim;
    const size_t rows = dim == 3 ? (size_t)N * N : (size_t)N;
    p->eng->prepare();  // the step workspace exists before any capture
    qlles_batch::Member x{p, it, u_hat, it ? r : nullptr, it ? np : 0, it ? u_r : nullptr,
                          3 * rows * (N / 2 + 1) * 2 * elem_bytes(p),
                          it ? (size_t)np * dim * sizeof(double) : 0, it ? (size_t)np * dim * elem_bytes(p) : 0};
    b->m.push_back(x);
    drop_graphs(b);
    b->kind_us.clear();
    if (index) *index = (int)b->m.size() - 1;
  });
}

int qlles_batch_configure(qlles_batch* b, int lanes, int chunk, int single_graph) {
  if (!b) return fail(QLLES_ERR_ARG, "null batch");
  if (lanes < 1 || lanes > 16 || chunk < 1 || chunk > 1000) return fail(QLLES_ERR_ARG, "need 1 <= lanes <= 16, 1 <= chunk <= 1000");
  if (single_graph < 0 || single_graph > 4) return fail(QLLES_ERR_ARG, "single_graph must be 0 ... 4");
  return guarded([&] {
    drop_graphs(b);
    b->lanes = lanes;
    b->chunk = chunk;
    b->single = single_graph;
  });
}

int qlles_batch_get_config(const qlles_batch* b, int* lanes, int* chunk, int* single_graph, double* us_per_traj_step) {
  if (!b) return fail(QLLES_ERR_ARG, "null batch");
  if (lanes) *lanes = b->lanes;
  if (chunk) *chunk = b->chunk;
  if (single_graph) *single_graph = b->single;
  if (us_per_traj_step) *us_per_traj_step = b->tuned_us_per_traj_step;
  return QLLES_OK;
}

int qlles_batch_kinds(const qlles_batch* b, int* nkinds, int* kind_of, double* us_per_step, int cap) {
  if (!b || !nkinds) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] {
    const auto parts = partition(b);
    *nkinds = (int)parts.size();
    for (size_t k = 0; k < parts.size(); ++k) {
      if (kind_of)
        for (int i : parts[k]) kind_of[i] = (int)k;
      if (us_per_step && (int)k < cap) us_per_step[k] = k < b->kind_us.size() ? b->kind_us[k] : 0.0;
    }
  });
}

int qlles_batch_run(qlles_batch* b, const int* nsteps, int nsteps_all, void* stream) {
  if (!b) return fail(QLLES_ERR_ARG, "null batch");
  if (!nsteps && nsteps_all < 0) return fail(QLLES_ERR_ARG, "step counts must be >= 0");
  std::vector<int> n(b->m.size(), nsteps_all);
  if (nsteps)
    for (size_t i = 0; i < n.size(); ++i) n[i] = nsteps[i];
  for (int v : n)
    if (v < 0) return fail(QLLES_ERR_ARG, "step counts must be >= 0");
  if (b->m.empty()) return QLLES_OK;
  return guarded([&] { run(b, n, (hipStream_t)stream); });
}

int qlles_batch_tune(qlles_batch* b, int trial_steps, void* stream) {
  if (!b) return fail(QLLES_ERR_ARG, "null batch");
  if (trial_steps < 1) return fail(QLLES_ERR_ARG, "trial_steps must be >= 1");
  if (b->m.empty()) return QLLES_OK;
  const hipStream_t caller = (hipStream_t)stream;
  std::vector<void*> scratch;
  int rc = guarded([&] {
    // trials run on copies of the member states, so tuning leaves every
    // trajectory where it was
    qlles_batch t;
    for (const auto& x : b->m) {
      qlles_batch::Member y = x;
      y.g = nullptr;
      void* p = nullptr;
      QL_HIP(hipMalloc(&p, x.uh_bytes));
      scratch.push_back(p);
      QL_HIP(hipMemcpyAsync(p, x.uh, x.uh_bytes, hipMemcpyDeviceToDevice, caller));
      y.uh = p;
      if (x.it) {
        QL_HIP(hipMalloc(&p, x.r_bytes));
        scratch.push_back(p);
        QL_HIP(hipMemcpyAsync(p, x.r, x.r_bytes, hipMemcpyDeviceToDevice, caller));
        y.r = static_cast<double*>(p);
        QL_HIP(hipMalloc(&p, std::max<size_t>(x.ur_bytes, 1)));
        scratch.push_back(p);
        y.ur = p;
      }
      t.m.push_back(y);
    }
    QL_HIP(hipStreamSynchronize(caller));
    // kernel variants first, member by member (plan state, shared with solo runs)
    for (const auto& x : b->m) qlapi::tune_plan(x.p, x.uh, std::max(2, trial_steps / 4), nullptr, caller);
    const int nm = (int)b->m.size();
    double best = -1.0;
    int best_l = 1, best_c = 1, best_s = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    QL_HIP(hipEventCreate(&e0));
    QL_HIP(hipEventCreate(&e1));
    try {
      // each kind's group alone, timed per step: the costs that spread the
      // groups over the streams in the grouped modes
      {
        t.single = 3;
        t.lanes = 1;
        t.chunk = 1;
        make_streams(&t);
        const std::vector<int> even(nm, trial_steps);
        make_groups(&t, even);
        std::vector<double> us(t.groups.size(), 0.0);
        for (size_t gi = 0; gi < t.groups.size(); ++gi) {
          std::vector<int> ns(t.groups[gi].mem.size(), trial_steps);
          group_steps(&t, t.groups[gi], ns, 0, 1, caller);  // warm-up
          QL_HIP(hipEventRecord(e0, caller));
          group_steps(&t, t.groups[gi], ns, 0, trial_steps, caller);
          QL_HIP(hipEventRecord(e1, caller));
          QL_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          QL_HIP(hipEventElapsedTime(&ms, e0, e1));
          us[gi] = 1e3 * ms / trial_steps;
        }
        t.kind_us = us;
        b->kind_us = us;
        drop_graphs(&t);
      }
      // the whole-run graph (single = 2) is not a candidate: the trials have
      // even lengths, where it times well, but members sharing a stream run
      // their whole trajectories back to back, which uneven lengths turn into
      // idle streams (24 uneven trajectories: 8357 against 9978
      // trajectory-steps/s with the chunked joint graph)
      // grouped modes: streams beyond one per group stay idle; chunk is the
      // interleaving of the eager mode and has no effect on the graph
      const int ngroups = (int)partition(b).size();
      for (int single : {0, 1, 3, 4})
      for (int lanes : {1, 2, 3, 4, 6, 8, 16}) {
        if (single < 3 && (lanes == 3 || lanes == 6)) continue;
        if (lanes > (single >= 3 ? ngroups : nm) && lanes > 1) break;
        for (int chunk : {1, 4, 16, 64}) {
          if (chunk > trial_steps && chunk > 1) break;
          if (single == 4 && chunk > 1) break;
          if (single == 3 && chunk > 16) break;
          t.lanes = lanes;
          t.chunk = chunk;
          t.single = single;
          drop_graphs(&t);
          const int n = std::max(chunk, trial_steps / chunk * chunk);
          // warm-up (also builds the graphs; the whole-run graph of mode 4 for
          // the timed step counts)
          run(&t, std::vector<int>(nm, single == 4 ? n : chunk), caller);
          QL_HIP(hipEventRecord(e0, caller));
          run(&t, std::vector<int>(nm, n), caller);
          QL_HIP(hipEventRecord(e1, caller));
          QL_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          QL_HIP(hipEventElapsedTime(&ms, e0, e1));
          const double us = 1e3 * ms / ((double)n * nm);
          if (best < 0 || us < best) {
            best = us;
            best_l = lanes;
            best_c = chunk;
            best_s = single;
          }
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      drop_graphs(&t);
      drop_streams(&t);
      if (t.fork) (void)hipEventDestroy(t.fork);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    drop_graphs(&t);
    drop_streams(&t);
    if (t.fork) (void)hipEventDestroy(t.fork);
    drop_graphs(b);
    b->lanes = best_l;
    b->chunk = best_c;
    b->single = best_s;
    b->tuned_us_per_traj_step = best;
  });
  (void)hipStreamSynchronize(caller);
  for (void* p : scratch) (void)hipFree(p);
  return rc;
}

}  // extern "C"
