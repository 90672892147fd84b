// Ensemble runner: many independent boxes stepped together.
//
// An ensemble operation (one step; a segment recorded with its adjoint tape;
// the adjoint over a tape) is recorded per member by running the member's
// entry point (qj_kdk_steps, qj_kdk_record, qj_kdk_adjoint) under a launch
// Recorder (qj_launch.h): nothing is issued, every launch is kept as
// (batched wrapper, grid, block, argument bytes).  Members whose recorded
// sequences match launch for launch (same kernels, grids and blocks: same
// grid size, precision, step kind and schedule options) form a group; each
// stage of a group becomes one k_many launch with blockIdx.z = member, whose
// argument packs sit in one device buffer.  Small boxes, which alone fill a
// fraction of the 256 CUs, so run as wide launches, and the per-step launch
// count no longer grows with the number of boxes.  The step program is
// cached by the members' entry states and replayed directly or from a
// hipGraph (one captured graph per cached program).
#include <string.h>

#include <vector>

#include "../../include/qj_api.h"
#include "qj_common.h"
#include "qj_kernels.h"

using namespace qj;

namespace {

constexpr int kMaxPrograms = 12;

size_t align_up(size_t b, size_t a) { return (b + a - 1) / a * a; }

struct BatchLaunch {
  const void* fn;
  dim3 grid;
  dim3 block;
  size_t off;  // of the group's pack array in the program's device buffer
};

struct Group {
  std::vector<int> members;
  std::vector<BatchLaunch> launches;
};

enum Op { kOpStep = 1, kOpRecord = 2, kOpAdjoint = 3 };

struct Program {
  std::vector<unsigned char> key;
  std::vector<qj_state> post;  // member states after the step
  std::vector<Group> groups;
  void* dpacks = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int exec_mode = -1;
  long long last_use = 0;

  void release_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    exec_mode = -1;
  }
  ~Program() {
    release_graph();
    if (dpacks) (void)hipFree(dpacks);
  }
};

bool same_shape(const std::vector<LaunchRec>& a, const std::vector<LaunchRec>& b) {
  if (a.size() != b.size()) return false;
  for (size_t j = 0; j < a.size(); ++j) {
    const LaunchRec& x = a[j];
    const LaunchRec& y = b[j];
    if (x.fn != y.fn || x.pack.size() != y.pack.size() || x.grid.x != y.grid.x ||
        x.grid.y != y.grid.y || x.grid.z != y.grid.z || x.block.x != y.block.x ||
        x.block.y != y.block.y || x.block.z != y.block.z)
      return false;
  }
  return true;
}

int hip_err(hipError_t e, const char* where) {
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  return set_error(QJ_EHIP, m.c_str());
}

#define QJ_ECHECK(expr, where)                       \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_err(e_, where); \
  } while (0)

}  // namespace

struct qj_ensemble {
  int n = 0;
  int mode = QJ_ENS_GRAPH;
  std::vector<qj_plan*> plans;
  std::vector<qj_state*> states;
  std::vector<qj_params> params;
  std::vector<void*> works;
  std::vector<Program*> progs;
  const Program* last = nullptr;
  long long clock = 0;
  hipStream_t cap = nullptr;            // capture stream (never the null stream)
  std::vector<hipStream_t> branch;      // fork branches of QJ_ENS_GRAPH_FORK
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;

  ~qj_ensemble() {
    for (Program* p : progs) delete p;
    for (hipStream_t s : branch) (void)hipStreamDestroy(s);
    for (hipEvent_t e : joins) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (cap) (void)hipStreamDestroy(cap);
  }

  // Entry key: the operation and its arguments (extra), then every member's
  // state, params and schedule options.
  std::vector<unsigned char> key(const std::vector<unsigned char>& extra = {}) const {
    std::vector<unsigned char> k(extra);
    k.reserve(extra.size() +
              (size_t)n * (sizeof(qj_state) + sizeof(qj_params) + 4 * QJ_OPT_COUNT));
    auto put = [&k](const void* p, size_t b) {
      const unsigned char* c = (const unsigned char*)p;
      k.insert(k.end(), c, c + b);
    };
    for (int m = 0; m < n; ++m) {
      put(states[m], sizeof(qj_state));
      put(&params[m], sizeof(qj_params));
      for (int o = 0; o < QJ_OPT_COUNT; ++o) {
        int v = 0;
        (void)qj_plan_get_option(plans[m], o, &v);
        put(&v, sizeof(v));
      }
    }
    return k;
  }

  // member(m, state copy, stream) issues member m's part of the operation
  template <class F>
  int build(Program* pr, F&& member, hipStream_t s) {
    std::vector<std::vector<LaunchRec>> recs(n);
    pr->post.resize(n);
    for (int m = 0; m < n; ++m) {
      qj_state st = *states[m];
      Recorder R;
      int rc;
      {
        RecordScope scope(&R);
        rc = member(m, &st, s);
      }
      if (rc) return rc;
      recs[m] = std::move(R.recs);
      pr->post[m] = st;
    }
    // group members with launch sequences of the same shape
    for (int m = 0; m < n; ++m) {
      Group* g = nullptr;
      for (Group& c : pr->groups)
        if (same_shape(recs[c.members[0]], recs[m])) {
          g = &c;
          break;
        }
      if (!g) {
        pr->groups.emplace_back();
        g = &pr->groups.back();
      }
      g->members.push_back(m);
    }
    // one pack array per (group, stage)
    size_t bytes = 0;
    for (Group& g : pr->groups) {
      const std::vector<LaunchRec>& r0 = recs[g.members[0]];
      for (const LaunchRec& r : r0) {
        bytes = align_up(bytes, 256);
        g.launches.push_back({r.fn, r.grid, r.block, bytes});
        bytes += r.pack.size() * g.members.size();
      }
    }
    std::vector<unsigned char> host(bytes > 0 ? bytes : 1);
    for (Group& g : pr->groups)
      for (size_t j = 0; j < g.launches.size(); ++j)
        for (size_t i = 0; i < g.members.size(); ++i) {
          const LaunchRec& r = recs[g.members[i]][j];
          memcpy(host.data() + g.launches[j].off + i * r.pack.size(), r.pack.data(),
                 r.pack.size());
        }
    QJ_ECHECK(hipMalloc(&pr->dpacks, host.size()), "ensemble packs");
    QJ_ECHECK(hipMemcpy(pr->dpacks, host.data(), host.size(), hipMemcpyHostToDevice),
              "ensemble packs");
    return QJ_OK;
  }

  template <class F>
  Program* find_or_build(const std::vector<unsigned char>& k, F&& member, hipStream_t s,
                         int* rc) {
    *rc = QJ_OK;
    for (Program* p : progs)
      if (p->key == k) return p;
    if ((int)progs.size() >= kMaxPrograms) {  // drop the least recently used
      size_t lru = 0;
      for (size_t i = 1; i < progs.size(); ++i)
        if (progs[i]->last_use < progs[lru]->last_use) lru = i;
      // its packs may still be read by queued launches
      hipError_t e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        *rc = hip_err(e, "ensemble cache");
        return nullptr;
      }
      delete progs[lru];
      progs.erase(progs.begin() + lru);
    }
    Program* p = new Program();
    p->key = k;
    if ((*rc = build(p, member, s))) {
      delete p;
      return nullptr;
    }
    progs.push_back(p);
    return p;
  }

  int issue_group(const Program* pr, const Group& g, hipStream_t s) {
    for (const BatchLaunch& L : g.launches) {
      void* packs = (char*)pr->dpacks + L.off;
      void* args[] = {&packs};
      dim3 grid(L.grid.x, L.grid.y, (unsigned)g.members.size());
      QJ_ECHECK(hipLaunchKernel(L.fn, grid, L.block, args, 0, s), "ensemble launch");
    }
    return QJ_OK;
  }

  int issue(const Program* pr, hipStream_t s) {
    for (const Group& g : pr->groups)
      if (int rc = issue_group(pr, g, s)) return rc;
    return QJ_OK;
  }

  int capture(Program* pr, int fork_mode) {
    pr->release_graph();
    if (!cap) QJ_ECHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "capture stream");
    const size_t ng = pr->groups.size();
    if (fork_mode) {
      if (!fork) QJ_ECHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "fork event");
      while (branch.size() < ng) {
        hipStream_t b;
        QJ_ECHECK(hipStreamCreateWithFlags(&b, hipStreamNonBlocking), "branch stream");
        branch.push_back(b);
        hipEvent_t ev;
        QJ_ECHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "join event");
        joins.push_back(ev);
      }
    }
    QJ_ECHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture");
    int rc = QJ_OK;
    if (fork_mode && ng > 1) {
      hipError_t e = hipEventRecord(fork, cap);
      for (size_t g = 0; g < ng && e == hipSuccess && !rc; ++g) {
        e = hipStreamWaitEvent(branch[g], fork, 0);
        if (e == hipSuccess) rc = issue_group(pr, pr->groups[g], branch[g]);
        if (e == hipSuccess && !rc) e = hipEventRecord(joins[g], branch[g]);
        if (e == hipSuccess && !rc) e = hipStreamWaitEvent(cap, joins[g], 0);
      }
      if (e != hipSuccess && !rc) rc = hip_err(e, "graph fork");
    } else {
      rc = issue(pr, cap);
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return hip_err(e, "end capture");
    pr->graph = graph;
    QJ_ECHECK(hipGraphInstantiate(&pr->exec, graph, nullptr, nullptr, 0), "graph instantiate");
    pr->exec_mode = fork_mode ? QJ_ENS_GRAPH_FORK : QJ_ENS_GRAPH;
    return QJ_OK;
  }

  // Runs one recorded operation on all members (batched, or from its graph).
  template <class F>
  int run(const std::vector<unsigned char>& extra, F&& member, hipStream_t s) {
    int rc;
    Program* pr = find_or_build(key(extra), member, s, &rc);
    if (!pr) return rc;
    pr->last_use = ++clock;
    if (mode == QJ_ENS_BATCHED) {
      if ((rc = issue(pr, s))) return rc;
    } else {
      if (pr->exec_mode != mode)
        if ((rc = capture(pr, mode == QJ_ENS_GRAPH_FORK))) return rc;
      QJ_ECHECK(hipGraphLaunch(pr->exec, s), "graph launch");
    }
    for (int m = 0; m < n; ++m) *states[m] = pr->post[m];
    last = pr;
    return QJ_OK;
  }

  int step(hipStream_t s) {
    const int op = kOpStep;
    std::vector<unsigned char> extra((const unsigned char*)&op, (const unsigned char*)(&op + 1));
    return run(
        extra,
        [this](int m, qj_state* st, hipStream_t ss) {
          return qj_kdk_steps(plans[m], st, &params[m], 1, works[m], ss);
        },
        s);
  }
};

namespace {

void put_bytes(std::vector<unsigned char>& k, const void* p, size_t b) {
  const unsigned char* c = (const unsigned char*)p;
  k.insert(k.end(), c, c + b);
}

}  // namespace

extern "C" {

int qj_ensemble_create(qj_ensemble** out, int nmembers, qj_plan* const* plans,
                       qj_state* const* states, const qj_params* params, void* const* works) {
  if (!out) return set_error(QJ_EARG, "qj_ensemble_create: null output pointer");
  *out = nullptr;
  if (nmembers <= 0 || !plans || !states || !params || !works)
    return set_error(QJ_EARG, "qj_ensemble_create: bad arguments");
  for (int m = 0; m < nmembers; ++m) {
    if (!plans[m] || !states[m] || !works[m])
      return set_error(QJ_EARG, "qj_ensemble_create: null member");
    for (int k = 0; k < m; ++k)
      if (plans[k] == plans[m] || states[k] == states[m] || works[k] == works[m])
        return set_error(QJ_EARG,
                         "qj_ensemble_create: members must not share plans, states or "
                         "workspaces (their launches run concurrently)");
  }
  qj_ensemble* e = new qj_ensemble();
  e->n = nmembers;
  e->plans.assign(plans, plans + nmembers);
  e->states.assign(states, states + nmembers);
  e->params.assign(params, params + nmembers);
  e->works.assign(works, works + nmembers);
  *out = e;
  return QJ_OK;
}

int qj_ensemble_destroy(qj_ensemble* e) {
  delete e;
  return QJ_OK;
}

int qj_ensemble_set_mode(qj_ensemble* e, int mode) {
  if (!e) return set_error(QJ_EARG, "null ensemble");
  if (mode < QJ_ENS_EAGER || mode > QJ_ENS_GRAPH_FORK)
    return set_error(QJ_EARG, "qj_ensemble_set_mode: unknown mode");
  e->mode = mode;
  return QJ_OK;
}

int qj_ensemble_steps(qj_ensemble* e, int nsteps, void* stream) {
  if (!e) return set_error(QJ_EARG, "null ensemble");
  if (nsteps < 0) return set_error(QJ_EARG, "qj_ensemble_steps: nsteps < 0");
  hipStream_t s = (hipStream_t)stream;
  if (e->mode == QJ_ENS_EAGER) {
    for (int m = 0; m < e->n; ++m)
      if (int rc = qj_kdk_steps(e->plans[m], e->states[m], &e->params[m], nsteps, e->works[m], s))
        return rc;
    return QJ_OK;
  }
  for (int k = 0; k < nsteps; ++k)
    if (int rc = e->step(s)) return rc;
  return QJ_OK;
}

int qj_ensemble_record(qj_ensemble* e, int nsteps, qj_tape_step* const* tapes,
                       void* const* grid0s, void* stream) {
  if (!e) return set_error(QJ_EARG, "null ensemble");
  if (nsteps < 0 || (nsteps > 0 && !tapes))
    return set_error(QJ_EARG, "qj_ensemble_record: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  auto member = [e, nsteps, tapes, grid0s](int m, qj_state* st, hipStream_t ss) {
    return qj_kdk_record(e->plans[m], st, &e->params[m], nsteps, tapes[m],
                         grid0s ? grid0s[m] : nullptr, e->works[m], ss);
  };
  if (e->mode == QJ_ENS_EAGER) {
    for (int m = 0; m < e->n; ++m)
      if (int rc = member(m, e->states[m], s)) return rc;
    return QJ_OK;
  }
  std::vector<unsigned char> extra;
  const int op = kOpRecord;
  put_bytes(extra, &op, sizeof(op));
  put_bytes(extra, &nsteps, sizeof(nsteps));
  for (int m = 0; m < e->n; ++m) {
    if (!tapes[m]) return set_error(QJ_EARG, "qj_ensemble_record: null tape");
    put_bytes(extra, tapes[m], sizeof(qj_tape_step) * (size_t)nsteps);
    const void* g0 = grid0s ? grid0s[m] : nullptr;
    put_bytes(extra, &g0, sizeof(g0));
  }
  return e->run(extra, member, s);
}

int qj_ensemble_adjoint(qj_ensemble* e, int nsteps, qj_tape_step* const* tapes,
                        void* const* grid0s, qj_adjoint* adjs, void* stream) {
  if (!e) return set_error(QJ_EARG, "null ensemble");
  if (nsteps < 0 || !adjs || (nsteps > 0 && !tapes))
    return set_error(QJ_EARG, "qj_ensemble_adjoint: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  auto member = [e, nsteps, tapes, grid0s, adjs](int m, qj_state* st, hipStream_t ss) {
    return qj_kdk_adjoint(e->plans[m], st, &e->params[m], nsteps, tapes[m],
                          grid0s ? grid0s[m] : nullptr, &adjs[m], e->works[m], ss);
  };
  if (e->mode == QJ_ENS_EAGER) {
    for (int m = 0; m < e->n; ++m)
      if (int rc = member(m, e->states[m], s)) return rc;
    return QJ_OK;
  }
  std::vector<unsigned char> extra;
  const int op = kOpAdjoint;
  put_bytes(extra, &op, sizeof(op));
  put_bytes(extra, &nsteps, sizeof(nsteps));
  for (int m = 0; m < e->n; ++m) {
    if (!tapes[m]) return set_error(QJ_EARG, "qj_ensemble_adjoint: null tape");
    put_bytes(extra, tapes[m], sizeof(qj_tape_step) * (size_t)nsteps);
    const void* g0 = grid0s ? grid0s[m] : nullptr;
    put_bytes(extra, &g0, sizeof(g0));
    put_bytes(extra, &adjs[m], sizeof(qj_adjoint));
  }
  return e->run(extra, member, s);
}

int qj_ensemble_finalize(qj_ensemble* e, void* stream) {
  if (!e) return set_error(QJ_EARG, "null ensemble");
  for (int m = 0; m < e->n; ++m)
    if (int rc = qj_kdk_finalize(e->plans[m], e->states[m], &e->params[m], stream)) return rc;
  return QJ_OK;
}

int qj_ensemble_group_count(const qj_ensemble* e) {
  return (e && e->last) ? (int)e->last->groups.size() : 0;
}

int qj_ensemble_group_info(const qj_ensemble* e, int group, int* members, int cap,
                           int* nmembers, int* launches) {
  if (!e || !e->last || group < 0 || group >= (int)e->last->groups.size())
    return set_error(QJ_EARG, "qj_ensemble_group_info: no such group");
  const Group& g = e->last->groups[group];
  for (int i = 0; i < cap && i < (int)g.members.size(); ++i)
    if (members) members[i] = g.members[i];
  if (nmembers) *nmembers = (int)g.members.size();
  if (launches) *launches = (int)g.launches.size();
  return QJ_OK;
}

}  // extern "C"
