// sph_server.cpp -- the serving form of the SPHENIX pass: a request server
// that accepts density+force step requests for independent periodic boxes
// while others are running, admits them by an HBM budget and a slot count,
// runs each admitted request on its own engine and HIP stream (so the work
// of concurrent requests overlaps on the device), and keeps every result on
// the host until the caller fetches it.  Requests are served first come,
// first served; a request waits at the head of the queue until enough HBM
// and a slot are free (continuous batching: a slot freed by one request is
// taken by the next at once).  Boxes submitted with `keep` stay resident on
// the device and can be stepped again without a new upload.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/swsph.h"

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch())
      .count();
}

// Device bytes of an engine for n particles on ncells cells: particle
// arrays, records, state copy, tile scratch and binning buffers (~600 B per
// particle), chunk tables and cell arrays, plus a fixed part.
size_t engine_bytes(int64_t n, int64_t ncells) {
  return (size_t)n * 600u + (size_t)ncells * 64u + ((size_t)4 << 20);
}

struct Box {
  int64_t id = 0;
  swsph_engine *eng = nullptr;
  int64_t n = 0;
  size_t bytes = 0;
  bool busy = false;
};

// Page-locked host blocks, reused from request to request: the host copies
// of submitted particles and the results live in them, so that uploads and
// result copies are DMA transfers that do not wait on the host.  A block is
// allocated (hipHostMalloc, rounded up to a power of two of at least 1 MiB)
// only when no free one is large enough.
struct PinnedPool {
  std::mutex m;
  std::multimap<size_t, unsigned char *> free;
  std::vector<unsigned char *> all;
  unsigned char *get(size_t bytes, size_t *cap) {
    std::lock_guard<std::mutex> lk(m);
    auto it = free.lower_bound(bytes);
    if (it != free.end()) {
      unsigned char *p = it->second;
      *cap = it->first;
      free.erase(it);
      return p;
    }
    size_t c = (size_t)1 << 20;
    while (c < bytes) c <<= 1;
    void *p = nullptr;
    if (hipHostMalloc(&p, c, 0) != hipSuccess) return nullptr;
    all.push_back((unsigned char *)p);
    *cap = c;
    return (unsigned char *)p;
  }
  void put(unsigned char *p, size_t cap) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(m);
    free.emplace(cap, p);
  }
  ~PinnedPool() {
    for (unsigned char *p : all) (void)hipHostFree(p);
  }
};

struct Request {
  int64_t id = 0;
  int64_t box = 0;        // resident box to step (0: upload from `parts`)
  bool keep = false;      // keep the box resident after the step
  swsph_props props{};
  int64_t n = 0;
  // host copy of the submitted particles: arrays in one pinned block
  unsigned char *in = nullptr;
  size_t in_cap = 0;
  swsph_host_parts parts{};
  size_t bytes = 0;
  // results: the fetched fields in one pinned block, at off[field]
  unsigned char *out = nullptr;
  size_t out_cap = 0;
  int64_t off[SWSPH_NUM_FIELDS];
  swsph_request_info info{};
  bool done = false;
  std::string err;
};

// Bytes per element of each swsph_field in the caller's order.
size_t field_bytes(int f) {
  switch (f) {
    case SWSPH_ROT_V:
    case SWSPH_A_HYDRO:
      return 12;
    case SWSPH_N_DENSITY:
    case SWSPH_N_GRADIENT:
    case SWSPH_N_FORCE:
      return 4;
    case SWSPH_MIN_NGB_TIME_BIN:
      return 1;
    default:
      return 4;
  }
}

}  // namespace

struct swsph_server {
  swsph_server_opts opts{};
  PinnedPool pool;
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<int64_t, std::shared_ptr<Request>> reqs;
  std::map<int64_t, std::shared_ptr<Box>> boxes;
  size_t bytes_in_use = 0;
  int inflight = 0;
  int64_t next_id = 1;
  bool stop = false;
  std::vector<std::thread> workers;
};

namespace {

thread_local std::string s_err;

int sfail(int code, const std::string &msg) {
  s_err = msg;
  return code;
}

// Runs one request on a resident box's engine, or on the worker's pooled
// engine `idle` (re-gridded, its buffers reused when large enough); fills
// the request's results.
void serve(swsph_server *s, const std::shared_ptr<Request> &r,
           std::shared_ptr<Box> box, swsph_engine *&idle) {
  auto fail_req = [&](const char *what) {
    r->err = std::string(what) + ": " + swsph_last_error();
  };
  r->info.t_start_ms = now_ms();
  swsph_engine *e = box ? box->eng : nullptr;
  if (!e) {
    if (idle) {
      e = idle;
      idle = nullptr;
      if (swsph_set_props(e, &r->props)) {
        swsph_engine_destroy(e);
        return fail_req("set_props");
      }
    } else if (swsph_engine_create(&r->props, &e)) {
      return fail_req("engine_create");
    }
    const swsph_host_parts hp = r->parts;
    const int up = swsph_set_particles(e, &hp);
    // the upload has completed: the input block goes back to the pool
    s->pool.put(r->in, r->in_cap);
    r->in = nullptr;
    if (up) {
      fail_req("set_particles");
      swsph_engine_destroy(e);
      return;
    }
    if (r->keep) {
      box = std::make_shared<Box>();
      box->eng = e;
      box->n = r->n;
      box->bytes = r->bytes;
    }
  }
  r->info.t_uploaded_ms = now_ms();
  swsph_pass_stats st{};
  const int rc = s->opts.use_graphs ? swsph_pass_graph(e, &st)
                                    : swsph_pass_batched(e, &st);
  r->info.t_passed_ms = now_ms();
  if (rc) {
    fail_req("pass");
  } else {
    // every requested field in one pinned block, fetched with one wait
    const unsigned int mask = s->opts.fields ? s->opts.fields : ~0u;
    std::vector<int> fl;
    std::vector<void *> outs;
    size_t total = 0;
    for (int f = 0; f < SWSPH_NUM_FIELDS; f++) {
      r->off[f] = -1;
      if (!(mask >> f & 1u)) continue;
      r->off[f] = (int64_t)total;
      total += ((size_t)r->n * field_bytes(f) + 255) & ~(size_t)255;
      fl.push_back(f);
    }
    r->out = s->pool.get(std::max<size_t>(total, 1), &r->out_cap);
    if (!r->out) {
      r->err = "page-locked result buffer: allocation failed";
    } else {
      for (int f : fl) outs.push_back(r->out + r->off[f]);
      if (r->n && swsph_get_fields(e, (int)fl.size(), fl.data(), outs.data()))
        fail_req("get_fields");
    }
    r->info.ghost_rounds = st.ghost_rounds;
    r->info.n_density = st.n_density;
    r->info.n_force = st.n_force;
  }
  if (box) {
    std::lock_guard<std::mutex> lk(s->m);
    if (!s->boxes.count(box->id)) {
      box->id = r->id;  // a new resident box takes the request's id
      s->boxes[box->id] = box;
    }
    r->info.box = box->id;
  } else {
    idle = e;  // kept for the worker's next request
  }
}

void worker(swsph_server *s) {
  swsph_engine *idle = nullptr;  // this worker's pooled engine
  struct Pool {
    swsph_engine *&e;
    ~Pool() {
      if (e) swsph_engine_destroy(e);
    }
  } pool{idle};
  for (;;) {
    std::shared_ptr<Request> r;
    std::shared_ptr<Box> box;
    {
      std::unique_lock<std::mutex> lk(s->m);
      // Admission: the head request runs when a slot is free and its bytes
      // fit the budget (or nothing else runs, so that any request can run).
      s->cv_work.wait(lk, [&] {
        if (s->stop) return true;
        // a step of a box that no longer exists fails at once instead of
        // blocking the queue behind it
        while (!s->queue.empty() && s->queue.front()->box &&
               s->boxes.find(s->queue.front()->box) == s->boxes.end()) {
          auto dead = s->queue.front();
          s->queue.pop_front();
          dead->err = "resident box was closed";
          dead->info.t_done_ms = now_ms();
          dead->done = true;
          s->cv_done.notify_all();
        }
        if (s->queue.empty() || s->inflight >= s->opts.max_inflight)
          return false;
        const auto &h = s->queue.front();
        if (h->box) {
          auto it = s->boxes.find(h->box);
          return it != s->boxes.end() && !it->second->busy;
        }
        return s->inflight == 0 ||
               s->bytes_in_use + h->bytes <= (size_t)s->opts.hbm_budget_bytes;
      });
      if (s->stop) return;
      r = s->queue.front();
      s->queue.pop_front();
      if (r->box) {
        box = s->boxes[r->box];
        box->busy = true;
      } else {
        s->bytes_in_use += r->bytes;
      }
      s->inflight++;
      r->info.t_admit_ms = now_ms();
    }
    serve(s, r, box, idle);
    {
      std::lock_guard<std::mutex> lk(s->m);
      if (box) {
        box->busy = false;
      } else if (!(r->keep && r->info.box)) {
        s->bytes_in_use -= r->bytes;  // a kept box holds its bytes
      }
      s->inflight--;
      r->info.t_done_ms = now_ms();
      r->done = true;
      // the request's host copy of the particles is no longer needed
      s->pool.put(r->in, r->in_cap);
      r->in = nullptr;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

}  // namespace

extern "C" {

const char *swsph_server_last_error(void) { return s_err.c_str(); }

int swsph_server_create(const swsph_server_opts *opts, swsph_server **out) {
  if (!opts || !out) return sfail(1, "null argument");
  if (opts->max_inflight < 1 || opts->max_inflight > 64)
    return sfail(1, "max_inflight outside [1, 64]");
  swsph_server *s = new swsph_server;
  s->opts = *opts;
  if (s->opts.hbm_budget_bytes <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete s;
      return sfail(2, "hipMemGetInfo failed");
    }
    s->opts.hbm_budget_bytes = 0.8 * (double)fr;
  }
  for (int k = 0; k < s->opts.max_inflight; k++)
    s->workers.emplace_back(worker, s);
  *out = s;
  return 0;
}

int swsph_server_destroy(swsph_server *s) {
  if (!s) return 0;
  {
    std::unique_lock<std::mutex> lk(s->m);
    // finish what was admitted; fail what still waits
    s->cv_done.wait(lk, [&] { return s->inflight == 0; });
    s->stop = true;
    for (auto &r : s->queue) {
      r->err = "server destroyed before the request was admitted";
      r->done = true;
    }
    s->queue.clear();
  }
  s->cv_done.notify_all();
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto &b : s->boxes) swsph_engine_destroy(b.second->eng);
  delete s;
  return 0;
}

int swsph_server_submit(swsph_server *s, const swsph_props *props,
                        const swsph_host_parts *p, int keep, int64_t *id) {
  if (!s || !props || !p || !id) return sfail(1, "null argument");
  if (p->n < 0 || !p->x || !p->v || !p->m || !p->u || !p->time_bin || !p->h)
    return sfail(1, "missing particle arrays");
  auto r = std::make_shared<Request>();
  r->props = *props;
  r->keep = keep != 0;
  const int64_t n = p->n;
  r->n = n;
  // the arrays, 256-byte aligned, in one pinned block
  const size_t un = (size_t)n;
  const size_t sizes[14] = {24 * un, 12 * un, 4 * un, 4 * un, un, 4 * un,
                            4 * un, 4 * un, 4 * un, 4 * un, 4 * un, 4 * un,
                            4 * un, 4 * un};
  size_t offs[14], total = 0;
  for (int k = 0; k < 14; k++) {
    offs[k] = total;
    total += (sizes[k] + 255) & ~(size_t)255;
  }
  r->in = s->pool.get(std::max<size_t>(total, 1), &r->in_cap);
  if (!r->in) return sfail(2, "page-locked request buffer: allocation failed");
  unsigned char *b = r->in;
  auto put = [&](int k, const void *src) {
    if (sizes[k]) std::memcpy(b + offs[k], src, sizes[k]);
    return (const void *)(b + offs[k]);
  };
  auto putf = [&](int k, const float *src, float dflt) {
    float *dst = (float *)(b + offs[k]);
    if (src) {
      if (sizes[k]) std::memcpy(dst, src, sizes[k]);
    } else {
      for (int64_t i = 0; i < n; i++) dst[i] = dflt;
    }
    return (const float *)dst;
  };
  swsph_host_parts &hp = r->parts;
  hp.n = n;
  hp.x = (const double *)put(0, p->x);
  hp.v = (const float *)put(1, p->v);
  hp.m = (const float *)put(2, p->m);
  hp.u = (const float *)put(3, p->u);
  hp.time_bin = (const int8_t *)put(4, p->time_bin);
  hp.h = (const float *)put(5, p->h);
  hp.rho = putf(6, p->rho, 0.f);
  hp.pressure = putf(7, p->pressure, 0.f);
  hp.soundspeed = putf(8, p->soundspeed, 0.f);
  hp.f = putf(9, p->f, 0.f);
  hp.balsara = putf(10, p->balsara, 0.f);
  hp.alpha_visc = putf(11, p->alpha_visc, 0.1f);
  hp.alpha_diff = putf(12, p->alpha_diff, 0.f);
  hp.div_v_prev = putf(13, p->div_v_prev, 0.f);
  const int64_t ncells =
      (int64_t)props->ncell[0] * props->ncell[1] * props->ncell[2];
  r->bytes = engine_bytes(n, ncells);
  {
    std::lock_guard<std::mutex> lk(s->m);
    r->id = s->next_id++;
    r->info.t_submit_ms = now_ms();
    s->reqs[r->id] = r;
    s->queue.push_back(r);
  }
  s->cv_work.notify_all();
  *id = r->id;
  return 0;
}

int swsph_server_step(swsph_server *s, int64_t box, int64_t *id) {
  if (!s || !id) return sfail(1, "null argument");
  auto r = std::make_shared<Request>();
  {
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->boxes.find(box);
    if (it == s->boxes.end()) return sfail(1, "no resident box with that id");
    r->box = box;
    r->keep = true;
    r->n = it->second->n;
    r->id = s->next_id++;
    r->info.t_submit_ms = now_ms();
    s->reqs[r->id] = r;
    s->queue.push_back(r);
  }
  s->cv_work.notify_all();
  *id = r->id;
  return 0;
}

int swsph_server_wait(swsph_server *s, int64_t id, swsph_request_info *info) {
  if (!s) return sfail(1, "null argument");
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return sfail(1, "unknown request id");
    r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
  }
  if (info) *info = r->info;
  if (!r->err.empty()) return sfail(3, r->err);
  return 0;
}

int swsph_server_get(swsph_server *s, int64_t id, int field, void *host_out) {
  if (!s || !host_out) return sfail(1, "null argument");
  if (field < 0 || field >= SWSPH_NUM_FIELDS) return sfail(1, "bad field");
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return sfail(1, "unknown request id");
    r = it->second;
    if (!r->done) return sfail(1, "request not finished");
  }
  if (!r->err.empty()) return sfail(3, r->err);
  if (!r->out || r->off[field] < 0)
    return sfail(1, "field not fetched (server option `fields`)");
  std::memcpy(host_out, r->out + r->off[field],
              (size_t)r->n * field_bytes(field));
  return 0;
}

int swsph_server_release(swsph_server *s, int64_t id) {
  if (!s) return sfail(1, "null argument");
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sfail(1, "unknown request id");
  if (!it->second->done) return sfail(1, "request not finished");
  s->pool.put(it->second->out, it->second->out_cap);
  it->second->out = nullptr;
  s->reqs.erase(it);
  return 0;
}

int swsph_server_close_box(swsph_server *s, int64_t box) {
  if (!s) return sfail(1, "null argument");
  std::shared_ptr<Box> b;
  {
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->boxes.find(box);
    if (it == s->boxes.end()) return sfail(1, "no resident box with that id");
    if (it->second->busy) return sfail(1, "box is being stepped");
    for (const auto &q : s->queue)
      if (q->box == box) return sfail(1, "box has queued steps");
    b = it->second;
    s->boxes.erase(it);
    s->bytes_in_use -= b->bytes;
  }
  swsph_engine_destroy(b->eng);
  s->cv_work.notify_all();
  return 0;
}

int swsph_server_stats(swsph_server *s, double *bytes_in_use, int *inflight,
                       int *queued) {
  if (!s) return sfail(1, "null argument");
  std::lock_guard<std::mutex> lk(s->m);
  if (bytes_in_use) *bytes_in_use = (double)s->bytes_in_use;
  if (inflight) *inflight = s->inflight;
  if (queued) *queued = (int)s->queue.size();
  return 0;
}

}  // extern "C"
