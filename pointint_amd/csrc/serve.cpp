// Serving engine: continuous batching of volume-matrix queries over resident
// geometry on one GPU.
//
// A query is a scene of n resident geometries, each under its own affine
// (rigid) transform; its answer is the n x n volume matrix.  Queries are
// admitted against an HBM budget (each holds its memory reservation from
// admission to completion; a query that does not fit is refused at once).
// A worker thread keeps up to kSlots batches in flight, each on its own HIP
// stream with its own workspaces: whenever a slot is free it takes every
// queued query that fits the slot's object limit, launches one fused
// form-factor pass over all objects of all those queries, then one Gram per
// query on its column range, and copies the matrices to pinned host memory.
// Queries that arrive while batches run join the next free slot, so the work
// of consecutive batches overlaps on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pii_capi_internal.hpp"

namespace pii {

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  int64_t id = 0;
  int n = 0;
  std::vector<const GeometryDev*> geoms;
  std::vector<double> dofs;   // n x 12
  size_t footprint = 0;
  double t_submit = 0, t_launch = 0, t_done = 0;
  std::vector<double> V;
  int state = 0;              // 0 queued, 1 running, 2 done
  int slot = -1;
  size_t v_off = 0;           // offset of V in the slot's pinned buffer
};

constexpr int kSlots = 2;

struct Slot {
  Workspace w;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Request*> reqs;
  Batch batch;
  double* pinned = nullptr;
  size_t pinned_n = 0;
  DevBuf Vdev;
};

}  // namespace

struct Server {
  int device = 0;
  const KGridDev* kg = nullptr;
  int gram_mode = PII_GRAM_FP16X3;
  bool series = true;                 // moment series (the creating context's setting)
  size_t budget = 0;
  size_t reserved = 0;
  int max_batch_objects = 1024;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  Slot slots[kSlots];
  // statistics
  int64_t n_submitted = 0, n_admitted = 0, n_refused = 0, n_completed = 0, n_batches = 0;
  int64_t n_objects = 0;
  double busy_since = 0;
  std::string worker_error;

  size_t footprint(int n, const std::vector<const GeometryDev*>& gs) const {
    const size_t ldX = align_up((size_t)kg->Qh, 64);
    const size_t ncol = align_up((size_t)n, 16);
    size_t b = ncol * ldX * sizeof(float2);                   // X columns
    if (gram_mode == PII_GRAM_FP64) {
      b += (size_t)n * n * sizeof(double) * (gram_fp64_splits(kg->Qh, n) + 1);
    } else {
      const int fp16 = gram_mode == PII_GRAM_FP16X3 ? 1 : 0;
      b += gram_plane_bytes(fp16, (int)(2 * ldX), (int)ncol);  // 16-bit planes
      b += (size_t)n * n * sizeof(double) *
           (gram_query(0, n, (int)(2 * ldX)).nslice + 1);       // partials + V
    }
    for (const GeometryDev* g : gs) {
      b += (size_t)g->n * sizeof(double) * kRecFastTri;       // element records
      if (series) b += (size_t)g->n * sizeof(double) * kRecMoment;  // moment records
      b += (size_t)g->nv * 3 * sizeof(double);                // world vertices
    }
    return b + 4096;
  }

  void launch(Slot& s) {
    // Concatenate the queries' objects into one batch.
    std::vector<const GeometryDev*> gs;
    std::vector<const double*> dp;
    std::vector<int> kinds, nd;
    for (Request* r : s.reqs) {
      for (int i = 0; i < r->n; ++i) {
        gs.push_back(r->geoms[i]);
        dp.push_back(r->dofs.data() + 12 * (size_t)i);
        kinds.push_back(PII_DOF_AFFINE);
        nd.push_back(12);
      }
    }
    const int N = (int)gs.size();
    // Each query's columns start at a multiple of 16 (the Gram's column groups).
    std::vector<int> cols(N);
    std::vector<int> c0s;
    int ncol = 0;
    {
      int k = 0;
      for (Request* r : s.reqs) {
        c0s.push_back(ncol);
        for (int i = 0; i < r->n; ++i) cols[k++] = ncol + i;
        ncol += (int)align_up((size_t)r->n, 16);
      }
    }
    s.batch = Batch();
    s.batch.build(s.w, s.stream, nullptr, N, gs.data(), kinds.data(), dp.data(), nd.data(), true,
                  series ? kg : nullptr, cols.data());
    const int Qh = kg->Qh;
    const int ldX = (int)align_up((size_t)Qh, 64);
    s.w.X.ensure(sizeof(float2) * (size_t)ldX * ncol);
    s.w.colmax.ensure(sizeof(unsigned) * (size_t)ncol);
    unsigned* colmax = s.w.colmax.as<unsigned>();
    PII_HIP(hipMemsetAsync(colmax, 0, sizeof(unsigned) * ncol, s.stream));
    launch_form_factors_fast(s.stream, s.batch, *kg, 1, nullptr, 0, s.w.X.as<float2>(), ldX,
                             colmax, s.w, nullptr);
    PII_HIP(hipGetLastError());
    size_t vtot = 0, maxpart = 0;
    for (Request* r : s.reqs) {
      r->v_off = vtot;
      vtot += (size_t)r->n * r->n;
      if (gram_mode == PII_GRAM_FP64)
        maxpart = std::max(maxpart, (size_t)gram_fp64_splits(Qh, r->n) * r->n * r->n);
    }
    s.Vdev.ensure(sizeof(double) * std::max<size_t>(vtot, 1));
    if (maxpart) s.w.part.ensure(sizeof(double) * maxpart);
    if (s.pinned_n < vtot) {
      if (s.pinned) (void)hipHostFree(s.pinned);
      s.pinned = nullptr;
      s.pinned_n = 0;
      const size_t want = std::max<size_t>(vtot, 4096);
      PII_HIP(hipHostMalloc((void**)&s.pinned, sizeof(double) * want, hipHostMallocDefault));
      s.pinned_n = want;
    }
    if (gram_mode == PII_GRAM_FP64) {
      for (size_t qi = 0; qi < s.reqs.size(); ++qi) {
        Request* r = s.reqs[qi];
        const float2* Xr = s.w.X.as<float2>() + (size_t)c0s[qi] * ldX;
        launch_gram_fp64_x(s.stream, Xr, Qh, r->n, ldX, s.w.part.as<double>(),
                           gram_fp64_splits(Qh, r->n), s.Vdev.as<double>() + r->v_off);
        PII_HIP(hipGetLastError());
      }
    } else {
      // one launch for the Grams of all queries of the batch
      std::vector<GramQuery> qs;
      for (size_t qi = 0; qi < s.reqs.size(); ++qi)
        qs.push_back(gram_query(c0s[qi], s.reqs[qi]->n, 2 * ldX));
      gram_split_run(s.stream, s.w, gram_mode == PII_GRAM_FP16X3 ? 1 : 0, s.w.X.as<float2>(),
                     ldX, ncol, colmax, qs, s.Vdev.as<double>());
      PII_HIP(hipGetLastError());
    }
    PII_HIP(hipMemcpyAsync(s.pinned, s.Vdev.p, sizeof(double) * vtot, hipMemcpyDeviceToHost,
                           s.stream));
    PII_HIP(hipEventRecord(s.done, s.stream));
    s.busy = true;
  }

  void retire(Slot& s) {
    const double t = now_s();
    for (Request* r : s.reqs) {
      r->V.assign(s.pinned + r->v_off, s.pinned + r->v_off + (size_t)r->n * r->n);
      r->t_done = t;
      r->state = 2;
      reserved -= r->footprint;
      ++n_completed;
    }
    s.reqs.clear();
    s.busy = false;
  }

  void run() {
    if (hipSetDevice(device) != hipSuccess) {
      std::lock_guard<std::mutex> lk(mu);
      worker_error = "hipSetDevice failed in the serving worker";
      return;
    }
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      bool progressed = false;
      for (Slot& s : slots) {
        if (s.busy && hipEventQuery(s.done) == hipSuccess) {
          retire(s);
          progressed = true;
          cv_done.notify_all();
        }
      }
      if (stop && queue.empty() && !slots[0].busy && !slots[1].busy) break;
      for (Slot& s : slots) {
        if (s.busy || queue.empty()) continue;
        int nobj = 0;
        while (!queue.empty()) {
          Request* r = queue.front();
          if (!s.reqs.empty() && nobj + r->n > max_batch_objects) break;
          queue.pop_front();
          s.reqs.push_back(r);
          r->state = 1;
          r->t_launch = now_s();
          nobj += r->n;
        }
        ++n_batches;
        n_objects += nobj;
        try {
          launch(s);
        } catch (const Error& e) {
          worker_error = e.msg;
          for (Request* r : s.reqs) {
            r->state = 2;
            r->t_done = now_s();
            reserved -= r->footprint;
          }
          s.reqs.clear();
          s.busy = false;
          cv_done.notify_all();
        }
        progressed = true;
      }
      if (!progressed) {
        const bool any_busy = slots[0].busy || slots[1].busy;
        if (any_busy) {
          // poll the in-flight batches without holding the lock
          lk.unlock();
          std::this_thread::sleep_for(std::chrono::microseconds(50));
          lk.lock();
        } else {
          cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
      }
    }
  }
};

}  // namespace pii

struct pii_server {
  pii::Server s;
};

using namespace pii;

extern "C" {

int pii_server_create(pii_context* ctx, const pii_kgrid* kg, int gram_mode, double hbm_budget,
                      int max_batch_objects, pii_server** out) {
  return guarded([&] {
    if (!ctx || !kg || !out || hbm_budget <= 0 || max_batch_objects <= 0)
      throw Error(PII_ERR_INVALID, "pii_server_create: bad arguments");
    if (gram_mode != PII_GRAM_FP64 && gram_mode != PII_GRAM_BF16X3 && gram_mode != PII_GRAM_FP16X3)
      throw Error(PII_ERR_INVALID, "pii_server_create: unknown Gram mode");
    PII_HIP(hipSetDevice(ctx->c.device));
    auto srv = std::make_unique<pii_server>();
    Server& S = srv->s;
    S.device = ctx->c.device;
    S.kg = &kg->k;
    S.series = ctx->c.series;
    S.gram_mode = gram_mode;
    S.budget = (size_t)hbm_budget;
    S.max_batch_objects = max_batch_objects;
    for (auto& sl : S.slots) {
      PII_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
      PII_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    S.worker = std::thread([p = &S] { p->run(); });
    *out = srv.release();
  });
}

int pii_server_destroy(pii_server* srv) {
  return guarded([&] {
    if (!srv) return;
    Server& S = srv->s;
    {
      std::lock_guard<std::mutex> lk(S.mu);
      S.stop = true;
    }
    S.cv_work.notify_all();
    if (S.worker.joinable()) S.worker.join();
    (void)hipSetDevice(S.device);
    for (auto& sl : S.slots) {
      if (sl.stream) (void)hipStreamSynchronize(sl.stream);
      if (sl.pinned) (void)hipHostFree(sl.pinned);
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete srv;
  });
}

int pii_server_submit(pii_server* srv, int n, const pii_geometry* const* geoms,
                      const double* affine_dofs, int64_t* id) {
  return guarded([&] {
    if (!srv || !geoms || !affine_dofs || !id || n <= 0)
      throw Error(PII_ERR_INVALID, "pii_server_submit: bad arguments");
    Server& S = srv->s;
    auto r = std::make_unique<Request>();
    r->n = n;
    r->geoms.resize(n);
    for (int i = 0; i < n; ++i) {
      if (!geoms[i]) throw Error(PII_ERR_INVALID, "pii_server_submit: NULL geometry");
      if (geoms[i]->g.type != PII_GEOM_TRIANGLE)
        throw Error(PII_ERR_UNSUPPORTED, "pii_server_submit: triangle geometry only");
      r->geoms[i] = &geoms[i]->g;
    }
    r->dofs.assign(affine_dofs, affine_dofs + 12 * (size_t)n);
    r->footprint = S.footprint(n, r->geoms);
    r->t_submit = now_s();
    std::lock_guard<std::mutex> lk(S.mu);
    ++S.n_submitted;
    if (!S.worker_error.empty()) throw Error(PII_ERR_HIP, "serving worker failed: " + S.worker_error);
    if (n > S.max_batch_objects || S.reserved + r->footprint > S.budget) {
      ++S.n_refused;
      throw Error(PII_ERR_BUDGET, "query refused: HBM budget (" + std::to_string(S.reserved) +
                                      " + " + std::to_string(r->footprint) + " > " +
                                      std::to_string(S.budget) + " bytes)");
    }
    ++S.n_admitted;
    S.reserved += r->footprint;
    r->id = S.next_id++;
    *id = r->id;
    S.queue.push_back(r.get());
    S.reqs[r->id] = std::move(r);
    S.cv_work.notify_one();
  });
}

int pii_server_wait(pii_server* srv, int64_t id, double timeout_s, double* V, double* latency_ms) {
  return guarded([&] {
    if (!srv) throw Error(PII_ERR_INVALID, "NULL server");
    Server& S = srv->s;
    std::unique_lock<std::mutex> lk(S.mu);
    auto it = S.reqs.find(id);
    if (it == S.reqs.end()) throw Error(PII_ERR_INVALID, "unknown query id " + std::to_string(id));
    Request* r = it->second.get();
    const auto deadline = std::chrono::steady_clock::now() +
                          std::chrono::duration<double>(timeout_s < 0 ? 1e9 : timeout_s);
    while (r->state != 2) {
      if (S.cv_done.wait_until(lk, deadline) == std::cv_status::timeout && r->state != 2)
        throw Error(PII_ERR_BUSY, "query " + std::to_string(id) + " not finished");
    }
    if (r->V.size() != (size_t)r->n * r->n)
      throw Error(PII_ERR_HIP, "query " + std::to_string(id) + " failed: " + S.worker_error);
    if (V) std::memcpy(V, r->V.data(), sizeof(double) * r->V.size());
    if (latency_ms) {
      latency_ms[0] = (r->t_done - r->t_submit) * 1e3;
    }
    S.reqs.erase(it);
  });
}

int pii_server_stats(pii_server* srv, double* out, int n) {
  return guarded([&] {
    if (!srv || !out) throw Error(PII_ERR_INVALID, "pii_server_stats: bad arguments");
    Server& S = srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    const double v[8] = {(double)S.n_submitted, (double)S.n_admitted, (double)S.n_refused,
                         (double)S.n_completed, (double)S.n_batches, (double)S.n_objects,
                         (double)S.reserved, (double)S.budget};
    for (int i = 0; i < n; ++i) out[i] = i < 8 ? v[i] : 0.0;
  });
}

}  // extern "C"
