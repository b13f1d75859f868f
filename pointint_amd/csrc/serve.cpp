// Serving engine: continuous batching of volume-matrix queries over resident
// geometry on one GPU.
//
// A query is a scene of n resident triangle geometries, each under its own
// affine (rigid) transform; its answer is the n x n volume matrix.
//
// Admission.  A query is refused at once (PII_ERR_BUDGET) when its HBM
// footprint does not fit the budget left by the queries admitted and not yet
// finished, or -- with a latency target set -- when the predicted completion
// (after the launched batches, then the queued work ahead of it, the 90th
// percentile of the recent queries' work in place of its own, so that the
// test does not favour small queries, and the largest batches of the slots
// in flight, priced by the measured throughput) would exceed the target.
//
// Batching.  A worker thread keeps up to kSlots batches in flight, each on its
// own HIP stream with its own workspaces.  Whenever a slot is free it takes
// the queued queries in arrival order, up to max_batch_objects objects and
// kMaxQueries queries (with a latency target, at most an eighth of the target
// of device work: a query completes with its whole batch, and the admission
// test counts that much work joining its batch behind it), and runs one fused
// pass over all their objects:
// pre-pass, moment series, per-face form factors, then one split-precision
// Gram launch over every query's column range, and the copy of the matrices
// to pinned host memory.  Queries that arrive while batches run join the next
// free slot, so consecutive batches overlap on the device.
//
// Graphs.  The launch sequence of a batch is laid out for a capacity (objects
// rounded up to a bucket, V room for any batch of that many): descriptors past
// the batch's objects are empty, the grid-stride and persistent kernels read
// their counts from device memory, and the Gram reads a device plan.  So the
// sequence of one capacity is captured once into a hipGraph and replayed for
// every later batch of that capacity, the host only rewriting the pinned
// staging it uploads; the copy of the results (exact size) follows it.  Replay and direct launch run the same kernels with the
// same arguments: results are bit-identical, and a query's result does not
// depend on the batch it is part of (per-object form factors; per-query Gram
// slices).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
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
  std::vector<pii_geometry*> handles;
  std::vector<const GeometryDev*> geoms;
  std::vector<double> dofs;   // n x 12
  size_t footprint = 0;
  double work = 0;            // the unit of the latency model: see kObjectWork
  double t_arrival = 0, t_submit = 0, t_launch = 0, t_done = 0;
  std::vector<double> V;
  int state = 0;              // 0 queued, 1 running, 2 done
  bool failed = false;
  size_t v_off = 0;           // offset of V in the slot's output
};

constexpr int kSlots = 2;
constexpr int kMaxQueries = 64;
// Work of a query in the latency model: its triangles plus kObjectWork per
// object, the per-object cost of the low-|k| series (cached harmonics or the
// direct series of small meshes) in triangles of the per-face kernel, as
// measured on the serving workload (Lebedev 1202 x L=64: 17 us per object
// against 8.4 us per thousand triangles).
constexpr double kObjectWork = 2000.0;
// batches in the throughput estimate (ratio of sums)
constexpr int kRateWindow = 16;   // per batch: bounds the column padding between queries
constexpr int kStamps = 8;
constexpr int kFixedGrid = 1024;  // workgroups of the grid-stride pre-pass
constexpr int kGramGrid = 512;    // workgroups of the persistent Gram

// Smallest of lo, 1.5 lo, 2 lo, 3 lo, 4 lo, ... that is >= n.
long long bucket(long long n, long long lo) {
  long long b = lo;
  while (true) {
    if (b >= n) return b;
    if (b + b / 2 >= n) return b + b / 2;
    b *= 2;
  }
}

constexpr int kBaked = 12;
struct Graph {
  hipGraphExec_t exec = nullptr;
  std::array<const void*, kBaked> baked{};
};

struct Slot {
  Workspace w;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Request*> reqs;
  Batch batch;
  HostBuf out;                // pinned: V of every query of the batch, then the stamps
  DevBuf Vdev, stamps;
  std::map<long long, Graph> graphs;     // by object capacity
  // the batch in flight
  long long v_cap = 0;
  double work = 0, tri_node = 0, face_dir = 0, t_launch = 0;
};

}  // namespace

struct Server {
  int device = 0;
  pii_kgrid* kgh = nullptr;
  const KGridDev* kg = nullptr;
  int gram_mode = PII_GRAM_FP16X3;
  bool series = true;                 // moment series (the creating context's setting)
  std::atomic<bool> use_graphs{true};
  size_t budget = 0;
  size_t reserved = 0;
  int max_batch_objects = 1024;
  double latency_target_ms = 0.0;     // 0: admission by memory only
  // latency model: ms of device time per unit of work, from the marginal
  // completion time of the recent batches
  double ms_per_work = 0.0;
  bool calibrated = false;
  double rate_dt[kRateWindow] = {}, rate_work[kRateWindow] = {};
  int rate_n = 0;
  double last_done = 0.0;
  double outstanding_work = 0.0;
  // admitted work not yet in a batch, and the predicted time (engine clock)
  // at which the device finishes every launched batch
  double queued_work = 0.0;
  double t_free = 0.0;
  // the recent submitted queries' work and its 90th percentile
  std::vector<double> recent_work;
  int64_t n_noted = 0;
  double work_q90 = 0.0;
  void note_query_work(double w) {
    constexpr size_t kRecent = 1024;
    if (recent_work.size() < kRecent) recent_work.push_back(w);
    else recent_work[(size_t)(n_noted % (int64_t)kRecent)] = w;
    ++n_noted;
    if (n_noted <= 64 || n_noted % 64 == 0) {
      std::vector<double> v(recent_work);
      const size_t k = (size_t)(0.9 * (double)(v.size() - 1));
      std::nth_element(v.begin(), v.begin() + (std::ptrdiff_t)k, v.end());
      work_q90 = v[k];
    }
  }
  // device time of one batch at most, with a latency target: an eighth of it
  double batch_ms() const { return 0.125 * latency_target_ms; }
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  Slot slots[kSlots];
  // statistics
  int64_t n_submitted = 0, n_admitted = 0, n_refused_mem = 0, n_refused_lat = 0;
  int64_t n_completed = 0, n_batches = 0, n_objects = 0;
  std::atomic<int64_t> n_replays{0}, n_captures{0};   // counted by the worker without the lock
  double ms_prepass = 0, ms_series = 0, ms_ff = 0, ms_gram = 0;
  double units_tri_node = 0, units_face_dir = 0;
  std::string worker_error;

  int ldX() const { return (int)align_up((size_t)kg->Qh, 64); }
  static int ncol_cap(int n_cap) {
    return (int)align_up((size_t)n_cap, 16) + 16 * std::min(kMaxQueries, n_cap);
  }

  size_t footprint(int n, const std::vector<const GeometryDev*>& gs) const {
    const size_t ld = (size_t)ldX();
    const size_t ncol = align_up((size_t)n, 16);
    size_t b = ncol * ld * sizeof(float2);                    // X columns
    if (gram_mode == PII_GRAM_FP64) {
      b += (size_t)n * n * sizeof(double) * (gram_fp64_splits(kg->Qh, n) + 1);
    } else {
      const int fp16 = gram_mode == PII_GRAM_FP16X3 ? 1 : 0;
      b += gram_plane_bytes(fp16, (int)(2 * ld), (int)ncol);  // 16-bit planes
      b += (size_t)n * n * sizeof(double) *
           (gram_query(0, n, (int)(2 * ld)).nslice + 1);       // partials + V
    }
    for (const GeometryDev* g : gs) {
      b += (size_t)g->n * sizeof(double) * kRecFastTri;        // element records
      if (series) b += (size_t)g->n * sizeof(double) * kRecMoment;  // moment records
      b += (size_t)g->nv * 3 * sizeof(double);                 // world vertices
    }
    return b + 4096;
  }

  // The launch sequence of a batch laid out for capacity (n_cap objects,
  // ncol columns, v_cap V entries); captured into a graph or run directly.
  void enqueue(Slot& s, int n_cap, int ncol, long long v_cap, size_t plan_bytes) {
    hipStream_t st = s.stream;
    const int fp16 = gram_mode == PII_GRAM_FP16X3 ? 1 : 0;
    const int ld = ldX();
    float2* X = s.w.X.as<float2>();
    unsigned* colmax = s.w.colmax.as<unsigned>();
    unsigned long long* stamps = s.stamps.as<unsigned long long>();
    const Batch& b = s.batch;
    PII_HIP(hipMemcpyAsync(s.w.misc.p, s.w.stage.p, b.stage_bytes, hipMemcpyHostToDevice, st));
    PII_HIP(hipMemcpyAsync(s.w.gplan.p, s.w.gplan_h.p, plan_bytes, hipMemcpyHostToDevice, st));
    launch_stamp(st, stamps, 0);
    launch_clear_u32(st, colmax, ncol);
    launch_prepass_fixed(st, b.d_objs, b.d_vprefix, b.d_eprefix, n_cap, 1, kFixedGrid);
    launch_stamp(st, stamps, 1);
    if (b.Mh > 0) {
      launch_series_sh(st, b.d_objs, n_cap, kg->rep.as<double>(), kg->Qh, b.Mh, 1, nullptr, 0, X,
                       ld, colmax);
      launch_series_tri(st, b.d_objs, n_cap, kg->rep.as<double>(), kg->Qh, b.Mh, 1, nullptr, 0,
                        X, ld, colmax, 1, nullptr);
    }
    launch_stamp(st, stamps, 2);
    launch_ff_fast(st, b.d_objs, n_cap, kg->rep.as<double>(), kg->Qh, 0, 1, nullptr, 0, X, ld,
                   colmax, 1, nullptr);
    launch_stamp(st, stamps, 3);
    launch_split_planes(st, fp16, reinterpret_cast<const float*>(X), 2 * ld, ncol, colmax,
                        s.w.planes.p, s.w.gplan.as<GramPlanHeader>());
    launch_gram_planes(st, fp16, s.w.planes.p, 2 * ld, ncol, s.w.gplan.as<GramPlanHeader>(),
                       kGramGrid, v_cap, s.w.part.as<double>(), colmax, s.Vdev.as<double>());
    launch_stamp(st, stamps, 4);
    PII_HIP(hipGetLastError());
  }

  // The batch's results to pinned memory (exact sizes, outside the graph).
  void copy_out(Slot& s, long long nentries) {
    double* out = s.out.as<double>();
    PII_HIP(hipMemcpyAsync(out, s.Vdev.p, sizeof(double) * nentries, hipMemcpyDeviceToHost,
                           s.stream));
    PII_HIP(hipMemcpyAsync(out + s.v_cap, s.stamps.p, sizeof(double) * kStamps,
                           hipMemcpyDeviceToHost, s.stream));
  }

  // FP64 Gram mode: the plain per-query FP64 Grams, launched directly.
  void launch_fp64(Slot& s, std::vector<const GeometryDev*>& gs, std::vector<const double*>& dp,
                   std::vector<int>& kinds, std::vector<int>& nd, std::vector<int>& cols,
                   std::vector<int>& c0s, int ncol) {
    const int N = (int)gs.size(), Qh = kg->Qh, ld = ldX();
    s.batch = Batch();
    s.batch.use_sh = true;
    s.batch.build(s.w, s.stream, nullptr, N, gs.data(), kinds.data(), dp.data(), nd.data(), true,
                  series ? kg : nullptr, cols.data());
    s.w.X.ensure(sizeof(float2) * (size_t)ld * ncol);
    s.w.colmax.ensure(sizeof(unsigned) * (size_t)ncol);
    launch_clear_u32(s.stream, s.w.colmax.as<unsigned>(), ncol);
    launch_form_factors_fast(s.stream, s.batch, *kg, 1, nullptr, 0, s.w.X.as<float2>(), ld,
                             s.w.colmax.as<unsigned>(), s.w, nullptr);
    size_t vtot = 0, maxpart = 1;
    for (Request* r : s.reqs) {
      r->v_off = vtot;
      vtot += (size_t)r->n * r->n;
      maxpart = std::max(maxpart, (size_t)gram_fp64_splits(Qh, r->n) * r->n * r->n);
    }
    s.v_cap = (long long)std::max<size_t>(vtot, 1);
    s.Vdev.ensure(sizeof(double) * s.v_cap);
    s.w.part.ensure(sizeof(double) * maxpart);
    s.out.ensure(sizeof(double) * (s.v_cap + kStamps));
    for (size_t qi = 0; qi < s.reqs.size(); ++qi) {
      Request* r = s.reqs[qi];
      launch_gram_fp64_x(s.stream, s.w.X.as<float2>() + (size_t)c0s[qi] * ld, Qh, r->n, ld,
                         s.w.part.as<double>(), gram_fp64_splits(Qh, r->n),
                         s.Vdev.as<double>() + r->v_off);
      PII_HIP(hipGetLastError());
    }
    std::memset(s.out.as<double>() + s.v_cap, 0, sizeof(double) * kStamps);
    PII_HIP(hipMemcpyAsync(s.out.p, s.Vdev.p, sizeof(double) * vtot, hipMemcpyDeviceToHost,
                           s.stream));
  }

  void launch(Slot& s) {
    // Concatenate the queries' objects; each query's columns start at a
    // multiple of 16 (the Gram's column groups).
    std::vector<const GeometryDev*> gs;
    std::vector<const double*> dp;
    std::vector<int> kinds, nd, cols, c0s;
    int ncol = 0;
    s.work = 0;
    for (Request* r : s.reqs) {
      c0s.push_back(ncol);
      for (int i = 0; i < r->n; ++i) {
        gs.push_back(r->geoms[i]);
        dp.push_back(r->dofs.data() + 12 * (size_t)i);
        kinds.push_back(PII_DOF_AFFINE);
        nd.push_back(12);
        cols.push_back(ncol + i);
      }
      ncol += (int)align_up((size_t)r->n, 16);
      s.work += r->work;
    }
    const int N = (int)gs.size();
    s.t_launch = now_s();
    if (gram_mode == PII_GRAM_FP64) {
      launch_fp64(s, gs, dp, kinds, nd, cols, c0s, ncol);
    } else {
      const int ld = ldX(), K = 2 * ld;
      const int fp16 = gram_mode == PII_GRAM_FP16X3 ? 1 : 0;
      const int n_cap = (int)bucket(N, 16);
      const int nc = ncol_cap(n_cap);
      s.batch = Batch();
      s.batch.use_sh = true;
      s.batch.build(s.w, s.stream, nullptr, N, gs.data(), kinds.data(), dp.data(), nd.data(),
                    true, series ? kg : nullptr, cols.data(), n_cap, false);
      std::vector<GramQuery> qs;
      for (size_t qi = 0; qi < s.reqs.size(); ++qi)
        qs.push_back(gram_query(c0s[qi], s.reqs[qi]->n, K));
      GramPlanHeader h{};
      gram_plan_layout(qs, h);
      for (size_t qi = 0; qi < s.reqs.size(); ++qi) s.reqs[qi]->v_off = (size_t)qs[qi].v;
      // V room for any batch of n_cap objects (queries of at most
      // max_batch_objects each): the graph of a capacity depends on n_cap only
      s.v_cap = (long long)n_cap * std::min(n_cap, max_batch_objects);
      // buffers whose addresses a captured graph holds
      s.w.X.ensure(sizeof(float2) * (size_t)ld * nc);
      s.w.colmax.ensure(sizeof(unsigned) * (size_t)nc);
      s.w.planes.ensure(gram_plane_bytes(fp16, K, nc));
      const size_t part_b = sizeof(double) * (size_t)std::max<long long>(h.part_doubles, 1);
      if (s.w.part.bytes < part_b) s.w.part.ensure(part_b + part_b / 2);
      const size_t plan_bytes = sizeof(GramPlanHeader) + sizeof(GramQuery) * kMaxQueries;
      s.w.gplan.ensure(plan_bytes);
      s.w.gplan_h.ensure(plan_bytes);
      s.Vdev.ensure(sizeof(double) * s.v_cap);
      s.stamps.ensure(sizeof(double) * kStamps);
      s.out.ensure(sizeof(double) * (s.v_cap + kStamps));
      std::memcpy(s.w.gplan_h.p, &h, sizeof(h));
      std::memcpy(static_cast<char*>(s.w.gplan_h.p) + sizeof(h), qs.data(),
                  sizeof(GramQuery) * qs.size());
      const std::array<const void*, kBaked> baked = {
          s.w.misc.p, s.w.stage.p, s.w.gplan.p, s.w.gplan_h.p, s.w.X.p,   s.w.colmax.p,
          s.w.planes.p, s.w.part.p, s.Vdev.p,  s.stamps.p,   s.out.p,   kg->rep.p};
      if (use_graphs) {
        const long long key = n_cap;
        auto it = s.graphs.find(key);
        if (it != s.graphs.end() && it->second.baked != baked) {
          (void)hipGraphExecDestroy(it->second.exec);
          s.graphs.erase(it);
          it = s.graphs.end();
        }
        if (it == s.graphs.end()) {
          hipGraph_t g = nullptr;
          PII_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
          try {
            enqueue(s, n_cap, nc, s.v_cap, plan_bytes);
          } catch (...) {
            (void)hipStreamEndCapture(s.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          PII_HIP(hipStreamEndCapture(s.stream, &g));
          Graph gr;
          gr.baked = baked;
          const hipError_t e = hipGraphInstantiate(&gr.exec, g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          PII_HIP(e);
          it = s.graphs.emplace(key, gr).first;
          ++n_captures;
        }
        PII_HIP(hipGraphLaunch(it->second.exec, s.stream));
        ++n_replays;
      } else {
        enqueue(s, n_cap, nc, s.v_cap, plan_bytes);
      }
      copy_out(s, h.nentries);
    }
    // work of the batch's form-factor kernels
    s.tri_node = s.face_dir = 0;
    const int Qh = kg->Qh;
    for (const ObjDesc& d : s.batch.h_objs) {
      s.tri_node += (double)d.n * std::max(0, Qh - d.qbeg);
      if (d.nser > 0) s.face_dir += (double)d.n * s.batch.Mh;
    }
    PII_HIP(hipEventRecord(s.done, s.stream));
  }

  void release(Request* r) {
    for (pii_geometry* g : r->handles) g->refs.fetch_sub(1);
    r->handles.clear();
    reserved -= r->footprint;
    outstanding_work -= r->work;
  }

  void retire(Slot& s) {
    const double t = now_s();
    const double* out = s.out.as<double>();
    for (Request* r : s.reqs) {
      r->V.assign(out + r->v_off, out + r->v_off + (size_t)r->n * r->n);
      r->t_done = t;
      r->state = 2;
      release(r);
      ++n_completed;
    }
    // phase times from the device clock stamps (100 MHz)
    const unsigned long long* st = reinterpret_cast<const unsigned long long*>(out + s.v_cap);
    if (gram_mode != PII_GRAM_FP64 && st[4] >= st[0]) {
      ms_prepass += (double)(st[1] - st[0]) * 1e-5;
      ms_series += (double)(st[2] - st[1]) * 1e-5;
      ms_ff += (double)(st[3] - st[2]) * 1e-5;
      ms_gram += (double)(st[4] - st[3]) * 1e-5;
      units_tri_node += s.tri_node;
      units_face_dir += s.face_dir;
    }
    // throughput: the time each batch added after the previous completion,
    // over its work, as a ratio of sums over the last kRateWindow batches
    if (s.work > 0) {
      const double dt = t - std::max(s.t_launch, last_done);
      rate_dt[rate_n % kRateWindow] = dt;
      rate_work[rate_n % kRateWindow] = s.work;
      ++rate_n;
      double sdt = 0.0, sw = 0.0;
      for (int i = 0; i < std::min(rate_n, kRateWindow); ++i) {
        sdt += rate_dt[i];
        sw += rate_work[i];
      }
      ms_per_work = sdt * 1e3 / sw;
      calibrated = true;
    }
    last_done = t;
    s.reqs.clear();
    s.busy = false;
    bool any_busy = false;
    for (const Slot& o : slots) any_busy = any_busy || o.busy;
    if (!any_busy) t_free = t;  // the device is idle now
  }

  void fail(Slot& s, const std::string& msg) {
    worker_error = msg;
    for (Request* r : s.reqs) {
      r->state = 2;
      r->failed = true;
      r->t_done = now_s();
      release(r);
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
      bool any_busy = false;
      for (Slot& s : slots) any_busy = any_busy || s.busy;
      if (stop && queue.empty() && !any_busy) break;
      for (Slot& s : slots) {
        if (s.busy || queue.empty()) continue;
        int nobj = 0;
        const double t = now_s();
        // With a latency target, a batch holds at most batch_ms() of work: a
        // query completes with its whole batch, so the work queued behind it
        // that joins its batch delays it by at most that much.
        const double wcap = (latency_target_ms > 0.0 && calibrated && ms_per_work > 0.0)
                                ? batch_ms() / ms_per_work : HUGE_VAL;
        const double start = std::max(t, t_free);
        double bwork = 0.0;
        while (!queue.empty() && (int)s.reqs.size() < kMaxQueries) {
          Request* r = queue.front();
          if (!s.reqs.empty() && (nobj + r->n > max_batch_objects || bwork + r->work > wcap))
            break;
          queue.pop_front();
          s.reqs.push_back(r);
          r->state = 1;
          r->t_launch = t;
          nobj += r->n;
          bwork += r->work;
          queued_work -= r->work;
        }
        t_free = calibrated ? start + bwork * ms_per_work * 1e-3 : t;
        ++n_batches;
        n_objects += nobj;
        s.busy = true;
        // host work of the batch (descriptors, plan, launches) without the lock
        lk.unlock();
        std::string err;
        try {
          launch(s);
        } catch (const Error& e) {
          err = e.msg.empty() ? "launch failed" : e.msg;
        } catch (const std::exception& e) {
          err = e.what();
        }
        lk.lock();
        if (!err.empty()) {
          fail(s, err);
          cv_done.notify_all();
        }
        progressed = true;
      }
      if (!progressed) {
        any_busy = false;
        for (Slot& s : slots) any_busy = any_busy || s.busy;
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

int pii_clock_seconds(double* t) {
  return guarded([&] {
    if (!t) throw Error(PII_ERR_INVALID, "pii_clock_seconds: NULL output");
    *t = now_s();
  });
}

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
    S.kgh = const_cast<pii_kgrid*>(kg);
    S.kg = &kg->k;
    S.series = ctx->c.series;
    S.gram_mode = gram_mode;
    S.budget = (size_t)hbm_budget;
    S.max_batch_objects = max_batch_objects;
    for (auto& sl : S.slots) {
      PII_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
      PII_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    S.kgh->refs.fetch_add(1);
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
      for (auto& kv : sl.graphs) (void)hipGraphExecDestroy(kv.second.exec);
      sl.graphs.clear();
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    // queries queued or running were all retired by the worker; release any
    // reference a failed launch left
    for (auto& kv : S.reqs) {
      for (pii_geometry* g : kv.second->handles) g->refs.fetch_sub(1);
      kv.second->handles.clear();
    }
    if (S.kgh) S.kgh->refs.fetch_sub(1);
    delete srv;
  });
}

int pii_server_set_latency_target(pii_server* srv, double target_ms) {
  return guarded([&] {
    if (!srv || !(target_ms >= 0.0))
      throw Error(PII_ERR_INVALID, "pii_server_set_latency_target: bad arguments");
    std::lock_guard<std::mutex> lk(srv->s.mu);
    srv->s.latency_target_ms = target_ms;
  });
}

int pii_server_set_graphs(pii_server* srv, int enable) {
  return guarded([&] {
    if (!srv) throw Error(PII_ERR_INVALID, "NULL server");
    std::lock_guard<std::mutex> lk(srv->s.mu);
    srv->s.use_graphs = enable != 0;
  });
}

int pii_server_submit_at(pii_server* srv, int n, const pii_geometry* const* geoms,
                         const double* affine_dofs, double t_arrival, int64_t* id) {
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
      if (geoms[i]->device != S.device)
        throw Error(PII_ERR_INVALID, "pii_server_submit: geometry on another device");
      r->geoms[i] = &geoms[i]->g;
      r->work += (double)geoms[i]->g.n + kObjectWork;
    }
    r->dofs.assign(affine_dofs, affine_dofs + 12 * (size_t)n);
    r->footprint = S.footprint(n, r->geoms);
    r->t_submit = now_s();
    r->t_arrival = t_arrival > 0.0 ? std::min(t_arrival, r->t_submit) : r->t_submit;
    std::lock_guard<std::mutex> lk(S.mu);
    ++S.n_submitted;
    if (!S.worker_error.empty()) throw Error(PII_ERR_HIP, "serving worker failed: " + S.worker_error);
    if (n > S.max_batch_objects || S.reserved + r->footprint > S.budget) {
      ++S.n_refused_mem;
      throw Error(PII_ERR_BUDGET, "query refused: HBM budget (" + std::to_string(S.reserved) +
                                      " + " + std::to_string(r->footprint) + " > " +
                                      std::to_string(S.budget) + " bytes)");
    }
    S.note_query_work(r->work);
    if (S.latency_target_ms > 0.0 && S.calibrated) {
      // completion if the queue drains in order: after the launched batches
      // and the work queued ahead of it, priced by the measured throughput,
      // plus the work that may join its batch behind it, its batch sharing
      // the device with the other batches in flight (kSlots batches' worth).
      // Its own work is taken as the 90th percentile of the recent queries',
      // so that the test does not depend on the query's size and the served
      // mix of scene sizes is the offered one.
      const double start = std::max(r->t_submit, S.t_free);
      const double predicted = (start - r->t_arrival) * 1e3 +
                               (S.queued_work + S.work_q90) * S.ms_per_work +
                               kSlots * S.batch_ms();
      if (predicted > S.latency_target_ms) {
        ++S.n_refused_lat;
        throw Error(PII_ERR_BUDGET, "query refused: predicted latency " +
                                        std::to_string(predicted) + " ms > target " +
                                        std::to_string(S.latency_target_ms) + " ms");
      }
    }
    ++S.n_admitted;
    S.reserved += r->footprint;
    S.outstanding_work += r->work;
    S.queued_work += r->work;
    r->handles.resize(n);
    for (int i = 0; i < n; ++i) {
      r->handles[i] = const_cast<pii_geometry*>(geoms[i]);
      r->handles[i]->refs.fetch_add(1);
    }
    r->id = S.next_id++;
    *id = r->id;
    S.queue.push_back(r.get());
    S.reqs[r->id] = std::move(r);
    S.cv_work.notify_one();
  });
}

int pii_server_submit(pii_server* srv, int n, const pii_geometry* const* geoms,
                      const double* affine_dofs, int64_t* id) {
  return pii_server_submit_at(srv, n, geoms, affine_dofs, 0.0, id);
}

int pii_server_wait_times(pii_server* srv, int64_t id, double timeout_s, double* V,
                          double* times) {
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
    if (r->failed || r->V.size() != (size_t)r->n * r->n) {
      const std::string msg = S.worker_error;
      S.reqs.erase(it);
      throw Error(PII_ERR_HIP, "query " + std::to_string(id) + " failed: " + msg);
    }
    if (V) std::memcpy(V, r->V.data(), sizeof(double) * r->V.size());
    if (times) {
      times[0] = r->t_arrival;
      times[1] = r->t_submit;
      times[2] = r->t_launch;
      times[3] = r->t_done;
    }
    S.reqs.erase(it);
  });
}

int pii_server_wait(pii_server* srv, int64_t id, double timeout_s, double* V, double* latency_ms) {
  double t[4];
  const int rc = pii_server_wait_times(srv, id, timeout_s, V, t);
  if (rc == PII_OK && latency_ms) latency_ms[0] = (t[3] - t[0]) * 1e3;
  return rc;
}

int pii_server_stats(pii_server* srv, double* out, int n) {
  return guarded([&] {
    if (!srv || !out) throw Error(PII_ERR_INVALID, "pii_server_stats: bad arguments");
    Server& S = srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    const double v[] = {(double)S.n_submitted, (double)S.n_admitted,
                        (double)(S.n_refused_mem + S.n_refused_lat), (double)S.n_completed,
                        (double)S.n_batches, (double)S.n_objects, (double)S.reserved,
                        (double)S.budget, (double)S.n_refused_mem, (double)S.n_refused_lat,
                        (double)S.n_replays.load(), (double)S.n_captures.load(), S.ms_prepass,
                        S.ms_series,
                        S.ms_ff, S.ms_gram, S.units_tri_node, S.units_face_dir,
                        S.calibrated ? S.ms_per_work : 0.0};
    const int nv = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n; ++i) out[i] = i < nv ? v[i] : 0.0;
  });
}

}  // extern "C"
