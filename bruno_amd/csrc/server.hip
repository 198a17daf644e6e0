/*
 * server.hip -- request loop for many independent assembly jobs on one
 * device: admission by estimated HBM footprint, continuous batching, and
 * overlapping batches on separate streams.
 *
 * Requests queue in arrival order.  Each worker thread owns a non-blocking
 * stream; when it is free it takes the oldest waiting request plus every
 * later waiting request with the same k and thresholds whose footprint
 * still fits the server's budget (up to max_batch_jobs), and runs them as
 * one batch (bruno_batch_create ... bruno_run_pipeline).  Requests that
 * arrive while batches are on the device wait for the next free worker and
 * the memory the running batches give back, so several batches overlap on
 * the device.  A request with a large table (2^26 slots or more) runs on
 * its own, counted by regions, which beats its share of a batch's atomic
 * count.  A request whose table proves too small is retried on its own
 * with more slots.  All device memory comes from the library's pool,
 * so a running server makes no hipMalloc / hipFree calls.
 *
 * The reference has no request loop (one assembly per MPI job,
 * compact_dbg_freq_construct_clean.cpp:431-840); each job's result is the
 * one bruno_run_pipeline gives for it alone.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bruno_hip.h"
#include "pool.h"

/* bruno_hip.hip: slots from which a one-job table counts by regions */
extern "C" uint64_t br_region_count_min_slots();

namespace {

__global__ void k_add_offset(uint64_t* o, uint64_t n, uint64_t add) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] += add;
}

struct Request {
  uint64_t ticket = 0;
  bruno_job_desc job{};
  bool synth = false;
  bruno_synth_desc sd{};
  std::vector<char> bases;       /* host reads (copied at submit) */
  std::vector<uint64_t> offsets; /* n_reads + 1, starting at 0 */
  uint64_t n_reads = 0, n_bases = 0;
  uint64_t capacity = 0;
  uint64_t footprint = 0;
  uint32_t attempts = 0;
  bool alone = false;
  double t_submit = 0;
};

struct Result {
  bool done = false;
  bruno_result_info info{};
  std::string unitigs, branch;
};

bool same_params(const bruno_job_desc& a, const bruno_job_desc& b) {
  if (a.k != b.k || a.thresholding != b.thresholding) return false;
  for (int i = 0; i < 6; ++i)
    if (a.thresholding && a.thresholds[i] != b.thresholds[i]) return false;
  return true;
}

}  // namespace

struct bruno_server {
  bruno_server_desc d{};
  int device = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<uint64_t, Result> results;
  uint64_t next_ticket = 1, next_batch = 1;
  uint64_t reserved = 0;
  bool stop = false;
  bruno_server_stats st{};
  std::vector<std::thread> workers;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

/* a request with a table of at least this many slots runs alone */
constexpr uint64_t SOLO_MIN_SLOTS = 1ULL << 26;

/* Estimated peak HBM of one job: table and side table, the present list
 * (8 B per distinct k-mer, at most one per slot), the chain, list-ranking
 * and output arrays (a quarter of the table for reads of one genome), the
 * reads themselves and a fixed allowance for small buffers. */
uint64_t footprint(const bruno_job_desc& j, uint64_t cap, uint64_t n_reads, uint64_t n_bases) {
  const uint64_t slot = j.k <= 31 ? 32 : 64;
  const uint64_t table = cap * slot + std::max<uint64_t>(1024, cap / 32) * 64;
  return table + table / 4 + cap * 8 + n_bases + 8 * (n_reads + 1) + (16ULL << 20);
}

void finish(bruno_server* S, const std::shared_ptr<Request>& r, const bruno_result_info& info, std::string&& u,
            std::string&& b) {
  Result& res = S->results[r->ticket];
  res.info = info;
  res.unitigs = std::move(u);
  res.branch = std::move(b);
  res.done = true;
  if (info.status == BRUNO_OK) ++S->st.jobs_done;
  else ++S->st.jobs_failed;
}

/* Run one batch on stream st.  Returns the batch status; per-job results in
 * out (only when the status is OK). */
bruno_status run_batch(bruno_server* S, hipStream_t st, const std::vector<std::shared_ptr<Request>>& batch,
                       std::vector<bruno_result_info>& info, std::vector<std::string>& uni,
                       std::vector<std::string>& bra) {
  const uint32_t n = (uint32_t)batch.size();
  uint64_t nb = 0, nr = 0;
  for (auto& r : batch) {
    nb += r->n_bases;
    nr += r->n_reads;
  }
  br::DevicePool& pool = br::DevicePool::current();
  char* d_bases = nullptr;
  uint64_t* d_offs = nullptr;
  if (pool.alloc((void**)&d_bases, nb + 1) != hipSuccess) return BRUNO_OOM;
  if (pool.alloc((void**)&d_offs, (nr + 1) * 8) != hipSuccess) {
    pool.release(d_bases);
    return BRUNO_OOM;
  }
  bruno_status status = BRUNO_OK;
  std::vector<uint64_t> first(n + 1, 0);
  std::vector<std::vector<uint64_t>> staged;
  uint64_t boff = 0, roff = 0;
  for (uint32_t i = 0; i < n && status == BRUNO_OK; ++i) {
    const Request& r = *batch[i];
    first[i] = roff;
    if (r.synth) {
      status = bruno_synth_reads_device(&r.sd, d_bases + boff, d_offs + roff, st);
      if (status == BRUNO_OK && boff)
        hipLaunchKernelGGL(k_add_offset, dim3((unsigned)((r.n_reads + 1 + 255) / 256)), dim3(256), 0, st,
                           d_offs + roff, r.n_reads + 1, boff);
    } else {
      staged.emplace_back(r.offsets);
      for (auto& x : staged.back()) x += boff;
      if ((r.n_bases && hipMemcpyAsync(d_bases + boff, r.bases.data(), r.n_bases, hipMemcpyHostToDevice, st) !=
                            hipSuccess) ||
          hipMemcpyAsync(d_offs + roff, staged.back().data(), (r.n_reads + 1) * 8, hipMemcpyHostToDevice, st) !=
              hipSuccess) {
        (void)hipGetLastError();
        status = BRUNO_INTERNAL;
      }
    }
    boff += r.n_bases;
    roff += r.n_reads;
  }
  first[n] = roff;
  if (status == BRUNO_OK && hipStreamSynchronize(st) != hipSuccess) {
    (void)hipGetLastError();
    status = BRUNO_INTERNAL;
  }
  bruno_job* job = nullptr;
  if (status == BRUNO_OK) {
    bruno_job_desc jd = batch[0]->job;
    jd.stream = st;
    jd.capacity = 0;
    std::vector<uint64_t> caps(n);
    for (uint32_t i = 0; i < n; ++i) caps[i] = batch[i]->capacity;
    status = bruno_batch_create(&jd, n, caps.data(), &job);
  }
  if (status == BRUNO_OK) status = bruno_batch_parse_count(job, d_bases, d_offs, nr, first.data());
  if (status == BRUNO_OK) status = bruno_run_pipeline(job, S->d.variant);
  info.assign(n, bruno_result_info{});
  uni.assign(n, std::string());
  bra.assign(n, std::string());
  for (uint32_t i = 0; i < n && status == BRUNO_OK; ++i) {
    bruno_stats s;
    status = bruno_batch_get_stats(job, i, &s);
    uint64_t lu = 0, lb = 0;
    if (status == BRUNO_OK) status = bruno_batch_get_unitigs(job, i, nullptr, &lu);
    if (status == BRUNO_OK) {
      uni[i].resize(lu);
      if (lu) status = bruno_batch_get_unitigs(job, i, &uni[i][0], &lu);
    }
    if (status == BRUNO_OK) status = bruno_batch_get_branch_fasta(job, i, nullptr, &lb);
    if (status == BRUNO_OK && lb) {
      bra[i].resize(lb);
      status = bruno_batch_get_branch_fasta(job, i, &bra[i][0], &lb);
    }
    bruno_result_info& o = info[i];
    o.n_reads = s.n_reads;
    o.n_instances = s.n_instances;
    o.n_unitigs = s.n_unitigs;
    o.n_unitig_nodes = s.n_unitig_nodes;
    o.n_branches = s.n_branches;
    o.iterations = (uint32_t)s.iterations;
    o.unitig_bytes = lu;
    o.branch_bytes = lb;
  }
  if (job) bruno_job_destroy(job);
  (void)hipStreamSynchronize(st);
  pool.release(d_bases);
  pool.release(d_offs);
  return status;
}

void worker(bruno_server* S) {
  (void)hipSetDevice(S->device);
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    st = nullptr;
  }
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch;
    uint64_t fp = 0, bid = 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        if (S->queue.empty()) {
          if (S->stop) break;
          S->cv_work.wait(lk);
          continue;
        }
        std::shared_ptr<Request> h = S->queue.front();
        if (h->footprint > S->d.hbm_budget) {
          /* can never fit */
          S->queue.pop_front();
          bruno_result_info o{};
          o.status = BRUNO_CAPACITY;
          o.submit_s = h->t_submit;
          o.start_s = o.end_s = S->now();
          o.attempts = h->attempts;
          finish(S, h, o, std::string(), std::string());
          S->cv_done.notify_all();
          continue;
        }
        if (S->reserved + h->footprint > S->d.hbm_budget) {
          S->cv_work.wait(lk);
          continue;
        }
        break;
      }
      if (S->queue.empty() && S->stop) break;
      /* the oldest request, then every later compatible one that fits */
      const std::shared_ptr<Request> head = S->queue.front();
      for (auto it = S->queue.begin(); it != S->queue.end() && batch.size() < S->d.max_batch_jobs;) {
        const std::shared_ptr<Request>& r = *it;
        const bool first = batch.empty();
        const bool ok = first || (!head->alone && !r->alone && same_params(r->job, head->job) &&
                                  S->reserved + fp + r->footprint <= S->d.hbm_budget);
        if (ok) {
          fp += r->footprint;
          batch.push_back(r);
          it = S->queue.erase(it);
          if (head->alone) break;
        } else {
          ++it;
        }
      }
      S->reserved += fp;
      bid = S->next_batch++;
      ++S->st.batches;
      S->st.max_batch_jobs = std::max<uint64_t>(S->st.max_batch_jobs, batch.size());
      S->st.peak_reserved = std::max(S->st.peak_reserved, S->reserved);
    }
    const double t_start = S->now();
    std::vector<bruno_result_info> info;
    std::vector<std::string> uni, bra;
    bruno_status s = run_batch(S, st, batch, info, uni, bra);
    const double t_end = S->now();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->reserved -= fp;
      if (s == BRUNO_OK) {
        for (size_t i = 0; i < batch.size(); ++i) {
          info[i].status = BRUNO_OK;
          info[i].batch_id = bid;
          info[i].batch_jobs = (uint32_t)batch.size();
          info[i].submit_s = batch[i]->t_submit;
          info[i].start_s = t_start;
          info[i].end_s = t_end;
          info[i].attempts = batch[i]->attempts + 1;
          finish(S, batch[i], info[i], std::move(uni[i]), std::move(bra[i]));
        }
      } else if ((s == BRUNO_CAPACITY || s == BRUNO_OOM) && batch.size() > 1) {
        /* which job overflowed is not known: each runs again on its own */
        for (auto it = batch.rbegin(); it != batch.rend(); ++it) {
          (*it)->alone = true;
          ++(*it)->attempts;
          S->queue.push_front(*it);
        }
        ++S->st.retries;
      } else if (s == BRUNO_CAPACITY && batch[0]->attempts < 3) {
        /* four times the slots, then a table that cannot fill: distinct
         * k-mers <= instances <= bases, at load 1/2 */
        std::shared_ptr<Request> r = batch[0];
        r->capacity *= 4;
        if (r->attempts >= 2) r->capacity = std::max<uint64_t>(r->capacity, 2 * r->n_bases + 4096);
        r->footprint = footprint(r->job, r->capacity, r->n_reads, r->n_bases);
        r->alone = true;
        ++r->attempts;
        S->queue.push_front(r);
        ++S->st.retries;
      } else {
        for (auto& r : batch) {
          bruno_result_info o{};
          o.status = s;
          o.batch_id = bid;
          o.batch_jobs = (uint32_t)batch.size();
          o.submit_s = r->t_submit;
          o.start_s = t_start;
          o.end_s = t_end;
          o.attempts = r->attempts + 1;
          finish(S, r, o, std::string(), std::string());
        }
      }
    }
    S->cv_work.notify_all();
    S->cv_done.notify_all();
  }
  if (st) (void)hipStreamDestroy(st);
}

}  // namespace

extern "C" {

bruno_status bruno_server_create(const bruno_server_desc* d, bruno_server** out) {
  if (!d || !out) return BRUNO_BAD_INPUT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return BRUNO_NO_DEVICE;
  }
  bruno_server* S = new (std::nothrow) bruno_server();
  if (!S) return BRUNO_OOM;
  S->d = *d;
  (void)hipGetDevice(&S->device);
  if (S->d.n_workers == 0) S->d.n_workers = 2;
  if (S->d.max_batch_jobs == 0) S->d.max_batch_jobs = 64;
  if (S->d.max_batch_jobs > BRUNO_MAX_BATCH_JOBS) S->d.max_batch_jobs = BRUNO_MAX_BATCH_JOBS;
  if (S->d.variant != 0) S->d.variant = 1;
  if (S->d.hbm_budget == 0) {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) {
      (void)hipGetLastError();
      delete S;
      return BRUNO_NO_DEVICE;
    }
    S->d.hbm_budget = (uint64_t)(0.85 * (double)f);
  }
  if (S->d.reserve_pool && br::DevicePool::current().reserve(S->d.hbm_budget) != hipSuccess) {
    delete S;
    return BRUNO_OOM;
  }
  S->t0 = std::chrono::steady_clock::now();
  for (uint32_t i = 0; i < S->d.n_workers; ++i) S->workers.emplace_back(worker, S);
  *out = S;
  return BRUNO_OK;
}

bruno_status bruno_server_submit(bruno_server* S, const bruno_request* q, uint64_t* ticket) {
  if (!S || !q || !ticket) return BRUNO_BAD_INPUT;
  const int k = q->job.k;
  if (k < 3 || k > 63) return BRUNO_BAD_INPUT;
  auto r = std::make_shared<Request>();
  r->job = q->job;
  r->job.stream = nullptr;
  if (q->synth) {
    r->synth = true;
    r->sd = *q->synth;
    if (r->sd.read_len == 0 || r->sd.genome_len < 2 || r->sd.read_len > r->sd.genome_len) return BRUNO_BAD_INPUT;
    r->n_reads = r->sd.n_reads;
    r->n_bases = r->sd.n_reads * r->sd.read_len;
  } else {
    if (q->n_reads && (!q->bases || !q->offsets)) return BRUNO_BAD_INPUT;
    r->n_reads = q->n_reads;
    const uint64_t o0 = q->n_reads ? q->offsets[0] : 0;
    r->n_bases = q->n_reads ? q->offsets[q->n_reads] - o0 : 0;
    r->bases.assign(q->bases ? q->bases + o0 : nullptr, q->bases ? q->bases + o0 + r->n_bases : nullptr);
    r->offsets.resize(q->n_reads + 1, 0);
    for (uint64_t i = 0; i <= q->n_reads && q->n_reads; ++i) r->offsets[i] = q->offsets[i] - o0;
  }
  /* without a capacity hint: one slot per k-mer instance at load 1/2 */
  r->capacity = q->job.capacity;
  if (r->capacity == 0) r->capacity = std::max<uint64_t>(4096, 2 * r->n_bases);
  r->footprint = footprint(r->job, r->capacity, r->n_reads, r->n_bases);
  /* a large table runs as a batch of its own, counted by regions: from
   * about 2^26 slots that beats the job's share of a batch (atomic count) */
  const uint64_t region = br_region_count_min_slots();
  r->alone = region && r->capacity >= std::max<uint64_t>(region, SOLO_MIN_SLOTS);
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stop) return BRUNO_STATE;
  r->ticket = S->next_ticket++;
  r->t_submit = S->now();
  S->results[r->ticket] = Result();
  S->queue.push_back(r);
  ++S->st.submitted;
  *ticket = r->ticket;
  S->cv_work.notify_one();
  return BRUNO_OK;
}

bruno_status bruno_server_poll(bruno_server* S, uint64_t ticket, int* done) {
  if (!S || !done) return BRUNO_BAD_INPUT;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->results.find(ticket);
  if (it == S->results.end()) return BRUNO_BAD_INPUT;
  *done = it->second.done ? 1 : 0;
  return BRUNO_OK;
}

bruno_status bruno_server_wait(bruno_server* S, uint64_t ticket, bruno_result_info* info) {
  if (!S) return BRUNO_BAD_INPUT;
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    auto it = S->results.find(ticket);
    if (it == S->results.end()) return BRUNO_BAD_INPUT;
    if (it->second.done) {
      if (info) *info = it->second.info;
      return BRUNO_OK;
    }
    S->cv_done.wait(lk);
  }
}

static bruno_status get_result_text(bruno_server* S, uint64_t ticket, bool unitigs, char* buf, uint64_t* n) {
  if (!S || !n) return BRUNO_BAD_INPUT;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->results.find(ticket);
  if (it == S->results.end()) return BRUNO_BAD_INPUT;
  if (!it->second.done) return BRUNO_STATE;
  const std::string& t = unitigs ? it->second.unitigs : it->second.branch;
  if (!buf) {
    *n = t.size();
    return BRUNO_OK;
  }
  if (*n < t.size()) return BRUNO_BAD_INPUT;
  if (!t.empty()) memcpy(buf, t.data(), t.size());
  *n = t.size();
  return BRUNO_OK;
}

bruno_status bruno_server_get_unitigs(bruno_server* S, uint64_t ticket, char* buf, uint64_t* n) {
  return get_result_text(S, ticket, true, buf, n);
}

bruno_status bruno_server_get_branch_fasta(bruno_server* S, uint64_t ticket, char* buf, uint64_t* n) {
  return get_result_text(S, ticket, false, buf, n);
}

bruno_status bruno_server_release(bruno_server* S, uint64_t ticket) {
  if (!S) return BRUNO_BAD_INPUT;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->results.find(ticket);
  if (it == S->results.end()) return BRUNO_BAD_INPUT;
  if (!it->second.done) return BRUNO_STATE;
  S->results.erase(it);
  return BRUNO_OK;
}

bruno_status bruno_server_get_stats(bruno_server* S, bruno_server_stats* out) {
  if (!S || !out) return BRUNO_BAD_INPUT;
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->st;
  out->queued = S->queue.size();
  out->reserved_now = S->reserved;
  out->hbm_budget = S->d.hbm_budget;
  return BRUNO_OK;
}

bruno_status bruno_server_destroy(bruno_server* S) {
  if (!S) return BRUNO_BAD_INPUT;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto& t : S->workers) t.join();
  delete S;
  return BRUNO_OK;
}

}  // extern "C"
