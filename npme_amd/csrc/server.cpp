// Serving loop: a persistent set of worker slots on one GPU that evaluates a
// stream of independent CalcV requests (charge sets of any size and kernel).
//
// Each slot owns one handle (its HIP stream, set-up state and buffers) and
// runs a request end to end: NPME_AddMissingKeywords (defaults and the RSEM
// grid for the requested tolerance), the device set-up of the geometry and
// the kernel grid, the evaluation, and the copies in and out.  Slots run
// concurrently, so requests that arrive while others are on the GPU are
// admitted as soon as a slot and enough device memory are free, and their
// kernels overlap on the device.  Admission is first come first served and
// reserves an estimate of the request's device footprint against the
// server's budget; buffers come from the caching allocator, so a finished
// request's memory is reused by the next without hipMalloc/hipFree.
//
// The result of a request does not depend on what else is running: every
// kernel is deterministic and a request touches only its own buffers.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <set>
#include <sstream>
#include <thread>
#include <unordered_map>

#include "npme_internal.h"
#include "request_board.h"

namespace npme {

namespace {

struct Job {
  long ticket = 0;
  std::string funcType;
  double alpha = 0, k0r = 0, k0i = 0, tol = 1e-6, Rdir = 8;
  long Bn = 0;
  bool fp64 = false;
  size_t n = 0;
  const double *coord = nullptr, *charge = nullptr;
  double *V = nullptr;
  double tArrive = 0, tSubmit = 0, tStart = 0, tEnd = 0;
  size_t reserve = 0;
  int status = NPME_OK;
  std::string err;
  long N[3] = {0, 0, 0};
  double deviceGB = 0, estGB = 0;
  int batch = 1;  // requests whose CalcVs ran in the same graph launch
  double ms[5] = {0, 0, 0, 0, 0};  // keywords, set-up, copy in, CalcV, copy out
};

}  // namespace

// the admission state (reserved bytes, running requests) is guarded by the
// board's mutex: it is read and changed only in the board's callbacks
struct Server {
  RequestBoard<Job> board;
  std::vector<std::thread> workers;
  size_t budget = 0, reserved = 0;
  int running = 0;  // admitted requests (a batch counts each member)
  int active = 0;   // slots holding admitted requests (a batch counts once)
  std::set<long> growing;  // running requests waiting to enlarge their reservation
  // small requests queued together run as one batch on a slot: their
  // set-ups in turn, then their CalcVs captured as one graph
  int maxBatch = 8;
  size_t batchMaxN = 32768;
  int device = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const
  {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

// Admission reserves device memory in three steps:
//  1. at submission, a provisional estimate from the charge count alone
//     (about 2.5 KB per charge at the density of the serving stream, plus
//     64 MB), which admits the request to a slot;
//  2. once its grid is chosen, the request's own footprint
//     (footprint_bytes): the reservation is lowered to it, or raised, waiting
//     for running requests to finish when the raise does not fit;
//  3. after the set-up, the handle's resident device bytes.
static size_t estimate_bytes(size_t n, bool complexK, bool fp64)
{
  const double perCharge = 2500.0 * (fp64 ? 1.5 : 1.0) * (complexK ? 1.5 : 1.0);
  return (size_t)(perCharge * (double)n) + (64u << 20);
}

// footprint once the grid is known: the FFT buffers and G (from the grid
// sizes), the FP64 kernel grid F of the set-up, the per-charge arrays, and
// the neighbour list at the average density of the charges' box (ordered
// pairs n rho 4/3 pi Rdir^3 of 4 bytes, 25% slice padding, 15% for the count
// pass's hit masks); a denser (clustered) set whose list would not fit
// falls back to the cell-row direct sum in the set-up
static size_t footprint_bytes(size_t n, const long N[3], bool fp64, bool cx, double Rdir,
                              const double X0[3])
{
  const double rs = fp64 ? 8.0 : 4.0, cs = 2.0 * rs;
  const double M1 = N[0] / 2.0, M2 = N[1] / 2.0, M3 = N[2] / 2.0;
  const double K3 = cx ? (double)N[2] : N[2] / 2.0 + 1.0, Kp = std::ceil(K3 / 8.0) * 8.0;
  const double grids = (cx ? cs : rs) * M1 * M2 * M3 + cs * M1 * M2 * Kp + cs * M1 * N[1] * Kp;
  const double Mf = (M1 + 1.0) * (M2 + 1.0) * (M3 + 1.0);
  const double G = (cx ? cs : rs) * (M1 + 1.0) * (M2 + 1.0) * std::ceil((M3 + 1.0) / 8.0) * 8.0 +
                   (cx ? 16.0 : 8.0) * Mf;
  const double perCharge = (fp64 ? 320.0 : 260.0) * (cx ? 1.4 : 1.0);
  const double vol = std::max(X0[0] * X0[1] * X0[2], 1.0);  // X0: the charges' extents
  const double nb = std::min((double)n, (double)n / vol * 4.18879020478639 * Rdir * Rdir * Rdir);
  const double list = (double)n * nb * 4.0 * 1.25 * 1.15;
  return (size_t)(grids + G + perCharge * (double)n + list) + (8u << 20);
}

// step 2: the reservation of a running request becomes need; a raise that
// does not fit waits until it fits, or its slot is the only one at work, or
// every other working slot is waiting to raise too (then the oldest goes on).
// Slots, not requests, are counted: a batch's members are set up one after
// another on one slot, so at most one of them waits at a time, and counting
// the others would leave every slot waiting for good
static void set_reservation(Server *S, Job &j, size_t need)
{
  if (need <= j.reserve) {
    S->board.adjust([] { return true; }, [&] {
      S->reserved -= j.reserve - need;
      j.reserve = need;
    });
    return;
  }
  S->board.adjust([] { return true; }, [&] { S->growing.insert(j.ticket); });
  S->board.adjust(
      [&] {
        const bool fits = S->reserved - j.reserve + need <= S->budget;
        const bool stuck = (int)S->growing.size() >= S->active && *S->growing.begin() == j.ticket;
        return fits || S->active == 1 || stuck;
      },
      [&] {
        S->growing.erase(j.ticket);
        S->reserved += need - j.reserve;
        j.reserve = need;
      });
}

// RSEM grid for r^alpha when the table has no row for alpha: the alpha = 1
// model at the same tolerance.  For alpha = 0.5 and 1.5 (N = 4000, density
// 0.1, Rdir 8) its grids reach relative RMS errors 100x below the requested
// 1e-4 and 1e-6 against the exact sum, i.e. the substitute is conservative.
static void add_keywords_for_request(Keywords &kw, const Job &j, const ChargeStats &st)
{
  const bool cx = j.funcType == "Helmholtz";
  try {
    add_missing_keywords(kw, j.n, !cx, j.coord, cx ? nullptr : j.charge, cx ? j.charge : nullptr,
                         &st);
  } catch (const Error &) {
    if (j.funcType != "Ralpha") throw;
    std::string s;
    if (kw.get("N1", s)) throw;  // the grid was given: a genuine error
    Keywords k2 = kw;
    k2.set("alpha", 1.0);
    add_missing_keywords(k2, j.n, true, j.coord, j.charge, nullptr, &st);
    const char *keys[] = {"N1", "N2", "N3", "n1", "n2", "n3", "nDeriv"};
    for (const char *k : keys) {
      long v;
      k2.get(k, v);
      kw.set(k, v);
    }
    add_missing_keywords(kw, j.n, true, j.coord, j.charge, nullptr, &st);
  }
}

using Clock = std::chrono::steady_clock;
static double lap(Clock::time_point &t)
{
  const auto t1 = Clock::now();
  const double ms = std::chrono::duration<double, std::milli>(t1 - t).count();
  t = t1;
  return ms;
}

// a request up to CalcV: positions to the device (their bounds are reduced
// there while the host sums the charges' squares, and the set-up uses the
// staged copy), NPME_AddMissingKeywords (the RSEM grid for its tolerance),
// its reservation, the device set-up and the charges to the device
static void job_prepare(Server *S, Handle &h, Job &j, Clock::time_point &t)
{
  Keywords kw;
  kw.set("calcType", std::string("pme"));
  kw.set("funcType", j.funcType);
  kw.set("Rdir", j.Rdir);
  kw.set("tol", j.tol);
  if (j.Bn > 0) kw.set("BnOrder", j.Bn);
  if (j.funcType == "Ralpha") kw.set("alpha", j.alpha);
  if (j.funcType == "Helmholtz") {
    kw.set("k0_r", j.k0r);
    kw.set("k0_i", j.k0i);
  }
  kw.set("precision", std::string(j.fp64 ? "fp64" : "mixed"));
  const bool cxq = j.funcType == "Helmholtz";
  handle_stage_coords(h, j.n, j.coord);
  ChargeStats st;
  st.Q0 = std::sqrt(sum_squares((cxq ? 2 : 1) * j.n, j.charge));
  double mn[3], mx[3];
  handle_staged_bounds(h, mn, mx);
  for (int d = 0; d < 3; d++)
    if (!std::isfinite(mn[d]) || !std::isfinite(mx[d]))
      throw Error{NPME_ERR_ARG, "coordinates must be finite"};
  box_from_bounds(mn, mx, st.X0, st.R0);
  add_keywords_for_request(kw, j, st);
  {
    long N[3];
    const char *nk[3] = {"N1", "N2", "N3"};
    for (int d = 0; d < 3; d++) kw.get(nk[d], N[d]);
    set_reservation(S, j, footprint_bytes(j.n, N, j.fp64, cxq, j.Rdir, st.X0));
  }
  j.ms[0] = lap(t);
  handle_setup(h, kw, j.n, j.coord);
  for (int d = 0; d < 3; d++) j.N[d] = h.N[d];
  j.deviceGB = h.deviceBytes / 1e9;
  j.estGB = j.reserve / 1e9;
  if (h.deviceBytes < j.reserve) set_reservation(S, j, h.deviceBytes);
  j.ms[1] = lap(t);
  const size_t qn = (h.complexK ? 2 : 1) * j.n;
  NPME_HIP_CHECK(hipMemcpyAsync(h.d_qin, j.charge, qn * sizeof(double), hipMemcpyHostToDevice, h.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));  // CalcV reads the charges from both streams
  j.ms[2] = lap(t);
}

static void job_copy_out(Handle &h, Job &j, Clock::time_point &t)
{
  const size_t vn = (h.complexK ? 8 : 4) * j.n;
  NPME_HIP_CHECK(hipMemcpyAsync(j.V, h.d_V, vn * sizeof(double), hipMemcpyDeviceToHost, h.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  j.ms[4] = lap(t);
}

static void run_job(Server *S, Handle &h, Job &j)
{
  std::lock_guard<std::mutex> lk(h.mtx);
  auto t = Clock::now();
  job_prepare(S, h, j, t);
  handle_calcv(h, h.d_qin, h.d_V);
  j.ms[3] = lap(t);
  job_copy_out(h, j, t);
}

// the per-slot state of batches: a handle per member, a stream and events
struct BatchState {
  std::vector<std::unique_ptr<Handle>> hs;
  hipStream_t sb = nullptr;
  hipEvent_t evStart = nullptr;
  std::vector<hipEvent_t> evDone;
  ~BatchState()
  {
    for (auto &h : hs) {
      handle_free_device(*h);
      handle_destroy_streams(*h);
    }
    for (hipEvent_t e : evDone) (void)hipEventDestroy(e);
    if (evStart) (void)hipEventDestroy(evStart);
    if (sb) (void)hipStreamDestroy(sb);
  }
};

// a batch of small requests: each member's set-up on its own handle, then
// the members' CalcVs as one graph launch (they run concurrently), then the
// copies out; a member whose set-up fails does not stop the others
static void run_batch(Server *S, BatchState &B, std::vector<Job *> &batch, std::vector<int> &rc,
                      std::vector<std::string> &err)
{
  const int m = (int)batch.size();
  while ((int)B.hs.size() < m) B.hs.emplace_back(new Handle());
  if (!B.sb) NPME_HIP_CHECK(hipStreamCreateWithFlags(&B.sb, hipStreamNonBlocking));
  if (!B.evStart) NPME_HIP_CHECK(hipEventCreateWithFlags(&B.evStart, hipEventDisableTiming));
  while ((int)B.evDone.size() < m) {
    hipEvent_t e;
    NPME_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    B.evDone.push_back(e);
  }
  std::vector<Handle *> ok;
  std::vector<int> okIdx;
  std::vector<Clock::time_point> t(m);
  for (int k = 0; k < m; k++) {
    t[k] = Clock::now();
    rc[k] = guarded([&] { job_prepare(S, *B.hs[k], *batch[k], t[k]); });
    if (rc[k] != NPME_OK) {
      err[k] = last_error();
      continue;
    }
    ok.push_back(B.hs[k].get());
    okIdx.push_back(k);
  }
  if (ok.empty()) return;
  auto tc = Clock::now();
  const int r = guarded([&] {
    handle_calcv_batch(ok.data(), (int)ok.size(), B.sb, B.evStart, B.evDone.data());
  });
  const double msCalc = lap(tc);
  const std::string e = r == NPME_OK ? std::string() : last_error();
  for (int k : okIdx) {
    batch[k]->batch = (int)ok.size();
    batch[k]->ms[3] = msCalc;
    if (r != NPME_OK) {
      rc[k] = r;
      err[k] = e;
      continue;
    }
    auto t2 = Clock::now();
    rc[k] = guarded([&] { job_copy_out(*B.hs[k], *batch[k], t2); });
    if (rc[k] != NPME_OK) err[k] = last_error();
  }
}

static void worker_main(Server *S)
{
  (void)hipSetDevice(S->device);
  Handle h;
  {
    BatchState B;
    auto onTake = [S](Job &jb) {
      S->reserved += jb.reserve;
      S->running++;
      jb.tStart = S->now();
    };
    for (;;) {
      // the head of the queue is admitted when its reservation fits the
      // budget, or when nothing else runs (a request larger than the
      // budget then runs alone)
      Job *j = S->board.take(
          [S](const Job &hd) { return S->running == 0 || S->reserved + hd.reserve <= S->budget; },
          [S, &onTake](Job &jb) {
            onTake(jb);
            S->active++;
          });
      if (!j) break;
      std::vector<Job *> batch{j};
      // small requests queued behind it join it (while the budget allows)
      S->board.take_more(
          [S, j, &batch](const Job &hd) {
            return S->maxBatch > 1 && j->n <= S->batchMaxN && (int)batch.size() < S->maxBatch &&
                   hd.n <= S->batchMaxN && S->reserved + hd.reserve <= S->budget;
          },
          onTake, batch);
      std::vector<int> rc(batch.size(), NPME_OK);
      std::vector<std::string> err(batch.size());
      if (batch.size() == 1) {
        rc[0] = guarded([&] { run_job(S, h, *j); });
        if (rc[0] != NPME_OK) err[0] = last_error();
      } else {
        const int r = guarded([&] { run_batch(S, B, batch, rc, err); });
        if (r != NPME_OK)
          for (size_t k = 0; k < batch.size(); k++) {
            rc[k] = r;
            err[k] = last_error();
          }
      }
      for (size_t k = 0; k < batch.size(); k++)
        S->board.finish(batch[k], [&](Job &jb) {
          jb.status = rc[k];
          jb.err = err[k];
          jb.tEnd = S->now();
          S->reserved -= jb.reserve;
          S->running--;
          if (k + 1 == batch.size()) S->active--;
        });
    }
  }
  handle_free_device(h);
  handle_destroy_streams(h);
}

}  // namespace npme

struct npme_server {
  npme::Server s;
};

using namespace npme;

extern "C" {

int npme_server_create(npme_server **out, int nSlots, double memGB)
{
  return guarded([&] {
    if (!out) throw Error{NPME_ERR_ARG, "npme_server_create: NULL argument"};
    if (nSlots < 1 || nSlots > 64) throw Error{NPME_ERR_ARG, "nSlots must be in [1, 64]"};
    *out = nullptr;
    std::unique_ptr<npme_server> srv(new npme_server());
    Server &S = srv->s;
    NPME_HIP_CHECK(hipGetDevice(&S.device));
    size_t freeB = 0, totalB = 0;
    NPME_HIP_CHECK(hipMemGetInfo(&freeB, &totalB));
    S.budget = memGB > 0 ? (size_t)(memGB * 1e9) : (size_t)(0.8 * (double)freeB);
    for (int i = 0; i < nSlots; i++) S.workers.emplace_back(worker_main, &S);
    *out = srv.release();
  });
}

int npme_server_submit(npme_server *srv, const npme_request *r, long *ticket)
{
  return guarded([&] {
    if (!srv || !r || !ticket || !r->funcType || !r->coord || !r->charge || !r->V)
      throw Error{NPME_ERR_ARG, "npme_server_submit: NULL argument"};
    const std::string ft = r->funcType;
    if (ft != "Laplace" && ft != "Ralpha" && ft != "Helmholtz")
      throw Error{NPME_ERR_KEYWORD, "funcType = " + ft + " is undefined"};
    if (r->nCharge < 2) throw Error{NPME_ERR_ARG, "nCharge must be at least 2"};
    Server &S = srv->s;
    std::unique_ptr<Job> j(new Job());
    j->funcType = ft;
    j->alpha = r->alpha;
    j->k0r = r->k0_r;
    j->k0i = r->k0_i;
    j->tol = r->tol > 0 ? r->tol : 1e-6;
    j->Rdir = r->Rdir;
    j->Bn = r->BnOrder;
    j->fp64 = r->fp64 != 0;
    j->n = r->nCharge;
    j->coord = r->coord;
    j->charge = r->charge;
    j->V = r->V;
    j->reserve = estimate_bytes(j->n, ft == "Helmholtz", j->fp64);
    j->tSubmit = S.now();
    j->tArrive = r->tArrival >= 0 ? r->tArrival : j->tSubmit;
    if (!S.board.submit(std::move(j), *ticket)) throw Error{NPME_ERR_ARG, "server is shutting down"};
  });
}

int npme_server_wait(npme_server *srv, long ticket, npme_request_info *info)
{
  if (!srv) {
    set_error("npme_server_wait: NULL server");
    return NPME_ERR_ARG;
  }
  Server &S = srv->s;
  std::unique_ptr<Job> j;
  const int w = S.board.wait(ticket, j);
  if (w == -1) {
    set_error("npme_server_wait: unknown ticket");
    return NPME_ERR_ARG;
  }
  if (w == -2) {
    set_error("npme_server_wait: the ticket is already being waited for");
    return NPME_ERR_ARG;
  }
  if (info) {
    info->status = j->status;
    info->tArrival = j->tArrive;
    info->tSubmit = j->tSubmit;
    info->tStart = j->tStart;
    info->tEnd = j->tEnd;
    for (int d = 0; d < 3; d++) info->N[d] = j->N[d];
    info->deviceGB = j->deviceGB;
    info->msKeywords = j->ms[0];
    info->msSetup = j->ms[1];
    info->msCopyIn = j->ms[2];
    info->msCalc = j->ms[3];
    info->msCopyOut = j->ms[4];
    info->reservedGB = j->estGB;
    info->batch = j->batch;
  }
  const int rc = j->status;
  if (rc != NPME_OK) set_error(j->err);
  return rc;
}

double npme_server_clock(const npme_server *srv) { return srv ? srv->s.now() : 0.0; }

int npme_host_alloc(size_t bytes, void **p)
{
  return guarded([&] {
    if (!p) throw Error{NPME_ERR_ARG, "npme_host_alloc: NULL argument"};
    *p = nullptr;
    if (hipHostMalloc(p, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
      *p = nullptr;
      throw Error{NPME_ERR_NOMEM, "npme_host_alloc: page-locked allocation failed"};
    }
  });
}

void npme_host_free(void *p)
{
  if (p) (void)hipHostFree(p);
}

int npme_server_set_batching(npme_server *srv, int maxBatch, size_t maxCharges)
{
  return guarded([&] {
    if (!srv) throw Error{NPME_ERR_ARG, "npme_server_set_batching: NULL server"};
    if (maxBatch < 1 || maxBatch > 64) throw Error{NPME_ERR_ARG, "maxBatch must be in [1, 64]"};
    Server &S = srv->s;
    S.board.adjust([] { return true; }, [&] {
      S.maxBatch = maxBatch;
      S.batchMaxN = maxCharges;
    });
  });
}

void npme_server_destroy(npme_server *srv)
{
  if (!srv) return;
  Server &S = srv->s;
  S.board.shutdown();  // the workers finish the queued requests, then stop
  for (auto &t : S.workers) t.join();
  delete srv;  // the board releases requests that were never waited for
}

}  // extern "C"
