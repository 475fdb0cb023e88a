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
  double deviceGB = 0;
  double ms[5] = {0, 0, 0, 0, 0};  // keywords, set-up, copy in, CalcV, copy out
};

}  // namespace

// the admission state (reserved bytes, running requests) is guarded by the
// board's mutex: it is read and changed only in the board's callbacks
struct Server {
  RequestBoard<Job> board;
  std::vector<std::thread> workers;
  size_t budget = 0, reserved = 0;
  int running = 0;
  int device = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const
  {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

// device footprint of a request: per-charge state and neighbour list (about
// 2.5 KB per charge at the densities of the serving stream) plus the grids
static size_t estimate_bytes(size_t n, bool complexK, bool fp64)
{
  const double perCharge = 2500.0 * (fp64 ? 1.5 : 1.0) * (complexK ? 1.5 : 1.0);
  return (size_t)(perCharge * (double)n) + (64u << 20);
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

static void run_job(Handle &h, Job &j)
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
  auto t = std::chrono::steady_clock::now();
  auto lap = [&t](double &ms) {
    const auto t1 = std::chrono::steady_clock::now();
    ms = std::chrono::duration<double, std::milli>(t1 - t).count();
    t = t1;
  };
  // the positions go to the device first and their bounds are reduced
  // there while the host sums the charges' squares; the grid choice then
  // reads no position on the host, and the set-up uses the staged copy
  std::lock_guard<std::mutex> lk(h.mtx);
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
  lap(j.ms[0]);
  handle_setup(h, kw, j.n, j.coord);
  for (int d = 0; d < 3; d++) j.N[d] = h.N[d];
  j.deviceGB = h.deviceBytes / 1e9;
  lap(j.ms[1]);
  const bool cx = h.complexK;
  const size_t qn = (cx ? 2 : 1) * j.n, vn = (cx ? 8 : 4) * j.n;
  NPME_HIP_CHECK(hipMemcpyAsync(h.d_qin, j.charge, qn * sizeof(double), hipMemcpyHostToDevice, h.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));  // CalcV reads the charges from both streams
  lap(j.ms[2]);
  handle_calcv(h, h.d_qin, h.d_V);
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  lap(j.ms[3]);
  NPME_HIP_CHECK(hipMemcpyAsync(j.V, h.d_V, vn * sizeof(double), hipMemcpyDeviceToHost, h.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  lap(j.ms[4]);
}

static void worker_main(Server *S)
{
  (void)hipSetDevice(S->device);
  Handle h;
  for (;;) {
    // the head of the queue is admitted when its reservation fits the
    // budget, or when nothing else runs (a request larger than the budget
    // then runs alone)
    Job *j = S->board.take(
        [S](const Job &hd) { return S->running == 0 || S->reserved + hd.reserve <= S->budget; },
        [S](Job &jb) {
          S->reserved += jb.reserve;
          S->running++;
          jb.tStart = S->now();
        });
    if (!j) break;
    std::string err;
    int rc = guarded([&] { run_job(h, *j); });
    if (rc != NPME_OK) err = last_error();
    S->board.finish(j, [&](Job &jb) {
      jb.status = rc;
      jb.err = err;
      jb.tEnd = S->now();
      S->reserved -= jb.reserve;
      S->running--;
    });
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

void npme_server_destroy(npme_server *srv)
{
  if (!srv) return;
  Server &S = srv->s;
  S.board.shutdown();  // the workers finish the queued requests, then stop
  for (auto &t : S.workers) t.join();
  delete srv;  // the board releases requests that were never waited for
}

}  // extern "C"
