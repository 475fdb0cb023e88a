// Host test of the serving loop's ticket bookkeeping (request_board.h), with
// threads and no GPU: a submitting thread, worker threads that finish jobs
// in random order, and waiting threads that each wait for a share of the
// tickets while submits continue (the ticket table rehashes meanwhile).
// Checks every result, that a ticket cannot be waited for twice, and that an
// unknown ticket is refused.  Exit 0 on success.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "request_board.h"

struct Job {
  long ticket = 0;
  long value = 0;
  long result = 0;
};

int main()
{
  npme::RequestBoard<Job> board;
  const long nJobs = 20000;
  const int nWorkers = 6, nWaiters = 4;
  std::atomic<long> inFlight{0};
  std::atomic<bool> hold{false};
  std::atomic<int> failures{0};
  std::vector<std::thread> workers;
  for (int w = 0; w < nWorkers; w++)
    workers.emplace_back([&, w] {
      std::mt19937 rng(w);
      for (;;) {
        Job *j = board.take([&](const Job &hd) { return inFlight.load() < 64 && !(hold && hd.value == 7); },
                            [&](Job &) { inFlight++; });
        if (!j) break;
        if (rng() % 8 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 50));
        board.finish(j, [&](Job &jb) {
          jb.result = 3 * jb.value + 1;
          inFlight--;
        });
      }
    });
  std::vector<long> tickets(nJobs, 0);
  std::atomic<long> submitted{0};
  std::thread submitter([&] {
    for (long i = 0; i < nJobs; i++) {
      std::unique_ptr<Job> j(new Job());
      j->value = i;
      long t = 0;
      if (!board.submit(std::move(j), t)) failures++;
      tickets[i] = t;
      submitted.store(i + 1, std::memory_order_release);
    }
  });
  std::vector<std::thread> waiters;
  for (int w = 0; w < nWaiters; w++)
    waiters.emplace_back([&, w] {
      for (long i = w; i < nJobs; i += nWaiters) {
        while (submitted.load(std::memory_order_acquire) <= i) std::this_thread::yield();
        std::unique_ptr<Job> j;
        if (board.wait(tickets[i], j) != 0 || !j || j->result != 3 * i + 1) failures++;
        // the ticket is gone once waited for
        std::unique_ptr<Job> again;
        if (board.wait(tickets[i], again) != -1) failures++;
      }
    });
  submitter.join();
  for (auto &t : waiters) t.join();
  // a ticket claimed by one waiter is refused to a second one: the job is
  // held back from the workers until the second wait has been refused
  {
    hold = true;
    std::unique_ptr<Job> j(new Job());
    j->value = 7;
    long t = 0;
    board.submit(std::move(j), t);
    std::unique_ptr<Job> first;
    std::thread w1([&] {
      if (board.wait(t, first) != 0) failures++;
    });
    while (board.state(t) != 1) std::this_thread::yield();
    std::unique_ptr<Job> second;
    if (board.wait(t, second) != -2 || second) failures++;
    {
      // under the board's lock, so that no worker misses the wake-up
      std::lock_guard<std::mutex> lk(board.mtx);
      hold = false;
    }
    board.cvWork.notify_all();
    w1.join();
    if (!first || first->result != 22) failures++;
  }
  std::unique_ptr<Job> none;
  if (board.wait(123456789, none) != -1) failures++;
  board.shutdown();
  for (auto &t : workers) t.join();
  std::printf("request board: %ld jobs, %d failures\n", nJobs, failures.load());
  return failures.load() == 0 ? 0 : 1;
}
