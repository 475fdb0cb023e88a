// Ticket bookkeeping of the serving loop: a FIFO of submitted requests and a
// table of tickets.  Plain C++ (no HIP), so the bookkeeping is tested on the
// host with threads (tests/cpp/request_board_host.cpp).
//
// A ticket is waited for at most once: wait() claims it under the lock, so
// a second waiter gets an error instead of a request that the first waiter
// is about to delete; the entry is erased by key after the wait, never
// through an iterator taken before it (a submit in between may rehash the
// table).
#pragma once

#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <unordered_map>

namespace npme {

template <typename Job> class RequestBoard {
 public:
  std::mutex mtx;
  std::condition_variable cvWork, cvDone;
  bool stop = false;

  // queues a job and returns its ticket (false: the board is stopping)
  bool submit(std::unique_ptr<Job> j, long &ticket)
  {
    {
      std::lock_guard<std::mutex> lk(mtx);
      if (stop) return false;
      ticket = next_++;
      j->ticket = ticket;
      Entry &e = table_[ticket];
      e.job = std::move(j);
      queue_.push_back(e.job.get());
    }
    cvWork.notify_all();
    return true;
  }

  // blocks until the head of the queue is admitted by ready (run under the
  // lock), then removes it and runs onTake on it under the lock; returns
  // nullptr when the board stops with an empty queue
  Job *take(const std::function<bool(const Job &)> &ready, const std::function<void(Job &)> &onTake)
  {
    std::unique_lock<std::mutex> lk(mtx);
    cvWork.wait(lk, [&] { return (stop && queue_.empty()) || (!queue_.empty() && ready(*queue_.front())); });
    if (queue_.empty()) return nullptr;
    Job *j = queue_.front();
    queue_.pop_front();
    onTake(*j);
    return j;
  }

  // further jobs at the head of the queue admitted by ready, without
  // waiting (a batch); onTake runs under the lock for each
  template <typename Out>
  void take_more(const std::function<bool(const Job &)> &ready, const std::function<void(Job &)> &onTake,
                 Out &out)
  {
    std::lock_guard<std::mutex> lk(mtx);
    while (!queue_.empty() && ready(*queue_.front())) {
      Job *j = queue_.front();
      queue_.pop_front();
      onTake(*j);
      out.push_back(j);
    }
  }

  // runs change under the lock once ok holds (both run under the lock;
  // finish and adjust wake the waiters), then wakes the workers
  void adjust(const std::function<bool()> &ok, const std::function<void()> &change)
  {
    {
      std::unique_lock<std::mutex> lk(mtx);
      cvWork.wait(lk, ok);
      change();
    }
    cvWork.notify_all();
  }

  // marks a job done (fin runs under the lock) and wakes its waiter and the
  // workers (admission may have changed)
  void finish(Job *j, const std::function<void(Job &)> &fin)
  {
    {
      std::lock_guard<std::mutex> lk(mtx);
      fin(*j);
      table_[j->ticket].done = true;
    }
    cvDone.notify_all();
    cvWork.notify_all();
  }

  // waits for a ticket and hands its job to the caller: 0 ok, -1 unknown
  // ticket, -2 already claimed by another waiter
  int wait(long ticket, std::unique_ptr<Job> &out)
  {
    std::unique_lock<std::mutex> lk(mtx);
    auto it = table_.find(ticket);
    if (it == table_.end()) return -1;
    if (it->second.claimed) return -2;
    it->second.claimed = true;
    cvDone.wait(lk, [&] { return table_.at(ticket).done; });
    out = std::move(table_.at(ticket).job);
    table_.erase(ticket);
    return 0;
  }

  void shutdown()
  {
    {
      std::lock_guard<std::mutex> lk(mtx);
      stop = true;
    }
    cvWork.notify_all();
  }

  // -1 unknown ticket, 0 not yet waited for, 1 claimed by a waiter
  int state(long ticket)
  {
    std::lock_guard<std::mutex> lk(mtx);
    auto it = table_.find(ticket);
    if (it == table_.end()) return -1;
    return it->second.claimed ? 1 : 0;
  }

  size_t queued()
  {
    std::lock_guard<std::mutex> lk(mtx);
    return queue_.size();
  }

 private:
  struct Entry {
    std::unique_ptr<Job> job;
    bool done = false, claimed = false;
  };
  std::deque<Job *> queue_;
  std::unordered_map<long, Entry> table_;
  long next_ = 1;
};

}  // namespace npme
