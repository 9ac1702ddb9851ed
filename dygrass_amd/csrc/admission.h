// admission.h -- memory-budget admission of the served loop (host code only,
// no HIP: it is compiled by itself for the host-side tests).
//
// A request is admitted when a slot is free and the resident device state
// plus the worst-case growth of every admitted request, its own included,
// fits the budget.  Resident state can grow while submitters wait (the
// worker grows the graph arenas), so a waiting request may come to need
// more than the budget can ever give it; the gate then wakes it and turns it
// away instead of letting it wait for memory that will never be released.
// A failed or closing server wakes every waiter too.
//
// The gate does no locking of its own: every call is made with the caller's
// mutex held (the mutex that also orders the queue of admitted requests, so
// admission order and application order stay one order), and `admit` waits
// on that mutex's lock.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/dygrass.h"

namespace dyg {

class AdmissionGate {
public:
    void open(uint32_t slots, uint64_t budget, uint64_t resident) {
        busy_.assign(slots, false);
        budget_ = budget;
        resident_ = resident;
        admitted_ = 0;
        failed_ = DYG_OK;
        closed_ = false;
    }

    // Blocks (releasing lk) until the request can be admitted, then takes a
    // slot and the request's `need` bytes.  DYG_ERR_ADMISSION when it can
    // never fit (at once, or once resident state has grown while it waited);
    // the server's failure status when it failed meanwhile; DYG_ERR_STATE
    // when the server is closing.  No slot is taken unless DYG_OK.
    dyg_status admit(std::unique_lock<std::mutex>& lk, uint64_t need, uint32_t* slot) {
        dyg_status st = DYG_OK;
        cv_.wait(lk, [&] {
            st = verdict(need);
            return st != kWait;
        });
        if (st != DYG_OK) return st;
        uint32_t k = 0;
        while (busy_[k]) ++k;  // verdict() saw a free slot
        busy_[k] = true;
        admitted_ += need;
        *slot = k;
        return DYG_OK;
    }

    // A request has finished: its slot and bytes are free again.
    void release(uint32_t slot, uint64_t need) {
        busy_[slot] = false;
        admitted_ -= need;
        cv_.notify_all();
    }

    // Resident state is at least r now (arena growth is resident from then
    // on); waiters re-check whether they can still fit.
    void raise_resident(uint64_t r) {
        if (r > resident_) resident_ = r;
        cv_.notify_all();
    }

    void fail(dyg_status st) {
        if (failed_ == DYG_OK) failed_ = st;
        cv_.notify_all();
    }

    void close() {
        closed_ = true;
        cv_.notify_all();
    }

    uint64_t resident() const { return resident_; }
    uint64_t admitted() const { return admitted_; }
    uint64_t budget() const { return budget_; }
    dyg_status failed() const { return failed_; }

private:
    static constexpr dyg_status kWait = 1;

    dyg_status verdict(uint64_t need) const {
        if (failed_ != DYG_OK) return failed_;
        if (closed_) return DYG_ERR_STATE;
        if (resident_ + need > budget_) return DYG_ERR_ADMISSION;
        if (resident_ + admitted_ + need > budget_) return kWait;
        for (bool b : busy_)
            if (!b) return DYG_OK;
        return kWait;
    }

    std::condition_variable cv_;
    std::vector<bool> busy_;
    uint64_t budget_ = 0, resident_ = 0, admitted_ = 0;
    dyg_status failed_ = DYG_OK;
    bool closed_ = false;
};

}  // namespace dyg
