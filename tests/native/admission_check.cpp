// Host-only checks of the served loop's admission gate (no GPU, no HIP):
// built with g++ by tests/test_serve_admission.py and run under a time limit,
// so a waiter that is never woken fails the test instead of hanging it.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <thread>

#include "admission.h"

using dyg::AdmissionGate;

namespace {

void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        std::exit(1);
    }
}

void pause_ms(int ms) { std::this_thread::sleep_for(std::chrono::milliseconds(ms)); }

constexpr int kPending = 1;  // the waiter has not returned yet

// Starts a thread that asks for `need` bytes; its status lands in *res.
std::thread waiter(std::mutex& mu, AdmissionGate& g, uint64_t need, std::atomic<int>* res,
                   uint32_t* slot) {
    return std::thread([&mu, &g, need, res, slot] {
        std::unique_lock<std::mutex> lk(mu);
        *res = g.admit(lk, need, slot);
    });
}

}  // namespace

int main() {
    // 1. Resident state grows while a request waits for memory: the waiter
    //    is turned away instead of waiting for a release that can never
    //    make room, and it takes no slot.
    {
        std::mutex mu;
        AdmissionGate g;
        g.open(2, 100, 10);
        uint32_t s1 = 99, s2 = 99;
        {
            std::unique_lock<std::mutex> lk(mu);
            expect(g.admit(lk, 50, &s1) == DYG_OK, "first request admitted");
        }
        std::atomic<int> res{kPending};
        std::thread t = waiter(mu, g, 50, &res, &s2);  // 10 + 50 + 50 > 100: waits
        pause_ms(50);
        expect(res == kPending, "second request waits for memory");
        {
            std::lock_guard<std::mutex> lk(mu);
            g.raise_resident(60);  // 60 + 50 > 100: it can never fit now
        }
        pause_ms(50);
        {
            std::lock_guard<std::mutex> lk(mu);
            g.release(s1, 50);  // in-flight work drains
        }
        t.join();
        expect(res == DYG_ERR_ADMISSION, "waiter turned away once it can never fit");
        expect(g.admitted() == 0, "the rejected request holds no bytes");
        std::unique_lock<std::mutex> lk(mu);
        uint32_t a = 9, b = 9;
        expect(g.admit(lk, 20, &a) == DYG_OK && g.admit(lk, 20, &b) == DYG_OK && a != b,
               "both slots still free after the rejection");
    }
    // 2. The same growth with the release first (the order the worker and
    //    the completer may take): still turned away, never left waiting.
    {
        std::mutex mu;
        AdmissionGate g;
        g.open(1, 100, 10);
        uint32_t s1 = 0, s2 = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            expect(g.admit(lk, 50, &s1) == DYG_OK, "admitted");
        }
        std::atomic<int> res{kPending};
        std::thread t = waiter(mu, g, 50, &res, &s2);  // no free slot: waits
        pause_ms(50);
        {
            std::lock_guard<std::mutex> lk(mu);
            g.release(s1, 50);
            g.raise_resident(70);
        }
        t.join();
        expect(res == DYG_ERR_ADMISSION || res == DYG_OK, "waiter returned");
    }
    // 3. Waiting for a slot: admitted when one frees, in its place.
    {
        std::mutex mu;
        AdmissionGate g;
        g.open(1, 1000, 0);
        uint32_t s1 = 5, s2 = 5;
        {
            std::unique_lock<std::mutex> lk(mu);
            expect(g.admit(lk, 10, &s1) == DYG_OK && s1 == 0, "slot 0 taken");
        }
        std::atomic<int> res{kPending};
        std::thread t = waiter(mu, g, 10, &res, &s2);
        pause_ms(50);
        expect(res == kPending, "waits for the slot");
        {
            std::lock_guard<std::mutex> lk(mu);
            g.release(s1, 10);
        }
        t.join();
        expect(res == DYG_OK && s2 == 0 && g.admitted() == 10, "admitted into the freed slot");
    }
    // 4. Waiting for memory: admitted once enough is released.
    {
        std::mutex mu;
        AdmissionGate g;
        g.open(4, 100, 10);
        uint32_t s1 = 0, s2 = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            expect(g.admit(lk, 80, &s1) == DYG_OK, "large request admitted");
        }
        std::atomic<int> res{kPending};
        std::thread t = waiter(mu, g, 30, &res, &s2);
        pause_ms(50);
        expect(res == kPending, "waits while 10 + 80 + 30 > 100");
        {
            std::lock_guard<std::mutex> lk(mu);
            g.release(s1, 80);
        }
        t.join();
        expect(res == DYG_OK && g.admitted() == 30, "admitted after the release");
    }
    // 5. A request that can never fit is refused at once.
    {
        std::mutex mu;
        AdmissionGate g;
        g.open(2, 100, 60);
        std::unique_lock<std::mutex> lk(mu);
        uint32_t s = 0;
        expect(g.admit(lk, 41, &s) == DYG_ERR_ADMISSION, "refused at once");
        expect(g.admit(lk, 40, &s) == DYG_OK, "exact fit admitted");
    }
    // 6. A failure or a close wakes every waiter with its status.
    for (int mode = 0; mode < 2; ++mode) {
        std::mutex mu;
        AdmissionGate g;
        g.open(1, 100, 0);
        uint32_t s1 = 0, s2 = 0, s3 = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            expect(g.admit(lk, 10, &s1) == DYG_OK, "admitted");
        }
        std::atomic<int> r2{kPending}, r3{kPending};
        std::thread t2 = waiter(mu, g, 10, &r2, &s2), t3 = waiter(mu, g, 10, &r3, &s3);
        pause_ms(50);
        expect(r2 == kPending && r3 == kPending, "both wait for the slot");
        {
            std::lock_guard<std::mutex> lk(mu);
            if (mode == 0)
                g.fail(DYG_ERR_HIP);
            else
                g.close();
        }
        t2.join();
        t3.join();
        const int want = mode == 0 ? DYG_ERR_HIP : DYG_ERR_STATE;
        expect(r2 == want && r3 == want, "waiters woken with the server's status");
        std::unique_lock<std::mutex> lk(mu);
        uint32_t s = 0;
        expect(g.admit(lk, 1, &s) == want, "later submits fail at once");
    }
    std::printf("admission checks passed\n");
    return 0;
}
