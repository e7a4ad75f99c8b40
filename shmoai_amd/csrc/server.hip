// Bootstrapping request server: continuous batching of S2C-first bootstraps.
//
// The reference has no serving loop: BatchBootstrapper (BatchBootstrap.cuh:
// 470-820) takes one fixed batch per call, and its callers bootstrap the
// ciphertexts of one layer together. Here requests (a few ciphertexts each, of
// uneven counts) arrive at any time from any thread. A worker thread keeps the
// rotation and relinearisation keys resident and repeatedly forms a group of up
// to `max_group` ciphertexts with one scale from the admitted requests in
// arrival order, slots of different requests side by side, and runs the
// grouped bootstrap on it (bootstrap_s2c_first_batch: one container for the
// linear transforms, one batched EvalMod). A request that arrives while a group
// is on the GPU joins the next group as soon as it is admitted. Admission is by
// memory: a request enters when the estimated device footprint of all admitted
// requests stays within the budget (a request larger than the budget enters
// alone). The grouped bootstrap gives every member exactly the limbs of a
// bootstrap of that ciphertext alone, so results do not depend on the mix.
//
// All device work runs on the legacy default stream, in stream order with the
// callers' own default-stream work; outputs are ordinary default-stream
// ciphertexts that the caller owns.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "boot.h"
#include "engine.h"

namespace shm {
namespace boot {

namespace {
// Device words the grouped pipeline holds per in-flight ciphertext, in units of
// one top-level ciphertext (2 nQ N words): the container share, the hoisted
// baby-step outputs (g - 1 = 7 rotated copies), the ModUp digits of the hoisted
// rotations (beta (nQ + P) / nQ ~ 8 polynomials) and the EvalMod basis T_1..T_32
// of its two parts at lower levels.
constexpr double kWorkFactor = 24.0;
}  // namespace

BootServer::BootServer(Bootstrapper &b, const GaloisKeys &gk, const KSwitchKey &rk, size_t mem_budget,
                       size_t max_group)
    : b_(b), gk_(gk), rk_(rk), budget_(mem_budget), max_group_(max_group) {
    if (!b.ctx) throw InvalidArgument("bootstrapping plan without a context cannot run");
    if (max_group == 0 || max_group > 8) throw InvalidArgument("server: max_group must be 1..8");
    if (!b.prepared) b.prepare(nullptr);
    SHM_HIP(hipStreamSynchronize(nullptr));
    const Context &c = *b.ctx;
    ct_bytes_ = (size_t)(kWorkFactor * 2.0 * (double)c.nQ * (double)c.N * 8.0);
    th_ = std::thread([this] { loop(); });
}

BootServer::~BootServer() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    if (th_.joinable()) th_.join();
}

uint64_t BootServer::submit(const std::vector<const Ciphertext *> &in, const std::vector<Ciphertext *> &out) {
    if (in.empty() || in.size() != out.size()) throw InvalidArgument("server: input and output counts differ");
    auto r = std::make_shared<Request>();
    for (size_t i = 0; i < in.size(); i++) {
        const Ciphertext *x = in[i];
        if (!x || !out[i]) throw InvalidArgument("server: null ciphertext");
        if (x->size != 2) throw InvalidArgument("bootstrapping takes size-2 ciphertexts");
        if (x->Ql() != b_.prm.s2c_input_primes)
            throw InvalidArgument("bootstrap_s2c_first: inputs must have exactly 4 primes (lS2C+1)");
        for (size_t j = 0; j < i; j++)
            if (out[j] == out[i]) throw InvalidArgument("server: an output ciphertext appears twice");
    }
    // private copies: the caller may reuse its inputs as soon as submit returns
    for (const Ciphertext *x : in) {
        auto cl = std::make_unique<Ciphertext>();
        ct_clone(*cl, *x, nullptr);
        r->in.push_back(std::move(cl));
    }
    r->out = out;
    r->bytes = in.size() * ct_bytes_;
    r->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) throw LogicError("server is shutting down");
        r->id = ++next_id_;
        waiting_.push_back(r);
        all_[r->id] = r;
        submitted_++;
    }
    cv_work_.notify_all();
    return r->id;
}

void BootServer::wait(uint64_t id, double *latency_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = all_.find(id);
    if (it == all_.end()) throw InvalidArgument("server: unknown request id");
    std::shared_ptr<Request> r = it->second;
    cv_done_.wait(lk, [&] { return r->finished; });
    all_.erase(id);
    if (latency_ms) *latency_ms = r->latency_ms;
    if (r->error) std::rethrow_exception(r->error);
}

ServerStats BootServer::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    ServerStats st{};
    st.submitted = submitted_;
    st.completed = completed_;
    st.failed = failed_;
    st.ciphertexts = cts_done_;
    st.groups = groups_;
    st.in_flight_bytes = in_flight_;
    st.busy_s = busy_s_;
    if (!lat_.empty()) {
        std::vector<double> v = lat_;
        std::sort(v.begin(), v.end());
        // nearest-rank percentiles
        auto pct = [&](double p) {
            size_t k = (size_t)std::ceil(p * (double)v.size());
            return v[std::min(v.size(), std::max<size_t>(k, 1)) - 1];
        };
        st.p50_ms = pct(0.50);
        st.p99_ms = pct(0.99);
    }
    return st;
}

void BootServer::admit_locked() {
    while (!waiting_.empty()) {
        auto &r = waiting_.front();
        if (in_flight_ != 0 && in_flight_ + r->bytes > budget_) break;
        in_flight_ += r->bytes;
        active_.push_back(r);
        waiting_.pop_front();
    }
}

void BootServer::loop() {
    for (;;) {
        std::vector<std::pair<std::shared_ptr<Request>, size_t>> group;
        {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;) {
                admit_locked();
                double scale = 0.0;
                for (auto &r : active_) {
                    while (r->next < r->in.size() && group.size() < max_group_) {
                        const double s = r->in[r->next]->scale;
                        if (!group.empty() && s != scale) break;
                        scale = s;
                        group.emplace_back(r, r->next++);
                    }
                    if (group.size() == max_group_) break;
                }
                if (!group.empty()) break;
                if (stop_ && waiting_.empty() && active_.empty()) return;
                cv_work_.wait(lk);
            }
        }
        std::vector<const Ciphertext *> gin;
        std::vector<Ciphertext *> gout;
        for (auto &[r, i] : group) {
            gin.push_back(r->in[i].get());
            gout.push_back(r->out[i]);
        }
        std::exception_ptr err;
        const auto t0 = std::chrono::steady_clock::now();
        try {
            bootstrap_s2c_first_batch(b_, gin, gout, gk_, rk_, 0.0, nullptr);
            SHM_HIP(hipStreamSynchronize(nullptr));
        } catch (...) {
            err = std::current_exception();
        }
        const auto t1 = std::chrono::steady_clock::now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            groups_++;
            busy_s_ += std::chrono::duration<double>(t1 - t0).count();
            for (auto &[r, i] : group) {
                r->in[i].reset();
                if (err && !r->error) r->error = err;
                r->done++;
                cts_done_++;
            }
            for (auto it = active_.begin(); it != active_.end();) {
                auto &r = *it;
                // a failed request stops scheduling its remaining slots
                if (r->error && r->next < r->in.size()) {
                    r->done += r->in.size() - r->next;
                    r->next = r->in.size();
                }
                if (r->done == r->in.size()) {
                    r->finished = true;
                    r->latency_ms = std::chrono::duration<double, std::milli>(t1 - r->t_submit).count();
                    if (r->error) {
                        failed_++;
                    } else {
                        completed_++;
                        lat_.push_back(r->latency_ms);
                    }
                    in_flight_ -= r->bytes;
                    it = active_.erase(it);
                } else {
                    ++it;
                }
            }
        }
        cv_done_.notify_all();
    }
}

}  // namespace boot
}  // namespace shm
