// Bootstrapping request server: continuous batching of S2C-first bootstraps.
//
// The reference has no serving loop: BatchBootstrapper (BatchBootstrap.cuh:
// 470-820) takes one fixed batch per call, and its callers bootstrap the
// ciphertexts of one layer together. Here requests arrive at any time from any
// thread, in two kinds:
//   - ordinary requests (a few ciphertexts each, of uneven counts): a worker forms
//     a group of up to `max_group` ciphertexts with one scale from the admitted
//     requests in arrival order, slots of different requests side by side, and
//     runs the grouped bootstrap on it (bootstrap_s2c_first_batch: one container
//     for the linear transforms, one batched EvalMod). Each member gets exactly
//     the limbs of a bootstrap of that ciphertext alone;
//   - packed requests (the column ciphertexts of one input in the MOAI layout,
//     data in slots [0, L), L = its token count, zeros elsewhere;
//     Batch_encode_encrypt.cuh:12-57): requests with the same column count and
//     scale share ciphertexts. Each takes a slot block of the next power of two at
//     or above L, aligned to its size, is rotated there (a key switch at the
//     4-prime input level) and added into the shared ciphertexts, which are
//     bootstrapped once for all of them; a request receives the bootstrapped
//     shared ciphertexts with its slots at [offset, offset + L), the slots its
//     bias_vec mask selects.
// A request that arrives while others run joins the next group or pack as soon
// as it is admitted. Admission is by memory: a request enters while the estimated
// device footprint of the admitted requests stays within the cap, the smaller of
// the caller's budget (if any) and the device's free memory (hipMemGetInfo) less
// a reserve; a request larger than the cap enters alone. Several workers, each
// with its own stream (default priority), run groups and packs concurrently, so
// one group's kernels overlap another's. Inputs are copied on the default stream
// at submit and outputs are delivered on the default stream, in stream order
// with the callers' own default-stream work.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>

#include "boot.h"
#include "engine.h"

namespace shm {
namespace boot {

namespace {
// Device words the grouped pipeline holds per in-flight ciphertext, in units of
// one top-level ciphertext (2 nQ N words): the container share, the hoisted
// baby-step outputs (g - 1 = 7 rotated copies), the ModUp digits of the hoisted
// rotations (beta (nQ + P) / nQ ~ 8 polynomials) and the EvalMod basis T_1..T_8,
// T_16, T_32 of its two parts at lower levels.
constexpr double kWorkFactor = 24.0;
// device memory kept free beyond the admitted footprint (at most 1/8 of the device)
constexpr size_t kReserveBytes = size_t(8) << 30;

size_t pow2_at_least(size_t v) {
    size_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// order work on `s` after everything queued so far on the default stream
void after_default_stream(hipStream_t s) {
    hipEvent_t ev;
    SHM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    SHM_HIP(hipEventRecord(ev, nullptr));
    SHM_HIP(hipStreamWaitEvent(s, ev, 0));
    SHM_HIP(hipEventDestroy(ev));
}
}  // namespace

BootServer::BootServer(Bootstrapper &b, const GaloisKeys &gk, const KSwitchKey &rk, size_t mem_budget,
                       size_t max_group, size_t workers)
    : b_(b), gk_(gk), rk_(rk), budget_(mem_budget), max_group_(max_group) {
    if (!b.ctx) throw InvalidArgument("bootstrapping plan without a context cannot run");
    if (max_group == 0 || max_group > 8) throw InvalidArgument("server: max_group must be 1..8");
    if (workers == 0 || workers > 4) throw InvalidArgument("server: 1..4 workers");
    if (!b.prepared) b.prepare(nullptr);
    SHM_HIP(hipStreamSynchronize(nullptr));
    const Context &c = *b.ctx;
    ct_bytes_ = (size_t)(kWorkFactor * 2.0 * (double)c.nQ * (double)c.N * 8.0);
    streams_.resize(workers, nullptr);
    for (auto &st : streams_) SHM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (size_t w = 0; w < workers; w++) th_.emplace_back([this, w] { loop(w); });
}

BootServer::~BootServer() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &t : th_)
        if (t.joinable()) t.join();
    // every temporary of the workers was released on its stream before they stopped
    for (hipStream_t st : streams_) {
        if (!st) continue;
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
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

uint64_t BootServer::submit_packed(const std::vector<const Ciphertext *> &in, size_t slots,
                                   const std::vector<Ciphertext *> &out) {
    const Context &c = *b_.ctx;
    const size_t n = c.N / 2;
    if (slots == 0 || slots > n) throw InvalidArgument("server: packed slots must be 1..N/2");
    if (in.empty() || in.size() != out.size()) throw InvalidArgument("server: input and output counts differ");
    // offsets are multiples of the block; moving a request there rotates by
    // n - offset, which rotate_vector splits into signed powers of two of at least
    // the block (evaluate.cu:1650-1706): every such key must be present
    const size_t blk = pow2_at_least(slots);
    for (size_t p = blk; p < n; p <<= 1)
        for (int sg : {1, -1})
            if (!gk_.keys.count(c.elt_from_step(sg * (int)p)))
                throw InvalidArgument("server: packed requests need Galois keys for the steps +-" +
                                      std::to_string(p) + " (signed powers of two from the slot block up)");
    const uint64_t id = submit(in, out);
    std::lock_guard<std::mutex> lk(mu_);
    auto it = all_.find(id);
    if (it != all_.end() && it->second->next == 0) {
        it->second->slots = slots;
        // its share of the shared ciphertexts
        it->second->bytes = std::max<size_t>(1, in.size() * ct_bytes_ * blk / n);
    }
    return id;
}

void BootServer::wait(uint64_t id, double *latency_ms, size_t *slot_offset) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = all_.find(id);
    if (it == all_.end()) throw InvalidArgument("server: unknown request id");
    std::shared_ptr<Request> r = it->second;
    cv_done_.wait(lk, [&] { return r->finished; });
    all_.erase(id);
    if (latency_ms) *latency_ms = r->latency_ms;
    if (slot_offset) *slot_offset = r->offset;
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
    st.packed_requests = packed_done_;
    st.shared_cts = shared_cts_;
    st.free_bytes = free_bytes_;
    st.admit_cap_bytes = cap_bytes_;
    st.workers = (uint32_t)th_.size();
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
    if (waiting_.empty()) return;
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) == hipSuccess) {
        free_bytes_ = freeb;
    } else {
        (void)hipGetLastError();
        freeb = free_bytes_;
    }
    const size_t reserve = total ? std::min(kReserveBytes, total / 8) : kReserveBytes;
    // what the device can still take: its free memory less the reserve, plus the
    // admitted requests' share (some of it may not be allocated yet)
    size_t cap = (freeb > reserve ? freeb - reserve : 0) + in_flight_;
    if (budget_) cap = std::min(cap, budget_);
    cap_bytes_ = cap;
    while (!waiting_.empty()) {
        auto &r = waiting_.front();
        if (in_flight_ != 0 && in_flight_ + r->bytes > cap) break;
        in_flight_ += r->bytes;
        active_.push_back(r);
        waiting_.pop_front();
    }
}

void BootServer::finish_locked(const std::chrono::steady_clock::time_point &t1) {
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
                if (r->slots) packed_done_++;
                lat_.push_back(r->latency_ms);
            }
            in_flight_ -= r->bytes;
            it = active_.erase(it);
        } else {
            ++it;
        }
    }
}

void BootServer::run_group(Group &group, hipStream_t s) {
    Context &c = *b_.ctx;
    after_default_stream(s);  // the inputs were copied there at submit
    std::vector<const Ciphertext *> gin;
    std::vector<std::unique_ptr<Ciphertext>> tmp;
    std::vector<Ciphertext *> gout;
    for (auto &[r, i] : group) {
        gin.push_back(r->in[i].get());
        tmp.push_back(std::make_unique<Ciphertext>());
        tmp.back()->ctx = &c;
        gout.push_back(tmp.back().get());
    }
    bootstrap_s2c_first_batch(b_, gin, gout, gk_, rk_, 0.0, s);
    SHM_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < group.size(); k++) ct_clone(*group[k].first->out[group[k].second], *tmp[k], nullptr);
    // the temporaries return to the pool on s, whose work has drained
}

void BootServer::run_pack(Pack &pack, hipStream_t s) {
    Context &c = *b_.ctx;
    const size_t n = c.N / 2, m = pack[0]->in.size();
    after_default_stream(s);
    std::vector<std::unique_ptr<Ciphertext>> shared(m);
    for (size_t i = 0; i < m; i++) {
        shared[i] = std::make_unique<Ciphertext>();
        shared[i]->ctx = &c;
        for (size_t k = 0; k < pack.size(); k++) {
            const Request &r = *pack[k];
            if (k == 0) {
                ct_clone(*shared[i], *r.in[i], s);
                if (r.offset) rotate_internal(c, *shared[i], (int)(n - r.offset), gk_, s, false);
                continue;
            }
            Ciphertext rot;
            ct_clone(rot, *r.in[i], s);
            if (r.offset) rotate_internal(c, rot, (int)(n - r.offset), gk_, s, false);
            ct_add(c, *shared[i], rot, s);
        }
    }
    for (size_t i0 = 0; i0 < m; i0 += max_group_) {
        const size_t cnt = std::min(max_group_, m - i0);
        std::vector<const Ciphertext *> gin;
        std::vector<std::unique_ptr<Ciphertext>> tmp;
        std::vector<Ciphertext *> gout;
        for (size_t i = i0; i < i0 + cnt; i++) {
            gin.push_back(shared[i].get());
            tmp.push_back(std::make_unique<Ciphertext>());
            tmp.back()->ctx = &c;
            gout.push_back(tmp.back().get());
        }
        bootstrap_s2c_first_batch(b_, gin, gout, gk_, rk_, 0.0, s);
        SHM_HIP(hipStreamSynchronize(s));
        for (auto &r : pack)
            for (size_t i = i0; i < i0 + cnt; i++) ct_clone(*r->out[i], *tmp[i - i0], nullptr);
    }
}

void BootServer::loop(size_t w) {
    const hipStream_t s = streams_[w];
    for (;;) {
        Group group;
        Pack pack;
        {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;) {
                admit_locked();
                // the oldest admitted request with unscheduled work decides the job
                std::shared_ptr<Request> first;
                for (auto &r : active_)
                    if (r->next < r->in.size()) {
                        first = r;
                        break;
                    }
                if (first && first->slots) {
                    // a pack: packed requests of the first one's column count and
                    // scale, in arrival order, at aligned offsets until the slots run out
                    const size_t n = b_.ctx->N / 2;
                    size_t cursor = 0;
                    for (auto &r : active_) {
                        if (!r->slots || r->next != 0) continue;
                        if (r->in.size() != first->in.size() || r->in[0]->scale != first->in[0]->scale) continue;
                        const size_t blk = pow2_at_least(r->slots);
                        const size_t o = (cursor + blk - 1) / blk * blk;
                        if (o + blk > n) break;
                        r->offset = o;
                        cursor = o + blk;
                        r->next = r->in.size();
                        pack.push_back(r);
                    }
                    break;
                }
                if (first) {
                    double scale = 0.0;
                    for (auto &r : active_) {
                        if (r->slots) continue;
                        while (r->next < r->in.size() && group.size() < max_group_) {
                            const double sc = r->in[r->next]->scale;
                            if (!group.empty() && sc != scale) break;
                            scale = sc;
                            group.emplace_back(r, r->next++);
                        }
                        if (group.size() == max_group_) break;
                    }
                    if (!group.empty()) break;
                }
                if (stop_ && waiting_.empty() && !first) return;
                cv_work_.wait(lk);
            }
        }
        std::exception_ptr err;
        const auto t0 = std::chrono::steady_clock::now();
        try {
            if (!pack.empty())
                run_pack(pack, s);
            else
                run_group(group, s);
        } catch (...) {
            err = std::current_exception();
            (void)hipStreamSynchronize(s);
        }
        const auto t1 = std::chrono::steady_clock::now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            busy_s_ += std::chrono::duration<double>(t1 - t0).count();
            if (!pack.empty()) {
                const size_t m = pack[0]->in.size();
                groups_ += (m + max_group_ - 1) / max_group_;
                shared_cts_ += m;
                cts_done_ += m;
                for (auto &r : pack) {
                    for (auto &x : r->in) x.reset();
                    if (err && !r->error) r->error = err;
                    r->done = r->in.size();
                }
            } else {
                groups_++;
                for (auto &[r, i] : group) {
                    r->in[i].reset();
                    if (err && !r->error) r->error = err;
                    r->done++;
                    cts_done_++;
                }
            }
            finish_locked(t1);
        }
        cv_done_.notify_all();
        cv_work_.notify_all();  // memory freed: admission may proceed for the others
    }
}

}  // namespace boot
}  // namespace shm
