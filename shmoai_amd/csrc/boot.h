// CKKS bootstrapping (slot-to-coefficient-first): plan (host FP64) and the
// device pipeline that executes it.
//
// Slot convention of the encoder: slot j of a plaintext m(X) is
// m(w_j) / scale with w_j = exp(2 pi i 5^j / 2N), j < n = N/2. Writing
// m = m_lo + X^n m_hi gives z = U0 (m_lo + i m_hi), U0[j][k] = w_j^k.
// U0 = B_logn ... B_1 R, where R is the bit reversal and B_s the radix-2
// butterfly stage of block length 2^s (fft.cu special FFT, fftSpecial
// form). Slot-to-coefficient applies T = B_logn ... B_1 (so the
// coefficients come out bit-reversed) and coefficient-to-slot applies
// T^-1; the bit reversals cancel because EvalMod works slot by slot.
// Each transform is split into three layers of consecutive stages
// (Bootstrapper.cu:2737-2777 sfl_full_3, 2891-2913 sflinv_full_3 level
// split), evaluated with baby-step / giant-step rotations.
#pragma once

#include <chrono>
#include <complex>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.h"

namespace shm {
namespace boot {

using cplx = std::complex<double>;

// Matrix on slot vectors by generalised diagonals:
// (A z)[j] = sum_d diag[d][j] * z[(j + d) mod n].
struct DiagMat {
    size_t n = 0;
    std::map<size_t, std::vector<cplx>> d;
};

DiagMat fft_stage(size_t n, int s, bool inverse);
DiagMat compose(const DiagMat &a, const DiagMat &b);  // a * b (b applied first)
DiagMat stages_product(size_t n, int first, int last, bool inverse);

// One linear-transform layer: terms k = k0 + g*i + j (j < g, i < n_giants)
// with slot offset k*base; plaintext (i, j) holds rot(diag_k, -(k0 + g*i)*base).
struct LtLayer {
    size_t base = 1;
    long k0 = 0;
    int g = 1;
    int n_giants = 1;
    double factor = 1.0;
    size_t chain = 0;          // chain index the layer is applied at
    std::vector<std::vector<std::vector<cplx>>> pre;  // [i][j] (empty = zero)
    std::vector<std::vector<DevBuf>> pts;             // encoded at scale q_last of `chain`
    int baby_step(int j) const { return (int)(j * base); }
    long giant_offset(int i) const { return (k0 + (long)g * i) * (long)base; }
};

// Build one layer from stages [first, last] (1-based, inclusive).
LtLayer make_layer(size_t n, int first, int last, bool inverse, double factor);

// p(y) ~ cos(2 pi (K y - 1/4) / 2^r) on [-1, 1]: Chebyshev coefficients.
std::vector<double> evalmod_chebyshev(int K, int r, int deg);
// p = q + T_m r (Chebyshev division by T_m); requires deg p < 2m.
void cheb_split(const std::vector<double> &p, size_t m, std::vector<double> &q, std::vector<double> &r);

struct BootParams {
    uint32_t s2c_input_primes = 4;  // lS2C + 1 (BatchBootstrap.cuh:727-728)
    double layer3_factor = 1.0 / 16.0;
    int K = 25;
    int deg = 63;
    int double_angles = 2;
    double input_bound = 1.0;  // bound on |Re z|, |Im z| of the input slots
};

struct Bootstrapper {
    Context *ctx = nullptr;
    BootParams prm;
    size_t n = 0;
    int logn = 0;
    LtLayer s2c[3], c2s[3];
    std::vector<double> cheb;
    bool prepared = false;

    std::vector<int> rotation_steps() const;  // 0 = conjugation
    void prepare(hipStream_t s);              // encode every diagonal on the device
};

std::unique_ptr<Bootstrapper> make_bootstrapper(Context &c, const BootParams &p);
// Host-only plan (no context: diagonals, schedule and coefficients, no device work).
std::unique_ptr<Bootstrapper> make_plan(size_t N, size_t nQ, const BootParams &p);

// Device pipeline. A RotateFn rotates a ciphertext (or shared-a batch) in place
// by a slot step; step 0 is complex conjugation.
using RotateFn = std::function<void(Ciphertext &, int)>;
// Several rotations of one input sharing one ModUp: outs[i] = rot(in, steps[i]).
using HoistFn = std::function<void(const Ciphertext &, const std::vector<int> &, const std::vector<Ciphertext *> &)>;
struct Rotator {
    RotateFn rot;
    HoistFn hoist;  // optional
};
void lt_apply(Context &c, const LtLayer &L, Ciphertext &ct, const Rotator &r, hipStream_t s);
void slot_to_coeff(Bootstrapper &b, Ciphertext &ct, const Rotator &r, hipStream_t s);
void slot_to_coeff(Bootstrapper &b, Ciphertext &ct, const GaloisKeys &gk, hipStream_t s);
void coeff_to_slot(Bootstrapper &b, Ciphertext &ct, Ciphertext &re, Ciphertext &im, const Rotator &r,
                   hipStream_t s);
void coeff_to_slot(Bootstrapper &b, Ciphertext &ct, Ciphertext &re, Ciphertext &im, const GaloisKeys &gk,
                   hipStream_t s);
void eval_mod(Bootstrapper &b, Ciphertext &x, const KSwitchKey &rk, double out_const, double out_scale,
              hipStream_t s);
// eval_mod of up to 32 ciphertexts at one level and scale, batched (bit-identical
// to eval_mod per ciphertext)
void eval_mod_batch(Bootstrapper &b, const std::vector<Ciphertext *> &xs, const KSwitchKey &rk, double out_const,
                    double out_scale, hipStream_t s);
// sum_k coeffs[k] T_k((2 x - lo - hi) / (hi - lo)) on up to 32 ciphertexts at one
// level and scale, in place, in lockstep (baby-step/giant-step Chebyshev, one
// batched key switch and rescale per product); the result keeps the input scale.
void chebyshev_eval_batch(Context &c, const std::vector<Ciphertext *> &xs, const std::vector<double> &coeffs,
                          double lo, double hi, const KSwitchKey &rk, hipStream_t s);
void bootstrap_s2c_first(Bootstrapper &b, const Ciphertext &in, Ciphertext &out, const GaloisKeys &gk,
                         const KSwitchKey &rk, double out_scale, hipStream_t s);
// bootstrap_s2c_first of several ciphertexts: groups of up to 8 inputs with one
// scale run their linear transforms on one container and share one batched
// EvalMod over their real and imaginary parts.
// out_scale 0 keeps each input's scale.
void bootstrap_s2c_first_batch(Bootstrapper &b, const std::vector<const Ciphertext *> &in,
                               const std::vector<Ciphertext *> &out, const GaloisKeys &gk, const KSwitchKey &rk,
                               double out_scale, hipStream_t s);
// ---- shared-a batches (shared_a.hip) -------------------------------------------------
// SharedACiphertextBatch (SharedACiphertext.cuh:22-99): ct has size n + 1;
// polynomials 0..n-1 are the b_i, polynomial n the shared a (NTT form).
struct SharedABatch {
    size_t n = 0;
    Ciphertext ct;
};

// SharedAGaloisKey (ShRotate.cuh:25-35): A [dnum][QP][N]; B [dnum][n][QP][N].
struct SharedGaloisKey {
    uint32_t elt = 0;
    DevBuf A, B;
};

// SharedAKeyGen + SharedAFormatKey (FmtSwitch.cuh:28-105): auxiliary secrets,
// the forward format key at fmt_chain (QlP, local limb order), the backward
// switching keys s'_i -> s and the shared-a Galois keys (all kept resident).
struct SharedAKeys {
    Context *ctx = nullptr;
    size_t n = 0, fmt_chain = 0;
    uint32_t hw = 0;
    uint64_t seed = 0;
    std::vector<SecretKey> aux;
    std::vector<std::vector<int8_t>> aux_coeffs;
    DevBuf a_fmt, B_fmt;                         // [n][QlP][N], [n][n][QlP][N] (index j * n + i)
    DevBuf ext_inv, ext_inv_s, ext_cg, ext_qmodp;  // exact Ql -> P extension constants
    std::vector<std::unique_ptr<KSwitchKey>> backward;
    std::map<uint32_t, std::unique_ptr<SharedGaloisKey>> galois;
};

std::unique_ptr<SharedAKeys> make_shared_keys(Context &c, const SecretKey &sk, size_t n, uint32_t hw,
                                              uint64_t seed, size_t fmt_chain, hipStream_t s);
void add_shared_galois_key(SharedAKeys &K, int step, uint64_t seed, hipStream_t s);
// `rounded`: unbiased ModDown (used inside bootstrapping); default = reference semantics
void fmt_switch_forward(SharedAKeys &K, const std::vector<const Ciphertext *> &cts, SharedABatch &out,
                        hipStream_t s, bool rounded = false);
void fmt_switch_backward(SharedAKeys &K, const SharedABatch &in, const std::vector<Ciphertext *> &out,
                         hipStream_t s, bool rounded = false);
void shared_apply_galois(SharedAKeys &K, SharedABatch &x, uint32_t elt, hipStream_t s, bool rounded = false);
void shared_rotate(SharedAKeys &K, SharedABatch &x, int step, hipStream_t s, bool rounded = false);
// hoisted Sh-Rot: one ModUp of a for all steps; outs are batch blocks of size n + 1
void shared_rotate_hoisted(SharedAKeys &K, const Ciphertext &x, const std::vector<int> &steps,
                           const std::vector<Ciphertext *> &outs, hipStream_t s, bool rounded);
void shared_slot_to_coeff(Bootstrapper &b, SharedAKeys &K, SharedABatch &x, hipStream_t s);
void shared_coeff_to_slot(Bootstrapper &b, SharedAKeys &K, SharedABatch &x, SharedABatch &re, SharedABatch &im,
                          hipStream_t s);
void bootstrap_shared_a_s2c_first(Bootstrapper &b, SharedAKeys &K, const std::vector<const Ciphertext *> &in,
                                  const std::vector<Ciphertext *> &out, const KSwitchKey &rk, double out_scale,
                                  hipStream_t s);

// Graph-captured chunks (boot_graph.hip): `body` bootstraps one chunk of fixed
// inputs into fixed outputs on a stream; it is run once eagerly, then recorded into
// a graph on that (created, non-default) stream. run_boot_graph replays the graph for
// every chunk of a larger input list (inputs copied into the staging buffers, outputs
// copied out), on the recording stream, enqueue-only.
using ChunkFn = std::function<void(const std::vector<const Ciphertext *> &, const std::vector<Ciphertext *> &,
                                   hipStream_t)>;
struct BootGraph {
    Context *ctx = nullptr;
    hipStream_t stream = nullptr;
    size_t chunk = 0, in_chain = 0, nodes = 0;
    double in_scale = 0.0;
    std::vector<std::unique_ptr<Ciphertext>> in, out;  // staging buffers the graph reads / writes
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BootGraph();
};
std::unique_ptr<BootGraph> make_boot_graph(Context &c, size_t chunk, size_t in_chain, double in_scale,
                                           const ChunkFn &body, hipStream_t s);
void run_boot_graph(BootGraph &g, const std::vector<const Ciphertext *> &in, const std::vector<Ciphertext *> &out);

// Request server with continuous batching of S2C-first bootstraps (server.hip).
struct ServerStats {
    uint64_t submitted, completed, failed, ciphertexts, groups;
    size_t in_flight_bytes;
    double p50_ms, p99_ms;  // request latency, submit to completion (completed requests)
    double busy_s;          // wall time spent in grouped bootstraps (summed over workers)
    uint64_t packed_requests, shared_cts;  // packed requests completed, shared ciphertexts bootstrapped
    size_t free_bytes, admit_cap_bytes;    // device memory free and the admission cap at the last admission
    uint32_t workers;
};

class BootServer {
   public:
    // mem_budget 0: admission by the device's free memory alone; workers 1..4, each
    // with its own stream, run groups concurrently
    BootServer(Bootstrapper &b, const GaloisKeys &gk, const KSwitchKey &rk, size_t mem_budget, size_t max_group,
               size_t workers = 1);
    ~BootServer();  // finishes every submitted request, then stops the workers
    BootServer(const BootServer &) = delete;
    BootServer &operator=(const BootServer &) = delete;
    // queue a request; out[i] receives the bootstrapped in[i] (4 primes in)
    uint64_t submit(const std::vector<const Ciphertext *> &in, const std::vector<Ciphertext *> &out);
    // queue a packed request: in[i] holds its data in slots [0, slots) and zeros
    // elsewhere; it is packed with other requests of the same count and scale into
    // shared ciphertexts, and out[i] receives the bootstrapped shared ciphertext i with
    // this request's slots at [offset, offset + slots) (offset from wait)
    uint64_t submit_packed(const std::vector<const Ciphertext *> &in, size_t slots,
                           const std::vector<Ciphertext *> &out);
    // block until the request has finished; rethrows its error
    void wait(uint64_t id, double *latency_ms, size_t *slot_offset = nullptr);
    ServerStats stats();

   private:
    struct Request {
        uint64_t id = 0;
        std::vector<std::unique_ptr<Ciphertext>> in;
        std::vector<Ciphertext *> out;
        size_t next = 0, done = 0, bytes = 0;
        size_t slots = 0, offset = 0;  // packed requests: slots > 0
        bool finished = false;
        double latency_ms = 0.0;
        std::exception_ptr error;
        std::chrono::steady_clock::time_point t_submit;
    };
    typedef std::vector<std::pair<std::shared_ptr<Request>, size_t>> Group;
    typedef std::vector<std::shared_ptr<Request>> Pack;
    void loop(size_t w);
    void admit_locked();
    void finish_locked(const std::chrono::steady_clock::time_point &t1);
    void run_group(Group &group, hipStream_t s);
    void run_pack(Pack &pack, hipStream_t s);

    Bootstrapper &b_;
    const GaloisKeys &gk_;
    const KSwitchKey &rk_;
    size_t budget_, max_group_, ct_bytes_ = 0, in_flight_ = 0, free_bytes_ = 0, cap_bytes_ = 0;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<Request>> waiting_, active_;
    std::map<uint64_t, std::shared_ptr<Request>> all_;
    std::vector<double> lat_;
    uint64_t next_id_ = 0, submitted_ = 0, completed_ = 0, failed_ = 0, cts_done_ = 0, groups_ = 0;
    uint64_t packed_done_ = 0, shared_cts_ = 0;
    double busy_s_ = 0.0;
    bool stop_ = false;
    std::vector<hipStream_t> streams_;
    std::vector<std::thread> th_;
};

// Multiply by the constant `value` encoded at scale `sigma` and accumulate:
// acc (+)= src * [round(value * sigma)]_q per limb over Ql limbs of src.
void const_lincomb(Context &c, Ciphertext &acc, const std::vector<const Ciphertext *> &src,
                   const std::vector<double> &values, const std::vector<double> &sigmas, size_t Ql, hipStream_t s);

}  // namespace boot
}  // namespace shm
