// C ABI (include/shmoai.h): thin, exception-safe wrappers over the engine.
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/shmoai.h"
#include "boot.h"
#include "engine.h"

using namespace shm;

struct shm_context {
    std::unique_ptr<Context> c;
};
struct shm_ciphertext {
    Ciphertext ct;
};
struct shm_plaintext {
    Plaintext pt;
};
struct shm_secret_key {
    SecretKey sk;
};
struct shm_kswitch_key {
    KSwitchKey k;
};
struct shm_galois_keys {
    GaloisKeys g;
};

namespace {

thread_local std::string g_err;

template <class F>
int guard(F &&f) {
    try {
        f();
        return SHM_OK;
    } catch (const InvalidArgument &e) {
        g_err = e.what();
        return SHM_ERR_INVALID_ARGUMENT;
    } catch (const LogicError &e) {
        g_err = e.what();
        return SHM_ERR_LOGIC;
    } catch (const DeviceError &e) {
        g_err = e.what();
        return SHM_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        g_err = "out of device or host memory";
        return SHM_ERR_OUT_OF_MEMORY;
    } catch (const std::invalid_argument &e) {
        g_err = e.what();
        return SHM_ERR_INVALID_ARGUMENT;
    } catch (const std::logic_error &e) {
        g_err = e.what();
        return SHM_ERR_LOGIC;
    } catch (const std::exception &e) {
        g_err = e.what();
        return SHM_ERR_RUNTIME;
    } catch (...) {
        g_err = "unknown error";
        return SHM_ERR_RUNTIME;
    }
}

inline hipStream_t S(shm_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

void need(const void *p, const char *what) {
    if (!p) throw InvalidArgument(std::string(what) + " is null");
}

void check_chain(const Context &c, uint32_t chain) {
    if (chain < 1 || chain > c.nQ) throw InvalidArgument("chain_index out of range");
}

void same_level(const Ciphertext &a, const Ciphertext &b) {
    if (a.chain != b.chain) throw InvalidArgument("encrypted1 and encrypted2 parameter mismatch");
}

}  // namespace

extern "C" {

const char *shm_last_error(void) { return g_err.c_str(); }
const char *shm_version(void) { return "shmoai-mi355x 0.1 (gfx950)"; }

int shm_context_create(uint32_t logN, const int *bits, uint32_t n, uint32_t nP, shm_context **out) {
    return guard([&] {
        need(bits, "coeff_bit_sizes");
        need(out, "out");
        std::vector<int> b(bits, bits + n);
        auto h = new shm_context();
        try {
            h->c = make_context((int)logN, b, nP);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int shm_context_destroy(shm_context *ctx) {
    return guard([&] {
        if (ctx) {
            (void)hipDeviceSynchronize();
            delete ctx;
        }
    });
}

int shm_context_get_info(const shm_context *ctx, shm_context_info *info) {
    return guard([&] {
        need(ctx, "ctx");
        need(info, "info");
        info->poly_degree = (uint32_t)ctx->c->N;
        info->size_q = (uint32_t)ctx->c->nQ;
        info->size_p = (uint32_t)ctx->c->nP;
        info->dnum = (uint32_t)ctx->c->dnum;
    });
}

int shm_context_coeff_modulus(const shm_context *ctx, uint64_t *out, size_t n) {
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        if (n < ctx->c->q.size()) throw InvalidArgument("output too small");
        std::memcpy(out, ctx->c->q.data(), ctx->c->q.size() * sizeof(uint64_t));
    });
}

int shm_galois_elt_from_step(const shm_context *ctx, int step, uint32_t *elt) {
    return guard([&] {
        need(ctx, "ctx");
        need(elt, "elt");
        *elt = ctx->c->elt_from_step(step);
    });
}

int shm_ntt_forward(shm_context *ctx, uint64_t *data, uint32_t n_polys, uint32_t n_limbs, uint32_t first,
                    shm_stream_t s) {
    return guard([&] {
        need(ctx, "ctx");
        need(data, "data");
        if (first + n_limbs > ctx->c->nQP()) throw InvalidArgument("prime range out of bounds");
        ntt_forward(*ctx->c, data, n_polys, n_limbs, first, S(s));
    });
}

int shm_synchronize(shm_stream_t s) {
    return guard([&] { SHM_HIP(hipStreamSynchronize(S(s))); });
}

// Streams made here are never destroyed: a buffer that was last used on a stream
// returns to the pool on that stream when it is freed, possibly after the caller
// has released the stream, so released streams are kept (drained) and handed out
// again by the next shm_stream_create.
namespace {
std::mutex g_stream_mu;
std::vector<hipStream_t> g_stream_free;
}  // namespace

int shm_stream_create(shm_stream_t *out) {
    return guard([&] {
        need(out, "out");
        {
            std::lock_guard<std::mutex> lk(g_stream_mu);
            if (!g_stream_free.empty()) {
                *out = (shm_stream_t)g_stream_free.back();
                g_stream_free.pop_back();
                return;
            }
        }
        hipStream_t st = nullptr;
        // default priority; non-blocking, so it never waits on the legacy stream implicitly
        SHM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        *out = (shm_stream_t)st;
    });
}

int shm_stream_destroy(shm_stream_t st) {
    return guard([&] {
        if (!st) return;
        SHM_HIP(hipStreamSynchronize(S(st)));
        std::lock_guard<std::mutex> lk(g_stream_mu);
        g_stream_free.push_back(S(st));
    });
}

struct shm_graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
};

int shm_graph_capture_begin(shm_stream_t st) {
    return guard([&] {
        if (!st) throw InvalidArgument("graph capture needs a created stream (not the default stream)");
        SHM_HIP(hipStreamBeginCapture(S(st), hipStreamCaptureModeThreadLocal));
    });
}

int shm_graph_capture_end(shm_stream_t st, shm_graph **out) {
    return guard([&] {
        need(out, "out");
        if (!st) throw InvalidArgument("graph capture needs a created stream (not the default stream)");
        auto h = std::make_unique<shm_graph>();
        SHM_HIP(hipStreamEndCapture(S(st), &h->g));
        SHM_HIP(hipGraphInstantiate(&h->x, h->g, nullptr, nullptr, 0));
        *out = h.release();
    });
}

int shm_graph_launch(shm_graph *g, shm_stream_t st) {
    return guard([&] {
        need(g, "graph");
        SHM_HIP(hipGraphLaunch(g->x, S(st)));
    });
}

int shm_graph_destroy(shm_graph *g) {
    return guard([&] {
        if (!g) return;
        if (g->x) SHM_HIP(hipGraphExecDestroy(g->x));
        if (g->g) SHM_HIP(hipGraphDestroy(g->g));
        delete g;
    });
}

int shm_ntt_inverse(shm_context *ctx, uint64_t *data, uint32_t n_polys, uint32_t n_limbs, uint32_t first,
                    shm_stream_t s) {
    return guard([&] {
        need(ctx, "ctx");
        need(data, "data");
        if (first + n_limbs > ctx->c->nQP()) throw InvalidArgument("prime range out of bounds");
        ntt_inverse(*ctx->c, data, n_polys, n_limbs, first, S(s));
    });
}

// ---- ciphertexts / plaintexts ------------------------------------------------
int shm_ciphertext_create(shm_context *ctx, uint32_t size, uint32_t chain, shm_ciphertext **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        check_chain(*ctx->c, chain);
        if (size < 2 || size > 3) throw InvalidArgument("ciphertext size must be 2 or 3");
        auto h = new shm_ciphertext();
        h->ct.ctx = ctx->c.get();
        try {
            h->ct.resize(size, chain);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int shm_ciphertext_destroy(shm_ciphertext *ct) {
    return guard([&] {
        // the block returns to the stream-ordered pool; work the caller queued on
        // other streams may still read it, so drain first
        if (ct) SHM_HIP(hipDeviceSynchronize());
        delete ct;
    });
}

int shm_ciphertext_info(const shm_ciphertext *ct, uint32_t *size, uint32_t *chain, uint32_t *cms, double *scale) {
    return guard([&] {
        need(ct, "ct");
        if (size) *size = (uint32_t)ct->ct.size;
        if (chain) *chain = (uint32_t)ct->ct.chain;
        if (cms) *cms = (uint32_t)ct->ct.Ql();
        if (scale) *scale = ct->ct.scale;
    });
}

int shm_ciphertext_set_scale(shm_ciphertext *ct, double scale) {
    return guard([&] {
        need(ct, "ct");
        ct->ct.scale = scale;
    });
}

uint64_t *shm_ciphertext_data(shm_ciphertext *ct) { return ct ? ct->ct.data() : nullptr; }

int shm_ciphertext_import(shm_ciphertext *ct, const uint64_t *host, size_t n, shm_stream_t s) {
    return guard([&] {
        need(ct, "ct");
        need(host, "host");
        const size_t words = ct->ct.size * ct->ct.Ql() * ct->ct.ctx->N;
        if (n != words) throw InvalidArgument("import size mismatch");
        SHM_HIP(hipMemcpyAsync(ct->ct.data(), host, words * 8, hipMemcpyHostToDevice, S(s)));
        SHM_HIP(hipStreamSynchronize(S(s)));
    });
}

int shm_ciphertext_export(const shm_ciphertext *ct, uint64_t *host, size_t n, shm_stream_t s) {
    return guard([&] {
        need(ct, "ct");
        need(host, "host");
        const size_t words = ct->ct.size * ct->ct.Ql() * ct->ct.ctx->N;
        if (n != words) throw InvalidArgument("export size mismatch");
        SHM_HIP(hipMemcpyAsync(host, ct->ct.data(), words * 8, hipMemcpyDeviceToHost, S(s)));
        SHM_HIP(hipStreamSynchronize(S(s)));
    });
}

int shm_ciphertext_copy(const shm_ciphertext *src, shm_ciphertext *dst, shm_stream_t s) {
    return guard([&] {
        need(src, "src");
        need(dst, "dst");
        dst->ct.ctx = src->ct.ctx;
        dst->ct.resize(src->ct.size, src->ct.chain, S(s), false);  // the copy writes every word
        dst->ct.scale = src->ct.scale;
        dst->ct.ntt = src->ct.ntt;
        launch_copy_words(dst->ct.data(), src->ct.data(), src->ct.size * src->ct.Ql() * src->ct.ctx->N, S(s));
    });
}

int shm_plaintext_create(shm_context *ctx, uint32_t chain, shm_plaintext **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        check_chain(*ctx->c, chain);
        auto h = new shm_plaintext();
        h->pt.ctx = ctx->c.get();
        h->pt.chain = chain;
        try {
            h->pt.buf.alloc(ctx->c->nQ * ctx->c->N * sizeof(uint64_t));
            SHM_HIP(hipMemset(h->pt.buf.as<void>(), 0, ctx->c->nQ * ctx->c->N * sizeof(uint64_t)));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int shm_plaintext_destroy(shm_plaintext *pt) {
    return guard([&] { delete pt; });
}

int shm_plaintext_info(const shm_plaintext *pt, uint32_t *chain, uint32_t *cms, double *scale) {
    return guard([&] {
        need(pt, "pt");
        if (chain) *chain = (uint32_t)pt->pt.chain;
        if (cms) *cms = (uint32_t)pt->pt.Ql();
        if (scale) *scale = pt->pt.scale;
    });
}

int shm_plaintext_set_scale(shm_plaintext *pt, double scale) {
    return guard([&] {
        need(pt, "pt");
        pt->pt.scale = scale;
    });
}

uint64_t *shm_plaintext_data(shm_plaintext *pt) { return pt ? pt->pt.data() : nullptr; }

int shm_plaintext_import(shm_plaintext *pt, const uint64_t *host, size_t n, shm_stream_t s) {
    return guard([&] {
        need(pt, "pt");
        need(host, "host");
        const size_t words = pt->pt.Ql() * pt->pt.ctx->N;
        if (n != words) throw InvalidArgument("import size mismatch");
        SHM_HIP(hipMemcpyAsync(pt->pt.data(), host, words * 8, hipMemcpyHostToDevice, S(s)));
        SHM_HIP(hipStreamSynchronize(S(s)));
    });
}

int shm_plaintext_export(const shm_plaintext *pt, uint64_t *host, size_t n, shm_stream_t s) {
    return guard([&] {
        need(pt, "pt");
        need(host, "host");
        const size_t words = pt->pt.Ql() * pt->pt.ctx->N;
        if (n != words) throw InvalidArgument("export size mismatch");
        SHM_HIP(hipMemcpyAsync(host, pt->pt.data(), words * 8, hipMemcpyDeviceToHost, S(s)));
        SHM_HIP(hipStreamSynchronize(S(s)));
    });
}

// ---- keys ----------------------------------------------------------------------
int shm_secret_key_create(shm_context *ctx, const int8_t *coeffs, shm_secret_key **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(coeffs, "coeffs");
        need(out, "out");
        for (size_t j = 0; j < ctx->c->N; j++)
            if (coeffs[j] < -1 || coeffs[j] > 1) throw InvalidArgument("secret key must be ternary");
        auto h = new shm_secret_key();
        h->sk.ctx = ctx->c.get();
        try {
            load_secret(*ctx->c, h->sk, coeffs);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int shm_secret_key_generate(shm_context *ctx, uint64_t seed, uint32_t hw, shm_secret_key **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        const std::vector<int8_t> coeff = ternary_secret(ctx->c->N, seed, hw);
        auto h = new shm_secret_key();
        h->sk.ctx = ctx->c.get();
        try {
            load_secret(*ctx->c, h->sk, coeff.data());
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int shm_secret_key_destroy(shm_secret_key *sk) {
    return guard([&] { delete sk; });
}

int shm_secret_key_export(const shm_secret_key *sk, uint64_t *host, size_t n) {
    return guard([&] {
        need(sk, "sk");
        need(host, "host");
        const size_t words = sk->sk.ctx->nQP() * sk->sk.ctx->N;
        if (n != words) throw InvalidArgument("export size mismatch");
        SHM_HIP(hipMemcpy(host, sk->sk.s.as<void>(), words * 8, hipMemcpyDeviceToHost));
    });
}

int shm_relin_key_generate(shm_context *ctx, const shm_secret_key *sk, uint64_t seed, shm_kswitch_key **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(out, "out");
        Context &c = *ctx->c;
        const size_t words = c.nQP() * c.N;
        Scratch s2(words * sizeof(uint64_t), nullptr);
        launch_mul_elementwise(c, sk->sk.s.as<uint64_t>(), sk->sk.s.as<uint64_t>(), s2.as<uint64_t>(), c.nQP(), true,
                               c.nQ, nullptr);
        auto h = new shm_kswitch_key();
        try {
            gen_kswitch(c, sk->sk, s2.as<uint64_t>(), seed, h->k);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int shm_kswitch_key_generate(shm_context *ctx, const shm_secret_key *sk, const uint64_t *new_key, uint64_t seed,
                             shm_kswitch_key **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(new_key, "new_key");
        need(out, "out");
        Context &c = *ctx->c;
        const size_t words = c.nQP() * c.N;
        Scratch nk(words * sizeof(uint64_t), nullptr);
        SHM_HIP(hipMemcpy(nk.as<void>(), new_key, words * 8, hipMemcpyHostToDevice));
        auto h = new shm_kswitch_key();
        try {
            gen_kswitch(c, sk->sk, nk.as<uint64_t>(), seed, h->k);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int shm_kswitch_key_destroy(shm_kswitch_key *key) {
    return guard([&] { delete key; });
}

int shm_kswitch_key_export(const shm_kswitch_key *key, uint64_t *host, size_t n) {
    return guard([&] {
        need(key, "key");
        need(host, "host");
        const Context &c = *key->k.ctx;
        const size_t words = c.dnum * 2 * c.nQP() * c.N;
        if (n != words) throw InvalidArgument("export size mismatch");
        SHM_HIP(hipMemcpy(host, key->k.k.as<void>(), words * 8, hipMemcpyDeviceToHost));
    });
}

int shm_galois_keys_generate(shm_context *ctx, const shm_secret_key *sk, const int *steps, size_t n_steps,
                             uint64_t seed, shm_galois_keys **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(out, "out");
        if (n_steps) need(steps, "steps");
        Context &c = *ctx->c;
        const size_t QP = c.nQP(), N = c.N;
        auto h = new shm_galois_keys();
        h->g.ctx = &c;
        try {
            Scratch rs(QP * N * sizeof(uint64_t), nullptr);
            for (size_t i = 0; i < n_steps; i++) {
                const uint32_t elt = c.elt_from_step(steps[i]);
                if (h->g.keys.count(elt)) continue;
                const uint32_t *table = c.galois_table(elt, nullptr);
                launch_permute(c, sk->sk.s.as<uint64_t>(), rs.as<uint64_t>(), table, QP, nullptr);
                auto k = std::make_unique<KSwitchKey>();
                gen_kswitch(c, sk->sk, rs.as<uint64_t>(), host_splitmix64(seed, elt), *k);
                h->g.keys[elt] = std::move(k);
            }
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int shm_galois_keys_destroy(shm_galois_keys *gk) {
    return guard([&] { delete gk; });
}

int shm_galois_keys_export(const shm_galois_keys *gk, uint32_t elt, uint64_t *host, size_t n) {
    return guard([&] {
        need(gk, "gk");
        need(host, "host");
        auto it = gk->g.keys.find(elt);
        if (it == gk->g.keys.end()) throw InvalidArgument("Galois elt not present");
        const Context &c = *gk->g.ctx;
        const size_t words = c.dnum * 2 * c.nQP() * c.N;
        if (n != words) throw InvalidArgument("export size mismatch");
        SHM_HIP(hipMemcpy(host, it->second->k.as<void>(), words * 8, hipMemcpyDeviceToHost));
    });
}

// ---- encryption -------------------------------------------------------------------
int shm_encrypt_symmetric(shm_context *ctx, const shm_secret_key *sk, const shm_plaintext *pt, uint64_t seed,
                          shm_ciphertext *out, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(pt, "pt");
        need(out, "out");
        Context &c = *ctx->c;
        const hipStream_t s = S(st);
        const size_t Ql = pt->pt.Ql(), N = c.N;
        out->ct.ctx = &c;
        out->ct.resize(2, pt->pt.chain, s);
        out->ct.scale = pt->pt.scale;
        out->ct.ntt = true;
        uint64_t *c0 = out->ct.poly(0), *c1 = out->ct.poly(1);
        launch_sample_uniform(c, host_splitmix64(seed, 0), c1, Ql, false, s);
        Scratch e(Ql * N * sizeof(uint64_t), s);
        launch_sample_gaussian_rns(c, host_splitmix64(seed, 1), e.as<uint64_t>(), Ql, false, s);
        ntt_forward(c, e.as<uint64_t>(), 1, Ql, 0, s);
        launch_encrypt_combine(c, sk->sk.s.as<uint64_t>(), c1, e.as<uint64_t>(), pt->pt.data(), c0, Ql, s);
    });
}

int shm_decrypt(shm_context *ctx, const shm_secret_key *sk, const shm_ciphertext *ct, shm_plaintext *out,
                shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(ct, "ct");
        need(out, "out");
        Context &c = *ctx->c;
        out->pt.chain = ct->ct.chain;
        out->pt.scale = ct->ct.scale;
        launch_decrypt(c, ct->ct.data(), ct->ct.size, sk->sk.s.as<uint64_t>(), out->pt.data(), ct->ct.Ql(), S(st));
    });
}

// ---- encoder ----------------------------------------------------------------------
int shm_encode(shm_context *ctx, const double *re, const double *im, size_t n_values, uint32_t chain, double scale,
               shm_plaintext *out, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        if (n_values) need(re, "re");
        Context &c = *ctx->c;
        check_chain(c, chain);
        if (n_values == 1) {
            // moai::Encoder forwards single values to the scalar path
            encode_scalar(c, re[0], c.ql_of_chain(chain), scale, out->pt.data(), S(st));
        } else {
            encode(c, re, im, n_values, c.ql_of_chain(chain), scale, out->pt.data(), S(st));
        }
        out->pt.chain = chain;
        out->pt.scale = scale;
    });
}

int shm_encode_scalar(shm_context *ctx, double value, uint32_t chain, double scale, shm_plaintext *out,
                      shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        Context &c = *ctx->c;
        check_chain(c, chain);
        encode_scalar(c, value, c.ql_of_chain(chain), scale, out->pt.data(), S(st));
        out->pt.chain = chain;
        out->pt.scale = scale;
    });
}

int shm_decode(shm_context *ctx, const shm_plaintext *pt, double *re, double *im, size_t n, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(pt, "pt");
        need(re, "re");
        decode(*ctx->c, pt->pt.data(), pt->pt.Ql(), pt->pt.scale, re, im, n, S(st));
    });
}

// ---- evaluator ----------------------------------------------------------------------
int shm_add_inplace(shm_context *ctx, shm_ciphertext *a, const shm_ciphertext *b, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(a, "a");
        need(b, "b");
        same_level(a->ct, b->ct);
        if (a->ct.size != b->ct.size) throw InvalidArgument("ciphertext sizes differ");
        launch_add(*ctx->c, a->ct.data(), b->ct.data(), a->ct.Ql(), a->ct.size, S(st));
    });
}

int shm_sub_inplace(shm_context *ctx, shm_ciphertext *a, const shm_ciphertext *b, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(a, "a");
        need(b, "b");
        same_level(a->ct, b->ct);
        if (a->ct.size != b->ct.size) throw InvalidArgument("ciphertext sizes differ");
        if (std::fabs(a->ct.scale - b->ct.scale) > 1e-6 * std::fabs(a->ct.scale))
            throw InvalidArgument("scale mismatch");
        launch_sub(*ctx->c, a->ct.data(), b->ct.data(), a->ct.Ql(), a->ct.size, S(st));
    });
}

int shm_negate_inplace(shm_context *ctx, shm_ciphertext *a, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(a, "a");
        launch_negate(*ctx->c, a->ct.data(), a->ct.Ql(), a->ct.size, S(st));
    });
}

int shm_add_plain_inplace(shm_context *ctx, shm_ciphertext *a, const shm_plaintext *pt, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(a, "a");
        need(pt, "pt");
        if (a->ct.chain != pt->pt.chain) throw InvalidArgument("encrypted and plain parameter mismatch");
        if (std::fabs(a->ct.scale - pt->pt.scale) > 1e-6 * std::fabs(a->ct.scale))
            throw InvalidArgument("scale mismatch");
        launch_add(*ctx->c, a->ct.data(), pt->pt.data(), a->ct.Ql(), 1, S(st));
    });
}

int shm_sub_plain_inplace(shm_context *ctx, shm_ciphertext *a, const shm_plaintext *pt, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(a, "a");
        need(pt, "pt");
        if (a->ct.chain != pt->pt.chain) throw InvalidArgument("encrypted and plain parameter mismatch");
        if (std::fabs(a->ct.scale - pt->pt.scale) > 1e-6 * std::fabs(a->ct.scale))
            throw InvalidArgument("scale mismatch");
        launch_sub(*ctx->c, a->ct.data(), pt->pt.data(), a->ct.Ql(), 1, S(st));
    });
}

int shm_multiply_plain_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_plaintext *pt, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        need(pt, "pt");
        if (ct->ct.chain != pt->pt.chain) throw InvalidArgument("encrypted and plain parameter mismatch");
        launch_mul_plain(*ctx->c, ct->ct.data(), pt->pt.data(), ct->ct.Ql(), ct->ct.size, S(st));
        ct->ct.scale = ct->ct.scale * pt->pt.scale;
    });
}

int shm_multiply_inplace(shm_context *ctx, shm_ciphertext *a, const shm_ciphertext *b, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(a, "a");
        need(b, "b");
        same_level(a->ct, b->ct);
        if (a->ct.size != 2 || b->ct.size != 2) throw InvalidArgument("multiply supports size-2 ciphertexts");
        Context &c = *ctx->c;
        const size_t Ql = a->ct.Ql(), S3 = 3 * Ql * c.N;
        const hipStream_t s = S(st);
        Scratch out(S3 * sizeof(uint64_t), s);
        if (a == b)
            launch_square(c, a->ct.data(), out.as<uint64_t>(), Ql, s);
        else
            launch_tensor(c, a->ct.data(), b->ct.data(), out.as<uint64_t>(), Ql, s);
        a->ct.resize(3, a->ct.chain, s);
        launch_copy_words(a->ct.data(), out.as<uint64_t>(), S3, s);
        a->ct.scale = a->ct.scale * b->ct.scale;
    });
}

int shm_square_inplace(shm_context *ctx, shm_ciphertext *a, shm_stream_t st) {
    return shm_multiply_inplace(ctx, a, a, st);
}

int shm_relinearize_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_kswitch_key *rk, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        need(rk, "relin_key");
        if (ct->ct.size != 3) throw InvalidArgument("destination_size must be 3");
        Context &c = *ctx->c;
        keyswitch(c, ct->ct.poly(2), rk->k.k.as<uint64_t>(), ct->ct.Ql(), ct->ct.poly(0), ct->ct.poly(1), true, S(st));
        ct->ct.size = 2;
    });
}

int shm_apply_kswitch_key_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_kswitch_key *key,
                                  shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        need(key, "key");
        if (ct->ct.size != 2) throw InvalidArgument("encrypted size must be 2");
        Context &c = *ctx->c;
        const size_t Ql = ct->ct.Ql();
        const hipStream_t s = S(st);
        Scratch t(Ql * c.N * sizeof(uint64_t), s);
        launch_copy_words(t.as<uint64_t>(), ct->ct.poly(1), Ql * c.N, s);
        keyswitch(c, t.as<uint64_t>(), key->k.k.as<uint64_t>(), Ql, ct->ct.poly(0), ct->ct.poly(1), false, s);
    });
}

int shm_rescale_to_next_inplace(shm_context *ctx, shm_ciphertext *ct, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        rescale_inplace(*ctx->c, ct->ct, S(st));
    });
}

int shm_rescale_to_next_inplace_noalloc(shm_context *ctx, shm_ciphertext *ct, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        rescale_inplace_noalloc(*ctx->c, ct->ct, S(st));
    });
}

int shm_mod_switch_to_next_inplace(shm_context *ctx, shm_ciphertext *ct, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        if (ct->ct.Ql() < 2) throw LogicError("end of modulus switching chain reached");
        mod_switch_drop(*ctx->c, ct->ct, ct->ct.chain + 1, S(st));
    });
}

int shm_mod_switch_to_inplace(shm_context *ctx, shm_ciphertext *ct, uint32_t chain, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        check_chain(*ctx->c, chain);
        mod_switch_drop(*ctx->c, ct->ct, chain, S(st));
    });
}

int shm_plain_mod_switch_to_inplace(shm_context *ctx, shm_plaintext *pt, uint32_t chain, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(pt, "pt");
        check_chain(*ctx->c, chain);
        if (chain < pt->pt.chain) throw InvalidArgument("cannot switch to a higher level");
        (void)st;
        pt->pt.chain = chain;  // limb-major: dropping trailing limbs needs no data movement
    });
}

int shm_apply_galois_inplace(shm_context *ctx, shm_ciphertext *ct, uint32_t elt, const shm_galois_keys *gk,
                             shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        need(gk, "galois_keys");
        apply_galois(*ctx->c, ct->ct, elt, gk->g, S(st));
    });
}

int shm_rotate_vector_inplace(shm_context *ctx, shm_ciphertext *ct, int step, const shm_galois_keys *gk,
                              shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        need(gk, "galois_keys");
        rotate_internal(*ctx->c, ct->ct, step, gk->g, S(st));
    });
}

int shm_hoisting_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_galois_keys *gk, const int *steps,
                         size_t n_steps, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        need(gk, "galois_keys");
        need(steps, "steps");
        hoisting_inplace(*ctx->c, ct->ct, std::vector<int>(steps, steps + n_steps), gk->g, S(st));
    });
}

int shm_complex_conjugate_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_galois_keys *gk, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        need(gk, "galois_keys");
        apply_galois(*ctx->c, ct->ct, ctx->c->elt_from_step(0), gk->g, S(st));
    });
}

// ---- MOAI matmul ------------------------------------------------------------------------
int shm_ct_pt_matrix_mul(shm_context *ctx, const shm_ciphertext *const *X, uint32_t row_W, const double *W,
                         uint32_t col_W, shm_ciphertext *const *out, int variant, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(X, "enc_X");
        need(W, "W");
        need(out, "out");
        if (variant < 0 || variant > 2) throw InvalidArgument("unknown matmul variant");
        std::vector<const Ciphertext *> xs(row_W);
        for (uint32_t j = 0; j < row_W; j++) {
            need(X[j], "enc_X[j]");
            xs[j] = &X[j]->ct;
        }
        std::vector<Ciphertext *> os(col_W);
        for (uint32_t i = 0; i < col_W; i++) {
            need(out[i], "out[i]");
            out[i]->ct.ctx = ctx->c.get();
            os[i] = &out[i]->ct;
        }
        ctpt_matmul(*ctx->c, xs, W, row_W, col_W, os, variant, S(st));
    });
}

int shm_ct_pt_matrix_mul_masked(shm_context *ctx, const shm_ciphertext *const *X, uint32_t row_W, const double *W,
                                uint32_t col_W, const int32_t *bias_vec, size_t n_bias, shm_ciphertext *const *out,
                                shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(X, "enc_X");
        need(W, "W");
        need(bias_vec, "bias_vec");
        need(out, "out");
        const size_t n = ctx->c->N / 2;
        if (n_bias != n) throw InvalidArgument("bias_vec must hold slot_count entries");
        std::vector<uint8_t> mask(n);
        for (size_t t = 0; t < n; t++) mask[t] = bias_vec[t] == 1 ? 1 : 0;
        std::vector<const Ciphertext *> xs(row_W);
        for (uint32_t j = 0; j < row_W; j++) {
            need(X[j], "enc_X[j]");
            xs[j] = &X[j]->ct;
        }
        std::vector<Ciphertext *> os(col_W);
        for (uint32_t i = 0; i < col_W; i++) {
            need(out[i], "out[i]");
            out[i]->ct.ctx = ctx->c.get();
            os[i] = &out[i]->ct;
        }
        ctpt_matmul_masked(*ctx->c, xs, W, row_W, col_W, mask.data(), os, S(st));
    });
}

double shm_last_matmul_kernel_ms(void) { return (double)last_matmul_kernel_ms(); }
double shm_last_matmul_rescale_tile_ms(void) { return (double)last_matmul_rescale_tile_ms(); }
int shm_last_matmul_variant(void) { return last_matmul_variant(); }

}  // extern "C"

// ---- bootstrapping ----------------------------------------------------------------------
struct shm_bootstrapper {
    std::unique_ptr<boot::Bootstrapper> b;
};

namespace {
const boot::LtLayer &pick_layer(const shm_bootstrapper *b, int transform, int layer) {
    if (transform < 0 || transform > 1 || layer < 0 || layer > 2)
        throw InvalidArgument("transform must be 0 or 1 and layer 0..2");
    return transform == 0 ? b->b->s2c[layer] : b->b->c2s[layer];
}
}  // namespace

namespace {
boot::BootParams boot_params(const shm_boot_params *p) {
    boot::BootParams bp;
    bp.s2c_input_primes = p->s2c_input_primes;
    bp.layer3_factor = p->layer3_factor;
    bp.K = p->K;
    bp.deg = p->deg;
    bp.double_angles = p->double_angles;
    bp.input_bound = p->input_bound;
    return bp;
}

boot::Bootstrapper &ready(shm_bootstrapper *b, hipStream_t s) {
    need(b, "bootstrapper");
    if (!b->b->ctx) throw InvalidArgument("bootstrapping plan without a context cannot run");
    if (!b->b->prepared) b->b->prepare(s);
    return *b->b;
}
}  // namespace

int shm_bootstrapper_create(shm_context *ctx, const shm_boot_params *p, shm_bootstrapper **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(p, "params");
        need(out, "out");
        auto h = std::make_unique<shm_bootstrapper>();
        h->b = boot::make_bootstrapper(*ctx->c, boot_params(p));
        *out = h.release();
    });
}

int shm_bootstrapper_create_plan(uint32_t log_poly_degree, uint32_t size_q, const shm_boot_params *p,
                                 shm_bootstrapper **out) {
    return guard([&] {
        need(p, "params");
        need(out, "out");
        if (log_poly_degree < 4 || log_poly_degree > 17) throw InvalidArgument("log_poly_degree out of range");
        auto h = std::make_unique<shm_bootstrapper>();
        h->b = boot::make_plan(size_t(1) << log_poly_degree, size_q, boot_params(p));
        *out = h.release();
    });
}

int shm_bootstrapper_destroy(shm_bootstrapper *b) {
    return guard([&] {
        if (b && b->b->prepared) SHM_HIP(hipDeviceSynchronize());
        delete b;
    });
}

int shm_bootstrapper_rotation_steps(const shm_bootstrapper *b, int *steps, size_t cap, size_t *count) {
    return guard([&] {
        need(b, "bootstrapper");
        const auto v = b->b->rotation_steps();
        if (count) *count = v.size();
        if (steps) {
            if (cap < v.size()) throw InvalidArgument("steps buffer too small");
            std::copy(v.begin(), v.end(), steps);
        }
    });
}

int shm_bootstrapper_layer_info(const shm_bootstrapper *b, int transform, int layer, int64_t info[5]) {
    return guard([&] {
        need(b, "bootstrapper");
        need(info, "info");
        const auto &L = pick_layer(b, transform, layer);
        info[0] = (int64_t)L.base;
        info[1] = L.k0;
        info[2] = L.g;
        info[3] = L.n_giants;
        info[4] = (int64_t)L.chain;
    });
}

int shm_bootstrapper_layer_diagonal(const shm_bootstrapper *b, int transform, int layer, int giant, int baby,
                                    double *re, double *im, size_t n) {
    return guard([&] {
        need(b, "bootstrapper");
        need(re, "re");
        need(im, "im");
        const auto &L = pick_layer(b, transform, layer);
        if (giant < 0 || giant >= L.n_giants || baby < 0 || baby >= L.g) throw InvalidArgument("diagonal index");
        if (n != b->b->n) throw InvalidArgument("n must be N/2");
        const auto &v = L.pre[giant][baby];
        for (size_t t = 0; t < n; t++) {
            re[t] = v.empty() ? 0.0 : v[t].real();
            im[t] = v.empty() ? 0.0 : v[t].imag();
        }
    });
}

int shm_bootstrapper_evalmod_coeffs(const shm_bootstrapper *b, double *out, size_t cap, size_t *count) {
    return guard([&] {
        need(b, "bootstrapper");
        const auto &c = b->b->cheb;
        if (count) *count = c.size();
        if (out) {
            if (cap < c.size()) throw InvalidArgument("buffer too small");
            std::copy(c.begin(), c.end(), out);
        }
    });
}

int shm_bootstrapper_prepare(shm_bootstrapper *b, shm_stream_t st) {
    return guard([&] {
        ready(b, S(st));
        SHM_HIP(hipStreamSynchronize(S(st)));
    });
}

int shm_modraise_inplace(shm_context *ctx, shm_ciphertext *ct, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        ct_modraise(*ctx->c, ct->ct, S(st));
    });
}

int shm_multiply_by_i_inplace(shm_context *ctx, shm_ciphertext *ct, int negate, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        if (!ct->ct.ntt) throw InvalidArgument("CKKS encrypted must be in NTT form");
        ct_mul_monomial_half(*ctx->c, ct->ct, negate != 0, S(st));
    });
}

int shm_bootstrap_slot_to_coeff(shm_bootstrapper *b, shm_ciphertext *ct, const shm_galois_keys *gk,
                                shm_stream_t st) {
    return guard([&] {
        need(ct, "ct");
        need(gk, "galois_keys");
        boot::slot_to_coeff(ready(b, S(st)), ct->ct, gk->g, S(st));
    });
}

int shm_bootstrap_coeff_to_slot(shm_bootstrapper *b, shm_ciphertext *ct, shm_ciphertext *re, shm_ciphertext *im,
                                const shm_galois_keys *gk, shm_stream_t st) {
    return guard([&] {
        need(ct, "ct");
        need(re, "re");
        need(im, "im");
        need(gk, "galois_keys");
        boot::coeff_to_slot(ready(b, S(st)), ct->ct, re->ct, im->ct, gk->g, S(st));
    });
}

int shm_bootstrap_eval_mod(shm_bootstrapper *b, shm_ciphertext *ct, const shm_kswitch_key *relin, shm_stream_t st) {
    return guard([&] {
        need(ct, "ct");
        need(relin, "relin_keys");
        const double sc = ct->ct.scale;
        boot::eval_mod(ready(b, S(st)), ct->ct, relin->k, 1.0, sc, S(st));
    });
}

int shm_bootstrap_eval_mod_batch(shm_bootstrapper *b, shm_ciphertext *const *cts, size_t count,
                                 const shm_kswitch_key *relin, shm_stream_t st) {
    return guard([&] {
        need(relin, "relin_keys");
        if (count == 0 || count > 32) throw InvalidArgument("eval_mod_batch: 1..32 ciphertexts");
        if (!cts) throw InvalidArgument("cts is null");
        std::vector<Ciphertext *> xs(count);
        for (size_t i = 0; i < count; i++) {
            need(cts[i], "cts[i]");
            xs[i] = &cts[i]->ct;
        }
        const double sc = xs[0]->scale;
        boot::eval_mod_batch(ready(b, S(st)), xs, relin->k, 1.0, sc, S(st));
    });
}

int shm_evaluate_chebyshev(shm_context *ctx, shm_ciphertext *const *cts, size_t count, const double *coeffs,
                           size_t n_coeffs, double lo, double hi, const shm_kswitch_key *relin, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(relin, "relin_keys");
        need(coeffs, "coeffs");
        if (!cts) throw InvalidArgument("cts is null");
        if (count == 0 || count > 32) throw InvalidArgument("evaluate_chebyshev: 1..32 ciphertexts");
        std::vector<Ciphertext *> xs(count);
        for (size_t i = 0; i < count; i++) {
            need(cts[i], "cts[i]");
            xs[i] = &cts[i]->ct;
        }
        boot::chebyshev_eval_batch(*ctx->c, xs, std::vector<double>(coeffs, coeffs + n_coeffs), lo, hi, relin->k,
                                   S(st));
    });
}

int shm_ct_ct_matrix_mul_colpacking(shm_context *ctx, const shm_ciphertext *const *enc_X,
                                    const shm_ciphertext *const *enc_W, int col_X, int row_X, int col_W, int row_W,
                                    int num_batch, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                                    shm_ciphertext *const *out, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(gk, "galois_keys");
        need(relin, "relin_keys");
        if (col_X != col_W || row_X != row_W) throw InvalidArgument("ct_ct_matrix_mul: bad dimensions of X or W");
        if (col_X <= 0 || row_X <= 0) throw InvalidArgument("ct_ct_matrix_mul: dimensions must be positive");
        if (!enc_X || !enc_W || !out) throw InvalidArgument("enc_X/enc_W/out is null");
        std::vector<const Ciphertext *> xs(col_X), ws(col_X);
        for (int j = 0; j < col_X; j++) {
            need(enc_X[j], "enc_X[j]");
            need(enc_W[j], "enc_W[j]");
            xs[j] = &enc_X[j]->ct;
            ws[j] = &enc_W[j]->ct;
        }
        std::vector<Ciphertext *> os(row_X);
        for (int i = 0; i < row_X; i++) {
            need(out[i], "out[i]");
            os[i] = &out[i]->ct;
        }
        ctct_matmul_colpacking(*ctx->c, xs, ws, num_batch, gk->g, relin->k, os, S(st));
    });
}

int shm_ct_ct_matrix_mul_diagpacking(shm_context *ctx, const shm_ciphertext *const *enc_X,
                                     const shm_ciphertext *const *enc_W, int col_X, int row_X, int col_W, int row_W,
                                     int num_batch, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                                     shm_ciphertext *const *out, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(gk, "galois_keys");
        need(relin, "relin_keys");
        (void)row_W;
        if (col_X <= 0 || col_W <= 0) throw InvalidArgument("ct_ct_matrix_mul_diag: dimensions must be positive");
        if (row_X < col_X) throw InvalidArgument("ct_ct_matrix_mul_diag: row_X must be at least col_X");
        if (!enc_X || !enc_W || !out) throw InvalidArgument("enc_X/enc_W/out is null");
        std::vector<const Ciphertext *> xs(col_X), ws(col_W);
        for (int j = 0; j < col_X; j++) {
            need(enc_X[j], "enc_X[j]");
            xs[j] = &enc_X[j]->ct;
        }
        std::vector<Ciphertext *> os(col_W);
        for (int i = 0; i < col_W; i++) {
            need(enc_W[i], "enc_W[i]");
            need(out[i], "out[i]");
            ws[i] = &enc_W[i]->ct;
            os[i] = &out[i]->ct;
        }
        ctct_matmul_diagpacking(*ctx->c, xs, ws, num_batch, gk->g, relin->k, os, S(st));
    });
}

int shm_bootstrap_s2c_first(shm_bootstrapper *b, const shm_ciphertext *const *in, size_t count,
                            shm_ciphertext *const *out, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                            double out_scale, shm_stream_t st) {
    return guard([&] {
        boot::Bootstrapper &bs = ready(b, S(st));
        need(gk, "galois_keys");
        need(relin, "relin_keys");
        if (count && (!in || !out)) throw InvalidArgument("in/out is null");
        for (size_t i = 0; i < count; i++) {
            need(in[i], "in[i]");
            need(out[i], "out[i]");
        }
        std::vector<const Ciphertext *> ins(count);
        std::vector<Ciphertext *> outs(count);
        for (size_t i = 0; i < count; i++) {
            ins[i] = &in[i]->ct;
            outs[i] = &out[i]->ct;
        }
        boot::bootstrap_s2c_first_batch(bs, ins, outs, gk->g, relin->k, out_scale, S(st));
        SHM_HIP(hipStreamSynchronize(S(st)));
    });
}

// ---- bootstrapping server --------------------------------------------------------------------
struct shm_server {
    std::unique_ptr<boot::BootServer> s;
};

int shm_server_create(shm_bootstrapper *b, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                      uint64_t mem_budget_bytes, uint32_t max_group, shm_server **out) {
    return guard([&] {
        need(gk, "galois_keys");
        need(relin, "relin_keys");
        need(out, "out");
        boot::Bootstrapper &bs = ready(b, nullptr);
        auto h = std::make_unique<shm_server>();
        h->s = std::make_unique<boot::BootServer>(bs, gk->g, relin->k, (size_t)mem_budget_bytes, max_group);
        *out = h.release();
    });
}

int shm_server_create_ex(shm_bootstrapper *b, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                         uint64_t mem_budget_bytes, uint32_t max_group, uint32_t workers, shm_server **out) {
    return guard([&] {
        need(gk, "galois_keys");
        need(relin, "relin_keys");
        need(out, "out");
        boot::Bootstrapper &bs = ready(b, nullptr);
        auto h = std::make_unique<shm_server>();
        h->s = std::make_unique<boot::BootServer>(bs, gk->g, relin->k, (size_t)mem_budget_bytes, max_group,
                                                  (size_t)workers);
        *out = h.release();
    });
}

int shm_server_submit_packed(shm_server *srv, const shm_ciphertext *const *in, size_t count, uint32_t slots,
                             shm_ciphertext *const *out, uint64_t *request_id) {
    return guard([&] {
        need(srv, "server");
        need(request_id, "request_id");
        if (count == 0 || !in || !out) throw InvalidArgument("server: empty request");
        std::vector<const Ciphertext *> ins(count);
        std::vector<Ciphertext *> outs(count);
        for (size_t i = 0; i < count; i++) {
            need(in[i], "in[i]");
            need(out[i], "out[i]");
            ins[i] = &in[i]->ct;
            outs[i] = &out[i]->ct;
        }
        *request_id = srv->s->submit_packed(ins, slots, outs);
    });
}

int shm_server_wait_packed(shm_server *srv, uint64_t request_id, double *latency_ms, uint32_t *slot_offset) {
    return guard([&] {
        need(srv, "server");
        size_t off = 0;
        srv->s->wait(request_id, latency_ms, &off);
        if (slot_offset) *slot_offset = (uint32_t)off;
    });
}

int shm_server_destroy(shm_server *srv) {
    return guard([&] { delete srv; });
}

int shm_server_submit(shm_server *srv, const shm_ciphertext *const *in, size_t count, shm_ciphertext *const *out,
                      uint64_t *request_id) {
    return guard([&] {
        need(srv, "server");
        need(request_id, "request_id");
        if (count == 0 || !in || !out) throw InvalidArgument("server: empty request");
        std::vector<const Ciphertext *> ins(count);
        std::vector<Ciphertext *> outs(count);
        for (size_t i = 0; i < count; i++) {
            need(in[i], "in[i]");
            need(out[i], "out[i]");
            ins[i] = &in[i]->ct;
            outs[i] = &out[i]->ct;
        }
        *request_id = srv->s->submit(ins, outs);
    });
}

int shm_server_wait(shm_server *srv, uint64_t request_id, double *latency_ms) {
    return guard([&] {
        need(srv, "server");
        srv->s->wait(request_id, latency_ms);
    });
}

int shm_server_get_stats(shm_server *srv, shm_server_stats *st) {
    return guard([&] {
        need(srv, "server");
        need(st, "stats");
        const boot::ServerStats v = srv->s->stats();
        st->submitted = v.submitted;
        st->completed = v.completed;
        st->failed = v.failed;
        st->ciphertexts = v.ciphertexts;
        st->groups = v.groups;
        st->in_flight_bytes = v.in_flight_bytes;
        st->p50_ms = v.p50_ms;
        st->p99_ms = v.p99_ms;
        st->packed_requests = v.packed_requests;
        st->shared_cts = v.shared_cts;
        st->free_bytes = v.free_bytes;
        st->admit_cap_bytes = v.admit_cap_bytes;
        st->workers = v.workers;
        st->busy_s = v.busy_s;
    });
}

// ---- shared-a batches ------------------------------------------------------------------------
struct shm_shared_a_keys {
    std::unique_ptr<boot::SharedAKeys> k;
};
struct shm_shared_a_batch {
    boot::SharedABatch b;
};

int shm_shared_a_keys_create(shm_context *ctx, const shm_secret_key *sk, uint32_t n, uint32_t hw, uint64_t seed,
                             uint32_t chain, shm_shared_a_keys **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(out, "out");
        check_chain(*ctx->c, chain);
        auto h = std::make_unique<shm_shared_a_keys>();
        h->k = boot::make_shared_keys(*ctx->c, sk->sk, n, hw, seed, chain, nullptr);
        *out = h.release();
    });
}

int shm_shared_a_keys_destroy(shm_shared_a_keys *keys) {
    return guard([&] {
        if (keys) SHM_HIP(hipDeviceSynchronize());
        delete keys;
    });
}

int shm_shared_a_keys_add_galois(shm_shared_a_keys *keys, const int *steps, size_t n_steps, uint64_t seed) {
    return guard([&] {
        need(keys, "keys");
        if (n_steps) need(steps, "steps");
        for (size_t i = 0; i < n_steps; i++) boot::add_shared_galois_key(*keys->k, steps[i], seed, nullptr);
    });
}

int shm_shared_a_keys_aux_coeffs(const shm_shared_a_keys *keys, uint32_t i, int8_t *out, size_t n) {
    return guard([&] {
        need(keys, "keys");
        need(out, "out");
        if (i >= keys->k->n) throw InvalidArgument("member index out of range");
        const auto &v = keys->k->aux_coeffs[i];
        if (n != v.size()) throw InvalidArgument("n must be N");
        std::copy(v.begin(), v.end(), out);
    });
}

int shm_shared_a_keys_export_format(const shm_shared_a_keys *keys, uint64_t *a_fmt, uint64_t *B, size_t nw) {
    return guard([&] {
        need(keys, "keys");
        const auto &K = *keys->k;
        const Context &c = *K.ctx;
        const size_t QlP = c.ql_of_chain(K.fmt_chain) + c.nP, words = K.n * K.n * QlP * c.N;
        if (nw != words) throw InvalidArgument("n_words_B must be n*n*(Ql+P)*N");
        if (a_fmt)
            SHM_HIP(hipMemcpy(a_fmt, K.a_fmt.as<void>(), K.n * QlP * c.N * 8, hipMemcpyDeviceToHost));
        if (B) SHM_HIP(hipMemcpy(B, K.B_fmt.as<void>(), words * 8, hipMemcpyDeviceToHost));
    });
}

int shm_shared_a_keys_export_galois(const shm_shared_a_keys *keys, uint32_t elt, uint64_t *A, uint64_t *B,
                                    size_t nw) {
    return guard([&] {
        need(keys, "keys");
        const auto &K = *keys->k;
        const Context &c = *K.ctx;
        auto it = K.galois.find(elt);
        if (it == K.galois.end()) throw InvalidArgument("shared-a Galois key not present");
        const size_t words = c.dnum * K.n * c.nQP() * c.N;
        if (nw != words) throw InvalidArgument("n_words_B must be dnum*n*QP*N");
        if (A) SHM_HIP(hipMemcpy(A, it->second->A.as<void>(), c.dnum * c.nQP() * c.N * 8, hipMemcpyDeviceToHost));
        if (B) SHM_HIP(hipMemcpy(B, it->second->B.as<void>(), words * 8, hipMemcpyDeviceToHost));
    });
}

int shm_shared_a_batch_create(shm_context *ctx, uint32_t n, uint32_t chain, shm_shared_a_batch **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        check_chain(*ctx->c, chain);
        if (n == 0 || n > 16) throw InvalidArgument("shared-a batches hold 1..16 members");
        auto h = std::make_unique<shm_shared_a_batch>();
        h->b.n = n;
        h->b.ct.ctx = ctx->c.get();
        h->b.ct.resize(n + 1, chain);
        *out = h.release();
    });
}

int shm_shared_a_batch_destroy(shm_shared_a_batch *b) {
    return guard([&] {
        if (b) SHM_HIP(hipDeviceSynchronize());
        delete b;
    });
}

int shm_shared_a_batch_info(const shm_shared_a_batch *b, uint32_t *n, uint32_t *chain, uint32_t *cms, double *scale) {
    return guard([&] {
        need(b, "batch");
        if (n) *n = (uint32_t)b->b.n;
        if (chain) *chain = (uint32_t)b->b.ct.chain;
        if (cms) *cms = (uint32_t)b->b.ct.Ql();
        if (scale) *scale = b->b.ct.scale;
    });
}

int shm_shared_a_batch_set_scale(shm_shared_a_batch *b, double scale) {
    return guard([&] {
        need(b, "batch");
        b->b.ct.scale = scale;
    });
}

uint64_t *shm_shared_a_batch_data(shm_shared_a_batch *b) { return b ? b->b.ct.data() : nullptr; }

int shm_shared_a_batch_import(shm_shared_a_batch *b, const uint64_t *host, size_t nw, shm_stream_t st) {
    return guard([&] {
        need(b, "batch");
        need(host, "host");
        const Ciphertext &ct = b->b.ct;
        if (nw != ct.size * ct.Ql() * ct.ctx->N) throw InvalidArgument("n_words must be (n+1)*Ql*N");
        SHM_HIP(hipMemcpyAsync(ct.data(), host, nw * 8, hipMemcpyHostToDevice, S(st)));
        SHM_HIP(hipStreamSynchronize(S(st)));
    });
}

int shm_shared_a_batch_export(const shm_shared_a_batch *b, uint64_t *host, size_t nw, shm_stream_t st) {
    return guard([&] {
        need(b, "batch");
        need(host, "host");
        const Ciphertext &ct = b->b.ct;
        if (nw != ct.size * ct.Ql() * ct.ctx->N) throw InvalidArgument("n_words must be (n+1)*Ql*N");
        SHM_HIP(hipMemcpyAsync(host, ct.data(), nw * 8, hipMemcpyDeviceToHost, S(st)));
        SHM_HIP(hipStreamSynchronize(S(st)));
    });
}

int shm_shared_a_batch_member(const shm_shared_a_batch *b, uint32_t i, shm_ciphertext *out, shm_stream_t st) {
    return guard([&] {
        need(b, "batch");
        need(out, "out");
        if (i >= b->b.n) throw InvalidArgument("SharedACiphertextBatch::assemble: index out of range");
        const Ciphertext &ct = b->b.ct;
        const size_t w = ct.Ql() * ct.ctx->N;
        out->ct.ctx = ct.ctx;
        out->ct.resize(2, ct.chain, S(st));
        out->ct.scale = ct.scale;
        out->ct.ntt = true;
        launch_copy_words(out->ct.poly(0), ct.poly(i), w, S(st));
        launch_copy_words(out->ct.poly(1), ct.poly(b->b.n), w, S(st));
    });
}

int shm_fmt_switch_forward(shm_shared_a_keys *keys, const shm_ciphertext *const *cts, size_t n,
                           shm_shared_a_batch *out, shm_stream_t st) {
    return guard([&] {
        need(keys, "keys");
        need(out, "out");
        if (n && !cts) throw InvalidArgument("cts is null");
        std::vector<const Ciphertext *> v(n);
        for (size_t i = 0; i < n; i++) {
            need(cts[i], "cts[i]");
            v[i] = &cts[i]->ct;
        }
        boot::fmt_switch_forward(*keys->k, v, out->b, S(st));
    });
}

int shm_fmt_switch_backward(shm_shared_a_keys *keys, const shm_shared_a_batch *batch, shm_ciphertext *const *out,
                            size_t n, shm_stream_t st) {
    return guard([&] {
        need(keys, "keys");
        need(batch, "batch");
        if (n && !out) throw InvalidArgument("out is null");
        std::vector<Ciphertext *> v(n);
        for (size_t i = 0; i < n; i++) {
            need(out[i], "out[i]");
            v[i] = &out[i]->ct;
        }
        boot::fmt_switch_backward(*keys->k, batch->b, v, S(st));
    });
}

int shm_shared_a_rotate_inplace(shm_shared_a_keys *keys, shm_shared_a_batch *b, int step, shm_stream_t st) {
    return guard([&] {
        need(keys, "keys");
        need(b, "batch");
        boot::shared_rotate(*keys->k, b->b, step, S(st));
    });
}

int shm_shared_a_add_inplace(shm_shared_a_batch *a, const shm_shared_a_batch *b, shm_stream_t st) {
    return guard([&] {
        need(a, "a");
        need(b, "b");
        if (a->b.n != b->b.n) throw InvalidArgument("batch sizes differ");
        ct_add(*a->b.ct.ctx, a->b.ct, b->b.ct, S(st));
    });
}

int shm_shared_a_multiply_plain_inplace(shm_shared_a_batch *b, const shm_plaintext *pt, shm_stream_t st) {
    return guard([&] {
        need(b, "batch");
        need(pt, "pt");
        Ciphertext &ct = b->b.ct;
        if (pt->pt.chain != ct.chain) throw InvalidArgument("encrypted and plain parameter mismatch");
        launch_mul_plain(*ct.ctx, ct.data(), pt->pt.data(), ct.Ql(), ct.size, S(st));
        ct.scale *= pt->pt.scale;
    });
}

int shm_shared_a_rescale_to_next_inplace(shm_shared_a_batch *b, shm_stream_t st) {
    return guard([&] {
        need(b, "batch");
        rescale_inplace(*b->b.ct.ctx, b->b.ct, S(st));
    });
}

int shm_shared_a_modraise_inplace(shm_shared_a_batch *b, shm_stream_t st) {
    return guard([&] {
        need(b, "batch");
        ct_modraise(*b->b.ct.ctx, b->b.ct, S(st));
    });
}

int shm_bootstrap_shared_a_s2c_first(shm_bootstrapper *b, shm_shared_a_keys *keys, const shm_ciphertext *const *in,
                                     size_t n, shm_ciphertext *const *out, const shm_kswitch_key *relin,
                                     double out_scale, shm_stream_t st) {
    return guard([&] {
        boot::Bootstrapper &bs = ready(b, S(st));
        need(keys, "keys");
        need(relin, "relin_keys");
        if (n && (!in || !out)) throw InvalidArgument("in/out is null");
        std::vector<const Ciphertext *> vi(n);
        std::vector<Ciphertext *> vo(n);
        for (size_t i = 0; i < n; i++) {
            need(in[i], "in[i]");
            need(out[i], "out[i]");
            vi[i] = &in[i]->ct;
            vo[i] = &out[i]->ct;
        }
        boot::bootstrap_shared_a_s2c_first(bs, *keys->k, vi, vo, relin->k, out_scale, S(st));
        SHM_HIP(hipStreamSynchronize(S(st)));
    });
}

int shm_bootstrap_shared_a_s2c_first_chunks(shm_bootstrapper *b, shm_shared_a_keys *keys,
                                            const shm_ciphertext *const *in, size_t count,
                                            shm_ciphertext *const *out, const shm_kswitch_key *relin,
                                            double out_scale, shm_stream_t st) {
    return guard([&] {
        boot::Bootstrapper &bs = ready(b, S(st));
        need(keys, "keys");
        need(relin, "relin_keys");
        const size_t n = keys->k->n;
        if (count % n) throw InvalidArgument("the ciphertext count must be a multiple of the batch size n");
        if (count && (!in || !out)) throw InvalidArgument("in/out is null");
        for (size_t i = 0; i < count; i++) {
            need(in[i], "in[i]");
            need(out[i], "out[i]");
        }
        std::vector<const Ciphertext *> vi(n);
        std::vector<Ciphertext *> vo(n);
        for (size_t c = 0; c < count; c += n) {
            for (size_t i = 0; i < n; i++) {
                vi[i] = &in[c + i]->ct;
                vo[i] = &out[c + i]->ct;
            }
            boot::bootstrap_shared_a_s2c_first(bs, *keys->k, vi, vo, relin->k, out_scale, S(st));
        }
        SHM_HIP(hipStreamSynchronize(S(st)));
    });
}

int shm_bootstrap_shared_slot_to_coeff(shm_bootstrapper *b, shm_shared_a_keys *keys, shm_shared_a_batch *batch,
                                       shm_stream_t st) {
    return guard([&] {
        boot::Bootstrapper &bs = ready(b, S(st));
        need(keys, "keys");
        need(batch, "batch");
        boot::shared_slot_to_coeff(bs, *keys->k, batch->b, S(st));
    });
}

int shm_bootstrap_shared_coeff_to_slot(shm_bootstrapper *b, shm_shared_a_keys *keys, shm_shared_a_batch *batch,
                                       shm_shared_a_batch *re, shm_shared_a_batch *im, shm_stream_t st) {
    return guard([&] {
        boot::Bootstrapper &bs = ready(b, S(st));
        need(keys, "keys");
        need(batch, "batch");
        need(re, "re");
        need(im, "im");
        boot::shared_coeff_to_slot(bs, *keys->k, batch->b, re->b, im->b, S(st));
    });
}

// ---- public-key encryption, input packing, facade operations ------------------------------
struct shm_public_key {
    PublicKey pk;
};

extern "C" {

int shm_public_key_generate(shm_context *ctx, const shm_secret_key *sk, uint64_t seed, shm_public_key **out) {
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(out, "out");
        auto h = std::make_unique<shm_public_key>();
        gen_public_key(*ctx->c, sk->sk, seed, h->pk, nullptr);
        SHM_HIP(hipStreamSynchronize(nullptr));
        *out = h.release();
    });
}

int shm_public_key_destroy(shm_public_key *pk) {
    return guard([&] {
        if (pk) SHM_HIP(hipDeviceSynchronize());
        delete pk;
    });
}

int shm_public_key_export(const shm_public_key *pk, uint64_t *host, size_t n_words) {
    return guard([&] {
        need(pk, "pk");
        need(host, "host");
        const Context &c = *pk->pk.ctx;
        const size_t words = 2 * c.nQP() * c.N;
        if (n_words != words) throw InvalidArgument("public key export: n_words must be 2 * (size_q + size_p) * N");
        SHM_HIP(hipMemcpy(host, pk->pk.pk.as<void>(), words * 8, hipMemcpyDeviceToHost));
    });
}

int shm_encrypt_asymmetric(shm_context *ctx, const shm_public_key *pk, const shm_plaintext *pt, uint64_t seed,
                           shm_ciphertext *out, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(pk, "pk");
        need(pt, "pt");
        need(out, "out");
        encrypt_asymmetric(*ctx->c, pk->pk, pt->pt.data(), pt->pt.chain, pt->pt.scale, seed, out->ct, S(st));
    });
}

int shm_batch_input(shm_context *ctx, const shm_public_key *pk, const double *X, int num_X, int num_row, int num_col,
                    double scale, uint64_t seed, shm_ciphertext *const *out, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(pk, "pk");
        if (num_X <= 0 || num_row <= 0 || num_col < 0) throw InvalidArgument("batch_input: bad dimensions");
        if (num_col > 0) {
            need(X, "X");
            need(out, "out");
        }
        std::vector<Ciphertext *> os((size_t)num_col);
        for (int i = 0; i < num_col; i++) {
            need(out[i], "out[i]");
            os[i] = &out[i]->ct;
        }
        batch_input(*ctx->c, pk->pk, X, (size_t)num_X, (size_t)num_row, (size_t)num_col, scale, seed, os, S(st));
    });
}

int shm_bias_vec(const int *lengths, int num_X, int num_row, int *out, size_t n) {
    return guard([&] {
        if (num_X <= 0 || num_row <= 0) throw InvalidArgument("bias_vec: bad dimensions");
        need(lengths, "lengths");
        need(out, "out");
        if (n != (size_t)num_X * (size_t)num_row) throw InvalidArgument("bias_vec: n must be num_X * num_row");
        std::fill(out, out + n, 0);
        for (int i = 0; i < num_X; i++) {
            if (lengths[i] < 0 || lengths[i] > num_row) throw InvalidArgument("bias_vec: length out of range");
            for (int j = 0; j < lengths[i]; j++) out[(size_t)j * num_X + i] = 1;
        }
    });
}

int shm_multiply_const_inplace(shm_context *ctx, shm_ciphertext *ct, double value, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        facade_multiply_const(*ctx->c, ct->ct, value, 0.0, S(st));
    });
}

int shm_multiply_const_complex_inplace(shm_context *ctx, shm_ciphertext *ct, double re, double im,
                                       shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        facade_multiply_const(*ctx->c, ct->ct, re, im, S(st));
    });
}

int shm_add_const_inplace(shm_context *ctx, shm_ciphertext *ct, double value, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        facade_add_const(*ctx->c, ct->ct, value, 0.0, S(st));
    });
}

int shm_multiply_vector_inplace_reduced_error(shm_context *ctx, shm_ciphertext *ct, const double *re,
                                              const double *im, size_t n, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        if (n) need(re, "re");
        facade_multiply_vector(*ctx->c, ct->ct, re, im, n, S(st));
    });
}

int shm_add_inplace_reduced_error(shm_context *ctx, shm_ciphertext *ct1, const shm_ciphertext *ct2,
                                  shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct1, "ct1");
        need(ct2, "ct2");
        facade_reduced_error(*ctx->c, ct1->ct, ct2->ct, 0, nullptr, S(st));
    });
}

int shm_sub_inplace_reduced_error(shm_context *ctx, shm_ciphertext *ct1, const shm_ciphertext *ct2,
                                  shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct1, "ct1");
        need(ct2, "ct2");
        facade_reduced_error(*ctx->c, ct1->ct, ct2->ct, 1, nullptr, S(st));
    });
}

int shm_multiply_inplace_reduced_error(shm_context *ctx, shm_ciphertext *ct1, const shm_ciphertext *ct2,
                                       const shm_kswitch_key *relin, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct1, "ct1");
        need(ct2, "ct2");
        need(relin, "relin_keys");
        facade_reduced_error(*ctx->c, ct1->ct, ct2->ct, 2, &relin->k, S(st));
    });
}

int shm_double_inplace(shm_context *ctx, shm_ciphertext *ct, shm_stream_t st) {
    return guard([&] {
        need(ctx, "ctx");
        need(ct, "ct");
        launch_add(*ctx->c, ct->ct.data(), ct->ct.data(), ct->ct.Ql(), ct->ct.size, S(st));
    });
}

}  // extern "C"

// ---- graph-captured bootstrapping ------------------------------------------------------------
struct shm_boot_graph {
    std::unique_ptr<boot::BootGraph> g;
};

extern "C" {

int shm_boot_graph_create_shared_a(shm_bootstrapper *b, shm_shared_a_keys *keys, const shm_kswitch_key *relin,
                                   double in_scale, double out_scale, shm_stream_t st, shm_boot_graph **out) {
    return guard([&] {
        boot::Bootstrapper &bs = ready(b, S(st));
        need(keys, "keys");
        need(relin, "relin_keys");
        need(out, "out");
        if (!(in_scale > 0)) throw InvalidArgument("in_scale must be positive");
        boot::SharedAKeys *K = keys->k.get();
        const KSwitchKey *rk = &relin->k;
        boot::ChunkFn body = [&bs, K, rk, out_scale](const std::vector<const Ciphertext *> &in,
                                                      const std::vector<Ciphertext *> &o, hipStream_t s) {
            boot::bootstrap_shared_a_s2c_first(bs, *K, in, o, *rk, out_scale, s);
        };
        auto h = std::make_unique<shm_boot_graph>();
        h->g = boot::make_boot_graph(*bs.ctx, K->n, K->fmt_chain, in_scale, body, S(st));
        *out = h.release();
    });
}

int shm_boot_graph_create_ordinary(shm_bootstrapper *b, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                                   uint32_t group, double in_scale, double out_scale, shm_stream_t st,
                                   shm_boot_graph **out) {
    return guard([&] {
        boot::Bootstrapper &bs = ready(b, S(st));
        need(gk, "galois_keys");
        need(relin, "relin_keys");
        need(out, "out");
        if (group < 1 || group > 8) throw InvalidArgument("group: 1..8 ciphertexts");
        if (!(in_scale > 0)) throw InvalidArgument("in_scale must be positive");
        const GaloisKeys *g = &gk->g;
        const KSwitchKey *rk = &relin->k;
        boot::ChunkFn body = [&bs, g, rk, out_scale](const std::vector<const Ciphertext *> &in,
                                                      const std::vector<Ciphertext *> &o, hipStream_t s) {
            boot::bootstrap_s2c_first_batch(bs, in, o, *g, *rk, out_scale, s);
        };
        const size_t chain = bs.ctx->chain_of_ql(bs.prm.s2c_input_primes);
        auto h = std::make_unique<shm_boot_graph>();
        h->g = boot::make_boot_graph(*bs.ctx, group, chain, in_scale, body, S(st));
        *out = h.release();
    });
}

int shm_boot_graph_run(shm_boot_graph *g, const shm_ciphertext *const *in, size_t count, shm_ciphertext *const *out) {
    return guard([&] {
        need(g, "graph");
        if (count && (!in || !out)) throw InvalidArgument("in/out is null");
        std::vector<const Ciphertext *> vi(count);
        std::vector<Ciphertext *> vo(count);
        for (size_t i = 0; i < count; i++) {
            need(in[i], "in[i]");
            need(out[i], "out[i]");
            vi[i] = &in[i]->ct;
            vo[i] = &out[i]->ct;
        }
        boot::run_boot_graph(*g->g, vi, vo);
    });
}

int shm_boot_graph_info(const shm_boot_graph *g, uint32_t *chunk, uint64_t *nodes) {
    return guard([&] {
        need(g, "graph");
        if (chunk) *chunk = (uint32_t)g->g->chunk;
        if (nodes) *nodes = (uint64_t)g->g->nodes;
    });
}

int shm_boot_graph_destroy(shm_boot_graph *g) {
    return guard([&] { delete g; });
}

}  // extern "C"

// ---- measurement hooks -------------------------------------------------------------------------
extern "C" {

int shm_probe_fused_ntt_begin(void) {
    return guard([&] { probe_fused_ntt_begin(); });
}

int shm_probe_fused_ntt_end(uint64_t *launches, double *ms, double *bytes) {
    return guard([&] {
        size_t n = 0;
        probe_fused_ntt_end(&n, ms, bytes);
        if (launches) *launches = n;
    });
}

int shm_probe_plain_ntt_last(uint64_t *launches, double *ms, double *bytes) {
    return guard([&] {
        size_t n = 0;
        probe_plain_ntt_last(&n, ms, bytes);
        if (launches) *launches = n;
    });
}

int shm_keyswitch_counters(uint64_t *modup, uint64_t *inner, uint64_t *moddown, size_t cap) {
    return guard([&] { ks_counters(modup, inner, moddown, cap); });
}

int shm_keyswitch_counters_reset(void) {
    return guard([&] { ks_counters_reset(); });
}

int shm_set_tuning(const char *name, int32_t value) {
    return guard([&] {
        if (!name) throw InvalidArgument("tuning knob name is null");
        const std::string n(name);
        if (n == "matmul_overlap") {
            if (value != 0 && value != 1) throw InvalidArgument("matmul_overlap: 0 or 1");
            ks_tuning().matmul_overlap = value;
            return;
        }
        if (n == "ntt_tile_threads") {
            if (value != 256 && value != 512) throw InvalidArgument("ntt_tile_threads: 256 or 512");
            ks_tuning().ntt_tile_threads = value;
            return;
        }
        if (value < -1 || value > 4) throw InvalidArgument("tuning knob value: -1 (default) or 0..4");
        if (n == "ks_batch_slots_log2")
            ks_tuning().batch_slots_log2 = value;
        else if (n == "ks_multi_slots_log2")
            ks_tuning().multi_slots_log2 = value;
        else
            throw InvalidArgument("unknown tuning knob: " + n);
    });
}

int shm_autotune_enable(int32_t on) {
    return guard([&] { ks_autotune_enable(on != 0); });
}

int shm_autotune_clear(void) {
    return guard([&] { ks_autotune_clear(); });
}

int shm_autotune_report(char *buf, size_t cap, size_t *needed) {
    return guard([&] {
        const std::string r = ks_autotune_report();
        if (needed) *needed = r.size() + 1;
        if (buf && cap) {
            const size_t n = std::min(cap - 1, r.size());
            std::memcpy(buf, r.data(), n);
            buf[n] = 0;
        }
    });
}

int shm_get_tuning(const char *name, int32_t *value) {
    return guard([&] {
        if (!name || !value) throw InvalidArgument("null argument");
        const std::string n(name);
        if (n == "ks_batch_slots_log2")
            *value = ks_tuning().batch_slots_log2;
        else if (n == "ks_multi_slots_log2")
            *value = ks_tuning().multi_slots_log2;
        else if (n == "matmul_overlap")
            *value = ks_tuning().matmul_overlap;
        else if (n == "ntt_tile_threads")
            *value = ks_tuning().ntt_tile_threads;
        else
            throw InvalidArgument("unknown tuning knob: " + n);
    });
}

}  // extern "C"
