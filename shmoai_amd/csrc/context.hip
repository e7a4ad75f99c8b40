// Context construction: RNS primes, NTT twiddles, per-level base-conversion
// and rescale constants, encoder roots. Host-side precompute (as in the
// reference: modulus.cu, util/ntt.cu, rns.cu, ckks.cu), uploaded once to HBM.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>

#include "engine.h"

namespace shm {

namespace {

typedef unsigned __int128 hu128;

uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((hu128)a * b) % q); }
uint64_t h_powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    for (; e; e >>= 1) {
        if (e & 1) r = h_mulmod(r, b, q);
        b = h_mulmod(b, b, q);
    }
    return r;
}
uint64_t h_inv(uint64_t a, uint64_t q) { return h_powmod(a % q, q - 2, q); }
uint64_t h_shoup(uint64_t w, uint64_t q) { return (uint64_t)(((hu128)w << 64) / q); }

uint32_t h_brv(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) r = (r << 1) | ((x >> i) & 1u);
    return r;
}

// Miller-Rabin with the first twelve prime bases: exact for all 64-bit inputs.
bool h_is_prime(uint64_t n) {
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t p : bases)
        if (n % p == 0) return n == p;
    uint64_t d = n - 1;
    int r = 0;
    while (!(d & 1)) d >>= 1, r++;
    for (uint64_t a : bases) {
        uint64_t x = h_powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        int i = 1;
        for (; i < r; i++) {
            x = h_mulmod(x, x, n);
            if (x == n - 1) break;
        }
        if (i == r) return false;
    }
    return true;
}

// modulus.cu:79-108 + util/numth.cu:209-235: candidates 2^b - 2N*k + 1,
// descending; each bit-size class is consumed from the smallest found prime.
std::vector<uint64_t> make_primes(size_t N, const std::vector<int> &bits) {
    std::map<int, size_t> need;
    for (int b : bits) {
        if (b < 2 || b > 61) throw InvalidArgument("coeff bit size must be in [2, 61]");
        need[b]++;
    }
    std::map<int, std::vector<uint64_t>> found;
    for (auto &kv : need) {
        const uint64_t step = 2 * (uint64_t)N;
        const uint64_t top = uint64_t(1) << kv.first;
        if (top <= step) throw InvalidArgument("bit size too small for poly degree");
        uint64_t v = top - step + 1;
        const uint64_t floor_v = uint64_t(1) << (kv.first - 1);
        std::vector<uint64_t> list;
        while (list.size() < kv.second && v > floor_v) {
            if (h_is_prime(v)) list.push_back(v);
            v -= step;
        }
        if (list.size() < kv.second) throw LogicError("failed to find enough qualifying primes");
        found[kv.first] = list;
    }
    std::vector<uint64_t> out;
    for (int b : bits) {
        out.push_back(found[b].back());
        found[b].pop_back();
    }
    return out;
}

// util/numth.cu:311-333: the smallest primitive 2N-th root of unity.
uint64_t min_root_2n(uint64_t q, uint64_t two_n) {
    const uint64_t cof = (q - 1) / two_n;
    uint64_t g = 0;
    for (uint64_t x = 2;; x++) {
        uint64_t c = h_powmod(x, cof, q);
        if (h_powmod(c, two_n / 2, q) == q - 1) {
            g = c;
            break;
        }
    }
    const uint64_t g2 = h_mulmod(g, g, q);
    uint64_t best = g, cur = g;
    for (uint64_t i = 0; i < two_n / 2; i++) {
        best = std::min(best, cur);
        cur = h_mulmod(cur, g2, q);
    }
    return best;
}

template <class T>
void upload(DevBuf &b, const std::vector<T> &v) {
    b.alloc(std::max<size_t>(v.size() * sizeof(T), 8));
    if (!v.empty()) SHM_HIP(hipMemcpy(b.as<void>(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

}  // namespace

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&ptr_, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        throw std::bad_alloc();
    }
    hip_check(e, "hipMalloc");
    bytes_ = bytes;
}

// Stream-ordered block cache for the default stream. A block released on the
// default stream is handed out again only on the default stream, so stream
// order alone puts every kernel of its previous user ahead of its next user;
// the per-op hipMallocAsync / hipFreeAsync pairs (a few microseconds of host
// time each, several per key switch) disappear from the hot loops. Blocks on
// other streams go straight back to the pool. Sizes are rounded up to a
// bucket (powers of two to 1 MiB, then whole MiB), so a cached block always
// holds its bucket.
namespace {

struct BlockCache {
    std::mutex mu;
    std::map<std::pair<int, size_t>, std::vector<void *>> free;
    size_t cached = 0;
};

BlockCache &block_cache() {
    static BlockCache *c = new BlockCache();  // outlives static destruction order
    return *c;
}

constexpr size_t kCacheCap = size_t(32) << 30;

size_t bucket_of(size_t bytes) {
    constexpr size_t MiB = size_t(1) << 20;
    if (bytes <= MiB) {
        size_t b = 256;
        while (b < bytes) b <<= 1;
        return b;
    }
    return (bytes + MiB - 1) / MiB * MiB;
}

void flush_block_cache() {
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lk(bc.mu);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto &kv : bc.free) {
        (void)hipSetDevice(kv.first.first);
        for (void *p : kv.second) (void)hipFreeAsync(p, nullptr);
        bc.cached -= kv.first.second * kv.second.size();
        kv.second.clear();
    }
    (void)hipSetDevice(cur);
}

}  // namespace

void *pool_alloc(size_t bytes, hipStream_t s, int *dev_out) {
    const size_t b = bucket_of(bytes == 0 ? 8 : bytes);
    int dev = 0;
    SHM_HIP(hipGetDevice(&dev));
    *dev_out = dev;
    if (s == nullptr) {
        BlockCache &bc = block_cache();
        std::lock_guard<std::mutex> lk(bc.mu);
        auto it = bc.free.find({dev, b});
        if (it != bc.free.end() && !it->second.empty()) {
            void *p = it->second.back();
            it->second.pop_back();
            bc.cached -= b;
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMallocAsync(&p, b, s);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        flush_block_cache();
        (void)hipDeviceSynchronize();
        e = hipMallocAsync(&p, b, s);
    }
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        throw std::bad_alloc();
    }
    hip_check(e, "hipMallocAsync");
    return p;
}

void pool_free(void *p, size_t bytes, hipStream_t s, int dev) {
    if (!p) return;
    const size_t b = bucket_of(bytes == 0 ? 8 : bytes);
    if (s == nullptr) {
        BlockCache &bc = block_cache();
        std::lock_guard<std::mutex> lk(bc.mu);
        if (bc.cached + b <= kCacheCap) {
            bc.free[{dev, b}].push_back(p);
            bc.cached += b;
            return;
        }
    }
    (void)hipFreeAsync(p, s);
}

void DevBuf::alloc_async(size_t bytes, hipStream_t s) {
    release();
    if (bytes == 0) bytes = 8;
    ptr_ = pool_alloc(bytes, s, &dev_);
    bytes_ = bytes;
    s_ = s;
    async_ = true;
}

void DevBuf::use_on(hipStream_t s) {
    if (!ptr_ || !async_ || s == s_) return;
    hipEvent_t ev;
    SHM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    SHM_HIP(hipEventRecord(ev, s_));
    SHM_HIP(hipStreamWaitEvent(s, ev, 0));
    SHM_HIP(hipEventDestroy(ev));
    s_ = s;
}

void DevBuf::release() {
    if (ptr_) {
        if (async_) {
            pool_free(ptr_, bytes_, s_, dev_);
        } else {
            // queued work may still read the buffer: drain the device before freeing
            (void)hipDeviceSynchronize();
            (void)hipFree(ptr_);
        }
    }
    async_ = false;
    s_ = nullptr;
    ptr_ = nullptr;
    bytes_ = 0;
}

Scratch::Scratch(size_t bytes, hipStream_t s) : bytes_(bytes), s_(s) { ptr_ = pool_alloc(bytes, s, &dev_); }

Scratch::~Scratch() { pool_free(ptr_, bytes_, s_, dev_); }

uint32_t Context::elt_from_step(int step) const {
    // galois.cu:47-90
    const uint64_t m = 2 * (uint64_t)N;
    if (step == 0) return (uint32_t)(m - 1);
    const uint32_t mag = (uint32_t)std::abs(step);
    if (mag >= N / 2) throw InvalidArgument("step count too large");
    uint64_t e = 1;
    uint32_t k = step < 0 ? (uint32_t)(N / 2) - mag : mag;
    for (uint32_t i = 0; i < k; i++) e = (e * 5) & (m - 1);
    return (uint32_t)e;
}

Context::AuxStream::~AuxStream() {
    if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
}

hipStream_t Context::aux_stream() {
    std::lock_guard<std::mutex> lk(aux.mu);
    if (!aux.s) SHM_HIP(hipStreamCreateWithFlags(&aux.s, hipStreamNonBlocking));
    return aux.s;
}

const uint32_t *Context::galois_table(uint32_t elt, hipStream_t s) {
    std::lock_guard<std::mutex> lk(galois_mu);
    auto it = galois_tables.find(elt);
    if (it != galois_tables.end()) return it->second->as<uint32_t>();
    if (!(elt & 1) || elt >= 2 * N) throw InvalidArgument("Galois element is not valid");
    // galois.cuh: NTT-domain permutation for X -> X^elt on bit-reversed slots.
    std::vector<uint32_t> t(N);
    for (size_t i = N; i < 2 * N; i++) {
        uint32_t r = h_brv((uint32_t)i, logN + 1);
        uint64_t raw = (((uint64_t)elt * r) >> 1) & (N - 1);
        t[i - N] = h_brv((uint32_t)raw, logN);
    }
    auto b = std::make_unique<DevBuf>(N * sizeof(uint32_t));
    SHM_HIP(hipMemcpyAsync(b->as<void>(), t.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    SHM_HIP(hipStreamSynchronize(s));
    const uint32_t *p = b->as<uint32_t>();
    galois_tables[elt] = std::move(b);
    return p;
}

std::unique_ptr<Context> make_context(int logN, const std::vector<int> &bits, size_t nP) {
    if (logN < 10 || logN > 17) throw InvalidArgument("poly_modulus_degree must be 2^10 .. 2^17");
    if (nP == 0 || nP >= bits.size()) throw InvalidArgument("special_modulus_size must be in [1, #primes)");
    auto c = std::make_unique<Context>();
    SHM_HIP(hipGetDevice(&c->device));
    {
        // keep freed pool blocks cached instead of returning them at every sync
        hipMemPool_t pool;
        SHM_HIP(hipDeviceGetDefaultMemPool(&pool, c->device));
        uint64_t thr = UINT64_MAX;
        SHM_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
    }
    c->logN = logN;
    c->N = size_t(1) << logN;
    c->nP = nP;
    c->nQ = bits.size() - nP;
    c->dnum = c->nQ / c->nP;
    c->q = make_primes(c->N, bits);
    c->primes52 = true;
    for (uint64_t qv : c->q) c->primes52 = c->primes52 && qv > (uint64_t(1) << 45) && qv < (uint64_t(1) << 52);
    const size_t N = c->N, QP = c->q.size();

    std::vector<DevMod> mods(QP);
    std::vector<uint64_t> tw(QP * N), tws(QP * N), itw(QP * N), itws(QP * N), ninv(QP), ninvs(QP);
    c->psi.resize(QP);
    for (size_t p = 0; p < QP; p++) {
        const uint64_t q = c->q[p];
        hu128 mu = (~(hu128)0) / q;  // floor((2^128 - 1)/q) == floor(2^128/q) for non-power-of-two q
        mods[p] = DevMod{q, (uint64_t)(mu >> 64), (uint64_t)mu};
        const uint64_t psi = min_root_2n(q, 2 * N), ipsi = h_inv(psi, q);
        c->psi[p] = psi;
        uint64_t pw = 1, ipw = 1;
        for (size_t i = 0; i < N; i++) {
            const uint32_t r = h_brv((uint32_t)i, logN);
            tw[p * N + r] = pw;
            itw[p * N + r] = ipw;
            pw = h_mulmod(pw, psi, q);
            ipw = h_mulmod(ipw, ipsi, q);
        }
        for (size_t i = 0; i < N; i++) {
            tws[p * N + i] = h_shoup(tw[p * N + i], q);
            itws[p * N + i] = h_shoup(itw[p * N + i], q);
        }
        ninv[p] = h_inv(N, q);
        ninvs[p] = h_shoup(ninv[p], q);
    }
    upload(c->d_mod, mods);
    std::vector<uint64_t> io(QP), ios(QP);
    for (size_t p = 0; p < QP; p++) {
        io[p] = h_powmod(c->psi[p], N / 2, c->q[p]);
        ios[p] = h_shoup(io[p], c->q[p]);
    }
    upload(c->iota, io);
    upload(c->iota_s, ios);
    upload(c->tw, tw);
    upload(c->tw_s, tws);
    upload(c->itw, itw);
    upload(c->itw_s, itws);
    upload(c->ninv, ninv);
    upload(c->ninv_s, ninvs);

    // ModDown constants (rns.cu:95-128, 187-190): P mod q_i, P^-1 mod q_i,
    // (P/p_j)^-1 mod p_j, (P/p_j) mod q_i.
    const size_t nQ = c->nQ;
    std::vector<uint64_t> pinv(nQ), pinvs(nQ), pmod(nQ), pmods(nQ), phinv(nP), phinvs(nP), phmod(nQ * nP);
    for (size_t i = 0; i < nQ; i++) {
        const uint64_t q = c->q[i];
        uint64_t P = 1;
        for (size_t j = 0; j < nP; j++) P = h_mulmod(P, c->q[nQ + j] % q, q);
        pmod[i] = P;
        pmods[i] = h_shoup(P, q);
        pinv[i] = h_inv(P, q);
        pinvs[i] = h_shoup(pinv[i], q);
        for (size_t j = 0; j < nP; j++) {
            uint64_t h = 1;
            for (size_t k = 0; k < nP; k++)
                if (k != j) h = h_mulmod(h, c->q[nQ + k] % q, q);
            phmod[i * nP + j] = h;
        }
    }
    for (size_t j = 0; j < nP; j++) {
        const uint64_t p = c->q[nQ + j];
        uint64_t h = 1;
        for (size_t k = 0; k < nP; k++)
            if (k != j) h = h_mulmod(h, c->q[nQ + k] % p, p);
        phinv[j] = h_inv(h, p);
        phinvs[j] = h_shoup(phinv[j], p);
    }
    upload(c->pinv, pinv);
    upload(c->pinv_s, pinvs);
    upload(c->pmod, pmod);
    upload(c->pmod_s, pmods);
    upload(c->phat_inv, phinv);
    upload(c->phat_inv_s, phinvs);
    upload(c->phat_mod, phmod);

    // Per-level constants (rns.cu:63-94 rescale; rns.cu:131-186 modup).
    c->levels.resize(nQ + 1);
    const size_t alpha = nP;
    for (size_t Ql = 1; Ql <= nQ; Ql++) {
        auto L = std::make_unique<LevelConsts>();
        L->Ql = Ql;
        L->beta = (Ql + alpha - 1) / alpha;
        const size_t QlP = Ql + nP;
        std::vector<uint64_t> qi(std::max<size_t>(Ql, 2) - 1), qis(std::max<size_t>(Ql, 2) - 1);
        for (size_t i = 0; i + 1 < Ql; i++) {
            qi[i] = h_inv(c->q[Ql - 1] % c->q[i], c->q[i]);
            qis[i] = h_shoup(qi[i], c->q[i]);
        }
        std::vector<uint64_t> hi(Ql), his(Ql), hm(L->beta * alpha * QlP, 0), dm(L->beta * QlP, 0);
        for (size_t g = 0; g < L->beta; g++) {
            const size_t start = g * alpha, size = std::min(alpha, Ql - start);
            for (size_t k = 0; k < QlP; k++) {
                const uint64_t p = c->q[k < Ql ? k : nQ + (k - Ql)];
                uint64_t d = 1;
                for (size_t i = 0; i < size; i++) d = h_mulmod(d, c->q[start + i] % p, p);
                dm[g * QlP + k] = d;
            }
            for (size_t i = 0; i < size; i++) {
                const uint64_t qv = c->q[start + i];
                uint64_t h = 1;
                for (size_t k = 0; k < size; k++)
                    if (k != i) h = h_mulmod(h, c->q[start + k] % qv, qv);
                hi[start + i] = h_inv(h, qv);
                his[start + i] = h_shoup(hi[start + i], qv);
                for (size_t k = 0; k < QlP; k++) {
                    if (k >= start && k < start + size) continue;
                    const uint64_t p = c->q[k < Ql ? k : nQ + (k - Ql)];
                    uint64_t hp = 1;
                    for (size_t kk = 0; kk < size; kk++)
                        if (kk != i) hp = h_mulmod(hp, c->q[start + kk] % p, p);
                    hm[(g * alpha + i) * QlP + k] = hp;
                }
            }
        }
        if (Ql >= 2) {
            // basis B = {q_last, p_0 .. p_(nP-1)} of the merged ModDown + rescale
            const size_t nb = nP + 1, L1 = Ql - 1;
            std::vector<uint64_t> bprime(nb);
            bprime[0] = c->q[Ql - 1];
            for (size_t j = 0; j < nP; j++) bprime[j + 1] = c->q[nQ + j];
            std::vector<uint64_t> bhi(nb), bhis(nb), bhm(std::max<size_t>(L1, 1) * nb), bm(std::max<size_t>(L1, 1)),
                bi(std::max<size_t>(L1, 1)), bis(std::max<size_t>(L1, 1));
            for (size_t j = 0; j < nb; j++) {
                uint64_t h = 1;
                for (size_t k = 0; k < nb; k++)
                    if (k != j) h = h_mulmod(h, bprime[k] % bprime[j], bprime[j]);
                bhi[j] = h_inv(h, bprime[j]);
                bhis[j] = h_shoup(bhi[j], bprime[j]);
            }
            for (size_t i = 0; i < L1; i++) {
                const uint64_t qv = c->q[i];
                uint64_t b = 1;
                for (size_t k = 0; k < nb; k++) b = h_mulmod(b, bprime[k] % qv, qv);
                bm[i] = b;
                bi[i] = h_inv(b, qv);
                bis[i] = h_shoup(bi[i], qv);
                for (size_t j = 0; j < nb; j++) {
                    uint64_t h = 1;
                    for (size_t k = 0; k < nb; k++)
                        if (k != j) h = h_mulmod(h, bprime[k] % qv, qv);
                    bhm[i * nb + j] = h;
                }
            }
            upload(L->bhat_inv, bhi);
            upload(L->bhat_inv_s, bhis);
            upload(L->bhat_mod, bhm);
            upload(L->bmod, bm);
            upload(L->binv, bi);
            upload(L->binv_s, bis);
        }
        upload(L->qlast_inv, qi);
        upload(L->qlast_inv_s, qis);
        upload(L->hat_inv, hi);
        upload(L->hat_inv_s, his);
        upload(L->hat_mod, hm);
        upload(L->dig_mod, dm);
        c->levels[Ql] = std::move(L);
    }

    // Encoder roots: ckks.cu:30-87, fft.cu:11-40 (1/8 of the 2N-th roots by
    // polar form, the rest by symmetry).
    const size_t m = 2 * N;
    const double PI_ = 3.1415926535897932384626433832795028842;
    std::vector<double> base_re(m / 8 + 1), base_im(m / 8 + 1);
    for (size_t i = 0; i <= m / 8; i++) {
        const double th = 2 * PI_ * (double)i / (double)m;
        base_re[i] = std::cos(th);
        base_im[i] = std::sin(th);
    }
    std::vector<double> roots(2 * m);
    for (size_t i0 = 0; i0 < m; i0++) {
        // resolve the symmetry chain iteratively
        // Map the index into [0, m/8] through conj(m - i), -(i - m/2),
        // -conj(m/2 - i) and mirror(m/4 - i), then undo the maps in reverse.
        size_t idx = i0;
        double re = 0, im = 0;
        {
            std::vector<int> ops;  // 1 conj, 2 neg, 3 negconj, 4 mirror
            for (int hop = 0; hop < 4; hop++) {
                idx &= m - 1;
                if (idx <= m / 8) break;
                if (idx <= m / 4) {
                    ops.push_back(4);
                    idx = m / 4 - idx;
                    break;
                }
                if (idx <= m / 2) {
                    ops.push_back(3);
                    idx = m / 2 - idx;
                } else if (idx <= 3 * m / 4) {
                    ops.push_back(2);
                    idx = idx - m / 2;
                } else {
                    ops.push_back(1);
                    idx = m - idx;
                }
            }
            re = base_re[idx];
            im = base_im[idx];
            for (size_t k = ops.size(); k-- > 0;) {
                switch (ops[k]) {
                    case 4: std::swap(re, im); break;
                    case 3: re = 0.0 - re; im = 0.0 - (-im); break;
                    case 2: re = 0.0 - re; im = 0.0 - im; break;
                    case 1: im = -im; break;
                }
            }
        }
        roots[2 * i0] = re;
        roots[2 * i0 + 1] = im;
    }
    upload(c->enc_roots, roots);
    std::vector<uint32_t> grp(std::max<size_t>(N / 4, 1));
    uint64_t pos = 1;
    for (size_t i = 0; i < N / 4; i++) {
        grp[i] = (uint32_t)pos;
        pos = (pos * 5) & (m - 1);
    }
    upload(c->enc_group, grp);
    SHM_HIP(hipDeviceSynchronize());
    return c;
}

// A fresh block is zero-filled unless the caller overwrites every word
// (zero = false).
void Ciphertext::resize(size_t new_size, size_t new_chain, hipStream_t s, bool zero) {
    const size_t nl = ctx->ql_of_chain(new_chain);
    if (new_size > cap_size || nl > cap_limbs) {
        const size_t bytes = new_size * nl * ctx->N * sizeof(uint64_t);
        DevBuf nb;
        nb.alloc_async(bytes, s);
        if (zero) launch_fill_zero(nb.as<uint64_t>(), bytes / sizeof(uint64_t), s);
        buf.use_on(s);
        buf = std::move(nb);  // the old block returns to the pool after the work queued on s
        cap_size = new_size;
        cap_limbs = nl;
    } else {
        buf.use_on(s);
    }
    size = new_size;
    chain = new_chain;
}

}  // namespace shm
