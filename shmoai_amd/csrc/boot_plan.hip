// Bootstrapping plan, host FP64: butterfly-stage factorisation of the slot
// transforms merged into three diagonal layers, their baby-step/giant-step
// schedule, the encoded diagonals, and the EvalMod Chebyshev coefficients.
//
// Reference counterparts: genfftcoeff_3 / geninvfftcoeff_3 and the
// orig-coefficient butterflies (Bootstrapper.cu:634-727, 1356-2119) build
// the same three merged layers; bsgs_linear_transform /
// rotated_bsgs_linear_transform (Bootstrapper.cu:2168-2339) pre-rotate each
// diagonal by the giant offset exactly as LtLayer::pre does here.
#include <cmath>

#include "boot.h"

namespace shm {
namespace boot {

namespace {

const long double kPi = 3.141592653589793238462643383279502884L;

uint64_t pow5_mod(uint64_t e, uint64_t m) {
    uint64_t r = 1 % m, b = 5 % m;
    while (e) {
        if (e & 1) r = (r * b) % m;
        b = (b * b) % m;
        e >>= 1;
    }
    return r;
}

cplx unit_root(uint64_t k, uint64_t m) {
    const long double th = 2.0L * kPi * (long double)k / (long double)m;
    return cplx((double)cosl(th), (double)sinl(th));
}

std::vector<cplx> rotated(const std::vector<cplx> &v, long r) {
    const long n = (long)v.size();
    std::vector<cplx> out(n);
    long sh = ((r % n) + n) % n;
    for (long t = 0; t < n; t++) out[t] = v[(t + sh) % n];
    return out;
}

long floor_div(long a, long b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

}  // namespace

// Stage s (block length 2^s): out[p] = in[p] + w in[p + h] for the upper half
// of each block, out[p] = in[p - h] - w in[p] for the lower half, with
// w = exp(2 pi i (5^j mod 2^(s+2)) / 2^(s+2)), j the position in the half.
DiagMat fft_stage(size_t n, int s, bool inverse) {
    const size_t len = size_t(1) << s, h = len >> 1, lq = len << 2;
    if (len > n) throw InvalidArgument("fft stage beyond the slot count");
    DiagMat m;
    m.n = n;
    auto &d0 = m.d[0], &dp = m.d[h % n], &dm = m.d[(n - h) % n];
    d0.assign(n, 0.0);
    dp.assign(n, 0.0);
    dm.assign(n, 0.0);
    for (size_t p = 0; p < n; p++) {
        const size_t j = p & (len - 1);
        if (j < h) {
            const cplx w = unit_root(pow5_mod(j, lq), lq);
            if (!inverse) {
                d0[p] = 1.0;
                dp[p] = w;
            } else {
                d0[p] = 0.5;
                dp[p] = 0.5;
            }
        } else {
            const cplx w = unit_root(pow5_mod(j - h, lq), lq);
            if (!inverse) {
                dm[p] = 1.0;
                d0[p] = -w;
            } else {
                dm[p] = std::conj(w) * 0.5;
                d0[p] = -std::conj(w) * 0.5;
            }
        }
    }
    return m;
}

DiagMat compose(const DiagMat &a, const DiagMat &b) {
    const size_t n = a.n;
    DiagMat out;
    out.n = n;
    for (const auto &ea : a.d) {
        for (const auto &eb : b.d) {
            const size_t off = (ea.first + eb.first) % n;
            auto &dst = out.d[off];
            if (dst.empty()) dst.assign(n, 0.0);
            const auto &va = ea.second, &vb = eb.second;
            for (size_t j = 0; j < n; j++) dst[j] += va[j] * vb[(j + ea.first) % n];
        }
    }
    for (auto it = out.d.begin(); it != out.d.end();) {
        bool zero = true;
        for (const auto &v : it->second)
            if (v != cplx(0.0, 0.0)) {
                zero = false;
                break;
            }
        it = zero ? out.d.erase(it) : std::next(it);
    }
    return out;
}

// forward: B_last ... B_first; inverse: B_first^-1 ... B_last^-1
DiagMat stages_product(size_t n, int first, int last, bool inverse) {
    if (!inverse) {
        DiagMat m = fft_stage(n, first, false);
        for (int s = first + 1; s <= last; s++) m = compose(fft_stage(n, s, false), m);
        return m;
    }
    DiagMat m = fft_stage(n, last, true);
    for (int s = last - 1; s >= first; s--) m = compose(fft_stage(n, s, true), m);
    return m;
}

LtLayer make_layer(size_t n, int first, int last, bool inverse, double factor) {
    const DiagMat m = stages_product(n, first, last, inverse);
    LtLayer L;
    L.base = size_t(1) << (first - 1);
    L.factor = factor;
    const long span = 1L << (last - first + 1);
    const bool wrapped = L.base * (size_t)span == n;
    const long period = (long)(n / L.base);
    std::map<long, const std::vector<cplx> *> byk;
    long kmin = 0, kmax = 0;
    bool any = false;
    for (const auto &e : m.d) {
        if (e.first % L.base) throw LogicError("diagonal offset is not a multiple of the layer step");
        long k = (long)(e.first / L.base);
        if (!wrapped && k > span - 1) k -= period;
        byk[k] = &e.second;
        kmin = any ? std::min(kmin, k) : k;
        kmax = any ? std::max(kmax, k) : k;
        any = true;
    }
    // baby count: fewest rotations (g - 1 babies + non-zero giants); ties -> more babies
    const long M = kmax - kmin + 1;
    int best_g = 1;
    long best_cost = -1;
    for (long g = 1; g <= std::min<long>(M, 16); g++) {
        const long k0 = g * floor_div(kmin, g);
        const long ng = (kmax - k0) / g + 1;
        long giants = 0;
        for (long i = 0; i < ng; i++)
            if (((k0 + g * i) % period + period) % period != 0) giants++;
        const long cost = (g - 1) + giants;
        if (best_cost < 0 || cost <= best_cost) {
            best_cost = cost;
            best_g = (int)g;
        }
    }
    L.g = best_g;
    L.k0 = (long)best_g * floor_div(kmin, best_g);
    L.n_giants = (int)((kmax - L.k0) / best_g + 1);
    L.pre.assign(L.n_giants, std::vector<std::vector<cplx>>(L.g));
    for (int i = 0; i < L.n_giants; i++) {
        for (int j = 0; j < L.g; j++) {
            const long k = L.k0 + (long)L.g * i + j;
            auto it = byk.find(k);
            if (it == byk.end()) continue;
            auto v = rotated(*it->second, -L.giant_offset(i));
            for (auto &x : v) x *= factor;
            L.pre[i][j] = std::move(v);
        }
    }
    return L;
}

// Chebyshev interpolation at 2048 first-kind nodes, truncated to `deg`.
std::vector<double> evalmod_chebyshev(int K, int r, int deg) {
    const int M = 2048;
    if (deg < 1 || deg >= M) throw InvalidArgument("EvalMod degree out of range");
    std::vector<long double> f(M), th(M);
    const long double sc = 2.0L * kPi / (long double)(1L << r);
    for (int i = 0; i < M; i++) {
        th[i] = kPi * ((long double)i + 0.5L) / (long double)M;
        const long double y = cosl(th[i]);
        f[i] = cosl(sc * ((long double)K * y - 0.25L));
    }
    std::vector<double> c(deg + 1);
    for (int k = 0; k <= deg; k++) {
        long double acc = 0;
        for (int i = 0; i < M; i++) acc += f[i] * cosl((long double)k * th[i]);
        acc *= 2.0L / (long double)M;
        if (k == 0) acc *= 0.5L;
        c[k] = (double)acc;
    }
    return c;
}

void cheb_split(const std::vector<double> &p, size_t m, std::vector<double> &q, std::vector<double> &r) {
    const size_t deg = p.size() - 1;
    // T_m T_m = (T_2m + T_0) / 2 covers deg = 2m
    if (deg < m || deg > 2 * m) throw InvalidArgument("cheb_split: degree must be in [m, 2m]");
    q.assign(p.begin(), p.begin() + m);
    r.assign(deg - m + 1, 0.0);
    r[0] = p[m];
    for (size_t k = m + 1; k <= deg; k++) {
        r[k - m] += 2.0 * p[k];
        q[2 * m - k] -= p[k];
    }
}

std::vector<int> Bootstrapper::rotation_steps() const {
    std::vector<int> steps;
    auto add = [&](long s) {
        s = ((s % (long)n) + (long)n) % (long)n;
        if (s == 0) return;
        if (std::find(steps.begin(), steps.end(), (int)s) == steps.end()) steps.push_back((int)s);
    };
    for (const LtLayer *L : {&s2c[0], &s2c[1], &s2c[2], &c2s[0], &c2s[1], &c2s[2]}) {
        for (int j = 1; j < L->g; j++) add(L->baby_step(j));
        for (int i = 0; i < L->n_giants; i++) add(L->giant_offset(i));
    }
    std::sort(steps.begin(), steps.end());
    steps.insert(steps.begin(), 0);  // complex conjugation
    return steps;
}

void Bootstrapper::prepare(hipStream_t s) {
    if (prepared) return;
    if (!ctx) throw LogicError("bootstrapping plan has no context");
    Context &c = *ctx;
    for (LtLayer *L : {&s2c[0], &s2c[1], &s2c[2], &c2s[0], &c2s[1], &c2s[2]}) {
        const size_t Ql = c.ql_of_chain(L->chain);
        const double scale = (double)c.q[Ql - 1];
        L->pts.clear();
        L->pts.resize(L->n_giants);
        std::vector<double> re(n), im(n);
        for (int i = 0; i < L->n_giants; i++) {
            L->pts[i].resize(L->g);
            for (int j = 0; j < L->g; j++) {
                const auto &v = L->pre[i][j];
                if (v.empty()) continue;
                for (size_t t = 0; t < n; t++) {
                    re[t] = v[t].real();
                    im[t] = v[t].imag();
                }
                L->pts[i][j].alloc(Ql * c.N * sizeof(uint64_t));
                encode(c, re.data(), im.data(), n, Ql, scale, L->pts[i][j].as<uint64_t>(), s);
                SHM_HIP(hipStreamSynchronize(s));  // re/im are reused
            }
        }
    }
    prepared = true;
}

std::unique_ptr<Bootstrapper> make_bootstrapper(Context &c, const BootParams &p) {
    auto b = make_plan(c.N, c.nQ, p);
    b->ctx = &c;
    return b;
}

std::unique_ptr<Bootstrapper> make_plan(size_t N, size_t nQ, const BootParams &p) {
    if (p.s2c_input_primes != 4)
        throw InvalidArgument("S2C-first bootstrapping takes ciphertexts with exactly 4 primes (lS2C+1)");
    if (p.K < 1 || p.deg < 16 || p.deg > 63 || p.double_angles < 0 || p.double_angles > 6)
        throw InvalidArgument("EvalMod parameters out of range (K >= 1, 16 <= deg <= 63, 0 <= r <= 6)");
    if (!(p.layer3_factor > 0) || !(p.input_bound > 0)) throw InvalidArgument("factors must be positive");
    if (N < 16 || (N & (N - 1))) throw InvalidArgument("poly degree must be a power of two >= 16");
    auto b = std::make_unique<Bootstrapper>();
    b->prm = p;
    b->n = N / 2;
    b->logn = 0;
    while ((size_t(1) << b->logn) < b->n) b->logn++;
    const int logn = b->logn;
    if (logn < 3) throw InvalidArgument("ring degree too small for three transform layers");
    // level budget: S2C 3, C2S 3, EvalMod 1 (x/K) + ceil(log2(deg+1)) + 1 + r, output >= 1 prime
    int dpoly = 0;
    while ((1 << dpoly) < p.deg + 1) dpoly++;
    const size_t need = 3 + (size_t)(dpoly + 1 + p.double_angles) + 1 + 1;
    if (nQ < need) throw InvalidArgument("not enough data primes for bootstrapping");

    // slot-to-coefficient split (sfl_full_3): layers of p1, p2, p3 stages from the bottom
    const int p3 = logn / 3, p2 = (logn - p3) / 2, p1 = logn - p3 - p2;
    const size_t chain_in = nQ - p.s2c_input_primes + 1;
    b->s2c[0] = make_layer(b->n, 1, p1, false, 1.0);
    b->s2c[1] = make_layer(b->n, p1 + 1, p1 + p2, false, 1.0);
    b->s2c[2] = make_layer(b->n, p1 + p2 + 1, logn, false, p.layer3_factor);
    for (int i = 0; i < 3; i++) b->s2c[i].chain = chain_in + i;
    // coefficient-to-slot split (coefftoslot_full_3): q1, q2 top stages first; the
    // first layer also carries the 1/2 of the real / imaginary part extraction
    const int q1 = logn / 3, q2 = (logn - q1) / 2;
    b->c2s[0] = make_layer(b->n, logn - q1 + 1, logn, true, 0.5);
    b->c2s[1] = make_layer(b->n, logn - q1 - q2 + 1, logn - q1, true, 1.0);
    b->c2s[2] = make_layer(b->n, 1, logn - q1 - q2, true, 1.0);
    for (int i = 0; i < 3; i++) b->c2s[i].chain = 1 + i;
    for (const LtLayer *L : {&b->s2c[0], &b->s2c[1], &b->s2c[2], &b->c2s[0], &b->c2s[1], &b->c2s[2]})
        if (L->g > 16) throw LogicError("baby count above 16");
    b->cheb = evalmod_chebyshev(p.K, p.double_angles, p.deg);
    return b;
}

}  // namespace boot
}  // namespace shm
