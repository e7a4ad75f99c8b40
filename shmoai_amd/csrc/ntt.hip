// Negacyclic NTT / iNTT on uint64 RNS limbs for CDNA4.
//
// Transform (same convention as the reference, fntt_2d.cu / intt_2d.cu):
//   forward Cooley-Tukey, stage s (m = 2^s groups, distance t = N/2^(s+1)),
//   group i uses tw[m + i] with tw[brv(k)] = psi^k; the output is the
//   evaluation at psi^(2 brv(j) + 1) in slot j, fully reduced to [0, q).
//   inverse Gentleman-Sande with itw[brv(k)] = psi^-k, then x N^-1.
//
// Decomposition: N = N1 * N2 with N2 = min(4096, N / 16). The first log N1
// stages (4 or 5) only pair elements whose indices differ by multiples of N2,
// so each "column" {c + N2 r} is an independent N1-point transform: phase A
// gives one column to each thread and runs it in registers. The last log N2
// stages act on contiguous blocks of N2 elements (phase B): a 4096-word tile
// is staged in LDS, then rounds of up to 4 stages run in registers (radix-16
// units). A sub-transform whose first-stage group id is G uses twiddle
// tw[(G << u) + i'] at its local stage u for local group i'.
//
// Lazy ranges (Harvey): forward keeps [0, 4q) between stages, inverse [0, 2q);
// primes are < 2^61 so nothing wraps. When every prime of a launch is below
// 2^64 / (5 + 4 log N) (about 2^57.9) the forward stages skip the per-butterfly
// reduction of a altogether and take a cheaper quotient in the twiddle product
// (values stay below (4 + 4 log N) q). Outputs are canonical residues.
#include <algorithm>

#include <mutex>
#include <vector>

#include "engine.h"

namespace shm {

struct LimbMap {
    uint32_t n_active;    // transformed limbs per polynomial
    uint32_t lo;          // first limb
    uint32_t skip_lo;     // limbs >= skip_lo are shifted by skip_width
    uint32_t skip_width;
    uint32_t Ql;          // limbs >= Ql use special primes nQ + (k - Ql)
    uint32_t nQ;
    uint32_t first_prime; // prime of limb k < Ql is first_prime + k
    uint64_t poly_stride; // words between polynomials
};

__device__ __forceinline__ void limb_of(const LimbMap &m, uint32_t z, uint32_t &poly, uint32_t &k,
                                        uint32_t &prime) {
    poly = z / m.n_active;
    const uint32_t j = z - poly * m.n_active;
    k = m.lo + j;
    if (k >= m.skip_lo) k += m.skip_width;
    prime = k < m.Ql ? m.first_prime + k : m.nQ + (k - m.Ql);
}

namespace {

constexpr int kTile = 4096;
constexpr int kThreads = 256;

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
// Pointers the kernel reads out of device memory or selects between are known to
// address global memory: the cast lets their accesses use global instead of flat
// instructions (flat ones also hold the LDS wait counter).
template <typename V>
__device__ __forceinline__ __attribute__((address_space(1))) V *global_ptr(V *p) {
    return (__attribute__((address_space(1))) V *)p;
}
template <typename V>
__device__ __forceinline__ const __attribute__((address_space(1))) V *global_ptr(const V *p) {
    return (const __attribute__((address_space(1))) V *)p;
}

// Per-prime constant as the butterflies see it.
struct IntK {
    uint64_t q;
};

// Twiddle and constant tables of one launch (forward or inverse set).
struct NttTables {
    const DevMod *mods;
    const uint64_t *tw, *tws;   // twiddles and their Shoup quotients
    const uint64_t *ninv, *ninv_s;
};

template <int R>
__device__ __forceinline__ void ct_unit(uint64_t *x, uint32_t G, const uint64_t *__restrict__ tw,
                                        const uint64_t *__restrict__ tws, IntK k) {
    const uint64_t q = k.q, q2 = q << 1;
#pragma unroll
    for (int u = 0; u < R; u++) {
        const int h = 1 << (R - 1 - u);
#pragma unroll
        for (int l = 0; l < (1 << R); l++) {
            if ((l & h) == 0) {
                const uint32_t ti = (G << u) + (uint32_t)(l >> (R - u));
                const uint64_t w = tw[ti], ws = tws[ti];
                uint64_t a = x[l];
                a = a >= q2 ? a - q2 : a;
                const uint64_t t = mul_shoup_lazy(x[l + h], w, ws, q);
                x[l] = a + t;
                x[l + h] = a + q2 - t;
            }
        }
    }
}

// Forward butterflies that leave a unreduced: with t = x w mod q in [0, 4q) (the
// quotient estimate of mul_shoup_lazy4, at most 2 short), both outputs stay below
// a + 4q, so log N stages from inputs below 4q end below (4 + 4 log N) q, which fits
// 64 bits for the primes the launcher allows here. Saves the compare-select-subtract
// of every butterfly and part of the quotient's multiplies; the final pass reduces once.
struct IntKLazy {
    uint64_t q;
};

template <int R>
__device__ __forceinline__ void ct_unit(uint64_t *x, uint32_t G, const uint64_t *__restrict__ tw,
                                        const uint64_t *__restrict__ tws, IntKLazy k) {
    const uint64_t q = k.q, q4 = q << 2;
#pragma unroll
    for (int u = 0; u < R; u++) {
        const int h = 1 << (R - 1 - u);
#pragma unroll
        for (int l = 0; l < (1 << R); l++) {
            if ((l & h) == 0) {
                const uint32_t ti = (G << u) + (uint32_t)(l >> (R - u));
                const uint64_t a = x[l];
                const uint64_t t = mul_shoup_lazy4(x[l + h], tw[ti], tws[ti], q);
                x[l] = a + t;
                x[l + h] = a + q4 - t;
            }
        }
    }
}

template <int R>
__device__ __forceinline__ void gs_unit(uint64_t *x, uint32_t G, const uint64_t *__restrict__ tw,
                                        const uint64_t *__restrict__ tws, IntK k);
template <int R>
__device__ __forceinline__ void gs_unit(uint64_t *x, uint32_t G, const uint64_t *__restrict__ tw,
                                        const uint64_t *__restrict__ tws, IntKLazy k) {
    gs_unit<R>(x, G, tw, tws, IntK{k.q});  // the inverse keeps its reductions
}

template <int R>
__device__ __forceinline__ void gs_unit(uint64_t *x, uint32_t G, const uint64_t *__restrict__ tw,
                                        const uint64_t *__restrict__ tws, IntK k) {
    const uint64_t q = k.q, q2 = q << 1;
#pragma unroll
    for (int u = R - 1; u >= 0; u--) {
        const int h = 1 << (R - 1 - u);
#pragma unroll
        for (int l = 0; l < (1 << R); l++) {
            if ((l & h) == 0) {
                const uint32_t ti = (G << u) + (uint32_t)(l >> (R - u));
                const uint64_t w = tw[ti], ws = tws[ti];
                const uint64_t a = x[l], b = x[l + h];
                uint64_t s = a + b;
                s = s >= q2 ? s - q2 : s;
                x[l] = s;
                x[l + h] = mul_shoup_lazy(a + q2 - b, w, ws, q);
            }
        }
    }
}

// LDS layout of a tile: rows of T words. Several rows (T < 4096) get one word
// of padding each so lanes walking down a column hit distinct banks; within a
// row of at least 256 words the low 4 index bits are XOR-ed with bits 4..7, so
// the last round (lanes 16 words apart) is spread over the banks too.
template <int T>
__host__ __device__ constexpr int row_stride() {
    return T >= kTile ? T : T + 1;
}
template <int T>
__device__ __forceinline__ int swz(int r) {
    return T >= 256 ? (r ^ ((r >> 4) & 15)) : r;
}

// Round split of LOGT stages into chunks of at most MAXR (4: radix-16 units of
// 256-thread blocks; 3: radix-8 units of 512-thread blocks).
__host__ __device__ constexpr int n_rounds(int logt, int maxr) { return (logt + maxr - 1) / maxr; }
__host__ __device__ constexpr int round_len(int logt, int r, int maxr) {
    // distribute as evenly as possible, larger rounds first
    return logt / n_rounds(logt, maxr) + (r < logt % n_rounds(logt, maxr) ? 1 : 0);
}
__host__ __device__ constexpr int round_start(int logt, int r, int maxr) {
    int s = 0;
    for (int i = 0; i < r; i++) s += round_len(logt, i, maxr);
    return s;
}

// One round over the LDS tile: SUBS independent sub-transforms of size 2^LOGT
// (row stride TP words). Sub-transform `sub` has first-stage group id
// gbase(sub) at local stage 0.
template <int LOGT, int R, bool FWD, int NT, typename E, typename K>
__device__ __forceinline__ void lds_round(E *lds, int s0, uint32_t gmul, uint32_t gadd_step, uint32_t gadd0,
                                          const E *__restrict__ tw, const E *__restrict__ tws, K k) {
    constexpr int T = 1 << LOGT;
    constexpr int SUBS = kTile / T;
    constexpr int TP = row_stride<T>();
    const int T0 = T >> s0;          // size of the sub-transforms at this round
    const int ups = T0 >> R;         // units per sub-transform
    const int units = (T >> R) * SUBS;
    for (int w = threadIdx.x; w < units; w += NT) {
        const int sub = w % SUBS;
        const int w2 = w / SUBS;
        const int kk = w2 / ups;
        const int o = w2 - kk * ups;
        // group id of the whole phase for this sub-transform, shifted by s0
        uint32_t G = ((gmul + gadd0 + gadd_step * (uint32_t)sub) << s0) + (uint32_t)kk;
        // one sub-transform of at least 64 units: the 64 lanes of a wave share kk,
        // so G is wave-uniform and the twiddles come through scalar loads
        if (SUBS == 1 && ups >= 64) G = __builtin_amdgcn_readfirstlane(G);
        E *row = lds + sub * TP;
        const int i0 = kk * T0 + o;
        E x[1 << R];
#pragma unroll
        for (int l = 0; l < (1 << R); l++) x[l] = row[swz<T>(i0 + l * ups)];
        if (FWD)
            ct_unit<R>(x, G, tw, tws, k);
        else
            gs_unit<R>(x, G, tw, tws, k);
#pragma unroll
        for (int l = 0; l < (1 << R); l++) row[swz<T>(i0 + l * ups)] = x[l];
    }
}

struct NoHook {
    __device__ void operator()() const {}
};

// before_last runs (forward only) just before the last round
template <int LOGT, bool FWD, int NT, typename E, typename K, typename H = NoHook>
__device__ __forceinline__ void run_rounds(E *lds, uint32_t gmul, uint32_t gadd_step, uint32_t gadd0, const E *tw,
                                           const E *tws, K k, H before_last = H{}) {
    constexpr int MAXR = NT >= 512 ? 3 : 4;
    constexpr int NR = n_rounds(LOGT, MAXR);
    if (FWD) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int R = round_len(LOGT, r, MAXR), s0 = round_start(LOGT, r, MAXR);
            if (r == NR - 1) before_last();
            switch (R) {
                case 4: lds_round<LOGT, 4, true, NT>(lds, s0, gmul, gadd_step, gadd0, tw, tws, k); break;
                case 3: lds_round<LOGT, 3, true, NT>(lds, s0, gmul, gadd_step, gadd0, tw, tws, k); break;
                case 2: lds_round<LOGT, 2, true, NT>(lds, s0, gmul, gadd_step, gadd0, tw, tws, k); break;
                default: lds_round<LOGT, 1, true, NT>(lds, s0, gmul, gadd_step, gadd0, tw, tws, k); break;
            }
            __syncthreads();
        }
    } else {
#pragma unroll
        for (int r = NR - 1; r >= 0; r--) {
            const int R = round_len(LOGT, r, MAXR), s0 = round_start(LOGT, r, MAXR);
            switch (R) {
                case 4: lds_round<LOGT, 4, false, NT>(lds, s0, gmul, gadd_step, gadd0, tw, tws, k); break;
                case 3: lds_round<LOGT, 3, false, NT>(lds, s0, gmul, gadd_step, gadd0, tw, tws, k); break;
                case 2: lds_round<LOGT, 2, false, NT>(lds, s0, gmul, gadd_step, gadd0, tw, tws, k); break;
                default: lds_round<LOGT, 1, false, NT>(lds, s0, gmul, gadd_step, gadd0, tw, tws, k); break;
            }
            __syncthreads();
        }
    }
}

// Extra I/O of fused variants.
//   IO_PLAIN   : in-place transform of `data`.
//   IO_REDUCE  : (phase A, forward) the input of limb k of polynomial p is
//                src[p * N + idx] reduced mod the limb's prime (a broadcast of
//                one coefficient-form limb to other primes, rns.cu:1125-1137).
//   IO_RESCALE : (phase B, forward) instead of storing v = NTT(delta), store
//                (acc - v) * qinv mod q to outs[p / ppo] + (p % ppo) L1 N + k N
//                (rns.cu:1138-1155 fused with the final NTT pass). With
//                out_packed, polynomial p reads acc at its acc_limbs stride and
//                writes out_p[p], adding onto it when bit p of acc_mask is set
//                (the ModDown combine, rns_bconv.cu:680-690 and 776-828, fused the same way).
//   IO_GATHER  : (phase B, inverse) limb k of polynomial p is read from
//                acc_p.p[p] + k N (separate input buffers) and the transform is
//                written to `data`: the gather copy fused into the first pass.
//   IO_ENCODE  : (phase A, forward) the input of limb k of polynomial p is the
//                rounded coefficient of the encoder's FFT output fft[p] (index
//                i < N/2: round(re), i >= N/2: round(im) of entry i - N/2) as a
//                residue, i.e. the encoder's decompose step fused into the load.
enum { IO_PLAIN = 0, IO_REDUCE = 1, IO_RESCALE = 2, IO_ENCODE = 3, IO_GATHER = 4 };

struct NttIO {
    const uint64_t *src;      // IO_REDUCE: [n_polys][N]
    const uint64_t *acc;      // IO_RESCALE: [n_polys][acc_limbs][N]
    uint64_t *const *outs;    // IO_RESCALE: output base per group of ppo polys (device array)
    uint64_t *out_single;     // IO_RESCALE: used instead of outs when outs is null
    const uint64_t *qinv;     // IO_RESCALE: per prime
    const uint64_t *qinv_s;
    uint32_t acc_limbs;
    uint32_t ppo;
    uint64_t center;          // IO_REDUCE: 0, or h = q_last >> 1: values above h stand for v - q_last
    uint32_t packed;          // IO_RESCALE: nonzero -> polynomial p reads acc_p[p], writes out_p[p]
    uint32_t out_packed;      // IO_RESCALE: nonzero -> polynomial p reads acc, writes out_p[p]
    uint64_t acc_mask;        // IO_RESCALE with out_packed: bit p set -> out_p[p] += result
    uint32_t out_limbs;       // IO_RESCALE with outs: limbs per output polynomial (0: the launch's)
    PolyPack acc_p, out_p;    // IO_RESCALE with packed / out_packed; IO_GATHER inputs in acc_p
    const uint32_t *add_perm; // IO_RESCALE with out_packed: add add_src.p[p] read through this
    PolyPack add_src;         // NTT-domain permutation instead of out_p[p] (acc_mask bit p set)
    const double2 *fft;       // IO_ENCODE: [n_polys][N/2]
};

// Phase A (columns, first log N1 stages, log N1 <= 5). Each thread owns one
// column {c + N2 r} and runs all its stages in registers: no LDS, no barrier;
// the 256 threads of a block read and write 256 consecutive words per row.
// FWD: forward stages 0..logN1-1 (output stays lazy).
// !FWD: inverse stages logN1-1..0, then x N^-1, output canonical.
// At most 80 VGPRs (6 waves per SIMD): the inverse form would otherwise take 99. The
// encoder's form keeps its registers (it would spill).
template <int LOGT, bool FWD, int IO, bool LAZY>
__global__ __launch_bounds__(kThreads, IO == IO_ENCODE ? 1 : 6) void k_ntt_cols(uint64_t *__restrict__ data, LimbMap mp, NttTables tb,
                                                       int logN, NttIO io) {
    constexpr int T = 1 << LOGT;
    const size_t N = size_t(1) << logN;
    const size_t N2 = N >> LOGT;
    const uint32_t groups = (uint32_t)(N2 / kThreads);
    const uint32_t z = blockIdx.x / groups;
    const size_t c = (size_t)(blockIdx.x % groups) * kThreads + threadIdx.x;
    uint32_t poly, k, prime;
    limb_of(mp, z, poly, k, prime);
    uint64_t *base = data + poly * mp.poly_stride + (size_t)k * N + c;
    // every column's group id at stage 0 is 1 (phase A starts at s = 0)
    {
        const uint64_t q = tb.mods[prime].q;
        uint64_t x[T];
        if (IO == IO_REDUCE) {
            const DevMod m = tb.mods[prime];
            const uint64_t *src = io.src + (size_t)poly * N + c;
#pragma unroll
            for (int r = 0; r < T; r++) {
                const uint64_t v = src[r * N2];
                uint64_t y = reduce64(v, m);
                if (io.center && v > io.center) y = sub_mod(y, reduce64(2 * io.center + 1, m), m.q);
                x[r] = y;
            }
        } else if (IO == IO_ENCODE) {
            const DevMod m = tb.mods[prime];
            const double2 *fb = io.fft + (size_t)poly * (N / 2) + c;
            // rows r < T/2 lie below N/2 (real parts), row r + T/2 is the
            // imaginary part of the same entry
#pragma unroll
            for (int r = 0; r < T / 2; r++) {
                const double2 v = fb[r * N2];
                x[r] = residue_of_rounded(round(v.x), m);
                x[r + T / 2] = residue_of_rounded(round(v.y), m);
            }
        } else {
#pragma unroll
            for (int r = 0; r < T; r++) x[r] = base[r * N2];
        }
        const uint64_t *tw = tb.tw + (size_t)prime * N, *tws = tb.tws + (size_t)prime * N;
        if (FWD) {
            if (LAZY)
                ct_unit<LOGT>(x, 1u, tw, tws, IntKLazy{q});
            else
                ct_unit<LOGT>(x, 1u, tw, tws, IntK{q});
            // streaming store: the tile pass reads these words back long after
#pragma unroll
            for (int r = 0; r < T; r++) __builtin_nontemporal_store(x[r], &base[r * N2]);
        } else {
            gs_unit<LOGT>(x, 1u, tw, tws, IntK{q});
            const uint64_t ni = tb.ninv[prime], nis = tb.ninv_s[prime];
#pragma unroll
            for (int r = 0; r < T; r++) __builtin_nontemporal_store(mul_shoup(x[r], ni, nis, q), &base[r * N2]);
        }
    }
}

// Phase B (contiguous blocks of N2, last log N2 stages). LOGT = log N2.
// FWD: input lazy, output canonical. !FWD: input canonical (< 2q accepted),
// output lazy.
// Work order: workgroups are dispatched to the 8 XCDs round-robin, so block b
// runs on XCD b % 8. Each XCD takes a contiguous range of the work index
// O = (limb * tiles + tile) * n_polys + poly: the polynomials that share one
// prime and tile (hence one twiddle slice) run back to back on the same L2.
// 512-thread blocks ask for 6 waves per SIMD (80 VGPRs), three 32 KB tiles per CU.
template <int LOGT, bool FWD, int IO, bool LAZY, int NT = kThreads>
__global__ __launch_bounds__(NT, NT >= 512 ? 6 : 1) void k_ntt_blocks(uint64_t *__restrict__ data, LimbMap mp, NttTables tb,
                                                         int logN, NttIO io, uint32_t n_polys, uint32_t per_xcd) {
    constexpr int T = 1 << LOGT;
    constexpr int BLKS = kTile / T;
    constexpr int TP = row_stride<T>();
    __shared__ __align__(16) uint64_t lds[BLKS * TP];
    const size_t N = size_t(1) << logN;
    const uint32_t tiles = (uint32_t)(N / kTile);
    const uint32_t O = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
    if (O >= n_polys * mp.n_active * tiles) return;  // whole block, before any barrier
    const uint32_t pz = O % n_polys, rest = O / n_polys;
    const uint32_t z = pz * mp.n_active + rest / tiles;
    const int t0 = (int)(rest % tiles) * kTile;
    uint32_t poly, k, prime;
    limb_of(mp, z, poly, k, prime);
    uint64_t *base = data + poly * mp.poly_stride + (size_t)k * N + t0;
    // block b (global index t0/T + b) has group id N1 + b at its first stage
    const uint32_t N1 = (uint32_t)(N >> LOGT);
    const uint64_t q = tb.mods[prime].q, q2 = q << 1;
    const uint64_t *ld = IO == IO_GATHER ? io.acc_p.p[poly] + (size_t)k * N + t0 : base;
    // LDS slot of tile word i
    auto slot = [](int i) { return (i / T) * TP + swz<T>(i % T); };
    // Word pair (i, i + 1), i even, to and from LDS. In a full-width tile the
    // swizzle keeps the pair in one aligned 16-byte slot, swapped when bit 4 of i
    // is set, so a lane moves it with one 16-byte LDS access (a wave covers 1 KB
    // of consecutive slots: no bank conflicts).
    auto lds_put = [&](int i, u64x2 v) {
        if (T >= kTile) {
            const bool sw = (i >> 4) & 1;
            u64x2 o;
            o.x = sw ? v.y : v.x;
            o.y = sw ? v.x : v.y;
            *reinterpret_cast<u64x2 *>(&lds[swz<T>(i) & ~1]) = o;
        } else {
            lds[slot(i)] = v.x;
            lds[slot(i + 1)] = v.y;
        }
    };
    auto lds_get = [&](int i) -> u64x2 {
        u64x2 v;
        if (T >= kTile) {
            const bool sw = (i >> 4) & 1;
            const u64x2 o = *reinterpret_cast<const u64x2 *>(&lds[swz<T>(i) & ~1]);
            v.x = sw ? o.y : o.x;
            v.y = sw ? o.x : o.y;
        } else {
            v.x = lds[slot(i)];
            v.y = lds[slot(i + 1)];
        }
        return v;
    };
    // the tile comes in two words per lane (16-byte loads; the launcher checks the
    // alignment). Plain outputs leave one word per lane: 16-byte stores measured slower.
    constexpr int PAIRS = kTile / (2 * NT);  // pairs per lane
#pragma unroll
    for (int j = 0; j < PAIRS; j++) {
        const int i = 2 * (threadIdx.x + j * NT);
        lds_put(i, __builtin_nontemporal_load(reinterpret_cast<const u64x2 *>(ld + i)));
    }
    // rescale: the accumulator words are loaded before the last round, so they
    // arrive during it (loading them before all rounds cost occupancy)
    constexpr int PER = kTile / NT;
    const uint64_t *accp = nullptr;
    uint64_t *out = nullptr;
    if (FWD && IO == IO_RESCALE) {
        accp = io.packed ? io.acc_p.p[poly] + (size_t)k * N + t0
                         : io.acc + ((size_t)poly * io.acc_limbs + k) * N + t0;
        if (io.packed || io.out_packed) {
            out = io.out_p.p[poly] + (size_t)k * N + t0;
        } else {
            uint64_t *ob = io.outs ? io.outs[poly / io.ppo] : io.out_single;
            out = ob + ((size_t)(poly % io.ppo) * (io.out_limbs ? io.out_limbs : mp.n_active) + k) * N + t0;
        }
    }
    // accumulator and output in 16-byte pairs when both are aligned (block-uniform);
    // lane word j is then tile word 2 (tid + (j / 2) NT) + j % 2, else tid + j NT
    const bool vec = (((uintptr_t)accp | (uintptr_t)out) & 15) == 0;
    auto word_of = [&](int j) {
        return vec ? 2 * ((int)threadIdx.x + (j >> 1) * NT) + (j & 1) : (int)threadIdx.x + j * NT;
    };
    uint64_t accv[IO == IO_RESCALE ? PER : 1];
    auto load_acc = [&] {
        if (FWD && IO == IO_RESCALE) {
            if (vec) {
#pragma unroll
                for (int j = 0; j < PER; j += 2) {
                    const u64x2 v = __builtin_nontemporal_load(global_ptr(reinterpret_cast<const u64x2 *>(accp + word_of(j))));
                    accv[j] = v.x;
                    accv[j + 1] = v.y;
                }
            } else {
#pragma unroll
                for (int j = 0; j < PER; j++) accv[j] = *global_ptr(accp + word_of(j));
            }
        }
    };
    __syncthreads();
    if (LAZY)
        run_rounds<LOGT, FWD, NT>(lds, N1, 1u, (uint32_t)(t0 / T), tb.tw + (size_t)prime * N,
                                  tb.tws + (size_t)prime * N, IntKLazy{q}, load_acc);
    else
        run_rounds<LOGT, FWD, NT>(lds, N1, 1u, (uint32_t)(t0 / T), tb.tw + (size_t)prime * N,
                                  tb.tws + (size_t)prime * N, IntK{q}, load_acc);
    // canonical residue of a forward output: below 4q, or below (4 + 4 log N) q when lazy
    auto canon = [&](uint64_t v) -> uint64_t {
        if (LAZY) return reduce64(v, tb.mods[prime]);
        v = v >= q2 ? v - q2 : v;
        return v >= q ? v - q : v;
    };
    if (FWD && IO == IO_RESCALE) {
        const uint64_t qi = io.qinv[prime], qis = io.qinv_s[prime];
        // unreduced outputs stay below (4 + 4 logN) q (< 2^64 for the lazy primes): acc - v
        // is taken as acc + (4 + 4 logN) q - v, which the Shoup product accepts as it is
        const uint64_t kq = (uint64_t)(4 + 4 * logN) * q;
        const bool addto = io.out_packed && poly < 64 && ((io.acc_mask >> poly) & 1);
        auto result = [&](uint64_t acc, uint64_t raw) -> uint64_t {
            return LAZY ? mul_shoup(acc + kq - raw, qi, qis, q) : mul_shoup(sub_mod(acc, canon(raw), q), qi, qis, q);
        };
        auto addend = [&](int i) -> uint64_t {
            return *global_ptr(io.add_src.p[poly] + (size_t)k * N + io.add_perm[t0 + i]);
        };
        if (vec) {
#pragma unroll
            for (int j = 0; j < PER; j += 2) {
                const int i = word_of(j);
                const u64x2 raw = lds_get(i);
                u64x2 y;
                y.x = result(accv[j], raw.x);
                y.y = result(accv[j + 1], raw.y);
                if (addto) {
                    if (io.add_perm) {
                        y.x = add_mod(y.x, addend(i), q);
                        y.y = add_mod(y.y, addend(i + 1), q);
                    } else {
                        const u64x2 o = *global_ptr(reinterpret_cast<const u64x2 *>(out + i));
                        y.x = add_mod(y.x, o.x, q);
                        y.y = add_mod(y.y, o.y, q);
                    }
                }
                __builtin_nontemporal_store(y, global_ptr(reinterpret_cast<u64x2 *>(out + i)));
            }
        } else {
#pragma unroll
            for (int j = 0; j < PER; j++) {
                const int i = word_of(j);
                uint64_t y = result(accv[j], lds[slot(i)]);
                if (addto) y = add_mod(y, io.add_perm ? addend(i) : *global_ptr(out + i), q);
                __builtin_nontemporal_store(y, global_ptr(out + i));
            }
        }
    } else if (FWD) {
#pragma unroll
        for (int j = 0; j < kTile / NT; j++) {
            const int i = threadIdx.x + j * NT;
            __builtin_nontemporal_store(canon(lds[slot(i)]), &base[i]);
        }
    } else {
#pragma unroll
        for (int j = 0; j < kTile / NT; j++) {
            const int i = threadIdx.x + j * NT;
            base[i] = lds[slot(i)];
        }
    }
}

// optional events around the tile pass of the fused rescale (timing of one launch)
hipEvent_t g_rescale_tile_ev[2] = {nullptr, nullptr};

// Probe of the fused last pass (IO_RESCALE tile pass of forward NTTs: the ModDown
// combine and the rescale): while on, every eager launch is bracketed by a pair of
// timing events and its algorithmic bytes are recorded.
struct FusedProbe {
    std::mutex mu;
    bool on = false;
    std::vector<hipEvent_t> ev;  // start, stop per launch of the fused last pass
    double bytes = 0.0;
    std::vector<hipEvent_t> ev_plain;  // the same for the plain forward tile pass
    double bytes_plain = 0.0;
    size_t last_plain_launches = 0;
    double last_plain_ms = 0.0, last_plain_bytes = 0.0;
} g_probe;

template <bool FWD, int IOA, int IOB>
void launch_ntt(const Context &c, uint64_t *data, const LimbMap &mp, size_t total_limbs, hipStream_t s,
                const NttIO &io = NttIO{}) {
    if (total_limbs == 0) return;
    if (c.N < (size_t)kTile) throw InvalidArgument("NTT requires N >= 4096");
    if (total_limbs % mp.n_active != 0) throw LogicError("NTT limb count is not a multiple of the polynomial size");
    const int logN = c.logN;
    // Phase B takes whole 4096-word tiles (12 stages) and phase A the rest:
    // with few column stages, phase A moves rows of >= 2 KB contiguous words
    // (N >> log1 >= 4096 / 2^log1 columns of a tile) instead of short strided
    // segments.
    const int log2b = std::min(12, logN - 4);
    const int log1 = logN - log2b;
    const size_t N = c.N;
    const unsigned grid_cols = (unsigned)(total_limbs * ((N >> log1) / kThreads));
    const uint32_t n_polys = (uint32_t)(total_limbs / mp.n_active);
    const uint32_t per_xcd = (uint32_t)((total_limbs * (N / kTile) + 7) / 8);
    const unsigned grid_blk = 8u * per_xcd;
    NttTables tb{c.mods(),
                 FWD ? c.tw.as<uint64_t>() : c.itw.as<uint64_t>(),
                 FWD ? c.tw_s.as<uint64_t>() : c.itw_s.as<uint64_t>(),
                 c.ninv.as<uint64_t>(),
                 c.ninv_s.as<uint64_t>()};
    // unreduced forward butterflies when (5 + 4 log N) q < 2^64 for every prime of the
    // launch: values below (4 + 4 log N) q, plus one q of the fused epilogue's operand
    bool lazy = FWD;
    const uint64_t lazy_max = ~uint64_t(0) / (uint64_t)(5 + 4 * c.logN);
    for (uint32_t j = 0; j < mp.n_active && lazy; j++) {
        uint32_t k = mp.lo + j;
        if (k >= mp.skip_lo) k += mp.skip_width;
        const uint32_t prime = k < mp.Ql ? mp.first_prime + k : mp.nQ + (k - mp.Ql);
        lazy = c.q.at(prime) < lazy_max;
    }
    // tile passes in 512-thread blocks of radix-8 rounds (the default): the lazy forward
    // passes and the inverse ones; the forward passes of primes above the lazy bound keep
    // 256-thread radix-16 rounds (their fused form would spill at 80 registers)
    const bool tile512 = ks_tuning().ntt_tile_threads == 512;
    // the tile pass reads its input in 16-byte pairs of words
    bool aligned = ((uintptr_t)data & 15) == 0 && mp.poly_stride % 2 == 0;
    if (IOB == IO_GATHER)
        for (uint32_t p = 0; p < n_polys; p++) aligned = aligned && ((uintptr_t)io.acc_p.p[p] & 15) == 0;
    if (!aligned) throw InvalidArgument("NTT operands must be 16-byte aligned");
#define SHM_COLS(L)                                                                                    \
    case L:                                                                                            \
        if (lazy)                                                                                      \
            k_ntt_cols<L, FWD, IOA, true><<<grid_cols, kThreads, 0, s>>>(data, mp, tb, logN, io);      \
        else                                                                                           \
            k_ntt_cols<L, FWD, IOA, false><<<grid_cols, kThreads, 0, s>>>(data, mp, tb, logN, io);     \
        break;
#define SHM_BLKS(L)                                                                                          \
    case L:                                                                                                  \
        if (lazy && tile512)                                                                                 \
            k_ntt_blocks<L, FWD, IOB, true, 512><<<grid_blk, 512, 0, s>>>(data, mp, tb, logN, io, n_polys,   \
                                                                          per_xcd);                          \
        else if (lazy)                                                                                       \
            k_ntt_blocks<L, FWD, IOB, true><<<grid_blk, kThreads, 0, s>>>(data, mp, tb, logN, io, n_polys,    \
                                                                          per_xcd);                          \
        else if (!FWD && tile512)                                                                            \
            k_ntt_blocks<L, FWD, IOB, false, 512><<<grid_blk, 512, 0, s>>>(data, mp, tb, logN, io, n_polys,  \
                                                                           per_xcd);                         \
        else                                                                                                 \
            k_ntt_blocks<L, FWD, IOB, false><<<grid_blk, kThreads, 0, s>>>(data, mp, tb, logN, io, n_polys,   \
                                                                           per_xcd);                         \
        break;
    auto cols = [&]() {
        switch (log1) {
            SHM_COLS(4)
            SHM_COLS(5)
            default: throw InvalidArgument("unsupported ring degree for NTT");
        }
    };
    auto blks = [&]() {
        switch (log2b) {
            SHM_BLKS(8)
            SHM_BLKS(9)
            SHM_BLKS(10)
            SHM_BLKS(11)
            SHM_BLKS(12)
            default: throw InvalidArgument("unsupported ring degree for NTT");
        }
    };
#undef SHM_COLS
#undef SHM_BLKS
    if (FWD) {
        cols();
        const bool timed = IOB == IO_RESCALE && g_rescale_tile_ev[0] && g_rescale_tile_ev[1];
        if (timed) SHM_HIP(hipEventRecord(g_rescale_tile_ev[0], s));
        hipEvent_t pe[2] = {nullptr, nullptr};
        if (IOB == IO_PLAIN) {
            std::lock_guard<std::mutex> lk(g_probe.mu);
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (g_probe.on && hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
                SHM_HIP(hipEventCreate(&pe[0]));
                SHM_HIP(hipEventCreate(&pe[1]));
                g_probe.ev_plain.push_back(pe[0]);
                g_probe.ev_plain.push_back(pe[1]);
                // the transform's algorithmic bytes: every limb read and written once
                g_probe.bytes_plain += 16.0 * N * total_limbs;
            }
        }
        if (IOB == IO_RESCALE) {
            std::lock_guard<std::mutex> lk(g_probe.mu);
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (g_probe.on && hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
                SHM_HIP(hipEventCreate(&pe[0]));
                SHM_HIP(hipEventCreate(&pe[1]));
                g_probe.ev.push_back(pe[0]);
                g_probe.ev.push_back(pe[1]);
                // per output limb: the result written and the combined operand read
                // (plus the output read where the pass accumulates onto it); the
                // transform's intermediate written by the column pass is not counted
                size_t acc_polys = 0;
                if (io.out_packed)
                    for (uint32_t p = 0; p < n_polys && p < 64; p++) acc_polys += (io.acc_mask >> p) & 1;
                g_probe.bytes += 16.0 * N * total_limbs + 8.0 * N * mp.n_active * acc_polys;
            }
        }
        if (pe[0]) SHM_HIP(hipEventRecord(pe[0], s));
        blks();
        if (pe[1]) SHM_HIP(hipEventRecord(pe[1], s));
        if (timed) SHM_HIP(hipEventRecord(g_rescale_tile_ev[1], s));
    } else {
        blks();
        cols();
    }
    SHM_HIP(hipGetLastError());
}

}  // namespace

void ntt_forward(const Context &c, uint64_t *data, size_t n_polys, size_t n_limbs, size_t first_prime,
                 hipStream_t s) {
    LimbMap mp{(uint32_t)n_limbs, 0, 0xffffffffu, 0, 0xffffffffu, (uint32_t)c.nQ, (uint32_t)first_prime,
               (uint64_t)n_limbs * c.N};
    launch_ntt<true, IO_PLAIN, IO_PLAIN>(c, data, mp, n_polys * n_limbs, s);
}

void probe_fused_ntt_begin() {
    std::lock_guard<std::mutex> lk(g_probe.mu);
    for (hipEvent_t e : g_probe.ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : g_probe.ev_plain) (void)hipEventDestroy(e);
    g_probe.ev.clear();
    g_probe.ev_plain.clear();
    g_probe.bytes = 0.0;
    g_probe.bytes_plain = 0.0;
    g_probe.on = true;
}

void probe_fused_ntt_end(size_t *launches, double *ms, double *bytes) {
    SHM_HIP(hipDeviceSynchronize());
    std::lock_guard<std::mutex> lk(g_probe.mu);
    g_probe.on = false;
    double total = 0.0;
    for (size_t i = 0; i + 1 < g_probe.ev.size(); i += 2) {
        float t = 0.0f;
        SHM_HIP(hipEventElapsedTime(&t, g_probe.ev[i], g_probe.ev[i + 1]));
        total += t;
    }
    if (launches) *launches = g_probe.ev.size() / 2;
    if (ms) *ms = total;
    if (bytes) *bytes = g_probe.bytes;
    for (hipEvent_t e : g_probe.ev) (void)hipEventDestroy(e);
    g_probe.ev.clear();
    double tp = 0.0;
    for (size_t i = 0; i + 1 < g_probe.ev_plain.size(); i += 2) {
        float t = 0.0f;
        SHM_HIP(hipEventElapsedTime(&t, g_probe.ev_plain[i], g_probe.ev_plain[i + 1]));
        tp += t;
    }
    g_probe.last_plain_launches = g_probe.ev_plain.size() / 2;
    g_probe.last_plain_ms = tp;
    g_probe.last_plain_bytes = g_probe.bytes_plain;
    for (hipEvent_t e : g_probe.ev_plain) (void)hipEventDestroy(e);
    g_probe.ev_plain.clear();
}

void probe_plain_ntt_last(size_t *launches, double *ms, double *bytes) {
    std::lock_guard<std::mutex> lk(g_probe.mu);
    if (launches) *launches = g_probe.last_plain_launches;
    if (ms) *ms = g_probe.last_plain_ms;
    if (bytes) *bytes = g_probe.last_plain_bytes;
}

void set_rescale_tile_events(hipEvent_t start, hipEvent_t stop) {
    g_rescale_tile_ev[0] = start;
    g_rescale_tile_ev[1] = stop;
}

void ntt_forward_encoded(const Context &c, const double2 *fft, uint64_t *out, size_t n_polys, size_t n_limbs,
                         hipStream_t s) {
    LimbMap mp{(uint32_t)n_limbs, 0, 0xffffffffu, 0, 0xffffffffu, (uint32_t)c.nQ, 0, (uint64_t)n_limbs * c.N};
    NttIO io{};
    io.fft = fft;
    launch_ntt<true, IO_ENCODE, IO_PLAIN>(c, out, mp, n_polys * n_limbs, s, io);
}

void ntt_inverse(const Context &c, uint64_t *data, size_t n_polys, size_t n_limbs, size_t first_prime,
                 hipStream_t s) {
    LimbMap mp{(uint32_t)n_limbs, 0, 0xffffffffu, 0, 0xffffffffu, (uint32_t)c.nQ, (uint32_t)first_prime,
               (uint64_t)n_limbs * c.N};
    launch_ntt<false, IO_PLAIN, IO_PLAIN>(c, data, mp, n_polys * n_limbs, s);
}

void ntt_inverse_gather(const Context &c, const PolyPack &in, uint64_t *out, size_t n_polys, size_t n_limbs,
                        size_t first_prime, hipStream_t s) {
    if (n_polys == 0) return;
    if (n_polys > (size_t)kPolyPack) throw InvalidArgument("inverse NTT: 1..64 polynomials per batch");
    LimbMap mp{(uint32_t)n_limbs, 0, 0xffffffffu, 0, 0xffffffffu, (uint32_t)c.nQ, (uint32_t)first_prime,
               (uint64_t)n_limbs * c.N};
    NttIO io{};
    io.acc_p = in;
    launch_ntt<false, IO_PLAIN, IO_GATHER>(c, out, mp, n_polys * n_limbs, s, io);
}

void ntt_forward_qlp(const Context &c, uint64_t *data, size_t Ql, size_t lo, size_t hi, size_t skip_lo,
                     size_t skip_hi, size_t n_polys, size_t poly_stride, hipStream_t s) {
    size_t active = hi - lo;
    uint32_t sl = 0xffffffffu, sw = 0;
    if (skip_hi > skip_lo && skip_lo >= lo && skip_hi <= hi) {
        active -= skip_hi - skip_lo;
        sl = (uint32_t)skip_lo;
        sw = (uint32_t)(skip_hi - skip_lo);
    }
    LimbMap mp{(uint32_t)active, (uint32_t)lo, sl, sw, (uint32_t)Ql, (uint32_t)c.nQ, 0, (uint64_t)poly_stride};
    launch_ntt<true, IO_PLAIN, IO_PLAIN>(c, data, mp, n_polys * active, s);
}

void ntt_inverse_qlp(const Context &c, uint64_t *data, size_t Ql, size_t lo, size_t hi, size_t n_polys,
                     size_t poly_stride, hipStream_t s) {
    LimbMap mp{(uint32_t)(hi - lo), (uint32_t)lo, 0xffffffffu, 0, (uint32_t)Ql, (uint32_t)c.nQ, 0,
               (uint64_t)poly_stride};
    launch_ntt<false, IO_PLAIN, IO_PLAIN>(c, data, mp, n_polys * (hi - lo), s);
}

void rescale_ntt_fused(const Context &c, const uint64_t *last, const uint64_t *acc, size_t acc_limbs,
                       uint64_t *delta, uint64_t *const *outs, uint64_t *out_single, size_t ppo, size_t n_polys,
                       size_t L1, const uint64_t *qinv, const uint64_t *qinv_s, hipStream_t s, bool rounded) {
    LimbMap mp{(uint32_t)L1, 0, 0xffffffffu, 0, 0xffffffffu, (uint32_t)c.nQ, 0, (uint64_t)L1 * c.N};
    NttIO io{};
    io.src = last;
    io.acc = acc;
    io.outs = outs;
    io.out_single = out_single;
    io.qinv = qinv;
    io.qinv_s = qinv_s;
    io.acc_limbs = (uint32_t)acc_limbs;
    io.ppo = (uint32_t)ppo;
    io.center = rounded ? (c.q[L1] >> 1) : 0;
    launch_ntt<true, IO_REDUCE, IO_RESCALE>(c, delta, mp, n_polys * L1, s, io);
}

void rescale_ntt_fused_window(const Context &c, const uint64_t *last, const uint64_t *acc, size_t acc_limbs,
                              uint64_t *delta, uint64_t *const *outs, size_t ppo, size_t n_polys, size_t L1,
                              size_t lo, size_t n, const uint64_t *qinv, const uint64_t *qinv_s, hipStream_t s) {
    if (lo + n > L1 || n == 0) throw InvalidArgument("rescale window outside the output limbs");
    LimbMap mp{(uint32_t)n, (uint32_t)lo, 0xffffffffu, 0, 0xffffffffu, (uint32_t)c.nQ, 0, (uint64_t)L1 * c.N};
    NttIO io{};
    io.src = last;
    io.acc = acc;
    io.outs = outs;
    io.qinv = qinv;
    io.qinv_s = qinv_s;
    io.acc_limbs = (uint32_t)acc_limbs;
    io.ppo = (uint32_t)ppo;
    io.out_limbs = (uint32_t)L1;
    launch_ntt<true, IO_REDUCE, IO_RESCALE>(c, delta, mp, n_polys * n, s, io);
}

void moddown_ntt_fused(const Context &c, uint64_t *delta, const uint64_t *cx, size_t Ql, size_t cx_limbs,
                       const PolyPack &out, size_t n_polys, uint64_t acc_mask, const uint64_t *pinv,
                       const uint64_t *pinv_s, hipStream_t s, const PolyPack *add_src, const uint32_t *add_perm) {
    if (n_polys == 0 || n_polys > (size_t)kPolyPack) throw InvalidArgument("ModDown: 1..64 polynomials per batch");
    if ((add_src == nullptr) != (add_perm == nullptr)) throw LogicError("ModDown: permuted addend needs both parts");
    LimbMap mp{(uint32_t)Ql, 0, 0xffffffffu, 0, 0xffffffffu, (uint32_t)c.nQ, 0, (uint64_t)Ql * c.N};
    NttIO io{};
    io.acc = cx;
    io.acc_limbs = (uint32_t)cx_limbs;
    io.qinv = pinv;
    io.qinv_s = pinv_s;
    io.out_packed = 1;
    io.acc_mask = acc_mask;
    io.out_p = out;
    if (add_src) {
        io.add_src = *add_src;
        io.add_perm = add_perm;
    }
    launch_ntt<true, IO_PLAIN, IO_RESCALE>(c, delta, mp, n_polys * Ql, s, io);
}

void rescale_ntt_fused_pack(const Context &c, const uint64_t *last, const PolyPack &in, const PolyPack &out,
                            uint64_t *delta, size_t n_polys, size_t L1, const uint64_t *qinv, const uint64_t *qinv_s,
                            hipStream_t s, bool rounded) {
    if (n_polys == 0 || n_polys > (size_t)kPolyPack) throw InvalidArgument("rescale: 1..64 polynomials per batch");
    LimbMap mp{(uint32_t)L1, 0, 0xffffffffu, 0, 0xffffffffu, (uint32_t)c.nQ, 0, (uint64_t)L1 * c.N};
    NttIO io{};
    io.src = last;
    io.qinv = qinv;
    io.qinv_s = qinv_s;
    io.center = rounded ? (c.q[L1] >> 1) : 0;
    io.packed = 1;
    io.acc_p = in;
    io.out_p = out;
    launch_ntt<true, IO_REDUCE, IO_RESCALE>(c, delta, mp, n_polys * L1, s, io);
}

}  // namespace shm
