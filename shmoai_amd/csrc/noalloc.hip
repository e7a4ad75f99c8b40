// Allocation-free rescale: the scratch of the rescale kernels (the gathered last
// limbs and the NTT correction) lives in one file-scope device array instead of
// the stream-ordered pool, so the call makes no device allocation. For callers
// that cannot allocate (code under graph capture, tight serving loops).
//
// The array is a per-device object: each GPU holds its own copy, looked up once
// per device with hipGetSymbolAddress and used only through that address. Every
// call on a GPU shares that GPU's array, so calls that overlap on one GPU (other
// streams, other threads) race; the kernels write every scratch word they read,
// so nothing needs clearing between calls.
#include <map>
#include <mutex>

#include "engine.h"

namespace shm {

namespace {

// 8 Mi words = 64 MiB per GPU: npolys x Ql x N words, e.g. a 2-polynomial
// ciphertext at N = 2^16 with up to 64 limbs.
constexpr size_t kScratchWords = size_t(8) << 20;
static __device__ __align__(256) uint64_t scratch[kScratchWords];

std::mutex g_mu;
std::map<int, uint64_t *> g_ptr;  // device -> address of that device's scratch array

uint64_t *scratch_on_current_device() {
    int dev = 0;
    SHM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_ptr.find(dev);
    if (it != g_ptr.end()) return it->second;
    void *p = nullptr;
    SHM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw DeviceError("hipGetSymbolAddress returned no address for the scratch array");
    g_ptr[dev] = static_cast<uint64_t *>(p);
    return static_cast<uint64_t *>(p);
}

}  // namespace

size_t noalloc_scratch_words() { return kScratchWords; }

void rescale_inplace_noalloc(Context &c, Ciphertext &ct, hipStream_t s) {
    const size_t Ql = ct.Ql(), N = c.N, np = ct.size;
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    if (!ct.ntt) throw InvalidArgument("CKKS encrypted must be in NTT form");
    const size_t need = np * N + np * (Ql - 1) * N;
    if (need > kScratchWords)
        throw InvalidArgument("rescale_to_next_inplace_noalloc: the ciphertext needs " + std::to_string(need * 8) +
                              " bytes of scratch, more than the " + std::to_string(kScratchWords * 8) +
                              " bytes of the static array");
    uint64_t *base = scratch_on_current_device();
    launch_rescale_scratch(c, ct.data(), ct.data(), Ql, np, s, false, base, base + np * N);
    ct.chain += 1;
    ct.scale = ct.scale / (double)c.q[Ql - 1];
}

}  // namespace shm
